"""ARAD descriptor generation (host, numpy): fml/arad.py:27-145.

For each atom i, with every atom k of the molecule (i itself included) as
neighbour at distance d_k and weight w_k = 1 - sin(pi d_k / (2 cut)):
the angle a_mk at i between neighbours m and k is arccos(sp / (d_m d_k))
(sp the dot product of the two bond vectors), set to 0 when
|sp - d_m d_k| <= eps or |d_m d_k| <= eps (eps = 100 * DBL_EPS);
neighbours are ordered by distance and kept while d < cut; the five
channels are d, PTP row, PTP column, sum_k cos(a_mk) w_k / W and
sum_k sin(a_mk) w_k / W with W the weight sum over kept neighbours; empty
distance slots hold 1e100 (arad.py:93).

The reference breaks distance ties with numpy's default (unstable) argsort;
here ties keep the input order (stable sort).  Tied neighbours have equal
distances, so this only reorders terms of equal weight.
"""
from __future__ import annotations

import numpy as np

from .ptp import PTP


class ARAD(object):

    def __init__(self, maxMolSize=30, maxAts=30, cut=5., debug=False):
        self.tag = 'coords'
        self.debug = debug
        self.maxMolSize = maxMolSize
        self.maxAts = maxAts
        self.cut = cut
        self.epsilon = 100.0 * np.finfo(float).eps
        self.PTP = PTP

    def getAngle(self, sp, norms):
        angles = np.zeros(sp.shape)
        mask = np.logical_and(np.abs(sp - norms) > self.epsilon, np.abs(norms) > self.epsilon)
        angles[mask] = np.arccos(sp[mask] / norms[mask])
        return angles

    def describe(self, coords, ocupationList, cell=None):
        if cell is not None:
            raise NotImplementedError("periodic ARAD (arad.py:76-87) is not part of this build")
        coords = np.asarray(coords, dtype=np.float64)
        charges = np.asarray(ocupationList)
        L = len(coords)
        M = np.zeros((self.maxMolSize, 5, self.maxAts))
        M[:, 0, :] = 1E+100
        types = np.asarray([self.PTP[int(o)] for o in charges], dtype=np.float64).reshape(-1, 2)
        for i in range(L):
            vec = coords - coords[i]
            sp = np.sum(vec[:, np.newaxis] * vec[np.newaxis, :], axis=2)
            dist = np.sqrt(np.sum(vec ** 2, axis=1))
            weight = 1. - np.sin(np.pi * dist / (2. * self.cut))
            ang = self.getAngle(sp, dist[:, np.newaxis] * dist[np.newaxis, :])
            cos_w = np.cos(ang) * weight[np.newaxis, :]
            sin_w = np.sin(ang) * weight[np.newaxis, :]
            order = np.argsort(dist, kind='stable')
            keep = order[dist[order] < self.cut]
            k = len(keep)
            dk = dist[keep]
            wsum = np.sum(1.0 - np.sin(np.pi * dk[np.newaxis, :] / (2.0 * self.cut)))
            M[i, 0, :k] = dk
            M[i, 1, :k] = types[keep, 0]
            M[i, 2, :k] = types[keep, 1]
            M[i, 3, :k] = np.sum(cos_w[np.ix_(keep, keep)], axis=1) / wsum
            M[i, 4, :k] = np.sum(sin_w[np.ix_(keep, keep)], axis=1) / wsum
        return M
