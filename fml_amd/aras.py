"""ARAS descriptor generation (host, numpy): fml/aras.py:27-131.

For each atom i, with every atom k of the molecule (i itself included) as
neighbour: the distances d_k, the angle at i between neighbours m and k,
arccos(clip(sp / (d_m d_k), -1, 1)) with sp the dot product of the two
bond vectors, set to 0 when |sp - d_m d_k| <= eps or |d_m d_k| <= eps
(eps = 10 * DBL_EPS, aras.py:29-33, 39); neighbours are ordered by distance
and kept while d < cut.  Row i of the (maxMolSize, 3 + maxAts, maxAts)
result holds the distances (row 0, 1e100 where empty), the neighbours'
nuclear charges (row 1) and the angle matrix of the kept neighbours
(rows 3.., aras.py:123-129); row 2 stays zero.

The reference orders neighbours with numpy's default (unstable) argsort;
here ties keep the input order (stable sort).  Tied neighbours have equal
distances, so this only permutes equal entries of row 0.
"""
from __future__ import annotations

import numpy as np


class ARAS(object):

    def __init__(self, maxMolSize=30, maxAts=30, cut=5., debug=False):
        self.tag = 'coords'
        self.debug = debug
        self.maxMolSize = maxMolSize
        self.maxAts = maxAts
        self.cut = cut
        self.epsilon = 10. * np.finfo(float).eps

    def getAngle(self, sp, norms):
        angles = np.zeros(sp.shape)
        mask = np.logical_and(np.abs(sp - norms) > self.epsilon, np.abs(norms) > self.epsilon)
        angles[mask] = np.arccos(np.clip(sp[mask] / norms[mask], -1.0, 1.0))
        return angles

    def describe(self, coords, ocupationList, cell=None):
        if cell is not None:
            raise NotImplementedError("periodic ARAS (aras.py:76-96) is not part of this build")
        coords = np.asarray(coords, dtype=np.float64)
        charges = np.asarray(ocupationList)
        L = len(coords)
        M = np.zeros((self.maxMolSize, 3 + self.maxAts, self.maxAts))
        M[:, 0, :] = 1E+100
        for i in range(L):
            cD = coords - coords[i]
            sp = np.sum(cD[:, np.newaxis] * cD[np.newaxis, :], axis=2)
            D1 = np.sqrt(np.sum(cD ** 2, axis=1))
            angs = self.getAngle(sp, D1[:, np.newaxis] * D1[np.newaxis, :])
            order = np.argsort(D1, kind='stable')
            keep = order[D1[order] < self.cut]
            k = len(keep)
            M[i, 0, :k] = D1[keep]
            M[i, 1, :k] = charges[keep]
            M[i, 3:k + 3, :k] = angs[np.ix_(keep, keep)]
        return M
