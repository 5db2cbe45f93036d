"""ARAD descriptor generation, fml/arad.py:27-145: the host class (numpy,
one molecule per call) and the device form (generate_arad_descriptors, many
molecules per launch).

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

from . import _native
from .device import check_offsets, empty, is_device_tensor, to_device, to_host
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


# ---------------------------------------------------------------------------
# Device form: one wavefront per molecule (fmlh_arad_describe)
# ---------------------------------------------------------------------------

MAX_DEVICE_ATOMS = 64  # atoms per molecule of the device form (one lane each)

_PTP_TABLES = {}


def ptp_table() -> np.ndarray:
    """(max Z + 1, 2) float64 PTP (row, column) per nuclear charge; charges
    without an entry have the row (0, 0)."""
    t = np.zeros((max(PTP) + 1, 2))
    for z, rc in PTP.items():
        t[z] = rc
    return t


def _ptp_device():
    """The PTP table in HBM, one copy per device, kept for the process."""
    t = _native.torch_mod()
    dev = t.cuda.current_device()
    tab = _PTP_TABLES.get(dev)
    if tab is None:
        tab = to_device(ptp_table(), np.float64)
        _PTP_TABLES[dev] = tab
    return tab


def describe_into(z, r, off, nmol, amax, cut, q, zt, n, status):
    """Enqueues fmlh_arad_describe on the current stream (device tensors)."""
    tab = _ptp_device()
    _native.call("fmlh_arad_describe", _native.ptr(z), _native.ptr(r), _native.ptr(off), nmol, amax,
                 float(cut), _native.ptr(tab), int(tab.shape[0]), _native.ptr(q), _native.ptr(zt),
                 _native.ptr(n), _native.ptr(status), _native.stream_ptr())


def generate_arad_descriptors(charges, coords, offsets, amax: int = 23, cut: float = 5.0,
                              on_device: bool = False, return_status: bool = False):
    """ARAD descriptors of many molecules in one launch: the device form of
    ARAD(maxMolSize=amax, maxAts=amax, cut=cut).describe(coords, charges) per
    molecule (fml/arad.py:70-145, entry Molecule.generate_arad_descriptor,
    fml/fml.py:80-86), together with the per-atom PTP types and atom counts
    the kernel functions take (kernels.py:261-292).

    charges (total,), coords (total, 3), offsets (nmol+1,) int: molecule m
    owns atoms offsets[m]:offsets[m+1] (at most 64 atoms).  Returns
    q (nmol, amax, 5, amax), zt (nmol, amax, 2) float64 and n (nmol,) int32,
    on the GPU when on_device, else as numpy arrays.  A molecule with more
    than amax atoms, or a charge that is not a key of the PTP table, raises
    ValueError (the reference fails there too: a broadcast error or a
    KeyError, arad.py:98,138); with return_status the call does not raise
    and returns a fourth array, 1 for such molecules (their block is all
    padding and their count 0).

    Distances, types and the neighbour order equal ARAD.describe's bit for
    bit (same operation order, numpy's pairwise summation order); channels
    3 and 4 differ from it only by the device sin/cos/acos rounding.
    """
    if not 1 <= int(amax) <= MAX_DEVICE_ATOMS:
        raise ValueError(f"amax must be in 1..{MAX_DEVICE_ATOMS}, got {amax}")
    if not float(cut) > 0.0:
        raise ValueError(f"cut must be positive, got {cut}")
    z = to_device(charges, np.float64)
    r = to_device(np.asarray(coords).reshape(-1, 3) if not is_device_tensor(coords) else coords,
                  np.float64)
    if int(r.shape[0]) != int(z.shape[0]):
        raise ValueError(f"{int(r.shape[0])} coordinates for {int(z.shape[0])} charges")
    check_offsets(offsets, int(z.shape[0]))
    off = to_device(offsets, np.int32)
    nmol = int(off.shape[0]) - 1
    amax = int(amax)
    q = empty((nmol, amax, 5, amax), np.float64)
    zt = empty((nmol, amax, 2), np.float64)
    n = empty((nmol,), np.int32)
    status = empty((max(nmol, 1),), np.int32)
    if nmol:
        describe_into(z, r, off, nmol, amax, cut, q, zt, n, status)
    st = status[:nmol]
    if not return_status and nmol and int(st.max().item()) != 0:
        bad = int(np.nonzero(to_host(st))[0][0])
        raise ValueError(f"molecule {bad} has more than amax={amax} atoms or a charge "
                         "without a PTP entry")
    outs = [q, zt, n] + ([st] if return_status else [])
    if not on_device:
        outs = [to_host(o) for o in outs]
    return tuple(outs)
