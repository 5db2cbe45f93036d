"""ARAS kernels with the reference's Python API (fml/kernels/kernels.py:68-98,
408-542 over fml/kernels/faras_kernels.f90:150-723) on the HIP library.

Descriptors are the (nm, amax, 3 + amax, amax) arrays ARAS.describe makes;
every molecule may hold at most 64 atoms.  Results are the reference's
Fortran-ordered (nsigmas, nm1, nm2) arrays (numpy in -> numpy out, CUDA
tensors in -> CUDA tensors out).
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _native
from ..device import empty, is_device_tensor, to_device, to_host, workspace
from ..ptp import PTP
from .kernels import _ret

_MAX_SIG = 8
_EMAX = 103            # the largest element of the PTP table
_PD = {}               # (r_width, c_width, device) -> device pd table


def periodic_distance(a, b, r_width, c_width):
    """kernels.py:68-82 -- element similarity of nuclear charges a and b:
    (rw^2 cw^2) / ((rw^2 + (row_a - row_b)^2) (cw^2 + (col_a - col_b)^2))."""
    ra, ca = PTP[int(a)]
    rb, cb = PTP[int(b)]
    return (r_width ** 2 * c_width ** 2) / ((r_width ** 2 + (ra - rb) ** 2) * (c_width ** 2 + (ca - cb) ** 2))


def gen_pd(emax=100, r_width=0.001, c_width=0.001):
    """kernels.py:84-98 -- (emax, emax) matrix of periodic_distance(i+1, j+1)."""
    pd = np.zeros((emax, emax))
    for i in range(emax):
        for j in range(emax):
            pd[i, j] = periodic_distance(i + 1, j + 1, r_width, c_width)
    return pd


def _pd_device(pd):
    """Device Fortran (emax, emax) copy of a host pd matrix."""
    return to_device(np.ascontiguousarray(np.asarray(pd, dtype=np.float64).T))


def _pd_cached(r_width, c_width):
    from ..device import torch
    key = (float(r_width), float(c_width), torch().cuda.current_device())
    t = _PD.get(key)
    if t is None:
        t = _pd_device(gen_pd(_EMAX, r_width, c_width))
        _PD[key] = t
    return t


def _neighbors(X, n, cut):
    """kernels.py:460-472 -- per atom, the number of distances < cut."""
    if is_device_tensor(X):
        t = _native.torch_mod()
        nn = (X[:, :, 0, :] < cut).sum(dim=2)
        idx = t.arange(X.shape[1], device="cuda")[None, :]
        na = t.as_tensor(np.asarray(n, dtype=np.int64), device="cuda")[:, None]
        return t.where(idx < na, nn, t.zeros_like(nn)).to(t.int32).contiguous()
    X = np.asarray(X)
    nn = np.zeros((X.shape[0], X.shape[1]), dtype=np.int32)
    for a in range(X.shape[0]):
        ni = int(n[a])
        nn[a, :ni] = np.sum(X[a, :ni, 0, :] < cut, axis=1)
    return nn


def aras_block_plan(n, amax, own=None):
    """Set-1 molecule blocks of at most 64 atoms (one wavefront lane each),
    on the host: (nblk + 1,) int32 block boundaries into the molecule order
    and that order ((nm,) int32, or None for the identity).  own: (nm, amax)
    nuclear charge of each atom (descriptor row 1, column 0) -- when given,
    molecules are ordered by their set of elements, then by size, so that a
    block's atoms share elements and neighbour counts; the matrix does not
    depend on the order."""
    from .atomic import plan_blocks
    n = np.clip(np.asarray(n, dtype=np.int64), 0, amax)
    order = None
    if own is not None and len(n):
        own = np.asarray(own)
        mask = np.arange(own.shape[1])[None, :] < n[:, None]
        elems = np.unique(own[mask])
        key = np.zeros(len(n), dtype=np.int64)
        for b, z in enumerate(elems[:62]):
            key |= np.any((own == z) & mask, axis=1).astype(np.int64) << b
        order = np.lexsort((n, key)).astype(np.int32)
        n = n[order]
    return plan_blocks(n, tile=64), order


def aras_blocks(n, amax, own=None):
    """aras_block_plan on the device: (blk, order or None, nblk)."""
    b, order = aras_block_plan(n, amax, own)
    return (to_device(b, np.int32), None if order is None else to_device(order, np.int32),
            int(b.shape[0]) - 1)


def _aras_call(mode, x1, n1, nn1, x2, n2, nn2, sigmas, t_width, d_width, cut, order, pd, scale):
    X1 = to_device(x1)
    nm1, amax = int(X1.shape[0]), int(X1.shape[1])
    if tuple(X1.shape[2:]) != (3 + amax, amax):
        raise ValueError(f"ARAS descriptors must be (nm, amax, 3 + amax, amax), got {tuple(X1.shape)}")
    if amax > 64:
        raise ValueError("the device form takes molecules of at most 64 atoms")
    N1 = to_device(np.asarray(n1, dtype=np.int32) if not is_device_tensor(n1) else n1, np.int32)
    NN1 = to_device(nn1, np.int32)
    if mode == 0:
        X2 = to_device(x2)
        if tuple(X2.shape[1:]) != tuple(X1.shape[1:]):
            raise ValueError("ERROR: Check ARAS decriptor sizes!")
        N2 = to_device(np.asarray(n2, dtype=np.int32) if not is_device_tensor(n2) else n2, np.int32)
        NN2 = to_device(nn2, np.int32)
        nm2 = int(X2.shape[0])
    else:
        X2 = N2 = NN2 = None
        nm2 = nm1
    sig = np.ascontiguousarray(np.asarray(sigmas, dtype=np.float64).reshape(-1))
    ns = sig.shape[0]
    out = empty((nm2, nm1, ns), np.float64)
    if nm1 == 0 or nm2 == 0 or ns == 0:
        return out
    emax = int(pd.shape[0])
    own = None
    if mode == 0:   # the symmetric form keeps the identity: it needs i >= j inside a block
        own = to_host(X1[:, :, 1, 0]) if is_device_tensor(X1) else np.asarray(X1)[:, :, 1, 0]
    blk, ordv, nblk = aras_blocks(to_host(n1) if is_device_tensor(n1) else n1, amax, own)
    lib = _native.load()
    wsb = lib.fmlh_aras_ws_bytes(nm1, nm2, amax, int(order), mode)
    ws = workspace(wsb)
    status = empty((1,), np.int32)
    p = _native.ptr
    for c0 in range(0, ns, _MAX_SIG):
        chunk = sig[c0:c0 + _MAX_SIG]
        arr = (ctypes.c_double * len(chunk))(*chunk.tolist())
        tmp = out if len(chunk) == ns else empty((nm2, nm1, len(chunk)), np.float64)
        _native.call("fmlh_aras_kernels", mode, p(X1), p(N1), p(NN1), nm1, p(X2), p(N2), p(NN2), nm2, amax,
                     arr, len(chunk), float(t_width), float(d_width), float(cut), int(order), p(pd), emax,
                     float(scale), p(blk), p(ordv), nblk, p(tmp), p(status), p(ws), wsb, _native.stream_ptr())
        if int(status.item()) != 0:
            raise ValueError("ARAS descriptors hold a charge that is not an element of the PTP table, "
                             "more than 16 distinct elements over both sets, or a molecule larger than amax")
        if tmp is not out:
            out[:, :, c0:c0 + len(chunk)] = tmp
    return out


def fget_kernels_aras(x1, x2, n1, n2, nneigh1, nneigh2, sigmas, nm1=None, nm2=None, nsigmas=None,
                      t_width=np.pi, r_width=1.0, c_width=0.5, d_width=0.2, cut_distance=5.0, order=1,
                      pd=None, angular_scale=0.1):
    """faras_kernels.f90:150 -> (nsigmas, nm1, nm2) Fortran order.  pd: the
    (emax, emax) element similarity (gen_pd(103, r_width, c_width) when
    None)."""
    pdd = _pd_cached(r_width, c_width) if pd is None else _pd_device(pd)
    out = _aras_call(0, x1, n1, nneigh1, x2, n2, nneigh2, sigmas, t_width, d_width, cut_distance, order,
                     pdd, angular_scale)
    return _ret(out, is_device_tensor(x1), (2, 1, 0))


def fget_symmetric_kernels_aras(x1, n1, nneigh1, sigmas, nm1=None, nsigmas=None, t_width=np.pi,
                                r_width=1.0, c_width=0.5, d_width=0.2, cut_distance=5.0, order=1,
                                pd=None, angular_scale=0.1):
    """faras_kernels.f90:486 -> (nsigmas, nm1, nm1) Fortran order."""
    pdd = _pd_cached(r_width, c_width) if pd is None else _pd_device(pd)
    out = _aras_call(1, x1, n1, nneigh1, None, None, None, sigmas, t_width, d_width, cut_distance, order,
                     pdd, angular_scale)
    return _ret(out, is_device_tensor(x1), (2, 1, 0))


def get_atomic_kernels_aras(X1, X2, Z1, Z2, sigmas, t_width=np.pi / 1.0, d_width=0.2, cut_distance=5.0,
                            r_width=1.0, order=1, c_width=0.5, scale_angular=0.1):
    """kernels.py:412-489 -- Gaussian kernel matrices of ARAS descriptors of
    two molecule sets for a list of sigmas: (nsigmas, nm1, nm2).  Z1, Z2:
    the molecules' nuclear charges (their lengths are the atom counts).
    The reference prints the descriptor shapes and neighbour arrays; this
    function prints nothing."""
    n1 = np.array([len(z) for z in Z1], dtype=np.int32)
    n2 = np.array([len(z) for z in Z2], dtype=np.int32)
    s1, s2 = tuple(X1.shape), tuple(X2.shape)
    if s1[0] != len(n1) or s2[0] != len(n2) or s2[1] != s1[1] or s2[3] != s1[3]:
        raise ValueError("ERROR: Check ARAS decriptor sizes!")
    nn1, nn2 = _neighbors(X1, n1, cut_distance), _neighbors(X2, n2, cut_distance)
    return fget_kernels_aras(X1, X2, n1, n2, nn1, nn2, sigmas, t_width=t_width, r_width=r_width,
                             c_width=c_width, d_width=d_width, cut_distance=cut_distance, order=order,
                             angular_scale=scale_angular)


def get_atomic_symmetric_kernels_aras(X1, Z1, sigmas, t_width=np.pi / 1.0, d_width=0.2, cut_distance=5.0,
                                      r_width=1.0, order=1, c_width=0.5, scale_angular=0.1):
    """kernels.py:492-542 -- the symmetric (nsigmas, nm1, nm1) form."""
    n1 = np.array([len(z) for z in Z1], dtype=np.int32)
    if tuple(X1.shape)[0] != len(n1):
        raise ValueError("ERROR: Check ARAS decriptor sizes!")
    nn1 = _neighbors(X1, n1, cut_distance)
    return fget_symmetric_kernels_aras(X1, n1, nn1, sigmas, t_width=t_width, r_width=r_width,
                                       c_width=c_width, d_width=d_width, cut_distance=cut_distance,
                                       order=order, angular_scale=scale_angular)
