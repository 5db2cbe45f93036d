"""Cholesky solve of a symmetric positive definite system on the GPU.

Drop-in for fml.math.cho_solve (fml/math/cho_solve.py:33-58 ->
fcho_solve.f90:23-47: LAPACK dpotrf('U') then dpotrs('U'), reading only the
upper triangle of A).

Deviations from the reference, by design:
  * the caller's A is never overwritten (the reference's F2PY layer passes a
    Fortran-ordered float64 A by reference and dpotrf overwrites it);
  * a failed factorisation (non-positive pivot) prints the reference's
    warning line and returns NaNs instead of the partially factored result.
"""
from __future__ import annotations

import numpy as np

from .. import _native
from ..device import empty, is_device_tensor, to_device, to_host, torch, workspace


class CholeskyFactor:
    """Device-resident upper Cholesky factor (row-major lower view)."""

    def __init__(self, m, info: int):
        self.m = m          # (n, n) tensor: Fortran memory of U, lower view
        self.info = info

    @property
    def n(self) -> int:
        return int(self.m.shape[0])

    def solve(self, y):
        return cho_solve_factored(self, y)


def _fortran_memory(A):
    """(n, n) contiguous device tensor holding A in Fortran order."""
    if is_device_tensor(A):
        t = torch()
        return A.t().to(t.float64).contiguous() if A.dtype != torch().float32 else A.t().contiguous()
    return to_device(np.asarray(A).T, np.float64 if np.asarray(A).dtype != np.float32 else np.float32)


def potrf_inplace(m, sync: bool = True):
    """Factor the (n, n) device tensor m (Fortran memory of A) in place.
    Returns info (0 ok, k>0 leading minor k not positive definite)."""
    t = torch()
    n = int(m.shape[0])
    info = t.zeros(1, dtype=t.int32, device="cuda")
    fn = "fmlh_potrf" if m.dtype == t.float64 else "fmlh_spotrf"
    _native.call(fn, _native.ptr(m), m.stride(0), n, _native.ptr(info), None, 0, _native.stream_ptr())
    if not sync:
        return info
    return int(info.item())


def potrf_ozaki_inplace(m, sync: bool = True):
    """potrf_inplace with the next-panel and trailing updates on INT8 matrix
    cores (fmlh_potrf_ozaki, FP64 data only): a factor for FP64 iterative
    refinement, within about 2^-40 per update of the FP64 factor."""
    t = torch()
    if m.dtype != t.float64:
        raise TypeError("the Ozaki-split factor takes float64 matrices")
    n = int(m.shape[0])
    info = t.zeros(1, dtype=t.int32, device="cuda")
    wsb = _native.load().fmlh_potrf_ozaki_ws_bytes(n)
    ws = workspace(wsb)
    _native.call("fmlh_potrf_ozaki", _native.ptr(m), m.stride(0), n, _native.ptr(info), _native.ptr(ws),
                 wsb, _native.stream_ptr())
    if not sync:
        return info
    return int(info.item())


def potrs_inplace(m, x):
    """Solve with the factor m for x (n,) or (nrhs, n) rows, in place."""
    t = torch()
    n = int(m.shape[0])
    nrhs = 1 if x.dim() == 1 else int(x.shape[0])
    lib = _native.load()
    if m.dtype == t.float64:
        wsb = lib.fmlh_potrs_ws_bytes(n, nrhs)
        fn = "fmlh_potrs"
    else:
        wsb = lib.fmlh_spotrs_ws_bytes(n, nrhs)
        fn = "fmlh_spotrs"
    ws = workspace(wsb)
    _native.call(fn, _native.ptr(m), m.stride(0), n, _native.ptr(x), n, nrhs, _native.ptr(ws), wsb,
                 _native.stream_ptr())
    return x


def potrs_correction_inplace(m, x):
    """The correction solve of iterative refinement with the FP64 factor m
    for one right-hand side x (n,), in place: full diagonal blocks applied as
    their inverses (fmlh_potrs_correction)."""
    n = int(m.shape[0])
    if m.dtype != torch().float64 or x.dim() != 1:
        raise ValueError("potrs_correction_inplace takes an FP64 factor and one right-hand side")
    wsb = _native.load().fmlh_potrs_ws_bytes(n, 1)
    ws = workspace(wsb)
    _native.call("fmlh_potrs_correction", _native.ptr(m), m.stride(0), n, _native.ptr(x), _native.ptr(ws),
                 wsb, _native.stream_ptr())
    return x


def cho_factor(A) -> CholeskyFactor:
    m = _fortran_memory(A)
    info = potrf_inplace(m)
    return CholeskyFactor(m, info)


def cho_solve_factored(F: CholeskyFactor, y):
    dev = is_device_tensor(y)
    x = to_device(y, np.float64 if F.m.dtype == torch().float64 else np.float32).clone()
    if F.info != 0:
        x.fill_(float("nan"))
    else:
        potrs_inplace(F.m, x)
    return x if dev else to_host(x)


def potri_inplace(m):
    """Inverse from the factor held in m (as left by potrf_inplace): m's
    row-major lower triangle (the upper triangle of the Fortran matrix)
    receives inv(A)'s; returns info."""
    t = torch()
    n = int(m.shape[0])
    info = t.zeros(1, dtype=t.int32, device="cuda")
    lib = _native.load()
    wsb = lib.fmlh_potri_ws_bytes(n)
    ws = workspace(wsb)
    _native.call("fmlh_potri", _native.ptr(m), m.stride(0), n, _native.ptr(info), _native.ptr(ws), wsb,
                 _native.stream_ptr())
    return int(info.item())


def cho_invert(A):
    """Inverse of a symmetric positive definite matrix through its Cholesky
    factor (cho_solve.py:27-30 -> fcho_solve.f90:49-68: dpotrf('U') then
    dpotri('U'), both on the upper triangle).  Returns the Fortran-ordered
    float64 result the reference's call leaves: inv(A) in the upper triangle,
    A's own values in the strictly lower triangle.

    Deviations, by design: the caller's A is never overwritten (the
    reference's asfortranarray + in-place call clobbers a Fortran-ordered
    float64 A); after a failed factorisation the reference's warning line is
    printed and the upper triangle is NaN instead of undefined."""
    if len(A.shape) != 2 or A.shape[0] != A.shape[1]:
        raise ValueError('expected square matrix')
    dev = is_device_tensor(A)
    m = _fortran_memory(A)
    if m.dtype != torch().float64:
        m = m.double()
    n = int(m.shape[0])
    if n:
        info = potrf_inplace(m)
        if info != 0:
            print(" WARNING: Cholesky decomposition DPOTRF() exited with error code:", info)
            tri = torch().tril(torch().ones(n, n, dtype=torch().bool, device="cuda"))
            m.masked_fill_(tri, float("nan"))
        else:
            info = potri_inplace(m)
            if info != 0:
                print(" WARNING: Cholesky Inversion DPOTRI() exited with error code:", info)
    return m.t() if dev else to_host(m).T


def cho_solve(A, y):
    """Solves A x = y for x with a Cholesky factorisation (cho_solve.py:33-58)."""
    if len(A.shape) != 2 or A.shape[0] != A.shape[1]:
        raise ValueError('expected square matrix')
    if len(y.shape) != 1 or y.shape[0] != A.shape[1]:
        raise ValueError('expected matrix and vector of same stride size')
    F = cho_factor(A)
    if F.info != 0:
        print(" WARNING: Cholesky decomposition DPOTRF() exited with error code:", F.info)
    return cho_solve_factored(F, y)
