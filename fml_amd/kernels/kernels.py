"""Gaussian, Laplacian and ARAD kernels with the reference's Python API
(fml/kernels/kernels.py:136-349) over the HIP library.

Layouts follow the reference: the dense kernels take descriptor matrices
with one sample per COLUMN, (d, n), and return Fortran-ordered (na, nb)
arrays; ARAD descriptors are (nm, amax, 5, amax) and kernels come back as
Fortran-ordered (nsigmas, nm1, nm2).  numpy in -> numpy out; CUDA tensors in
-> CUDA tensors out (same shapes, Fortran order expressed as a transposed
view of a contiguous tensor).
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _native
from ..device import empty, is_device_tensor, to_device, to_host, torch, workspace
from ..ptp import PTP

_MAX_SIG = 8


# ---------------------------------------------------------------------------
# dense kernels
# ---------------------------------------------------------------------------

_PRECISION = {"fp64": None, "fp32": 0, "bf16x3": 1}


def _rows(x, dtype):
    """(d, n) column-per-sample -> contiguous (n, d) device rows."""
    if is_device_tensor(x):
        t = torch()
        tdt = t.float64 if dtype == np.float64 else t.float32
        return x.t().to(tdt).contiguous()
    return to_device(np.asarray(x).T, dtype)


def gaussian_kernel_rows(X1, X2, sigma, precision="fp64", out=None, out_dtype=None, recomputed=None):
    """Device form: X1 (na, d), X2 (nb, d) contiguous CUDA tensors ->
    tensor (nb, na) whose [i, j] is K(j, i) (the memory of the Fortran
    (na, nb) kernel).  precision: fp64 | fp32 (FP32 MFMA) | bf16x3.
    recomputed: optional device int64 tensor of one element, incremented by
    the number of near pairs the bf16x3 form recomputed."""
    na, d = X1.shape
    nb = X2.shape[0]
    mode = _PRECISION[precision]
    if mode is None:
        k = out if out is not None else empty((nb, na), np.float64)
        _native.call("fmlh_gaussian_kernel", _native.ptr(X1), X1.stride(0), na, _native.ptr(X2),
                     X2.stride(0), nb, d, float(sigma), _native.ptr(k), k.stride(0),
                     _native.stream_ptr())
        return k
    od = out_dtype or np.float32
    k = out if out is not None else empty((nb, na), od)
    is64 = k.dtype == torch().float64
    same = int(X1.data_ptr() == X2.data_ptr() and X1.stride(0) == X2.stride(0) and na == nb)
    wsb = _native.load().fmlh_gaussian_kernel_f32_ws_bytes(na, nb, d, mode, same)
    ws = workspace(wsb) if wsb else None
    _native.call("fmlh_gaussian_kernel_f32", _native.ptr(X1), X1.stride(0), na, _native.ptr(X2),
                 X2.stride(0), nb, d, float(sigma), mode, None if is64 else _native.ptr(k),
                 _native.ptr(k) if is64 else None, k.stride(0), _native.ptr(ws), wsb,
                 _native.ptr(recomputed), _native.stream_ptr())
    return k


def laplacian_kernel_rows(X1, X2, sigma, out=None, out_dtype=None):
    na, d = X1.shape
    nb = X2.shape[0]
    t = torch()
    if X1.dtype == t.float64:
        k = out if out is not None else empty((nb, na), np.float64)
        _native.call("fmlh_laplacian_kernel", _native.ptr(X1), X1.stride(0), na, _native.ptr(X2),
                     X2.stride(0), nb, d, float(sigma), _native.ptr(k), k.stride(0),
                     _native.stream_ptr())
        return k
    od = out_dtype or np.float32
    k = out if out is not None else empty((nb, na), od)
    is64 = k.dtype == t.float64
    _native.call("fmlh_laplacian_kernel_f32", _native.ptr(X1), X1.stride(0), na, _native.ptr(X2),
                 X2.stride(0), nb, d, float(sigma), None if is64 else _native.ptr(k),
                 _native.ptr(k) if is64 else None, k.stride(0), _native.stream_ptr())
    return k


def _dense(fn, A, B, sigma):
    dev = is_device_tensor(A)
    X1 = _rows(A, np.float64)
    X2 = _rows(B, np.float64)
    if X1.shape[1] != X2.shape[1]:
        raise ValueError("A and B must have the same descriptor length")
    k = fn(X1, X2, sigma)
    return k.t() if dev else to_host(k).T


def gaussian_kernel(A, B, sigma):
    """K_ij = exp(-0.5 sigma^-2 ||A_i - B_j||^2) with the squared norm, as
    fkernels.f90:334-342 computes it (kernels.py:167-196).  A (d, na),
    B (d, nb) -> K (na, nb), Fortran order."""
    return _dense(gaussian_kernel_rows, A, B, sigma)


def laplacian_kernel(A, B, sigma):
    """K_ij = exp(-sigma^-1 ||A_i - B_j||_1) (kernels.py:136-164)."""
    return _dense(laplacian_kernel_rows, A, B, sigma)


# ---------------------------------------------------------------------------
# ARAD
# ---------------------------------------------------------------------------

def arad_types(Z, amax):
    """(nm, amax, 2) PTP (row, col) per atom and (nm,) int32 atom counts,
    as kernels.py:261-292 builds Z*_arad and N*."""
    nm = len(Z)
    za = np.zeros((nm, amax, 2))
    n = np.zeros(nm, dtype=np.int32)
    for i, charges in enumerate(Z):
        n[i] = len(charges)
        for j, q in enumerate(charges):
            za[i, j] = PTP[int(q)]
    return za, n


def _sig_host(sigmas):
    s = np.ascontiguousarray(np.asarray(sigmas, dtype=np.float64).reshape(-1))
    return s


def _arad_call(mode, q1, z1, n1, q2, z2, n2, sigmas, width, cut, rw, cw, precision=0, tiling=None):
    """Runs one ARAD molecular-mode launch per chunk of <= 8 sigmas.
    Returns the device output tensor in its C-order storage shape.
    tiling: an _native.AradTiling, or None for the library's current choice
    (autotuned when fml_amd.tuning has run for this shape)."""
    dq1, dz1, dn1 = to_device(q1), to_device(z1), to_device(n1, np.int32)
    nm1, amax = int(dq1.shape[0]), int(dq1.shape[1])
    if q2 is None:
        dq2 = dz2 = dn2 = None
        nm2 = nm1
    else:
        dq2, dz2, dn2 = to_device(q2), to_device(z2), to_device(n2, np.int32)
        nm2 = int(dq2.shape[0])
    lib = _native.load()
    wsb = lib.fmlh_arad_ws_bytes(nm1, nm2, amax, mode)
    ws = workspace(wsb)
    tp = ctypes.byref(tiling) if tiling is not None else None
    if mode == _native.ARAD_DISTANCE:
        out = empty((amax, amax, nm2, nm1), np.float64)
        _native.call("fmlh_arad_kernels", mode, _native.ptr(dq1), _native.ptr(dz1), _native.ptr(dn1),
                     nm1, _native.ptr(dq2), _native.ptr(dz2), _native.ptr(dn2), nm2, amax, None, 0,
                     width, cut, rw, cw, precision, _native.ptr(out), _native.ptr(ws), wsb, tp,
                     _native.stream_ptr())
        return out
    sig = _sig_host(sigmas)
    if mode == _native.ARAD_KERNEL:
        out = empty((nm1, nm1), np.float64)
        arr = (ctypes.c_double * 1)(float(sig[0]))
        _native.call("fmlh_arad_kernels", mode, _native.ptr(dq1), _native.ptr(dz1), _native.ptr(dn1),
                     nm1, None, None, None, nm1, amax, arr, 1, width, cut, rw, cw, precision,
                     _native.ptr(out), _native.ptr(ws), wsb, tp, _native.stream_ptr())
        return out
    ns = sig.shape[0]
    out = empty((nm2, nm1, ns), np.float64)
    for c0 in range(0, ns, _MAX_SIG):
        chunk = sig[c0:c0 + _MAX_SIG]
        arr = (ctypes.c_double * len(chunk))(*chunk.tolist())
        tmp = out if (c0 == 0 and len(chunk) == ns) else empty((nm2, nm1, len(chunk)), np.float64)
        _native.call("fmlh_arad_kernels", mode, _native.ptr(dq1), _native.ptr(dz1), _native.ptr(dn1),
                     nm1, _native.ptr(dq2), _native.ptr(dz2), _native.ptr(dn2), nm2, amax, arr,
                     len(chunk), width, cut, rw, cw, precision, _native.ptr(tmp), _native.ptr(ws),
                     wsb, tp, _native.stream_ptr())
        if tmp is not out:
            out[:, :, c0:c0 + len(chunk)] = tmp
    return out


def fget_alpha_arad(q, z, n, y, sigma, llambda, alpha=None):
    """farad_kernels.f90:231-436 -- regression coefficients of molecular ARAD
    KRR, one per atom: the molecular kernel K of the training set (the
    kernel of fget_kernel_arad: width .2, cut 6.0, r_width 1, c_width .5),
    lambda added to its diagonal, (K + lambda I)^-1 y from dpotrf('U') +
    dpotri('U') + dsymv('U') (here: fmlh_potrf, fmlh_potri and the
    upper-triangle product), and molecule i's value repeated for each of its
    n(i) atoms.

    q (nm, amax, 5, amax), z (nm, amax, 2), n (nm,), y (nm,).  alpha: the
    caller's (sum(n),) float64 array, filled in place as F2PY fills the
    reference's intent(inout) argument; a new array when None.  Returns
    alpha.  A failed factorisation or inversion prints the reference's
    warning line (the coefficients are then NaN)."""
    dev = is_device_tensor(q)
    counts = to_host(n).astype(np.int64) if is_device_tensor(n) else np.asarray(n, dtype=np.int64)
    K = _arad_call(_native.ARAD_KERNEL, q, z, n, None, None, None, [sigma], 0.2, 6.0, 1.0, 0.5)
    return solve_expand(K, counts, y, llambda, alpha, dev)


def solve_expand(K, counts, y, llambda, alpha, dev):
    """The regression step shared by fget_alpha, fget_alpha_from_distance
    (fkernels.f90:446-476) and fget_alpha_arad (farad_kernels.f90:408-434):
    lambda added to the diagonal of the (nm, nm) device kernel K (Fortran
    memory; overwritten), dpotrf('U') + dpotri('U') (fmlh_potrf, fmlh_potri),
    ly = inv(K + lambda I) y from the upper triangle (dsymv('U')), and
    molecule i's value repeated for its counts[i] atoms.  A failed
    factorisation or inversion prints the reference's warning line (the
    coefficients are then NaN).  alpha: filled in place when given."""
    from ..krr import add_diagonal, sym_residual
    from ..math.cho_solve import potrf_inplace, potri_inplace
    nm = int(K.shape[0])
    add_diagonal(K, float(llambda))
    info = potrf_inplace(K)
    if info > 0:
        print(" WARNING: Cholesky decomposition DPOTRF() exited with error code:", info)
        K.fill_(float("nan"))
    else:
        info = potri_inplace(K)
        if info > 0:
            print(" WARNING: Cholesky inversion DPOTRI() exited with error code:", info)
    yd = to_device(y, np.float64)
    zero = torch().zeros(nm, dtype=torch().float64, device="cuda")
    ly = empty((nm,), np.float64)
    sym_residual(K, 0.0, yd, zero, ly)   # ly = 0 - K^-1 y
    ly.neg_()
    per_atom = torch().repeat_interleave(ly, torch().as_tensor(np.asarray(counts, dtype=np.int64),
                                                                device="cuda"))
    if alpha is None:
        return per_atom if dev else to_host(per_atom)
    if is_device_tensor(alpha):
        alpha.copy_(per_atom)
    else:
        alpha[...] = to_host(per_atom)
    return alpha


def _ret(t, dev, perm):
    return t.permute(*perm) if dev else to_host(t).transpose(perm)


def fget_kernels_arad(q1, q2, z1, z2, n1, n2, sigmas, nm1=None, nm2=None, nsigmas=None,
                      width=0.2, cut_distance=6.0, r_width=1.0, c_width=0.5, precision=0):
    """farad_kernels.f90:439 -> (nsigmas, nm1, nm2) Fortran order."""
    dev = is_device_tensor(q1)
    out = _arad_call(_native.ARAD_GLOBAL, q1, z1, n1, q2, z2, n2, sigmas, width, cut_distance,
                     r_width, c_width, precision)
    return _ret(out, dev, (2, 1, 0))


def fget_symmetric_kernels_arad(q1, z1, n1, sigmas, nm1=None, nsigmas=None, width=0.2,
                                cut_distance=6.0, r_width=1.0, c_width=0.5, precision=0):
    """farad_kernels.f90:619 -> (nsigmas, nm1, nm1) Fortran order."""
    dev = is_device_tensor(q1)
    out = _arad_call(_native.ARAD_SYMMETRIC, q1, z1, n1, None, None, None, sigmas, width,
                     cut_distance, r_width, c_width, precision)
    return _ret(out, dev, (2, 1, 0))


def fget_kernel_arad(q, z, n, sigma, nm=None):
    """farad_kernels.f90:95 (width .2, cut 6.0, r_width 1, c_width .5 fixed)
    -> (nm, nm) Fortran order."""
    dev = is_device_tensor(q)
    out = _arad_call(_native.ARAD_KERNEL, q, z, n, None, None, None, [sigma], 0.2, 6.0, 1.0, 0.5)
    return _ret(out, dev, (1, 0))


def _atomic_call(mode, q1, q2, z1, z2, n1, n2, sigmas, width, cut, rw, cw):
    dq1, dq2 = to_device(q1), to_device(q2)
    dz1, dz2 = to_device(z1), to_device(z2)
    dn1, dn2 = to_device(n1, np.int32), to_device(n2, np.int32)
    na1, na2, amax = int(dq1.shape[0]), int(dq2.shape[0]), int(dq1.shape[2])
    if int(dq2.shape[2]) != amax:
        raise ValueError("ARAD descriptor sizes differ")
    lib = _native.load()
    wsb = lib.fmlh_arad_atomic_ws_bytes(na1, na2, amax)
    ws = workspace(wsb)
    if mode == _native.ARAD_ATOMIC_DISTANCE:
        out = empty((na2, na1), np.float64)
        _native.call("fmlh_arad_atomic", mode, _native.ptr(dq1), _native.ptr(dz1), _native.ptr(dn1),
                     na1, _native.ptr(dq2), _native.ptr(dz2), _native.ptr(dn2), na2, amax, None, 0,
                     width, cut, rw, cw, _native.ptr(out), _native.ptr(ws), wsb, _native.stream_ptr())
        return out
    sig = _sig_host(sigmas)
    ns = sig.shape[0]
    out = empty((na2, na1, ns), np.float64)
    for c0 in range(0, ns, _MAX_SIG):
        chunk = sig[c0:c0 + _MAX_SIG]
        arr = (ctypes.c_double * len(chunk))(*chunk.tolist())
        tmp = out if (c0 == 0 and len(chunk) == ns) else empty((na2, na1, len(chunk)), np.float64)
        _native.call("fmlh_arad_atomic", mode, _native.ptr(dq1), _native.ptr(dz1), _native.ptr(dn1),
                     na1, _native.ptr(dq2), _native.ptr(dz2), _native.ptr(dn2), na2, amax, arr,
                     len(chunk), width, cut, rw, cw, _native.ptr(tmp), _native.ptr(ws), wsb,
                     _native.stream_ptr())
        if tmp is not out:
            out[:, :, c0:c0 + len(chunk)] = tmp
    return out


def fget_atomic_kernels_arad(q1, q2, z1, z2, n1, n2, sigmas, na1=None, na2=None, nsigmas=None,
                             width=0.2, cut_distance=6.0, r_width=1.0, c_width=0.5):
    """farad_kernels.f90:757 -> (nsigmas, na1, na2) Fortran order.
    q: (na, 5, amax) per-atom descriptors, z: (na, 2), n: (na,) atom counts."""
    dev = is_device_tensor(q1)
    out = _atomic_call(_native.ARAD_ATOMIC_KERNELS, q1, q2, z1, z2, n1, n2, sigmas, width,
                       cut_distance, r_width, c_width)
    return _ret(out, dev, (2, 1, 0))


def fget_atomic_distance_arad(q1, q2, z1, z2, n1, n2, na1=None, na2=None, width=0.2,
                              cut_distance=6.0, r_width=1.0, c_width=0.5):
    """farad_kernels.f90:902 -> D (na1, na2) Fortran order with D[i, j] the
    distance of atom i (set 1) and atom j (set 2).  The reference stores that
    value at kernel(j, i) of an (na1, na2) array (farad_kernels.f90:932,1019),
    which is only in bounds for na1 == na2; there its array is the transpose
    of this one."""
    dev = is_device_tensor(q1)
    out = _atomic_call(_native.ARAD_ATOMIC_DISTANCE, q1, q2, z1, z2, n1, n2, None, width,
                       cut_distance, r_width, c_width)
    return _ret(out, dev, (1, 0))


def _check_arad_shapes(X1, X2, Z1, Z2):
    # kernels.py:249-259
    amax = X1.shape[1]
    assert X1.shape[3] == amax, "ERROR: Check ARAD decriptor sizes! code = 1"
    if X2 is not None:
        assert X2.shape[1] == amax, "ERROR: Check ARAD decriptor sizes! code = 2"
        assert X2.shape[3] == amax, "ERROR: Check ARAD decriptor sizes! code = 3"
    assert X1.shape[0] == len(Z1), "ERROR: Check ARAD decriptor sizes! code = 4"
    if X2 is not None:
        assert X2.shape[0] == len(Z2), "ERROR: Check ARAD decriptor sizes! code = 5"
    return amax


def get_atomic_kernels_arad(X1, X2, Z1, Z2, sigmas, width=0.2, cut_distance=6.0, r_width=1.0,
                            c_width=0.5, precision=0):
    """Molecular ARAD kernels for a list of sigmas (kernels.py:229-297) ->
    (nsigmas, N1, N2).  precision=1 selects the FP32 inner-overlap variant."""
    amax = _check_arad_shapes(X1, X2, Z1, Z2)
    z1, n1 = arad_types(Z1, amax)
    z2, n2 = arad_types(Z2, amax)
    return fget_kernels_arad(X1, X2, z1, z2, n1, n2, sigmas, width=width, cut_distance=cut_distance,
                             r_width=r_width, c_width=c_width, precision=precision)


def get_atomic_symmetric_kernels_arad(X1, Z1, sigmas, width=0.2, cut_distance=6.0, r_width=1.0,
                                      c_width=0.5, precision=0):
    """kernels.py:300-349.  The reference wrapper raises NameError because it
    never imports fget_symmetric_kernels_arad (kernels.py:26-39, 348); this
    one calls the symmetric kernel it intends."""
    amax = _check_arad_shapes(X1, None, Z1, None)
    z1, n1 = arad_types(Z1, amax)
    return fget_symmetric_kernels_arad(X1, z1, n1, sigmas, width=width, cut_distance=cut_distance,
                                       r_width=r_width, c_width=c_width, precision=precision)
