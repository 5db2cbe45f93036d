"""The atomic KRR path with the reference's Python API: vector (atomic)
kernels over per-atom descriptors, get_alpha / get_prediction and their
distance forms (fml/kernels/kernels.py:101-134, 351-406 over
fml/kernels/fkernels.f90:23-162, 380-702), on the HIP library.

Atoms are handed to the device as packed rows, molecule after molecule;
every molecule may hold at most 64 atoms.  Arguments may be numpy arrays
(copied to the GPU, results returned as numpy) or CUDA tensors (results
stay on the GPU).  Kernels come back as the reference's Fortran-ordered
(nsigmas, nm1, nm2) arrays.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _native
from ..device import empty, is_device_tensor, to_device, to_host, torch, workspace
from .kernels import _ret, solve_expand

MAX_ATOMS = 64   # largest molecule of the device form
_MAX_SIG = 8     # sigmas per launch


def _counts(n):
    c = to_host(n) if is_device_tensor(n) else np.asarray(n)
    return np.asarray(c, dtype=np.int64).reshape(-1)


def _offsets(counts):
    return np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)


def plan_blocks(counts, tile=128):
    """Molecules packed greedily, in order, into blocks of at most `tile`
    atoms (the device's tile side): (nb + 1,) int32 block starts."""
    off = np.concatenate([[0], np.cumsum(np.asarray(counts, dtype=np.int64))])
    nm = len(off) - 1
    starts = [0]
    m = 0
    while m < nm:
        m2 = int(np.searchsorted(off, off[m] + tile, side="right")) - 1
        m = min(max(m2, m + 1), nm)
        starts.append(m)
    return np.asarray(starts, dtype=np.int32)


def _blocks(counts):
    b = plan_blocks(counts)
    return to_device(b, np.int32), int(b.shape[0]) - 1


def _amax(*counts):
    m = max([int(c.max()) if c.size else 0 for c in counts] + [1])
    if m > MAX_ATOMS:
        raise ValueError(f"a molecule has {m} atoms; the device form takes at most {MAX_ATOMS}")
    return m


def _rows_from_q(q, counts):
    """The reference's q (d, amax, nm) -- q(:, ia, i) is atom ia of molecule
    i -- as packed atom rows (sum(counts), d)."""
    if is_device_tensor(q):
        t = q.permute(2, 1, 0)
        mask = torch().arange(t.shape[1], device="cuda")[None, :] < \
            torch().as_tensor(counts, device="cuda")[:, None]
        return t[mask].to(torch().float64).contiguous()
    t = np.transpose(np.asarray(q, dtype=np.float64), (2, 1, 0))
    if t.shape[0] != counts.shape[0] or (counts.size and counts.max() > t.shape[1]):
        raise ValueError(f"descriptor shape {t.shape[::-1]} does not hold {counts.shape[0]} molecules "
                         f"of up to {int(counts.max()) if counts.size else 0} atoms")
    mask = np.arange(t.shape[1])[None, :] < counts[:, None]
    return to_device(np.ascontiguousarray(t[mask]))


def _rows_from_cols(Q, counts=None):
    """Q (d, na) with one atom per column -> device rows (na, d)."""
    if is_device_tensor(Q):
        x = Q.t().to(torch().float64).contiguous()
    else:
        x = to_device(np.ascontiguousarray(np.asarray(Q, dtype=np.float64).T))
    if counts is not None and int(counts.sum()) != int(x.shape[0]):
        raise ValueError(f"{int(x.shape[0])} atom columns for molecules of {int(counts.sum())} atoms")
    return x


def _transposed(D):
    """A Fortran (m, n) array D(k, a) as device rows a: (n, m) C order."""
    if is_device_tensor(D):
        return D.t().to(torch().float64).contiguous()
    return to_device(np.ascontiguousarray(np.asarray(D, dtype=np.float64).T))


def _check_status(status):
    if int(status.item()) != 0:
        raise RuntimeError("a molecule is larger than the atom bound passed to the device")


def vector_kernels(kind, x1, counts1, x2, counts2, sigmas):
    """Device (nm2, nm1, nsig) molecular sums of the atom-pair Gaussian
    (kind 0; kind 2 with the matrix-core distances) or Laplacian (kind 1)
    kernel over device atom rows x1, x2."""
    nm1, nm2 = int(counts1.shape[0]), int(counts2.shape[0])
    amax = _amax(counts1, counts2)
    sig = np.ascontiguousarray(np.asarray(sigmas, dtype=np.float64).reshape(-1))
    if np.any(~(sig > 0)):
        raise ValueError("sigmas must be positive")
    ns = sig.shape[0]
    out = empty((nm2, nm1, ns), np.float64)
    if nm1 == 0 or nm2 == 0 or ns == 0:
        return out
    d = int(x1.shape[1])
    if int(x2.shape[1]) != d:
        raise ValueError(f"descriptor lengths differ: {d} and {int(x2.shape[1])}")
    off1 = to_device(_offsets(counts1), np.int32)
    off2 = to_device(_offsets(counts2), np.int32)
    blk1, nb1 = _blocks(counts1)
    blk2, nb2 = (blk1, nb1) if counts2 is counts1 else _blocks(counts2)
    status = empty((1,), np.int32)
    for c0 in range(0, ns, _MAX_SIG):
        chunk = sig[c0:c0 + _MAX_SIG]
        arr = (ctypes.c_double * len(chunk))(*chunk.tolist())
        tmp = out if len(chunk) == ns else empty((nm2, nm1, len(chunk)), np.float64)
        _native.call("fmlh_vector_kernels", kind, _native.ptr(x1), x1.stride(0), _native.ptr(off1), nm1,
                     int(x1.shape[0]), _native.ptr(x2), x2.stride(0), _native.ptr(off2), nm2,
                     int(x2.shape[0]), d, amax, arr, len(chunk), _native.ptr(blk1), nb1, _native.ptr(blk2),
                     nb2, _native.ptr(tmp), _native.ptr(status), _native.stream_ptr())
        _check_status(status)
        if tmp is not out:
            out[:, :, c0:c0 + len(chunk)] = tmp
    return out


_GAUSS_KIND = {"fp64": 0, "mfma": 2}


def fget_vector_kernels_gaussian(q1, q2, n1, n2, sigmas, nm1=None, nm2=None, nsigmas=None,
                                 precision="fp64"):
    """fkernels.f90:93 -- q1 (d, amax1, nm1), q2 (d, amax2, nm2) ->
    (nsigmas, nm1, nm2): sum over atom pairs of exp(-|a - b|^2 / (2 sigma^2)).

    precision "fp64": the squared distances summed in the reference's order
    (the reference's distances bit for bit); "mfma": |a|^2 + |b|^2 - 2 a.b
    with the products on the matrix cores -- faster, each distance within
    about 1e-13 (|a|^2 + |b|^2), so an entry K carries a relative error of
    about 1e-12 + 1e-13 |ln K| (noticeable only for entries far below 1)."""
    c1, c2 = _counts(n1), _counts(n2)
    out = vector_kernels(_GAUSS_KIND[precision], _rows_from_q(q1, c1), c1, _rows_from_q(q2, c2), c2,
                         sigmas)
    return _ret(out, is_device_tensor(q1), (2, 1, 0))


def fget_vector_kernels_laplacian(q1, q2, n1, n2, sigmas, nm1=None, nm2=None, nsigmas=None):
    """fkernels.f90:23 -- as fget_vector_kernels_gaussian with
    exp(-sum |a - b| / sigma)."""
    c1, c2 = _counts(n1), _counts(n2)
    out = vector_kernels(1, _rows_from_q(q1, c1), c1, _rows_from_q(q2, c2), c2, sigmas)
    return _ret(out, is_device_tensor(q1), (2, 1, 0))


def get_atomic_kernels_laplacian(x1, x2, N1, N2, sigmas):
    """kernels.py:351-362 -- x1 (d, amax1, nm1), x2 (d, amax2, nm2) in the
    layout the reference passes to fget_vector_kernels_laplacian."""
    return fget_vector_kernels_laplacian(x1, x2, N1, N2, sigmas)


def get_atomic_kernels_gaussian(mols1, mols2, sigmas, precision="fp64"):
    """kernels.py:380-406 -- the molecules' local_coulomb_matrix rows
    (Molecule.generate_local_coulomb_matrix) -> (nsigmas, nm1, nm2);
    precision as fget_vector_kernels_gaussian."""
    c1 = np.array([m.natoms for m in mols1], dtype=np.int64)
    c2 = np.array([m.natoms for m in mols2], dtype=np.int64)
    r1 = np.concatenate([np.asarray(m.local_coulomb_matrix, dtype=np.float64) for m in mols1])
    r2 = np.concatenate([np.asarray(m.local_coulomb_matrix, dtype=np.float64) for m in mols2])
    if r1.shape[0] != c1.sum() or r2.shape[0] != c2.sum():
        raise ValueError("local_coulomb_matrix rows differ from natoms")
    out = vector_kernels(_GAUSS_KIND[precision], to_device(r1), c1, to_device(r2), c2, sigmas)
    return _ret(out, False, (2, 1, 0))


def _mol_kernel_laplacian(x, counts, sigma):
    """(nm, nm) device molecular Laplacian kernel of atom rows x (Fortran
    memory: K(i, j) at [j][i]), fget_alpha's ltcl (fkernels.f90:420-435)."""
    out = vector_kernels(1, x, counts, x, counts, [sigma])
    return out[:, :, 0]


def get_alpha(Q, N, Y, sigma, llambda):
    """kernels.py:119-125 / fkernels.f90:380-478 -- Q (d, na) one atom per
    column, N (nm,) atoms per molecule, Y (nm,) -> alpha (na,): the
    molecular Laplacian kernel (exp(-sum |q_a - q_b| / sigma) summed over
    the atoms of each molecule pair), lambda on its diagonal, inverted
    through its Cholesky factor and applied to Y; each molecule's value
    repeated for its atoms."""
    c = _counts(N)
    K = _mol_kernel_laplacian(_rows_from_cols(Q, c), c, float(sigma))
    return solve_expand(K, c, Y, llambda, None, is_device_tensor(Q))


def fget_alpha(q, n, y, sigma, llambda, alpha):
    """fkernels.f90:380 -- fills alpha (sum(n),) in place."""
    c = _counts(n)
    K = _mol_kernel_laplacian(_rows_from_cols(q, c), c, float(sigma))
    return solve_expand(K, c, y, llambda, alpha, is_device_tensor(q))


def _predict(xt, counts_t, x, alpha, sigma):
    nm_t = int(counts_t.shape[0])
    y = empty((nm_t,), np.float64)
    if nm_t == 0:
        return y
    d = int(xt.shape[1])
    if int(x.shape[1]) != d:
        raise ValueError(f"descriptor lengths differ: {int(x.shape[1])} and {d}")
    na = int(x.shape[0])
    al = to_device(alpha, np.float64)
    if int(al.shape[0]) != na:
        raise ValueError(f"{int(al.shape[0])} coefficients for {na} training atoms")
    amax = _amax(counts_t)
    off = to_device(_offsets(counts_t), np.int32)
    lib = _native.load()
    wsb = lib.fmlh_atomic_predict_ws_bytes(nm_t, na)
    ws = workspace(wsb)
    status = empty((1,), np.int32)
    blk, nb = _blocks(counts_t)
    _native.call("fmlh_atomic_predict", _native.ptr(xt), xt.stride(0), _native.ptr(off), nm_t,
                 int(xt.shape[0]), _native.ptr(x), x.stride(0), na, d, amax, _native.ptr(al),
                 float(sigma), _native.ptr(blk), nb, _native.ptr(y), _native.ptr(status), _native.ptr(ws),
                 wsb, _native.stream_ptr())
    _check_status(status)
    return y


def get_prediction(Q, Q2, N2, alpha, sigma):
    """kernels.py:101-108 / fkernels.f90:480-539 -- training atoms Q (d, na),
    test atoms Q2 (d, na2) in molecules N2, alpha (na,) -> Y (nm2,):
    Y(m) = sum over test atoms a of m and training atoms b of
    exp(-sum |Q(:, b) - Q2(:, a)| / sigma) alpha(b)."""
    c2 = _counts(N2)
    y = _predict(_rows_from_cols(Q2, c2), c2, _rows_from_cols(Q), alpha, sigma)
    return y if is_device_tensor(Q) else to_host(y)


def fget_prediction(Q, Q2, N2, alpha, sigma, Y):
    """fkernels.f90:480 -- fills Y (nm2,) in place."""
    y = get_prediction(Q, Q2, N2, alpha, sigma)
    if is_device_tensor(Y):
        Y.copy_(y if is_device_tensor(y) else to_device(y))
    else:
        Y[...] = to_host(y) if is_device_tensor(y) else y
    return Y


def _mol_kernel_from_distance(D, counts, gamma):
    na = int(np.sum(counts))
    dm = _transposed(D)
    if tuple(dm.shape) != (na, na):
        raise ValueError(f"distance matrix {tuple(dm.shape)} for {na} atoms")
    nm = int(counts.shape[0])
    K = empty((nm, nm), np.float64)
    if nm == 0:
        return K
    status = empty((1,), np.int32)
    blk, nb = _blocks(counts)
    _native.call("fmlh_distance_mol_kernel", _native.ptr(dm), dm.stride(0),
                 _native.ptr(to_device(_offsets(counts), np.int32)), nm, na, _amax(counts),
                 float(gamma), _native.ptr(blk), nb, _native.ptr(K), _native.ptr(status),
                 _native.stream_ptr())
    _check_status(status)
    return K


def get_alpha_from_distance(D, N, Y, sigma, llambda):
    """kernels.py:127-134 / fkernels.f90:541-641 -- D (na, na) atom
    distances -> alpha (na,): the molecular kernel sums exp(-D(k, a) / sigma)
    over the atoms of each molecule pair, then as get_alpha."""
    c = _counts(N)
    K = _mol_kernel_from_distance(D, c, float(sigma))
    return solve_expand(K, c, Y, llambda, None, is_device_tensor(D))


def fget_alpha_from_distance(d, n, y, dgamma, llambda, alpha):
    """fkernels.f90:541 -- fills alpha in place."""
    c = _counts(n)
    K = _mol_kernel_from_distance(d, c, float(dgamma))
    return solve_expand(K, c, y, llambda, alpha, is_device_tensor(d))


def get_prediction_from_distance(D2, N2, alpha, sigma):
    """kernels.py:110-117 / fkernels.f90:645-702 -- D2 (na, na2) distances
    between training atoms (rows) and test atoms (columns) -> Y (nm2,):
    Y(m) = sum over test atoms a of m and training atoms b of
    exp(-D2(b, a) / sigma) alpha(b).  The reference accumulates into an
    uninitialised array (fkernels.f90:680-686), so its values are undefined;
    here every sum starts at 0."""
    c2 = _counts(N2)
    dm = _transposed(D2)                 # rows: test atoms
    na_t, na = int(dm.shape[0]), int(dm.shape[1])
    if na_t != int(c2.sum()):
        raise ValueError(f"{na_t} test atoms for molecules of {int(c2.sum())} atoms")
    nm_t = int(c2.shape[0])
    y = empty((nm_t,), np.float64)
    if nm_t:
        al = to_device(alpha, np.float64)
        if int(al.shape[0]) != na:
            raise ValueError(f"{int(al.shape[0])} coefficients for {na} training atoms")
        lib = _native.load()
        wsb = lib.fmlh_atomic_predict_ws_bytes(nm_t, na)
        ws = workspace(wsb)
        status = empty((1,), np.int32)
        blk, nb = _blocks(c2)
        _native.call("fmlh_distance_predict", _native.ptr(dm), dm.stride(0),
                     _native.ptr(to_device(_offsets(c2), np.int32)), nm_t, na_t, na, _amax(c2),
                     _native.ptr(al), float(sigma), _native.ptr(blk), nb, _native.ptr(y),
                     _native.ptr(status), _native.ptr(ws), wsb, _native.stream_ptr())
        _check_status(status)
    return y if is_device_tensor(D2) else to_host(y)


def fget_prediction_from_distance(D2, N2, alpha, dgamma, Y):
    """fkernels.f90:645 -- fills Y in place."""
    y = get_prediction_from_distance(D2, N2, alpha, dgamma)
    if is_device_tensor(Y):
        Y.copy_(y if is_device_tensor(y) else to_device(y))
    else:
        Y[...] = to_host(y) if is_device_tensor(y) else y
    return Y
