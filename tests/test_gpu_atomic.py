"""GPU parity of the atomic KRR path against the oracle (oracle/fml_oracle.c,
bitwise equal to the reference Fortran build -- tests/test_oracle.py) and
the committed reference outputs (tests/golden/atomic.npz).

Tolerances (stated per check):
  * unsorted / local / atomic Coulomb matrices: bit-exact.
  * vector kernels: atom distances are the oracle's bit for bit (same
    sequential order over the descriptor, no fused multiply-add); entries
    differ by the device exp (<= 1 ulp) and the order of the per-pair sum:
    1e-13 relative, entries below 1e-290 compared as underflow.
  * the Gaussian's matrix-core form (precision "mfma"): distances from
    |a|^2 + |b|^2 - 2 a.b carry ~1e-13 (|a|^2 + |b|^2) absolute error, and an
    exponent error dx moves exp(x) by dx relative: 1e-12 + 1e-13 |ln K|.
  * alpha: kernel entries within 1e-13 move alpha by up to cond(K) times
    that; the bound used is 10 * cond(K + lambda I) * 1e-15 relative.
  * predictions: 1e-12 relative to sum |terms| of each prediction.
"""
import os

import numpy as np
import pytest
from scipy.spatial.distance import cdist

import oracle
from fml_amd import synth
from fml_amd.fml import Molecule

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def gold():
    return np.load(os.path.join(GOLD, "atomic.npz"))


def _close(got, ref, rel, floor=1e-290, per_log=0.0):
    """Relative error <= rel + per_log * |ln ref|."""
    got, ref = np.asarray(got), np.asarray(ref)
    assert got.shape == ref.shape
    big = np.abs(ref) > floor
    assert np.all(np.abs(got[~big]) <= 1e-280)
    if big.any():
        err = np.abs(got[big] - ref[big]) / np.abs(ref[big])
        tol = rel + per_log * np.abs(np.log(np.abs(ref[big])))
        assert np.all(err <= tol), np.max(err / tol)


def _q_layout(rows, counts, amax):
    """Packed atom rows -> the reference's q (d, amax, nm)."""
    nm, d = len(counts), rows.shape[1]
    q = np.zeros((nm, amax, d))
    off = np.concatenate([[0], np.cumsum(counts)])
    for i in range(nm):
        q[i, :counts[i]] = rows[off[i]:off[i + 1]]
    return np.transpose(q, (2, 1, 0)), q


# ------------------------------------------------------------ representations

def test_coulomb_forms_golden(gpu):
    from fml_amd.representations import (generate_atomic_coulomb_matrices,
                                         generate_local_coulomb_matrices,
                                         generate_unsorted_coulomb_matrices)
    d = gold()
    z, r, off = d["charges"], d["coords"], d["offsets"]
    assert np.array_equal(generate_unsorted_coulomb_matrices(z, r, off, 23), d["cm_unsorted"])
    assert np.array_equal(generate_local_coulomb_matrices(z, r, off, 23), d["cm_local"])
    assert np.array_equal(generate_atomic_coulomb_matrices(z, r, off, 23), d["cm_atomic"])


@pytest.mark.parametrize("nmax", [29, 40])
def test_coulomb_forms_vs_oracle(gpu, nmax):
    from fml_amd.representations import (generate_atomic_coulomb_matrices,
                                         generate_local_coulomb_matrices,
                                         generate_unsorted_coulomb_matrices)
    ms = synth.generate(120, seed=31, nmin=1, nmax=29)
    args = (ms.charges, ms.coords, ms.offsets, nmax)
    assert np.array_equal(generate_unsorted_coulomb_matrices(*args), oracle.coulomb_forms(0, *args))
    assert np.array_equal(generate_local_coulomb_matrices(*args), oracle.coulomb_forms(1, *args))
    assert np.array_equal(generate_atomic_coulomb_matrices(*args), oracle.coulomb_forms(2, *args))


def test_coulomb_forms_single_and_errors(gpu):
    from fml_amd.representations import (fgenerate_atomic_coulomb_matrix,
                                         fgenerate_local_coulomb_matrix,
                                         fgenerate_unsorted_coulomb_matrix,
                                         generate_local_coulomb_matrices)
    m = Molecule()
    m.read_xyz(os.path.join(GOLD, "xyz", "03.xyz"))
    Z, R = m.nuclear_charges, m.coordinates
    assert np.array_equal(fgenerate_unsorted_coulomb_matrix(Z, R, m.natoms, 23),
                          oracle.coulomb_form(0, Z, R, 23))
    assert np.array_equal(fgenerate_local_coulomb_matrix(Z, R, m.natoms, 23),
                          oracle.coulomb_form(1, Z, R, 23))
    assert np.array_equal(fgenerate_atomic_coulomb_matrix(Z, R, m.natoms, 23),
                          oracle.coulomb_form(2, Z, R, 23))
    m.generate_local_coulomb_matrix(size=23)
    m.generate_atomic_coulomb_matrix(size=23)
    m.generate_unsorted_coulomb_matrix(size=23)
    assert m.local_coulomb_matrix.shape == (m.natoms, 276)
    assert np.array_equal(m.atomic_coulomb_matrix, oracle.coulomb_form(2, Z, R, 23))
    with pytest.raises(ValueError):
        fgenerate_local_coulomb_matrix(Z, R, m.natoms, 2)
    with pytest.raises(ValueError):
        fgenerate_local_coulomb_matrix(Z, R[:-1], m.natoms, 23)
    # an oversize molecule fails the batch call, naming it
    ms = synth.generate(5, seed=32, nmin=3, nmax=12)
    with pytest.raises(ValueError, match="molecule"):
        generate_local_coulomb_matrices(ms.charges, ms.coords, ms.offsets, 8)


# ------------------------------------------------------------ vector kernels

def test_vector_kernels_golden(gpu):
    from fml_amd.kernels import fget_vector_kernels_gaussian, fget_vector_kernels_laplacian
    d = gold()
    q1, q2 = np.transpose(d["q1"], (2, 1, 0)), np.transpose(d["q2"], (2, 1, 0))
    kg = fget_vector_kernels_gaussian(q1, q2, d["n1"], d["n2"], d["sigmas_gauss"])
    kl = fget_vector_kernels_laplacian(q1, q2, d["n1"], d["n2"], d["sigmas_lap"])
    _close(kg, d["vec_gauss"], 1e-13)
    _close(kl, d["vec_lap"], 1e-13)
    km = fget_vector_kernels_gaussian(q1, q2, d["n1"], d["n2"], d["sigmas_gauss"], precision="mfma")
    _close(km, d["vec_gauss"], 1e-12, per_log=1e-13)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_vector_kernels_multiblock_vs_oracle(gpu, kind):
    """Hundreds of molecules (tens of 128-atom blocks per side), molecules
    of 1-29 atoms, local Coulomb matrices as descriptors."""
    from fml_amd.kernels import fget_vector_kernels_gaussian, fget_vector_kernels_laplacian
    from fml_amd.representations import generate_local_coulomb_matrices
    ms1 = synth.generate(160, seed=33, nmin=1, nmax=29)
    ms2 = synth.generate(130, seed=34, nmin=1, nmax=29)
    r1 = generate_local_coulomb_matrices(ms1.charges, ms1.coords, ms1.offsets, 29)
    r2 = generate_local_coulomb_matrices(ms2.charges, ms2.coords, ms2.offsets, 29)
    q1, q1o = _q_layout(r1, ms1.natoms, 29)
    q2, q2o = _q_layout(r2, ms2.natoms, 29)
    sig = [300.0, 3000.0] if kind == 1 else [30.0, 300.0]
    if kind == 1:
        got = fget_vector_kernels_laplacian(q1, q2, ms1.natoms, ms2.natoms, sig)
    else:
        got = fget_vector_kernels_gaussian(q1, q2, ms1.natoms, ms2.natoms, sig,
                                           precision="mfma" if kind == 2 else "fp64")
    ref = oracle.vector_kernels(1 if kind == 1 else 0, q1o, q2o, ms1.natoms, ms2.natoms, sig)
    if kind == 2:
        _close(got, ref, 1e-12, per_log=1e-13)
    else:
        _close(got, ref, 1e-13)


def test_vector_kernels_edges(gpu):
    """Empty molecules, 64-atom molecules, more than 8 sigmas (two
    launches), one-atom molecules equal to the dense kernels."""
    from fml_amd.kernels import fget_vector_kernels_laplacian, laplacian_kernel
    rng = np.random.default_rng(35)
    d = 11
    n1 = np.array([0, 64, 3, 0, 17, 64, 1, 0])
    n2 = np.array([5, 0, 64, 2])
    x1 = rng.standard_normal((n1.sum(), d))
    x2 = rng.standard_normal((n2.sum(), d))
    q1, q1o = _q_layout(x1, n1, 64)
    q2, q2o = _q_layout(x2, n2, 64)
    sig = list(np.linspace(1.0, 12.0, 10))
    got = fget_vector_kernels_laplacian(q1, q2, n1, n2, sig)
    ref = oracle.vector_kernels(1, q1o, q2o, n1, n2, sig)
    _close(got, ref, 1e-13)
    assert np.all(got[:, n1 == 0, :] == 0) and np.all(got[:, :, n2 == 0] == 0)
    ones = np.ones(6, dtype=np.int64)
    a = rng.standard_normal((6, d))
    b = rng.standard_normal((6, d))
    qa, _ = _q_layout(a, ones, 1)
    qb, _ = _q_layout(b, ones, 1)
    k1 = fget_vector_kernels_laplacian(qa, qb, ones, ones, [2.5])[0]
    _close(k1, laplacian_kernel(a.T, b.T, 2.5), 1e-15)


def test_vector_kernels_batch_invariance(gpu):
    """A molecule pair's value is bitwise the same in any batch."""
    from fml_amd.kernels import fget_vector_kernels_gaussian
    ms = synth.generate(90, seed=36, nmin=1, nmax=23)
    from fml_amd.representations import generate_local_coulomb_matrices
    rows = generate_local_coulomb_matrices(ms.charges, ms.coords, ms.offsets, 23)
    q, _ = _q_layout(rows, ms.natoms, 23)
    full = fget_vector_kernels_gaussian(q, q, ms.natoms, ms.natoms, [50.0])
    sub = fget_vector_kernels_gaussian(q[:, :, 7:41], q[:, :, 60:83], ms.natoms[7:41], ms.natoms[60:83],
                                       [50.0])
    assert np.array_equal(sub, full[:, 7:41, 60:83])


def test_block_lists_do_not_change_results(gpu):
    """Greedy blocks, one molecule per block and the default fixed-window
    blocks give bitwise the same kernel; a block list wider than 128 atoms
    is reported through status, not computed."""
    import ctypes
    from fml_amd import _native
    from fml_amd.device import empty, to_device
    from fml_amd.kernels.atomic import plan_blocks
    from fml_amd.representations import generate_local_coulomb_matrices
    ms1 = synth.generate(70, seed=40, nmin=1, nmax=23)
    ms2 = synth.generate(50, seed=41, nmin=1, nmax=23)
    x1 = generate_local_coulomb_matrices(ms1.charges, ms1.coords, ms1.offsets, 23, on_device=True)
    x2 = generate_local_coulomb_matrices(ms2.charges, ms2.coords, ms2.offsets, 23, on_device=True)
    o1, o2 = to_device(ms1.offsets, np.int32), to_device(ms2.offsets, np.int32)
    sig = (ctypes.c_double * 2)(20.0, 200.0)
    status = empty((1,), np.int32)

    def run(kind, b1, b2):
        out = empty((ms2.nmol, ms1.nmol, 2), np.float64)
        d1 = to_device(b1, np.int32) if b1 is not None else None
        d2 = to_device(b2, np.int32) if b2 is not None else None
        _native.call("fmlh_vector_kernels", kind, _native.ptr(x1), x1.stride(0), _native.ptr(o1), ms1.nmol,
                     int(x1.shape[0]), _native.ptr(x2), x2.stride(0), _native.ptr(o2), ms2.nmol,
                     int(x2.shape[0]), int(x1.shape[1]), 23, sig, 2, _native.ptr(d1),
                     0 if b1 is None else len(b1) - 1, _native.ptr(d2), 0 if b2 is None else len(b2) - 1,
                     _native.ptr(out), _native.ptr(status), _native.stream_ptr())
        return out.cpu().numpy(), int(status.item())

    for kind in (0, 1, 2):
        ref, st = run(kind, None, None)
        assert st == 0
        got, st = run(kind, plan_blocks(ms1.natoms), plan_blocks(ms2.natoms))
        assert st == 0 and np.array_equal(got, ref)
        one1 = np.arange(ms1.nmol + 1, dtype=np.int32)
        got, st = run(kind, one1, plan_blocks(ms2.natoms))
        assert st == 0 and np.array_equal(got, ref)
    _, st = run(1, np.array([0, ms1.nmol], dtype=np.int32), None)   # one block, > 128 atoms
    assert st == 1


def test_get_atomic_kernels_gaussian_molecules(gpu):
    """kernels.py:380-406 from Molecule objects."""
    from fml_amd.kernels import get_atomic_kernels_gaussian
    mols = []
    for f in sorted(os.listdir(os.path.join(GOLD, "xyz"))):
        m = Molecule()
        m.read_xyz(os.path.join(GOLD, "xyz", f))
        m.generate_local_coulomb_matrix(size=23)
        mols.append(m)
    K = get_atomic_kernels_gaussian(mols[:6], mols[4:], [100.0, 1000.0])
    n1 = np.array([m.natoms for m in mols[:6]])
    n2 = np.array([m.natoms for m in mols[4:]])
    _, q1 = _q_layout(np.concatenate([m.local_coulomb_matrix for m in mols[:6]]), n1, 23)
    _, q2 = _q_layout(np.concatenate([m.local_coulomb_matrix for m in mols[4:]]), n2, 23)
    _close(K, oracle.vector_kernels(0, q1, q2, n1, n2, [100.0, 1000.0]), 1e-13)


# ------------------------------------------------------------ alpha / prediction

def _alpha_tol(K, lam):
    return 10.0 * np.linalg.cond(K + lam * np.eye(K.shape[0])) * 1e-15


def test_get_alpha_prediction_golden(gpu):
    from fml_amd.kernels import get_alpha, get_alpha_from_distance, get_prediction
    d = gold()
    x1, x2, n1, n2, y = d["x1"], d["x2"], d["n1"], d["n2"], d["y1"]
    E = np.exp(-cdist(x1, x1, "cityblock") / 200.0)
    idx = np.concatenate([[0], np.cumsum(n1)[:-1]])
    K = np.add.reduceat(np.add.reduceat(E, idx, axis=0), idx, axis=1)
    tol = _alpha_tol(K, 1e-4)
    alpha = get_alpha(x1.T, n1, y, 200.0, 1e-4)
    assert np.max(np.abs(alpha - d["alpha"])) <= tol * np.max(np.abs(d["alpha"]))
    alpha_d = get_alpha_from_distance(d["dmat"], n1, y, 200.0, 1e-4)
    assert np.max(np.abs(alpha_d - d["alpha_d"])) <= tol * np.max(np.abs(d["alpha_d"]))
    pred = get_prediction(x1.T, x2.T, n2, d["alpha"], 200.0)
    E2 = np.exp(-cdist(x2, x1, "cityblock") / 200.0)
    scale = np.add.reduceat(np.abs(E2 * d["alpha"][None, :]).sum(1), np.concatenate([[0], np.cumsum(n2)[:-1]]))
    assert np.all(np.abs(pred - d["pred"]) <= 1e-12 * scale)


def test_alpha_prediction_vs_oracle(gpu):
    """A larger synthetic case through every entry, numpy in, numpy out."""
    from fml_amd.kernels import (fget_alpha, fget_prediction, get_alpha_from_distance,
                                 get_prediction, get_prediction_from_distance)
    from fml_amd.representations import generate_local_coulomb_matrices
    ms1 = synth.generate(70, seed=37, nmin=2, nmax=23)
    ms2 = synth.generate(30, seed=38, nmin=2, nmax=23)
    x1 = generate_local_coulomb_matrices(ms1.charges, ms1.coords, ms1.offsets, 23)
    x2 = generate_local_coulomb_matrices(ms2.charges, ms2.coords, ms2.offsets, 23)
    D = cdist(x1, x1, "cityblock")
    D2 = cdist(x1, x2, "cityblock")
    rng = np.random.default_rng(39)
    y = rng.standard_normal(ms1.nmol)
    sigma, lam = 3000.0, 1e-3
    ref, info = oracle.get_alpha(x1.T, ms1.natoms, y, sigma, lam)
    assert info == 0
    alpha = np.zeros(x1.shape[0])
    fget_alpha(x1.T, ms1.natoms, y, sigma, lam, alpha)
    E = np.exp(-D / sigma)
    idx = np.concatenate([[0], np.cumsum(ms1.natoms)[:-1]])
    K = np.add.reduceat(np.add.reduceat(E, idx, axis=0), idx, axis=1)
    assert np.max(np.abs(alpha - ref)) <= _alpha_tol(K, lam) * np.max(np.abs(ref))
    pref = oracle.get_prediction(x1.T, x2.T, ms2.natoms, ref, sigma)
    Y = np.zeros(ms2.nmol)
    fget_prediction(x1.T, x2.T, ms2.natoms, ref, sigma, Y)
    E2 = np.exp(-D2.T / sigma)
    idx2 = np.concatenate([[0], np.cumsum(ms2.natoms)[:-1]])
    scale = np.add.reduceat(np.abs(E2 * ref[None, :]).sum(1), idx2)
    assert np.all(np.abs(Y - pref) <= 1e-12 * scale)
    assert np.array_equal(get_prediction(x1.T, x2.T, ms2.natoms, ref, sigma), Y)
    # distance forms on the L1 distance matrices: the same model
    aref, _ = oracle.get_alpha_from_distance(D, ms1.natoms, y, sigma, lam)
    ad = get_alpha_from_distance(D, ms1.natoms, y, sigma, lam)
    assert np.max(np.abs(ad - aref)) <= _alpha_tol(K, lam) * np.max(np.abs(aref))
    pd = get_prediction_from_distance(D2, ms2.natoms, ref, sigma)
    assert np.all(np.abs(pd - oracle.get_prediction_from_distance(D2, ms2.natoms, ref, sigma))
                  <= 1e-12 * scale)


def test_atomic_path_device_tensors(gpu):
    """CUDA tensors in: results stay on the GPU and equal the numpy path."""
    from fml_amd.kernels import fget_vector_kernels_laplacian, get_alpha, get_prediction
    torch = gpu
    d = gold()
    q1 = np.transpose(d["q1"], (2, 1, 0))
    q2 = np.transpose(d["q2"], (2, 1, 0))
    kh = fget_vector_kernels_laplacian(q1, q2, d["n1"], d["n2"], d["sigmas_lap"])
    kd = fget_vector_kernels_laplacian(torch.from_numpy(np.ascontiguousarray(q1)).cuda(),
                                       torch.from_numpy(np.ascontiguousarray(q2)).cuda(),
                                       d["n1"], d["n2"], d["sigmas_lap"])
    assert kd.is_cuda and np.array_equal(kd.cpu().numpy(), kh)
    X1 = torch.from_numpy(np.ascontiguousarray(d["x1"].T)).cuda()
    X2 = torch.from_numpy(np.ascontiguousarray(d["x2"].T)).cuda()
    a = get_alpha(X1, d["n1"], d["y1"], 200.0, 1e-4)
    assert a.is_cuda and np.array_equal(a.cpu().numpy(), get_alpha(d["x1"].T, d["n1"], d["y1"], 200.0, 1e-4))
    p = get_prediction(X1, X2, d["n2"], a, 200.0)
    assert p.is_cuda and np.array_equal(p.cpu().numpy(),
                                        get_prediction(d["x1"].T, d["x2"].T, d["n2"], a.cpu().numpy(), 200.0))
