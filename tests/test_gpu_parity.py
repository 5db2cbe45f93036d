"""GPU parity of every hot-path routine against the oracle (the C
restatement, itself bitwise equal to the flang build of the reference
Fortran -- see tests/test_oracle.py).  Sizes are chosen so the oracle
finishes in seconds.

Tolerances (stated per check):
  * Coulomb matrices: bit-exact values and permutations.
  * Laplacian: the L1 sums are bit-identical, entries differ only by the
    device exp's rounding: 4e-16 relative.
  * Gaussian FP64: MFMA dot products + norm form vs sequential sums:
    1e-12 relative.
  * ARAD: every overlap term is formed in the reference's operation order
    and its exp and sin round as the host libm's, so D = self1 + self2 -
    2 S T -- whose absolute error 1/sigma^2 amplifies in the kernel -- is
    the reference's; 1e-12 relative at every sigma (0.1 included) and for
    the distance forms (absolute, scaled by the self terms).
"""
import numpy as np
import pytest

import oracle
from fml_amd import synth
from fml_amd.arad import ARAD
from fml_amd.kernels.kernels import arad_types

pytestmark = pytest.mark.gpu


def _rel(a, b, floor=1e-300):
    a = np.asarray(a)
    b = np.asarray(b)
    m = np.abs(b) > floor
    assert np.array_equal(m, np.abs(a) > floor) or np.max(np.abs(a[~m])) < 1e-290
    return float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m]))) if m.any() else 0.0


@pytest.fixture(scope="module")
def mols():
    return synth.generate(160, seed=3, nmin=3, nmax=29)


@pytest.fixture(scope="module")
def arad_set():
    ms = synth.generate(36, seed=5, nmin=3, nmax=23)
    ar = ARAD(maxMolSize=23, maxAts=23)
    X = np.array([ar.describe(ms.molecule(m)[1], ms.molecule(m)[0]) for m in range(ms.nmol)])
    Z = [ms.molecule(m)[0] for m in range(ms.nmol)]
    z, n = arad_types(Z, 23)
    return X, Z, z, n


def test_coulomb_bitwise(gpu, mols):
    from fml_amd.representations import generate_coulomb_matrices
    cm, perm = generate_coulomb_matrices(mols.charges, mols.coords, mols.offsets, nmax=29,
                                         return_perm=True)
    ocm, operm = oracle.coulomb_matrices(mols.charges, mols.coords, mols.offsets, 29)
    assert np.array_equal(cm, ocm)
    for m in range(mols.nmol):
        n = mols.natoms[m]
        assert np.array_equal(perm[m, :n], operm[m, :n])
        assert np.all(perm[m, n:] == -1)


def test_coulomb_ties_and_padding(gpu):
    """Symmetric molecules (exact row-norm ties) and nmax padding."""
    from fml_amd.representations import fgenerate_coulomb_matrix
    # square of 4 H atoms + a centred C: all H norms tie exactly
    z = np.array([1.0, 1.0, 6.0, 1.0, 1.0])
    r = np.array([[1, 0, 0], [-1, 0, 0], [0, 0, 0], [0, 1, 0], [0, -1, 0]], dtype=float)
    for nmax in (5, 7, 12):
        a = fgenerate_coulomb_matrix(z, r, 5, nmax)
        b, _ = oracle.coulomb_matrix(z, r, nmax)
        assert np.array_equal(a, b)
        assert np.all(a[15:] == 0.0)


def test_coulomb_rejects_oversize(gpu):
    from fml_amd.representations import generate_coulomb_matrices
    z = np.ones(6)
    r = np.arange(18, dtype=float).reshape(6, 3)
    with pytest.raises(ValueError):
        generate_coulomb_matrices(z, r, np.array([0, 6], dtype=np.int32), nmax=5)


def _cm(mols, nmax=29):
    return oracle.coulomb_matrices(mols.charges, mols.coords, mols.offsets, nmax)[0]


def test_gaussian_fp64(gpu, mols):
    from fml_amd.kernels import gaussian_kernel
    X = _cm(mols)
    A, B = X[:100].T, X[100:].T
    for sigma in (1000.0, 50.0):
        k = gaussian_kernel(A, B, sigma)
        ko = oracle.gaussian_kernel(A, B, sigma)
        assert k.shape == (100, 60) and k.flags.f_contiguous
        assert _rel(k, ko) < 1e-12
    # symmetric case: exact ones on the diagonal (near pairs recomputed)
    kk = gaussian_kernel(A, A, 700.0)
    assert np.all(np.diag(kk) == 1.0)
    assert _rel(kk, oracle.gaussian_kernel(A, A, 700.0)) < 1e-12


def test_gaussian_ragged_shapes(gpu):
    from fml_amd.kernels import gaussian_kernel, laplacian_kernel
    rng = np.random.default_rng(0)
    for (d, na, nb) in ((1, 1, 1), (5, 3, 70), (17, 65, 129), (33, 64, 64)):
        A = rng.normal(size=(d, na))
        B = rng.normal(size=(d, nb))
        assert _rel(gaussian_kernel(A, B, 2.0), oracle.gaussian_kernel(A, B, 2.0)) < 1e-12
        assert _rel(laplacian_kernel(A, B, 2.0), oracle.laplacian_kernel(A, B, 2.0)) < 4e-16


def test_kernels_of_a_set_with_itself(gpu):
    """A kernel of a device set with itself runs only the lower-triangle
    tiles and mirrors their stores; it is bitwise the kernel the full grid
    writes for the same rows held in a second buffer, for the FP64 and FP32
    Gaussian and Laplacian, at ragged sizes (600 and 1300 rows: 128 and
    64 tiles with partial edges; Coulomb matrices, so near pairs occur)."""
    torch = gpu
    from fml_amd.kernels.kernels import gaussian_kernel_rows, laplacian_kernel_rows
    from fml_amd import synth
    from fml_amd.representations import generate_coulomb_matrices
    for nmol in (600, 1300):
        ms = synth.generate(nmol, seed=nmol, nmin=3, nmax=29)
        X = torch.from_numpy(generate_coulomb_matrices(ms.charges, ms.coords, ms.offsets, nmax=29)).cuda()
        Xc = X.clone()
        for dt in (torch.float64, torch.float32):
            Xs, Xo = X.to(dt), Xc.to(dt)
            prec = "fp64" if dt == torch.float64 else "fp32"
            ks, kf = gaussian_kernel_rows(Xs, Xs, 1000.0, precision=prec), gaussian_kernel_rows(Xs, Xo, 1000.0, precision=prec)
            assert torch.equal(ks, kf), (nmol, dt, "gaussian")
            assert torch.equal(ks, ks.T)
            ls, lf = laplacian_kernel_rows(Xs, Xs, 10000.0), laplacian_kernel_rows(Xs, Xo, 10000.0)
            assert torch.equal(ls, lf), (nmol, dt, "laplacian")


def test_laplacian_fp64(gpu, mols):
    from fml_amd.kernels import laplacian_kernel
    X = _cm(mols)
    A, B = X[:100].T, X[100:].T
    k = laplacian_kernel(A, B, 10000.0)
    assert _rel(k, oracle.laplacian_kernel(A, B, 10000.0)) < 4e-16


def test_arad_global(gpu, arad_set):
    from fml_amd.kernels import get_atomic_kernels_arad
    X, Z, z, n = arad_set
    sig = [0.1, 1.0, 10.0]
    k = get_atomic_kernels_arad(X[:20], X[20:], Z[:20], Z[20:], sig, cut_distance=5.0)
    ko = oracle.kernels_arad(X[:20], X[20:], z[:20], z[20:], n[:20], n[20:], sig, 0.2, 5.0, 1.0, 0.5)
    assert k.shape == (3, 20, 16)
    assert _rel(k[0], ko[0]) < 1e-12
    assert _rel(k[1], ko[1]) < 1e-12
    assert _rel(k[2], ko[2]) < 1e-12


def test_arad_tiles_uneven_sizes(gpu):
    """Size-sorted tiles: molecules of 1-29 atoms (truncated geometric), amax
    29, a lone 1-atom molecule, general, symmetric and single-sigma forms."""
    from fml_amd.kernels import (get_atomic_kernels_arad, get_atomic_symmetric_kernels_arad,
                                 fget_kernel_arad)
    ms = synth.generate(44, seed=21, nmin=1, nmax=29, dist="geometric")
    ar = ARAD(maxMolSize=29, maxAts=29)
    X = np.array([ar.describe(ms.molecule(m)[1], ms.molecule(m)[0]) for m in range(ms.nmol)])
    Z = [ms.molecule(m)[0] for m in range(ms.nmol)]
    z, n = arad_types(Z, 29)
    assert n.min() <= 2 and n.max() >= 20
    sig = [1.0, 4.0]
    k = get_atomic_kernels_arad(X[:30], X[30:], Z[:30], Z[30:], sig, cut_distance=5.0)
    ko = oracle.kernels_arad(X[:30], X[30:], z[:30], z[30:], n[:30], n[30:], sig, 0.2, 5.0)
    assert _rel(k, ko) < 1e-12
    ks = get_atomic_symmetric_kernels_arad(X, Z, sig)
    assert _rel(ks, oracle.symmetric_kernels_arad(X, z, n, sig)) < 1e-12
    k1 = fget_kernel_arad(X, z, n, 1.5)
    assert _rel(k1, oracle.kernel_arad(X, z, n, 1.5)) < 1e-12


def test_arad_generic_type_path(gpu):
    """More than 16 distinct element types among the neighbours: the kernels
    take the raw-channel path (T evaluated per term, not from the type
    table)."""
    from fml_amd.kernels import get_atomic_kernels_arad, get_atomic_symmetric_kernels_arad
    ms = synth.generate(24, seed=8, nmin=3, nmax=23)
    rng = np.random.default_rng(8)
    Z = [rng.integers(1, 41, size=len(ms.molecule(m)[0])) for m in range(ms.nmol)]
    assert len(np.unique(np.concatenate(Z))) > 16
    ar = ARAD(maxMolSize=23, maxAts=23)
    X = np.array([ar.describe(ms.molecule(m)[1], Z[m]) for m in range(ms.nmol)])
    z, n = arad_types(Z, 23)
    sig = [1.0, 4.0]
    k = get_atomic_kernels_arad(X[:14], X[14:], Z[:14], Z[14:], sig, cut_distance=5.0)
    ko = oracle.kernels_arad(X[:14], X[14:], z[:14], z[14:], n[:14], n[14:], sig, 0.2, 5.0)
    assert _rel(k, ko) < 1e-12
    ks = get_atomic_symmetric_kernels_arad(X, Z, sig)
    assert _rel(ks, oracle.symmetric_kernels_arad(X, z, n, sig)) < 1e-12


def test_arad_symmetric_and_kernel(gpu, arad_set):
    from fml_amd.kernels import get_atomic_symmetric_kernels_arad, fget_kernel_arad
    X, Z, z, n = arad_set
    sig = [1.0, 4.0]
    k = get_atomic_symmetric_kernels_arad(X, Z, sig)
    ko = oracle.symmetric_kernels_arad(X, z, n, sig, 0.2, 6.0, 1.0, 0.5)
    assert _rel(k, ko) < 1e-12
    assert np.array_equal(k, np.transpose(k, (0, 2, 1)))
    k1 = fget_kernel_arad(X, z, n, 2.0)
    assert _rel(k1, oracle.kernel_arad(X, z, n, 2.0)) < 1e-12


def test_arad_distance_forms(gpu, arad_set):
    from fml_amd.math import get_l2_distance_arad
    from fml_amd.kernels import fget_atomic_kernels_arad, fget_atomic_distance_arad
    X, Z, z, n = arad_set
    d = get_l2_distance_arad(X[:7], X[7:12], Z[:7], Z[7:12])
    do = oracle.arad_l2_distance_all(X[:7], X[7:12], z[:7], z[7:12], n[:7], n[7:12])
    assert d.shape == do.shape
    scale = np.max(np.abs(do))
    assert np.max(np.abs(d - do)) <= 1e-13 * scale
    # padding is exactly zero
    assert np.all(d[np.abs(do) == 0] == 0)
    # per-atom forms
    qa = np.concatenate([X[m, :n[m]] for m in range(5)])
    za = np.concatenate([z[m, :n[m]] for m in range(5)])
    na = np.concatenate([[n[m]] * n[m] for m in range(5)]).astype(np.int32)
    qb, zb, nb = qa[3:40], za[3:40], na[3:40]
    ka = fget_atomic_kernels_arad(qa, qb, za, zb, na, nb, [1.0, 4.0])
    assert _rel(ka, oracle.atomic_kernels_arad(qa, qb, za, zb, na, nb, [1.0, 4.0])) < 1e-12
    da = fget_atomic_distance_arad(qa, qb, za, zb, na, nb)
    dao = oracle.atomic_distance_arad(qa, qb, za, zb, na, nb)
    assert np.max(np.abs(da - dao)) <= 1e-13 * np.max(np.abs(dao))
    # self distances vanish exactly
    dd = fget_atomic_distance_arad(qa, qa, za, za, na, na)
    assert np.all(np.diag(dd) == 0.0)


def test_cho_solve(gpu, mols):
    from fml_amd.math import cho_solve
    X = _cm(mols)[:150]
    K = oracle.gaussian_kernel(X.T, X.T, 700.0)
    K[np.diag_indices_from(K)] += 1e-4
    y = mols.energies[:150]
    Kc = K.copy()
    x = cho_solve(K, y)
    assert np.array_equal(K, Kc)  # caller's matrix untouched
    xo, info = oracle.cho_solve(K, y)
    assert info == 0
    # backward error of a stable Cholesky solve
    r = K @ x - y
    be = np.linalg.norm(r) / (np.linalg.norm(K, 2) * np.linalg.norm(x))
    assert be < 1e-14
    cond = np.linalg.cond(K)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 50 * cond * 2.2e-16


def test_cho_solve_sizes_and_failure(gpu):
    from fml_amd.math import cho_solve
    rng = np.random.default_rng(1)
    for n in (1, 2, 63, 64, 65, 200, 257, 300):
        # integer, well conditioned: strictly lower entries in {-1,0,1},
        # diagonal n so every product below is exact
        L = np.tril(rng.integers(-1, 2, size=(n, n)), -1).astype(float) + n * np.eye(n)
        A = L @ L.T
        xt = rng.integers(-3, 4, size=n).astype(float)
        y = A @ xt
        x = cho_solve(A, y)
        assert np.max(np.abs(x - xt)) < 1e-8 * max(1.0, np.max(np.abs(xt))) * n
    A = np.eye(5)
    A[3, 3] = -1.0
    x = cho_solve(A, np.ones(5))
    assert np.all(np.isnan(x))
    with pytest.raises(ValueError):
        cho_solve(np.ones((3, 4)), np.ones(3))


def test_spotrs_against_fp64_solve_of_same_factor(gpu):
    """FP32 solve (one right-hand side: the chained kernel, whose full 64-row
    blocks multiply by the diagonal block's inverse and whose ragged last
    block substitutes; three: the blocked solve) against an FP64 solve with
    the same FP32 factor U (A = U^T U).  Bound: 2e-5 relative in the max
    norm on matrices of condition number below 10; a wrong inverse entry or
    a skipped block gives errors of order one."""
    import scipy.linalg as sl
    import torch
    from fml_amd.math.cho_solve import potrf_inplace, potrs_inplace
    rng = np.random.default_rng(7)
    for n in (1, 63, 64, 65, 128, 200, 257, 1000, 4100):
        B = rng.standard_normal((n, n))
        A = (B @ B.T / n + 2.0 * np.eye(n)).astype(np.float32)
        m = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
        assert potrf_inplace(m) == 0
        # Fortran memory: U(i, j), i <= j, sits at m[j, i]
        U = np.triu(m.cpu().numpy().T).astype(np.float64)
        for nrhs in (1, 3):
            y = rng.standard_normal((nrhs, n)).astype(np.float32)
            x = potrs_inplace(m, torch.from_numpy(y.copy()).cuda().squeeze(0) if nrhs == 1
                              else torch.from_numpy(y.copy()).cuda())
            x = x.cpu().numpy().reshape(nrhs, n).astype(np.float64)
            for j in range(nrhs):
                z = sl.solve_triangular(U, y[j].astype(np.float64), trans="T", lower=False)
                xr = sl.solve_triangular(U, z, lower=False)
                err = np.max(np.abs(x[j] - xr)) / np.max(np.abs(xr))
                assert err < 2e-5, (n, nrhs, j, err)


# ------------------------------------------------------------ precision forms

def test_gaussian_precision_forms(gpu, mols):
    """FP32-MFMA and split-BF16x3 forms on float descriptors.  Stated error
    bounds against the FP64 oracle (relative, entries > 1e-300):
    FP32 5e-5, BF16x3 5e-4 at sigma = 1000 on nmax-29 Coulomb matrices."""
    import torch
    from fml_amd.kernels import gaussian_kernel_rows
    X = _cm(mols)
    X32 = torch.from_numpy(X.astype(np.float32)).cuda()
    ko = oracle.gaussian_kernel(X[:100].T, X[100:].T, 1000.0)
    for prec, tol in (("fp32", 5e-5), ("bf16x3", 5e-4)):
        for od in (np.float32, np.float64):
            k = gaussian_kernel_rows(X32[:100], X32[100:], 1000.0, precision=prec, out_dtype=od)
            kh = k.cpu().numpy().T
            assert _rel(kh, ko) < tol, (prec, od)
        kk = gaussian_kernel_rows(X32, X32, 1000.0, precision=prec).cpu().numpy()
        assert np.all(np.diag(kk) == 1.0)
        # a set with itself: lower-triangle tiles, mirrored -- exactly symmetric
        assert np.array_equal(kk, kk.T), prec
        assert _rel(kk[:100, 100:].astype(np.float64), ko) < tol, prec


def test_laplacian_fp32(gpu, mols):
    import torch
    from fml_amd.kernels import laplacian_kernel_rows
    X = _cm(mols)
    X32 = torch.from_numpy(X.astype(np.float32)).cuda()
    ko = oracle.laplacian_kernel(X[:100].T, X[100:].T, 1e4)
    k = laplacian_kernel_rows(X32[:100], X32[100:], 1e4).cpu().numpy().T
    assert _rel(k, ko) < 1e-5


def test_arad_fp32_variant(gpu, arad_set):
    from fml_amd.kernels import get_atomic_kernels_arad
    X, Z, z, n = arad_set
    k = get_atomic_kernels_arad(X[:20], X[20:], Z[:20], Z[20:], [1.0, 4.0], precision=1)
    ko = oracle.kernels_arad(X[:20], X[20:], z[:20], z[20:], n[:20], n[20:], [1.0, 4.0], 0.2, 6.0)
    assert _rel(k[1], ko[1]) < 1e-4
    assert _rel(k[0], ko[0]) < 1e-3


def test_krr_fp64_and_mixed(gpu):
    """KRR end to end on the device: the plain FP64 fit reproduces the
    LAPACK solution's predictions; the mixed form (FP32 factor as CG
    preconditioner, FP64 residuals) reaches a small FP64 backward error."""
    from fml_amd.krr import KRR
    ms = synth.generate(400, seed=0, nmin=5, nmax=23)
    X = oracle.coulomb_matrices(ms.charges, ms.coords, ms.offsets, 23)[0]
    y = ms.energies
    Xtr, Xte, ytr = X[:320], X[320:], y[:320]
    K = oracle.gaussian_kernel(Xtr.T, Xtr.T, 700.0)
    K[np.diag_indices_from(K)] += 1e-8
    ao, info = oracle.cho_solve(K, ytr)
    po = oracle.gaussian_kernel(Xtr.T, Xte.T, 700.0).T @ ao
    m = KRR(700.0, 1e-8).fit(Xtr, ytr)
    p = m.predict(Xte).cpu().numpy()
    scale = np.max(np.abs(po))
    assert np.max(np.abs(p - po)) < 1e-6 * scale
    mae_o = np.mean(np.abs(po - y[320:]))
    assert abs(np.mean(np.abs(p - y[320:])) - mae_o) <= 1e-3 * mae_o
    # kappa(K + 1e-8 I) ~ 1e11 here: the FP32-preconditioned iteration reaches
    # a relative residual of 1e-7 (1.3e-8 after 300 iterations, measured)
    mm = KRR(700.0, 1e-8, precision="mixed", tol=1e-7, max_iter=300).fit(Xtr, ytr)
    pm = mm.predict(Xte).cpu().numpy()
    assert mm.report.backward_error < 1e-12
    assert mm.report.residual <= 1e-7
    assert abs(np.mean(np.abs(pm - y[320:])) - mae_o) <= 1e-2 * mae_o


def test_krr_mixed_failure_is_visible(gpu):
    """A mixed fit that cannot reach its tolerance within max_iter raises,
    carrying the last iterate and the report (it used to return alpha
    silently)."""
    from fml_amd.krr import KRR, ConvergenceError
    ms = synth.generate(400, seed=0, nmin=5, nmax=23)
    X = oracle.coulomb_matrices(ms.charges, ms.coords, ms.offsets, 23)[0]
    with pytest.raises(ConvergenceError) as ei:
        KRR(700.0, 1e-8, precision="mixed", tol=1e-14, max_iter=2, check_every=1).fit(X[:320], ms.energies[:320])
    assert ei.value.report.iterations == 2
    assert ei.value.report.residual > 1e-14
    assert ei.value.alpha is not None and ei.value.alpha.shape[0] == 320
