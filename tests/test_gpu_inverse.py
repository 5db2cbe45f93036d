"""GPU inverse from the Cholesky factor (fmlh_potri: cho_invert,
fml/math/fcho_solve.f90:49-68) and the all-in-one ARAD regression
coefficients (fget_alpha_arad, fml/kernels/farad_kernels.f90:231-436),
against the oracle (scipy's dpotrf/dpotri/dsymv, bitwise equal to the
reference's Fortran -- tests/test_oracle.py) and the golden vectors.

Tolerances, relative to the largest entry, from kappa * n * u (two
backward-stable algorithms agree to about that): random SPD matrices with
kappa <= 6: 1e-13; the golden Gaussian system (kappa 1.5e6, n 40): 1e-7;
the golden ARAD system (sigma 10, lambda 1e-3; kappa 2.9e4, n 10): 1e-10;
50 uneven molecules (sigma 20, lambda 1e-2; kappa 1e6): 1e-7."""
import os

import numpy as np
import pytest

import oracle
from exact_residual import exact_refined_solution, exact_residual
from fml_amd import synth
from fml_amd.arad import ARAD
from fml_amd.kernels.kernels import arad_types

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    B = rng.normal(size=(n, n))
    return B @ B.T / n + np.eye(n)


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 130, 600])
def test_cho_invert_matches_lapack(gpu, n):
    from fml_amd.math import cho_invert
    A = np.asfortranarray(_spd(n, 50 + n))
    A_before = A.copy()
    got = cho_invert(A)
    assert np.array_equal(A, A_before)  # the caller's matrix is not overwritten
    ref, info = oracle.cho_invert(A)
    assert info == 0
    up = np.triu_indices(n)
    scale = np.max(np.abs(ref[up]))
    assert np.max(np.abs(got[up] - ref[up])) <= 1e-13 * scale
    assert np.array_equal(np.tril(got, -1), np.tril(A, -1))  # strictly lower part: A's own
    full = np.triu(got) + np.triu(got, 1).T
    assert np.max(np.abs(full @ A - np.eye(n))) < 1e-12


@pytest.mark.parametrize("n", [2200, 4300])
def test_cho_invert_split_product(gpu, n):
    """Sizes whose triangular-inverse products run split over k (block rows
    deeper than 2048 columns; 4300 gives three chunks): against LAPACK's
    inverse (numpy) and through inv(A) A = I."""
    from fml_amd.math import cho_invert
    A = np.asfortranarray(_spd(n, 7 + n))
    got = cho_invert(A)
    full = np.triu(got) + np.triu(got, 1).T
    ref = np.linalg.inv(A)
    assert np.max(np.abs(full - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.max(np.abs(full @ A - np.eye(n))) < 1e-11


def test_cho_invert_golden_and_failure(gpu, capsys):
    from fml_amd.math import cho_invert
    d = np.load(os.path.join(GOLD, "inverse.npz"))
    got = cho_invert(d["A"])
    up = np.triu_indices(d["A"].shape[0])
    ref = d["inv"][up]
    # the golden system is the 40-molecule Gaussian kernel + 1e-6 I: kappa 1.5e6
    assert np.max(np.abs(got[up] - ref)) <= 1e-7 * np.max(np.abs(ref))
    A = np.eye(6)
    A[3, 3] = -1.0
    bad = cho_invert(A)
    assert "DPOTRF() exited with error code: 4" in capsys.readouterr().out
    assert np.isnan(bad[np.triu_indices(6)]).all()


def test_fget_alpha_arad(gpu):
    from fml_amd.kernels import fget_alpha_arad
    a = np.load(os.path.join(GOLD, "arad.npz"))
    d = np.load(os.path.join(GOLD, "inverse.npz"))
    got = fget_alpha_arad(a["X"], a["z"], a["n"], d["y_arad"], 10.0, 1e-3)
    ref = d["alpha_arad"]
    assert got.shape == ref.shape == (int(a["n"].sum()),)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    # the reference fills the caller's alpha in place
    alpha = np.zeros_like(ref)
    out = fget_alpha_arad(a["X"], a["z"], a["n"], d["y_arad"], 10.0, 1e-3, alpha)
    assert out is alpha and np.array_equal(alpha, got)
    # per molecule it equals the cho_solve of the same system
    from fml_amd.kernels import fget_kernel_arad
    from fml_amd.math import cho_solve
    K = fget_kernel_arad(a["X"], a["z"], a["n"], 10.0) + 1e-3 * np.eye(len(a["n"]))
    x = cho_solve(K, d["y_arad"])
    per_mol = got[np.cumsum(a["n"]) - 1]
    assert np.max(np.abs(per_mol - x)) <= 1e-9 * np.max(np.abs(x))


def test_fget_alpha_arad_uneven_molecules(gpu):
    from fml_amd.kernels import fget_alpha_arad
    ms = synth.generate(50, seed=51, nmin=1, nmax=29, dist="geometric")
    ar = ARAD(maxMolSize=29, maxAts=29)
    X = np.array([ar.describe(ms.molecule(m)[1], ms.molecule(m)[0]) for m in range(ms.nmol)])
    z, n = arad_types([ms.molecule(m)[0] for m in range(ms.nmol)], 29)
    got = fget_alpha_arad(X, z, n, ms.energies, 20.0, 1e-2)
    ref, info = oracle.alpha_arad(X, z, n, ms.energies, 20.0, 1e-2)
    assert info == 0
    assert np.max(np.abs(got - ref)) <= 1e-7 * np.max(np.abs(ref))


def test_potrf_lookahead_concurrent_threads(gpu):
    """Factorisations large enough for the panel look-ahead (3 outer panels),
    enqueued at the same time from two host threads on two streams, give
    the same factors bit for bit as one at a time."""
    import threading
    import torch
    from fml_amd.math.cho_solve import potrf_inplace
    mats = [torch.from_numpy(_spd(3000, 90 + k)).cuda() for k in range(2)]
    ref = []
    for m in mats:
        f = m.clone()
        assert potrf_inplace(f) == 0
        ref.append(f)
    out = [m.clone() for m in mats]
    streams = [torch.cuda.Stream() for _ in mats]
    for st in streams:  # the clones (on the current stream) come first
        st.wait_stream(torch.cuda.current_stream())
    infos = [None, None]
    barrier = threading.Barrier(2)

    def run(k):
        with torch.cuda.stream(streams[k]):
            barrier.wait()
            infos[k] = potrf_inplace(out[k], sync=False)
    ths = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    torch.cuda.synchronize()
    for k in range(2):
        assert int(infos[k].item()) == 0
        assert torch.equal(torch.tril(out[k]), torch.tril(ref[k]))


def test_ozaki_factor_and_refinement(gpu):
    """fmlh_potrf_ozaki: the next-panel and trailing updates from the exact
    five-digit INT8 split (n = 3000: two panels of look-ahead updates).  The
    factor is within 1e-10 of the FP64 factor and |L L^T - A| / |A| <= 1e-10
    (the dropped digit products are below ~2^-38 of the row and column
    scales per update); KRR(precision="ozaki") refines to a relative residual
    no larger than the FP64 direct solve's, and its predictions agree with
    the FP64 model's to 1e-9 relative."""
    torch = gpu
    from fml_amd import synth
    from fml_amd.krr import KRR, kernel_matrix, predict_from_kernel
    from fml_amd.math.cho_solve import potrf_inplace, potrf_ozaki_inplace
    from fml_amd.representations import generate_coulomb_matrices
    rng = np.random.default_rng(41)
    n = 3000
    B = rng.normal(size=(n, n))
    A = B @ B.T / n + np.eye(n)
    M1 = torch.from_numpy(A).cuda()
    M2 = M1.clone()
    assert potrf_inplace(M1) == 0 and potrf_ozaki_inplace(M2) == 0
    L1, L2 = torch.tril(M1), torch.tril(M2)
    Ad = torch.from_numpy(A).cuda()
    diff = float((L1 - L2).abs().max() / L1.abs().max())
    berr = float((L2 @ L2.T - Ad).abs().max() / Ad.abs().max())
    print(f"\nOzaki factor n={n}: vs FP64 factor {diff:.3g}, |LL^T - A|/|A| {berr:.3g}")
    assert 0 < diff <= 1e-10 and berr <= 1e-10
    # a KRR system: 3000 molecules, 2600 train
    ms = synth.generate(3000, seed=42, nmin=3, nmax=29)
    X = generate_coulomb_matrices(ms.charges, ms.coords, ms.offsets, nmax=29)
    y = ms.energies
    nt = 2600
    m64 = KRR(1000.0, 1e-6).fit(X[:nt], y[:nt])
    moz = KRR(1000.0, 1e-6, precision="ozaki", tol=1e-12, max_iter=20).fit(X[:nt], y[:nt])
    K = kernel_matrix(m64.X, m64.X, 1000.0)
    yd = torch.from_numpy(y[:nt].copy()).cuda()
    r = torch.empty_like(yd)

    def rel_res(a):
        predict_from_kernel(K, a, out=r)
        r.add_(a, alpha=1e-6)
        return float(torch.linalg.norm(yd - r) / torch.linalg.norm(yd))

    r64, roz = rel_res(m64.alpha), rel_res(moz.alpha)
    p64, poz = m64.predict(X[nt:]), moz.predict(X[nt:])
    pdiff = float((p64 - poz).abs().max() / p64.abs().max())
    # alpha against the solution of the same FP64 system, refined on the CPU
    # with exactly rounded residuals
    xs = exact_refined_solution(K.cpu().numpy(), 1e-6, y[:nt])
    e_oz = float(np.linalg.norm(moz.alpha.cpu().numpy() - xs) / np.linalg.norm(xs))
    e_64 = float(np.linalg.norm(m64.alpha.cpu().numpy() - xs) / np.linalg.norm(xs))
    # the same refinement from the FP64 factor, on the device
    from fml_amd.krr import refine_dd
    F64 = K.clone()
    F64.diagonal().add_(1e-6)
    assert potrf_inplace(F64) == 0
    a2 = m64.alpha.clone()
    h2, _, _ = refine_dd(K, 1e-6, F64, yd, a2, tol=1e-14)
    e_2 = float(np.linalg.norm(a2.cpu().numpy() - xs) / np.linalg.norm(xs))
    print(f"KRR 2600: relative residual FP64 {r64:.3g}, Ozaki refined {roz:.3g} "
          f"({moz.report.iterations} steps: corrections {moz.report.history}), prediction diff "
          f"{pdiff:.3g}; alpha vs the exact solution: Ozaki {e_oz:.3g}, plain FP64 {e_64:.3g}, "
          f"FP64 factor refined on the device {e_2:.3g} (corrections {h2})")
    assert roz <= max(r64, 1e-12)
    assert pdiff <= 1e-9
    assert e_oz <= 1e-8


def test_residual_dd_against_exact(gpu):
    """fmlh_residual_dd equals the exactly rounded residual to a few units
    in the last place (Dot2 bound) on a KRR system whose FP64 residual is
    dominated by cancellation, and differs from the plain FP64 residual."""
    torch = gpu
    from fml_amd import synth
    from fml_amd.krr import kernel_matrix, residual_dd, sym_residual
    from fml_amd.math.cho_solve import potrf_inplace, potrs_inplace
    from fml_amd.representations import generate_coulomb_matrices
    ms = synth.generate(1500, seed=7, nmin=3, nmax=29)
    X = torch.from_numpy(generate_coulomb_matrices(ms.charges, ms.coords, ms.offsets, nmax=29)).cuda()
    K = kernel_matrix(X, X, 1000.0)
    n = int(K.shape[0])
    y = torch.from_numpy(ms.energies.copy()).cuda()
    F = K.clone()
    F.diagonal().add_(1e-6)
    assert potrf_inplace(F) == 0
    x = y.clone()
    potrs_inplace(F, x)
    r = torch.empty_like(y)
    residual_dd(K, 1e-6, x, y, r)
    ex = exact_residual(K.cpu().numpy(), 1e-6, x.cpu().numpy(), ms.energies)
    err = float(np.max(np.abs(r.cpu().numpy() - ex)) / np.max(np.abs(ex)))
    r64 = torch.empty_like(y)
    sym_residual(K, 1e-6, x, y, r64)
    err64 = float(np.max(np.abs(r64.cpu().numpy() - ex)) / np.max(np.abs(ex)))
    print(f"\nresidual n={n}: double-length vs exact {err:.3g}, plain FP64 vs exact {err64:.3g}")
    assert err <= 1e-14
    assert err64 > 10 * err


def test_potrs_correction_against_triangular_solves(gpu):
    """fmlh_potrs_correction (full diagonal blocks applied as inverses, the
    refinement's correction solve) against two LAPACK triangular solves
    (scipy solve_triangular) with the same factor: to working accuracy on
    well-conditioned matrices, including a ragged last block, and close to
    the substitution solve's own error on an ill-conditioned KRR factor."""
    torch = gpu
    from scipy.linalg import solve_triangular
    from fml_amd import synth
    from fml_amd.krr import kernel_matrix
    from fml_amd.math.cho_solve import potrf_inplace, potrs_correction_inplace, potrs_inplace
    from fml_amd.representations import generate_coulomb_matrices
    rng = np.random.default_rng(43)
    cases = []
    for n in (3000, 1037, 64, 50):
        B = rng.normal(size=(n, n))
        cases.append((f"wishart {n}", torch.from_numpy(B @ B.T / n + np.eye(n)).cuda(), 1e-13))
    ms = synth.generate(1500, seed=44, nmin=3, nmax=29)
    X = torch.from_numpy(generate_coulomb_matrices(ms.charges, ms.coords, ms.offsets, nmax=29)).cuda()
    K = kernel_matrix(X, X, 1000.0)
    K.diagonal().add_(1e-6)
    cases.append(("KRR 1500, lambda 1e-6", K, None))
    for name, A, tol in cases:
        n = int(A.shape[0])
        F = A.clone()
        assert potrf_inplace(F) == 0
        b = rng.normal(size=n)
        xs = potrs_inplace(F, torch.from_numpy(b.copy()).cuda()).cpu().numpy()
        xc = potrs_correction_inplace(F, torch.from_numpy(b.copy()).cuda()).cpu().numpy()
        L = torch.tril(F).cpu().numpy()
        xr = solve_triangular(L.T, solve_triangular(L, b, lower=True), lower=False)
        e_s = float(np.linalg.norm(xs - xr) / np.linalg.norm(xr))
        e_c = float(np.linalg.norm(xc - xr) / np.linalg.norm(xr))
        print(f"\n{name}: substitution vs LAPACK {e_s:.3g}, inverse diagonal blocks vs LAPACK {e_c:.3g}")
        if tol is not None:
            assert e_s <= tol and e_c <= tol
        else:
            assert e_c <= max(100 * e_s, 1e-12)


def test_factor_leaves_the_other_triangle(gpu):
    """fmlh_potrf and fmlh_potrf_ozaki write only the factor's triangle
    (row-major lower): the strictly upper part is bitwise the caller's, on
    sizes with several 1024-column panels and ragged edges, so K can share
    its storage with its factor."""
    torch = gpu
    from fml_amd.math.cho_solve import potrf_inplace, potrf_ozaki_inplace
    rng = np.random.default_rng(45)
    for n in (2100, 3000):
        B = rng.normal(size=(n, n))
        A = torch.from_numpy(B @ B.T / n + np.eye(n)).cuda()
        for fn in (potrf_inplace, potrf_ozaki_inplace):
            F = A.clone()
            assert fn(F) == 0
            assert torch.equal(torch.triu(F, 1), torch.triu(A, 1)), (fn.__name__, n)


def test_potrf_ragged_middle_blocks(gpu):
    """Outer panels factor their 64-column inner blocks in 256-column middle
    blocks, the rest of the panel updated once per middle block.  Sizes whose
    last panel ends in a short middle block after full ones (1324: 256 + 44;
    2900: 3 x 256 + 84) against LAPACK's factor, FP64 and the Ozaki form."""
    torch = gpu
    from fml_amd.math.cho_solve import potrf_inplace, potrf_ozaki_inplace
    rng = np.random.default_rng(46)
    for n in (1324, 2900):
        B = rng.normal(size=(n, n))
        A = B @ B.T / n + np.eye(n)
        ref = np.linalg.cholesky(A)
        Ad = torch.from_numpy(A).cuda()
        for fn, tol in ((potrf_inplace, 1e-12), (potrf_ozaki_inplace, 1e-10)):
            F = Ad.clone()
            assert fn(F) == 0
            L = torch.tril(F)
            err = np.max(np.abs(L.cpu().numpy() - ref)) / np.max(np.abs(ref))
            berr = float((L @ L.T - Ad).abs().max() / Ad.abs().max())
            assert err <= tol and berr <= tol, (fn.__name__, n, err, berr)


def test_residual_dd_lower_matches_full(gpu):
    """fmlh_residual_dd_lower (K's strictly lower triangle in Fortran order
    plus its diagonal, the other triangle holding the factor) against the
    exactly rounded residual, on a KRR system whose plain FP64 residual is
    dominated by cancellation, and on ragged sizes; then an Ozaki fit on the
    shared storage reaches the exact solution."""
    torch = gpu
    from fml_amd import synth
    from fml_amd.krr import KRR, kernel_matrix, residual_dd, residual_dd_lower
    from fml_amd.math.cho_solve import potrf_inplace, potrs_inplace
    from fml_amd.representations import generate_coulomb_matrices
    for nmol, seed in ((1500, 7), (1000, 8), (63, 9), (1, 10)):
        ms = synth.generate(nmol, seed=seed, nmin=3, nmax=29)
        X = torch.from_numpy(generate_coulomb_matrices(ms.charges, ms.coords, ms.offsets, nmax=29)).cuda()
        K = kernel_matrix(X, X, 1000.0)
        n = int(K.shape[0])
        y = torch.from_numpy(ms.energies.copy()).cuda()
        F = K.clone()
        F.diagonal().add_(1e-6)
        assert potrf_inplace(F) == 0
        x = y.clone()
        potrs_inplace(F, x)
        ex = exact_residual(K.cpu().numpy(), 1e-6, x.cpu().numpy(), ms.energies)
        # the shared matrix: factor in the lower triangle, K above it
        S = torch.tril(F) + torch.triu(K, 1)
        d = K.diagonal().clone()
        r = torch.empty_like(y)
        residual_dd_lower(S, d, 1e-6, x, y, r)
        r_full = torch.empty_like(y)
        residual_dd(K, 1e-6, x, y, r_full)
        scale = max(np.max(np.abs(ex)), 1e-16 * np.max(np.abs(ms.energies)))  # n = 1 may cancel exactly
        err = float(np.max(np.abs(r.cpu().numpy() - ex)) / scale)
        err_full = float(np.max(np.abs(r_full.cpu().numpy() - ex)) / scale)
        print(f"\nresidual n={n}: triangle form vs exact {err:.3g}, full form {err_full:.3g}")
        assert err <= 1e-14
    # the Ozaki fit (shared storage) against the exact solution
    ms = synth.generate(2000, seed=11, nmin=3, nmax=29)
    X = generate_coulomb_matrices(ms.charges, ms.coords, ms.offsets, nmax=29)
    m = KRR(1000.0, 1e-6, precision="ozaki", tol=1e-12, max_iter=20).fit(X, ms.energies)
    K = kernel_matrix(m.X, m.X, 1000.0)
    xs = exact_refined_solution(K.cpu().numpy(), 1e-6, ms.energies)
    e = float(np.linalg.norm(m.alpha.cpu().numpy() - xs) / np.linalg.norm(xs))
    print(f"Ozaki fit on the shared matrix, n=2000: alpha vs exact {e:.3g} ({m.report.iterations} steps)")
    assert e <= 1e-8
