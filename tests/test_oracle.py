"""Pins the oracle (CPU, no GPU needed).

The C restatement (oracle/fml_oracle.c) must equal, bit for bit, the golden
vectors written by the reference's own Fortran (tests/golden/make_golden.py)
and, where it is built, the reference library itself.  Closed forms and
invariants cover what a plausible mistake would break: a transposed output,
a wrong tie rule, a missing sine guard, a dropped factor in the ARAD overlap,
the wrong eps literal.
"""
import glob
import math
import os

import numpy as np
import pytest

import oracle
from oracle.pyref import arad_atomic_distance_py, arad_describe
from fml_amd import synth
from fml_amd.arad import ARAD
from fml_amd.fml import Molecule
from fml_amd.kernels.kernels import arad_types

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def g(name):
    return np.load(os.path.join(GOLD, name))


def xyz_mols():
    out = []
    for f in sorted(glob.glob(os.path.join(GOLD, "xyz", "*.xyz"))):
        m = Molecule()
        m.read_xyz(f)
        out.append(m)
    return out


# --------------------------------------------------------------- Coulomb

def test_coulomb_golden_bitwise():
    d = g("coulomb.npz")
    for m, ref in zip(xyz_mols(), d["cm_xyz"]):
        cm, _ = oracle.coulomb_matrix(m.nuclear_charges, m.coordinates, 23)
        assert np.array_equal(cm, ref)
    off = d["syn_offsets"]
    cm, _ = oracle.coulomb_matrices(d["syn_charges"], d["syn_coords"], off, 29)
    assert np.array_equal(cm, d["cm_syn"])


def test_coulomb_closed_form_and_ties():
    # H2 at 0.74: diag 0.5, pair 1/0.74
    cm, perm = oracle.coulomb_matrix([1, 1], [[0, 0, 0], [0, 0, 0.74]], 3)
    assert np.array_equal(cm[:3], [0.5, 1 / 0.74, 0.5]) and np.all(cm[3:] == 0)
    # exact tie: higher original index first (frepresentations.f90:87-91)
    assert list(perm) == [1, 0]
    # water-like: O first (largest norm), then the two H (tied) high index first
    z = [1, 8, 1]
    r = [[0.0, 0.0, 1.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
    cm, perm = oracle.coulomb_matrix(z, r, 3)
    assert list(perm) == [1, 2, 0]
    o = 0.5 * 8 ** 2.4
    assert cm[0] == o and cm[1] == 8.0 and cm[2] == 0.5
    assert cm[3] == 8.0 and cm[4] == 1 / math.sqrt(2.0) and cm[5] == 0.5
    with pytest.raises(ValueError):
        oracle.coulomb_matrix([1] * 4, np.zeros((4, 3)) + np.arange(4)[:, None], 3)


def test_coulomb_invariants():
    ms = synth.generate(30, seed=2, nmin=3, nmax=20)
    for i in range(ms.nmol):
        z, r = ms.molecule(i)
        cm, perm = oracle.coulomb_matrix(z, r, 20)
        n = len(z)
        M = np.zeros((n, n))
        # rebuild the sorted matrix from the packed row-major lower triangle
        t = 0
        for a in range(n):
            for b in range(a + 1):
                M[a, b] = M[b, a] = cm[t]
                t += 1
        norms = M.sum(axis=1)
        assert np.all(np.diff(norms) <= 1e-9 * np.abs(norms[:-1]))
        # permuting the input atoms leaves the matrix unchanged (no ties here)
        p = np.random.default_rng(i).permutation(n)
        cm2, _ = oracle.coulomb_matrix(z[p], r[p], 20)
        assert np.allclose(cm2, cm, rtol=1e-15, atol=0)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
def test_coulomb_vs_reference_library():
    ms = synth.generate(60, seed=9, nmin=3, nmax=29)
    for i in range(ms.nmol):
        z, r = ms.molecule(i)
        assert np.array_equal(oracle.coulomb_matrix(z, r, 29)[0], oracle.ref_coulomb_matrix(z, r, 29))


# ---------------------------------------------------------- dense kernels

def test_dense_golden_bitwise():
    d = g("dense.npz")
    A, B = d["A"], d["B"]
    assert np.array_equal(oracle.gaussian_kernel(A, B, 700.0), d["gauss_700"])
    assert np.array_equal(oracle.gaussian_kernel(A, B, 50.0), d["gauss_50"])
    assert np.array_equal(oracle.laplacian_kernel(A, B, 1e4), d["lap_1e4"])
    assert np.array_equal(oracle.laplacian_kernel(A, B, 300.0), d["lap_300"])


def test_dense_brute_force():
    rng = np.random.default_rng(4)
    A = rng.normal(size=(7, 5))
    B = rng.normal(size=(7, 3))
    Kg = oracle.gaussian_kernel(A, B, 1.3)
    Kl = oracle.laplacian_kernel(A, B, 1.3)
    for j in range(5):
        for i in range(3):
            assert Kg[j, i] == pytest.approx(math.exp(-0.5 / 1.69 * np.sum((A[:, j] - B[:, i]) ** 2)), rel=1e-14)
            assert Kl[j, i] == pytest.approx(math.exp(-np.sum(np.abs(A[:, j] - B[:, i])) / 1.3), rel=1e-14)
    assert Kg.flags.f_contiguous and Kg.shape == (5, 3)
    assert np.all(np.diag(oracle.gaussian_kernel(A, A, 0.1)) == 1.0)


# ----------------------------------------------------------------- ARAD

def _arad():
    d = g("arad.npz")
    return d, d["X"], d["z"], d["n"], list(d["sigmas"])


def test_arad_golden_bitwise():
    d, X, z, n, sig = _arad()
    assert np.array_equal(oracle.kernels_arad(X[:5], X[5:], z[:5], z[5:], n[:5], n[5:], sig),
                          d["kernels_default"])
    assert np.array_equal(oracle.kernels_arad(X[:5], X[5:], z[:5], z[5:], n[:5], n[5:], sig, 0.2, 5.0),
                          d["kernels_cut5"])
    assert np.array_equal(oracle.symmetric_kernels_arad(X, z, n, sig), d["symmetric"])
    assert np.array_equal(oracle.kernel_arad(X, z, n, 1.0), d["kernel_s1"])
    assert np.array_equal(oracle.arad_l2_distance_all(X[:5], X[5:], z[:5], z[5:], n[:5], n[5:]),
                          d["distance_all"])
    qa, za, na = d["qa"], d["za"], d["na"]
    assert np.array_equal(oracle.atomic_kernels_arad(qa, qa, za, za, na, na, sig), d["atomic_kernels"])
    # the reference stores D(i,j) at kernel(j,i) (farad_kernels.f90:1019)
    assert np.array_equal(oracle.atomic_distance_arad(qa, qa, za, za, na, na),
                          d["atomic_distance_ref_layout"].T)


def test_arad_against_reference_test_restatement():
    """tests/test_arad.py:278-308 restates the overlap independently (sines
    inline, |d2-d1| < 8w); its distances must match the oracle's."""
    d, X, z, n, sig = _arad()
    mols = xyz_mols()
    for a, b in ((0, 5), (3, 8), (2, 2)):
        Za, Zb = mols[a].nuclear_charges, mols[b].nuclear_charges
        Dp = arad_atomic_distance_py(X[a], X[b], Za, Zb)
        Do = oracle.arad_l2_distance_all(X[a:a + 1], X[b:b + 1], z[a:a + 1], z[b:b + 1],
                                         n[a:a + 1], n[b:b + 1])[0, 0, :len(Za), :len(Zb)]
        # the oracle clamps |D| < eps to zero (farad_kernels.f90:594)
        mask = np.abs(Dp) >= 1e-10
        assert np.allclose(Do[mask], Dp[mask], rtol=1e-12, atol=1e-12)
        assert np.all(np.abs(Do[~mask]) < 1e-10)


def test_arad_eps_literal():
    """The clamp constants are single-precision literals widened to double
    (farad_kernels.f90:148, 505)."""
    assert float(np.float32(5.0e-12)) == 4.999999980020986e-12
    assert float(np.float32(1.0e-11)) == 9.999999960041972e-12


def test_arad_invariants():
    d, X, z, n, sig = _arad()
    # D(self, self) == 0 exactly; kernel of a molecule with itself, sigma=0.1:
    # only identical atom environments contribute, so K >= natoms
    D = oracle.arad_l2_distance_all(X, X, z, z, n, n)
    for i in range(X.shape[0]):
        assert np.all(np.diag(D[i, i, :n[i], :n[i]]) == 0.0)
    K = oracle.kernels_arad(X, X, z, z, n, n, sig)
    for i in range(X.shape[0]):
        assert K[0, i, i] >= n[i] - 1e-12
    # molecular kernel == sum of the atomic kernel over the molecule-pair block
    qa = np.concatenate([X[m, :n[m]] for m in range(X.shape[0])])
    za = np.concatenate([z[m, :n[m]] for m in range(X.shape[0])])
    na = np.concatenate([[n[m]] * n[m] for m in range(X.shape[0])]).astype(np.int32)
    Ka = oracle.atomic_kernels_arad(qa, qa, za, za, na, na, sig)
    Ks = oracle.symmetric_kernels_arad(X, z, n, sig)
    off = np.concatenate([[0], np.cumsum(n)])
    for i in range(X.shape[0]):
        for j in range(X.shape[0]):
            blk = Ka[:, off[i]:off[i + 1], off[j]:off[j + 1]].sum(axis=(1, 2))
            assert np.allclose(blk, Ks[:, i, j], rtol=1e-12, atol=0)
    # general X-vs-X equals the symmetric form to rounding (not bitwise:
    # the loop orientation changes the summation order)
    assert np.allclose(K, oracle.symmetric_kernels_arad(X, z, n, sig, 0.2, 6.0), rtol=1e-13)
    # kernel == sum exp(-D / sigma^2) (tests/test_arad.py:382-390)
    for s, sg in enumerate(sig):
        # i <= j: the symmetric kernel computes that orientation (and mirrors)
        for i in range(X.shape[0]):
            for j in range(i, X.shape[0]):
                kd = np.sum(np.exp(D[i, j, :n[i], :n[j]] * (-1.0 / sg ** 2)))
                assert kd == pytest.approx(Ks[s, i, j], rel=1e-13, abs=0)


def test_arad_overlap_closed_form():
    """Two one-neighbour atoms at distances d1, d2 (same element): S is one
    term exp(-(d2-d1)^2/(4w^2)) s1 s2 (1 + a1 a2 + b1 b2); beyond 8w it is 0;
    a neighbour beyond the cut ends the list."""
    def atom(d, a, b):
        X = np.zeros((1, 2, 5, 2))
        X[0, 0, 0] = [d, 1e100]
        X[0, 0, 1] = [2, 0]
        X[0, 0, 2] = [4, 0]
        X[0, 0, 3] = [a, 0]
        X[0, 0, 4] = [b, 0]
        return X
    z = np.zeros((1, 2, 2))
    z[0, 0] = [2, 4]
    n = np.array([1], dtype=np.int32)
    w, cut = 0.2, 5.0
    s = lambda d: 1.0 - math.sin(d * math.pi / (2 * cut))
    for d1, d2, a1, a2, b1, b2 in ((1.0, 1.3, 0.2, 0.5, 0.1, 0.3), (2.0, 2.0, 0.7, 0.1, 0.0, 0.9)):
        X1, X2 = atom(d1, a1, b1), atom(d2, a2, b2)
        S11 = s(d1) ** 2 * (1 + a1 * a1 + b1 * b1)
        S22 = s(d2) ** 2 * (1 + a2 * a2 + b2 * b2)
        S12 = math.exp(-(d2 - d1) ** 2 / (4 * w * w)) * s(d1) * s(d2) * (1 + a1 * a2 + b1 * b2)
        D = oracle.arad_l2_distance_all(X1, X2, z, z, n, n, w, cut)[0, 0, 0, 0]
        assert D == pytest.approx(S11 + S22 - 2 * S12, rel=1e-13, abs=1e-14)
    # outside the 8w window the cross term vanishes
    X1, X2 = atom(1.0, 0.2, 0.1), atom(1.0 + 8 * w + 1e-9, 0.5, 0.3)
    D = oracle.arad_l2_distance_all(X1, X2, z, z, n, n, w, cut)[0, 0, 0, 0]
    S11 = s(1.0) ** 2 * (1 + 0.04 + 0.01)
    S22 = s(1.0 + 8 * w + 1e-9) ** 2 * (1 + 0.25 + 0.09)
    assert D == pytest.approx(S11 + S22, rel=1e-13)
    # a first neighbour beyond the cut: empty list, D = 0
    X1 = atom(cut + 0.5, 0.2, 0.1)
    assert oracle.arad_l2_distance_all(X1, X1, z, z, n, n, w, cut)[0, 0, 0, 0] == 0.0


def test_arad_descriptor_closed_form():
    """Right-angle triatomic: angles are 0, pi/2, pi/4 closed forms."""
    co = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
    M = arad_describe(co, [6, 1, 1], size=3, cut=5.0)
    w = lambda d: 1.0 - math.sin(math.pi * d / 10.0)
    W = w(0) + 2 * w(1)
    # atom 0: neighbours self (d 0), then atoms 1, 2 (d 1, stable order)
    assert list(M[0, 0]) == [0.0, 1.0, 1.0]
    assert list(M[0, 1]) == [2, 1, 1] and list(M[0, 2]) == [4, 1, 1]
    # channel 3 for neighbour 1: cos(0)*w(0) [self: angle 0] + cos(0)w(1) + cos(90)w(1)
    assert M[0, 3, 1] == pytest.approx((w(0) + w(1) + 0.0) / W, rel=1e-14)
    assert M[0, 4, 1] == pytest.approx((0.0 + 0.0 + 1.0 * w(1)) / W, rel=1e-14)
    # atom 1: neighbours self, atom 0 (d 1), atom 2 (d sqrt 2); angle 0-1-2 = 45 deg
    assert M[1, 0, 2] == pytest.approx(math.sqrt(2.0))
    W1 = w(0) + w(1) + w(math.sqrt(2))
    assert M[1, 4, 1] == pytest.approx(math.sin(math.pi / 4) * w(math.sqrt(2)) / W1, rel=1e-14)
    assert np.all(M[:, 0, 3:] == 1e100) if M.shape[2] > 3 else True


def test_arad_descriptor_product_vs_restatement():
    """fml_amd.ARAD (vectorised, the product's host code) against the loop
    restatement on the reference's xyz molecules (incl. linear C2H2)."""
    ar = ARAD(maxMolSize=30, maxAts=30)
    for m in xyz_mols():
        a = ar.describe(m.coordinates, m.nuclear_charges)
        b = arad_describe(m.coordinates, m.nuclear_charges, size=30)
        fin = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), fin)
        assert np.allclose(a[fin], b[fin], rtol=1e-12, atol=1e-13)
    # invariants: ascending distances, self first, |ch3|,|ch4| <= 1
    for m in xyz_mols():
        a = ar.describe(m.coordinates, m.nuclear_charges)
        for i in range(len(m.nuclear_charges)):
            k = np.sum(a[i, 0] < 1e99)
            assert a[i, 0, 0] == 0.0 and np.all(np.diff(a[i, 0, :k]) >= 0)
            assert np.all(np.abs(a[i, 3:5, :k][np.isfinite(a[i, 3:5, :k])]) <= 1 + 1e-12)


# ------------------------------------------------------------- cho_solve

def test_cho_solve_golden():
    d = g("cho_solve.npz")
    x, info = oracle.cho_solve(d["A"], d["y"])
    assert info == 0
    assert np.array_equal(x, d["x"])  # scipy LAPACK == reference fcho_solve


def test_cho_solve_closed_forms():
    rng = np.random.default_rng(0)
    for n in (1, 3, 17):
        L = np.tril(rng.integers(-1, 2, size=(n, n)), -1).astype(float) + 2 * np.eye(n)
        A = L @ L.T
        xt = rng.integers(-5, 6, size=n).astype(float)
        x, info = oracle.cho_solve(A, A @ xt)
        assert info == 0 and np.allclose(x, xt, atol=1e-12)
        x2, info2 = oracle.cho_solve_c(A, A @ xt)
        assert info2 == 0 and np.allclose(x2, xt, atol=1e-12)
    A = np.eye(4)
    A[2, 2] = -1
    assert oracle.cho_solve(A, np.ones(4))[1] == 3
    assert oracle.cho_solve_c(A, np.ones(4))[1] == 3
    # only the upper triangle is read (fcho_solve.f90:35 'U')
    A = np.array([[4.0, 2.0], [999.0, 3.0]])
    x, _ = oracle.cho_solve(A, np.array([1.0, 2.0]))
    assert np.allclose(np.array([[4, 2], [2, 3.0]]) @ x, [1, 2])


def test_cho_invert_and_alpha_arad_golden():
    """scipy's dpotrf + dpotri (+ dsymv) equal the reference's fcho_invert
    and fget_alpha_arad bit for bit (same LAPACK), the latter on the C
    oracle's molecular kernel."""
    d = g("inverse.npz")
    inv, info = oracle.cho_invert(d["A"])
    assert info == 0
    assert np.array_equal(np.triu(inv), np.triu(d["inv"]))
    assert np.array_equal(np.tril(d["inv"], -1), np.tril(d["A"], -1))  # dpotri('U') leaves the rest
    a = g("arad.npz")
    alpha, info = oracle.alpha_arad(a["X"], a["z"], a["n"], d["y_arad"], 10.0, 1e-3)
    assert info == 0
    assert np.array_equal(alpha, d["alpha_arad"])


def test_cho_invert_and_alpha_arad_closed_forms():
    # inverse of L L^T with a small-integer unit-diagonal L is an integer matrix
    rng = np.random.default_rng(5)
    for n in (1, 4, 9):
        L = np.tril(rng.integers(-2, 3, size=(n, n)), -1).astype(float) + np.eye(n)
        A = L @ L.T
        inv, info = oracle.cho_invert(A)
        Li = np.linalg.inv(L)
        assert info == 0 and np.allclose(np.triu(inv), np.triu(np.round(Li.T @ Li)), atol=1e-9)
    # one molecule: alpha = y / (K + lambda) for each of its atoms
    ms = synth.generate(1, seed=6, nmin=7, nmax=7)
    X = ARAD(maxMolSize=7, maxAts=7).describe(ms.molecule(0)[1], ms.molecule(0)[0])[None]
    z, n = arad_types([ms.molecule(0)[0]], 7)
    k = oracle.kernel_arad(X, z, n, 3.0)[0, 0]
    alpha, info = oracle.alpha_arad(X, z, n, [2.5], 3.0, 0.25)
    assert info == 0 and alpha.shape == (7,)
    assert np.allclose(alpha, 2.5 / (k + 0.25), rtol=1e-14)


def test_oracle_ptp_and_charges():
    from oracle.pyref import NUCLEAR_CHARGE_ORACLE, ptp_types
    from fml_amd.data import NUCLEAR_CHARGE
    from fml_amd.ptp import PTP
    assert {k: NUCLEAR_CHARGE[k] for k in ("H", "C", "N", "O", "S", "Cl")} == \
        {"H": 1, "C": 6, "N": 7, "O": 8, "S": 16, "Cl": 17}
    assert NUCLEAR_CHARGE == {k: v for k, v in NUCLEAR_CHARGE_ORACLE.items()}
    for zz, rc in {1: (1, 1), 6: (2, 4), 7: (2, 5), 8: (2, 6), 16: (3, 6)}.items():
        assert tuple(PTP[zz]) == rc == ptp_types(zz)
    for zz in range(1, 119):
        assert tuple(PTP[zz]) == ptp_types(zz)
    assert PTP[101] == [7, 32] and PTP[102] == [7, 14]


# ------------------------------------------------------- atomic KRR path

def test_coulomb_forms_golden_bitwise():
    """Unsorted, local and atomic Coulomb matrices of the reference's xyz
    molecules against the reference Fortran's outputs."""
    d = g("atomic.npz")
    z, r, off = d["charges"], d["coords"], d["offsets"]
    assert np.array_equal(oracle.coulomb_forms(oracle.UNSORTED, z, r, off, 23), d["cm_unsorted"])
    assert np.array_equal(oracle.coulomb_forms(oracle.LOCAL, z, r, off, 23), d["cm_local"])
    assert np.array_equal(oracle.coulomb_forms(oracle.ATOMIC, z, r, off, 23), d["cm_atomic"])


def test_coulomb_forms_closed_form():
    """O at the origin, H at (1,0,0) and (0,2,0): pair values 8, 4, 1/sqrt5;
    row norms order O, H1, H2; distances from H1 order H1, O, H2."""
    Z = np.array([8.0, 1.0, 1.0])
    R = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 2.0, 0]])
    d0, dh = 0.5 * 8.0 ** 2.4, 0.5
    h12 = 1.0 / math.sqrt(5.0)
    un = oracle.coulomb_form(oracle.UNSORTED, Z, R, 4)
    assert un[:6] == pytest.approx([d0, 8.0, dh, 4.0, h12, dh], rel=1e-15) and np.all(un[6:] == 0)
    loc = oracle.coulomb_form(oracle.LOCAL, Z, R, 3)
    # row of H2 (k=2): order H2, O, H1
    assert loc[2] == pytest.approx([dh, 4.0, d0, h12, 8.0, dh], rel=1e-15)
    # row of O (k=0): order O, H1, H2 (the sorted list itself)
    assert loc[0] == pytest.approx([d0, 8.0, dh, 4.0, h12, dh], rel=1e-15)
    at = oracle.coulomb_form(oracle.ATOMIC, Z, R, 3)
    # row of H1: order H1 (0), O (1), H2 (sqrt5)
    assert at[1] == pytest.approx([dh, 8.0, d0, h12, 4.0, dh], rel=1e-15)
    # row of H2: order H2 (0), O (2), H1 (sqrt5)
    assert at[2] == pytest.approx([dh, 4.0, d0, h12, 8.0, dh], rel=1e-15)
    with pytest.raises(ValueError):
        oracle.coulomb_form(oracle.LOCAL, Z, R, 2)


def test_coulomb_forms_invariants():
    """Row k of the local and atomic forms starts with atom k's diagonal; the
    atomic form's first column lists pair values by ascending distance."""
    ms = synth.generate(12, seed=21, nmin=2, nmax=16)
    for i in range(ms.nmol):
        Z, R = ms.molecule(i)
        loc = oracle.coulomb_form(oracle.LOCAL, Z, R, 16)
        at = oracle.coulomb_form(oracle.ATOMIC, Z, R, 16)
        diag = 0.5 * Z ** 2.4
        assert np.allclose(loc[:, 0], diag, rtol=1e-15) and np.allclose(at[:, 0], diag, rtol=1e-15)
        n = len(Z)
        tri = [m * (m + 1) // 2 for m in range(n)]   # (m, 0) entries: pair with atom k
        for k in range(n):
            dk = np.sqrt(((R - R[k]) ** 2).sum(axis=1))
            order = np.argsort(dk, kind="stable")
            expect = [diag[k]] + [Z[k] * Z[j] / dk[j] for j in order[1:]]
            assert np.allclose(at[k, tri], expect, rtol=1e-14)


def test_vector_kernels_golden_bitwise():
    d = g("atomic.npz")
    assert np.array_equal(oracle.vector_kernels(0, d["q1"], d["q2"], d["n1"], d["n2"], d["sigmas_gauss"]),
                          d["vec_gauss"])
    assert np.array_equal(oracle.vector_kernels(1, d["q1"], d["q2"], d["n1"], d["n2"], d["sigmas_lap"]),
                          d["vec_lap"])


def test_vector_kernels_closed_form():
    """Brute force with numpy, and one-atom molecules = the dense kernels."""
    rng = np.random.default_rng(22)
    nm1, nm2, amax, d = 5, 4, 6, 9
    q1, q2 = rng.standard_normal((nm1, amax, d)), rng.standard_normal((nm2, amax, d))
    n1, n2 = rng.integers(1, amax + 1, nm1), rng.integers(1, amax + 1, nm2)
    sig = [0.7, 3.0]
    for kind in (0, 1):
        got = oracle.vector_kernels(kind, q1, q2, n1, n2, sig)
        for s, sg in enumerate(sig):
            for i in range(nm1):
                for j in range(nm2):
                    a, b = q1[i, :n1[i]], q2[j, :n2[j]]
                    diff = a[:, None, :] - b[None, :, :]
                    dist = (diff ** 2).sum(-1) if kind == 0 else np.abs(diff).sum(-1)
                    inv = -0.5 / sg ** 2 if kind == 0 else -1.0 / sg
                    assert got[s, i, j] == pytest.approx(np.exp(dist * inv).sum(), rel=1e-13)
    ones1, ones2 = np.ones(nm1, np.int32), np.ones(nm2, np.int32)
    A, B = q1[:, 0, :].T, q2[:, 0, :].T
    assert np.array_equal(oracle.vector_kernels(0, q1, q2, ones1, ones2, [2.0])[0],
                          oracle.gaussian_kernel(A, B, 2.0))
    assert np.array_equal(oracle.vector_kernels(1, q1, q2, ones1, ones2, [2.0])[0],
                          oracle.laplacian_kernel(A, B, 2.0))


def _laplacian_mol_kernel(x, n, sigma):
    E = np.exp(-np.abs(x[:, None, :] - x[None, :, :]).sum(-1) / sigma)
    idx = np.concatenate([[0], np.cumsum(n)[:-1]])
    return np.add.reduceat(np.add.reduceat(E, idx, axis=0), idx, axis=1)


def test_get_alpha_prediction_golden_bitwise():
    d = g("atomic.npz")
    alpha, info = oracle.get_alpha(d["x1"].T, d["n1"], d["y1"], 200.0, 1e-4)
    assert info == 0 and np.array_equal(alpha, d["alpha"])
    assert np.array_equal(oracle.get_prediction(d["x1"].T, d["x2"].T, d["n2"], d["alpha"], 200.0),
                          d["pred"])
    alpha_d, info = oracle.get_alpha_from_distance(d["dmat"], d["n1"], d["y1"], 200.0, 1e-4)
    assert info == 0 and np.array_equal(alpha_d, d["alpha_d"])


def test_get_alpha_prediction_closed_form():
    """alpha = (K + lambda I)^-1 y per molecule, repeated over its atoms,
    K the molecular Laplacian kernel; predictions = segment sums of the
    atom kernel times alpha.  numpy's solve differs from dpotri + dsymv by
    a few ulps times the condition number."""
    d = g("atomic.npz")
    x1, x2, n1, n2, y = d["x1"], d["x2"], d["n1"], d["n2"], d["y1"]
    K = _laplacian_mol_kernel(x1, n1, 200.0)
    ref = np.repeat(np.linalg.solve(K + 1e-4 * np.eye(len(n1)), y), n1)
    alpha, _ = oracle.get_alpha(x1.T, n1, y, 200.0, 1e-4)
    cond = np.linalg.cond(K + 1e-4 * np.eye(len(n1)))
    assert np.max(np.abs(alpha - ref)) <= 1e-15 * cond * np.max(np.abs(ref)) * 10
    E = np.exp(-np.abs(x2[:, None, :] - x1[None, :, :]).sum(-1) / 200.0)
    pred = np.add.reduceat(E @ alpha, np.concatenate([[0], np.cumsum(n2)[:-1]]))
    assert np.allclose(oracle.get_prediction(x1.T, x2.T, n2, alpha, 200.0), pred, rtol=1e-13, atol=1e-14)
    # the distance forms on the L1 distance matrix are the same model
    D2 = np.abs(x1[:, None, :] - x2[None, :, :]).sum(-1)
    assert np.allclose(oracle.get_prediction_from_distance(D2, n2, alpha, 200.0), pred, rtol=1e-13,
                       atol=1e-14)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
def test_atomic_path_vs_reference_build():
    """Fresh random cases against the reference Fortran, bit for bit
    (fget_prediction_from_distance excluded: the reference accumulates
    into an uninitialised array, fkernels.f90:680-686)."""
    rng = np.random.default_rng(23)
    ms = synth.generate(15, seed=24, nmin=1, nmax=20)
    for i in range(ms.nmol):
        Z, R = ms.molecule(i)
        for mode in (oracle.UNSORTED, oracle.LOCAL, oracle.ATOMIC):
            assert np.array_equal(oracle.coulomb_form(mode, Z, R, 20), oracle.ref_coulomb_form(mode, Z, R, 20))
    nm1, nm2, amax, d = 6, 5, 7, 13
    q1, q2 = rng.standard_normal((nm1, amax, d)), rng.standard_normal((nm2, amax, d))
    n1, n2 = rng.integers(0, amax + 1, nm1), rng.integers(1, amax + 1, nm2)
    for kind in (0, 1):
        assert np.array_equal(oracle.vector_kernels(kind, q1, q2, n1, n2, [0.4, 5.0]),
                              oracle.ref_vector_kernels(kind, q1, q2, n1, n2, [0.4, 5.0]))
    n1 = np.maximum(n1, 1)
    Q = rng.standard_normal((d, int(n1.sum())))
    Q2 = rng.standard_normal((d, int(n2.sum())))
    y = rng.standard_normal(nm1)
    a, info = oracle.get_alpha(Q, n1, y, 6.0, 1e-2)
    assert info == 0 and np.array_equal(a, oracle.ref_get_alpha(Q, n1, y, 6.0, 1e-2))
    assert np.array_equal(oracle.get_prediction(Q, Q2, n2, a, 6.0), oracle.ref_get_prediction(Q, Q2, n2, a, 6.0))
    P = rng.standard_normal((int(n1.sum()), 3))
    D = np.sqrt(((P[:, None] - P[None]) ** 2).sum(-1))
    a, info = oracle.get_alpha_from_distance(D, n1, y, 1.0, 1e-2)
    assert info == 0 and np.array_equal(a, oracle.ref_get_alpha_from_distance(D, n1, y, 1.0, 1e-2))


# ------------------------------------------------------------------- ARAS

def test_aras_golden_bitwise():
    d = g("aras.npz")
    X, n, sig = d["X"], d["n"], d["sigmas"]
    from fml_amd.aras import ARAS
    ar = ARAS(maxMolSize=11, maxAts=11)
    Xh = np.array([ar.describe(m.coordinates, m.nuclear_charges) for m in xyz_mols()])
    assert np.array_equal(Xh, X)
    assert np.array_equal(oracle.kernels_aras(X[:5], X[5:], n[:5], n[5:], sig), d["kernels"])
    assert np.array_equal(oracle.kernels_aras(X, None, n, None, sig), d["symmetric"])
    assert np.array_equal(oracle.kernels_aras(X[:5], X[5:], n[:5], n[5:], sig, t_width=2.0, d_width=0.3,
                                              cut=4.0, order=2), d["kernels_o2"])


def test_aras_closed_forms():
    """One-atom molecules: l2 = pd(p,p) + pd(q,q) - 2 pd(p,q)^2, so K = 1 for
    equal elements and exp(-(2 - 2 pd^2)/sigma^2) otherwise.  A diatomic
    A-B against itself: each atom's self distance is 1 + ksi^2 with
    ksi = qA qB / d^6 (no angular term without two neighbours), the A-B
    distance pd(A,B) (1 + ksi^2), so the two cross pairs have
    l2 = 2 (1 + ksi^2)(1 - pd^2) and K = 2 + 2 exp(-l2 / sigma^2)."""
    from fml_amd.aras import ARAS
    ar = ARAS(maxMolSize=2, maxAts=2)
    Zs = [1, 6, 8]
    X = np.array([ar.describe(np.zeros((1, 3)), [z]) for z in Zs])
    n = np.ones(3, dtype=np.int32)
    pd = oracle.aras_pd()
    K = oracle.kernels_aras(X, None, n, None, [1.5])[0]
    for i, zi in enumerate(Zs):
        for j, zj in enumerate(Zs):
            p = pd[zi - 1, zj - 1]
            assert K[i, j] == pytest.approx(math.exp(-(2.0 - 2.0 * p * p) / 2.25), rel=1e-14)
    co = np.array([[0.0, 0.0, 0.0], [1.1, 0.0, 0.0]])
    for za, zb, sigma in ((6, 8, 100.0), (1, 1, 0.7)):
        Xd = ar.describe(co, [za, zb])[None]
        Kd = oracle.kernels_aras(Xd, None, np.array([2], dtype=np.int32), None, [sigma])[0, 0, 0]
        ksi = za * zb / 1.1 ** 6
        p = pd[za - 1, zb - 1]
        expect = 2.0 + 2.0 * math.exp(-2.0 * (1.0 + ksi * ksi) * (1.0 - p * p) / sigma ** 2)
        assert Kd == pytest.approx(expect, rel=1e-12)
    # global against itself equals the symmetric form for i >= j, bitwise
    ms = synth.generate(6, seed=25, nmin=2, nmax=7)
    ar = ARAS(maxMolSize=7, maxAts=7)
    Xs = np.array([ar.describe(ms.molecule(i)[1], ms.molecule(i)[0]) for i in range(ms.nmol)])
    g_ = oracle.kernels_aras(Xs, Xs, ms.natoms, ms.natoms, [1.0, 3.0])
    s_ = oracle.kernels_aras(Xs, None, ms.natoms, None, [1.0, 3.0])
    low = np.tril(np.ones((ms.nmol, ms.nmol), dtype=bool))
    assert np.array_equal(g_[:, low], s_[:, low])


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
def test_aras_vs_reference_build():
    from fml_amd.aras import ARAS
    ms = synth.generate(8, seed=26, nmin=3, nmax=9)
    ar = ARAS(maxMolSize=9, maxAts=9, cut=4.0)
    X = np.array([ar.describe(ms.molecule(i)[1], ms.molecule(i)[0]) for i in range(ms.nmol)])
    kw = dict(t_width=1.5, d_width=0.25, cut=4.0, order=3, r_width=0.8, c_width=0.6, scale=0.3)
    assert np.array_equal(oracle.kernels_aras(X[:3], X[3:], ms.natoms[:3], ms.natoms[3:], [0.8, 4.0], **kw),
                          oracle.ref_kernels_aras(X[:3], X[3:], ms.natoms[:3], ms.natoms[3:], [0.8, 4.0], **kw))
    assert np.array_equal(oracle.kernels_aras(X, None, ms.natoms, None, [2.0], **kw),
                          oracle.ref_kernels_aras(X, None, ms.natoms, None, [2.0], **kw))
