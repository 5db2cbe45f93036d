"""Host-side logic that needs no GPU: XYZ reading, tables, the synthetic
generator, argument checks of the drop-in API."""
import glob
import os

import numpy as np
import pytest

from fml_amd import synth
from fml_amd.fml import Molecule
from fml_amd.kernels.kernels import arad_types
from fml_amd.ptp import PTP

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_read_xyz():
    m = Molecule()
    m.read_xyz(os.path.join(GOLD, "xyz", "04.xyz"))
    assert m.natoms == 4
    assert m.atomtypes == ["C", "C", "H", "H"]
    assert m.nuclear_charges == [6, 6, 1, 1]
    assert m.coordinates.shape == (4, 3)
    assert m.coordinates[0, 0] == 0.899231
    for f in glob.glob(os.path.join(GOLD, "xyz", "*.xyz")):
        m = Molecule()
        m.read_xyz(f)
        assert m.natoms == len(m.nuclear_charges) == m.coordinates.shape[0]


def test_arad_types():
    z, n = arad_types([[1, 6], [8, 7, 16]], 4)
    assert list(n) == [2, 3]
    assert z.shape == (2, 4, 2)
    assert list(z[1, 2]) == PTP[16] and list(z[0, 3]) == [0, 0]


def test_synth_generator():
    ms = synth.generate(50, seed=0, nmin=5, nmax=23)
    assert ms.nmol == 50 and ms.offsets[-1] == ms.charges.shape[0]
    assert ms.natoms.min() >= 5 and ms.natoms.max() <= 23
    assert set(np.unique(ms.charges)) <= {1, 6, 7, 8, 16}
    for i in range(ms.nmol):
        z, r = ms.molecule(i)
        edge = 1.2 * len(z) ** (1 / 3)
        assert r.min() >= 0 and r.max() <= edge
        d = np.sqrt(((r[:, None] - r[None]) ** 2).sum(-1)) + np.eye(len(z)) * 10
        assert d.min() >= 0.9
    # deterministic
    ms2 = synth.generate(50, seed=0, nmin=5, nmax=23)
    assert np.array_equal(ms.coords, ms2.coords) and np.array_equal(ms.energies, ms2.energies)
    g = synth.generate(300, seed=1, nmin=3, nmax=29, dist="geometric")
    assert 10 < g.natoms.mean() < 18


def test_cho_solve_shape_errors():
    from fml_amd.math import cho_solve
    with pytest.raises(ValueError):
        cho_solve(np.ones((3, 4)), np.ones(3))
    with pytest.raises(ValueError):
        cho_solve(np.eye(3), np.ones(4))


def test_arad_shape_asserts():
    from fml_amd.kernels import get_atomic_kernels_arad
    X = np.zeros((2, 5, 5, 4))
    with pytest.raises(AssertionError):
        get_atomic_kernels_arad(X, X, [[1], [1]], [[1], [1]], [1.0])


def test_ptp_table_matches_dictionary():
    from fml_amd.arad import ptp_table
    t = ptp_table()
    assert t.shape == (max(PTP) + 1, 2)
    for z in range(t.shape[0]):
        if z in PTP:
            assert list(t[z]) == PTP[z] and t[z, 0] > 0
        else:
            assert list(t[z]) == [0.0, 0.0]


def test_check_offsets():
    from fml_amd.device import check_offsets
    check_offsets(np.array([0, 3, 3, 7]), 7)
    check_offsets(np.array([0]), 0)
    for bad, n in (([0, 4, 2], 5), ([-1, 2], 5), ([0, 6], 5), ([], 0)):
        with pytest.raises(ValueError):
            check_offsets(np.array(bad, dtype=np.int64), n)


def test_describe_argument_checks():
    """Rejected on the host before any device work."""
    from fml_amd.arad import generate_arad_descriptors
    z, r, off = np.array([6.0, 1.0]), np.zeros((2, 3)), np.array([0, 2])
    for kw in ({"amax": 0}, {"amax": 65}, {"cut": 0.0}, {"cut": -1.0}):
        with pytest.raises(ValueError):
            generate_arad_descriptors(z, r, off, **kw)


def test_plan_blocks_greedy():
    """Molecules in order, whole, at most 128 atoms per block, and a block
    closes only when the next molecule would not fit."""
    from fml_amd.kernels.atomic import plan_blocks
    rng = np.random.default_rng(3)
    for trial in range(20):
        counts = rng.integers(0 if trial % 3 == 0 else 1, 65, rng.integers(1, 400))
        b = plan_blocks(counts)
        off = np.concatenate([[0], np.cumsum(counts)])
        assert b[0] == 0 and b[-1] == len(counts) and np.all(np.diff(b) >= 1)
        sizes = off[b[1:]] - off[b[:-1]]
        assert np.all(sizes <= 128)
        for k in range(len(b) - 2):  # every closed block could not take the next molecule
            assert sizes[k] + counts[b[k + 1]] > 128
    assert list(plan_blocks(np.zeros(5, dtype=np.int64))) == [0, 5]


def test_gen_pd_matches_oracle_table():
    """kernels.py gen_pd over the product's PTP table equals the oracle's,
    built from its own restatement of the table."""
    import oracle
    from fml_amd.kernels import gen_pd, periodic_distance
    assert np.array_equal(gen_pd(103, 1.0, 0.5), oracle.aras_pd(103, 1.0, 0.5))
    assert periodic_distance(6, 6, 1.0, 0.5) == 1.0
    assert periodic_distance(6, 8, 1.0, 0.5) == (1.0 * 0.25) / ((1.0 + 0.0) * (0.25 + 4.0))


def test_aras_argument_checks():
    from fml_amd.kernels import get_atomic_kernels_aras
    X = np.zeros((2, 3, 6, 3))
    with pytest.raises(ValueError):
        get_atomic_kernels_aras(X, np.zeros((2, 4, 7, 4)), [[1, 1], [6]], [[1], [8]], [1.0])


def test_aras_block_plan():
    """Set-1 blocks for the ARAS tiles: at most 64 atoms, every molecule in
    exactly one block, molecules grouped by element set, then by size."""
    from fml_amd.kernels.aras_kernels import aras_block_plan
    rng = np.random.default_rng(3)
    nm, amax = 200, 23
    n = rng.integers(1, amax + 1, size=nm)
    own = np.zeros((nm, amax))
    sets = [[1, 6], [1, 6, 8], [1, 6, 7, 8]]
    choice = rng.integers(0, 3, size=nm)
    for m in range(nm):
        own[m, :n[m]] = rng.choice(sets[choice[m]], size=n[m])
        own[m, 0] = 1
        own[m, n[m]:] = 99.0          # padding past the atom count is ignored
    for o in (None, own):
        blk, order = aras_block_plan(n, amax, o)
        order = np.arange(nm) if order is None else order
        assert sorted(order.tolist()) == list(range(nm))
        assert blk[0] == 0 and blk[-1] == nm and np.all(np.diff(blk) >= 1)
        for b in range(len(blk) - 1):
            assert n[order[blk[b]:blk[b + 1]]].sum() <= 64
    blk, order = aras_block_plan(n, amax, own)
    key = [frozenset(own[m, :n[m]].astype(int).tolist()) for m in order]
    runs = [k for i, k in enumerate(key) if i == 0 or k != key[i - 1]]
    assert len(runs) == len(set(key))                       # each element set is one run
    for k in set(key):
        sizes = [n[m] for m, kk in zip(order, key) if kk == k]
        assert sizes == sorted(sizes)


def test_fml_import_path():
    """The reference's own import lines resolve through the ``fml`` name
    (tests/test_arad.py:8-11, examples/example_arad.py:9-12)."""
    import fml
    from fml.kernels import get_atomic_kernels_arad
    from fml.math.distance import get_l2_distance_arad
    import fml_amd
    assert fml.kernels is fml_amd.kernels
    assert get_atomic_kernels_arad is fml_amd.kernels.get_atomic_kernels_arad
    assert get_l2_distance_arad is fml_amd.math.distance.get_l2_distance_arad
    from fml.math import cho_solve
    from fml.kernels import gaussian_kernel, laplacian_kernel
    assert cho_solve is fml_amd.math.cho_solve and gaussian_kernel and laplacian_kernel
    import fml.fml
    assert fml.fml.Molecule is fml.Molecule is fml_amd.Molecule
    from fml.arad import ARAD
    assert ARAD is fml_amd.ARAD
    m = fml.Molecule()
    m.read_xyz(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "xyz", "01.xyz"))
    assert m.natoms > 0


def test_exact_residual_helper():
    """The exactly rounded residual used to check refined KRR solutions
    equals y - (A + lam I) x evaluated in rational arithmetic, rounded once;
    the refined solution it yields solves a nearly singular system to
    working accuracy (its exact residual is below one unit in the last
    place of the terms)."""
    from fractions import Fraction
    from exact_residual import exact_refined_solution, exact_residual, two_product
    rng = np.random.default_rng(3)
    a, b = rng.normal(size=300) * 1e4, rng.normal(size=300)
    p, e = two_product(a, b)
    assert all(Fraction(p[i]) + Fraction(e[i]) == Fraction(a[i]) * Fraction(b[i]) for i in range(300))
    n = 60
    B = rng.normal(size=(n, n))
    A = B @ B.T / n
    y = rng.normal(size=n)
    x = np.linalg.solve(A + 1e-3 * np.eye(n), y)
    r = exact_residual(A, 1e-3, x, y)
    lam = Fraction(1e-3)
    for i in range(n):
        ex = Fraction(y[i]) - sum(Fraction(A[i, j]) * Fraction(x[j]) for j in range(n)) - lam * Fraction(x[i])
        assert r[i] == float(ex)
    # kappa ~ 1e12: plain LAPACK's solution is off by ~1e-5, the refined one is not
    w = np.linspace(1.0, 1e-12, n)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A2 = (Q * w) @ Q.T
    A2 = (A2 + A2.T) / 2
    xs = exact_refined_solution(A2, 0.0, y)
    rs = exact_residual(A2, 0.0, xs, y)
    assert np.max(np.abs(rs)) <= 1e-15 * np.max(np.abs(A2) @ np.abs(xs))


def test_predict_request_packing():
    """A predict request packs its molecules into contiguous charges,
    coordinates and atom counts on construction, and keeps the error of a
    request that cannot be served instead of raising."""
    from fml_amd.serving import PredictRequest
    ms = synth.generate(12, seed=21, nmin=3, nmax=29)
    ch = [ms.molecule(i)[0] for i in range(ms.nmol)]
    co = [ms.molecule(i)[1] for i in range(ms.nmol)]
    r = PredictRequest(ch, co, 29)
    assert r.error is None and r.nmol == 12
    assert r.z.dtype == np.float64 and np.array_equal(r.z, ms.charges)
    assert r.r.shape == (len(ms.charges), 3) and np.array_equal(r.r, ms.coords)
    assert np.array_equal(np.concatenate([[0], np.cumsum(r.counts)]), ms.offsets)
    # flat coordinate lists are read as rows of three
    assert np.array_equal(PredictRequest([[1, 6]], [[0, 0, 0, 1, 1, 1]]).r, [[0, 0, 0], [1, 1, 1]])
    assert PredictRequest([], []).nmol == 0
    big = int(r.counts.max())
    assert "nmax" in str(PredictRequest(ch, co, big - 1).error)
    assert "match" in str(PredictRequest(ch, co[:-1]).error)
    assert "match" in str(PredictRequest([[1, 6]], [[0, 0, 0]]).error)
