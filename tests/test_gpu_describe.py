"""GPU ARAD descriptors (fmlh_arad_describe) against the oracle's loop
restatement of ARAD.describe (oracle/pyref.py, fml/arad.py:70-145), the
committed descriptors of the reference's xyz molecules, and the host class.

Tolerances (stated per check):
  * channel 0 (distances), channels 1-2 (PTP types), padding, atom counts
    and the per-atom type array: bit-exact (same operation order,
    contraction off);
  * channels 3-4 (weighted cos/sin sums divided by the weight sum, each
    value in [-1, 1]): 1e-14 absolute -- the device sin/cos/acos round
    differently from the host's; NaN (arccos of a cosine rounded past 1,
    not clipped in the reference) in the same places;
  * ARAD kernels computed from device descriptors against kernels from
    host descriptors: 1e-12 relative.
"""
import os

import numpy as np
import pytest

from fml_amd import synth
from fml_amd.arad import ARAD, generate_arad_descriptors
from fml_amd.fml import Molecule
from fml_amd.kernels.kernels import arad_types
from oracle.pyref import arad_describe

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _pack(mols):
    """charges, coords, offsets of a list of (Z, R) molecules."""
    z = np.concatenate([np.asarray(m[0], dtype=np.float64) for m in mols]) if mols else np.zeros(0)
    r = np.concatenate([np.asarray(m[1], dtype=np.float64).reshape(-1, 3) for m in mols]) \
        if mols else np.zeros((0, 3))
    off = np.concatenate([[0], np.cumsum([len(m[0]) for m in mols])]).astype(np.int32)
    return z, r, off


def _check(q, ref):
    """Channels 0-2 bit-exact, 3-4 within 1e-14 with NaNs in the same places."""
    q = np.asarray(q)
    ref = np.asarray(ref)
    assert q.shape == ref.shape
    assert np.array_equal(q[..., 0:3, :], ref[..., 0:3, :])
    a, b = q[..., 3:5, :], ref[..., 3:5, :]
    assert np.array_equal(np.isnan(a), np.isnan(b))
    fin = ~np.isnan(b)
    assert np.max(np.abs(a[fin] - b[fin]), initial=0.0) <= 1e-14


def _xyz():
    out = []
    for f in sorted(os.listdir(os.path.join(GOLD, "xyz"))):
        m = Molecule()
        m.read_xyz(os.path.join(GOLD, "xyz", f))
        out.append((m.nuclear_charges, m.coordinates))
    return out


def test_describe_reference_molecules(gpu):
    """The reference's xyz molecules at size 30 against the committed
    descriptors and the oracle, with types and counts as arad_types."""
    mols = _xyz()
    z, r, off = _pack(mols)
    q, zt, n = generate_arad_descriptors(z, r, off, amax=30, cut=5.0)
    gold = np.load(os.path.join(GOLD, "arad.npz"))
    _check(q, gold["X"])
    assert np.array_equal(zt, gold["z"]) and np.array_equal(n, gold["n"])
    for m, (zc, rc) in enumerate(mols):
        _check(q[m], arad_describe(rc, zc, size=30, cut=5.0))


def test_describe_synthetic_vs_oracle_and_host(gpu):
    ms = synth.generate(40, seed=11, nmin=1, nmax=23)
    q, zt, n = generate_arad_descriptors(ms.charges, ms.coords, ms.offsets, amax=23)
    ar = ARAD(maxMolSize=23, maxAts=23)
    host = np.array([ar.describe(ms.molecule(m)[1], ms.molecule(m)[0]) for m in range(ms.nmol)])
    _check(q, host)
    for m in range(0, ms.nmol, 5):
        Z, R = ms.molecule(m)
        _check(q[m], arad_describe(R, Z, size=23, cut=5.0))
    zr, nr = arad_types([ms.molecule(m)[0] for m in range(ms.nmol)], 23)
    assert np.array_equal(zt, zr) and np.array_equal(n, nr)


@pytest.mark.parametrize("cut", [1.2, 3.0])
def test_describe_short_cut(gpu, cut):
    """A cut inside the molecule: neighbour lists end early; at cut 3.0
    both fewer and more than 8 kept neighbours occur (the two branches of
    the pairwise summation order)."""
    ms = synth.generate(24, seed=12, nmin=2, nmax=29)
    q, _, _ = generate_arad_descriptors(ms.charges, ms.coords, ms.offsets, amax=29, cut=cut)
    ar = ARAD(maxMolSize=29, maxAts=29, cut=cut)
    host = np.array([ar.describe(ms.molecule(m)[1], ms.molecule(m)[0]) for m in range(ms.nmol)])
    _check(q, host)
    kept = np.sum(host[:, :, 0, :] < 1e99, axis=2)
    assert kept[kept > 0].min() < 8
    if cut == 3.0:
        assert kept.max() > 8


def test_describe_64_atoms(gpu):
    """The largest molecule of the device form: 64 atoms at amax 64."""
    ms = synth.generate(3, seed=13, nmin=60, nmax=64, min_dist=0.8)
    q, _, n = generate_arad_descriptors(ms.charges, ms.coords, ms.offsets, amax=64)
    ar = ARAD(maxMolSize=64, maxAts=64)
    host = np.array([ar.describe(ms.molecule(m)[1], ms.molecule(m)[0]) for m in range(ms.nmol)])
    _check(q, host)
    assert np.array_equal(n, ms.natoms)


def test_describe_degenerate_geometry(gpu):
    """Linear and coincident-direction triples: the eps branches of
    getAngle (angle 0) and arccos at +-1."""
    lin = ([6, 6, 1, 1], [[0, 0, 0], [1.2, 0, 0], [-1.06, 0, 0], [2.26, 0, 0]])
    tri = ([8, 1, 1], [[0, 0, 0], [0.96, 0, 0], [-0.24, 0.93, 0]])
    one = ([1], [[0.3, -0.2, 0.1]])
    empty_mol = ([], np.zeros((0, 3)))
    mols = [lin, tri, one, empty_mol, lin]
    z, r, off = _pack(mols)
    q, zt, n = generate_arad_descriptors(z, r, off, amax=6)
    ar = ARAD(maxMolSize=6, maxAts=6)
    for m, (zc, rc) in enumerate(mols):
        _check(q[m], ar.describe(np.asarray(rc, dtype=np.float64).reshape(-1, 3), zc))
        _check(q[m], arad_describe(rc, zc, size=6))
    assert list(n) == [4, 3, 1, 0, 4]
    assert np.all(q[3, :, 0, :] == 1e100) and np.all(q[3, :, 1:, :] == 0.0)
    assert np.array_equal(q[0], q[4])


def test_describe_batch_invariance(gpu):
    """A molecule's block is the same alone and in any batch."""
    ms = synth.generate(30, seed=14, nmin=3, nmax=23)
    q, _, _ = generate_arad_descriptors(ms.charges, ms.coords, ms.offsets, amax=23)
    for m in (0, 17, 29):
        Z, R = ms.molecule(m)
        z1, r1, off1 = _pack([(Z, R)])
        q1, _, _ = generate_arad_descriptors(z1, r1, off1, amax=23)
        assert np.array_equal(q1[0], q[m])


def test_describe_bad_molecules(gpu):
    """Too many atoms or a charge that is not a PTP key: ValueError, or with
    return_status a flagged all-padding block; the other molecules are
    unaffected."""
    good = synth.generate(4, seed=15, nmin=3, nmax=10)
    mols = [good.molecule(m) for m in range(4)]
    big = synth.generate(1, seed=16, nmin=12, nmax=12).molecule(0)
    odd = (np.array([6.0, 6.5, 1.0]), mols[0][1][:3])
    unknown = (np.array([6.0, 0.0]), mols[1][1][:2])
    heavy = (np.array([200.0, 1.0]), mols[2][1][:2])
    batch = [mols[0], big, mols[1], odd, mols[2], unknown, heavy, mols[3]]
    z, r, off = _pack(batch)
    with pytest.raises(ValueError):
        generate_arad_descriptors(z, r, off, amax=10)
    q, zt, n, st = generate_arad_descriptors(z, r, off, amax=10, return_status=True)
    assert list(st) == [0, 1, 0, 1, 0, 1, 1, 0]
    zg, rg, offg = _pack(mols)
    qg, ztg, ng = generate_arad_descriptors(zg, rg, offg, amax=10)
    for k, m in enumerate((0, 2, 4, 7)):
        assert np.array_equal(q[m], qg[k]) and np.array_equal(zt[m], ztg[k]) and n[m] == ng[k]
    for m in (1, 3, 5, 6):
        assert n[m] == 0 and np.all(q[m, :, 0, :] == 1e100) and np.all(q[m, :, 1:, :] == 0.0)
        assert np.all(zt[m] == 0.0)


def test_describe_no_molecules(gpu):
    q, zt, n = generate_arad_descriptors(np.zeros(0), np.zeros((0, 3)), np.array([0]), amax=5)
    assert q.shape == (0, 5, 5, 5) and zt.shape == (0, 5, 2) and n.shape == (0,)


def test_kernels_from_device_descriptors(gpu):
    """Global ARAD kernels from device descriptors kept in HBM against the
    same kernels from host descriptors: 1e-12 relative."""
    from fml_amd.kernels import fget_kernels_arad
    ms1 = synth.generate(20, seed=17, nmin=3, nmax=23)
    ms2 = synth.generate(14, seed=18, nmin=3, nmax=23)
    d1 = generate_arad_descriptors(ms1.charges, ms1.coords, ms1.offsets, amax=23, on_device=True)
    d2 = generate_arad_descriptors(ms2.charges, ms2.coords, ms2.offsets, amax=23, on_device=True)
    sig = [0.5, 1.0, 10.0]
    kd = fget_kernels_arad(d1[0], d2[0], d1[1], d2[1], d1[2], d2[2], sig)
    kd = kd.cpu().numpy() if hasattr(kd, "cpu") else np.asarray(kd)
    ar = ARAD(maxMolSize=23, maxAts=23)
    h1 = np.array([ar.describe(ms1.molecule(m)[1], ms1.molecule(m)[0]) for m in range(ms1.nmol)])
    h2 = np.array([ar.describe(ms2.molecule(m)[1], ms2.molecule(m)[0]) for m in range(ms2.nmol)])
    z1, n1 = arad_types([ms1.molecule(m)[0] for m in range(ms1.nmol)], 23)
    z2, n2 = arad_types([ms2.molecule(m)[0] for m in range(ms2.nmol)], 23)
    kh = fget_kernels_arad(h1, h2, z1, z2, n1, n2, sig)
    assert kd.shape == kh.shape
    assert np.max(np.abs(kd - kh) / np.abs(kh)) <= 1e-12


def test_describe_and_kernel_in_one_graph(gpu):
    """Prepared describe + kernel calls captured in one HIP graph: every
    replay, with new coordinates between replays, equals eager calls."""
    from fml_amd.batch import AradDescribeCall, AradKernelCall
    from fml_amd.graphs import Graph
    ms = synth.generate(24, seed=19, nmin=3, nmax=23)
    dc = AradDescribeCall(ms.charges, ms.coords, ms.offsets, amax=23)
    dc.run()
    kc = AradKernelCall(dc.q, dc.zt, dc.n, sigmas=(1.0, 4.0), mode="symmetric")
    assert kc.q1.data_ptr() == dc.q.data_ptr()
    kc.run()
    gpu.cuda.synchronize()
    eager = kc.result().cpu().numpy().copy()

    def pipeline():
        dc.run()
        kc.run()

    g = Graph.capture(pipeline, keep=[dc, kc])
    for rep in range(3):
        scale = 1.0 + 0.01 * rep
        dc.set_coords(ms.coords * scale)
        kc.out.fill_(-1.0)
        gpu.cuda.synchronize()
        g.launch()
        gpu.cuda.synchronize()
        got = kc.result().cpu().numpy().copy()
        pipeline()
        gpu.cuda.synchronize()
        assert np.array_equal(got, kc.result().cpu().numpy()), rep
        if rep == 0:
            assert np.array_equal(got, eager)
        _ = (gpu.ones(1000, device="cuda") * 2).sum().item()  # eager work between replays
    g.destroy()
