"""Batching form on the GPU: the device-planned size-sorted tiles under
every tile shape the autotuner may pick, the autotuner itself, and HIP-graph
replays of Coulomb-matrix + ARAD-kernel batches.

Tolerances: tile shapes only change which lanes compute which atom pairs;
each molecule pair's sum runs over its atom pairs in the reference's order
whatever the shape, so results are compared BITWISE across shapes and
between eager calls and graph replays.  Against the oracle: 1e-12 relative
(sigma >= 1), as in test_gpu_parity.py.
"""
import ctypes

import numpy as np
import pytest

import oracle
from fml_amd import synth
from fml_amd.arad import ARAD
from fml_amd.kernels.kernels import arad_types

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    m = np.abs(b) > 1e-300
    return float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m])))


def _desc(ms, amax):
    ar = ARAD(maxMolSize=amax, maxAts=amax)
    X = np.array([ar.describe(ms.molecule(m)[1], ms.molecule(m)[0]) for m in range(ms.nmol)])
    z, n = arad_types([ms.molecule(m)[0] for m in range(ms.nmol)], amax)
    return X, z, n


@pytest.fixture(scope="module")
def uneven():
    ms1 = synth.generate(40, seed=31, nmin=1, nmax=29, dist="geometric")
    ms2 = synth.generate(52, seed=32, nmin=1, nmax=29, dist="geometric")
    return _desc(ms1, 29), _desc(ms2, 29)


@pytest.mark.parametrize("mode", ["global", "symmetric", "kernel"])
def test_tile_shapes_bitwise_equal(gpu, uneven, mode):
    import fml_amd._native as nat
    from fml_amd.batch import AradKernelCall
    (X1, z1, n1), (X2, z2, n2) = uneven
    sig = [1.0, 4.0, 9.0]
    ref = None
    shapes = [(p, w, g, o) for p in (64, 128, 256, 384, 512, 768, 1024) for w in (4, 5)
              for g in (1, 16) for o in (0, 1)]
    for pairs, waves, gpc, order in shapes:
        if True:
            if True:
                t = nat.AradTiling(pairs, waves, gpc, order)
                c = AradKernelCall(X1, z1, n1, X2, z2, n2, sig, mode=mode, cut_distance=5.0, tiling=t)
                c.run()
                r = c.result().cpu().numpy()
                if ref is None:
                    ref = r
                    if mode == "global":
                        ko = oracle.kernels_arad(X1, X2, z1, z2, n1, n2, sig, 0.2, 5.0)
                    elif mode == "symmetric":
                        ko = oracle.symmetric_kernels_arad(X1, z1, n1, sig, 0.2, 5.0)
                    else:
                        ko = oracle.kernel_arad(X1, z1, n1, 1.0)
                    assert _rel(r, ko) < 1e-12, mode
                else:
                    assert np.array_equal(r, ref), (mode, pairs, waves, gpc, order)


def test_autotune_picks_a_candidate_and_keeps_results(gpu, uneven, tmp_path):
    import fml_amd._native as nat
    from fml_amd import tuning
    from fml_amd.batch import AradKernelCall
    (X1, z1, n1), (X2, z2, n2) = uneven
    sig = [1.0, 4.0]
    c = AradKernelCall(X1, z1, n1, X2, z2, n2, sig, cut_distance=5.0)
    c.run()
    before = c.result().cpu().numpy()
    best, ms = c.autotune(sample_rows=16, reps=2)
    assert best.pairs_per_tile in (128, 256, 384, 512, 768, 1024)
    assert best.waves_per_simd in (4, 5) and best.pair_order in (0, 1) and ms > 0
    assert np.array_equal(c.result().cpu().numpy(), before)
    # the library now uses the tuned tiling for this key by default
    assert tuning.current(nat.ARAD_GLOBAL, 29, 0, 2).as_tuple() == best.as_tuple()
    path = str(tmp_path / "tiling.json")
    tuning.save(path)
    nat.call("fmlh_arad_tiling_set", nat.ARAD_GLOBAL, 29, 0, 2, ctypes.byref(nat.AradTiling(256, 0, 16, 0)))
    assert tuning.load(path) >= 1
    assert tuning.current(nat.ARAD_GLOBAL, 29, 0, 2).as_tuple() == best.as_tuple()


def test_graph_replays_coulomb_and_arad(gpu, uneven):
    """One HIP graph holding a Coulomb-matrix batch and a global ARAD kernel;
    replayed several times with eager work and synchronisation between the
    replays, each replay checked against the eager result; new inputs copied
    in between replays give the new results."""
    torch = gpu
    from fml_amd.batch import AradKernelCall, CoulombCall
    from fml_amd.graphs import Graph
    (X1, z1, n1), (X2, z2, n2) = uneven
    ms = synth.generate(64, seed=33, nmin=3, nmax=29)
    cmc = CoulombCall(ms.charges, ms.coords, ms.offsets, nmax=29)
    ac = AradKernelCall(X1, z1, n1, X2, z2, n2, [1.0, 4.0], cut_distance=5.0)
    cmc.run()
    ac.run()
    torch.cuda.synchronize()
    cm_eager = cmc.cm.cpu().numpy()
    k_eager = ac.result().cpu().numpy()
    assert np.array_equal(cm_eager, oracle.coulomb_matrices(ms.charges, ms.coords, ms.offsets, 29)[0])

    def batch():
        cmc.run()
        ac.run()

    g = Graph.capture(batch, keep=[cmc, ac])
    for rep in range(4):
        cmc.cm.fill_(-1.0)
        ac.out.fill_(-1.0)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(cmc.cm.cpu().numpy(), cm_eager), rep
        assert np.array_equal(ac.result().cpu().numpy(), k_eager), rep
        # eager work between replays
        _ = (torch.ones(1000, device="cuda") * 2).sum().item()
    # new coordinates and descriptors for the same molecules (same sizes)
    rng = np.random.default_rng(34)
    coords2 = ms.coords * 1.05 + rng.normal(0, 1e-3, ms.coords.shape)
    cmc.set_coords(coords2)
    ms1 = synth.generate(40, seed=31, nmin=1, nmax=29, dist="geometric")
    ar = ARAD(maxMolSize=29, maxAts=29)
    X1b = np.array([ar.describe(ms1.molecule(m)[1] * 0.97, ms1.molecule(m)[0]) for m in range(ms1.nmol)])
    assert X1b.shape == X1.shape and not np.array_equal(X1b, X1)
    ac.set_descriptors(q1=X1b)
    g.launch()
    torch.cuda.synchronize()
    assert np.array_equal(cmc.cm.cpu().numpy(), oracle.coulomb_matrices(ms.charges, coords2, ms.offsets, 29)[0])
    ko = oracle.kernels_arad(X1b, X2, z1, z2, n1, n2, [1.0, 4.0], 0.2, 5.0)
    assert _rel(ac.result().cpu().numpy(), ko) < 1e-12
    g.destroy()


def test_krr_call_graph_replays(gpu):
    """The prepared FP64 KRR chain (Coulomb matrices -> kernel -> K + lambda I
    -> cho_solve -> test kernel -> predictions): eager run bitwise equal to
    the KRR class on the same molecules, predictions within 1e-6 of the
    LAPACK solution's, and HIP-graph replays (with eager work between them,
    outputs cleared before each) bitwise equal to the eager run, also after
    new targets are copied in."""
    torch = gpu
    from fml_amd.batch import KRRCall
    from fml_amd.graphs import Graph
    from fml_amd.krr import KRR
    from fml_amd.representations import generate_coulomb_matrices
    ms = synth.generate(400, seed=35, nmin=5, nmax=23)
    n = 320
    y = ms.energies
    call = KRRCall(ms.charges, ms.coords, ms.offsets, n, y[:n], 700.0, 1e-8, nmax=23)
    call.run()
    torch.cuda.synchronize()
    assert int(call.info.item()) == 0
    pred, alpha = call.pred.cpu().numpy(), call.alpha.cpu().numpy()
    X = generate_coulomb_matrices(ms.charges, ms.coords, ms.offsets, nmax=23)
    m = KRR(700.0, 1e-8).fit(X[:n], y[:n])
    assert np.array_equal(m.alpha.cpu().numpy(), alpha)
    assert np.array_equal(m.predict(X[n:]).cpu().numpy(), pred)
    Xo = oracle.coulomb_matrices(ms.charges, ms.coords, ms.offsets, 23)[0]
    K = oracle.gaussian_kernel(Xo[:n].T, Xo[:n].T, 700.0)
    K[np.diag_indices_from(K)] += 1e-8
    ao, info = oracle.cho_solve(K, y[:n])
    assert info == 0
    po = oracle.gaussian_kernel(Xo[:n].T, Xo[n:].T, 700.0).T @ ao
    assert np.max(np.abs(pred - po)) <= 1e-6 * np.max(np.abs(po))

    g = Graph.capture(call.run, keep=[call])
    for rep in range(3):
        call.alpha.fill_(-1.0)
        call.pred.fill_(-1.0)
        call.info.fill_(-1)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert int(call.info.item()) == 0, rep
        assert np.array_equal(call.alpha.cpu().numpy(), alpha), rep
        assert np.array_equal(call.pred.cpu().numpy(), pred), rep
        _ = (torch.ones(1000, device="cuda") * 2).sum().item()
    y2 = y[:n] + np.random.default_rng(36).normal(0.0, 0.5, n)
    call.set_targets(y2)
    g.launch()
    torch.cuda.synchronize()
    pred_g = call.pred.cpu().numpy()
    call.run()
    torch.cuda.synchronize()
    assert np.array_equal(pred_g, call.pred.cpu().numpy())
    assert not np.array_equal(pred_g, pred)
    g.destroy()
