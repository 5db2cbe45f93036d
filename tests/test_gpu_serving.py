"""Serving loop on the GPU: continuous batching returns, for every request,
exactly (bitwise) what the request gets when served alone; fits run beside
predictions and are admitted by the HBM budget."""
import threading
import time

import numpy as np
import pytest

import oracle
from fml_amd import synth

pytestmark = pytest.mark.gpu


def _model(n_train=600):
    from fml_amd.krr import KRR
    ms = synth.generate(n_train, seed=4, nmin=3, nmax=29)
    X = oracle.coulomb_matrices(ms.charges, ms.coords, ms.offsets, 29)[0]
    m = KRR(1000.0, 1e-6).fit(X, ms.energies)
    return X, m


def test_serving_batch_invariance(gpu):
    from fml_amd.serving import KRRServer, latency_percentiles
    X, m = _model()
    srv = KRRServer(X, m.alpha, 1000.0, nmax=29, max_batch_molecules=512).start()
    rng = np.random.default_rng(7)
    reqs = []
    futs = []
    for k in range(40):
        nm = int(np.exp(rng.uniform(0, np.log(64))))
        ms = synth.generate(max(nm, 1), seed=100 + k, nmin=3, nmax=29)
        ch = [ms.molecule(i)[0] for i in range(ms.nmol)]
        co = [ms.molecule(i)[1] for i in range(ms.nmol)]
        reqs.append((ch, co))
        futs.append(srv.submit(ch, co))
        if k % 5 == 0:
            time.sleep(0.002)
    results = [f.result(timeout=120) for f in futs]
    srv.stop()
    assert len(srv.batches) < len(reqs)  # requests were batched together
    # each request alone
    from fml_amd.serving import PredictRequest
    for (ch, co), (y, lat, _) in zip(reqs, results):
        alone = srv.predict_batch([PredictRequest(ch, co)])[0]
        assert np.array_equal(alone, y)
        assert lat > 0
    # and against the oracle prediction (FP64 kernel + same alpha)
    ch, co = reqs[3]
    off = np.concatenate([[0], np.cumsum([len(c) for c in ch])]).astype(np.int32)
    Xq = oracle.coulomb_matrices(np.concatenate(ch), np.concatenate(co), off, 29)[0]
    po = oracle.gaussian_kernel(X.T, Xq.T, 1000.0).T @ m.alpha.cpu().numpy()
    assert np.max(np.abs(results[3][0] - po)) <= 1e-9 * np.max(np.abs(po))
    st = latency_percentiles([r[1] for r in results])
    assert st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_serving_fit_admission(gpu):
    from fml_amd.serving import KRRServer
    X, m = _model(300)
    srv = KRRServer(X, m.alpha, 1000.0, hbm_budget_bytes=1024).start()
    ms = synth.generate(200, seed=9, nmin=3, nmax=29)
    Xf = oracle.coulomb_matrices(ms.charges, ms.coords, ms.offsets, 29)[0]
    fut = srv.submit_fit(Xf, ms.energies, 1000.0, 1e-6)
    time.sleep(0.5)
    assert not fut.done()          # waiting for HBM
    # predictions keep flowing while the fit waits
    y, _, _ = srv.submit([ms.molecule(0)[0]], [ms.molecule(0)[1]]).result(timeout=60)
    assert y.shape == (1,)
    srv.hbm_budget = 1 << 40
    model = fut.result(timeout=120)
    srv.close()
    assert model.alpha.shape[0] == 200


def test_serving_rejects_oversize_molecule_only(gpu):
    """A request holding a molecule larger than nmax fails alone; requests
    batched with it are served, and a batch larger than the staging buffers
    grows them."""
    from fml_amd.krr import KRR
    from fml_amd.serving import KRRServer, PredictRequest
    ms = synth.generate(300, seed=12, nmin=3, nmax=23)
    X = oracle.coulomb_matrices(ms.charges, ms.coords, ms.offsets, 23)[0]
    m = KRR(700.0, 1e-8).fit(X, ms.energies)
    srv = KRRServer(X, m.alpha, 700.0, nmax=23)
    big = synth.generate(1, seed=13, nmin=24, nmax=24)
    ok = synth.generate(1500, seed=14, nmin=3, nmax=23)
    reqs = [PredictRequest([ok.molecule(i)[0] for i in range(700)], [ok.molecule(i)[1] for i in range(700)]),
            PredictRequest([big.molecule(0)[0]], [big.molecule(0)[1]]),
            PredictRequest([ok.molecule(i)[0] for i in range(700, 1500)],
                           [ok.molecule(i)[1] for i in range(700, 1500)])]
    out = srv.predict_batch(reqs)
    assert isinstance(out[1], ValueError)
    assert out[0].shape == (700,) and out[2].shape == (800,)
    # same values as the FP64 oracle kernel times alpha
    Xq = oracle.coulomb_matrices(ok.charges, ok.coords, ok.offsets, 23)[0]
    Kq = oracle.gaussian_kernel(X.T, Xq.T, 700.0)  # (ntrain, nq), the reference's Fortran K
    po = Kq.T @ m.alpha.cpu().numpy()
    got = np.concatenate([out[0], out[2]])
    assert np.max(np.abs(got - po)) <= 1e-9 * np.max(np.abs(po))
    # batch invariance across the buffer growth
    alone = srv.predict_batch([reqs[2]])[0]
    assert np.array_equal(alone, out[2])
    srv.close()


def test_predict_empty_training_set(gpu):
    """fmlh_krr_predict against an empty training set: zeros, through the C
    ABI with exactly the workspace the size query reports."""
    import torch
    import fml_amd._native as nat
    ms = synth.generate(5, seed=15, nmin=3, nmax=23)
    z = torch.tensor(ms.charges, dtype=torch.float64, device="cuda")
    r = torch.tensor(ms.coords, dtype=torch.float64, device="cuda")
    off = torch.tensor(ms.offsets, dtype=torch.int32, device="cuda")
    y = torch.full((5,), 7.0, dtype=torch.float64, device="cuda")
    status = torch.zeros(5, dtype=torch.int32, device="cuda")
    wsb = nat.load().fmlh_krr_predict_ws_bytes(5, 0, 23)
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    nat.call("fmlh_krr_predict", nat.ptr(z), nat.ptr(r), nat.ptr(off), 5, 23, None, 276, 0, None, 700.0,
             nat.ptr(y), nat.ptr(status), nat.ptr(ws), wsb, nat.stream_ptr())
    torch.cuda.synchronize()
    assert torch.equal(y, torch.zeros_like(y))
    assert int(status.sum().item()) == 0
