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
