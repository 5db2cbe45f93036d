"""Continuous-batching KRR serving on one GPU.

No counterpart exists in the reference (fml has no serving code); this is
the serving form built from the hot-path pieces:

* the model (training Coulomb matrices X, coefficients alpha, sigma) stays
  resident in HBM;
* predict requests (lists of molecules) queue up; a scheduler thread drains
  everything pending into one batch, so requests that arrive while a batch
  runs are admitted into the next one -- continuous batching, not a queue
  served one request at a time;
* a batch is one pass of the GPU path: sorted Coulomb matrices for all its
  molecules (one workgroup per molecule), the Gaussian kernel block against
  the training set, and K^T alpha with one workgroup per test molecule;
* fit requests (new training sets) run on a second stream, admitted only
  when the free-HBM budget covers their kernel matrix and factor; predict
  batches keep flowing on the first stream meanwhile, so the two overlap on
  the device.  The fit stream's kernels are confined to all but
  ``reserved_cus`` compute units (a CU-masked stream with a hardware queue of
  its own), so a prediction never waits for a fit's workgroups to drain.

Batch invariance: every number a request gets back depends only on its own
molecules -- per-molecule Coulomb matrices, per-entry kernel values (MFMA dot
products over k in a fixed order, per-row norms, per-entry near-pair
recompute) and a fixed-order reduction per test molecule -- so a request's
result is bitwise the same alone and in any mix.
"""
from __future__ import annotations

import ctypes
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .device import to_device, torch
from .krr import KRR, kernel_matrix, predict_from_kernel
from .representations import generate_coulomb_matrices


@dataclass
class PredictRequest:
    charges: list            # per molecule: array of nuclear charges
    coords: list             # per molecule: (n, 3) array
    t_arrival: float = field(default_factory=time.perf_counter)
    future: Future = field(default_factory=Future)


@dataclass
class FitRequest:
    X: np.ndarray            # (n, d) descriptors (float32 or float64)
    y: np.ndarray
    sigma: float
    lam: float
    precision: str = "fp64"
    t_arrival: float = field(default_factory=time.perf_counter)
    future: Future = field(default_factory=Future)

    def hbm_bytes(self) -> int:
        n = int(self.X.shape[0])
        # FP64 kernel matrix (factored in place) + an FP32 copy for the mixed form
        return n * n * 8 + (n * n * 4 if self.precision == "mixed" else 0) + 16 * n * self.X.shape[1]


class KRRServer:
    """Serve predictions of a trained Gaussian-kernel KRR model.

    X_train: (n, d) sorted-Coulomb-matrix rows; alpha: (n,); nmax: the
    Coulomb-matrix size the rows were built with."""

    def __init__(self, X_train, alpha, sigma, nmax=29, max_batch_molecules=4096,
                 hbm_budget_bytes=None, reserved_cus=32):
        t = torch()
        self.X = to_device(X_train, np.float64)
        self.alpha = to_device(alpha, np.float64)
        self.sigma = float(sigma)
        self.nmax = int(nmax)
        self.max_batch = int(max_batch_molecules)
        self.hbm_budget = hbm_budget_bytes
        cus = _native.load().fmlh_device_cu_count()
        self._raw_streams = []
        self.predict_stream = self._stream(t, 0)
        self.fit_stream = self._stream(t, max(cus - int(reserved_cus), 1) if reserved_cus else 0)
        self._q: "queue.Queue[PredictRequest]" = queue.Queue()
        self._fits: "queue.Queue[FitRequest]" = queue.Queue()
        self._stop = threading.Event()
        self._threads = []
        self.batches = []        # (n_requests, n_molecules, seconds) per batch
        self.fit_log = []        # (n, start, end) perf_counter seconds per admitted fit

    # -- client API ----------------------------------------------------------
    def submit(self, charges, coords) -> Future:
        r = PredictRequest(list(charges), list(coords))
        self._q.put(r)
        return r.future

    def submit_fit(self, X, y, sigma, lam, precision="fp64") -> Future:
        r = FitRequest(np.asarray(X), np.asarray(y), float(sigma), float(lam), precision)
        self._fits.put(r)
        return r.future

    def start(self):
        self._threads = [threading.Thread(target=self._predict_loop, daemon=True),
                         threading.Thread(target=self._fit_loop, daemon=True)]
        for th in self._threads:
            th.start()
        return self

    def _stream(self, t, cu_count):
        p = ctypes.c_void_p(0)
        _native.call("fmlh_stream_create", int(cu_count), ctypes.byref(p))
        self._raw_streams.append(p)
        return t.cuda.ExternalStream(p.value)

    def stop(self):
        self._stop.set()
        for th in self._threads:
            th.join(timeout=30)
        self.predict_stream.synchronize()
        self.fit_stream.synchronize()

    def close(self):
        """Stop the loops and release the two streams."""
        self.stop()
        for p in self._raw_streams:
            _native.call("fmlh_stream_destroy", p)
        self._raw_streams = []

    # -- batching ------------------------------------------------------------
    def _drain(self):
        """Block for the first request, then take everything already pending
        (up to the batch size in molecules)."""
        try:
            first = self._q.get(timeout=0.05)
        except queue.Empty:
            return []
        batch = [first]
        nmol = len(first.charges)
        while nmol < self.max_batch:
            try:
                r = self._q.get_nowait()
            except queue.Empty:
                break
            batch.append(r)
            nmol += len(r.charges)
        return batch

    def predict_batch(self, requests):
        """Run one batch synchronously on the predict stream; returns the
        per-request prediction arrays."""
        t = torch()
        zs, rs, counts = [], [], []
        for r in requests:
            for z, c in zip(r.charges, r.coords):
                zs.append(np.asarray(z, dtype=np.float64))
                rs.append(np.asarray(c, dtype=np.float64).reshape(-1, 3))
                counts.append(len(z))
        off = np.zeros(len(counts) + 1, dtype=np.int32)
        off[1:] = np.cumsum(counts)
        with t.cuda.stream(self.predict_stream):
            Xq = generate_coulomb_matrices(np.concatenate(zs), np.concatenate(rs), off,
                                           nmax=self.nmax, on_device=True)
            Kt = kernel_matrix(self.X, Xq, self.sigma)
            y = predict_from_kernel(Kt, self.alpha)
            yh = y.cpu().numpy()
        out, k = [], 0
        for r in requests:
            m = len(r.charges)
            out.append(yh[k:k + m])
            k += m
        return out

    def _predict_loop(self):
        while not self._stop.is_set():
            batch = self._drain()
            if not batch:
                continue
            t0 = time.perf_counter()
            try:
                res = self.predict_batch(batch)
            except Exception as e:  # deliver the error to every waiter
                for r in batch:
                    r.future.set_exception(e)
                continue
            t1 = time.perf_counter()
            self.batches.append((len(batch), sum(len(r.charges) for r in batch), t1 - t0))
            for r, y in zip(batch, res):
                r.future.set_result((y, t1 - r.t_arrival, t1))

    # -- fits ----------------------------------------------------------------
    def free_hbm(self) -> int:
        t = torch()
        if self.hbm_budget is not None:
            return int(self.hbm_budget)
        free, _ = t.cuda.mem_get_info()
        return int(free)

    def _fit_loop(self):
        pending = []
        while not self._stop.is_set():
            try:
                pending.append(self._fits.get(timeout=0.05))
            except queue.Empty:
                pass
            admitted = [r for r in pending if r.hbm_bytes() <= self.free_hbm()]
            for r in admitted:
                pending.remove(r)
                t0 = time.perf_counter()
                try:
                    t = torch()
                    with t.cuda.stream(self.fit_stream):
                        m = KRR(r.sigma, r.lam, precision=r.precision).fit(r.X.astype(np.float64), r.y)
                        self.fit_stream.synchronize()
                    self.fit_log.append((int(r.X.shape[0]), t0, time.perf_counter()))
                    r.future.set_result(m)
                except Exception as e:
                    r.future.set_exception(e)
        for r in pending:
            r.future.set_exception(RuntimeError("server stopped before the fit was admitted"))


def latency_percentiles(latencies_s):
    a = np.sort(np.asarray(latencies_s)) * 1e3
    return {"p50_ms": float(np.percentile(a, 50)), "p99_ms": float(np.percentile(a, 99)),
            "mean_ms": float(a.mean()), "n": int(a.size)}
