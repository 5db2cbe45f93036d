"""Continuous-batching KRR serving on one GPU.

No counterpart exists in the reference (fml has no serving code); this is
the serving form built from the hot-path pieces:

* the model (training Coulomb matrices X, coefficients alpha, sigma) stays
  resident in HBM;
* predict requests (lists of molecules) queue up; a scheduler thread drains
  everything pending into one batch, so requests that arrive while a batch
  runs are admitted into the next one -- continuous batching, not a queue
  served one request at a time;
* a batch is one pass of the GPU path, one native call (fmlh_krr_predict):
  sorted Coulomb matrices for all its molecules (one workgroup per
  molecule), the Gaussian kernel block against the training set, and
  K^T alpha with one workgroup per test molecule; inputs and outputs move
  through reused pinned buffers, one event wait per batch;
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
import gc
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .device import to_device, torch
from .krr import KRR


class PredictRequest:
    """One predict request: its molecules packed once, on the submitting
    thread, into contiguous arrays (charges z, coordinates r and atoms per
    molecule), so a batch costs a few concatenations of whole requests
    instead of per-molecule work on the serving thread.  A request that
    cannot be served (a molecule above the server's nmax atoms, coordinates
    not matching the charges) keeps the error and fails alone."""

    __slots__ = ("z", "r", "counts", "error", "t_arrival", "future")

    def __init__(self, charges, coords, nmax=None):
        self.t_arrival = time.perf_counter()
        self.future = Future()
        self.error = None
        charges, coords = list(charges), list(coords)
        self.counts = np.fromiter(map(len, charges), dtype=np.int64, count=len(charges))
        if not charges:
            self.z, self.r = np.empty(0), np.empty((0, 3))
            return
        try:
            self.z = np.concatenate(charges, dtype=np.float64)
            self.r = np.concatenate([np.reshape(c, (-1, 3)) for c in coords], axis=0, dtype=np.float64)
        except ValueError as e:
            self.error = ValueError(f"malformed molecules: {e}")
            return
        if len(coords) != len(charges) or self.r.shape[0] != self.z.shape[0]:
            self.error = ValueError("coordinates do not match the charges")
        elif nmax is not None and self.counts.max() > nmax:
            self.error = ValueError(f"a molecule has more than nmax={nmax} atoms")

    @property
    def nmol(self) -> int:
        return int(self.counts.shape[0])


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


# CU-partitioned streams, one per (device, CU count), created on first use
# and never destroyed: torch's pinned-memory allocator records an event on
# every stream a pinned block was used on when the block is freed, possibly
# long after the server that used the stream has stopped, so the stream has
# to outlive every pinned buffer it ever touched.
_CU_STREAMS: dict = {}
_CU_STREAMS_LOCK = threading.Lock()


def _cu_stream(t, cu_count):
    key = (t.cuda.current_device(), int(cu_count))
    with _CU_STREAMS_LOCK:
        st = _CU_STREAMS.get(key)
        if st is None:
            p = ctypes.c_void_p(0)
            _native.call("fmlh_stream_create", int(cu_count), ctypes.byref(p))
            st = t.cuda.ExternalStream(p.value)
            _CU_STREAMS[key] = st
        return st


class _Staging:
    """Pinned host and device buffers of one predict batch (atoms, molecules,
    the kernel block's workspace), grown on demand and reused, so a batch
    costs three asynchronous uploads, one native call, one download and one
    event wait."""

    def __init__(self, t, ntrain, nmax):
        self.t, self.ntrain, self.nmax = t, ntrain, nmax
        self.cap_atoms = self.cap_mols = 0
        self.done = t.cuda.Event()

    def ensure(self, natoms, nmol):
        if natoms <= self.cap_atoms and nmol <= self.cap_mols:
            return
        t = self.t
        ca = max(natoms, 2 * self.cap_atoms)
        cm = max(nmol, 2 * self.cap_mols)
        f64, i32 = t.float64, t.int32
        self.h_z = t.empty(ca, dtype=f64, pin_memory=True)
        self.h_r = t.empty((ca, 3), dtype=f64, pin_memory=True)
        self.h_off = t.empty(cm + 1, dtype=i32, pin_memory=True)
        self.h_y = t.empty(cm, dtype=f64, pin_memory=True)
        self.hz, self.hr, self.hoff, self.hy = (self.h_z.numpy(), self.h_r.numpy(), self.h_off.numpy(),
                                                self.h_y.numpy())
        self.d_z = t.empty(ca, dtype=f64, device="cuda")
        self.d_r = t.empty((ca, 3), dtype=f64, device="cuda")
        self.d_off = t.empty(cm + 1, dtype=i32, device="cuda")
        self.d_y = t.empty(cm, dtype=f64, device="cuda")
        self.wsb = _native.load().fmlh_krr_predict_ws_bytes(cm, self.ntrain, self.nmax)
        self.ws = t.empty(max(self.wsb, 1), dtype=t.uint8, device="cuda")
        self.cap_atoms, self.cap_mols = ca, cm


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
        self.predict_stream = self._stream(t, 0)
        self.fit_stream = self._stream(t, max(cus - int(reserved_cus), 1) if reserved_cus else 0)
        self._q: "queue.Queue[PredictRequest]" = queue.Queue()
        self._fits: "queue.Queue[FitRequest]" = queue.Queue()
        self._stop = threading.Event()
        self._threads = []
        self._lock = threading.Lock()
        self._staging = _Staging(t, int(self.X.shape[0]), self.nmax)
        self._staging.ensure(64 * self.nmax, 1024)
        self.batches = []        # (n_requests, n_molecules, seconds) per batch
        self.fit_log = []        # (n, start, end) perf_counter seconds per admitted fit

    # -- client API ----------------------------------------------------------
    def submit(self, charges, coords) -> Future:
        r = PredictRequest(charges, coords, self.nmax)
        self._q.put(r)
        return r.future

    def submit_fit(self, X, y, sigma, lam, precision="fp64") -> Future:
        r = FitRequest(np.asarray(X), np.asarray(y), float(sigma), float(lam), precision)
        self._fits.put(r)
        return r.future

    def start(self, freeze_gc=True, warm_fits=("fp64",)):
        """Start the predict and fit loops.

        warm_fits: precision forms to run once on a tiny fit (64 of the
        served training rows, from float32 and float64 host arrays) on the
        fit stream before serving starts.  The first fit in a process pays
        one-time runtime costs (first launches of its kernels and of the
        fit stream), during which a concurrent predict batch waited: a
        256-row first fit took 65 ms instead of 0.5 ms and held a
        one-molecule batch for 54 ms.  Paid here, it does not land on the
        predictions that overlap the first real fit.

        freeze_gc: collect, then move every object alive now to the
        collector's permanent generation (gc.freeze), so the periodic full
        collections during serving scan only what serving allocates instead
        of pausing all threads for the whole heap (torch's modules alone
        hold hundreds of thousands of objects)."""
        # one predict of a single-atom molecule loads the predict path's code
        self.predict_batch([PredictRequest([[1.0]], [[0.0, 0.0, 0.0]], self.nmax)])
        n = min(64, int(self.X.shape[0]))
        if warm_fits and n > 0:
            x64 = self.X[:n].cpu().numpy()
            for prec in warm_fits:
                for xh in (x64.astype(np.float32), x64):
                    self._run_fit(FitRequest(xh, np.zeros(n), self.sigma, 1.0, prec))
        if freeze_gc:
            gc.collect()
            gc.freeze()
        self._threads = [threading.Thread(target=self._predict_loop, daemon=True),
                         threading.Thread(target=self._fit_loop, daemon=True)]
        for th in self._threads:
            th.start()
        return self

    def _stream(self, t, cu_count):
        return _cu_stream(t, cu_count)

    def stop(self):
        self._stop.set()
        for th in self._threads:
            th.join(timeout=30)
        self.predict_stream.synchronize()
        self.fit_stream.synchronize()

    def close(self):
        """Stop the loops and drop the staging buffers.  The two streams are
        shared process-wide and stay alive (see _cu_stream)."""
        self.stop()
        self._staging = None

    # -- batching ------------------------------------------------------------
    def _drain(self):
        """Block for the first request, then take everything already pending
        (up to the batch size in molecules)."""
        try:
            first = self._q.get(timeout=0.05)
        except queue.Empty:
            return []
        batch = [first]
        nmol = first.nmol
        while nmol < self.max_batch:
            try:
                r = self._q.get_nowait()
            except queue.Empty:
                break
            batch.append(r)
            nmol += r.nmol
        return batch

    def predict_batch(self, requests):
        """Run one batch on the predict stream and wait for it; returns the
        per-request prediction arrays (a ValueError in place of the array for
        a request holding a molecule with more than nmax atoms)."""
        t = torch()
        out = [None] * len(requests)
        live = []
        for k, r in enumerate(requests):
            err = r.error
            if err is None and r.nmol and int(r.counts.max()) > self.nmax:
                err = ValueError(f"a molecule has more than nmax={self.nmax} atoms")
            if err is not None:
                out[k] = err
            elif r.nmol == 0:
                out[k] = np.empty(0)
            else:
                live.append(k)
        if not live:
            return out
        reqs = [requests[k] for k in live]
        counts = np.concatenate([r.counts for r in reqs])
        nmol = int(counts.shape[0])
        with self._lock:
            st = self._staging
            natoms = int(counts.sum())
            st.ensure(natoms, nmol)
            np.concatenate([r.z for r in reqs], out=st.hz[:natoms])
            np.concatenate([r.r for r in reqs], axis=0, out=st.hr[:natoms])
            st.hoff[0] = 0
            np.cumsum(counts, out=st.hoff[1:nmol + 1])
            ps = self.predict_stream
            with t.cuda.stream(ps):
                st.d_z[:natoms].copy_(st.h_z[:natoms], non_blocking=True)
                st.d_r[:natoms].copy_(st.h_r[:natoms], non_blocking=True)
                st.d_off[:nmol + 1].copy_(st.h_off[:nmol + 1], non_blocking=True)
                _native.call("fmlh_krr_predict", _native.ptr(st.d_z), _native.ptr(st.d_r),
                             _native.ptr(st.d_off), nmol, self.nmax, _native.ptr(self.X), self.X.stride(0),
                             int(self.X.shape[0]), _native.ptr(self.alpha), self.sigma, _native.ptr(st.d_y),
                             None, _native.ptr(st.ws), st.wsb, ctypes.c_void_p(ps.cuda_stream))
                st.h_y[:nmol].copy_(st.d_y[:nmol], non_blocking=True)
                st.done.record(ps)
            st.done.synchronize()
            yh = st.hy[:nmol].copy()
        i = 0
        for k, r in zip(live, reqs):
            out[k] = yh[i:i + r.nmol]
            i += r.nmol
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
            self.batches.append((len(batch), sum(r.nmol for r in batch), t1 - t0))
            for r, y in zip(batch, res):
                if isinstance(y, Exception):
                    r.future.set_exception(y)
                else:
                    r.future.set_result((y, t1 - r.t_arrival, t1))

    # -- fits ----------------------------------------------------------------
    def free_hbm(self) -> int:
        t = torch()
        if self.hbm_budget is not None:
            return int(self.hbm_budget)
        free, _ = t.cuda.mem_get_info()
        return int(free)

    def _run_fit(self, r):
        t = torch()
        with t.cuda.stream(self.fit_stream):
            # widen on the device: torch calls release the GIL, so the
            # predict thread keeps running meanwhile
            Xd = t.from_numpy(np.ascontiguousarray(r.X)).to("cuda").to(t.float64)
            m = KRR(r.sigma, r.lam, precision=r.precision).fit(Xd, r.y)
            self.fit_stream.synchronize()
        return m

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
                    m = self._run_fit(r)
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
