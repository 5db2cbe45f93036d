#!/usr/bin/env python3
"""Benchmark of the fml KRR hot path on one MI355X.

Default workload (the global ARAD kernel case): 500 synthetic molecules from
seed 0 (5-23 atoms, H/C/N/O/S), ARAD descriptors (500, 23, 5, 23) FP64,
width 0.2, cut 5.0, r_width 1.0, c_width 0.5, the full 500 x 500 molecular
kernel for sigma in {0.1, 1, 10} in one launch.  One step = one call of the
GPU entry point (descriptor preparation, self terms, molecule-pair tiles) on
inputs already resident in HBM.  Metric: kernel entries per second
(3 sigmas x 500 x 500 entries per step).

Printed: one JSON line with the metric, the roofline of the dominant kernel
(k_arad_tiles, timed with HIP events on its stream) and a CPU baseline (the
reference Fortran built by oracle/Makefile, on a bounded row block, all
host threads).

    python bench.py [--steps K] [--warmup W] [--workload NAME]

Other workloads: arad4096 / arad4096_fp32 (4096 x 4096 uneven molecules),
krr (1000-molecule FP64 KRR), gaussian64 / gaussian32 / gaussianbf16x3 /
laplacian64 / laplacian32 (36k x 36k kernels), cholesky64 / cholesky32 /
cholesky_oz, krr36k_fp64 / krr36k_mixed / krr36k_ozaki, serving, copy
(counter calibration).
"""
from __future__ import annotations

import argparse
import collections
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6            # vector and matrix, FMA = 2 flop
PEAK_FP32_TFLOPS = 157.3
PEAK_INT8_POPS = 5.0               # dense INT8 matrix
PEAK_FP64_VALU_TOPS = 39.3         # FP64 VALU lane-operations/s: the FP64 vector peak / 2 (FMA = 1 op)
PEAK_FP32_VALU_TOPS = 78.6         # FP32 VALU lane-operations/s: the FP32 vector peak / 2

# FP64 operations per inner (m1, m2) evaluation of the overlap term: the
# reference's 12 arithmetic operations (difference, square, scale, the
# products by s1, s2 and the type factor T, the angular (1 + a1 a2) + b1 b2,
# the product and the accumulation) plus the band test, and the exponential
# at the 12 FP64 instructions its compiled form issues.
ARAD_OPS_PER_EVAL = 25

# HBM bytes per launch of k_arad_tiles (FP64, tile shape 1024 pairs, 5 waves,
# set-1 atom fastest), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of
# this script, each scaled by the counter's reading for a device copy of known
# size (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0): profiles/hbm_traffic_counters.md.
ARAD_TRAFFIC_BYTES = {"arad500": 2.0 * 457764.4 * 1024 + 18797.98 * 1024}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="arad",
                    choices=["arad", "krr", "gaussian64", "gaussian32", "gaussianbf16x3",
                             "laplacian64", "laplacian32", "cholesky64", "cholesky32", "cholesky_oz",
                             "krr36k_fp64", "krr36k_mixed", "krr36k_ozaki", "serving", "arad4096", "arad4096_fp32", "copy",
                             "atomic", "atomic_lap", "atomic_mfma", "aras", "inverse64", "solve64", "solve32"])
    ap.add_argument("--duration", type=float, default=5.0, help="serving run length in seconds")
    ap.add_argument("--rate", type=float, default=2000.0, help="serving: Poisson predict arrivals per second")
    ap.add_argument("--n", type=int, default=0, help="problem size override for the 36k workloads")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-clauses", action="store_true",
                    help="default workload: skip the fit-time and serving clauses of the metric")
    ap.add_argument("--clause-duration", type=float, default=5.0,
                    help="serving clause of the default workload: run length in seconds")
    ap.add_argument("--tol", type=float, default=1e-10, help="mixed KRR fit: relative residual to reach")
    ap.add_argument("--max-iter", type=int, default=400, help="mixed KRR fit: iteration cap")
    ap.add_argument("--tiling", default=None,
                    help="ARAD tile shape pairs,waves,groups_per_cu,pair_order instead of autotuning")
    ap.add_argument("--eager", action="store_true", help="plain calls instead of HIP-graph replays")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="row block of the ARAD CPU baseline sample (0: the whole 500-molecule set, 100 rows at 4096)")
    return ap.parse_args()


def host_threads():
    v = os.environ.get("OMP_NUM_THREADS")
    try:
        return int(v) if v else (os.cpu_count() or 1)
    except ValueError:
        return os.cpu_count() or 1


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def median_time(fn, runs=5, warmup=1):
    """Median wall seconds of `runs` calls after `warmup` untimed ones; the
    last call's result."""
    out = None
    for _ in range(warmup):
        out = fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts, out


def cpu_record(value, unit, kind, sample, ts):
    return {"value": value, "unit": unit, "cores": host_threads(), "kind": kind, "sample": sample,
            "host_cpu": host_cpu_model(), "runs_s": [round(t, 4) for t in ts]}


def arad_inner_evaluations(X1, n1, X2, n2, cut, width):
    """Inner (m1, m2) evaluations of the global kernel of set 1 against set 2:
    every ordered pair of neighbour slots (u from an atom of set 1, v from an
    atom of set 2) with (v-u)^2 < (8 width)^2, u, v <= cut.  Counted on the
    sorted multisets of all neighbour distances (band edges to within one
    rounding)."""
    def dists(X, n):
        ds = []
        for i in range(X.shape[0]):
            for a in range(n[i]):
                d = X[i, a, 0, :n[i]]
                c = np.argmax(d > cut) if np.any(d > cut) else n[i]
                ds.append(d[:c])
        return np.sort(np.concatenate(ds))
    u = dists(X1, n1)
    v = u if X2 is X1 else dists(X2, n2)
    w = 8.0 * width
    lo = np.searchsorted(v, u - w, side="right")
    hi = np.searchsorted(v, u + w, side="left")
    return int(np.sum(hi - lo))


def arad_workload(args, big=False, precision=0):
    """big=False: the 500 x 500 global kernel (sigmas 0.1, 1, 10, amax 23).
    big=True: 4096 x 4096 molecules with 3-29 atoms from a truncated geometric
    distribution (mean ~14; set 1 seed 0, set 2 seed 1), amax 29, sigmas
    {1, 4}; precision 1 = FP32 inner overlap."""
    import torch
    import fml_amd._native as nat
    from fml_amd import synth
    from fml_amd.arad import ARAD
    from fml_amd.kernels.kernels import arad_types

    width, cut, rw, cw = 0.2, 5.0, 1.0, 0.5
    workload_key = None if precision else ("arad4096" if big else "arad500")
    if big:
        nm, amax, sig = args.n or 4096, 29, [1.0, 4.0]
        ms1 = synth.generate(nm, seed=0, nmin=3, nmax=29, dist="geometric")
        ms2 = synth.generate(nm, seed=1, nmin=3, nmax=29, dist="geometric")
    else:
        nm, amax, sig = 500, 23, [0.1, 1.0, 10.0]
        ms1 = ms2 = synth.generate(nm, seed=0, nmin=5, nmax=23)
    ar = ARAD(maxMolSize=amax, maxAts=amax, cut=5.0)
    t0 = time.perf_counter()
    X1 = np.array([ar.describe(ms1.molecule(m)[1], ms1.molecule(m)[0]) for m in range(nm)])
    X2 = X1 if ms2 is ms1 else np.array([ar.describe(ms2.molecule(m)[1], ms2.molecule(m)[0])
                                         for m in range(nm)])
    host_describe_s = time.perf_counter() - t0
    z1, n1 = arad_types([ms1.molecule(m)[0] for m in range(nm)], amax)
    z2, n2 = (z1, n1) if ms2 is ms1 else arad_types([ms2.molecule(m)[0] for m in range(nm)], amax)
    evals = arad_inner_evaluations(X1, n1, X2, n2, cut, width)

    from fml_amd.batch import AradKernelCall
    call = AradKernelCall(X1, z1, n1, X2, z2, n2, sig, mode="global", width=width, cut_distance=cut,
                          r_width=rw, c_width=cw, precision=precision)
    out = call.out
    # tile shape: the autotuner times its candidates on the device (on the
    # first 512 set-1 molecules for the 4096 x 4096 set) before the timed steps
    t0 = time.perf_counter()
    if args.tiling:
        tiling = nat.AradTiling(*[int(v) for v in args.tiling.split(",")])
        call.tiling = tiling
        call.run()
        tuned_ms = float("nan")
    else:
        tiling, tuned_ms = call.autotune(sample_rows=512 if big else None, reps=3)
    tune_s = time.perf_counter() - t0
    torch.cuda.synchronize()

    def step():
        call.run()
    step.graph_capable = True

    units_per_step = len(sig) * nm * nm
    if big:
        desc = (f"global ARAD kernel, {nm}x{nm} molecules (3-29 atoms, truncated geometric, seeds 0/1), "
                f"descriptors ({nm},29,5,29), sigmas {{1,4}}, {'FP32' if precision else 'FP64'} overlap")
    else:
        desc = ("global ARAD kernel, 500x500 molecules (5-23 atoms, seed 0), descriptors (500,23,5,23), "
                "sigmas {0.1,1,10}, width 0.2, cut 5.0, r_width 1.0, c_width 0.5")
    info = {
        "metric": "kernel entries/s",
        "unit": "entries/s",
        "dtype": "f32" if precision else "f64",
        "config": {"workload": desc, "molecules": nm, "amax": amax,
                   "sigmas": sig, "inner_evaluations_per_step": evals,
                   "tiling": {"pairs_per_tile": tiling.pairs_per_tile,
                              "waves_per_simd": tiling.waves_per_simd,
                              "groups_per_cu": tiling.groups_per_cu,
                              "pair_order": tiling.pair_order,
                              "autotuned_on": (f"{512 if big else nm} x {nm} molecules" if not args.tiling
                                               else "fixed by --tiling"),
                              "autotune_s": round(tune_s, 2),
                              "tuned_ms_per_call": round(tuned_ms, 3)}},
        "kernel": "k_arad_tiles",
    }

    def roofline(t_main_s):
        ops = evals * ARAD_OPS_PER_EVAL
        ach = ops / t_main_s / 1e12
        peak = PEAK_FP32_VALU_TOPS if precision else PEAK_FP64_VALU_TOPS
        # the call's compulsory bytes: both descriptor sets and type channels
        # read, the kernel written
        alg = (X1.nbytes + X2.nbytes + z1.nbytes + z2.nbytes + n1.nbytes + n2.nbytes
               + 8 * len(sig) * X1.shape[0] * X2.shape[0])
        return {"bound": "valu", "achieved": round(ach, 3), "peak": peak,
                "unit": "T FP32-VALU-ops/s" if precision else "T FP64-VALU-ops/s",
                "frac": round(ach / peak, 4),
                "traffic": ARAD_TRAFFIC_BYTES.get(workload_key),
                "algorithmic_bytes": int(alg),
                "inner_evaluations_per_s": round(evals / t_main_s, 1),
                "ops_per_evaluation": ARAD_OPS_PER_EVAL}

    def cpu_baseline(rows):
        import oracle
        kind = "reference" if oracle.ref_available() else "port"
        fn = oracle.ref_kernels_arad if kind == "reference" else oracle.kernels_arad
        rows = min(rows or (100 if big else nm), nm)
        runs, warm = (5, 1) if not big else (1, 0)
        dt, ts, kc = median_time(lambda: fn(X1[:rows], X2, z1[:rows], z2, n1[:rows], n2, sig, width, cut,
                                            rw, cw), runs=runs, warmup=warm)
        g = torch.stack([out[:, :rows, k] for k in range(len(sig))]).cpu().numpy()  # (s, nm, rows)
        g = np.transpose(g, (0, 2, 1))
        m = kc > 1e-300
        err = float(np.max(np.abs(g[m] - kc[m]) / kc[m]))
        what = "the full kernel" if rows == nm else f"rows 0..{rows - 1} x all {nm} molecules"
        return (cpu_record(round(len(sig) * rows * nm / dt, 1), "entries/s", kind,
                           f"{what} ({len(sig) * rows * nm} entries), median of {runs} runs"
                           f"{' after a warm-up' if warm else ''}: {dt:.2f} s", ts),
                {"max_rel": err, "bitwise_equal_fraction": float(np.mean(g == kc))})

    def finish(res):
        # the same call through the Python API with host (numpy) inputs and
        # output: PCIe transfers included (outside the timed region)
        from fml_amd.kernels import fget_kernels_arad
        reps = 1 if big else 3
        fget_kernels_arad(X1, X2, z1, z2, n1, n2, sig, width=width, cut_distance=cut, r_width=rw,
                          c_width=cw, precision=precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fget_kernels_arad(X1, X2, z1, z2, n1, n2, sig, width=width, cut_distance=cut, r_width=rw,
                              c_width=cw, precision=precision)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res["host_buffers"] = {"ms_per_call": round(dt * 1e3, 3),
                               "entries_per_s": round(units_per_step / dt, 1),
                               "note": "numpy in, numpy out: descriptor upload and kernel download included"}
        # the descriptors themselves on the device (fmlh_arad_describe), both
        # sets, against the host numpy ARAD.describe loop that made X1/X2
        from fml_amd.batch import AradDescribeCall
        sets = [(ms1, X1)] if ms2 is ms1 else [(ms1, X1), (ms2, X2)]
        calls = [AradDescribeCall(m.charges, m.coords, m.offsets, amax=amax, cut=5.0) for m, _ in sets]
        for c in calls:
            c.run()
        torch.cuda.synchronize()
        diff = 0.0
        for c, (_, X) in zip(calls, sets):
            q = c.q.cpu().numpy()
            if not np.array_equal(q[:, :, 0:3], X[:, :, 0:3]):
                raise RuntimeError("device ARAD descriptors differ from the host's in channels 0-2")
            a, b = q[:, :, 3:5], X[:, :, 3:5]
            fin = ~np.isnan(b)
            diff = max(diff, float(np.max(np.abs(a[fin] - b[fin]), initial=0.0)))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            for c in calls:
                c.run()
        e1.record()
        torch.cuda.synchronize()
        res["descriptors_on_device"] = {
            "ms_per_call": round(e0.elapsed_time(e1) / 5, 3),
            "molecules": nm * len(sets),
            "host_numpy_s": round(host_describe_s, 3),
            "max_abs_diff_channels_3_4": diff,
            "note": "fmlh_arad_describe of every set, one wavefront per atom row; channels 0-2 "
                    "bit-identical to the host ARAD.describe"}

    return info, step, units_per_step, roofline, cpu_baseline, finish


_CM_SETS = {}


def _cm_set(nmol, seed, nmin, nmax_atoms, nmax):
    """A synthetic molecule set (generated once per process and key)."""
    from fml_amd import synth
    key = (nmol, seed, nmin, nmax_atoms)
    if key not in _CM_SETS:
        _CM_SETS[key] = synth.generate(nmol, seed=seed, nmin=nmin, nmax=nmax_atoms)
    return _CM_SETS[key]


def krr_workload(args):
    """Plain FP64 KRR (the CPU-runnable case): 1000 molecules, seed 0, 5-23
    atoms, sorted CM nmax 23; Gaussian sigma 700 on 800 train, lambda 1e-8,
    predict 200.  One step = CM generation + train kernel + Cholesky fit +
    test kernel + prediction, all on the GPU."""
    from fml_amd.batch import KRRCall
    ms = _cm_set(1000, 0, 5, 23, 23)
    # inputs, kernels, factor, alpha and predictions resident; one step is the
    # whole chain (coordinates -> predictions), also replayable as one graph
    call = KRRCall(ms.charges, ms.coords, ms.offsets, 800, ms.energies[:800], 700.0, 1e-8, nmax=23)
    state = {"pred": call.pred, "info": call.info}

    def step():
        call.run()
    step.graph_capable = True

    info = {"metric": "KRR fit+predict time", "unit": "ms", "dtype": "f64", "kernel": "k_potf2..",
            "higher_is_better": False, "roofline_on_step": True,
            "config": {"workload": "plain FP64 KRR: 1000 molecules (seed 0, 5-23 atoms), sorted CM "
                                   "nmax 23, Gaussian sigma 700, 800 train / 200 test, lambda 1e-8"}}

    # FP64 flop per step: two Gaussian blocks (2 D per entry), the factor
    # (n^3/3), two triangular solves and the prediction
    flop = 2.0 * 276 * (800 * 800 + 800 * 200) + 800 ** 3 / 3.0 + 2 * 800 ** 2 + 2 * 800 * 200

    def roofline(t):
        ach = flop / t / 1e12
        return {"bound": "mfma", "achieved": round(ach, 4), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP64_TFLOPS, 5), "traffic": None, "flop_per_step": flop,
                "note": "latency-bound: ~40 dependent launches on an 800-row system; the whole step is timed"}

    def cpu_baseline(rows):
        import oracle
        kind = "reference" if oracle.ref_available() else "port"
        cm = oracle.ref_coulomb_matrix if kind == "reference" else oracle.coulomb_matrix
        gk = oracle.ref_gaussian_kernel if kind == "reference" else oracle.gaussian_kernel
        cs = oracle.ref_cho_solve if kind == "reference" else (lambda A, y: oracle.cho_solve(A, y)[0])

        def run():
            X = np.array([cm(*ms.molecule(i), 23) for i in range(ms.nmol)])
            K = gk(X[:800].T, X[:800].T, 700.0)
            K[np.diag_indices_from(K)] += 1e-8
            a = cs(K, ms.energies[:800])
            return gk(X[:800].T, X[800:].T, 700.0).T @ a
        dt, ts, pred_cpu = median_time(run)
        pred = state["pred"].cpu().numpy()
        err = float(np.max(np.abs(pred - pred_cpu)) / np.max(np.abs(pred_cpu)))
        return cpu_record(round(dt * 1e3, 3), "ms", kind,
                          "the whole fit+predict: Coulomb matrices (one call per molecule), Gaussian "
                          "800x800 and 800x200, fcho_solve, predictions; median of 5 runs after a "
                          f"warm-up: {dt * 1e3:.1f} ms", ts), err

    def finish(res):
        import oracle
        pred = state["pred"].cpu().numpy()
        X = oracle.coulomb_matrices(ms.charges, ms.coords, ms.offsets, 23)[0]
        K = oracle.gaussian_kernel(X[:800].T, X[:800].T, 700.0)
        K[np.diag_indices_from(K)] += 1e-8
        ao, _ = oracle.cho_solve(K, ms.energies[:800])
        po = oracle.gaussian_kernel(X[:800].T, X[800:].T, 700.0).T @ ao
        res["mae_gpu"] = float(np.mean(np.abs(pred - ms.energies[800:])))
        res["mae_oracle"] = float(np.mean(np.abs(po - ms.energies[800:])))
        res["prediction_max_rel_vs_oracle"] = float(np.max(np.abs(pred - po)) / np.max(np.abs(po)))
        res["potrf_info"] = int(state["info"].item())
        res["cond_estimate"] = float(np.linalg.cond(K))

    return info, step, None, roofline, cpu_baseline, finish


def dense36k_workload(args, kind):
    """Kernel construction at scale: 36,000 x 36,000 training kernel on
    nmax-29 Coulomb matrices (435 features; 40k molecules seed 0, 3-29 atoms,
    first 36k).  kind: gaussian64 | gaussian32 | gaussianbf16x3 | laplacian64
    | laplacian32."""
    import torch
    from fml_amd.device import to_device
    from fml_amd.krr import kernel_matrix
    from fml_amd.representations import generate_coulomb_matrices
    n = args.n or 36000
    ms = _cm_set(n, 0, 3, 29, 29)
    X64 = generate_coulomb_matrices(to_device(ms.charges), to_device(ms.coords),
                                    to_device(ms.offsets, np.int32), nmax=29, on_device=True)
    X32 = X64.to(torch.float32)
    d = X64.shape[1]
    is_g = kind.startswith("gaussian")
    sigma = 1000.0 if is_g else 10000.0
    prec = {"gaussian64": "fp64", "gaussian32": "fp32", "gaussianbf16x3": "bf16x3"}.get(kind)
    Xin = X64 if kind.endswith("64") else X32
    out = torch.empty((n, n), dtype=torch.float64 if kind.endswith("64") else torch.float32,
                      device="cuda")

    def step():
        if is_g:
            kernel_matrix(Xin, Xin, sigma, "gaussian", precision=prec, out=out)
        else:
            kernel_matrix(Xin, Xin, sigma, "laplacian", out=out)

    units = n * n
    info = {"metric": "kernel entries/s", "unit": "entries/s",
            "dtype": {"gaussian64": "f64", "gaussian32": "f32", "gaussianbf16x3": "bf16x3/f32",
                      "laplacian64": "f64", "laplacian32": "f32"}[kind],
            "config": {"workload": f"{kind} kernel {n}x{n}, D={d}, sigma {sigma}"}}

    def roofline(t):
        fl = 2.0 * n * n * d
        if kind in ("gaussian64", "gaussian32", "gaussianbf16x3"):
            # a set against itself runs only the lower-triangle 128-tiles (the
            # diagonal tiles whole) and mirrors their stores
            g = -(-n // 128)
            fl = 2.0 * d * (g * (g + 1) // 2) * 128 * 128
        if kind in ("gaussian64", "gaussian32", "gaussianbf16x3"):
            peak = {"gaussian64": PEAK_FP64_TFLOPS, "gaussian32": PEAK_FP32_TFLOPS,
                    "gaussianbf16x3": 2500.0 / 3.0}[kind]
            ach = fl / t / 1e12
            return {"bound": "mfma", "achieved": round(ach, 2), "peak": round(peak, 1), "unit": "TFLOP/s",
                    "frac": round(ach / peak, 4), "traffic": None}
        g = -(-n // 64)   # a set against itself: the lower-triangle 64-tiles run
        ach = (2.0 * d * (g * (g + 1) // 2) * 64 * 64) / t / 1e12   # sub + abs-add per feature
        peak = PEAK_FP64_VALU_TOPS if kind == "laplacian64" else PEAK_FP32_VALU_TOPS
        return {"bound": "valu", "achieved": round(ach, 2), "peak": peak,
                "unit": "T VALU-ops/s", "frac": round(ach / peak, 4), "traffic": None}

    def finish(res):
        if kind != "gaussianbf16x3":
            return
        # near pairs the split form recomputed by direct differences (tau 2e-2)
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        kernel_matrix(Xin, Xin, sigma, "gaussian", precision=prec, out=out, recomputed=cnt)
        c = int(cnt.item())
        res["near_pairs_recomputed"] = {"count": c, "fraction": c / float(n * n),
                                        "note": "entries with d2 <= 2e-2 (|a|^2 + |b|^2), the diagonal "
                                                "included, recomputed from FP32 differences"}

    return info, step, units, roofline, None, finish


def atomic_workload(args, kind=0):
    """The atomic (vector) kernels of the atomic KRR path: kind 0 the
    Gaussian (fget_vector_kernels_gaussian, fkernels.f90:93-162, sigmas
    {30, 300, 3000}), kind 1 the Laplacian (fget_vector_kernels_laplacian,
    :23-91, sigmas {300, 3000, 30000}); local Coulomb matrices (nmax 23, 276
    features) of two 4000-molecule sets (5-23 atoms, seeds 0 and 1); one
    step = one fmlh_vector_kernels call on HBM-resident atom rows."""
    import torch
    import fml_amd._native as nat
    from fml_amd import synth
    from fml_amd.device import empty, to_device
    from fml_amd.representations import generate_local_coulomb_matrices
    nm = args.n or 4000
    nmax = 23
    sig = [300.0, 3000.0, 30000.0] if kind == 1 else [30.0, 300.0, 3000.0]
    ms1 = synth.generate(nm, seed=0, nmin=5, nmax=23)
    ms2 = synth.generate(nm, seed=1, nmin=5, nmax=23)
    x1 = generate_local_coulomb_matrices(ms1.charges, ms1.coords, ms1.offsets, nmax, on_device=True)
    x2 = generate_local_coulomb_matrices(ms2.charges, ms2.coords, ms2.offsets, nmax, on_device=True)
    off1 = to_device(ms1.offsets, np.int32)
    off2 = to_device(ms2.offsets, np.int32)
    na1, na2, d = int(x1.shape[0]), int(x2.shape[0]), int(x1.shape[1])
    amax = int(max(ms1.natoms.max(), ms2.natoms.max()))
    out = empty((nm, nm, len(sig)), np.float64)
    status = empty((1,), np.int32)
    arr = (ctypes.c_double * len(sig))(*sig)
    # molecules packed greedily into blocks of at most 128 atoms (host plan)
    from fml_amd.kernels.atomic import plan_blocks
    b1, b2 = plan_blocks(ms1.natoms), plan_blocks(ms2.natoms)
    blk1, blk2 = to_device(b1, np.int32), to_device(b2, np.int32)
    args_c = (kind, nat.ptr(x1), x1.stride(0), nat.ptr(off1), nm, na1, nat.ptr(x2), x2.stride(0),
              nat.ptr(off2), nm, na2, d, amax, arr, len(sig), nat.ptr(blk1), len(b1) - 1, nat.ptr(blk2),
              len(b2) - 1, nat.ptr(out), nat.ptr(status))

    def step():
        nat.call("fmlh_vector_kernels", *args_c, nat.stream_ptr())
    step.graph_capable = True

    units = nm * nm * len(sig)
    pairs = float(na1) * float(na2)
    name = {0: "Gaussian (fget_vector_kernels_gaussian)",
            1: "Laplacian (fget_vector_kernels_laplacian)",
            2: "Gaussian (fget_vector_kernels_gaussian, precision mfma)"}[kind]
    info = {"metric": "kernel entries/s", "unit": "entries/s", "dtype": "f64",
            "config": {"workload": f"atomic {name} kernel, {nm}x{nm} molecules (5-23 atoms, seeds 0/1), "
                                   f"local Coulomb matrices nmax {nmax} (D={d}), {na1}x{na2} atoms, "
                                   f"sigmas {{{','.join(f'{v:g}' for v in sig)}}}",
                       "molecules": nm, "atoms": [na1, na2], "features": d,
                       "atom_blocks": [len(b1) - 1, len(b2) - 1]},
            "kernel": {0: "k_atom_tiles<L2>", 1: "k_atom_tiles<L1>", 2: "k_atom_tiles<L2 MFMA>"}[kind]}

    def roofline(t):
        # matrix-core form: 2 D flop per atom pair (the dot product) on the
        # FP64 MFMA; VALU forms: per feature a difference and an abs-add
        # (Laplacian, 2 ops) or a difference, a product and a sum (Gaussian,
        # 3 ops, the reference's order)
        if kind == 2:
            fl = pairs * d * 2.0
            ach = fl / t / 1e12
            return {"bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(ach / PEAK_FP64_TFLOPS, 4), "traffic": None,
                    "atom_pairs_per_s": round(pairs / t, 1), "flop_per_atom_pair": 2 * d}
        per = 3 if kind == 0 else 2
        ops = pairs * d * per
        ach = ops / t / 1e12
        return {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP64_VALU_TOPS,
                "unit": "T FP64-VALU-ops/s", "frac": round(ach / PEAK_FP64_VALU_TOPS, 4), "traffic": None,
                "atom_pairs_per_s": round(pairs / t, 1), "ops_per_atom_pair": per * d}

    def cpu_baseline(rows):
        import oracle
        if int(status.item()) != 0:
            raise RuntimeError("atomic kernel reported an oversize block")
        rows = max(1, min(rows or 100, nm))
        bkind = "reference" if oracle.ref_available() else "port"
        fn = oracle.ref_vector_kernels if bkind == "reference" else oracle.vector_kernels
        r1, r2 = x1.cpu().numpy(), x2.cpu().numpy()
        q1 = np.zeros((rows, nmax, d))
        q2 = np.zeros((nm, nmax, d))
        for i in range(rows):
            q1[i, :ms1.natoms[i]] = r1[ms1.offsets[i]:ms1.offsets[i + 1]]
        for j in range(nm):
            q2[j, :ms2.natoms[j]] = r2[ms2.offsets[j]:ms2.offsets[j + 1]]
        t0 = time.perf_counter()
        kc = fn(1 if kind == 1 else 0, q1, q2, ms1.natoms[:rows], ms2.natoms, sig)
        dt = time.perf_counter() - t0
        g = np.transpose(out[:, :rows, :].cpu().numpy(), (2, 1, 0))
        m = kc > 1e-290
        err = float(np.max(np.abs(g[m] - kc[m]) / kc[m]))
        return ({"value": round(len(sig) * rows * nm / dt, 1), "unit": "entries/s",
                 "cores": host_threads(), "kind": bkind,
                 "sample": f"set-1 molecules 0..{rows - 1} x all {nm} set-2 molecules "
                           f"({len(sig) * rows * nm} entries), {dt:.2f} s"}, err)

    return info, step, units, roofline, cpu_baseline, None


def aras_workload(args):
    """The global ARAS kernel (fget_kernels_aras, faras_kernels.f90:150-483):
    500 x 500 molecules (5-23 atoms, seeds 0 and 1), descriptors
    (500, 23, 26, 23), default widths (t_width pi, d_width 0.2, cut 5,
    r_width 1, c_width 0.5, order 1, angular scale 0.1), sigmas {0.5, 2, 10};
    one step = one fmlh_aras_kernels call (per-set preparation, self
    distances and the molecule-pair kernel) on HBM-resident descriptors."""
    import torch
    import fml_amd._native as nat
    from fml_amd import synth
    from fml_amd.aras import ARAS
    from fml_amd.device import empty, to_device, workspace
    from fml_amd.kernels.aras_kernels import _neighbors, _pd_cached, aras_blocks
    nm, amax, sig = args.n or 500, 23, [0.5, 2.0, 10.0]
    ms1 = synth.generate(nm, seed=0, nmin=5, nmax=23)
    ms2 = synth.generate(nm, seed=1, nmin=5, nmax=23)
    ar = ARAS(maxMolSize=amax, maxAts=amax)
    X1 = np.array([ar.describe(ms1.molecule(m)[1], ms1.molecule(m)[0]) for m in range(nm)])
    X2 = np.array([ar.describe(ms2.molecule(m)[1], ms2.molecule(m)[0]) for m in range(nm)])
    nn1, nn2 = _neighbors(X1, ms1.natoms, 5.0), _neighbors(X2, ms2.natoms, 5.0)
    dX1, dX2 = to_device(X1), to_device(X2)
    dn1, dn2 = to_device(ms1.natoms, np.int32), to_device(ms2.natoms, np.int32)
    dnn1, dnn2 = to_device(nn1, np.int32), to_device(nn2, np.int32)
    pd = _pd_cached(1.0, 0.5)
    out = empty((nm, nm, len(sig)), np.float64)
    status = empty((1,), np.int32)
    lib = nat.load()
    wsb = lib.fmlh_aras_ws_bytes(nm, nm, amax, 1, 0)
    ws = workspace(wsb)
    arr = (ctypes.c_double * len(sig))(*sig)
    p = nat.ptr
    blk, ordv, nblk = aras_blocks(ms1.natoms, amax, X1[:, :, 1, 0])
    args_c = (0, p(dX1), p(dn1), p(dnn1), nm, p(dX2), p(dn2), p(dnn2), nm, amax, arr, len(sig), np.pi, 0.2,
              5.0, 1, p(pd), int(pd.shape[0]), 0.1, p(blk), p(ordv), nblk, p(out), p(status), p(ws), wsb)

    def step():
        nat.call("fmlh_aras_kernels", *args_c, nat.stream_ptr())
    step.graph_capable = True

    # VALU work of the molecule-pair kernel, counted exactly on set-1 rows
    # 0..19 against all of set 2 and scaled to the whole set: per neighbour
    # pair inside the Gaussian window 28 ops (window test, exp counted as
    # 20, radial products and sums) + 5 per element pair (two products, a
    # sum, the pd weight, the accumulation) + 2; atom pairs: 4 + 3 per sigma
    # (the exp of l2 counted as 20 is in the exponent stage, not here).
    def nt_of(X, n):
        out_ = np.zeros((X.shape[0], amax), dtype=np.int64)
        nb_ = np.zeros((X.shape[0], amax), dtype=np.int64)
        for a in range(X.shape[0]):
            for i in range(int(n[a])):
                k = int(np.sum(X[a, i, 0] < 5.0))
                nb_[a, i] = k
                out_[a, i] = len(set(X[a, i, 1, :k].astype(int).tolist()))
        return nb_, out_
    nb1, nt1 = nt_of(X1[:20], ms1.natoms[:20])
    nb2, nt2 = nt_of(X2, ms2.natoms)
    ops_sample = 0.0
    evals_sample = 0
    for i in range(20):
        for i1 in range(int(ms1.natoms[i])):
            d1 = X1[i, i1, 0, 1:nb1[i, i1]]
            for j in range(nm):
                nj = int(ms2.natoms[j])
                d2 = X2[j, :nj, 0, :]
                w = ((d2[:, None, :] - d1[None, :, None]) ** 2 < 2.56)
                mask = np.arange(amax)[None, None, :] >= 1
                mask = mask & (np.arange(amax)[None, None, :] < nb2[j, :nj, None, None])
                cnt = np.sum(w & mask, axis=(1, 2))
                evals_sample += int(cnt.sum())
                ops_sample += float(np.sum(cnt * (30 + 5 * nt1[i, i1] * nt2[j, :nj])))
    rows_frac = float(np.sum(ms1.natoms[:20])) / float(np.sum(ms1.natoms))
    ops = ops_sample / rows_frac
    evals = evals_sample / rows_frac
    units = len(sig) * nm * nm
    info = {"metric": "kernel entries/s", "unit": "entries/s", "dtype": "f64",
            "config": {"workload": f"global ARAS kernel, {nm}x{nm} molecules (5-23 atoms, seeds 0/1), "
                                   f"descriptors ({nm},{amax},{3 + amax},{amax}), sigmas {{0.5,2,10}}, "
                                   "t_width pi, d_width 0.2, cut 5, order 1",
                       "molecules": nm, "amax": amax,
                       "inner_evaluations_per_step_est": round(evals, 0)},
            "kernel": "k_aras_tiles"}

    def roofline(t):
        ach = ops / t / 1e12
        return {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP64_VALU_TOPS,
                "unit": "T FP64-VALU-ops/s", "frac": round(ach / PEAK_FP64_VALU_TOPS, 4), "traffic": None,
                "ops_per_step_est": round(ops, 0),
                "note": "ops counted on set-1 rows 0-19 against all of set 2, scaled by atom count"}

    def cpu_baseline(rows):
        import oracle
        if int(status.item()) != 0:
            raise RuntimeError("ARAS kernel reported bad descriptors")
        rows = max(1, min((rows or 100) // 4, nm))
        bkind = "reference" if oracle.ref_available() else "port"
        fn = oracle.ref_kernels_aras if bkind == "reference" else oracle.kernels_aras
        t0 = time.perf_counter()
        kc = fn(X1[:rows], X2, ms1.natoms[:rows], ms2.natoms, sig)
        dt = time.perf_counter() - t0
        g = np.transpose(out[:, :rows, :].cpu().numpy(), (2, 1, 0))
        m = kc > 1e-290
        err = float(np.max(np.abs(g[m] - kc[m]) / kc[m]))
        return ({"value": round(len(sig) * rows * nm / dt, 1), "unit": "entries/s",
                 "cores": host_threads(), "kind": bkind,
                 "sample": f"set-1 molecules 0..{rows - 1} x all {nm} set-2 molecules "
                           f"({len(sig) * rows * nm} entries), {dt:.2f} s"}, err)

    return info, step, units, roofline, cpu_baseline, None


def copy_workload(args):
    """Plain device-to-device copy of 512 MiB (torch's elementwise copy
    kernel): the known-size calibration for the HBM FETCH_SIZE / WRITE_SIZE
    counters."""
    import torch
    nbytes = 512 << 20
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)

    def step():
        dst.copy_(src)

    info = {"metric": "copy bandwidth", "unit": "GB/s", "dtype": "f64", "kernel": "copy",
            "config": {"workload": "device copy, 512 MiB read + 512 MiB written"}}

    def roofline(t):
        ach = 2 * nbytes / t / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": 2 * nbytes}

    return info, step, 2 * nbytes / 1e9, roofline, None, None


def cholesky_workload(args, kind):
    """Cholesky of K + lambda I for the 36k Gaussian kernel (lambda 1e-6):
    FP64 (cholesky64), FP32 of the shifted matrix (cholesky32) or the Ozaki
    form with its updates on INT8 cores (cholesky_oz, the factor of the Ozaki
    KRR fit)."""
    import torch
    from fml_amd.device import to_device
    from fml_amd.krr import add_diagonal, kernel_matrix
    from fml_amd.math.cho_solve import potrf_inplace, potrf_ozaki_inplace
    from fml_amd.representations import generate_coulomb_matrices
    n = args.n or 36000
    ms = _cm_set(n, 0, 3, 29, 29)
    X = generate_coulomb_matrices(to_device(ms.charges), to_device(ms.coords),
                                  to_device(ms.offsets, np.int32), nmax=29, on_device=True)
    K = kernel_matrix(X, X, 1000.0)
    add_diagonal(K, 1e-6)
    if kind == "cholesky32":
        K = K.to(torch.float32)
        idx = torch.arange(n, device="cuda")
        K[idx, idx] += 1e-3   # shift so the FP32 factor exists
    work = torch.empty_like(K)
    state = {}

    factor = potrf_ozaki_inplace if kind == "cholesky_oz" else potrf_inplace

    def step():
        work.copy_(K)
        state["info"] = factor(work, sync=False)

    info = {"metric": "Cholesky TFLOP/s", "unit": "TFLOP/s",
            "dtype": {"cholesky64": "f64", "cholesky32": "f32", "cholesky_oz": "i8+f64"}[kind],
            "config": {"workload": f"{kind}: potrf of the {n}x{n} Gaussian K + lambda I"}}
    flops = n ** 3 / 3.0

    def roofline(t):
        ach = flops / t / 1e12
        if kind == "cholesky_oz":
            # 15 INT8 digit products per FP64-equivalent multiply-add of the updates
            return {"bound": "mfma", "achieved": round(ach * 15 / 1e3, 3), "peak": PEAK_INT8_POPS,
                    "unit": "POP/s", "frac": round(ach * 15 / 1e3 / PEAK_INT8_POPS, 4), "traffic": None,
                    "note": "n^3/3 FP64-equivalent flop x 15 digit products over the whole factorisation"}
        peak = PEAK_FP64_TFLOPS if kind == "cholesky64" else PEAK_FP32_TFLOPS
        return {"bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 4), "traffic": None}

    def finish(res):
        res["potrf_info"] = int(state["info"].item())
        res["value"] = round(flops / (res["dominant_kernel_ms"] / 1e3) / 1e12, 3)

    return info, step, None, roofline, None, finish


def solve_workload(args, kind):
    """Triangular solves of cho_solve (dpotrs) with the factor of the 36k
    Gaussian K + lambda I: one right-hand side, forward and backward (solve64:
    FP64 factor; solve32: the FP32 factor of the shifted matrix, the mixed
    fit's preconditioner).  HBM-bound: each direction reads the factor's
    triangle once."""
    import torch
    from fml_amd.device import to_device
    from fml_amd.krr import add_diagonal, kernel_matrix
    from fml_amd.math.cho_solve import potrf_inplace
    from fml_amd.representations import generate_coulomb_matrices
    import fml_amd._native as nat
    n = args.n or 36000
    ms = _cm_set(n, 0, 3, 29, 29)
    X = generate_coulomb_matrices(to_device(ms.charges), to_device(ms.coords),
                                  to_device(ms.offsets, np.int32), nmax=29, on_device=True)
    K = kernel_matrix(X, X, 1000.0)
    add_diagonal(K, 1e-6)
    f32 = kind == "solve32"
    if f32:
        K = K.to(torch.float32)
        idx = torch.arange(n, device="cuda")
        K[idx, idx] += 1e-3
    assert potrf_inplace(K) == 0
    y = torch.from_numpy(ms.energies.astype(np.float32 if f32 else np.float64)).cuda()
    x = torch.empty_like(y)
    fn = "fmlh_spotrs" if f32 else "fmlh_potrs"
    wsb = getattr(nat.load(), fn + "_ws_bytes")(n, 1)
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device="cuda")

    def step():
        x.copy_(y)
        nat.call(fn, nat.ptr(K), n, n, nat.ptr(x), n, 1, nat.ptr(ws), wsb, nat.stream_ptr())
    step.graph_capable = True
    esz = 4 if f32 else 8
    nbytes = 2 * (n * (n + 1) // 2) * esz  # the triangle, once per direction
    info = {"metric": "triangular solve GB/s", "unit": "GB/s", "dtype": "f32" if f32 else "f64",
            "config": {"workload": f"{kind}: forward + backward solve with the {n}x{n} Cholesky factor, "
                                   "one right-hand side"}}

    def roofline(t):
        ach = nbytes / t / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None, "algorithmic_bytes": nbytes}

    def finish(res):
        res["value"] = round(nbytes / (res["ms_per_step"] / 1e3) / 1e9, 1)
        res["solution_head"] = x[:4].double().cpu().numpy().tolist()

    return info, step, None, roofline, None, finish


def inverse_workload(args):
    """cho_invert's second half at 36k (cho_solve.py:27-30 -> fcho_solve.f90:
    49-68, dpotri('U')): one step = one fmlh_potri of the FP64 Cholesky
    factor of the Gaussian K + lambda I (factored once, copied in each step):
    L^{-1} by blocked substitution, then L^{-T} L^{-1} on the lower triangle."""
    import torch
    import fml_amd._native as nat
    from fml_amd.device import to_device, workspace
    from fml_amd.krr import add_diagonal, kernel_matrix
    from fml_amd.math.cho_solve import potrf_inplace
    from fml_amd.representations import generate_coulomb_matrices
    n = args.n or 36000
    ms = _cm_set(n, 0, 3, 29, 29)
    X = generate_coulomb_matrices(to_device(ms.charges), to_device(ms.coords),
                                  to_device(ms.offsets, np.int32), nmax=29, on_device=True)
    A = kernel_matrix(X, X, 1000.0)
    add_diagonal(A, 1e-6)
    F = A.clone()
    info0 = potrf_inplace(F, sync=True)
    del X
    work = torch.empty_like(F)
    lib = nat.load()
    wsb = lib.fmlh_potri_ws_bytes(n)
    ws = workspace(wsb)
    infot = torch.zeros(1, dtype=torch.int32, device="cuda")

    def step():
        work.copy_(F)
        nat.call("fmlh_potri", nat.ptr(work), work.stride(0), n, nat.ptr(infot), nat.ptr(ws), wsb,
                 nat.stream_ptr())

    info = {"metric": "inverse TFLOP/s", "unit": "TFLOP/s", "dtype": "f64",
            "config": {"workload": f"inverse64: potri from the FP64 factor of the {n}x{n} Gaussian K + "
                                   "lambda I (sigma 1000, lambda 1e-6)"}}
    flops = 2.0 * n ** 3 / 3.0   # triangular inverse n^3/3 + L^-T L^-1 n^3/3

    def roofline(t):
        ach = flops / t / 1e12
        return {"bound": "mfma", "achieved": round(ach, 2), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP64_TFLOPS, 4), "traffic": None}

    def finish(res):
        res["potrf_info"] = int(info0 if isinstance(info0, int) else info0.item())
        res["potri_info"] = int(infot.item())
        res["value"] = round(flops / (res["dominant_kernel_ms"] / 1e3) / 1e12, 3)
        # a sample of rows of inv(A) A - I; inv(A) sits in work's row-major
        # lower triangle, so row r is work[r, :r+1] followed by work[r+1:, r]
        rows = torch.arange(0, n, max(1, n // 64), device="cuda")
        cols = torch.arange(n, device="cuda")
        inv_rows = torch.where(cols[None, :] <= rows[:, None], work[rows], work[:, rows].t())
        eye = (cols[None, :] == rows[:, None]).to(torch.float64)
        r = inv_rows @ A - eye
        res["sampled_rows_max_abs_inv_a_minus_i"] = float(r.abs().max().item())

    return info, step, None, roofline, None, finish


def _train_36k(n_total, n_train, precision="fp64", sigma=1000.0, lam=1e-6):
    """Coulomb features (nmax 29, 40k molecules seed 0 3-29 atoms) and a fitted model."""
    import torch
    from fml_amd.device import to_device
    from fml_amd.krr import KRR
    from fml_amd.representations import generate_coulomb_matrices
    ms = _cm_set(n_total, 0, 3, 29, 29)
    X = generate_coulomb_matrices(to_device(ms.charges), to_device(ms.coords),
                                  to_device(ms.offsets, np.int32), nmax=29, on_device=True)
    y = ms.energies
    m = KRR(sigma, lam, precision=precision).fit(X[:n_train], y[:n_train])
    torch.cuda.synchronize()
    return ms, X, y, m


def krr36k_workload(args, precision):
    """Precision form at scale: fit a 36,000-molecule Gaussian KRR (sigma
    1000, lambda 1e-6) and predict 4,000 held-out molecules, FP64 plain form
    vs the mixed form (FP32 factor preconditioning FP64-residual CG) or the
    Ozaki form (INT8-core factor refined with FP64 residuals)."""
    import torch
    from fml_amd.krr import KRR
    n = args.n or 40000
    ntr = int(n * 0.9)
    ms = _cm_set(n, 0, 3, 29, 29)
    from fml_amd.device import to_device
    from fml_amd.representations import generate_coulomb_matrices
    X = generate_coulomb_matrices(to_device(ms.charges), to_device(ms.coords),
                                  to_device(ms.offsets, np.int32), nmax=29, on_device=True)
    y = ms.energies
    state = {}

    def step():
        from fml_amd.krr import ConvergenceError
        try:
            m = KRR(1000.0, 1e-6, precision=precision, tol=args.tol, max_iter=args.max_iter).fit(X[:ntr], y[:ntr])
            state["error"] = None
        except ConvergenceError as e:  # reported, with the last iterate's predictions
            m = KRR(1000.0, 1e-6, precision=precision)
            m.X, m.alpha, m.report = X[:ntr].contiguous(), e.alpha, e.report
            state["error"] = str(e)
        state["m"] = m
        state["pred"] = m.predict(X[ntr:])

    info = {"metric": "KRR fit+predict time", "unit": "ms",
            "dtype": {"fp64": "f64", "mixed": "f32+f64", "ozaki": "i8+f64"}[precision],
            "higher_is_better": False,
            "config": {"workload": f"Gaussian KRR fit {ntr} + predict {n - ntr} molecules (nmax 29, seed 0), "
                                   f"sigma 1000, lambda 1e-6, precision {precision}"}}

    def finish(res):
        m = state["m"]
        pred = state["pred"].cpu().numpy()
        res["mae"] = float(np.mean(np.abs(pred - y[ntr:])))
        rep = m.report
        res["fit_breakdown_s"] = {"kernel": round(rep.kernel_s, 4), "factor": round(rep.factor_s, 4),
                                  "solve": round(rep.solve_s, 4), "total": round(rep.seconds, 4)}
        res["iterations"] = rep.iterations
        res["alpha_error_estimate"] = rep.alpha_error_estimate
        res["shift"] = rep.shift
        res["relative_residual"] = rep.residual
        res["converged"] = state["error"] is None
        if state["error"]:
            res["error"] = state["error"]
        res["alpha_head"] = m.alpha.cpu().numpy()[:8].tolist()
        if precision != "fp64":
            # the plain FP64 form on the same inputs: time, MAE and alpha
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            m64 = KRR(1000.0, 1e-6).fit(X[:ntr], y[:ntr], keep_kernel=True)
            p64 = m64.predict(X[ntr:])
            torch.cuda.synchronize()
            t64 = time.perf_counter() - t0
            mae64 = float(np.mean(np.abs(p64.cpu().numpy() - y[ntr:])))
            a64 = m64.alpha
            res["vs_fp64"] = {
                "fp64_fit_predict_ms": round(t64 * 1e3, 2),
                "speedup": round(t64 / (res["ms_per_step"] / 1e3), 3),
                "mae_fp64": mae64,
                "mae_rel_diff": abs(res["mae"] - mae64) / mae64,
                "alpha_rel_err": float((torch.linalg.norm(m.alpha - a64) / torch.linalg.norm(a64)).item()),
                "prediction_max_rel_diff": float(((state["pred"] - p64).abs().max() / p64.abs().max()).item())}
            # lambda_max(K + lambda I) by power iteration (its lambda_min is >= lambda: K is PSD)
            from fml_amd.krr import kernel_matrix, predict_from_kernel, refine_dd
            K = kernel_matrix(X[:ntr], X[:ntr], 1000.0)
            # the exact solution of the FP64 system: the FP64 factor's solution refined
            # with double-length residuals (independent of the form being measured)
            yd = torch.from_numpy(np.ascontiguousarray(y[:ntr])).cuda()
            ax = a64.clone()
            dx_ref, _, _ = refine_dd(K, 1e-6, m64.factor, yd, ax, tol=1e-14, max_iter=30)
            m64.factor = None
            axn = torch.linalg.norm(ax)
            res["vs_exact"] = {
                "reference": "the FP64 factor's solution refined with double-length residuals",
                "reference_steps": len(dx_ref), "reference_last_correction_rel": dx_ref[-1],
                "alpha_rel_err": float((torch.linalg.norm(m.alpha - ax) / axn).item()),
                "fp64_plain_alpha_rel_err": float((torch.linalg.norm(a64 - ax) / axn).item())}
            del ax
            v = torch.ones(ntr, dtype=torch.float64, device="cuda") / ntr ** 0.5
            w = torch.empty_like(v)
            lam_max = 0.0
            for _ in range(40):
                predict_from_kernel(K, v, out=w)
                w.add_(v, alpha=1e-6)
                lam_max = float(torch.dot(v, w).item())
                v = w / torch.linalg.norm(w)
            res["conditioning"] = {"lambda_max": lam_max, "lambda": 1e-6, "kappa_upper_bound": lam_max / 1e-6,
                                   "note": "lambda_max by 40 power iterations; K is positive semidefinite, so "
                                           "lambda_min(K + lambda I) >= lambda"}
            del K
        # the dominant kernel work: the Cholesky factorisation, n^3/3 flop
        fl = ntr ** 3 / 3.0
        ach = fl / rep.factor_s / 1e12
        if precision == "ozaki":
            # INT8 work: 15 digit products per FP64 multiply-add of the updates
            res["roofline"] = {"bound": "mfma", "achieved": round(ach * 15 / 1e3, 3), "peak": PEAK_INT8_POPS,
                               "unit": "POP/s", "frac": round(ach * 15 / 1e3 / PEAK_INT8_POPS, 4), "traffic": None,
                               "note": "the factorisation of K + lambda I with its updates on INT8 cores: "
                                       "n^3/3 FP64-equivalent flop x 15 digit products, timed inside the fit"}
        else:
            peak = PEAK_FP64_TFLOPS if precision == "fp64" else PEAK_FP32_TFLOPS
            res["roofline"] = {"bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                               "frac": round(ach / peak, 4), "traffic": None,
                               "note": f"the {'FP64' if precision == 'fp64' else 'FP32'} factorisation of "
                                       f"K + lambda I (n^3/3 flop), timed inside the fit"}

    def cpu_baseline(rows):
        # the reference on the host: Gaussian kernel and fcho_solve on a 12,000-molecule
        # block, extrapolated to 36k (kernel by entries, factorisation by n^3)
        import oracle
        kind = "reference" if oracle.ref_available() else "port"
        gk = oracle.ref_gaussian_kernel if kind == "reference" else oracle.gaussian_kernel
        cs = oracle.ref_cho_solve if kind == "reference" else (lambda A, b: oracle.cho_solve(A, b)[0])
        nb = 12000
        Xh = X[:nb].cpu().numpy()
        t0 = time.perf_counter()
        K = gk(Xh.T, Xh.T, 1000.0)
        tk = time.perf_counter() - t0
        K[np.diag_indices_from(K)] += 1e-6
        t0 = time.perf_counter()
        cs(K, y[:nb])
        tc = time.perf_counter() - t0
        est = tk * (ntr * ntr + ntr * (n - ntr)) / (nb * nb) + tc * (ntr / nb) ** 3
        return cpu_record(round(est * 1e3, 1), "ms", kind,
                          f"Gaussian kernel {nb}x{nb} ({tk:.2f} s) and fcho_solve n={nb} ({tc:.2f} s), "
                          f"one run each, extrapolated to fit {ntr} + predict {n - ntr} (kernel by entries, "
                          "factorisation by n^3)", [tk, tc]), None

    return info, step, None, None, cpu_baseline, finish


def serving_workload(args):
    """Serving loop: 36,000-molecule Gaussian KRR model resident in HBM;
    Poisson predict arrivals at 2,000 req/s, each 1-256 molecules
    (log-uniform) of 3-29 atoms; fit requests of 1,000-16,000 molecules
    interleaved, admitted by free HBM.  Reports p50/p99 latency."""
    import threading
    import torch
    from fml_amd import synth
    from fml_amd.serving import KRRServer, latency_percentiles
    ntr = args.n or 36000
    ms, X, y, m = _train_36k(ntr, ntr)
    pool = synth.generate(4000, seed=2, nmin=3, nmax=29)
    pool_mols = [pool.molecule(i) for i in range(pool.nmol)]
    rng = np.random.default_rng(2)
    duration = args.duration
    rate = float(getattr(args, "rate", 2000.0))
    srv = KRRServer(X[:ntr], m.alpha, 1000.0, nmax=29, max_batch_molecules=16384).start()
    # fit requests: sizes drawn up front (their own stream of numbers) and the
    # float32 training rows prepared on the host before the timed loop
    fit_sizes = [int(v) for v in np.random.default_rng(3).integers(1000, 16001, size=2)]
    X32 = X.cpu().numpy().astype(np.float32)
    y_h = np.asarray(y)
    # warm-up
    srv.submit([pool_mols[0][0]], [pool_mols[0][1]]).result(timeout=120)
    # the client keeps only the latency and completion time of each answered
    # request, as a remote client would: holding tens of thousands of futures
    # in this process would make the interpreter's full garbage collections,
    # which stop every thread, scan them all
    futs = collections.deque()
    lats, dones = [], []

    def harvest(block=False):
        while futs and (block or futs[0].done()):
            _, lat_s, t_done = futs.popleft().result(timeout=300)
            lats.append(lat_s)
            dones.append(t_done)
    t_start = time.perf_counter()
    t_next = t_start
    fit_times = [t_start + duration * f for f in (0.25, 0.6)]
    fit_futs = []
    nmol_total = 0
    while time.perf_counter() - t_start < duration:
        t_next += rng.exponential(1.0 / rate)
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        k = int(np.floor(np.exp(rng.uniform(0, np.log(257)))))
        k = min(max(k, 1), 256)
        idx = rng.integers(0, len(pool_mols), k)
        futs.append(srv.submit([pool_mols[i][0] for i in idx], [pool_mols[i][1] for i in idx]))
        nmol_total += k
        harvest()
        if fit_times and time.perf_counter() >= fit_times[0]:
            fit_times.pop(0)
            nf = fit_sizes.pop(0)
            fit_futs.append((nf, srv.submit_fit(X32[:nf], y_h[:nf], 1000.0, 1e-6)))
    harvest(block=True)
    t_end = time.perf_counter()
    fits = [(nf, f.result(timeout=600)) for nf, f in fit_futs]
    srv.close()
    lat = latency_percentiles(lats)
    # requests whose time in the server overlapped a fit, and the rest
    fit_iv = [(a, b) for _, a, b in srv.fit_log]
    during = [any(d - l < b and d > a for a, b in fit_iv) for l, d in zip(lats, dones)]
    lat_fit = latency_percentiles([l for l, f in zip(lats, during) if f]) if any(during) else None
    lat_nofit = latency_percentiles([l for l, f in zip(lats, during) if not f])
    wall = t_end - t_start
    # the dominant device work per batch is reading the resident training
    # rows (ntrain x 435 FP64) for the Gaussian block: HBM-bound
    d = int(X.shape[1])
    bsec = float(np.sum([b[2] for b in srv.batches]))
    bbytes = float(len(srv.batches)) * ntr * d * 8 + float(np.sum([b[1] for b in srv.batches])) * ntr * 8 * 2
    ach = bbytes / bsec / 1e9
    roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None,
            "note": "per batch: the training rows (ntrain x D x 8 B) plus the kernel block written and read "
                    "(2 x b x ntrain x 8 B), over the batches' host-measured wall time (staging, launch and "
                    "one event wait included)"}
    # CPU: the reference's routines predicting a sample of the same requests
    import oracle
    kind = "reference" if oracle.ref_available() else "port"
    cm = oracle.ref_coulomb_matrix if kind == "reference" else oracle.coulomb_matrix
    gk = oracle.ref_gaussian_kernel if kind == "reference" else oracle.gaussian_kernel
    Xh = X[:ntr].cpu().numpy()
    al = m.alpha.cpu().numpy()
    crng = np.random.default_rng(4)
    cts, cmols = [], 0
    for _ in range(12):
        k = min(max(int(np.floor(np.exp(crng.uniform(0, np.log(257))))), 1), 256)
        idx = crng.integers(0, len(pool_mols), k)
        t0 = time.perf_counter()
        Xq = np.array([cm(*pool_mols[i], 29) for i in idx])
        gk(Xh.T, Xq.T, 1000.0).T @ al
        cts.append(time.perf_counter() - t0)
        cmols += k
    cpu = cpu_record(round(float(np.median(cts)) * 1e3, 3), "ms (median per request)", kind,
                     f"12 requests of the same size law ({cmols} molecules), one at a time: Coulomb "
                     "matrices and the Gaussian block against the 36k training rows, then K^T alpha", cts)
    return {"metric": "serving p50/p99 predict latency", "value": round(lat["p99_ms"], 3), "unit": "ms (p99)",
            "p50_ms": round(lat["p50_ms"], 3), "p99_ms": round(lat["p99_ms"], 3), "mean_ms": round(lat["mean_ms"], 3),
            "requests": len(lats), "molecules": nmol_total, "wall_s": round(wall, 3),
            "throughput_req_s": round(len(lats) / wall, 1), "throughput_mol_s": round(nmol_total / wall, 1),
            "batches": len(srv.batches), "mean_batch_requests": round(float(np.mean([b[0] for b in srv.batches])), 2),
            "fits": [{"n": nf, "seconds": round(b - a, 3)} for nf, a, b in srv.fit_log],
            "overlapping_fits": lat_fit, "outside_fits": lat_nofit,
            "roofline": roof, "cpu_baseline": cpu,
            "higher_is_better": False, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"serving: {ntr}-molecule Gaussian KRR model resident, Poisson {rate:g} req/s, "
                                   "1-256 molecules per request, fit requests 1k-16k admitted by free HBM",
                       "duration_s": duration}}


def krr_forms_clause(args):
    """The metric's fit-time clause at the 36k configuration: Gaussian KRR
    fit of 36,000 molecules + prediction of 4,000 (nmax 29, sigma 1000,
    lambda 1e-6) in the plain FP64 form (timed after one untimed run) and
    in the Ozaki and mixed forms, each against the FP64 model and against
    the exact solution of the FP64 system (the FP64 factor's solution
    refined with double-length residuals: a reference independent of the
    INT8 factor).  Residuals are double-length (exact to working accuracy)."""
    import torch
    from fml_amd.device import to_device
    from fml_amd.krr import KRR, ConvergenceError, kernel_matrix, refine_dd, residual_dd
    from fml_amd.representations import generate_coulomb_matrices
    n, ntr = 40000, 36000
    ms = _cm_set(n, 0, 3, 29, 29)
    X = generate_coulomb_matrices(to_device(ms.charges), to_device(ms.coords),
                                  to_device(ms.offsets, np.int32), nmax=29, on_device=True)
    y = ms.energies

    def run(prec, keep=False):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        err = None
        try:
            m = KRR(1000.0, 1e-6, precision=prec, tol=args.tol, max_iter=args.max_iter).fit(
                X[:ntr], y[:ntr], keep_kernel=keep)
        except ConvergenceError as e:
            m = KRR(1000.0, 1e-6, precision=prec)
            m.X, m.alpha, m.report = X[:ntr].contiguous(), e.alpha, e.report
            err = str(e)
        p = m.predict(X[ntr:])
        torch.cuda.synchronize()
        return m, p, time.perf_counter() - t0, err

    run("fp64")  # first-call set-up outside the comparison
    m64, p64, t64, _ = run("fp64", keep=True)
    ph = p64.cpu().numpy()
    mae64 = float(np.mean(np.abs(ph - y[ntr:])))
    K = kernel_matrix(m64.X, m64.X, 1000.0)
    yd = torch.from_numpy(np.ascontiguousarray(y[:ntr])).cuda()
    r = torch.empty_like(yd)

    def rel_res(a):
        residual_dd(K, 1e-6, a, yd, r)
        return float((torch.linalg.norm(r) / torch.linalg.norm(yd)).item())

    # the exact solution of the FP64 system, from the FP64 factor
    ax = m64.alpha.clone()
    dx_ref, _, _ = refine_dd(K, 1e-6, m64.factor, yd, ax, tol=1e-14, max_iter=30)
    m64.factor = None
    axn = torch.linalg.norm(ax)

    def err_exact(a):
        return float((torch.linalg.norm(a - ax) / axn).item())

    out = {"workload": "Gaussian KRR fit 36000 + predict 4000 molecules (nmax 29, 40k seed 0), sigma 1000, "
                       "lambda 1e-6; times are fit + predict on HBM-resident features",
           "exact_alpha": {"reference": "the FP64 factor's solution refined with double-length residuals "
                                        "(fmlh_residual_dd), independent of the INT8 factor",
                           "steps": len(dx_ref), "last_correction_rel": dx_ref[-1]},
           "fp64": {"fit_predict_ms": round(t64 * 1e3, 2), "mae": mae64,
                    "relative_residual": rel_res(m64.alpha), "alpha_rel_err_vs_exact": err_exact(m64.alpha)}}
    run("ozaki")  # its first call's set-up (kernel loading, the digit planes' allocation) likewise
    for prec in ("ozaki", "mixed"):
        m, p, t, err = run(prec)
        pe = p.cpu().numpy()
        mae = float(np.mean(np.abs(pe - y[ntr:])))
        out[prec] = {"fit_predict_ms": round(t * 1e3, 2), "speedup_vs_fp64": round(t64 / t, 3),
                     "mae": mae, "mae_rel_diff": abs(mae - mae64) / mae64,
                     "alpha_rel_diff_vs_fp64": float((torch.linalg.norm(m.alpha - m64.alpha)
                                                      / torch.linalg.norm(m64.alpha)).item()),
                     "alpha_rel_err_vs_exact": err_exact(m.alpha),
                     "alpha_error_estimate": getattr(m.report, "alpha_error_estimate", None),
                     "relative_residual": rel_res(m.alpha), "iterations": m.report.iterations,
                     "converged": err is None}
        del m, p
    del K, m64, X, ax
    torch.cuda.empty_cache()
    return out


def metric_clauses(args):
    """The metric's other two clauses, measured in the same run as the
    headline kernel: KRR fit time against FP64 (36k) and serving p50/p99."""
    import torch
    out = {}
    for name, fn in (("krr_fit_vs_fp64", krr_forms_clause), ("serving", None)):
        print(f"[bench] clause {name} ...", file=sys.stderr, flush=True)
        try:
            if fn is not None:
                out[name] = fn(args)
            else:
                keys = ("p50_ms", "p99_ms", "mean_ms", "requests", "throughput_req_s", "mean_batch_requests",
                        "fits", "overlapping_fits", "outside_fits", "cpu_baseline", "config")
                sa = argparse.Namespace(**vars(args))
                sa.duration, sa.n, sa.rate = args.clause_duration, 0, 2000.0
                srv = serving_workload(sa)
                out[name] = {k: srv[k] for k in keys if k in srv}
                # the same loop at four times the arrival rate: continuous batching under load
                sa.duration, sa.rate = 2.0, 8000.0
                srv = serving_workload(sa)
                out[name]["at_8000_req_s"] = {k: srv[k] for k in keys[:8] if k in srv}
        except Exception as e:  # a clause that fails is reported, the headline line still prints
            out[name] = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); every GPU box here has "
              "one GPU and multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "serving":
        import torch
        torch.cuda.set_device(0)
        out = serving_workload(args)
        out.update({"n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": out["p50_ms"], "scaling": "weak",
                    "vs_baseline": None})
        print(json.dumps(out), flush=True)
        return
    import torch
    import fml_amd._native as nat
    from fml_amd.graphs import Graph
    torch.cuda.set_device(0)
    # a side stream: HIP graphs cannot be captured on the null stream
    torch.cuda.set_stream(torch.cuda.Stream())
    lib = nat.load()
    lib.fmlh_timing_enable.argtypes = [ctypes.c_int]
    lib.fmlh_timing_enable.restype = None
    lib.fmlh_timing_collect.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]

    wl = args.workload
    if wl == "arad":
        info, step, units, roofline, cpu_baseline, finish = arad_workload(args)
    elif wl.startswith("arad4096"):
        info, step, units, roofline, cpu_baseline, finish = arad_workload(
            args, big=True, precision=1 if wl.endswith("fp32") else 0)
    elif wl == "krr":
        info, step, units, roofline, cpu_baseline, finish = krr_workload(args)
    elif wl.startswith("krr36k"):
        info, step, units, roofline, cpu_baseline, finish = krr36k_workload(args, wl.split("_")[1])
    elif wl.startswith("solve"):
        info, step, units, roofline, cpu_baseline, finish = solve_workload(args, wl)
    elif wl == "inverse64":
        info, step, units, roofline, cpu_baseline, finish = inverse_workload(args)
    elif wl.startswith("cholesky"):
        info, step, units, roofline, cpu_baseline, finish = cholesky_workload(args, wl)
    elif wl == "copy":
        info, step, units, roofline, cpu_baseline, finish = copy_workload(args)
    elif wl == "aras":
        info, step, units, roofline, cpu_baseline, finish = aras_workload(args)
    elif wl.startswith("atomic"):
        info, step, units, roofline, cpu_baseline, finish = atomic_workload(
            args, {"atomic": 0, "atomic_lap": 1, "atomic_mfma": 2}[wl])
    else:
        info, step, units, roofline, cpu_baseline, finish = dense36k_workload(args, wl)
    print(f"[bench] workload ready: {info['config']['workload']}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.fmlh_timing_collect(None, None)
    lib.fmlh_timing_enable(1)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lib.fmlh_timing_enable(0)
    ms_total = e0.elapsed_time(e1)
    kms = ctypes.c_double(0.0)
    kcnt = ctypes.c_int(0)
    lib.fmlh_timing_collect(ctypes.byref(kms), ctypes.byref(kcnt))
    t_step = ms_total / 1e3 / args.steps
    t_kernel = kms.value / 1e3 / kcnt.value if kcnt.value else t_step
    print(f"[bench] {args.steps} steps: {ms_total:.3f} ms (wall {wall * 1e3:.1f} ms), "
          f"dominant kernel avg {t_kernel * 1e3:.3f} ms over {kcnt.value}", file=sys.stderr, flush=True)
    graph_ms = None
    if getattr(step, "graph_capable", False) and not args.eager:
        # the same step captured once into a HIP graph, replayed K times
        g = Graph.capture(step, stream=torch.cuda.current_stream())
        g.launch()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.steps):
            g.launch()
        e1.record()
        torch.cuda.synchronize()
        graph_ms = e0.elapsed_time(e1) / args.steps
        g.destroy()
        print(f"[bench] {args.steps} graph replays: {graph_ms:.3f} ms per step", file=sys.stderr, flush=True)

    res = {
        "metric": info["metric"],
        "value": round(units / t_step, 1) if units else round(t_step * 1e3, 4),
        "unit": info["unit"],
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": info.get("higher_is_better", True),
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": info["dtype"],
        "data": "synthetic",
        "config": info["config"],
        "roofline": (roofline(t_step if info.get("roofline_on_step") else t_kernel)
                     if roofline else None),
        "dominant_kernel_ms": round(t_kernel * 1e3, 4),
    }
    if graph_ms is not None:
        res["graph_replay"] = {"ms_per_step": round(graph_ms, 4),
                               "value": round(units / (graph_ms / 1e3), 1) if units else None,
                               "note": "the step captured once with fmlh_graph_begin/end, replayed per step"}
    if finish:
        finish(res)
    if not args.no_cpu and cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cb, err = cpu_baseline(args.cpu_rows)
        res["cpu_baseline"] = cb
        if isinstance(err, dict):
            res["parity_vs_cpu_baseline_max_rel"] = err["max_rel"]
            res["parity_vs_cpu_baseline"] = err
        elif err is not None:
            res["parity_vs_cpu_baseline_max_rel"] = err
    if wl == "arad" and not args.no_clauses:
        res["metric_clauses"] = metric_clauses(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
