"""Kernel ridge regression composed from the hot-path pieces, device resident.

The reference has no KRR driver; users chain the routines (README.md:33-69):
representations -> kernel -> K + lambda I -> cho_solve -> K_test^T alpha.
This module is that chain kept on the GPU end to end:

* plain form (``precision="fp64"``): FP64 Gaussian/Laplacian kernel, FP64
  blocked Cholesky (MFMA trailing updates), FP64 triangular solves;
* mixed form (``precision="mixed"``): the FP64 kernel matrix is kept for the
  residual, an FP32 copy (optionally shifted until it factors) is factored
  on FP32 MFMA, and the FP64 system is solved by conjugate gradients
  preconditioned with that FP32 factor, every residual and update in FP64
  (iterative refinement is the special case of one-step convergence);
* Ozaki form (``precision="ozaki"``): K + lambda I is factored in place in
  K's lower triangle with its large updates on INT8 matrix cores
  (fmlh_potrf_ozaki: an exact five-digit split of each panel, the factor
  within ~1e-12 of the FP64 one), K itself staying in the other triangle
  and its diagonal aside; alpha is then refined with residuals
  y - (K + lambda I) alpha computed in double length from that triangle
  (fmlh_residual_dd_lower) and corrections from the factor
  (fmlh_potrs_correction) until alpha's estimated error falls below tol:
  alpha converges to the solution of the FP64 system to working accuracy,
  where a plain FP64 solve is off by up to kappa times the unit roundoff.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .device import empty, to_device, torch, workspace
from .kernels.kernels import gaussian_kernel_rows, laplacian_kernel_rows
from .math.cho_solve import potrf_inplace, potrf_ozaki_inplace, potrs_correction_inplace, potrs_inplace


def kernel_matrix(X1, X2, sigma, kernel="gaussian", precision="fp64", out=None, out_dtype=None,
                  recomputed=None):
    """Device tensor (n2, n1) whose [i, j] is K(x1_j, x2_i) (the memory of
    the reference's Fortran (n1, n2) kernel)."""
    if kernel == "gaussian":
        return gaussian_kernel_rows(X1, X2, sigma, precision=precision, out=out, out_dtype=out_dtype,
                                    recomputed=recomputed)
    if kernel == "laplacian":
        return laplacian_kernel_rows(X1, X2, sigma, out=out, out_dtype=out_dtype)
    raise ValueError(kernel)


def add_diagonal(K, lam):
    n = int(K.shape[0])
    _native.call("fmlh_add_diagonal", _native.ptr(K), K.stride(0), n, float(lam), _native.stream_ptr())


def sym_residual(K, lam, x, y, r):
    """r = y - (K + lam I) x from the upper triangle of the (Fortran) K."""
    n = int(K.shape[0])
    _native.call("fmlh_sym_residual", _native.ptr(K), K.stride(0), n, float(lam), _native.ptr(x),
                 _native.ptr(y), _native.ptr(r), _native.stream_ptr())
    return r


def residual_dd(K, lam, x, y, r):
    """r = y - (K + lam I) x for the full symmetric device kernel K (n, n),
    each entry accumulated in double length and rounded once."""
    n = int(K.shape[0])
    _native.call("fmlh_residual_dd", _native.ptr(K), K.stride(0), n, float(lam), _native.ptr(x),
                 _native.ptr(y), _native.ptr(r), _native.stream_ptr())
    return r


def residual_dd_lower(A, diag, lam, x, y, r):
    """r = y - (K + lam I) x for K held as A's strictly upper triangle (the
    row-major view of the Fortran strictly lower one) plus its diagonal
    `diag`; A's lower triangle (where a factor of K + lam I may sit) is not
    read.  Double length as residual_dd."""
    n = int(A.shape[0])
    wsb = _native.load().fmlh_residual_dd_lower_ws_bytes(n)
    ws = workspace(wsb)
    _native.call("fmlh_residual_dd_lower", _native.ptr(A), A.stride(0), n, _native.ptr(diag), float(lam),
                 _native.ptr(x), _native.ptr(y), _native.ptr(r), _native.ptr(ws), wsb, _native.stream_ptr())
    return r


def refine_dd(K, lam, F, y, x, tol=1e-12, max_iter=30, diag=None):
    """Iterative refinement of x in place for (K + lam I) x = y with the
    factor F (fmlh_potrf layout) as the solver (its correction solve, the
    diagonal blocks applied as inverses) and double-length residuals:
    x += F^-1 (y - (K + lam I) x).  The corrections shrink geometrically, by
    rho = |d_k| / |d_(k-1)| a step, so the error left after a step is about
    |d_k| rho / (1 - rho); refinement stops when that estimate (or the
    correction itself) is at most tol times |x|, or when a correction no
    longer halves (the working-precision floor).  Returns (history of
    |correction| / |x|, history of |residual| / |y|, the final error
    estimate), one host read per step.  diag: K's diagonal when K holds the
    matrix in its strictly upper triangle only (F may then be K itself, its
    lower triangle the factor): residuals by residual_dd_lower."""
    t = torch()
    r = empty((int(K.shape[0]),), np.float64)
    ynorm = t.linalg.norm(y)
    dx, res = [], []
    est = float("inf")
    for _ in range(max_iter):
        if diag is None:
            residual_dd(K, lam, x, y, r)
        else:
            residual_dd_lower(K, diag, lam, x, y, r)
        rn = t.linalg.norm(r) / ynorm
        potrs_correction_inplace(F, r)
        x.add_(r)
        v = t.stack([t.linalg.norm(r) / t.linalg.norm(x), rn]).cpu()  # one host read per step
        dx.append(float(v[0]))
        res.append(float(v[1]))
        if len(dx) > 1:
            rho = dx[-1] / dx[-2] if dx[-2] > 0 else 0.0
            if rho > 0.5:  # no longer converging: the floor
                est = dx[-1]
                break
            est = dx[-1] * rho / (1.0 - rho)
        if dx[-1] <= tol or est <= tol:
            est = min(est, dx[-1])
            break
    return dx, res, est


def predict_from_kernel(Kt, alpha, out=None):
    """y_i = sum_j K(j, i) alpha_j for the device tensor Kt (nt, n)."""
    nt, n = int(Kt.shape[0]), int(Kt.shape[1])
    y = out if out is not None else empty((nt,), np.float64)
    _native.call("fmlh_gemv_t", _native.ptr(Kt), Kt.stride(0), n, nt, _native.ptr(alpha), _native.ptr(y),
                 _native.stream_ptr())
    return y


@dataclass
class FitReport:
    precision: str
    n: int
    seconds: float = 0.0
    kernel_s: float = 0.0
    factor_s: float = 0.0
    solve_s: float = 0.0
    iterations: int = 0
    shift: float = 0.0
    backward_error: float = float("nan")
    residual: float = float("nan")
    alpha_error_estimate: float = float("nan")
    history: list = field(default_factory=list)


class ConvergenceError(RuntimeError):
    """The mixed-precision fit stopped short of its tolerance; .alpha holds
    the last iterate and .report the fit's figures."""

    def __init__(self, msg, alpha=None, report=None):
        super().__init__(msg)
        self.alpha = alpha
        self.report = report


class KRR:
    """Device-resident KRR model.  X rows are samples (n, d)."""

    def __init__(self, sigma, lam, kernel="gaussian", precision="fp64", tol=1e-12, max_iter=400,
                 check_every=10):
        self.sigma = float(sigma)
        self.lam = float(lam)
        self.kernel = kernel
        self.precision = precision
        self.tol = tol
        self.max_iter = max_iter
        self.check_every = max(1, int(check_every))
        self.alpha = None
        self.X = None
        self.report = None

    # -- fit ---------------------------------------------------------------
    def fit(self, X, y, keep_kernel=False):
        t = torch()
        self.X = to_device(X, np.float64)
        yd = to_device(y, np.float64)
        n = int(self.X.shape[0])
        rep = FitReport(self.precision, n)
        t.cuda.current_stream().synchronize()
        t0 = time.perf_counter()
        K = kernel_matrix(self.X, self.X, self.sigma, self.kernel)
        t.cuda.current_stream().synchronize()
        rep.kernel_s = time.perf_counter() - t0
        if self.precision == "fp64":
            t1 = time.perf_counter()
            add_diagonal(K, self.lam)
            info = potrf_inplace(K)
            t.cuda.current_stream().synchronize()
            rep.factor_s = time.perf_counter() - t1
            if info != 0:
                raise np.linalg.LinAlgError(f"K + lambda I is not positive definite (leading minor {info})")
            t2 = time.perf_counter()
            alpha = yd.clone()
            potrs_inplace(K, alpha)
            t.cuda.current_stream().synchronize()
            rep.solve_s = time.perf_counter() - t2
            self.factor = K if keep_kernel else None
        elif self.precision == "mixed":
            alpha = self._fit_mixed(K, yd, rep)
        elif self.precision == "ozaki":
            alpha = self._fit_ozaki(K, yd, rep)
        else:
            raise ValueError(self.precision)
        rep.seconds = time.perf_counter() - t0
        self.alpha = alpha
        self.report = rep
        return self

    def _fit_mixed(self, K, yd, rep):
        """FP32 factor of K + (lambda + mu) I (mu the smallest shift, from 0
        up, with which it factors) as the preconditioner of conjugate
        gradients on the FP64 system (K + lambda I) alpha = y, every residual
        and update in FP64.  The iteration stays on the device: step lengths
        are device scalars, and the host reads the residual norm once every
        `check_every` iterations.  Raises ConvergenceError when the
        residual has not reached tol * |y| after max_iter iterations (or
        the preconditioned operator stops being positive definite)."""
        t = torch()
        n = int(K.shape[0])
        t1 = time.perf_counter()
        K32 = K.to(t.float32)
        shift = 0.0
        idx = t.arange(n, device="cuda")
        dmax = float(K32.diagonal().abs().max().item())
        for attempt in range(10):
            F = K32.clone()  # potrf overwrites its argument
            F[idx, idx] += self.lam + shift
            info = potrf_inplace(F)
            if info == 0:
                break
            shift = max(shift * 10.0, 1e-6 * dmax)
        else:
            raise np.linalg.LinAlgError("FP32 factor failed even with a shift")
        del K32
        t.cuda.current_stream().synchronize()
        rep.factor_s = time.perf_counter() - t1
        rep.shift = shift
        t2 = time.perf_counter()
        z32 = t.empty(n, dtype=t.float32, device="cuda")

        def precond(r, out):
            z32.copy_(r)
            potrs_inplace(F, z32)
            out.copy_(z32)
            return out

        # the products with K + lam I read the full (symmetric) FP64 kernel row by
        # row, coalesced: K x + lam x
        def apply(v, out):
            predict_from_kernel(K, v, out=out)
            out.add_(v, alpha=self.lam)
            return out

        def residual(v, out):  # y - (K + lam I) v
            apply(v, out)
            out.neg_().add_(yd)
            return out

        # preconditioned CG on (K + lam I) x = y, FP64 residuals
        x = empty((n,), np.float64)
        precond(yd, x)
        r = empty((n,), np.float64)
        residual(x, r)
        ynorm = t.linalg.norm(yd)
        z = empty((n,), np.float64)
        precond(r, z)
        p = z.clone()
        rz = t.dot(r, z)
        Ap = empty((n,), np.float64)
        tol = float(self.tol)
        it = 0
        converged = False
        pap_ok = t.ones((), dtype=t.bool, device="cuda")
        for it in range(1, self.max_iter + 1):
            apply(p, Ap)
            pAp = t.dot(p, Ap)
            pap_ok &= pAp > 0
            a = rz / pAp
            x.add_(p * a)
            r.sub_(Ap * a)
            if it % self.check_every == 0 or it == self.max_iter:
                residual(x, r)  # re-anchor on the true FP64 residual
                rel = float((t.linalg.norm(r) / ynorm).item())  # one host read per check
                rep.history.append(rel)
                if not bool(pap_ok.item()):
                    break
                if rel <= tol:
                    converged = True
                    break
            precond(r, z)
            rz_new = t.dot(r, z)
            p.mul_(rz_new / rz).add_(z)
            rz = rz_new
        rep.iterations = it
        t.cuda.current_stream().synchronize()
        rep.solve_s = time.perf_counter() - t2
        residual(x, r)
        rep.residual = float((t.linalg.norm(r) / ynorm).item())
        knorm = float(t.linalg.matrix_norm(K, ord="fro").item()) if n <= 20000 else float("nan")
        rep.backward_error = float(t.linalg.norm(r).item()) / (knorm * float(t.linalg.norm(x).item()))
        if not converged:
            why = ("the preconditioned operator was not positive definite" if not bool(pap_ok.item())
                   else f"{it} iterations")
            raise ConvergenceError(
                f"mixed-precision fit did not reach |y - (K + lambda I) alpha| <= {tol:g} |y| "
                f"({why}; relative residual {rep.residual:.3g}, FP32 shift {shift:g}); "
                "raise max_iter or use precision='fp64'", alpha=x, report=rep)
        return x

    def _fit_ozaki(self, K, yd, rep):
        """Factor of K + lambda I with INT8-core updates, in place in K's lower
        triangle (K itself stays in the strictly upper one and its diagonal
        aside, for the residuals), then iterative refinement with double-length
        residuals (refine_dd) until alpha's correction is at most tol times
        |alpha|.  Raises ConvergenceError if the refinement stalls (or runs
        out of steps) with a correction above 1e-8 |alpha|: the factor is then
        too inexact for the system's conditioning."""
        t = torch()
        t1 = time.perf_counter()
        # one matrix for both: the factor of K + lambda I overwrites the lower
        # triangle (row-major view) in place, K stays in the strictly upper
        # one, and its diagonal is kept aside for the residuals
        d = K.diagonal().clone()
        add_diagonal(K, self.lam)
        info = potrf_ozaki_inplace(K)
        t.cuda.current_stream().synchronize()
        rep.factor_s = time.perf_counter() - t1
        if info != 0:
            raise np.linalg.LinAlgError(f"K + lambda I is not positive definite (leading minor {info})")
        t2 = time.perf_counter()
        x = yd.clone()
        potrs_correction_inplace(K, x)
        dx, res, est = refine_dd(K, self.lam, K, yd, x, tol=float(self.tol), max_iter=self.max_iter, diag=d)
        rep.history = dx
        rep.iterations = len(dx)
        rep.residual = res[-1] if res else float("nan")
        rep.alpha_error_estimate = est
        t.cuda.current_stream().synchronize()
        rep.solve_s = time.perf_counter() - t2
        if not dx or est > max(float(self.tol), 1e-8):
            raise ConvergenceError(
                f"Ozaki-factor refinement stopped with alpha's estimated error at {est:.3g} of |alpha| "
                f"after {rep.iterations} steps (the factor is too inexact for this system's "
                "conditioning); use precision='fp64'", alpha=x, report=rep)
        return x

    # -- predict -----------------------------------------------------------
    def predict(self, Xt, precision="fp64"):
        Xd = to_device(Xt, np.float64)
        Kt = kernel_matrix(self.X, Xd, self.sigma, self.kernel)
        return predict_from_kernel(Kt, self.alpha)
