"""Exactly rounded residuals of a linear system, for checking refined
solutions: error-free products (Dekker) summed by math.fsum, so each entry
is the correctly rounded value of the exact residual of the FP64 data."""
import numpy as np


def two_product(a, b):
    """Dekker's error-free product: a * b = p + e exactly (no overflow)."""
    split = 134217729.0  # 2^27 + 1
    def halves(v):
        c = split * v
        hi = c - (c - v)
        return hi, v - hi
    ah, al = halves(a)
    bh, bl = halves(b)
    p = a * b
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl
    return p, e


def exact_residual(A, lam, x, y):
    """y - (A + lam I) x with every entry the correctly rounded value of the
    exact sum (error-free products, math.fsum)."""
    import math
    p, e = two_product(A, x[None, :])
    lp, le = two_product(np.full_like(x, lam), x)
    return np.array([math.fsum(np.concatenate(([y[i], -lp[i], -le[i]], -p[i], -e[i])))
                     for i in range(A.shape[0])])


def exact_refined_solution(A, lam, y, steps=8):
    """The solution of (A + lam I) x = y for the FP64 matrix A: LAPACK's
    Cholesky solution refined with exactly rounded residuals (its limiting
    accuracy is the working precision, not kappa times it)."""
    import scipy.linalg as sl
    cf = sl.cho_factor(A + lam * np.eye(A.shape[0]))
    x = sl.cho_solve(cf, y)
    for _ in range(steps):
        x = x + sl.cho_solve(cf, exact_residual(A, lam, x, y))
    return x
