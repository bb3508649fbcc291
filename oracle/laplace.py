"""Laplace (T2) inversion oracle in numpy/scipy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Restates
B/Laplace/src/nmrinv_core.cpp, B/Laplace/src/TikhonovSolver.cpp and
B/Laplace/src/vector_funcs.cpp of RandomWalk_NMR; the NNLS step is solved
with scipy.optimize.nnls (Lawson-Hanson) on the same augmented system
[K; lambda I] x = [b; 0] that B/Laplace/include/NNLS.hpp solves.

Eigen's JacobiSVD (TikhonovSolver.cpp:18-20) is replaced by LAPACK's SVD:
the L-curve only uses squares of U^T b and the singular values, so the
singular vectors' sign convention does not matter.
"""
from __future__ import annotations

import math
import numpy as np
from scipy.optimize import nnls as _scipy_nnls


def linspace_accumulated(start: float, end: float, size: int, include_final: bool = True):
    """value += step accumulation (B/Laplace/src/vector_funcs.cpp:85-100)."""
    step = (end - start) / (size - 1) if include_final else (end - start) / size
    out = np.empty(size)
    v = start
    for i in range(size):
        out[i] = v
        v += step
    return out


def logspace_accumulated(start: float, end: float, size: int, include_final: bool = True):
    """10**linspace_accumulated (B/Laplace/src/vector_funcs.cpp:102-107), by
    the C library's pow as the reference (numpy's vectorised power differs
    from it by an ulp on some values)."""
    return np.array([math.pow(10.0, v) for v in linspace_accumulated(start, end, size, include_final)])


def t2_bins(min_t2: float, max_t2: float, num: int, use_log: bool = True):
    """B/Laplace/src/nmrinv_core.cpp:129-141."""
    if use_log:
        return logspace_accumulated(np.log10(min_t2), np.log10(max_t2), num)
    return linspace_accumulated(min_t2, max_t2, num)


def kernel_matrix(raw_bins, t2b, lam: float = 0.0):
    """K_ij = exp(-t_i / T2_j), optionally stacked over lam * I
    (B/Laplace/src/nmrinv_core.cpp:99-126, 317-336)."""
    raw_bins = np.asarray(raw_bins, float)
    t2b = np.asarray(t2b, float)
    K = np.exp(-raw_bins[:, None] / t2b[None, :])
    if lam:
        K = np.vstack([K, lam * np.eye(len(t2b))])
    return K


def pruning_slices(raw_bins, num_pruning: int):
    """Log-spaced pruning slices (B/Laplace/src/nmrinv_core.cpp:145-179).
    Returns (start_index, [(p_index, begin, end), ...])."""
    raw_bins = np.asarray(raw_bins, float)
    n = len(raw_bins)
    logb = np.log10(raw_bins)
    lo = logb[0]
    diff = (logb[n - 1] - lo) / (num_pruning - 1)
    newb = np.array([lo + diff * i for i in range(num_pruning)])
    p = 1
    while (10.0 ** newb[p]) - raw_bins[p] <= 0.0:
        p += 1
    p -= 1
    ib = 0
    while logb[ib] - newb[p] <= 0.0:
        ib += 1
    slices = []
    for pi in range(p, num_pruning - 1):
        start = ib
        bi, bf = newb[pi], newb[pi + 1]
        while ib < n and logb[ib] >= bi and logb[ib] < bf:
            ib += 1
        slices.append((pi, start, ib))
    return p, slices


def prune(data, num_pruning: int, start: int, slices, use_log: bool):
    """B/Laplace/src/nmrinv_core.cpp:182-239 for a single curve."""
    data = np.asarray(data, float)
    out = np.zeros(num_pruning)
    out[:start] = data[:start]
    with np.errstate(invalid="ignore", divide="ignore"):
        for pi, b, e in slices:
            seg = data[b:e]
            if use_log:
                out[pi] = 10.0 ** (np.sum(np.log10(seg)) / (e - b)) if e > b else np.nan
            else:
                out[pi] = np.sum(seg) / (e - b) if e > b else np.nan
    out[num_pruning - 1] = data[-1]
    return out


def lcurve(K, b, min_lambda=-4.0, max_lambda=2.0, num_lambdas=512):
    """Hansen's L-curve curvature over log-spaced lambdas
    (B/Laplace/src/TikhonovSolver.cpp:14-116).  Returns dict with lambdas,
    solution_norms, residual_norms, curvature, best_lambda, best_index.

    The curvature keeps the reference's denominator (lam^2 eta^2 + rho^2)^1.5
    (TikhonovSolver.cpp:111), where Hansen's closed form has lam^4 eta^2, so
    that the same lambda is chosen; with lam^4 the expression is exactly twice
    the geometric curvature of (ln rho, ln eta)."""
    U, s, _ = np.linalg.svd(np.asarray(K, float), full_matrices=False)
    keep = 0
    while keep < len(s) and s[keep] != 0.0:
        keep += 1
    s = s[:keep]
    uatb = U[:, :keep].T @ np.asarray(b, float)
    lambdas = logspace_accumulated(min_lambda, max_lambda, num_lambdas)
    q = uatb / s
    q2 = q * q
    sol = np.empty(num_lambdas)
    res = np.empty(num_lambdas)
    curv = np.empty(num_lambdas)
    s2 = s * s
    for il, lam in enumerate(lambdas):
        sql = lam * lam
        f = s2 / (s2 + sql)
        omf = 1.0 - f
        eta = np.sum((f * q) ** 2)
        rho = np.sum((omf * uatb) ** 2)
        deta = np.sum(omf * f * f * q2) * (-4.0 / lam)
        mult = -2.0 * eta * rho / deta
        num = sql * deta * rho + 2.0 * lam * eta * rho + sql * sql * eta * deta
        den = (sql * eta * eta + rho * rho) ** 1.5
        sol[il] = np.sqrt(eta)
        res[il] = np.sqrt(rho)
        curv[il] = mult * num / den
    best = int(np.argmax(curv))
    return dict(lambdas=lambdas, solution_norms=sol, residual_norms=res, curvature=curv,
                best_index=best, best_lambda=float(lambdas[best]))


def nnls_tikhonov(K, b, lam: float):
    """argmin_{x>=0} |[K; lam I] x - [b; 0]|  (nmrinv_core.cpp:402-480)."""
    K = np.asarray(K, float)
    A = kernel_augment(K, lam)
    rhs = np.concatenate([np.asarray(b, float), np.zeros(A.shape[0] - K.shape[0])])
    x, _ = _scipy_nnls(A, rhs, maxiter=50 * A.shape[1])
    return x


def kernel_augment(K, lam):
    if not lam:
        return K
    return np.vstack([K, lam * np.eye(K.shape[1])])


def invert_decay(times, decay, *, min_t2=0.1, max_t2=1e4, num_t2=128, use_log=True,
                 prune_num=512, lam=None, min_lambda=-4.0, max_lambda=2.0, num_lambdas=512):
    """The CPMG inversion pipeline as NMR_cpmg::applyLaplace runs it
    (B/NMR_Simulation/NMR_cpmg.cpp:387-453): the caller has already dropped
    the t = 0 point.  lam=None selects lambda by the L-curve."""
    times = np.asarray(times, float)
    decay = np.asarray(decay, float)
    bins = t2_bins(min_t2, max_t2, num_t2, use_log)
    if prune_num and prune_num < len(times):
        start, sl = pruning_slices(times, prune_num)
        used_t = prune(times, prune_num, start, sl, True)
        used_b = prune(decay, prune_num, start, sl, False)
    else:
        used_t, used_b = times, decay
    K = kernel_matrix(used_t, bins)
    lc = None
    if lam is None:
        lc = lcurve(K, used_b, min_lambda, max_lambda, num_lambdas)
        lam = lc["best_lambda"]
    amps = nnls_tikhonov(K, used_b, lam)
    return dict(t2_bins=bins, t2_amps=amps, used_times=used_t, used_decay=used_b, lam=lam,
                lcurve=lc)


def stejskal_tanner(rhs, lhs, n_points=None):
    """No-intercept least squares D = sum(x y) / sum(x^2) with the standard
    error sqrt(MSE n/(n-1)) (B/Math/LeastSquareAdjust.cpp:77-170,
    B/NMR_Simulation/NMR_pfgse.cpp:641-692)."""
    x = np.asarray(rhs, float)
    y = np.asarray(lhs, float)
    if n_points is not None:
        x, y = x[:n_points], y[:n_points]
    n = len(x)
    # sequential sums in index order, as the reference's loops
    sxy = sxx = 0.0
    for xi, yi in zip(x.tolist(), y.tolist()):
        sxy += xi * yi
        sxx += xi * xi
    B = sxy / sxx
    res = 0.0
    for xi, yi in zip(x.tolist(), y.tolist()):
        res += (yi - (B * xi + 0.0)) ** 2
    mse = res / n
    return float(B), float(math.sqrt(mse * (n / (n - 1.0))))
