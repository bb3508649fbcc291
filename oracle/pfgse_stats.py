"""PFGSE walker-sampling statistics restated in numpy (TEST INFRASTRUCTURE
ONLY: used by tests/ as the checker).

Follows NMR_PFGSE::recoverDsatWithSampling (src/backend/NMR_Simulation/
NMR_pfgse.cpp:876-1060 of RandomWalk_NMR) after its bulk / noise terms:
per-sample normalisation by M(k = 0), LHS = log(max(ratio, 1e-6)), mean and
population standard deviation over samples, a no-intercept least-squares
slope per sample over the first n_points points (LeastSquareAdjust.cpp:77-170)
with error sqrt(MSE n / (n - 1)) (NMR_pfgse.cpp:1027-1028).  Pinned in
tests/test_host.py by identical samples (zero spread, single fit) and exact
Gaussian attenuation (slope recovered exactly).
"""
import numpy as np


def sample_stats(M_samples, rhs, n_points):
    Ms = np.asarray(M_samples, float)
    ratio = (Ms / Ms[:, :1]) / (Ms[:, :1] / Ms[:, :1])
    lhs = np.where(ratio > 1e-6, np.log(np.maximum(ratio, 1e-300)), np.log(1e-6))
    out = dict(mean_M=Ms.mean(axis=0), std_M=Ms.std(axis=0), mean_lhs=lhs.mean(axis=0), std_lhs=lhs.std(axis=0))
    r = np.asarray(rhs, float)[:n_points]
    D, err = [], []
    for s in range(Ms.shape[0]):
        y = lhs[s, :n_points]
        b = float(np.dot(r, y) / np.dot(r, r))
        mse = float(np.mean((y - b * r) ** 2))
        D.append(b)
        err.append(np.sqrt(mse * n_points / (n_points - 1.0)))
    D = np.array(D)
    out.update(D_samples=D, D_mean=float(D.mean()), D_std=float(D.std()), D_err_mean=float(np.mean(err)))
    return out
