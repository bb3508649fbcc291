"""T2 inversion through the reference-named C API of librwnmr_amd.so
(initialize_inverter / find_best_lambda / set_lambda / invert ...,
B/Laplace/include/nmrinv.h:13-39), plus the CPMG pipeline step
NMR_cpmg::applyLaplace (B/NMR_Simulation/NMR_cpmg.cpp:387-453).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from ._lib import load


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


class Inverter:
    """Thin object view of the library's global (mutex-guarded) inverter."""

    def __init__(self, raw_bins, *, min_t2=0.1, max_t2=1e4, num_t2_bins=128, num_lambdas=512, prune_num=512,
                 noise_amp=0.0, min_lambda=-4.0, max_lambda=2.0, t2_bins=None):
        self.L = load()
        rb = np.ascontiguousarray(raw_bins, dtype=np.float64)
        self.num_raw = len(rb)
        self.L.rwnmr_inverter_set_lambda_range(float(min_lambda), float(max_lambda))
        if t2_bins is None:
            self.L.initialize_inverter(len(rb), _p(rb), float(min_t2), float(max_t2), int(num_t2_bins),
                                       int(num_lambdas), int(prune_num), float(noise_amp))
        else:
            tb = np.ascontiguousarray(t2_bins, dtype=np.float64)
            self.L.initialize_inverter_with_t2_bins(len(rb), _p(rb), len(tb), _p(tb), int(num_lambdas),
                                                    int(prune_num), float(noise_amp))

    @property
    def t2_bins(self):
        out = np.zeros(self.L.rwnmr_inverter_num_t2_bins())
        self.L.get_used_t2_bins(_p(out))
        return out

    @property
    def used_raw_bins(self):
        out = np.zeros(self.L.rwnmr_inverter_num_used_raw_bins())
        self.L.get_used_raw_bins(_p(out))
        return out

    def find_best_lambda(self, amps):
        a = np.ascontiguousarray(amps, dtype=np.float64)
        return float(self.L.find_best_lambda(len(a), _p(a)))

    def set_lambda(self, lam):
        self.L.set_lambda(float(lam))

    @property
    def lam(self):
        return float(self.L.get_tikhonov_lambda())

    def processed(self, amps):
        a = np.ascontiguousarray(amps, dtype=np.float64)
        out = np.zeros((len(a) // self.num_raw) * self.L.rwnmr_inverter_num_used_raw_bins())
        self.L.get_processed_raw_amps(len(a), _p(a), _p(out))
        return out

    def invert(self, amps):
        a = np.ascontiguousarray(amps, dtype=np.float64)
        nd = len(a) // self.num_raw
        out = np.zeros(nd * self.L.rwnmr_inverter_num_t2_bins())
        if not self.L.invert(len(a), _p(a), _p(out)):
            raise RuntimeError("inversion failed (lambda not set, or NNLS did not converge)")
        return out if nd > 1 else out.copy()

    def lcurve(self):
        n = self.L.rwnmr_inverter_num_lambdas()
        sn, rn, la, cu = (np.zeros(n) for _ in range(4))
        if not self.L.get_lcurve_values(_p(sn), _p(rn), _p(la), _p(cu)):
            return None
        return dict(solution_norms=sn, residual_norms=rn, lambdas=la, curvature=cu)


def apply_laplace(times, signal, *, min_t2=0.1, max_t2=1e4, num_t2_bins=128, min_lambda=-4.0, max_lambda=2.0,
                  num_lambdas=512, prune_num=512, noise_amp=0.0, lam=None):
    """NMR_cpmg::applyLaplace: drop the t = 0 point, set up the inverter,
    choose lambda by the L-curve (or use `lam`), invert.  Returns
    (t2_bins, t2_amps, lambda)."""
    times = np.asarray(times, float)[1:]
    decay = np.asarray(signal, float)[1:]
    inv = Inverter(times, min_t2=min_t2, max_t2=max_t2, num_t2_bins=num_t2_bins, num_lambdas=num_lambdas,
                   prune_num=prune_num, noise_amp=noise_amp, min_lambda=min_lambda, max_lambda=max_lambda)
    if lam is None:
        inv.find_best_lambda(decay)
    else:
        inv.set_lambda(lam)
    amps = inv.invert(decay)
    return inv.t2_bins, amps, inv.lam


def nnls(A, b, lam=0.0, tol=1e-10, max_iter=0):
    """min |[A; lam I] x - [b; 0]| subject to x >= 0 (Lawson-Hanson, FP64)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m, n = A.shape
    x = np.zeros(n)
    rc = load().rwnmr_laplace_nnls(m, n, _p(A), _p(b), _p(x), float(lam), float(tol), int(max_iter))
    if rc != 0:
        raise RuntimeError(f"rwnmr_laplace_nnls failed ({rc})")
    return x
