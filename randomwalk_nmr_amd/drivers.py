"""Sequence drivers over the C ABI (include/rwnmr/rwnmr_drivers.h): the
CPMG sequence (NMR_cpmg::run_simulation), the PFGSE sequence (NMR_PFGSE::run)
and the multitau series (NMR_multitau::run).  Host walker arrays."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import core
from ._lib import CpmgParams, MultitauParams, PfgseParams, PfgseResult, check, load


def _p(a):
    return None if a is None else ct.c_void_p(a.ctypes.data)


def _walkers(x, y, z, seed, penalty=None):
    x, y, z = (np.ascontiguousarray(a, np.int32).copy() for a in (x, y, z))
    seed = np.ascontiguousarray(seed, np.uint64).copy()
    pen = None if penalty is None else np.ascontiguousarray(np.broadcast_to(penalty, (len(x),)), np.float64)
    w = core._walker_struct(x, y, z, seed, None, None, pen, 1.0, len(x), 0)
    return w, (x, y, z, seed, pen)


def cpmg_params(*, observation_time, steps_per_echo, voxel_size, D0, bulk_time=2800.0, apply_bulk=True, rho=0.0,
                sigmoid=None, bc="mirror", shift=0):
    sg = None if sigmoid is None else np.ascontiguousarray(sigmoid, np.float64)
    p = CpmgParams(float(observation_time), int(steps_per_echo), float(voxel_size), float(D0), float(bulk_time),
                   int(bool(apply_bulk)), float(rho), _p(sg), core.bc_code(bc), int(shift))
    return p, sg


def cpmg_simulate(ctx, img, x, y, z, seed, *, penalty=None, **kw):
    """NMR_cpmg::run_simulation: returns dict(times, signal, x, y, z, seed)
    (signal normalised, bulk applied when asked)."""
    p, keep = cpmg_params(**kw)
    echoes, _ = core.time_framework(p.observation_time, p.voxel_size, p.D0, p.steps_per_echo)
    w, arrs = _walkers(x, y, z, seed, penalty)
    times, sig = np.zeros(echoes + 1), np.zeros(echoes + 1)
    npts = ct.c_uint32()
    check(load().rwnmr_cpmg_simulate(ctx.handle, img.handle, ct.byref(w), ct.byref(p), echoes + 1, _p(times), _p(sig),
                                     ct.byref(npts)), ctx.handle)
    return dict(times=times[:npts.value], signal=sig[:npts.value], x=arrs[0], y=arrs[1], z=arrs[2], seed=arrs[3])


def pfgse_simulate(ctx, img, x, y, z, seed, *, exposure_times, kx, ky, kz, voxel_size, D0, bulk_time=2800.0,
                   apply_bulk=True, steps_per_echo=1, bc="mirror", pulse_width=0.0, rho=0.0, sigmoid=None,
                   rho_sets=None, threshold_samples=0, threshold_lhs=0.0, threshold_window=1):
    """NMR_PFGSE::run on host walkers; k arrays [n_k] or [n_times, n_k]."""
    T = len(exposure_times)
    kx, ky, kz = (np.ascontiguousarray(k, np.float64) for k in (kx, ky, kz))
    K = kx.shape[-1]
    sg = None if sigmoid is None else np.ascontiguousarray(sigmoid, np.float64)
    rs = None if rho_sets is None else np.ascontiguousarray(rho_sets, np.float64)
    S = 1 if rs is None else len(rs)
    p = PfgseParams(float(voxel_size), float(D0), float(bulk_time), int(bool(apply_bulk)), int(steps_per_echo),
                    core.bc_code(bc), float(pulse_width), float(rho), _p(sg), 0 if rs is None else len(rs), _p(rs),
                    int(kx.ndim == 2), int(threshold_samples), float(threshold_lhs), int(threshold_window))
    out = dict(Mkt=np.zeros((S, T, K)), rhs=np.zeros((S, T, K)), lhs=np.zeros((S, T, K)), D_sat=np.zeros((S, T)),
               D_sat_err=np.zeros((S, T)), points=np.zeros((S, T), np.uint32), msd=np.zeros((T, 3)),
               D_msd=np.zeros(T), steps=np.zeros(T, np.uint64))
    res = PfgseResult(*[_p(out[k]) for k in ("Mkt", "rhs", "lhs", "D_sat", "D_sat_err", "points", "msd", "D_msd",
                                             "steps")])
    et = np.ascontiguousarray(exposure_times, np.float64)
    w, arrs = _walkers(x, y, z, seed)
    check(load().rwnmr_pfgse_simulate(ctx.handle, img.handle, ct.byref(w), ct.byref(p), _p(et), T, _p(kx), _p(ky),
                                      _p(kz), K, ct.byref(res)), ctx.handle)
    out.update(x=arrs[0], y=arrs[1], z=arrs[2], seed=arrs[3])
    return out


def multitau_params(*, tau_min=0.01, tau_max=10.0, tau_points=10, scale="log", tau_values=None,
                    complete_decay=False):
    tv = None if tau_values is None else np.ascontiguousarray(tau_values, np.float64)
    code = {"linear": 0, "log": 1, "manual": 2}[scale]
    n = len(tv) if code == 2 else tau_points
    return MultitauParams(float(tau_min), float(tau_max), int(n), code, _p(tv), int(bool(complete_decay))), tv


def multitau_sequence(time_interval, **kw):
    p, keep = multitau_params(**kw)
    steps, times = np.zeros(p.tau_points, np.uint32), np.zeros(p.tau_points)
    check(load().rwnmr_multitau_sequence(ct.byref(p), float(time_interval), _p(steps), _p(times)))
    return steps, times


def multitau_simulate(ctx, img, x, y, z, seed, *, tau, cpmg):
    """NMR_multitau::run: tau = multitau_params kwargs, cpmg = cpmg_params
    kwargs.  Returns (echo_times, signal); the walkers are not changed."""
    mp, keep1 = multitau_params(**tau)
    cp, keep2 = cpmg_params(**cpmg)
    w, arrs = _walkers(x, y, z, seed)
    times, sig = np.zeros(mp.tau_points), np.zeros(mp.tau_points)
    check(load().rwnmr_multitau_simulate(ctx.handle, img.handle, ct.byref(w), ct.byref(mp), ct.byref(cp), _p(times),
                                         _p(sig)), ctx.handle)
    return times, sig
