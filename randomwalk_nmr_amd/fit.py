"""Relaxivity fit of per-label surface relaxivities rho_l and the bulk T2:
a least-squares fit (Levenberg-Marquardt) with the exact Jacobian from one
walk (rwnmr_cpmg_forward_jacobian), a differentiable CPMG decay
(torch.autograd.Function over rwnmr_cpmg_forward_ckpt / rwnmr_cpmg_backward)
driving a gradient fit, and a genetic algorithm kept as fallback and comparator.

The reference fits relaxivity with an (unbuilt) MPI island GA
(B/GA/ga_core.cpp:101-131 run, :133-166 reproduction, :231-357 roulette and
parents, :535-657 diversity check and reset, :706-770 blend crossover,
individual.cpp:22-44 mutation) whose fitness is a full forward simulation
per individual; its operators are restated in `genetic_fit` below, scoring
each batch of individuals on one GPU trajectory (CpmgModel.decays).

Decay model (NMR_cpmg.cpp:346-385): S(t_n) = exp(-t_n / T2b) M(n) / M(0),
M(n) = sum_w E_w(n), delta_l = 1 - (2/3) rho_l eps / (D0 1e3)
(B/Walker/walker.h:203-209).
"""
from __future__ import annotations

import ctypes as ct
import time

import numpy as np

from . import core
from ._lib import check, load


def _spread3(v):
    """Bits 0..20 of v moved to bits 0, 3, 6, ... (uint64)."""
    v = v & np.uint64(0x1FFFFF)
    for shift, mask in ((32, 0x1F00000000FFFF), (16, 0x1F0000FF0000FF), (8, 0x100F00F00F00F00F),
                        (4, 0x10C30C30C30C30C3), (2, 0x1249249249249249)):
        v = (v | (v << np.uint64(shift))) & np.uint64(mask)
    return v


def morton_order(x, y, z):
    """Permutation sorting walkers by the Morton code of their position, so
    that walkers walked by neighbouring lanes and blocks share image lines
    in the caches."""
    c = [np.asarray(a, np.int64) for a in (x, y, z)]
    c = [(a - a.min()).astype(np.uint64) if a.size else a.astype(np.uint64) for a in c]
    key = _spread3(c[0]) | (_spread3(c[1]) << np.uint64(1)) | (_spread3(c[2]) << np.uint64(2))
    return np.argsort(key, kind="stable")


class CpmgModel:
    """Fixed labelled geometry and walkers resident on the GPU; evaluates the
    echo sums M(n) for label deltas and their gradient by checkpoint replay.
    With locality=True the walkers are held in the Morton order of their
    start positions (`order`: the permutation applied to the inputs), which
    keeps the image lines of neighbouring walkers in L2 for the forward walk
    and the replay; the echo sums then differ from the input order's only by
    summation order."""

    def __init__(self, ctx: core.Context, img: core.Image, x, y, z, seed, *, echoes, steps_per_echo, voxel, D0,
                 bc="mirror", ckpt_every=500, locality=True):
        import torch
        if img.n_labels == 0:
            raise ValueError("the image needs solid labels")
        self.torch = torch
        self.ctx, self.img = ctx, img
        self.echoes, self.spe = int(echoes), int(steps_per_echo)
        self.voxel, self.D0, self.bc = float(voxel), float(D0), core.bc_code(bc)
        self.every = int(ckpt_every)
        self.n = len(x)
        self.n_labels = img.n_labels
        dev = torch.device(f"cuda:{ctx.device}")
        self.dev = dev
        self.order = morton_order(x, y, z) if locality else np.arange(len(x))
        x, y, z, seed = (np.asarray(a)[self.order] for a in (x, y, z, seed))
        self.x0, self.y0, self.z0 = (torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev) for a in (x, y, z))
        self.s0 = torch.from_numpy(np.ascontiguousarray(seed, np.uint64).view(np.int64)).to(dev)
        self.x, self.y, self.z, self.s = (t.clone() for t in (self.x0, self.y0, self.z0, self.s0))
        self.e = torch.ones(self.n, dtype=torch.float64, device=dev)
        self.sums = torch.zeros(self.echoes + 1, dtype=torch.float64, device=dev)
        self.delta_dev = torch.zeros(self.n_labels, dtype=torch.float64, device=dev)
        nseg = (self.echoes + self.every - 1) // self.every
        nbytes = int(load().rwnmr_ckpt_bytes(self.n, nseg, self.n_labels))
        self.ckpt = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        dt = core.time_step(self.voxel, self.D0)
        t, times = 0.0, [0.0]
        for _ in range(self.echoes):
            t += self.spe * dt
            times.append(t)
        self.times = torch.tensor(times, dtype=torch.float64, device=dev)
        self.forwards = 0
        self.backwards = 0

    def delta_of_rho(self, rho):
        return 1.0 - (2.0 / 3.0) * ((rho * self.voxel) / (self.D0 * 1e03))

    @property
    def ddelta_drho(self):
        return -(2.0 / 3.0) * (self.voxel / (self.D0 * 1e03))

    def forward_sums(self, delta):
        """Echo sums M(0..E) (device float64) from the start state."""
        torch = self.torch
        self.delta_dev.copy_(torch.as_tensor(delta, dtype=torch.float64))
        self.x.copy_(self.x0)
        self.y.copy_(self.y0)
        self.z.copy_(self.z0)
        self.s.copy_(self.s0)
        self.e.fill_(1.0)
        torch.cuda.current_stream().synchronize()
        w = core._walker_struct(self.x, self.y, self.z, self.s, self.e, None, None, 1.0, self.n, 1)
        L = load()
        check(L.rwnmr_cpmg_forward_ckpt(self.ctx.handle, self.img.handle, ct.byref(w), self.echoes, self.spe,
                                        self.bc, core._ptr(self.delta_dev), self.every, core._ptr(self.ckpt),
                                        core._ptr(self.sums)), self.ctx.handle)
        self.ctx.synchronize()
        self.forwards += 1
        return self.sums.clone()

    def backward_delta(self, grad_sums):
        """dL/d delta_l (numpy) for dL/dM(n) = grad_sums, replaying the
        checkpoints of the last forward_sums call."""
        torch = self.torch
        g = grad_sums.to(device=self.dev, dtype=torch.float64).contiguous()
        torch.cuda.current_stream().synchronize()
        out = np.zeros(self.n_labels)
        check(load().rwnmr_cpmg_backward(self.ctx.handle, self.img.handle, core._ptr(self.ckpt), self.n,
                                         self.echoes, self.spe, self.bc, core._ptr(self.delta_dev), self.every,
                                         core._ptr(g), out.ctypes.data_as(ct.c_void_p)), self.ctx.handle)
        self.backwards += 1
        return out

    def decay(self, rho, t2b):
        """S(t_n) as a differentiable torch function of rho[n_labels] and T2b."""
        return CpmgDecay.apply(rho, t2b, self)

    def forward_jacobian(self, delta):
        """Echo sums M(0..E) and J[l, n] = dM(n)/d log delta_l (device
        float64, [E+1] and [n_labels, E+1]) from the start state, in one
        walk (rwnmr_cpmg_forward_jacobian); M equals forward_sums(delta)."""
        torch = self.torch
        if not hasattr(self, "jac"):
            self.jac = torch.zeros((self.n_labels, self.echoes + 1), dtype=torch.float64, device=self.dev)
        self.delta_dev.copy_(torch.as_tensor(delta, dtype=torch.float64))
        self.x.copy_(self.x0)
        self.y.copy_(self.y0)
        self.z.copy_(self.z0)
        self.s.copy_(self.s0)
        self.e.fill_(1.0)
        torch.cuda.current_stream().synchronize()
        w = core._walker_struct(self.x, self.y, self.z, self.s, self.e, None, None, 1.0, self.n, 1)
        check(load().rwnmr_cpmg_forward_jacobian(self.ctx.handle, self.img.handle, ct.byref(w), self.echoes,
                                                 self.spe, self.bc, core._ptr(self.delta_dev), None,
                                                 core._ptr(self.sums), core._ptr(self.jac)), self.ctx.handle)
        self.ctx.synchronize()
        self.forwards += 1
        return self.sums.clone(), self.jac.clone()

    def decay_and_jacobian(self, rho, t2b):
        """S(t_n) = exp(-t_n/T2b) M(n)/M(0) and its Jacobian with respect to
        theta = (log rho_1 .. log rho_L, log T2b), numpy [E+1] and
        [E+1, L+1]: dS/d log rho_l = exp(-t/T2b) J_l / M(0) x
        (d delta/d rho) rho_l / delta_l (M(0), the start energies, does not
        depend on rho), dS/d log T2b = S t / T2b."""
        torch = self.torch
        rho = np.asarray(rho, float)
        delta = self.delta_of_rho(rho)
        M, J = self.forward_jacobian(delta)
        bulk = torch.exp(-self.times / float(t2b))
        S = bulk * M / M[0]
        dlog = torch.as_tensor(self.ddelta_drho * rho / delta, dtype=torch.float64, device=self.dev)
        jac = torch.cat([(bulk / M[0])[:, None] * J.T * dlog[None, :], (S * self.times / float(t2b))[:, None]], 1)
        return S.cpu().numpy(), jac.cpu().numpy()

    def population_sums(self, deltas):
        """Echo sums [P, E+1] (device float64) of P rows of label deltas, all
        on the trajectories of the start state (rwnmr_cpmg_run_label_sets:
        8 rows per walk); row p equals forward_sums(deltas[p]) per walker bit
        for bit, in its sums to rounding."""
        torch = self.torch
        d = torch.as_tensor(np.ascontiguousarray(deltas, np.float64).reshape(-1, self.n_labels)).to(self.dev)
        P = d.shape[0]
        out = torch.empty((P, self.echoes + 1), dtype=torch.float64, device=self.dev)
        self.x.copy_(self.x0)
        self.y.copy_(self.y0)
        self.z.copy_(self.z0)
        self.s.copy_(self.s0)
        self.e.fill_(1.0)
        torch.cuda.current_stream().synchronize()
        w = core._walker_struct(self.x, self.y, self.z, self.s, self.e, None, None, 1.0, self.n, 1)
        check(load().rwnmr_cpmg_run_label_sets(self.ctx.handle, self.img.handle, ct.byref(w), self.echoes, self.spe,
                                               self.bc, P, core._ptr(d), core._ptr(out)), self.ctx.handle)
        self.ctx.synchronize()
        self.forwards += (P + 7) // 8
        return out

    def decays(self, rho, t2b):
        """S_p(t_n) = exp(-t_n / T2b_p) M_p(n) / M_p(0) (numpy [P, E+1]) for
        rows rho[P, n_labels] and t2b[P], scored on one trajectory."""
        torch = self.torch
        rho = np.asarray(rho, float).reshape(-1, self.n_labels)
        t2b = np.asarray(t2b, float).reshape(-1)
        M = self.population_sums(self.delta_of_rho(rho))
        bulk = torch.exp(-self.times[None, :] / torch.as_tensor(t2b, dtype=torch.float64, device=self.dev)[:, None])
        return (bulk * M / M[:, :1]).cpu().numpy()


def _autograd_function():
    import torch

    class _CpmgDecay(torch.autograd.Function):
        """S(t_n) = exp(-t_n/T2b) M(n)/M(0) with M from the GPU walk.
        Backward: dL/dM(n) = g_n exp(-t_n/T2b)/M(0) (M(0) = sum of start
        energies, independent of rho), dL/d delta from the replay kernel,
        chain rule to rho; dS/dT2b = S t / T2b^2."""

        @staticmethod
        def forward(fctx, rho, t2b, model):
            delta = model.delta_of_rho(rho.detach().to(torch.float64))
            M = model.forward_sums(delta.cpu().numpy())
            bulk = torch.exp(-model.times / t2b.detach().to(model.dev))
            S = bulk * M / M[0]
            fctx.model = model
            fctx.save_for_backward(t2b, M, S)
            return S.to(rho.device)

        @staticmethod
        def backward(fctx, gS):
            t2b, M, S = fctx.saved_tensors
            model = fctx.model
            gS = gS.to(model.dev, torch.float64)
            bulk = torch.exp(-model.times / t2b.to(model.dev))
            gM = gS * bulk / M[0]
            gdelta = model.backward_delta(gM)
            grho = torch.as_tensor(gdelta * model.ddelta_drho, dtype=torch.float64)
            gt2b = torch.sum(gS * S * model.times) / (t2b.to(model.dev) ** 2)
            return grho, gt2b.cpu(), None

    return _CpmgDecay


class _Lazy:
    fn = None


def _get_fn():
    if _Lazy.fn is None:
        _Lazy.fn = _autograd_function()
    return _Lazy.fn


class CpmgDecay:
    @staticmethod
    def apply(rho, t2b, model):
        return _get_fn().apply(rho, t2b, model)


def rms(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.sqrt(np.mean((a - b) ** 2)))


def gradient_fit(model: CpmgModel, target, *, rho0, t2b0, truth=None, tol=1e-3, max_evals=80,
                 rho_bounds=(0.5, 200.0), t2b_bounds=(50.0, 1e5), log=None):
    """L-BFGS on log(rho_l), log(T2b) minimising the mean squared decay
    misfit to `target`.  Stops when the RMS error against `truth` (the
    noise-free decay, if given) falls below tol.  Returns a dict with the
    fitted parameters, history and GPU wall time."""
    import torch
    target_t = torch.as_tensor(np.asarray(target, float), dtype=torch.float64)
    truth_np = None if truth is None else np.asarray(truth, float)
    theta = torch.tensor(np.concatenate([np.log(rho0), [np.log(t2b0)]]), dtype=torch.float64, requires_grad=True)
    lo = torch.tensor(np.concatenate([np.full(len(rho0), np.log(rho_bounds[0])), [np.log(t2b_bounds[0])]]))
    hi = torch.tensor(np.concatenate([np.full(len(rho0), np.log(rho_bounds[1])), [np.log(t2b_bounds[1])]]))
    opt = torch.optim.LBFGS([theta], lr=1.0, max_iter=max_evals, max_eval=max_evals, tolerance_grad=1e-14,
                            tolerance_change=1e-16, history_size=20, line_search_fn="strong_wolfe")
    hist = []
    state = {"done": False, "best": None}
    t0 = time.perf_counter()

    def closure():
        opt.zero_grad()
        th = torch.clamp(theta, lo, hi)
        rho = torch.exp(th[:-1])
        t2b = torch.exp(th[-1])
        S = model.decay(rho, t2b).cpu()
        loss = torch.mean((S - target_t) ** 2)
        loss.backward()
        Sn = S.detach().numpy()
        err = rms(Sn, truth_np) if truth_np is not None else float(np.sqrt(loss.item()))
        hist.append(dict(t=time.perf_counter() - t0, loss=float(loss.item()), rms_truth=err,
                         rho=torch.exp(th[:-1]).detach().numpy().copy(), t2b=float(t2b.item())))
        if log:
            log(f"fit eval {len(hist)} at {hist[-1]['t']:.3f} s: rms vs truth {err:.3e} loss {loss.item():.3e}")
        if err <= tol and not state["done"]:
            state["done"] = True
            state["best"] = hist[-1]
        if state["done"]:
            # freeze the optimiser once the goal is met
            theta.grad.zero_()
        return loss

    try:
        opt.step(closure)
    except RuntimeError:
        pass
    best = state["best"] or min(hist, key=lambda h: h["rms_truth"])
    return dict(converged=state["done"], evals=len(hist), seconds=best["t"] if state["done"] else hist[-1]["t"],
                rho=best["rho"], t2b=best["t2b"], rms_truth=best["rms_truth"], history=hist)


def lm_fit(model: CpmgModel, target, *, rho0, t2b0, truth=None, tol=1e-3, max_evals=40,
           rho_bounds=(0.5, 200.0), t2b_bounds=(50.0, 1e5), log=None):
    """Levenberg-Marquardt on theta = (log rho_l, log T2b) minimising the
    squared decay misfit to `target`, with the exact Jacobian from one GPU
    walk per evaluation (CpmgModel.decay_and_jacobian).  The damping mu
    follows the gain ratio (Nielsen's update: on a successful step mu *=
    max(1/3, 1 - (2 r - 1)^3), else mu *= nu, nu *= 2), starting from 1e-3
    max diag(J^T J).  Steps are clipped to the bounds.  Stops when the RMS
    error against `truth` (the noise-free decay, if given) falls below tol
    at an accepted step.  Returns the dict of gradient_fit."""
    target = np.asarray(target, float)
    truth_np = None if truth is None else np.asarray(truth, float)
    lo = np.concatenate([np.full(len(rho0), np.log(rho_bounds[0])), [np.log(t2b_bounds[0])]])
    hi = np.concatenate([np.full(len(rho0), np.log(rho_bounds[1])), [np.log(t2b_bounds[1])]])
    th = np.clip(np.concatenate([np.log(rho0), [np.log(t2b0)]]), lo, hi)
    hist = []
    t0 = time.perf_counter()

    def evaluate(theta):
        S, J = model.decay_and_jacobian(np.exp(theta[:-1]), float(np.exp(theta[-1])))
        r = S - target
        err = rms(S, truth_np) if truth_np is not None else float(np.sqrt(np.mean(r ** 2)))
        hist.append(dict(t=time.perf_counter() - t0, loss=float(np.mean(r ** 2)), rms_truth=err,
                         rho=np.exp(theta[:-1]).copy(), t2b=float(np.exp(theta[-1]))))
        if log:
            log(f"lm eval {len(hist)} at {hist[-1]['t']:.3f} s: rms vs truth {err:.3e} loss {hist[-1]['loss']:.3e}")
        return r, J, err

    r, J, err = evaluate(th)
    best = hist[-1] if err <= tol else None
    A = J.T @ J
    mu, nu = 1e-3 * float(np.max(np.diag(A))), 2.0
    while best is None and len(hist) < max_evals:
        A = J.T @ J
        g = J.T @ r
        h = np.linalg.solve(A + mu * np.eye(len(th)), -g)
        th_new = np.clip(th + h, lo, hi)
        h = th_new - th
        r_new, J_new, err_new = evaluate(th_new)
        F, F_new = 0.5 * float(r @ r), 0.5 * float(r_new @ r_new)
        pred = 0.5 * float(h @ (mu * h - g))
        gain = (F - F_new) / pred if pred > 0 else -1.0
        if gain > 0:
            th, r, J = th_new, r_new, J_new
            mu *= max(1.0 / 3.0, 1.0 - (2.0 * gain - 1.0) ** 3)
            nu = 2.0
            if err_new <= tol:
                best = hist[-1]
        else:
            mu *= nu
            nu *= 2.0
    done = best is not None
    if best is None:
        best = min(hist, key=lambda e: e["rms_truth"])
    return dict(converged=done, evals=len(hist), seconds=best["t"] if done else hist[-1]["t"], rho=best["rho"],
                t2b=best["t2b"], rms_truth=best["rms_truth"], history=hist)


def ga_offspring_count(population: int, offspring_proportion: float) -> int:
    """numberOfOffsprings = (proportion * (population / 2)) * 2 with the
    integer halving first (B/GA/ga_core.cpp:80)."""
    return int((offspring_proportion * (population // 2)) * 2)


def genotype_deviation(pop) -> float:
    """Mean over genes of the population standard deviation divided by the
    gene mean (checkGenotypeDeviation with GA_MEAN_DEVIATION,
    B/GA/ga_core.cpp:535-564, findStdDev :799-811)."""
    pop = np.asarray(pop, float)
    mean = pop.mean(axis=0)
    std = np.sqrt(((pop - mean) ** 2).mean(axis=0))
    return float(np.mean(std / mean))


def roulette_index(cum, r: float) -> int:
    """First index whose cumulative probability reaches r (findIndex,
    B/GA/ga_core.cpp:842-852), kept inside the population."""
    return int(min(np.searchsorted(cum, r, side="left"), len(cum) - 1))


def t2_distribution(times, decay, **laplace_kw):
    """T2 amplitudes of a decay (the Laplace inversion of the CPMG path)."""
    from .laplace import apply_laplace
    _, amps, _ = apply_laplace(np.asarray(times, float), np.asarray(decay, float), **laplace_kw)
    return amps


def genetic_fit(model, target, *, bounds_lo, bounds_hi, truth=None, tol=1e-3, population=64,
                offspring_proportion=0.5, gamma=0.5, mutation_ratio=0.1, mutation_deviation=1.0, diversity=0.9,
                mutations_per_reset=5, top_size=0.25, fitness="rms", laplace_kw=None, max_generations=200,
                time_limit=None, seed=1234, log=None):
    """Genetic algorithm over the genotype (rho_1..rho_L, T2b), restating the
    reference's run loop (B/GA/ga_core.cpp:110-166): per generation,
    reproduction (fitness-proportional roulette :231-247 with distinct
    couples :249-296, blend crossover alpha ~ U(0, gamma) per gene :713-733,
    per-gene mutation g += g dev U(-1, 1) with probability mutation_ratio,
    individual.cpp:22-44, clamping to the bounds :735-742), selection (the
    2 |offspring| worst replaced, population sorted by fitness :510-533) and
    diversification (:535-657: when the mean relative gene deviation falls
    below 1 - diversity, all genes of individuals 1 .. P/2-1 are mutated, or
    after `mutations_per_reset` such events the population is reset except
    its top `top_size` fraction).

    The population is scored in one batch per step (model.decays: all
    individuals on one trajectory) instead of one forward simulation per
    individual (ga_core.cpp:463-490).  Fitness: "rms" (default) is 1 / RMS
    misfit of the decay to `target`; "t2_correlation" is the squared Pearson
    correlation of the Laplace-inverted T2 distributions of the decay and
    the target (the reference's fitNMRT2 squares a T2 correlation whose
    method, leastSquaresT2, is absent from its sources).  Deviations:
    imposed mutations are clamped to the bounds (the reference's mutate
    leaves them unbounded); the generator is numpy's seeded PCG64 instead
    of a random_device reseed per draw (so runs repeat).

    Stops when the fittest individual's decay is within `tol` RMS of `truth`
    (the noise-free decay, if given; else of `target`)."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(bounds_lo, float)
    hi = np.asarray(bounds_hi, float)
    G = len(lo)
    P = int(population)
    target = np.asarray(target, float)
    truth_np = np.asarray(truth, float) if truth is not None else target
    times = model.times.cpu().numpy() if hasattr(model.times, "cpu") else np.asarray(model.times)
    lkw = dict(laplace_kw or {})
    t2_target = t2_distribution(times, target, **lkw) if fitness == "t2_correlation" else None
    t0 = time.perf_counter()
    n_eval = [0]

    def score(g):
        """Fitness and RMS against the truth of every row of g."""
        g = np.atleast_2d(g)
        S = model.decays(g[:, :-1], g[:, -1])
        n_eval[0] += len(g)
        err_truth = np.sqrt(np.mean((S - truth_np[None, :]) ** 2, axis=1))
        if fitness == "t2_correlation":
            fit = np.array([np.corrcoef(t2_distribution(times, s, **lkw), t2_target)[0, 1] ** 2 for s in S])
            fit = np.nan_to_num(fit, nan=0.0)
        else:
            fit = 1.0 / np.maximum(np.sqrt(np.mean((S - target[None, :]) ** 2, axis=1)), 1e-300)
        return fit, err_truth

    def sort(pop, fit, err):
        order = np.argsort(-fit, kind="stable")
        return pop[order], fit[order], err[order]

    pop = lo + (hi - lo) * rng.random((P, G))
    fit, err = score(pop)
    pop, fit, err = sort(pop, fit, err)
    n_off = ga_offspring_count(P, offspring_proportion)
    imposed = 0
    resets = 0
    hist = []
    for gen in range(max_generations + 1):
        hist.append(dict(t=time.perf_counter() - t0, gen=gen, rms_truth=float(err[0]), evals=n_eval[0],
                         deviation=genotype_deviation(pop)))
        if log:
            log(f"ga gen {gen}: fittest rms vs truth {err[0]:.3e} ({n_eval[0]} evaluations)")
        if err[0] <= tol:
            return dict(converged=True, seconds=hist[-1]["t"], generations=gen, evals=n_eval[0], genotype=pop[0],
                        rms_truth=float(err[0]), resets=resets, history=hist)
        if gen == max_generations or (time_limit and hist[-1]["t"] > time_limit):
            break
        # reproduction: couples by roulette on fitness, blend crossover, mutation, bounds
        cum = np.cumsum(fit / fit.sum()) if fit.sum() != 0 else np.cumsum(np.full(P, 1.0 / P))
        children = []
        for _ in range(n_off // 2):
            p1 = roulette_index(cum, rng.random())
            p2 = roulette_index(cum, rng.random())
            while p2 == p1:
                p2 = roulette_index(cum, rng.random())
            alpha = rng.uniform(0.0, gamma, G)
            children.append(alpha * pop[p1] + (1.0 - alpha) * pop[p2])
            children.append(alpha * pop[p2] + (1.0 - alpha) * pop[p1])
        children = np.array(children)
        for c in children:
            m = rng.random(G) < mutation_ratio
            c[m] += c[m] * mutation_deviation * rng.uniform(-1.0, 1.0, int(m.sum()))
        children = np.clip(children, lo, hi)
        cfit, cerr = score(children)
        # selection: the worst 2 |offspring| make room for the children
        keep = P - len(children)
        pop = np.concatenate([pop[:keep], children])
        fit = np.concatenate([fit[:keep], cfit])
        err = np.concatenate([err[:keep], cerr])
        pop, fit, err = sort(pop, fit, err)
        # diversification
        if genotype_deviation(pop) < 1.0 - diversity:
            if imposed == mutations_per_reset:
                top = int(np.ceil(top_size * P))
                fresh = lo + (hi - lo) * rng.random((P - top, G))
                ffit, ferr = score(fresh)
                pop = np.concatenate([pop[:top], fresh])
                fit = np.concatenate([fit[:top], ffit])
                err = np.concatenate([err[:top], ferr])
                imposed = 0
                resets += 1
            else:
                ids = np.arange(1, P // 2)
                mutant = pop[ids] + pop[ids] * mutation_deviation * rng.uniform(-1.0, 1.0, (len(ids), G))
                pop[ids] = np.clip(mutant, lo, hi)
                fit[ids], err[ids] = score(pop[ids])
                imposed += 1
            pop, fit, err = sort(pop, fit, err)
    return dict(converged=False, seconds=time.perf_counter() - t0, generations=len(hist) - 1, evals=n_eval[0],
                genotype=pop[0], rms_truth=float(err[0]), resets=resets, history=hist)
