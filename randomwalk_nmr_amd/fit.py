"""Relaxivity fit: a differentiable CPMG decay (torch.autograd.Function over
rwnmr_cpmg_forward_ckpt / rwnmr_cpmg_backward) driving a gradient fit of
per-label surface relaxivities rho_l and the bulk T2, with a genetic
algorithm kept as fallback and judge.

The reference fits relaxivity with an (unbuilt) MPI island GA
(B/GA/ga_core.cpp:101-131 run, :133-166 reproduction, :231-357 roulette and
parents, :706-770 blend crossover, individual.cpp:22-44 mutation) whose
fitness is a full forward simulation; its operators are restated in
`genetic_fit` below over the GPU forward.

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
            log(f"fit eval {len(hist)}: rms vs truth {err:.3e} loss {loss.item():.3e}")
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


def genetic_fit(model: CpmgModel, target, *, bounds_lo, bounds_hi, truth=None, tol=1e-3, population=64,
                offspring_proportion=0.5, gamma=0.5, mutation_ratio=0.1, mutation_deviation=1.0, max_generations=200,
                time_limit=None, seed=1234, log=None):
    """GA over genotype (rho_1..rho_L, T2b) with the reference's operators:
    uniform initial population, fitness-proportional (roulette) parents with
    distinct couples, blend crossover alpha ~ U(0, gamma), per-gene mutation
    g += g dev U(-1, 1) with probability mutation_ratio, clamping to bounds,
    worst individuals replaced by the offspring.  Fitness = 1 / misfit RMS.
    Stops when the best individual's RMS against `truth` <= tol."""
    import torch
    rng = np.random.default_rng(seed)
    lo = np.asarray(bounds_lo, float)
    hi = np.asarray(bounds_hi, float)
    G = len(lo)
    target = np.asarray(target, float)
    truth_np = None if truth is None else np.asarray(truth, float)
    t0 = time.perf_counter()
    n_eval = [0]

    def evaluate(g):
        with torch.no_grad():
            S = model.decay(torch.tensor(g[:-1]), torch.tensor(g[-1])).cpu().numpy()
        n_eval[0] += 1
        return rms(S, target), (rms(S, truth_np) if truth_np is not None else rms(S, target))

    pop = lo + (hi - lo) * rng.random((population, G))
    scores = [evaluate(g) for g in pop]
    hist = []
    n_off = max(2, int(round(offspring_proportion * population)) // 2 * 2)
    for gen in range(max_generations):
        fit = np.array([1.0 / max(s[0], 1e-300) for s in scores])
        best = int(np.argmin([s[1] for s in scores]))
        hist.append(dict(t=time.perf_counter() - t0, gen=gen, rms_truth=scores[best][1], evals=n_eval[0]))
        if log:
            log(f"ga gen {gen}: best rms vs truth {scores[best][1]:.3e} ({n_eval[0]} evaluations)")
        if scores[best][1] <= tol:
            return dict(converged=True, seconds=hist[-1]["t"], generations=gen, evals=n_eval[0], genotype=pop[best],
                        rms_truth=scores[best][1], history=hist)
        if time_limit and hist[-1]["t"] > time_limit:
            break
        p = fit / fit.sum()
        children = []
        while len(children) < n_off:
            a = rng.choice(population, p=p)
            b = rng.choice(population, p=p)
            while b == a:
                b = rng.choice(population, p=p)
            alpha = rng.uniform(0.0, gamma, G)
            c1 = alpha * pop[a] + (1 - alpha) * pop[b]
            c2 = alpha * pop[b] + (1 - alpha) * pop[a]
            for c in (c1, c2):
                m = rng.random(G) < mutation_ratio
                c[m] += c[m] * mutation_deviation * rng.uniform(-1, 1, m.sum())
                children.append(np.clip(c, lo, hi))
        child_scores = [evaluate(c) for c in children]
        worst = np.argsort([s[0] for s in scores])[::-1][:len(children)]
        for idx, c, s in zip(worst, children, child_scores):
            pop[idx] = c
            scores[idx] = s
    best = int(np.argmin([s[1] for s in scores]))
    return dict(converged=False, seconds=time.perf_counter() - t0, generations=len(hist), evals=n_eval[0],
                genotype=pop[best], rms_truth=scores[best][1], history=hist)
