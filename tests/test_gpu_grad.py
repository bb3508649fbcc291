"""Differentiable CPMG (GPU): checkpointed forward == plain labelled run,
backward == the analytic per-walker hit-count gradient from the oracle, and
== central finite differences with common random numbers."""
import numpy as np
import pytest

import oracle as O
import randomwalk_nmr_amd as R
from randomwalk_nmr_amd.fit import CpmgModel

pytestmark = pytest.mark.gpu

VOXEL, D0 = 1.0, 2.5


def _model(ctx, n=3000, echoes=60, every=16, side=32, seed=5):
    solid, _ = R.sphere_pack(side, side, side, 4, 6, 0.3, seed)
    labels = R.voronoi_labels(solid, 12, 8, seed + 1)
    seeds, x, y, z = R.seed_and_place(solid, seed, n)
    img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
    m = CpmgModel(ctx, img, x, y, z, seeds, echoes=echoes, steps_per_echo=3, voxel=VOXEL, D0=D0, ckpt_every=every)
    # the model walks the walkers in the Morton order of their start positions
    assert sorted(m.order.tolist()) == list(range(n))
    return m, solid, labels, tuple(a[m.order] for a in (x, y, z, seeds))


def test_forward_ckpt_equals_plain_labelled_run(ctx):
    print("[test] forward ckpt", flush=True)
    m, solid, labels, (x, y, z, seeds) = _model(ctx)
    rho = np.linspace(5, 40, 8)
    delta = m.delta_of_rho(rho)
    M = m.forward_sums(delta).cpu().numpy()
    g = R.cpmg_run(ctx, m.img, x, y, z, seeds, echoes=m.echoes, steps_per_echo=3, bc="mirror", label_delta=delta,
                   label_hits=True)
    assert np.array_equal(M, g["echo_sums"])
    assert np.array_equal(m.e.cpu().numpy(), g["energy"])
    o = O.walk(O.bitblock_pack(solid), solid.shape[::-1], x, y, z, seeds, echoes=m.echoes, steps_per_echo=3,
               bc="mirror", labels=labels, label_delta=delta, label_hits=True)
    assert np.array_equal(g["energy"], o["energy"])
    assert np.all(np.abs(M - o["echo_sums"]) <= 1e-12 * o["echo_sums"])


def test_backward_equals_hit_count_gradient(ctx):
    """L = M(E): dL/d delta_l = sum_w E_w(E) h_wl(E) / delta_l exactly."""
    print("[test] backward analytic", flush=True)
    import torch
    m, solid, labels, (x, y, z, seeds) = _model(ctx)
    delta = m.delta_of_rho(np.linspace(5, 40, 8))
    m.forward_sums(delta)
    g = torch.zeros(m.echoes + 1, dtype=torch.float64)
    g[-1] = 1.0
    got = m.backward_delta(g)
    o = O.walk(O.bitblock_pack(solid), solid.shape[::-1], x, y, z, seeds, echoes=m.echoes, steps_per_echo=3,
               bc="mirror", labels=labels, label_delta=delta, label_hits=True)
    import math
    ref = np.array([math.fsum(o["energy"] * o["label_hits"][:, l]) / delta[l] for l in range(8)])
    assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref).max())


def test_backward_matches_central_finite_differences(ctx):
    """Common random numbers: trajectories do not depend on rho, so central
    differences (h = 1e-3 relative on rho) match to O(h^2)."""
    print("[test] backward FD", flush=True)
    import torch
    m, *_ = _model(ctx, n=4000, echoes=80, every=20)
    rho = np.linspace(5, 40, 8)
    rng = np.random.default_rng(0)
    c = rng.standard_normal(m.echoes + 1)
    m.forward_sums(m.delta_of_rho(rho))
    grad_delta = m.backward_delta(torch.tensor(c))
    grad_rho = grad_delta * m.ddelta_drho
    for l in range(8):
        h = 1e-3 * rho[l]
        rp, rm = rho.copy(), rho.copy()
        rp[l] += h
        rm[l] -= h
        Lp = float(np.dot(c, m.forward_sums(m.delta_of_rho(rp)).cpu().numpy()))
        Lm = float(np.dot(c, m.forward_sums(m.delta_of_rho(rm)).cpu().numpy()))
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - grad_rho[l]) <= 1e-4 * max(abs(grad_rho[l]), 1e-3 * np.abs(grad_rho).max()), (l, fd,
                                                                                                       grad_rho[l])


def test_autograd_function_gradients(ctx):
    print("[test] autograd", flush=True)
    import torch
    m, *_ = _model(ctx, n=2000, echoes=40, every=8)
    rho = torch.tensor(np.linspace(8, 30, 8), dtype=torch.float64, requires_grad=True)
    t2b = torch.tensor(900.0, dtype=torch.float64, requires_grad=True)
    S = m.decay(rho, t2b)
    w = torch.linspace(0, 1, m.echoes + 1, dtype=torch.float64)
    (S * w).sum().backward()
    gr, gt = rho.grad.clone(), t2b.grad.clone()
    with torch.no_grad():
        eps = 1e-3
        for l in (0, 3, 7):
            rp = rho.detach().clone()
            rm = rho.detach().clone()
            rp[l] += eps * rp[l]
            rm[l] -= eps * rm[l]
            fd = ((m.decay(rp, t2b) * w).sum() - (m.decay(rm, t2b) * w).sum()) / (2 * eps * rho[l])
            assert abs(fd - gr[l]) <= 1e-4 * abs(gr).max()
        fd_t = ((m.decay(rho, t2b + 0.5) * w).sum() - (m.decay(rho, t2b - 0.5) * w).sum()) / 1.0
        assert abs(fd_t - gt) <= 1e-5 * abs(gt)


@pytest.mark.parametrize("bc", ["periodic", "mirror"])
def test_backward_unfolded_equals_folded_replay(ctx, bc):
    """The replay on the unfolded image (non-power-of-two periods) gives the
    same label gradient as the folded replay."""
    print(f"[test] backward unfolded vs folded bc={bc}", flush=True)
    import torch
    solid, _ = R.sphere_pack(30, 21, 17, 3, 5, 0.3, 8)
    labels = R.voronoi_labels(solid, 10, 8, 9)
    seeds, x, y, z = R.seed_and_place(solid, 10, 2500)
    grads, sums = [], []
    for mode in (3, 0):
        ctx.set_unfolded(mode)
        try:
            img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
            m = CpmgModel(ctx, img, x, y, z, seeds, echoes=50, steps_per_echo=3, voxel=VOXEL, D0=D0, bc=bc,
                          ckpt_every=12)
            delta = m.delta_of_rho(np.linspace(5, 40, 8))
            sums.append(m.forward_sums(delta).cpu().numpy())
            c = torch.tensor(np.random.default_rng(3).standard_normal(m.echoes + 1))
            grads.append(np.asarray(m.backward_delta(c)))
        finally:
            ctx.set_unfolded(3)
    assert np.array_equal(sums[0], sums[1])
    assert np.all(np.abs(grads[0] - grads[1]) <= 1e-12 * np.abs(grads[1]).max())
