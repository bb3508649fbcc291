"""Jacobian forward (GPU, rwnmr_cpmg_forward_jacobian): the walk and echo
sums equal the plain labelled run; each echo's dM(n)/d log delta_l equals
the oracle's sum_w h_wl(n) E_w(n) (oracle walks of n echoes) and central
finite differences of the echo sums; a run split in two with the label hits
carried equals the whole run; the chain rule to (log rho_l, log T2b)
matches finite differences of the decay; the Levenberg-Marquardt fit
reaches the decay tolerance."""
import math

import numpy as np
import pytest

import oracle as O
import randomwalk_nmr_amd as R
from randomwalk_nmr_amd.fit import CpmgModel, lm_fit, rms

pytestmark = pytest.mark.gpu

VOXEL, D0 = 1.0, 2.5
ECHO_RTOL = 1e-12


def _case(ctx, n=3000, side=32, seed=5):
    solid, _ = R.sphere_pack(side, side, side, 4, 6, 0.3, seed)
    labels = R.voronoi_labels(solid, 12, 8, seed + 1)
    seeds, x, y, z = R.seed_and_place(solid, seed, n)
    img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
    return solid, labels, img, seeds, x, y, z


def _delta(rho):
    return 1.0 - (2.0 / 3.0) * np.asarray(rho, float) * VOXEL / (D0 * 1e3)


@pytest.mark.parametrize("shape", [(256, 1), (1024, 1), (256, 2)])
def test_forward_jacobian_equals_labelled_run_and_oracle(ctx, shape):
    print(f"[test] forward jacobian launch={shape}", flush=True)
    solid, labels, img, seeds, x, y, z = _case(ctx)
    n, E = len(x), 60
    delta = _delta(np.linspace(5, 40, 8))
    ctx.set_launch(*shape)
    try:
        j = R.cpmg_forward_jacobian(ctx, img, x, y, z, seeds, echoes=E, steps_per_echo=3, label_delta=delta,
                                    label_hits=np.zeros((8, n), np.uint32))
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=E, steps_per_echo=3, bc="mirror", label_delta=delta,
                       label_hits=True)
    finally:
        ctx.set_launch(256, 1)
    for k in ("x", "y", "z", "seed"):
        assert np.array_equal(j[k], g[k]), k
    assert np.array_equal(j["energy"].view(np.uint64), g["energy"].view(np.uint64))
    assert np.array_equal(j["label_hits"], g["label_hits"])
    if shape[1] == 1:
        # the same walk and the same grouped echo reduction
        assert np.array_equal(j["echo_sums"], g["echo_sums"])
    else:
        assert np.all(np.abs(j["echo_sums"] - g["echo_sums"]) <= ECHO_RTOL * np.abs(g["echo_sums"]))
    jac = j["jac"]
    assert jac.shape == (8, E + 1)
    assert np.all(jac[:, 0] == 0.0)   # no hits before the run
    blocks = O.bitblock_pack(solid)
    for e in (1, 2, 17, 33, E):
        o = O.walk(blocks, solid.shape[::-1], x, y, z, seeds, echoes=e, steps_per_echo=3, bc="mirror",
                   labels=labels, label_delta=delta, label_hits=True)
        ref = np.array([math.fsum(o["energy"] * o["label_hits"][:, l]) for l in range(8)])
        assert np.all(np.abs(jac[:, e] - ref) <= ECHO_RTOL * np.abs(ref).max()), (e, jac[:, e], ref)
    img.close()


def test_forward_jacobian_matches_central_differences(ctx):
    print("[test] forward jacobian vs finite differences", flush=True)
    solid, labels, img, seeds, x, y, z = _case(ctx, seed=7)
    E, h = 80, 1e-5
    delta = _delta(np.linspace(8, 30, 8))
    j = R.cpmg_forward_jacobian(ctx, img, x, y, z, seeds, echoes=E, steps_per_echo=3, label_delta=delta)
    for l in range(8):
        dp, dm = delta.copy(), delta.copy()
        dp[l] *= math.exp(h)
        dm[l] *= math.exp(-h)
        # the trajectories do not depend on the deltas: common random numbers
        mp = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=E, steps_per_echo=3, bc="mirror", label_delta=dp)
        mm = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=E, steps_per_echo=3, bc="mirror", label_delta=dm)
        fd = (mp["echo_sums"] - mm["echo_sums"]) / (2 * h)
        scale = np.abs(j["jac"][l]).max()
        assert scale > 0
        assert np.all(np.abs(j["jac"][l] - fd) <= 1e-5 * scale), l
    img.close()


def test_forward_jacobian_carries_label_hits(ctx):
    """Two runs of 30 echoes with walkers, energies and label hits carried
    give the Jacobian rows of one 60-echo run (the hits before a run count)."""
    print("[test] forward jacobian split run", flush=True)
    solid, labels, img, seeds, x, y, z = _case(ctx, seed=9)
    n = len(x)
    delta = _delta(np.linspace(5, 40, 8))
    whole = R.cpmg_forward_jacobian(ctx, img, x, y, z, seeds, echoes=60, steps_per_echo=3, label_delta=delta)
    a = R.cpmg_forward_jacobian(ctx, img, x, y, z, seeds, echoes=30, steps_per_echo=3, label_delta=delta,
                                label_hits=np.zeros((8, n), np.uint32))
    b = R.cpmg_forward_jacobian(ctx, img, a["x"], a["y"], a["z"], a["seed"], echoes=30, steps_per_echo=3,
                                label_delta=delta, energy=a["energy"], label_hits=a["label_hits"])
    assert np.all(np.abs(a["jac"] - whole["jac"][:, :31]) <= ECHO_RTOL * np.abs(whole["jac"]).max())
    assert np.all(np.abs(b["jac"] - whole["jac"][:, 30:]) <= ECHO_RTOL * np.abs(whole["jac"]).max())
    img.close()


def _model(ctx, echoes=200, seed=11):
    solid, labels, img, seeds, x, y, z = _case(ctx, n=4000, seed=seed)
    return CpmgModel(ctx, img, x, y, z, seeds, echoes=echoes, steps_per_echo=3, voxel=VOXEL, D0=D0, ckpt_every=50)


def test_decay_jacobian_chain_rule_matches_finite_differences(ctx):
    print("[test] decay jacobian chain rule", flush=True)
    m = _model(ctx)
    rho, t2b, h = np.linspace(6, 33, 8), 2500.0, 1e-5
    S, J = m.decay_and_jacobian(rho, t2b)
    assert J.shape == (m.echoes + 1, 9)
    th = np.concatenate([np.log(rho), [np.log(t2b)]])
    for i in range(9):
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        Sp, _ = m.decay_and_jacobian(np.exp(tp[:-1]), float(np.exp(tp[-1])))
        Sm, _ = m.decay_and_jacobian(np.exp(tm[:-1]), float(np.exp(tm[-1])))
        fd = (Sp - Sm) / (2 * h)
        assert np.all(np.abs(J[:, i] - fd) <= 1e-6 * np.abs(J[:, i]).max() + 1e-12), i


def test_lm_fit_reaches_decay_tolerance(ctx):
    print("[test] lm fit", flush=True)
    m = _model(ctx)
    rho_true, t2b_true = np.array([12.0, 30.0, 9.0, 23.0, 6.0, 16.0, 32.0, 10.0]), 2800.0
    truth, _ = m.decay_and_jacobian(rho_true, t2b_true)
    r = lm_fit(m, truth, rho0=np.full(8, 20.0), t2b0=2000.0, truth=truth, tol=1e-3, max_evals=30)
    assert r["converged"], r["history"][-1]
    assert r["evals"] <= 30
    S, _ = m.decay_and_jacobian(r["rho"], r["t2b"])
    assert rms(S, truth) <= 1e-3
