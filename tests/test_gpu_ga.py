"""Label-set CPMG (GPU): n sets of per-label relaxivities on one trajectory
equal n labelled runs from the same start state (walkers and energies bit
for bit, echo sums within 1e-12 relative), through every path of
rwnmr_cpmg_run_label_sets: groups of 8 sets, energies carried between
launches, and the per-set fallback without an unfolded image.  Then the
batched population decays of the relaxivity fit and the genetic algorithm
that scores with them."""
import numpy as np
import pytest

import oracle as O
import randomwalk_nmr_amd as R
from randomwalk_nmr_amd.fit import CpmgModel, genetic_fit

pytestmark = pytest.mark.gpu

ECHO_RTOL = 1e-12
VOXEL, D0 = 1.0, 2.5


def _case(W, H, D, n, seed, n_labels=8):
    solid, _ = R.sphere_pack(W, H, D, 3.0, 6.0, 0.3, seed)
    labels = R.voronoi_labels(solid, 12, n_labels, seed + 1)
    seeds, x, y, z = R.seed_and_place(solid, seed + 2, n)
    return solid, labels, seeds, x, y, z


def _sets(n_sets, n_labels, seed):
    rho = np.random.default_rng(seed).uniform(2.0, 60.0, (n_sets, n_labels))
    return 1.0 - (2.0 / 3.0) * ((rho * VOXEL) / (D0 * 1e03))


def _check_against_single_runs(ctx, img, x, y, z, seeds, sets, echoes, bc, got):
    n_sets = len(sets)
    last_group_first = 8 * ((n_sets - 1) // 8)
    for q in range(n_sets):
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=echoes, steps_per_echo=3, bc=bc, label_delta=sets[q])
        for k in ("x", "y", "z", "seed"):
            assert np.array_equal(got[k], g[k]), (q, k)
        if q == last_group_first:
            assert np.array_equal(got["energy"].view(np.uint64), g["energy"].view(np.uint64)), q
        e, w = got["echo_sums"][q], g["echo_sums"]
        assert np.all(np.abs(e - w) <= ECHO_RTOL * np.abs(w)), (q, np.max(np.abs(e - w) / np.abs(w)))


@pytest.mark.parametrize("dims,bc", [((32, 32, 32), "mirror"), ((30, 21, 17), "mirror"), ((32, 32, 32), "noflux"),
                                     ((30, 21, 17), "periodic")])
def test_label_sets_equal_per_set_labelled_runs(ctx, dims, bc):
    print(f"[test] label sets dims={dims} bc={bc}", flush=True)
    solid, labels, seeds, x, y, z = _case(*dims, 3000, 31)
    img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
    sets = _sets(11, 8, 5)
    got = R.cpmg_run_label_sets(ctx, img, x, y, z, seeds, echoes=60, steps_per_echo=3, bc=bc, set_label_delta=sets)
    assert got["echo_sums"].shape == (11, 61)
    _check_against_single_runs(ctx, img, x, y, z, seeds, sets, 60, bc, got)
    # and one set against the oracle directly
    o = O.walk(O.bitblock_pack(solid), dims, x, y, z, seeds, echoes=60, steps_per_echo=3, bc=bc, labels=labels,
               label_delta=sets[3])
    assert np.all(np.abs(got["echo_sums"][3] - o["echo_sums"]) <= ECHO_RTOL * o["echo_sums"])
    img.close()


def test_label_sets_carry_energies_between_launches(ctx):
    """200,000 walkers at 64 threads per block: 3125 blocks, so 8 sets x
    (echoes + 1) rows of partials exceed the 1 GiB bound at 6000 echoes and
    the run takes two launches, the 8 set energies carried in between."""
    print("[test] label sets, 200000 walkers, 6000 echoes, two launches", flush=True)
    solid, labels, seeds, x, y, z = _case(96, 80, 72, 200_000, 41)
    img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
    blocks = -(-200_000 // 64)
    assert 8 * 8 * blocks * 6001 > (1 << 30)
    sets = _sets(9, 8, 6)
    ctx.set_launch(64, 1)
    try:
        got = R.cpmg_run_label_sets(ctx, img, x, y, z, seeds, echoes=6000, steps_per_echo=3, bc="mirror",
                                    set_label_delta=sets)
    finally:
        ctx.set_launch(256, 1)
    _check_against_single_runs(ctx, img, x, y, z, seeds, sets, 6000, "mirror", got)
    img.close()


def test_label_sets_without_unfolded_image(ctx):
    print("[test] label sets on the folded kernels", flush=True)
    solid, labels, seeds, x, y, z = _case(24, 20, 16, 2000, 51)
    ctx.set_unfolded(0)
    try:
        img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
        sets = _sets(5, 8, 7)
        got = R.cpmg_run_label_sets(ctx, img, x, y, z, seeds, echoes=40, steps_per_echo=3, bc="mirror",
                                    set_label_delta=sets)
        _check_against_single_runs(ctx, img, x, y, z, seeds, sets, 40, "mirror", got)
        img.close()
    finally:
        ctx.set_unfolded(3)


def test_population_decays_equal_single_decays(ctx):
    import torch
    print("[test] population decays vs single decays", flush=True)
    solid, labels, seeds, x, y, z = _case(32, 32, 32, 3000, 61)
    img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
    m = CpmgModel(ctx, img, x, y, z, seeds, echoes=80, steps_per_echo=3, voxel=VOXEL, D0=D0, ckpt_every=20)
    rng = np.random.default_rng(8)
    rho = rng.uniform(5.0, 40.0, (10, 8))
    t2b = rng.uniform(1000.0, 4000.0, 10)
    S = m.decays(rho, t2b)
    assert S.shape == (10, 81)
    for p in range(10):
        with torch.no_grad():
            s1 = m.decay(torch.tensor(rho[p]), torch.tensor(t2b[p])).cpu().numpy()
        assert np.max(np.abs(S[p] - s1)) <= 1e-12, p
    img.close()


def test_genetic_fit_converges_on_small_labelled_problem(ctx):
    print("[test] genetic fit, 32^3, 4 labels", flush=True)
    solid, labels, seeds, x, y, z = _case(32, 32, 32, 4000, 71, n_labels=4)
    img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=4)
    m = CpmgModel(ctx, img, x, y, z, seeds, echoes=200, steps_per_echo=3, voxel=VOXEL, D0=D0, ckpt_every=50)
    truth_g = np.array([8.0, 25.0, 15.0, 35.0, 2800.0])
    truth = m.decays(truth_g[None, :-1], truth_g[-1:])[0]
    lo = np.r_[np.full(4, 1.0), 500.0]
    hi = np.r_[np.full(4, 60.0), 5000.0]
    walks0 = m.forwards
    r = genetic_fit(m, truth, bounds_lo=lo, bounds_hi=hi, truth=truth, tol=2e-3, population=32, max_generations=60,
                    log=lambda s: print(s, flush=True))
    assert r["converged"], r["rms_truth"]
    # every batch of up to 8 individuals is one walk
    assert m.forwards - walks0 <= -(-r["evals"] // 8) + 2 * r["generations"] + 1
    img.close()
