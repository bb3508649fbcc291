"""Batched persistent CPMG vs one launch per instance and vs the oracle (GPU).

With the single-instance launch shape equal to the batch tile (256 threads x
K walkers), both paths produce the same per-tile partials and sum them in the
same order, so walkers AND echo sums agree bit for bit.
"""
import numpy as np
import pytest

import oracle as O
import randomwalk_nmr_amd as R

pytestmark = pytest.mark.gpu


def _instances(ctx, torch, spec, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i, (side, phi, r, n, rho, echoes, bc) in enumerate(spec):
        solid, _ = R.sphere_pack(side, side, side, r, r + 2, phi, seed * 100 + i)
        seeds, x, y, z = R.seed_and_place(solid, seed * 100 + i, n)
        img = R.Image.from_solid(ctx, solid)
        d = R.penalty(rho, 1.0, 2.5)
        dev = torch.device("cuda:0")
        q = dict(img=img, solid=solid, echoes=echoes, steps_per_echo=3, bc=bc, delta=d,
                 host=(x, y, z, seeds))
        q["x"], q["y"], q["z"] = (torch.from_numpy(a).to(dev) for a in (x, y, z))
        q["seed"] = torch.from_numpy(seeds.view(np.int64)).to(dev)
        q["energy"] = torch.ones(n, dtype=torch.float64, device=dev)
        q["echo_sums"] = torch.zeros(echoes + 1, dtype=torch.float64, device=dev)
        out.append(q)
    return out


SPEC = [(24, 0.3, 3, 700, 10.0, 40, "mirror"), (36, 0.25, 4, 3000, 40.0, 25, "periodic"),
        (16, 0.4, 3, 150, 5.0, 60, "noflux"), (48, 0.2, 5, 5000, 20.0, 30, "mirror"),
        (20, 0.35, 3, 257, 30.0, 45, "mirror")]


@pytest.mark.parametrize("k", [1, 2])
def test_batch_equals_individual_runs(ctx, k):
    import torch
    print(f"[test] batch vs individual k={k}", flush=True)
    ctx.set_launch(256, k)
    try:
        a = _instances(ctx, torch, SPEC, seed=1)
        b = _instances(ctx, torch, SPEC, seed=1)
        R.cpmg_batch_run(ctx, a)
        R.cpmg_batch_run(ctx, b, sequential=True)
    finally:
        ctx.set_launch(256, 1)
    for qa, qb in zip(a, b):
        for key in ("x", "y", "z", "seed", "energy", "echo_sums"):
            assert torch.equal(qa[key], qb[key]), key


def test_batch_matches_oracle(ctx):
    import torch
    print("[test] batch vs oracle", flush=True)
    qs = _instances(ctx, torch, SPEC, seed=2)
    R.cpmg_batch_run(ctx, qs)
    for q in qs:
        x, y, z, seeds = q["host"]
        D = q["solid"].shape[0]
        o = O.walk(O.bitblock_pack(q["solid"]), (D, D, D), x, y, z, seeds, echoes=q["echoes"], steps_per_echo=3,
                   bc=q["bc"], penalty=q["delta"])
        assert np.array_equal(q["x"].cpu().numpy(), o["x"])
        assert np.array_equal(q["seed"].cpu().numpy().view(np.uint64), o["seed"])
        assert np.array_equal(q["energy"].cpu().numpy(), o["energy"])
        es = q["echo_sums"].cpu().numpy()
        assert np.all(np.abs(es - o["echo_sums"]) <= 1e-12 * np.abs(o["echo_sums"]))
