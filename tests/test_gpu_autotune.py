"""The autotuner (rwnmr_ctx_autotune): it settles on a valid launch shape,
image form and re-sort interval, and runs with the tuned context stay
bit-exact against the oracle."""
import numpy as np
import pytest

import oracle as O
import randomwalk_nmr_amd as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [20_000, 70_000])
def test_autotuned_context_stays_exact(n):
    print(f"[test] autotune n={n}", flush=True)
    solid, _ = R.sphere_pack(48, 48, 48, 3.0, 6.0, 0.3, 41)
    seeds, x, y, z = R.seed_and_place(solid, 42, n)
    delta = R.penalty(20.0, 1.0, 2.5)
    with R.Context(0) as ctx:
        img = R.Image.from_solid(ctx, solid)
        tpb, k, ms = ctx.autotune(img, x, y, z, seeds, steps_per_echo=3, bc="mirror", trial_echoes=40, delta=delta)
        t2, k2, form, resort = ctx.get_tuning()
        assert (t2, k2) == (tpb, k) and tpb in (64, 128, 256, 512, 1024) and k in (1, 2, 4) and ms > 0
        # the speculative walk (5) is timed for runs of at most 65,536
        # walkers; mirror walks also time the 8-byte chunk cache (6)
        assert form in ((0, 3, 5, 6) if n <= 65536 else (0, 3, 6))
        # 70,000 walkers use the locality ordering, so the re-sort interval was tuned too
        assert resort in ((10, 20, 40) if n >= 65536 else (1000,))
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=60, steps_per_echo=3, bc="mirror", delta=delta)
        img.close()
    o = O.walk(O.bitblock_pack(solid), (48, 48, 48), x, y, z, seeds, echoes=60, steps_per_echo=3, bc="mirror",
               penalty=delta)
    for key in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(g[key], o[key]), key
    assert np.array_equal(g["energy"].view(np.uint64), o["energy"].view(np.uint64))
    assert np.all(np.abs(g["echo_sums"] - o["echo_sums"]) <= 1e-12 * o["echo_sums"])
