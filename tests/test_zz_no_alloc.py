"""rwnmr_cpmg_run_noalloc (the CPMG entry point that allocates nothing) vs
rwnmr_cpmg_run, both called through the C ABI on device-resident walkers.

Bar: positions, seeds, hit counts and energies bit-exact; echo sums within
1e-12 relative.  Two calls in a row, the second on fewer walkers, so the
shared device scratch keeps the first call's contents when the second runs.
The first call walks an image whose unfolded copy was never built (folded
kernels) and takes several launches (its partials exceed the scratch); the
second walks after rwnmr_cpmg_run built the unfolded copy.
"""
import ctypes as ct

import numpy as np
import pytest

import randomwalk_nmr_amd as R
from randomwalk_nmr_amd._lib import load
from randomwalk_nmr_amd.core import _walker_struct, bc_code

pytestmark = pytest.mark.gpu

ECHO_RTOL = 1e-12
SCRATCH_PARTIAL_DOUBLES = ((16 << 20) - 256) // 8


def _device_walkers(torch, solid, seed, n, with_hits):
    seeds, x, y, z = R.seed_and_place(solid, seed, n)
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(v).to(dev) for k, v in (("x", x), ("y", y), ("z", z))}
    t["seed"] = torch.from_numpy(seeds.view(np.int64)).to(dev)
    t["energy"] = torch.ones(n, dtype=torch.float64, device=dev)
    t["hits"] = torch.zeros(n, dtype=torch.int32, device=dev) if with_hits else None
    return t


def _clone(t):
    return {k: (None if v is None else v.clone()) for k, v in t.items()}


def _run(ctx, img, entry, t, echoes, spe, bc, delta):
    import torch
    L = load()
    n = t["x"].numel()
    sums = torch.full((echoes + 1,), -1.0, dtype=torch.float64, device="cuda:0")
    w = _walker_struct(t["x"], t["y"], t["z"], t["seed"], t["energy"], t["hits"], None, delta, n, 1)
    torch.cuda.synchronize()
    fn = L.rwnmr_cpmg_run_noalloc if entry == "noalloc" else L.rwnmr_cpmg_run
    code = fn(ctx.handle, img.handle, ct.byref(w), echoes, spe, bc_code(bc), 0, ct.c_void_p(sums.data_ptr()))
    assert code == 0, L.rwnmr_last_error(ctx.handle)
    ctx.synchronize()
    return sums.cpu().numpy()


def _compare(a, b, sa, sb):
    for k in ("x", "y", "z", "seed", "hits"):
        if a[k] is not None:
            assert np.array_equal(a[k].cpu().numpy(), b[k].cpu().numpy()), k
    assert np.array_equal(a["energy"].cpu().numpy().view(np.uint64), b["energy"].cpu().numpy().view(np.uint64))
    assert np.all(np.abs(sa - sb) <= ECHO_RTOL * np.abs(sb)), np.max(np.abs(sa - sb) / np.abs(sb))


def test_noalloc_matches_cpmg_run_twice():
    import torch
    ctx = R.Context(0)
    try:
        # call 1: 200,000 walkers at 64 threads x 1 -> 3125 blocks; 671 partial rows fit the
        # scratch, so 1500 echoes take three launches; the image has no unfolded copy yet
        solid, _ = R.sphere_pack(96, 80, 72, 4.0, 9.0, 0.3, 5)
        img = R.Image.from_solid(ctx, solid)
        ctx.set_launch(64, 1)
        n1, e1 = 200_000, 1500
        blocks = -(-n1 // 64)
        assert SCRATCH_PARTIAL_DOUBLES // blocks - 1 < e1
        base = _device_walkers(torch, solid, 17, n1, with_hits=True)
        a, b = _clone(base), _clone(base)
        delta = R.penalty(20.0, 1.0, 2.5)
        print("[test] noalloc call 1: 200000 walkers, 1500 echoes, mirror, 64x1", flush=True)
        sa = _run(ctx, img, "noalloc", a, e1, 3, "mirror", delta)
        sb = _run(ctx, img, "run", b, e1, 3, "mirror", delta)
        _compare(a, b, sa, sb)
        # call 2: fewer walkers, another launch shape and boundary; rwnmr_cpmg_run above built
        # the unfolded images it uses, so this call walks them
        ctx.set_launch(256, 1)
        for bc, form in (("mirror", 3), ("noflux", 3), ("mirror", 6)):
            # image form 6: the 8-byte chunk cache of mirror walks
            ctx.set_unfolded(form)
            base = _device_walkers(torch, solid, 23, 3000, with_hits=False)
            a, b = _clone(base), _clone(base)
            sb = _run(ctx, img, "run", b, 200, 3, bc, delta)
            print(f"[test] noalloc call 2: 3000 walkers, 200 echoes, {bc}, 256x1, form {form}", flush=True)
            sa = _run(ctx, img, "noalloc", a, 200, 3, bc, delta)
            _compare(a, b, sa, sb)
        ctx.set_unfolded(3)
        img.close()
    finally:
        ctx.close()


def test_noalloc_rejects_what_it_cannot_run():
    import torch
    L = load()
    ctx = R.Context(0)
    try:
        solid, _ = R.sphere_pack(32, 32, 32, 4.0, 4.0, 0.3, 3)
        img = R.Image.from_solid(ctx, solid)
        seeds, x, y, z = R.seed_and_place(solid, 3, 100)
        # host arrays
        w = _walker_struct(x, y, z, seeds, None, None, None, 0.99, 100, 0)
        assert L.rwnmr_cpmg_run_noalloc(ctx.handle, img.handle, ct.byref(w), 10, 3, bc_code("mirror"), 0,
                                        None) == 1
        assert b"device memory" in L.rwnmr_last_error(ctx.handle)
        # a multi-launch run without an energy array (64 x 1 launch: 200,000 walkers, 3125 blocks)
        ctx.set_launch(64, 1)
        t = _device_walkers(torch, solid, 3, 200_000, with_hits=False)
        sums = torch.zeros(1501, dtype=torch.float64, device="cuda:0")
        w = _walker_struct(t["x"], t["y"], t["z"], t["seed"], None, None, None, 0.99, 200_000, 1)
        assert L.rwnmr_cpmg_run_noalloc(ctx.handle, img.handle, ct.byref(w), 1500, 3, bc_code("mirror"), 0,
                                        ct.c_void_p(sums.data_ptr())) == 1
        assert b"energy array" in L.rwnmr_last_error(ctx.handle)
        # more blocks than the scratch holds two partial rows for: 67,200,000 walkers at 64 x 1
        # are 1,050,000 blocks (the limit is 1,048,560); rejected before any launch
        n = 67_200_000
        big = {k: torch.zeros(n, dtype=torch.int32, device="cuda:0") for k in ("x", "y", "z")}
        bseed = torch.ones(n, dtype=torch.int64, device="cuda:0")
        w = _walker_struct(big["x"], big["y"], big["z"], bseed, None, None, None, 0.99, n, 1)
        assert L.rwnmr_cpmg_run_noalloc(ctx.handle, img.handle, ct.byref(w), 10, 3, bc_code("mirror"), 0,
                                        ct.c_void_p(sums.data_ptr())) == 1
        assert b"scratch holds" in L.rwnmr_last_error(ctx.handle)
        ctx.synchronize()
        del big, bseed
        img.close()
    finally:
        ctx.close()
