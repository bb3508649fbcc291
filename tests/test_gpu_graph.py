"""HIP-graph CPMG runs (rwnmr_cpmg_graph_create / rwnmr_graph_launch): every
replay, with eager library and torch work and synchronisations between the
replays, equals the eager rwnmr_cpmg_run from the same state (walkers and
energies bit for bit, echo sums within 1e-12 relative)."""
import numpy as np
import pytest

import randomwalk_nmr_amd as R

pytestmark = pytest.mark.gpu

ECHO_RTOL = 1e-12


def _tensors(torch, solid, seed, n):
    seeds, x, y, z = R.seed_and_place(solid, seed, n)
    dev = torch.device("cuda:0")
    return [torch.from_numpy(a).to(dev) for a in (x, y, z)] + [torch.from_numpy(seeds.view(np.int64)).to(dev)]


def _eager(torch, ctx, img, start, echoes, bc, delta):
    x, y, z, s = (t.clone() for t in start)
    e = torch.ones(len(x), dtype=torch.float64, device="cuda:0")
    h = torch.zeros(len(x), dtype=torch.int32, device="cuda:0")
    sums = torch.zeros(echoes + 1, dtype=torch.float64, device="cuda:0")
    R.cpmg_run(ctx, img, x, y, z, s, echoes=echoes, steps_per_echo=3, bc=bc, delta=delta, energy=e, hits=h,
               echo_sums=sums)
    return [t.cpu().numpy() for t in (x, y, z, s, e, h, sums)]


def _same(got, want):
    for g, w in zip(got[:6], want[:6]):
        assert np.array_equal(g.view(np.uint64) if g.dtype == np.float64 else g,
                              w.view(np.uint64) if w.dtype == np.float64 else w)
    assert np.all(np.abs(got[6] - want[6]) <= ECHO_RTOL * np.abs(want[6]))


@pytest.mark.parametrize("bc", ["mirror", "noflux", "periodic"])
@pytest.mark.parametrize("unfolded", [3, 6])
def test_graph_replays_equal_eager_runs(ctx, bc, unfolded):
    import torch
    print(f"[test] cpmg graph bc={bc} unfolded={unfolded}", flush=True)
    solid, _ = R.sphere_pack(48, 40, 36, 3.0, 7.0, 0.3, 21)
    ctx.set_unfolded(unfolded)
    img = R.Image.from_solid(ctx, solid)
    delta = R.penalty(15.0, 1.0, 2.5)
    E = 300
    start = _tensors(torch, solid, 5, 7000)
    want = _eager(torch, ctx, img, start, E, bc, delta)
    x, y, z, s = (torch.empty_like(t) for t in start)
    e = torch.empty(7000, dtype=torch.float64, device="cuda:0")
    h = torch.empty(7000, dtype=torch.int32, device="cuda:0")
    sums = torch.zeros(E + 1, dtype=torch.float64, device="cuda:0")
    g = R.CpmgGraph(ctx, img, x, y, z, s, echoes=E, steps_per_echo=3, bc=bc, delta=delta, energy=e, hits=h,
                    echo_sums=sums, start=start)
    other = _tensors(torch, solid, 9, 3000)
    try:
        for rep in range(4):
            sums.fill_(-1.0)
            g.launch(sync=True)
            _same([t.cpu().numpy() for t in (x, y, z, s, e, h, sums)], want)
            # eager work between replays: another run through the library, torch ops, a sync
            R.cpmg_run(ctx, img, *[t.clone() for t in other], echoes=50, steps_per_echo=3, bc=bc, delta=delta)
            x.add_(1)
            e.mul_(0.5)
            torch.cuda.synchronize()
    finally:
        g.close()
        ctx.set_unfolded(3)
    img.close()


def test_graph_without_start_continues_the_walk(ctx):
    import torch
    print("[test] cpmg graph, continuing replays", flush=True)
    solid, _ = R.sphere_pack(40, 40, 40, 3.0, 6.0, 0.3, 23)
    img = R.Image.from_solid(ctx, solid)
    delta = R.penalty(25.0, 1.0, 2.5)
    start = _tensors(torch, solid, 7, 5000)
    # eager: two consecutive runs of 200 echoes
    x, y, z, s = (t.clone() for t in start)
    e = torch.ones(5000, dtype=torch.float64, device="cuda:0")
    sums_eager = []
    for _ in range(2):
        es = torch.zeros(201, dtype=torch.float64, device="cuda:0")
        R.cpmg_run(ctx, img, x, y, z, s, echoes=200, steps_per_echo=3, bc="mirror", delta=delta, energy=e,
                   echo_sums=es)
        sums_eager.append(es.cpu().numpy())
    want = [t.cpu().numpy() for t in (x, y, z, s, e)]
    gx, gy, gz, gs = (t.clone() for t in start)
    ge = torch.ones(5000, dtype=torch.float64, device="cuda:0")
    gsums = torch.zeros(201, dtype=torch.float64, device="cuda:0")
    g = R.CpmgGraph(ctx, img, gx, gy, gz, gs, echoes=200, steps_per_echo=3, bc="mirror", delta=delta, energy=ge,
                    echo_sums=gsums)
    try:
        for rep in range(2):
            g.launch(sync=True)
            got = gsums.cpu().numpy()
            assert np.all(np.abs(got - sums_eager[rep]) <= ECHO_RTOL * np.abs(sums_eager[rep])), rep
            torch.cuda.synchronize()
    finally:
        g.close()
    for a, b in zip([t.cpu().numpy() for t in (gx, gy, gz, gs, ge)], want):
        assert np.array_equal(a.view(np.uint64) if a.dtype == np.float64 else a,
                              b.view(np.uint64) if b.dtype == np.float64 else b)
    img.close()
