"""HIP path vs the CPU oracle on the same seeded inputs (GPU tests).

Bar: walker positions, seeds, hit counts and per-walker energies bit-exact
(energy is a sequential product of the same doubles in the same order);
echo sums within 1e-12 relative of the oracle's compensated long-double sum;
PFGSE M(k, Delta) within 1e-11 of M(0, Delta) (GPU and libm cos/pow may
differ in the last ulp per term).
"""
import numpy as np
import pytest

import oracle as O
import randomwalk_nmr_amd as R

pytestmark = pytest.mark.gpu

ECHO_RTOL = 1e-12


def _case(W, H, D, r, phi, n, seed, rho_spread=True):
    solid, _ = R.sphere_pack(W, H, D, r, r, phi, seed)
    seeds, x, y, z = R.seed_and_place(solid, seed + 7, n)
    rng = np.random.default_rng(seed)
    pen = 1.0 - rng.uniform(0.001, 0.02, n) if rho_spread else np.full(n, 0.99)
    return solid, seeds, x, y, z, pen


def _assert_walkers(g, o):
    for k in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(np.asarray(g[k]), np.asarray(o[k])), k
    if g.get("energy") is not None and o.get("energy") is not None:
        assert np.array_equal(np.asarray(g["energy"]).view(np.uint64),
                              np.asarray(o["energy"]).view(np.uint64)), "energy bits"


def _assert_echoes(ge, oe):
    ge, oe = np.asarray(ge), np.asarray(oe)
    assert ge.shape == oe.shape
    assert np.all(np.abs(ge - oe) <= ECHO_RTOL * np.abs(oe) + 1e-300), np.max(np.abs(ge - oe) / np.abs(oe))


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
@pytest.mark.parametrize("shape", [(256, 1), (1024, 1), (64, 2), (512, 2), (128, 4), (256, 4)])
def test_cpmg_matches_oracle(ctx, bc, shape):
    print(f"[test] cpmg bc={bc} launch={shape}", flush=True)
    solid, seeds, x, y, z, pen = _case(30, 21, 17, 3.5, 0.35, 1037, 11)
    img = R.Image.from_solid(ctx, solid)
    ctx.set_launch(*shape)
    try:
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=40, steps_per_echo=3, bc=bc, penalty=pen)
    finally:
        ctx.set_launch(256, 1)
    o = O.walk(O.bitblock_pack(solid), (30, 21, 17), x, y, z, seeds, echoes=40, steps_per_echo=3, bc=bc,
               penalty=pen)
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
@pytest.mark.parametrize("layout", [(0, 0, 0), (1, 1, 1), (2, 1, 0)])
@pytest.mark.parametrize("shape", [(256, 1), (128, 2)])
def test_locality_order_and_brick_layout_match_oracle(ctx, bc, layout, shape):
    """Morton-sorted walkers with XCD-striped tiles (re-sorted every 7
    echoes) on a brick-layout image: walkers bit-exact, sums to 1e-12."""
    print(f"[test] locality bc={bc} layout={layout} launch={shape}", flush=True)
    solid, seeds, x, y, z, pen = _case(37, 22, 29, 3.5, 0.35, 5003, 17)
    ctx.set_image_layout(*layout)
    ctx.set_locality(True, resort_every=7, min_walkers=0)
    ctx.set_launch(*shape)
    try:
        img = R.Image.from_solid(ctx, solid)
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=30, steps_per_echo=3, bc=bc, penalty=pen)
    finally:
        ctx.set_image_layout(1, 1, 1)
        ctx.set_locality(True, resort_every=1000, min_walkers=65536)
        ctx.set_launch(256, 1)
    o = O.walk(O.bitblock_pack(solid), (37, 22, 29), x, y, z, seeds, echoes=30, steps_per_echo=3, bc=bc, penalty=pen)
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
@pytest.mark.parametrize("dims", [(32, 16, 8), (16, 16, 16), (30, 21, 17), (9, 1, 5), (2, 3, 64)])
@pytest.mark.parametrize("unfolded", [6, 5, 4, 3, 2, 1, 0])
@pytest.mark.parametrize("shape", [(256, 1), (512, 2), (256, 4)])
def test_unfolded_and_folded_walks_match_oracle(ctx, bc, dims, unfolded, shape):
    """Noflux / periodic / mirror CPMG and map walks on the unfolded image
    with and without the per-walker chunk cache (power-of-two periods: plain
    dilated wrap; others: wrap fix; noflux: face test and turn-around on the
    periodic image, folded kernel when a dimension is 1), the speculative
    two-candidate CPMG walk (form 5, one walker per lane), the 8-byte chunk
    cache of mirror walks (form 6; other walks run form 3) and the folded
    kernel: walkers bit-exact, echo sums to 1e-12, for every launch shape."""
    print(f"[test] unfolded={unfolded} bc={bc} dims={dims} launch={shape}", flush=True)
    W, H, D = dims
    rng = np.random.default_rng(W * 10000 + H * 100 + D)
    solid = (rng.random((D, H, W)) < 0.3).astype(np.uint8)
    n = 1500
    seeds, x, y, z = R.seed_and_place(solid, 31 + W, n)
    pen = 1.0 - rng.uniform(0.001, 0.02, n)
    ctx.set_unfolded(unfolded)
    ctx.set_launch(*shape)
    try:
        img = R.Image.from_solid(ctx, solid)
        if bc == "noflux" and min(dims) < 2:
            # noflux is refused on an image with a dimension of 1 (the walk
            # would step outside it)
            with pytest.raises(R.RwnmrError):
                R.cpmg_run(ctx, img, x, y, z, seeds, echoes=2, steps_per_echo=1, bc=bc, penalty=pen)
            return
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=37, steps_per_echo=5, bc=bc, penalty=pen)
        m = R.map_run(ctx, img, x, y, z, seeds, steps=37 * 5, bc=bc)
    finally:
        ctx.set_unfolded(3)
        ctx.set_launch(256, 1)
    o = O.walk(O.bitblock_pack(solid), dims, x, y, z, seeds, echoes=37, steps_per_echo=5, bc=bc, penalty=pen)
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])
    for k in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(m[k], o[k]), k


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
@pytest.mark.parametrize("spe", [1, 600])
@pytest.mark.parametrize("unfolded", [6, 5, 4, 3, 2])
def test_unfolded_long_echoes_and_locality_match_oracle(ctx, bc, spe, unfolded):
    """Steps per echo above the packed-displacement flush interval (511) and
    the Z-order re-sort of the unfolded walk on a power-of-two image."""
    print(f"[test] unfolded locality bc={bc} spe={spe} unfolded={unfolded}", flush=True)
    solid, seeds, x, y, z, pen = _case(32, 32, 32, 3.5, 0.35, 4001, 23)
    echoes = 9 if spe > 1 else 400
    ctx.set_locality(True, resort_every=4, min_walkers=0)
    ctx.set_unfolded(unfolded)
    try:
        img = R.Image.from_solid(ctx, solid)
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=echoes, steps_per_echo=spe, bc=bc, penalty=pen)
    finally:
        ctx.set_locality(True, resort_every=1000, min_walkers=65536)
        ctx.set_unfolded(3)
    o = O.walk(O.bitblock_pack(solid), (32, 32, 32), x, y, z, seeds, echoes=echoes, steps_per_echo=spe, bc=bc,
               penalty=pen)
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])


def _mirror_images(solid, x, y, z, seed):
    """The walkers moved into random mirror images (-2 .. 2 per axis) of
    their start voxels: an odd image holds the reflection W - 1 - x, so every
    start folds back to the same pore voxel and carries nonzero reflection
    flags."""
    D, H, W = solid.shape
    rng = np.random.default_rng(seed)
    out = []
    for c, L in ((x, W), (y, H), (z, D)):
        m = rng.integers(-2, 3, len(c))
        out.append((m * L + np.where(m % 2 == 0, c, L - 1 - c)).astype(np.int32))
    return out


@pytest.mark.parametrize("dims", [(32, 16, 8), (30, 21, 17)])
@pytest.mark.parametrize("shape", [(256, 1), (512, 1), (1024, 1), (512, 2)])
@pytest.mark.parametrize("unfolded", [3, 6])
def test_mirror_reflection_flags_match_oracle(ctx, dims, shape, unfolded):
    """Mirror CPMG and map walks from starts in every mirror image, so each
    launch folds its walkers with all 8 reflection-flag combinations, for
    the launch shapes of both direction-table forms (blocks of 512 threads
    or more read the flag-keyed tables) and periods with and without the
    wrap fix-up; a locality re-sort every 5 echoes makes later launches
    start from the walkers' moved positions."""
    print(f"[test] mirror flags dims={dims} launch={shape} unfolded={unfolded}", flush=True)
    W, H, D = dims
    solid, seeds, x, y, z, pen = _case(W, H, D, 3.5, 0.35, 3001, 41)
    x, y, z = _mirror_images(solid, x, y, z, 43)
    ctx.set_launch(*shape)
    ctx.set_locality(True, resort_every=5, min_walkers=0)
    ctx.set_unfolded(unfolded)
    try:
        img = R.Image.from_solid(ctx, solid)
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=23, steps_per_echo=7, bc="mirror", penalty=pen)
        m = R.map_run(ctx, img, x, y, z, seeds, steps=23 * 7, bc="mirror")
    finally:
        ctx.set_locality(True, resort_every=1000, min_walkers=65536)
        ctx.set_launch(256, 1)
        ctx.set_unfolded(3)
    o = O.walk(O.bitblock_pack(solid), dims, x, y, z, seeds, echoes=23, steps_per_echo=7, bc="mirror", penalty=pen)
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])
    for k in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(m[k], o[k]), k


@pytest.mark.parametrize("dims", [(512, 512, 512), (512, 520, 512)])
@pytest.mark.parametrize("unfolded", [3, 6])
def test_mirror_address_displacements_match_oracle(ctx, dims, unfolded):
    """Mirror CPMG of 1024-thread blocks on images whose mirror periods all
    exceed twice the flush interval (1022 steps): the unbounded moves come
    from the walkers' address changes between flushes (power-of-two periods,
    and a period of 1040 with the wrap fix-up), walkers starting in mirror
    images from -2 to 2 per axis, re-sorts splitting the run into launches:
    walkers bit-exact, echo sums to 1e-12.  The same walk on device arrays
    with no hit array (the kernel then counts no hits, as the reference's
    CPMG kernels) gives the same positions, seeds and energies."""
    import torch
    print(f"[test] mirror address displacements dims={dims} unfolded={unfolded}", flush=True)
    W, H, D = dims
    solid, seeds, x, y, z, pen = _case(W, H, D, 8.0, 0.25, 3000, 47)
    x, y, z = _mirror_images(solid, x, y, z, 49)
    ctx.set_launch(1024, 1)
    ctx.set_locality(True, resort_every=300, min_walkers=0)
    ctx.set_unfolded(unfolded)
    try:
        img = R.Image.from_solid(ctx, solid)
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=700, steps_per_echo=3, bc="mirror", penalty=pen)
        dev = [torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") for a in (x, y, z, seeds.view(np.int64), pen)]
        en = torch.ones(len(x), dtype=torch.float64, device="cuda:0")
        sums = torch.zeros(701, dtype=torch.float64, device="cuda:0")
        R.cpmg_run(ctx, img, dev[0], dev[1], dev[2], dev[3], echoes=700, steps_per_echo=3, bc="mirror",
                   penalty=dev[4], energy=en, echo_sums=sums)
        img.close()
    finally:
        ctx.set_locality(True, resort_every=1000, min_walkers=65536)
        ctx.set_launch(256, 1)
        ctx.set_unfolded(3)
    o = O.walk(O.bitblock_pack(solid), dims, x, y, z, seeds, echoes=700, steps_per_echo=3, bc="mirror", penalty=pen)
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])
    for k, T in (("x", dev[0]), ("y", dev[1]), ("z", dev[2])):
        assert np.array_equal(T.cpu().numpy(), o[k]), k
    assert np.array_equal(dev[3].cpu().numpy().view(np.uint64), o["seed"])
    assert np.array_equal(en.cpu().numpy().view(np.uint64), np.asarray(o["energy"]).view(np.uint64))
    _assert_echoes(sums.cpu().numpy(), o["echo_sums"])


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
def test_map_matches_oracle(ctx, bc):
    print(f"[test] map bc={bc}", flush=True)
    solid, seeds, x, y, z, _ = _case(24, 24, 24, 4, 0.3, 777, 5)
    img = R.Image.from_solid(ctx, solid)
    g = R.map_run(ctx, img, x, y, z, seeds, steps=501, bc=bc)
    o = O.walk(O.bitblock_pack(solid), (24, 24, 24), x, y, z, seeds, echoes=1, steps_per_echo=501, bc=bc,
               echo_sums=False)
    for k in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(g[k], o[k]), k


def test_map_hits_equal_cpmg_wall_events(ctx):
    """The map replay and the CPMG run share trajectories (same seeds), so
    the hit counts agree and E == delta^hits by sequential product."""
    print("[test] map vs cpmg hits", flush=True)
    solid, seeds, x, y, z, _ = _case(32, 32, 32, 4, 0.3, 500, 9)
    img = R.Image.from_solid(ctx, solid)
    d = 0.995
    c = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=30, steps_per_echo=4, bc="mirror", delta=d)
    m = R.map_run(ctx, img, x, y, z, seeds, steps=120, bc="mirror")
    assert np.array_equal(c["hits"], m["hits"])
    table = np.ones(int(m["hits"].max()) + 1)
    for i in range(1, len(table)):
        table[i] = table[i - 1] * d
    assert np.array_equal(c["energy"], table[m["hits"]])


def test_cpmg_labels_match_oracle(ctx):
    print("[test] cpmg labels", flush=True)
    solid, seeds, x, y, z, _ = _case(28, 28, 28, 4, 0.3, 900, 3)
    labels = R.voronoi_labels(solid, 16, 8, 99)
    img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
    ld = np.linspace(0.97, 0.999, 8)
    g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=25, steps_per_echo=4, bc="mirror", label_delta=ld,
                   label_hits=True)
    o = O.walk(O.bitblock_pack(solid), (28, 28, 28), x, y, z, seeds, echoes=25, steps_per_echo=4, bc="mirror",
               labels=labels, label_delta=ld, label_hits=True)
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])
    assert np.array_equal(g["label_hits"], o["label_hits"].T)


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
@pytest.mark.parametrize("dims", [(28, 28, 28), (32, 16, 8)])
@pytest.mark.parametrize("unfolded", [6, 4, 3, 2, 1, 0])
def test_cpmg_labels_unfolded_and_folded_match_oracle(ctx, bc, dims, unfolded):
    """Per-label CPMG on the unfolded image (4-bit label volume) and on the
    folded kernel: walkers, energies and per-label hit counts bit-exact."""
    print(f"[test] labels unfolded={unfolded} bc={bc} dims={dims}", flush=True)
    W, H, D = dims
    solid, seeds, x, y, z, _ = _case(W, H, D, 3.5, 0.3, 1100, 13)
    labels = R.voronoi_labels(solid, 16, 8, 7)
    ld = np.linspace(0.97, 0.999, 8)
    ctx.set_unfolded(unfolded)
    try:
        img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=21, steps_per_echo=5, bc=bc, label_delta=ld,
                       label_hits=True)
    finally:
        ctx.set_unfolded(3)
    o = O.walk(O.bitblock_pack(solid), dims, x, y, z, seeds, echoes=21, steps_per_echo=5, bc=bc,
               labels=labels, label_delta=ld, label_hits=True)
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])
    assert np.array_equal(g["label_hits"], o["label_hits"].T)


@pytest.mark.parametrize("side", [16, 64, 512])
def test_cpmg_labels_octant_fold_match_oracle(ctx, side):
    """Mirror images of equal power-of-two sides keep one byte per voxel of
    the first octant only and read a hit's label at the folded address (one
    extract, a 24-bit multiply-add and a masked xor; at 512^3 the multiplier's
    top term is added as a shift).  Walkers start next to the faces so that
    their candidate voxels cross into the other octants of the mirror image;
    walkers, energies and per-label hit counts bit-exact."""
    print(f"[test] labels octant fold side={side}", flush=True)
    solid, _ = R.sphere_pack(side, side, side, 3.0 if side < 64 else 6.0, 3.0 if side < 64 else 12.0, 0.3, 31)
    labels = R.voronoi_labels(solid, 16, 8, 37)
    seeds, x, y, z = R.seed_and_place(solid, 41, 20000 if side < 512 else 200000)
    if side == 512:
        near = np.minimum.reduce([x, y, z, side - 1 - x, side - 1 - y, side - 1 - z]) < 3
        keep = np.flatnonzero(near)[:3000]
        seeds, x, y, z = seeds[keep], x[keep], y[keep], z[keep]
    else:
        seeds, x, y, z = seeds[:3000], x[:3000], y[:3000], z[:3000]
    ld = np.linspace(0.95, 0.999, 8)
    img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
    g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=30, steps_per_echo=4, bc="mirror", label_delta=ld,
                   label_hits=True)
    img.close()
    o = O.walk(O.bitblock_pack(solid), (side, side, side), x, y, z, seeds, echoes=30, steps_per_echo=4,
               bc="mirror", labels=labels, label_delta=ld, label_hits=True)
    # the walks left the first octant (reflected copies) and hit walls there
    assert np.any((np.asarray(o["x"]) < 0) | (np.asarray(o["x"]) >= side) | (np.asarray(o["y"]) < 0) |
                  (np.asarray(o["z"]) >= side))
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])
    assert np.array_equal(g["label_hits"], o["label_hits"].T)


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
@pytest.mark.parametrize("shape", [(256, 1), (1024, 1), (512, 2), (256, 4)])
def test_cpmg_labels_launch_shapes_match_oracle(ctx, bc, shape):
    """Per-label CPMG on the unfolded image for every launch shape: one
    walker per lane keeps the per-label counts in LDS, wider lanes in
    registers; both bit-exact."""
    print(f"[test] labels launch={shape} bc={bc}", flush=True)
    solid, seeds, x, y, z, _ = _case(30, 21, 17, 3.0, 0.3, 2600, 17)
    labels = R.voronoi_labels(solid, 16, 8, 5)
    ld = np.linspace(0.95, 0.999, 8)
    ctx.set_launch(*shape)
    try:
        img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
        g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=19, steps_per_echo=6, bc=bc, label_delta=ld,
                       label_hits=True)
    finally:
        ctx.set_launch(256, 1)
    o = O.walk(O.bitblock_pack(solid), (30, 21, 17), x, y, z, seeds, echoes=19, steps_per_echo=6, bc=bc,
               labels=labels, label_delta=ld, label_hits=True)
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])
    assert np.array_equal(g["label_hits"], o["label_hits"].T)


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
@pytest.mark.parametrize("shape", [(256, 1), (512, 2), (256, 4)])
@pytest.mark.parametrize("spe", [4, 7])
def test_internal_field_cpmg_matches_oracle(ctx, bc, shape, spe):
    """Internal-field CPMG: positions, seeds, hits, p_alive and phases
    bit-exact (the phase update is one fused multiply-add on both sides);
    echo sums of p_alive cos(phase) within 1e-12 relative."""
    print(f"[test] field bc={bc} launch={shape} spe={spe}", flush=True)
    solid, seeds, x, y, z, pen = _case(26, 19, 23, 3.5, 0.35, 1200, 51)
    rng = np.random.default_rng(spe)
    field = rng.normal(0.0, 1e-3, solid.shape)
    field[solid == 1] = 0.0
    gyro, dt_ms = 42.576, 0.0333
    gt = 0.5 * (1.0e6 * gyro) * (1.0e-3 * dt_ms)
    img = R.Image.from_solid(ctx, solid)
    ctx.set_launch(*shape)
    try:
        g = R.cpmg_field_run(ctx, img, x, y, z, seeds, echoes=33, steps_per_echo=spe, field=field,
                             gyromagnetic_ratio=gyro, time_interval_ms=dt_ms, bc=bc, penalty=pen)
    finally:
        ctx.set_launch(256, 1)
    o = O.walk_field(O.bitblock_pack(solid), (26, 19, 23), x, y, z, seeds, echoes=33, steps_per_echo=spe,
                     field=field, gammatau=gt, bc=bc, penalty=pen)
    for k in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(g["p_alive"], o["p_alive"])
    assert np.array_equal(g["phase"], o["phase"])
    _assert_echoes(g["echo_sums"], o["echo_sums"])


def test_internal_field_state_carries_across_calls(ctx):
    """Two calls of 10 + 15 echoes continuing p_alive / phase == one call of 25."""
    print("[test] field continuation", flush=True)
    solid, seeds, x, y, z, pen = _case(20, 20, 20, 3, 0.35, 700, 61)
    field = np.random.default_rng(1).normal(0.0, 1e-3, solid.shape)
    img = R.Image.from_solid(ctx, solid)
    kw = dict(field=field, gyromagnetic_ratio=42.576, time_interval_ms=0.05, bc="mirror", penalty=pen)
    a = R.cpmg_field_run(ctx, img, x, y, z, seeds, echoes=10, steps_per_echo=6, **kw)
    b = R.cpmg_field_run(ctx, img, a["x"], a["y"], a["z"], a["seed"], echoes=15, steps_per_echo=6,
                         p_alive=a["p_alive"], phase=a["phase"], hits=a["hits"], **kw)
    c = R.cpmg_field_run(ctx, img, x, y, z, seeds, echoes=25, steps_per_echo=6, **kw)
    for k in ("x", "y", "z", "seed", "hits", "p_alive", "phase"):
        assert np.array_equal(b[k], c[k]), k
    assert np.array_equal(b["echo_sums"], c["echo_sums"][10:])


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
@pytest.mark.parametrize("log_scale", [False, True])
@pytest.mark.parametrize("device", [False, True])
def test_histogram_method_matches_oracle(ctx, bc, log_scale, device):
    """Collision-histogram CPMG: the GPU map walk per segment (fresh counts,
    walkers carried over) feeding the histogram arithmetic == oracle map
    walks feeding the oracle's restatement of the histogram method."""
    print(f"[test] histogram method bc={bc} log={log_scale} device={device}", flush=True)
    from oracle import histogram as OH
    solid, seeds, x, y, z, _ = _case(24, 22, 20, 3.5, 0.35, 3000, 71)
    echoes, spe, nh, size = 40, 4, 3, 16
    img = R.Image.from_solid(ctx, solid)
    if device:
        import torch
        dev = torch.device("cuda:0")
        X, Y, Z = (torch.from_numpy(a.copy()).to(dev) for a in (x, y, z))
        S = torch.from_numpy(seeds.view(np.int64).copy()).to(dev)
        g = R.cpmg_histogram_simulate(ctx, img, X, Y, Z, S, echoes=echoes, steps_per_echo=spe, bc=bc,
                                      n_histograms=nh, size=size, log_scale=log_scale, rho=15.0, step_length=1.0,
                                      D0=2.5)
        gx = X.cpu().numpy()
    else:
        g = R.cpmg_histogram_simulate(ctx, img, x, y, z, seeds, echoes=echoes, steps_per_echo=spe, bc=bc,
                                      n_histograms=nh, size=size, log_scale=log_scale, rho=15.0, step_length=1.0,
                                      D0=2.5)
        gx = g["x"]
    segs = OH.histogram_list(echoes, nh)
    blocks = O.bitblock_pack(solid)
    cx, cy, cz, cs = x, y, z, seeds
    total = np.zeros(len(x), np.int64)
    amps = []
    for f, l in segs:
        o = O.walk(blocks, (24, 22, 20), cx, cy, cz, cs, echoes=1, steps_per_echo=spe * (l - f), bc=bc,
                   echo_sums=False)
        cx, cy, cz, cs = o["x"], o["y"], o["z"], o["seed"]
        total += o["hits"]
        amps.append(OH.collision_histogram([int(h) for h in o["hits"]], spe * (l - f), size, log_scale)[1])
    bins, _ = OH.collision_histogram([int(h) for h in total], segs[-1][1] * spe, size, log_scale)
    sig = OH.decay(amps, segs, OH.penalties(bins, 15.0, 1.0, 2.5, spe))
    assert np.array_equal(gx, cx)
    assert np.array_equal(g["bins"], bins) and np.array_equal(g["amps"], np.array(amps))
    assert np.allclose(g["signal"], sig, rtol=1e-13, atol=0)


@pytest.mark.parametrize("bc", ["noflux", "mirror"])
@pytest.mark.parametrize("n,n_samples", [(3000, 10), (2999, 7), (700, 3)])
@pytest.mark.parametrize("unfolded", [3, 0])
def test_pfgse_sampled_matches_oracle_per_sample(ctx, bc, n, n_samples, unfolded):
    """Walker samples of N // n_samples consecutive walkers (the remainder
    walks in no sample): each sample's M(k, Delta) == the oracle on that
    sample alone (1e-11 M0); all-walker M and every walker's state exact."""
    print(f"[test] pfgse sampled bc={bc} n={n} samples={n_samples} unfolded={unfolded}", flush=True)
    solid, seeds, x, y, z, _ = _case(28, 24, 20, 4, 0.35, n, 81)
    steps_inc = np.array([3, 9, 27], np.int64)
    kx = np.stack([np.linspace(0.0, 1.2 / np.sqrt(c), 10) for c in np.cumsum(steps_inc)])
    zero = np.zeros_like(kx)
    sets = np.array([1.0, 0.98])
    ctx.set_unfolded(unfolded)
    try:
        img = R.Image.from_solid(ctx, solid)
        g = R.pfgse_run_sampled(ctx, img, x, y, z, seeds, n_samples=n_samples, steps_inc=steps_inc, kx=kx, ky=zero,
                                kz=zero, voxel=1.0, bc=bc, set_delta=sets)
    finally:
        ctx.set_unfolded(3)
    blocks = O.bitblock_pack(solid)
    full = O.pfgse(blocks, (28, 24, 20), x, y, z, seeds, steps_inc=steps_inc, kx=kx, ky=zero, kz=zero, voxel=1.0,
                   bc=bc, set_delta=sets)
    for k in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(g[k], full[k]), k
    assert np.all(np.abs(g["M"] - full["M"]) <= 1e-11 * np.abs(full["M"][..., :1]))
    wps = n // n_samples
    for smp in range(n_samples):
        sl = slice(smp * wps, (smp + 1) * wps)
        o = O.pfgse(blocks, (28, 24, 20), x[sl], y[sl], z[sl], seeds[sl], steps_inc=steps_inc, kx=kx, ky=zero,
                    kz=zero, voxel=1.0, bc=bc, set_delta=sets)
        assert np.all(np.abs(g["M_samples"][smp] - o["M"]) <= 1e-11 * np.abs(o["M"][..., :1])), smp


@pytest.mark.parametrize("bc", ["noflux", "mirror"])
def test_pfgse_matches_oracle(ctx, bc):
    print(f"[test] pfgse bc={bc}", flush=True)
    solid, seeds, x, y, z, pen = _case(32, 32, 32, 5, 0.35, 1500, 21)
    img = R.Image.from_solid(ctx, solid)
    steps_inc = np.array([3, 5, 9, 20, 40], np.int64)
    kmax = 1.3
    kx = np.array([j * kmax / 15 for j in range(16)])
    ky = np.zeros(16)
    kz = 0.05 * np.arange(16)
    g = R.pfgse_run(ctx, img, x, y, z, seeds, steps_inc=steps_inc, kx=kx, ky=ky, kz=kz, voxel=0.7, bc=bc,
                    penalty=pen)
    o = O.pfgse(O.bitblock_pack(solid), (32, 32, 32), x, y, z, seeds, steps_inc=steps_inc, kx=kx, ky=ky, kz=kz,
                voxel=0.7, bc=bc, penalty=pen)
    for k in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(g[k], o[k]), k
    assert np.allclose(g["energy"], o["energy"], rtol=1e-13, atol=0)
    M0 = o["M"][..., :1]
    assert np.all(np.abs(g["M"] - o["M"]) <= 1e-11 * np.abs(M0))
    assert np.array_equal(g["msd"], o["msd"])  # integer-valued sums: exact


@pytest.mark.parametrize("dims", [(32, 32, 32), (30, 21, 17)])
@pytest.mark.parametrize("kind", ["arithmetic", "random"])
def test_pfgse_mirror_reflection_flags_match_oracle(ctx, dims, kind):
    """Mirror PFGSE from starts in every mirror image: the walk folds each
    walker into the first octant with its reflection flags (8 flag
    combinations; periods with and without the wrap fix-up), two energy
    sets; walkers bit-exact, M within 1e-11 M(0), msd exact."""
    print(f"[test] pfgse mirror flags dims={dims} kind={kind}", flush=True)
    W, H, D = dims
    solid, seeds, x, y, z, _ = _case(W, H, D, 4, 0.35, 2500, 51)
    x, y, z = _mirror_images(solid, x, y, z, 53)
    img = R.Image.from_solid(ctx, solid)
    steps_inc = np.array([2, 7, 15, 30, 46], np.int64)
    if kind == "arithmetic":
        kx = np.linspace(0.0, 1.1, 16)
    else:
        kx = np.random.default_rng(3).uniform(0.0, 1.1, 16)
    ky = 0.02 * np.arange(16)
    kz = np.zeros(16)
    sd = np.array([1.0, 0.985])
    g = R.pfgse_run(ctx, img, x, y, z, seeds, steps_inc=steps_inc, kx=kx, ky=ky, kz=kz, voxel=0.9, bc="mirror",
                    set_delta=sd)
    o = O.pfgse(O.bitblock_pack(solid), dims, x, y, z, seeds, steps_inc=steps_inc, kx=kx, ky=ky, kz=kz,
                voxel=0.9, bc="mirror", set_delta=sd)
    for k in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(g[k], o[k]), k
    assert np.all(np.abs(g["M"] - o["M"]) <= 1e-11 * np.abs(o["M"][..., :1]))
    assert np.array_equal(g["msd"], o["msd"])
    img.close()


def test_pfgse_k_set_per_diffusion_time(ctx):
    """k_max scaled per diffusion time (k sqrt(D0 Delta) = const)."""
    print("[test] pfgse k per delta", flush=True)
    solid, seeds, x, y, z, _ = _case(32, 32, 32, 5, 0.35, 1200, 31)
    img = R.Image.from_solid(ctx, solid)
    steps_inc = np.array([4, 8, 16, 32], np.int64)
    cum = np.cumsum(steps_inc)
    kmax = 3.0 / np.sqrt(cum.astype(float))
    kx = np.stack([np.linspace(0, km, 12) for km in kmax])
    zeros = np.zeros_like(kx)
    g = R.pfgse_run(ctx, img, x, y, z, seeds, steps_inc=steps_inc, kx=kx, ky=zeros, kz=zeros, voxel=1.0,
                    set_delta=np.array([1.0, 0.99]))
    o = O.pfgse(O.bitblock_pack(solid), (32, 32, 32), x, y, z, seeds, steps_inc=steps_inc, kx=kx, ky=zeros,
                kz=zeros, voxel=1.0, set_delta=np.array([1.0, 0.99]))
    assert np.all(np.abs(g["M"] - o["M"]) <= 1e-11 * np.abs(o["M"][..., :1]))
    assert np.array_equal(g["msd"], o["msd"])


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
@pytest.mark.parametrize("kind", ["arithmetic", "random"])
@pytest.mark.parametrize("unfolded", [3, 0])
def test_pfgse_phase_paths_and_kernels_match_oracle(ctx, bc, kind, unfolded):
    """Arithmetic k rows (phases by complex rotation) and irregular k rows
    (cos per k), on the unfolded-image and folded PFGSE kernels, with a
    non-power-of-two image and two energy sets: M within 1e-11 M(0), walkers
    and msd exact."""
    print(f"[test] pfgse phase={kind} bc={bc} unfolded={unfolded}", flush=True)
    solid, seeds, x, y, z, pen = _case(27, 19, 33, 4, 0.35, 1300, 41)
    steps_inc = np.array([5, 7, 30, 300, 600], np.int64)
    rng = np.random.default_rng(5)
    if kind == "arithmetic":
        kx = np.stack([np.linspace(0.0, 2.0 / np.sqrt(c), 20) for c in np.cumsum(steps_inc)])
        ky = 0.3 * kx
        kz = np.stack([np.linspace(0.1, 0.4, 20)] * len(steps_inc))
    else:
        kx = rng.uniform(-1, 1, (len(steps_inc), 20))
        ky = rng.uniform(-1, 1, (len(steps_inc), 20))
        kz = rng.uniform(-1, 1, (len(steps_inc), 20))
    sets = np.array([1.0, 0.985])
    ctx.set_unfolded(unfolded)
    try:
        img = R.Image.from_solid(ctx, solid)
        g = R.pfgse_run(ctx, img, x, y, z, seeds, steps_inc=steps_inc, kx=kx, ky=ky, kz=kz, voxel=0.8, bc=bc,
                        set_delta=sets)
    finally:
        ctx.set_unfolded(3)
    o = O.pfgse(O.bitblock_pack(solid), (27, 19, 33), x, y, z, seeds, steps_inc=steps_inc, kx=kx, ky=ky, kz=kz,
                voxel=0.8, bc=bc, set_delta=sets)
    for k in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(g[k], o[k]), k
    assert np.all(np.abs(g["M"] - o["M"]) <= 1e-11 * np.abs(o["M"][..., :1]))
    assert np.array_equal(g["msd"], o["msd"])


@pytest.mark.parametrize("n_sets", [3, 8])
def test_pfgse_energy_sets(ctx, n_sets):
    """Several uniform deltas on one trajectory equal separate runs (sets
    summed over the wave in pairs, a lone last set on its own)."""
    print(f"[test] pfgse sets {n_sets}", flush=True)
    solid, seeds, x, y, z, _ = _case(24, 24, 24, 4, 0.3, 800, 4)
    img = R.Image.from_solid(ctx, solid)
    si = np.array([4, 8, 16], np.int64)
    kx = np.linspace(0, 1.0, 12)
    zeros = np.zeros(12)
    sets = np.linspace(1.0, 0.95, n_sets)
    g = R.pfgse_run(ctx, img, x, y, z, seeds, steps_inc=si, kx=kx, ky=zeros, kz=zeros, voxel=1.0,
                    set_delta=sets)
    for q, d in enumerate(sets):
        o = O.pfgse(O.bitblock_pack(solid), (24, 24, 24), x, y, z, seeds, steps_inc=si, kx=kx, ky=zeros, kz=zeros,
                    voxel=1.0, penalty=d)
        assert np.all(np.abs(g["M"][q] - o["M"][0]) <= 1e-11 * np.abs(o["M"][0][:, :1]))


def test_noflux_walker_outside_image_is_an_error(ctx):
    print("[test] noflux outside", flush=True)
    solid, seeds, x, y, z, _ = _case(16, 16, 16, 3, 0.4, 64, 2)
    img = R.Image.from_solid(ctx, solid)
    x = x.copy()
    x[5] = 16
    with pytest.raises(R.RwnmrError):
        R.cpmg_run(ctx, img, x, y, z, seeds, echoes=2, steps_per_echo=1, bc="noflux")


def test_config1_small_cpmg(ctx):
    """Plain CPMG, small: 128^3 sphere pack r=8 porosity 0.25, 10,000 walkers,
    eps 1 um, D0 2.5, rho 10 um/s, TE 0.2 ms (3 steps), 2000 echoes, mirror."""
    print("[test] config1 128^3", flush=True)
    solid, phi = R.sphere_pack(128, 128, 128, 8, 8, 0.25, 2024)
    seeds, x, y, z = R.seed_and_place(solid, 2024, 10_000)
    img = R.Image.from_solid(ctx, solid)
    d = R.penalty(10.0, 1.0, 2.5)
    g = R.cpmg_run(ctx, img, x, y, z, seeds, echoes=2000, steps_per_echo=3, bc="mirror", delta=d)
    o = O.walk(O.bitblock_pack(solid), (128, 128, 128), x, y, z, seeds, echoes=2000, steps_per_echo=3,
               bc="mirror", penalty=d)
    _assert_walkers(g, o)
    _assert_echoes(g["echo_sums"], o["echo_sums"])
