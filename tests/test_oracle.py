"""Pin the CPU oracle before trusting it (CPU tests, no GPU).

Each oracle routine is checked against something it does not contain:
independent numpy / pure-Python restatements, brute force on small domains,
closed forms of the random walk, the C++ standard's mt19937 value, libstdc++'s
uniform_int_distribution, and the reference's own PNG stacks.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import randomwalk_nmr_amd as R

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def py_xorshift64(s):
    s ^= (s << 13) & M64
    s ^= s >> 7
    s ^= (s << 17) & M64
    return s


# ------------------------------------------------------------------- RNG

def test_xorshift64_known_answers():
    s = 88172645463325252
    for _ in range(1000):
        s_py = py_xorshift64(s)
        assert O.xorshift64(s) == s_py
        s = s_py
    # Marsaglia (2003) example seed; value computed by the restatement above
    assert O.xorshift64(88172645463325252) == py_xorshift64(88172645463325252)
    assert O.xorshift64(1) == 1082269761


def test_mod6_digit_sum_is_remainder():
    rng = np.random.default_rng(0)
    vals = list(range(0, 5000)) + [int(v) for v in rng.integers(0, 2**63, 20000, dtype=np.int64)]
    vals += [M64, M64 - 1, 1 << 63, (1 << 63) - 1, 0xAAAAAAAAAAAAAAAA, 0x5555555555555555, 11, 12]
    for v in vals:
        assert O.mod6_digit_sum(v) == v % 6, v


# ------------------------------------------------------------------ folds

def _reflect(g, d):
    t = g % (2 * d)
    return t if t < d else 2 * d - 1 - t


@pytest.mark.parametrize("d", [1, 2, 3, 4, 7, 10])
def test_folds_brute_force(d):
    for g in range(-50, 51):
        assert O.fold_periodic(g, 0, d) == g % d
        assert O.fold_mirror(g, 0, d) == _reflect(g, d), (g, d)
    # voxel division: the image voxel is the arithmetic shift of the position
    for p in range(-64, 65):
        assert O.fold_mirror(p, 2, d) == _reflect(p >> 2, d)


# --------------------------------------------------------------- bitBlock

def numpy_pack(solid):
    D, H, W = solid.shape
    bC, bR, bD = (W + 3) // 4, (H + 3) // 4, (D + 3) // 4
    out = np.zeros(bC * bR * bD, np.uint64)
    z, y, x = np.nonzero(solid)
    blk = (x // 4) + (y // 4) * bC + (z // 4) * bC * bR
    bit = (x % 4) + 4 * (y % 4) + 16 * (z % 4)
    np.bitwise_or.at(out, blk, np.left_shift(np.uint64(1), bit.astype(np.uint64)))
    return out


def numpy_unpack(blocks, W, H, D):
    bC, bR = (W + 3) // 4, (H + 3) // 4
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    blk = (x // 4) + (y // 4) * bC + (z // 4) * bC * bR
    bit = ((x % 4) + 4 * (y % 4) + 16 * (z % 4)).astype(np.uint64)
    return ((blocks[blk] >> bit) & np.uint64(1)).astype(np.uint8)


@pytest.mark.parametrize("dims", [(5, 7, 9), (4, 4, 4), (13, 1, 6), (17, 9, 11)])
def test_bitblock_round_trip(dims):
    W, H, D = dims
    rng = np.random.default_rng(W * 100 + H)
    solid = (rng.random((D, H, W)) < 0.4).astype(np.uint8)
    b = O.bitblock_pack(solid)
    assert np.array_equal(b, numpy_pack(solid))
    assert np.array_equal(numpy_unpack(b, W, H, D), solid)
    # padding bits stay zero
    full = O.bitblock_pack(np.ones((D, H, W), np.uint8))
    assert int(np.sum([bin(int(v)).count("1") for v in full])) == W * H * D


def test_reference_png_stacks_pack():
    data = np.load(os.path.join(HERE, "golden", "images.npz"))
    for key in ("tiny_3D", "sphere_r2_5", "antisphere_r2_5"):
        solid = data[f"{key}_solid"]
        assert np.array_equal(O.bitblock_pack(solid), data[f"{key}_blocks"]), key


# ------------------------------------------------------------------- walk

def py_walk(solid, x, y, z, s, steps, bc, delta):
    """Independent pure-Python restatement of one walker (Appendix A of the
    step spec: draw, noflux flip, move, fold, wall test)."""
    D, H, W = solid.shape
    dims = (W, H, D)
    e, h = 1.0, 0
    moves = {1: (0, -1, 0), 2: (-1, 0, 0), 3: (0, 1, 0), 4: (1, 0, 0), 5: (0, 0, 1), 6: (0, 0, -1)}
    opposite = {1: 3, 3: 1, 2: 4, 4: 2, 5: 6, 6: 5}
    for _ in range(steps):
        s = py_xorshift64(s)
        d = s % 6 + 1
        p = [x, y, z]
        if bc == "noflux":
            ax = [i for i, v in enumerate(moves[d]) if v][0]
            if (moves[d][ax] < 0 and p[ax] == 0) or (moves[d][ax] > 0 and p[ax] == dims[ax] - 1):
                d = opposite[d]
        n = [p[i] + moves[d][i] for i in range(3)]
        if bc == "periodic":
            img = [n[i] % dims[i] for i in range(3)]
        elif bc == "mirror":
            img = [_reflect(n[i], dims[i]) for i in range(3)]
        else:
            img = n
        if solid[img[2], img[1], img[0]]:
            h += 1
            e *= delta
        else:
            x, y, z = n
    return x, y, z, s, e, h


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
def test_walk_matches_pure_python(bc):
    rng = np.random.default_rng(3)
    D, H, W = 6, 5, 7
    solid = (rng.random((D, H, W)) < 0.3).astype(np.uint8)
    pz, py, px = np.nonzero(solid == 0)
    idx = rng.choice(len(px), 12)
    x, y, z = px[idx].astype(np.int32), py[idx].astype(np.int32), pz[idx].astype(np.int32)
    seeds = rng.integers(1, 2**63, 12, dtype=np.int64).astype(np.uint64)
    out = O.walk(O.bitblock_pack(solid), (W, H, D), x, y, z, seeds, echoes=3, steps_per_echo=41, bc=bc,
                 penalty=0.9)
    for i in range(12):
        ref = py_walk(solid, int(x[i]), int(y[i]), int(z[i]), int(seeds[i]), 123, bc, 0.9)
        got = (out["x"][i], out["y"][i], out["z"][i], int(out["seed"][i]), out["energy"][i], out["hits"][i])
        assert got[:4] == ref[:4], (bc, i)
        assert got[4] == ref[4] and got[5] == ref[5]


def _pore_walkers(solid, n, seed):
    rng = np.random.default_rng(seed)
    pz, py, px = np.nonzero(solid == 0)
    idx = rng.choice(len(px), n)
    seeds = rng.integers(1, 2**63, n, dtype=np.int64).astype(np.uint64)
    return px[idx].astype(np.int32), py[idx].astype(np.int32), pz[idx].astype(np.int32), seeds


def test_all_pore_image_keeps_energy_and_msd():
    """No solid voxel: no hits, E == 1, echo sums == N exactly, and the mean
    squared displacement after n unit steps is n (lattice walk)."""
    solid = np.zeros((16, 16, 16), np.uint8)
    x, y, z, seeds = _pore_walkers(solid, 20000, 1)
    out = O.walk(O.bitblock_pack(solid), (16, 16, 16), x, y, z, seeds, echoes=20, steps_per_echo=5,
                 bc="periodic", penalty=0.5)
    assert np.all(out["hits"] == 0) and np.all(out["energy"] == 1.0)
    assert np.all(out["echo_sums"] == 20000.0)
    msd = np.mean((out["x"] - x) ** 2.0 + (out["y"] - y) ** 2.0 + (out["z"] - z) ** 2.0)
    assert abs(msd - 100.0) < 3.0


def test_energy_is_sequential_product_and_map_hits_agree():
    rng = np.random.default_rng(5)
    solid = (rng.random((12, 12, 12)) < 0.35).astype(np.uint8)
    x, y, z, seeds = _pore_walkers(solid, 300, 6)
    b = O.bitblock_pack(solid)
    d = 0.993
    c = O.walk(b, (12, 12, 12), x, y, z, seeds, echoes=10, steps_per_echo=7, bc="mirror", penalty=d)
    m = O.walk(b, (12, 12, 12), x, y, z, seeds, echoes=1, steps_per_echo=70, bc="mirror", echo_sums=False)
    assert np.array_equal(c["hits"], m["hits"])
    table = [1.0]
    for _ in range(int(m["hits"].max())):
        table.append(table[-1] * d)
    assert np.array_equal(c["energy"], np.array(table)[m["hits"]])


def test_noflux_positions_stay_inside():
    rng = np.random.default_rng(8)
    solid = (rng.random((6, 7, 8)) < 0.2).astype(np.uint8)
    x, y, z, seeds = _pore_walkers(solid, 500, 9)
    out = O.walk(O.bitblock_pack(solid), (8, 7, 6), x, y, z, seeds, echoes=50, steps_per_echo=10, bc="noflux")
    assert out["x"].min() >= 0 and out["x"].max() < 8
    assert out["y"].min() >= 0 and out["y"].max() < 7
    assert out["z"].min() >= 0 and out["z"].max() < 6


def test_echo_sums_are_exact_sums():
    rng = np.random.default_rng(11)
    solid = (rng.random((10, 10, 10)) < 0.3).astype(np.uint8)
    x, y, z, seeds = _pore_walkers(solid, 400, 12)
    b = O.bitblock_pack(solid)
    pen = 1 - rng.uniform(0.001, 0.05, 400)
    full = O.walk(b, (10, 10, 10), x, y, z, seeds, echoes=6, steps_per_echo=3, bc="mirror", penalty=pen)
    # echo e sum == math.fsum of per-walker energies after e echoes
    import math
    for e in range(7):
        part = O.walk(b, (10, 10, 10), x, y, z, seeds, echoes=e, steps_per_echo=3, bc="mirror", penalty=pen)
        assert full["echo_sums"][e] == math.fsum(part["energy"])


def test_label_walk_reduces_to_uniform():
    """All solid voxels labelled alike equals the per-walker-delta walk."""
    rng = np.random.default_rng(13)
    solid = (rng.random((9, 9, 9)) < 0.3).astype(np.uint8)
    x, y, z, seeds = _pore_walkers(solid, 200, 14)
    b = O.bitblock_pack(solid)
    lab = np.zeros_like(solid)
    a = O.walk(b, (9, 9, 9), x, y, z, seeds, echoes=5, steps_per_echo=5, bc="mirror", penalty=0.97)
    l = O.walk(b, (9, 9, 9), x, y, z, seeds, echoes=5, steps_per_echo=5, bc="mirror", labels=lab,
               label_delta=[0.97], label_hits=True)
    assert np.array_equal(a["energy"], l["energy"])
    assert np.array_equal(l["label_hits"][:, 0], a["hits"])


# ------------------------------------------------------------------ PFGSE

def test_pfgse_free_diffusion_characteristic_function():
    """All-pore image: a step moves +-1 along x with probability 1/6 each, so
    E[cos(k eps dx)] after n steps is exactly (2/3 + cos(k eps)/3)^n."""
    solid = np.zeros((8, 8, 8), np.uint8)
    n = 40000
    x, y, z, seeds = _pore_walkers(solid, n, 21)
    kx = np.linspace(0.0, 1.2, 7)
    zeros = np.zeros(7)
    steps = np.array([5, 15, 30], np.int64)
    out = O.pfgse(O.bitblock_pack(solid), (8, 8, 8), x, y, z, seeds, steps_inc=steps, kx=kx, ky=zeros, kz=zeros,
                  voxel=1.0, bc="periodic")
    M = out["M"][0]
    cum = np.cumsum(steps)
    for i, nst in enumerate(cum):
        expect = (2.0 / 3.0 + np.cos(kx) / 3.0) ** nst
        assert np.all(np.abs(M[i] / n - expect) < 4.0 / np.sqrt(n)), (i, M[i] / n, expect)
        assert M[i, 0] == n


def test_pfgse_k0_column_is_energy_sum_with_pow():
    rng = np.random.default_rng(22)
    solid = (rng.random((10, 10, 10)) < 0.3).astype(np.uint8)
    x, y, z, seeds = _pore_walkers(solid, 300, 23)
    b = O.bitblock_pack(solid)
    steps = np.array([3, 7], np.int64)
    out = O.pfgse(b, (10, 10, 10), x, y, z, seeds, steps_inc=steps, kx=[0.0, 0.3], ky=[0, 0], kz=[0, 0], voxel=1.0,
                  penalty=0.98)
    m1 = O.walk(b, (10, 10, 10), x, y, z, seeds, echoes=1, steps_per_echo=3, bc="mirror", echo_sums=False)
    m2 = O.walk(b, (10, 10, 10), m1["x"], m1["y"], m1["z"], m1["seed"], echoes=1, steps_per_echo=7, bc="mirror",
                echo_sums=False)
    e = np.power(0.98, m1["hits"].astype(float)) * np.power(0.98, m2["hits"].astype(float))
    import math
    assert out["M"][0, 1, 0] == math.fsum(e)
    assert np.array_equal(out["x"], m2["x"]) and np.array_equal(out["hits"], m1["hits"] + m2["hits"])


# ------------------------------------------------------------- seeding

def test_mt19937_standard_value_and_numpy_stream():
    r = O.MT19937(5489)
    for _ in range(9999):
        r.next()
    assert r.next() == 4123659995  # required by the C++ standard [rand.predef]
    rs = np.random.RandomState(20240607)
    m = O.MT19937(20240607)
    ref = rs.randint(0, 2**32, size=2000, dtype=np.uint64)
    assert [m.next() for _ in range(2000)] == [int(v) for v in ref]


@pytest.fixture(scope="module")
def lemire_bin(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("native") / "lemire_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "native", "lemire_check.cpp")],
                   check=True)
    return exe


@pytest.mark.parametrize("seed,cap", [(1, 10), (424242, 1_000_003), (7, 12345), (99, 2**31 + 5)])
def test_seeds_and_pore_draws_match_libstdcxx(lemire_bin, seed, cap):
    out = subprocess.run([lemire_bin, str(seed), "50", str(cap), "500"], check=True, capture_output=True,
                         text=True).stdout.split()
    vals = [int(v) for v in out]
    m = O.MT19937(seed)
    for _ in range(4096):
        m.next()
    seeds = [(m.next() << 32) | m.next() for _ in range(50)]
    assert seeds == vals[:50]
    draws = [m.lemire(cap + 1) for _ in range(500)]
    assert draws == vals[50:]


def test_seed_and_place_uses_pore_list():
    rng = np.random.default_rng(31)
    solid = (rng.random((6, 6, 6)) < 0.5).astype(np.uint8)
    pores = O.pore_list(solid)
    seeds, x, y, z = O.seed_and_place(5, 1000, pores, solid.size)
    assert np.all(solid[z, y, x] == 0)
    assert len(set(seeds.tolist())) == 1000


# ---------------------------------------------------------- internal field

def py_walk_field(solid, field, x, y, z, s, echoes, spe, bc, delta, gt):
    """Independent pure-Python internal-field walker: per echo two phase
    inversions of spe // 2 steps, phase += gt * field(image voxel) before and
    after each step, echo value p_alive * cos(phase)."""
    D, H, W = solid.shape
    dims = (W, H, D)

    def img(p):
        if bc == "periodic":
            return [p[i] % dims[i] for i in range(3)]
        if bc == "mirror":
            return [_reflect(p[i], dims[i]) for i in range(3)]
        return list(p)

    pa, ph, vals = 1.0, 0.0, [1.0]
    for _ in range(echoes):
        for _inv in range(2):
            ph = -ph
            for _ in range(spe // 2):
                v = img((x, y, z))
                ph = ph + gt * field[v[2], v[1], v[0]]
                x, y, z, s, e, h = py_walk(solid, x, y, z, s, 1, bc, delta)
                pa *= e
                v = img((x, y, z))
                ph = ph + gt * field[v[2], v[1], v[0]]
        vals.append(pa * math.cos(ph))
    return x, y, z, s, pa, ph, vals


@pytest.mark.parametrize("bc", ["noflux", "periodic", "mirror"])
@pytest.mark.parametrize("spe", [2, 5])
def test_field_walk_matches_pure_python(bc, spe):
    """Field values are small integers and gt a power of two, so every
    product and partial sum is exact and fused or separate arithmetic agree:
    positions, p_alive, phases bit-exact; echo values to libm rounding."""
    rng = np.random.default_rng(11 + spe)
    D, H, W = 5, 6, 7
    solid = (rng.random((D, H, W)) < 0.3).astype(np.uint8)
    field = rng.integers(-40, 41, (D, H, W)).astype(np.float64)
    pz, py, px = np.nonzero(solid == 0)
    idx = rng.choice(len(px), 9)
    x, y, z = px[idx].astype(np.int32), py[idx].astype(np.int32), pz[idx].astype(np.int32)
    seeds = rng.integers(1, 2**63, 9, dtype=np.int64).astype(np.uint64)
    gt = 2.0 ** -7
    out = O.walk_field(O.bitblock_pack(solid), (W, H, D), x, y, z, seeds, echoes=4, steps_per_echo=spe,
                       field=field, gammatau=gt, bc=bc, penalty=0.8)
    tot = np.zeros(5)
    for i in range(9):
        rx, ry, rz, rs, pa, ph, vals = py_walk_field(solid, field, int(x[i]), int(y[i]), int(z[i]), int(seeds[i]),
                                                     4, spe, bc, 0.8, gt)
        assert (out["x"][i], out["y"][i], out["z"][i], int(out["seed"][i])) == (rx, ry, rz, rs)
        assert out["p_alive"][i] == pa and out["phase"][i] == ph
        tot += vals
    assert np.allclose(out["echo_sums"], tot, rtol=1e-14, atol=1e-14)


def test_field_walk_zero_field_is_the_plain_walk():
    """No field: phase 0, echo value = p_alive = the plain walk's energy after
    2 * (spe // 2) steps per echo."""
    solid, _ = R.sphere_pack(20, 18, 16, 3, 5, 0.35, 4)
    x, y, z, seeds = _pore_walkers(solid, 300, 2)
    f = O.walk_field(O.bitblock_pack(solid), (20, 18, 16), x, y, z, seeds, echoes=15, steps_per_echo=7,
                     field=np.zeros((16, 18, 20)), gammatau=3.0, bc="mirror", penalty=0.97)
    p = O.walk(O.bitblock_pack(solid), (20, 18, 16), x, y, z, seeds, echoes=15, steps_per_echo=6, bc="mirror",
               penalty=0.97)
    for k in ("x", "y", "z", "seed", "hits"):
        assert np.array_equal(f[k], p[k]), k
    assert np.array_equal(f["p_alive"], p["energy"]) and not np.any(f["phase"])
    assert np.array_equal(f["echo_sums"], p["echo_sums"])


def test_field_walk_constant_field_refocuses():
    """A uniform field dephases during the first half-echo and exactly
    refocuses during the second (CPMG refocusing): phase 0 at every echo."""
    solid, _ = R.sphere_pack(16, 16, 16, 3, 4, 0.4, 6)
    x, y, z, seeds = _pore_walkers(solid, 200, 3)
    f = O.walk_field(O.bitblock_pack(solid), (16, 16, 16), x, y, z, seeds, echoes=9, steps_per_echo=4,
                     field=np.full((16, 16, 16), 3.0), gammatau=0.25, bc="periodic", penalty=0.9)
    assert not np.any(f["phase"])
    plain = O.walk(O.bitblock_pack(solid), (16, 16, 16), x, y, z, seeds, echoes=9, steps_per_echo=4,
                   bc="periodic", penalty=0.9)
    assert np.array_equal(f["echo_sums"], plain["echo_sums"])
