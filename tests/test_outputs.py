"""Output tree and walker set-up (CPU): the reference's result-file formats
written by the library, the walker placement modes and voxel division, the
windowed LHS threshold.

Pins: file texts against the reference's layouts restated here (headers
copied from the reference's writers as data, numbers in C++ ostream form:
"%.17g" for max_digits10, "%g" for the default precision) and a float
round trip; placement against the oracle's one-draw-at-a-time restatement,
itself pinned by brute-force reachable sets on tiny images.
"""
import math

import numpy as np
import pytest

import oracle as O
import randomwalk_nmr_amd as R
from randomwalk_nmr_amd import outputs as OUT


def g17(v):
    return "%.17g" % v


# ---------------------------------------------------------------- walker set-up

def _rock(seed, shape=(16, 18, 20)):
    D, H, W = shape
    solid, _ = R.sphere_pack(W, H, D, 3.0, 3.0, 0.4, seed)
    return solid


@pytest.mark.parametrize("placement,deviation", [("random", 0), ("cubic", 4), ("cubic", 7), ("point", 0)])
@pytest.mark.parametrize("shift", [0, 1, 2])
def test_setup_walkers_matches_oracle(placement, deviation, shift):
    solid = _rock(3)
    if placement == "point":
        D, H, W = solid.shape
        solid[D // 2, H // 2, W // 2] = 0          # the centre must be pore
    o = O.setup_walkers(solid, 987654321, 700, placement, deviation, shift)
    d = R.setup_walkers(solid, 987654321, 700, placement=placement, deviation=deviation, shift=shift)
    seeds, x, y, z, draws = o
    assert np.array_equal(d["seed"], seeds)
    assert np.array_equal(d["x"], x) and np.array_equal(d["y"], y) and np.array_equal(d["z"], z)
    assert d["draws"] == draws
    assert d["occupancy"] == 700 / float((solid == 0).sum())
    if placement == "random":
        s2, x2, y2, z2 = R.seed_and_place(solid, 987654321, 700)
        f = 1 << shift
        assert np.array_equal(s2, seeds) and np.array_equal(x2, x >> shift) and np.array_equal(z2, z >> shift)
        assert np.all((x >= x2 * f) & (x < (x2 + 1) * f))


def test_setup_walkers_rejects_what_the_reference_cannot_place():
    solid = _rock(4)
    D, H, W = solid.shape
    solid[D // 2, H // 2, W // 2] = 1
    assert O.setup_walkers(solid, 5, 10, "point") is None
    with pytest.raises(R.RwnmrError):
        R.setup_walkers(solid, 5, 10, placement="point")
    # deviation 0: the cube [c, c) is empty; deviation past the centre: a
    # negative lower corner (an unsigned wrap in the reference) leaves it empty
    for dev in (0, 20):
        assert O.setup_walkers(solid, 5, 10, "cubic", dev) is None
        with pytest.raises(R.RwnmrError):
            R.setup_walkers(solid, 5, 10, placement="cubic", deviation=dev)


def test_oracle_cubic_and_voxel_division_reach_exactly_their_sets():
    rng = np.random.default_rng(1)
    solid = (rng.random((5, 6, 7)) < 0.3).astype(np.uint8)
    D, H, W = solid.shape
    dev = 2
    c = ((W - 1) // 2, (H - 1) // 2, (D - 1) // 2)
    want = {(x, y, z) for z in range(c[2] - dev, min(c[2] + dev, D)) for y in range(c[1] - dev, min(c[1] + dev, H))
            for x in range(c[0] - dev, min(c[0] + dev, W)) if not solid[z, y, x]}
    _, x, y, z, _ = O.setup_walkers(solid, 77, 4000, "cubic", dev, 0)
    assert set(zip(x.tolist(), y.tolist(), z.tolist())) == want
    # voxel division: every fine site of every start voxel is reached
    _, x, y, z, _ = O.setup_walkers(solid, 78, 6000, "cubic", dev, 1)
    fine = set(zip(x.tolist(), y.tolist(), z.tolist()))
    assert {(a >> 1, b >> 1, e >> 1) for a, b, e in fine} == want
    assert len(fine) == 8 * len(want)
    # the number of generator outputs: 4096 + 2n seeds + one per draw when
    # no draw is rejected (power-of-two ranges never reject)
    seeds, *_ , draws = O.setup_walkers(solid, 79, 50, "point" if not solid[D // 2, H // 2, W // 2] else "cubic",
                                        dev, 2)
    if not solid[D // 2, H // 2, W // 2]:
        assert draws == 4096 + 2 * 50 + 3 * 50


# ---------------------------------------------------------------- writers

def test_cpmg_and_histogram_files(tmp_path):
    t = np.array([0.0, 0.1, 0.30000000000000004, 1e-300, 12345.678])
    a = np.array([1.0, 0.9, 1 / 3, 2.5e-7, -0.0])
    nz = np.array([0.0, 1e-3, -2e-3, 0.5, 1.0])
    p = tmp_path / "cpmg_decay.csv"
    OUT.write_cpmg_decay(p, t, a, nz)
    rows = [f"{g17(t[i])},{g17(a[i] + nz[i])},{g17(nz[i])},{g17(a[i])}" for i in range(len(t))]
    assert p.read_text() == "time,signal,noise,noiseless\n" + "".join(r + "\n" for r in rows)
    back = np.loadtxt(p, delimiter=",", skiprows=1)
    assert np.array_equal(back[:, 0], t) and np.array_equal(back[:, 3], a)
    OUT.write_cpmg_decay(p, t[:2], a[:2])
    assert p.read_text().splitlines()[2] == f"{g17(0.1)},{g17(0.9)},0,{g17(0.9)}"
    q = tmp_path / "cpmg_T2.csv"
    OUT.write_two_columns(q, "NMRT2_bins,NMRT2_amps", t, a)
    assert q.read_text() == "NMRT2_bins,NMRT2_amps\n" + "".join(f"{g17(u)},{g17(v)}\n" for u, v in zip(t, a))
    h = tmp_path / "cpmg_histList.csv"
    bins = np.arange(6.0).reshape(2, 3) / 7
    amps = np.arange(6.0).reshape(2, 3) / 11
    OUT.write_histogram_list(h, bins, amps)
    lines = h.read_text().split("\n")
    assert lines[0] == "Bins0,Amps0,Bins1,Amps1,"
    assert lines[1] == f"{g17(bins[0, 0])},{g17(amps[0, 0])},{g17(bins[1, 0])},{g17(amps[1, 0])},"
    assert len(lines) == 5 and lines[4] == ""


def test_pfgse_files(tmp_path):
    p = tmp_path / "PFGSE_parameters.txt"
    OUT.write_pfgse_parameters(p, D0=2.5, pulse_width=0.1, gamma=42.576 * 2 * math.pi, gmax=(3.0, 4.0, 0.0),
                               gradients=10, times=[0.5, 1.25])
    assert p.read_text() == ("RWNMR-PFGSE Parameters\nD_0: 2.5\nPulse width: 0.10000000000000001\n"
                             f"Giromagnetic Ratio: {g17(42.576 * 2 * math.pi)}\n"
                             f"Gradient direction: {{0.59999999999999998, 0.80000000000000004, 0}}\n"
                             "Gradients: 10\nTimes: [0.5, 1.25]\n")
    g = np.array([[0.0, 0, 0], [1.5, 0, 0.25]])
    k = g * 1e-3
    q = tmp_path / "PFGSE_gradient.csv"
    OUT.write_pfgse_gradient(q, g, k)
    assert q.read_text() == ("Id,Gx,Gy,Gz,Kx,Ky,Kz\n0,0,0,0,0,0,0\n"
                             f"1,1.5,0,0.25,{g17(1.5e-3)},0,{g17(0.25e-3)}\n")
    e = tmp_path / "echoes_0.csv"
    cols = [np.array([0.0, 0.5]), np.array([100.0, 90.0]), np.zeros(2), np.array([0.0, 1.5]),
            np.array([0.0, -0.1]), np.zeros(2), np.array([-0.0, -0.25])]
    OUT.write_pfgse_echoes(e, gradient=cols[0], Mkt=cols[1], noise=cols[2], Mkt_std=cols[3], lhs=cols[4],
                           lhs_std=cols[5], rhs=cols[6])
    assert e.read_text() == ("Idx,Gradient,NMR_signal(mean),NMR_signal(noise),NMR_signal(std),SAT_lhs(mean),"
                             "SAT_lhs(std),SAT_rhs\n0,0,100,0,0,0,0,-0\n"
                             f"1,0.5,90,0,1.5,{g17(-0.1)},0,-0.25\n")
    m = tmp_path / "msd_0.csv"
    OUT.write_pfgse_msd(m, [1, 2, 3], [0.5, 0, 0], [4, 5, 6], [0, 0, 0.25])
    assert m.read_text() == ("msdX(mean),msdX(std),msdY(mean),msdY(std),msdZ(mean),msdZ(std),DmsdX(mean),"
                             "DmsdX(std),DmsdY(mean),DmsdY(std),DmsdZ(mean),DmsdZ(std)\n"
                             "1,0.5,2,0,3,0,4,0,5,0,6,0.25")
    r = tmp_path / "PFGSE_results.csv"
    OUT.write_pfgse_results_header(r)
    OUT.append_pfgse_results(r, 0.5, [2.1, 0.01, 0.0, 2.2, 0.0, 0.7, 0, 0.8, 0, 0.9, 0], 8)
    OUT.append_pfgse_results(r, 1.5, [1.9, 0.02, 0.1, 2.0, 0.1, 0.6, 0, 0.7, 0, 0.8, 0.05], 6)
    lines = r.read_text().splitlines()
    assert lines[0] == ("Time,D_sat,D_sat(error),D_sat(std),D_sat(pts),D_msd,D_msd(std),D_msdX,D_msdX(std),"
                        "D_msdY,D_msdY(std),D_msdZ,D_msdZ(std)")
    assert lines[1] == f"0.5,{g17(2.1)},{g17(0.01)},0,8,{g17(2.2)},0,{g17(0.7)},0,{g17(0.8)},0,{g17(0.9)},0"
    assert lines[2].split(",")[4] == "6" and len(lines) == 3


def test_sim_info_and_bin_image(tmp_path):
    p = tmp_path / "SimInfo.txt"
    OUT.write_sim_info(p, name="rock", data_path="./db/rock", image_path="/img/s_00.png", voxel_resolution=0.625,
                       D0=2.5, n_images=10, occupancy=0.123456789, seed=18446744073709551615, gpu=True, bc="mirror")
    rule = "-" * 54
    assert p.read_text() == (f"{rule}\n>>> NMR SIMULATION 3D PARAMETERS: rock\n{rule}\nData path: ./db/rock\n"
                             "Image path: /img/s_00.png\nImage resolution (um/voxel): 0.625\n"
                             "Diffusion coefficient (um^2/ms): 2.5\nNumber of images: 10\n"
                             "Walkers pore occupancy in simulation: 12.3457%\n"
                             "Initial seed: 18446744073709551615\nGPU usage: ON\nBC: mirror\n" + rule + "\n\n")
    solid = (np.random.default_rng(2).random((5, 6, 9)) < 0.5).astype(np.uint8)
    blocks = R.pack_bitblocks(solid)
    b = tmp_path / "NMR_binImage.txt"
    OUT.write_bin_image(b, blocks, solid.shape)
    lines = b.read_text().splitlines()
    assert lines[0] == "blocks, bRows, bColumns, bDepth, imgRows, imgColumns, imgDepth, "
    assert lines[1] == f"{len(blocks)}, 2, 3, 2, 6, 9, 5"
    assert lines[2] == "" and lines[3] == "blockID, blockData"
    assert [int(l.split(", ")[1]) for l in lines[4:]] == [int(v) for v in blocks]


def test_make_dirs(tmp_path):
    d = tmp_path / "db" / "name" / "/NMR_pfgse".lstrip("/") / "timesamples"
    OUT.make_dirs(d)
    OUT.make_dirs(d)                 # existing: fine
    assert d.is_dir()
    f = tmp_path / "file"
    f.write_text("x")
    with pytest.raises(R.RwnmrError):
        OUT.make_dirs(f / "sub")


# ---------------------------------------------------------------- threshold

def _window_threshold(lhs, value, window):
    """setThresholdFromLHSWindow restated (B/NMR_Simulation/NMR_pfgse.cpp:405-439)."""
    if len(lhs) < window:
        return len(lhs)
    win = list(lhs[:window])
    idx, greater, lv = window, True, math.log(value)
    if sum(win) / len(win) < lv:
        greater = False
    while idx < len(lhs) and greater:
        win[idx % window] = lhs[idx]
        if sum(win) / len(win) < lv:
            greater = False
        else:
            idx += 1
    return idx - 1 if greater else idx


@pytest.mark.parametrize("window", [1, 2, 3, 5])
def test_threshold_lhs_window(window):
    rng = np.random.default_rng(window)
    for _ in range(50):
        lhs = np.sort(np.log(rng.uniform(0.05, 1.0, 12)))[::-1] + rng.normal(0, 0.05, 12)
        for value in (0.1, 0.3, 0.5, 0.9):
            assert OUT.threshold_lhs_window(lhs, value, window) == _window_threshold(list(lhs), value, window)
    # known answers: a straight decay log(0.9^k) crosses log(0.5) at k = 7
    lhs = np.log(0.9 ** np.arange(12))
    assert OUT.threshold_lhs_window(lhs, 0.5, 1) == 7
    assert OUT.threshold_lhs_window(lhs, 1e-9, 1) == 11          # never below: the last point
