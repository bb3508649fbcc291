"""The measured configurations at their full sizes on the GPU, checked
against the CPU oracle on walker subsamples and by size-independent
properties of the whole run.

* 512^3 CPMG, 1M walkers, 10,000 echoes x 12 steps, mirror: ten 1000-echo
  launches with locality re-sorts on the 128 MiB unfolded mirror image.
* 256^3 PFGSE, 1M walkers, 16 diffusion times x 64 k x 2 relaxivities.
* 64 uneven CPMG instances (64^3..256^3, 10k..500k walkers, 4000 echoes) in
  one persistent launch.
* 256^3 labelled CPMG (8 Voronoi labels), 1M walkers, 4000 echoes with
  checkpoints every 500 echoes: forward, backward and finite differences.

Bars: walker positions, seeds, hits and (CPMG) energies bit-exact; echo
sums within 1e-12 relative; PFGSE M within 1e-11 M(0) and pow-based
energies within 1e-13 relative; gradients as the small-size gradient tests.
"""
import math

import numpy as np
import pytest

import oracle as O
import randomwalk_nmr_amd as R
from randomwalk_nmr_amd.fit import CpmgModel

pytestmark = pytest.mark.gpu

ECHO_RTOL = 1e-12
SEED = 20250101


def _dev(torch, a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a).to("cuda:0")


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def test_cpmg_512_1m_walkers_10000_echoes(ctx):
    import torch
    print("[test] cpmg 512^3, 1M walkers, 10000 echoes x 12, mirror", flush=True)
    S, N, E, voxel, D0, rho = 512, 1_000_000, 10_000, 0.5, 2.5, 25.0
    spe = int(round(0.2 / R.time_step(voxel, D0)))
    assert spe == 12
    solid, _ = R.sphere_pack(S, S, S, 6.0, 16.0, 0.20, SEED)
    seeds, x, y, z = R.seed_and_place(solid, SEED, N)
    delta = R.penalty(rho, voxel, D0)
    img = R.Image.from_solid(ctx, solid)
    X, Y, Z, SD = _dev(torch, x), _dev(torch, y), _dev(torch, z), _dev(torch, seeds)
    EN = torch.ones(N, dtype=torch.float64, device="cuda:0")
    HT = torch.zeros(N, dtype=torch.int32, device="cuda:0")
    SUMS = torch.zeros(E + 1, dtype=torch.float64, device="cuda:0")
    R.cpmg_run(ctx, img, X, Y, Z, SD, echoes=E, steps_per_echo=spe, bc="mirror", delta=delta, energy=EN, hits=HT,
               echo_sums=SUMS)
    sums = SUMS.cpu().numpy()
    en = EN.cpu().numpy()
    # whole run: sum of initial energies, non-increasing decay, last echo == sum of final energies
    assert sums[0] == N
    assert np.all(np.diff(sums) <= 0)
    assert abs(sums[-1] - math.fsum(en)) <= ECHO_RTOL * sums[-1]
    # first 10,000 walkers of the full run vs the oracle, bit for bit
    ns = 10_000
    blocks = O.bitblock_pack(solid)
    print("[test] oracle on the first 10000 walkers", flush=True)
    o = O.walk(blocks, (S, S, S), x[:ns], y[:ns], z[:ns], seeds[:ns], echoes=E, steps_per_echo=spe, bc="mirror",
               penalty=delta)
    for k, T in (("x", X), ("y", Y), ("z", Z)):
        assert np.array_equal(T[:ns].cpu().numpy(), o[k]), k
    assert np.array_equal(SD[:ns].cpu().numpy().view(np.uint64), o["seed"])
    assert np.array_equal(HT[:ns].cpu().numpy().view(np.uint32), o["hits"])
    assert _bits_equal(en[:ns], o["energy"])
    # the same subsample walked alone: its echo sums vs the oracle's compensated sums
    g = R.cpmg_run(ctx, img, x[:ns], y[:ns], z[:ns], seeds[:ns], echoes=E, steps_per_echo=spe, bc="mirror",
                   delta=delta)
    assert np.all(np.abs(g["echo_sums"] - o["echo_sums"]) <= ECHO_RTOL * np.abs(o["echo_sums"]))
    # the whole run again in 1024-thread blocks (flag-keyed tables; on this
    # image, whose periods exceed twice the flush interval, the unbounded
    # moves come from address displacements; 1M walkers are more than one
    # round of resident blocks, so the tiles are balanced: 977 walkers in
    # each of two full rounds of blocks), with both chunk caches,
    # position-ordered and (form 6) in the caller's order: the same walkers
    # bit for bit (runs without a hit array count no hits: the headline's
    # kernel variant)
    for form, local, count in ((3, True, True), (3, True, False), (6, True, True), (6, False, False)):
        print(f"[test] cpmg 512^3 full run, 1024 x 1, image form {form}, locality {local}, hits {count}",
              flush=True)
        X2, Y2, Z2, SD2 = _dev(torch, x), _dev(torch, y), _dev(torch, z), _dev(torch, seeds)
        EN2 = torch.ones(N, dtype=torch.float64, device="cuda:0")
        HT2 = torch.zeros(N, dtype=torch.int32, device="cuda:0") if count else None
        SUMS2 = torch.zeros(E + 1, dtype=torch.float64, device="cuda:0")
        ctx.set_launch(1024, 1)
        ctx.set_unfolded(form)
        ctx.set_locality(local)
        try:
            R.cpmg_run(ctx, img, X2, Y2, Z2, SD2, echoes=E, steps_per_echo=spe, bc="mirror", delta=delta,
                       energy=EN2, hits=HT2, echo_sums=SUMS2)
        finally:
            ctx.set_launch(256, 1)
            ctx.set_unfolded(3)
            ctx.set_locality(True)
        for A, B in ((X, X2), (Y, Y2), (Z, Z2), (SD, SD2)) + (((HT, HT2),) if count else ()):
            assert torch.equal(A, B)
        assert _bits_equal(en, EN2.cpu().numpy())
        assert np.all(np.abs(SUMS2.cpu().numpy() - sums) <= ECHO_RTOL * np.abs(sums))
    img.close()


def test_pfgse_256_1m_walkers_16_delta_64_k_2_rho(ctx):
    import torch
    print("[test] pfgse 256^3, 1M walkers, 16 Delta x 64 k x 2 rho", flush=True)
    S, N, ND, NK, voxel, D0 = 256, 1_000_000, 16, 64, 1.0, 2.5
    dt = R.time_step(voxel, D0)
    deltas = np.logspace(np.log10(0.5), np.log10(40.0), ND)
    steps = np.array([int(round(d / dt)) for d in deltas])
    steps_inc = np.diff(np.concatenate([[0], steps])).astype(np.uint32)
    K = np.linspace(0.0, 1.0, NK)[None, :] * (3.0 / np.sqrt(D0 * deltas))[:, None]
    zero = np.zeros_like(K)
    solid, _ = R.sphere_pack(S, S, S, 10.0, 10.0, 0.30, SEED)
    seeds, x, y, z = R.seed_and_place(solid, SEED, N)
    # set 0: rho 10 um/s (its energies are returned), set 1: rho 0
    set_delta = np.array([R.penalty(10.0, voxel, D0), 1.0])
    img = R.Image.from_solid(ctx, solid)
    X0, Y0, Z0 = _dev(torch, x), _dev(torch, y), _dev(torch, z)
    X, Y, Z, SD = X0.clone(), Y0.clone(), Z0.clone(), _dev(torch, seeds)
    EN = torch.ones(N, dtype=torch.float64, device="cuda:0")
    M = torch.zeros((2, ND, NK), dtype=torch.float64, device="cuda:0")
    msd = torch.zeros((ND, 3), dtype=torch.float64, device="cuda:0")
    R.pfgse_run(ctx, img, X, Y, Z, SD, steps_inc=_dev(torch, steps_inc.view(np.int32)), kx=_dev(torch, K),
                ky=_dev(torch, zero), kz=_dev(torch, zero), voxel=voxel, bc="mirror", x0=X0, y0=Y0, z0=Z0,
                set_delta=_dev(torch, set_delta), energy=EN, M=M, msd=msd)
    Mn, msdn, en = M.cpu().numpy(), msd.cpu().numpy(), EN.cpu().numpy()
    xf, yf, zf = X.cpu().numpy(), Y.cpu().numpy(), Z.cpu().numpy()
    # whole run: rho = 0 keeps every energy at 1; k = 0 of the last time == sum of the pow energies;
    # the last time's squared displacements are exact integer sums of the final positions
    assert np.all(Mn[1, :, 0] == N)
    assert abs(Mn[0, -1, 0] - math.fsum(en)) <= 1e-12 * Mn[0, -1, 0]
    for c, (f, i0) in enumerate(((xf, x), (yf, y), (zf, z))):
        d = f.astype(np.int64) - i0.astype(np.int64)
        assert msdn[-1, c] == float(np.sum(d * d))
    # first 2,000 walkers vs the oracle
    ns = 2000
    blocks = O.bitblock_pack(solid)
    o = O.pfgse(blocks, (S, S, S), x[:ns], y[:ns], z[:ns], seeds[:ns], steps_inc=steps_inc, kx=K, ky=zero, kz=zero,
                voxel=voxel, bc="mirror", set_delta=set_delta)
    assert np.array_equal(xf[:ns], o["x"]) and np.array_equal(yf[:ns], o["y"]) and np.array_equal(zf[:ns], o["z"])
    assert np.array_equal(SD[:ns].cpu().numpy().view(np.uint64), o["seed"])
    assert np.all(np.abs(en[:ns] - o["energy"]) <= 1e-13 * o["energy"])
    g = R.pfgse_run(ctx, img, x[:ns], y[:ns], z[:ns], seeds[:ns], steps_inc=steps_inc, kx=K, ky=zero, kz=zero,
                    voxel=voxel, bc="mirror", set_delta=set_delta)
    err = np.max(np.abs(g["M"] - o["M"]) / np.abs(o["M"][..., :1]))
    assert err <= 1e-11, err
    img.close()


def test_batched_64_uneven_instances(ctx):
    import torch
    print("[test] batched cpmg: 64 uneven instances, 4000 echoes", flush=True)
    rng = np.random.default_rng(SEED)
    voxel, D0, E = 1.0, 2.5, 4000
    spe = int(round(0.2 / R.time_step(voxel, D0)))
    inst, start, host = [], [], []
    for i in range(64):
        side = int(rng.choice([64, 96, 128, 192, 256]))
        phi = float(rng.uniform(0.10, 0.40))
        r = float(rng.uniform(4.0, 12.0))
        n = int(rng.integers(10_000, 500_001))
        rho = float(rng.uniform(5.0, 50.0))
        solid, _ = R.sphere_pack(side, side, side, r, r, phi, SEED + i)
        seeds, x, y, z = R.seed_and_place(solid, SEED + 1000 + i, n)
        t = [_dev(torch, a) for a in (x, y, z, seeds)]
        start.append(t)
        host.append((solid, x, y, z, seeds))
        inst.append(dict(img=R.Image.from_solid(ctx, solid), x=t[0].clone(), y=t[1].clone(), z=t[2].clone(),
                         seed=t[3].clone(), energy=torch.ones(n, dtype=torch.float64, device="cuda:0"),
                         hits=torch.zeros(n, dtype=torch.int32, device="cuda:0"),
                         echo_sums=torch.zeros(E + 1, dtype=torch.float64, device="cuda:0"), echoes=E,
                         steps_per_echo=spe, bc="mirror", delta=R.penalty(rho, voxel, D0)))

    def run(sequential):
        for q, t in zip(inst, start):
            q["x"].copy_(t[0]), q["y"].copy_(t[1]), q["z"].copy_(t[2]), q["seed"].copy_(t[3])
            q["energy"].fill_(1.0), q["hits"].zero_()
        R.cpmg_batch_run(ctx, inst, sequential=sequential)
        return [{k: q[k].cpu().numpy().copy() for k in ("x", "y", "z", "seed", "energy", "hits", "echo_sums")}
                for q in inst]

    batched = run(False)
    print("[test] batched launch done; sequential launches", flush=True)
    seq = run(True)
    for i, (b, s) in enumerate(zip(batched, seq)):
        for k in ("x", "y", "z", "seed", "hits"):
            assert np.array_equal(b[k], s[k]), (i, k)
        assert _bits_equal(b["energy"], s["energy"]), i
        assert np.all(np.abs(b["echo_sums"] - s["echo_sums"]) <= ECHO_RTOL * s["echo_sums"]), i
    print("[test] oracle on 2000 walkers of every instance", flush=True)
    ns = 2000
    for i, ((solid, x, y, z, seeds), b, q) in enumerate(zip(host, batched, inst)):
        side = solid.shape[0]
        o = O.walk(O.bitblock_pack(solid), (side, side, side), x[:ns], y[:ns], z[:ns], seeds[:ns], echoes=E,
                   steps_per_echo=spe, bc="mirror", penalty=q["delta"], echo_sums=False)
        for k in ("x", "y", "z"):
            assert np.array_equal(b[k][:ns], o[k]), (i, k)
        assert np.array_equal(b["seed"][:ns].view(np.uint64), o["seed"]), i
        assert np.array_equal(b["hits"][:ns].view(np.uint32), o["hits"]), i
        assert _bits_equal(b["energy"][:ns], o["energy"]), i
        # whole instance: last echo == sum of its final energies
        assert abs(b["echo_sums"][-1] - math.fsum(b["energy"])) <= ECHO_RTOL * b["echo_sums"][-1], i
    for q in inst:
        q["img"].close()


def test_labelled_fit_256_1m_walkers_4000_echoes(ctx):
    import torch
    print("[test] labelled cpmg 256^3, 1M walkers, 4000 echoes, checkpoint every 500", flush=True)
    S, N, E, voxel, D0 = 256, 1_000_000, 4000, 1.0, 2.5
    spe = int(round(0.2 / R.time_step(voxel, D0)))
    solid, _ = R.sphere_pack(S, S, S, 6.0, 12.0, 0.25, SEED)
    labels = R.voronoi_labels(solid, 64, 8, SEED + 1)
    seeds, x, y, z = R.seed_and_place(solid, SEED + 2, N)
    img = R.Image.from_solid(ctx, solid, labels=labels, n_labels=8)
    m = CpmgModel(ctx, img, x, y, z, seeds, echoes=E, steps_per_echo=spe, voxel=voxel, D0=D0, ckpt_every=500)
    xo, yo, zo, so = (np.asarray(a)[m.order] for a in (x, y, z, seeds))
    rho = np.random.default_rng(SEED).uniform(5.0, 40.0, 8)
    delta = m.delta_of_rho(rho)
    Mf = m.forward_sums(delta).cpu().numpy()
    ef = m.e.cpu().numpy()
    # checkpointed forward == the plain labelled run, bit for bit
    X, Y, Z, SD = (_dev(torch, a) for a in (xo, yo, zo, so))
    EN = torch.ones(N, dtype=torch.float64, device="cuda:0")
    LH = torch.zeros((8, N), dtype=torch.int32, device="cuda:0")
    SUMS = torch.zeros(E + 1, dtype=torch.float64, device="cuda:0")
    R.cpmg_run(ctx, img, X, Y, Z, SD, echoes=E, steps_per_echo=spe, bc="mirror", energy=EN, echo_sums=SUMS,
               label_delta=_dev(torch, delta), label_hits=LH)
    assert np.array_equal(Mf, SUMS.cpu().numpy())
    assert _bits_equal(ef, EN.cpu().numpy())
    lh = LH.cpu().numpy().view(np.uint32)
    # first 2,000 walkers (model order) vs the oracle
    ns = 2000
    o = O.walk(O.bitblock_pack(solid), (S, S, S), xo[:ns], yo[:ns], zo[:ns], so[:ns], echoes=E, steps_per_echo=spe,
               bc="mirror", labels=labels, label_delta=delta, label_hits=True, echo_sums=False)
    assert np.array_equal(X[:ns].cpu().numpy(), o["x"]) and np.array_equal(SD[:ns].cpu().numpy().view(np.uint64),
                                                                           o["seed"])
    assert _bits_equal(ef[:ns], o["energy"])
    assert np.array_equal(lh[:, :ns].T, o["label_hits"])
    # backward for L = M(E): dL/d delta_l = sum_w E_w h_wl / delta_l over all 1M walkers
    g = torch.zeros(E + 1, dtype=torch.float64)
    g[-1] = 1.0
    got = m.backward_delta(g)
    ref = np.array([math.fsum(ef * lh[l]) / delta[l] for l in range(8)])
    assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref).max()), (got, ref)
    # the Jacobian forward at full size: the same sums bit for bit, and its
    # last row dM(E)/d log delta_l = delta_l dM(E)/d delta_l = the
    # replay's gradient for L = M(E) times delta_l, = sum_w E_w h_wl
    Mj, Jj = (t.cpu().numpy() for t in m.forward_jacobian(delta))
    assert np.array_equal(Mj, Mf)
    assert np.all(np.abs(Jj[:, -1] - ref * delta) <= 1e-12 * np.abs(ref * delta).max()), (Jj[:, -1], ref * delta)
    assert np.all(Jj[:, 0] == 0.0)
    # central differences with common random numbers, h = 1e-3 relative on rho
    c = np.random.default_rng(1).standard_normal(E + 1)
    m.forward_sums(delta)
    grad_rho = m.backward_delta(torch.tensor(c)) * m.ddelta_drho
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
    img.close()
