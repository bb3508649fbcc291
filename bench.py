#!/usr/bin/env python3
"""Benchmark: CPMG walker-steps/s on one MI355X.

Workload (the headline configuration): 512^3 seeded sphere pack (radii
uniform 6-16 voxels, porosity 0.20), voxel 0.5 um, D0 2.5 um^2/ms,
rho 25 um/s, 1,000,000 walkers, 10,000 echoes at TE 0.2 ms (12 steps per
echo), mirror boundaries, FP64 energies and echo sums.  One bench step = one
full CPMG run (walkers reset to their start state, then 1.2e11 walker-steps)
with all inputs resident in HBM.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E bytes/s (published)
BYTES_PER_STEP = 8         # one uint64 bitBlock word read per walker-step
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "walk_kernel_cpmg512_traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--walkers", type=int, default=1_000_000)
    ap.add_argument("--echoes", type=int, default=10_000)
    ap.add_argument("--seed", type=int, default=20250101)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--launch", default="", help="threads,k (skips the autotuner)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run one device-to-device copy of this many bytes (HBM counter scale)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(args.copy_calibration, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_calibration}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build runs on one GPU "
              "(walkers would shard across GPUs with one all-reduce of the echo sums)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import randomwalk_nmr_amd as R

    voxel, D0, rho, te = 0.5, 2.5, 25.0, 0.2
    dt = R.time_step(voxel, D0)
    spe = int(round(te / dt))                       # 12 steps per echo
    E, N, S = args.echoes, args.walkers, args.size
    t = time.time()
    solid, phi = R.sphere_pack(S, S, S, 6.0, 16.0, 0.20, args.seed)
    log(f"sphere pack {S}^3 porosity {phi:.4f} in {time.time() - t:.1f}s")
    t = time.time()
    seeds, x0, y0, z0 = R.seed_and_place(solid, args.seed, N)
    log(f"placed {N} walkers in {time.time() - t:.1f}s")
    delta = R.penalty(rho, voxel, D0)

    ctx = R.Context(0)
    img = R.Image.from_solid(ctx, solid)
    dev = torch.device("cuda:0")
    X0, Y0, Z0 = (torch.from_numpy(a).to(dev) for a in (x0, y0, z0))
    SEED0 = torch.from_numpy(seeds.view(np.int64)).to(dev)
    X, Y, Z, SEED = X0.clone(), Y0.clone(), Z0.clone(), SEED0.clone()
    EN = torch.ones(N, dtype=torch.float64, device=dev)
    SUMS = torch.zeros(E + 1, dtype=torch.float64, device=dev)

    if args.launch:
        tpb, k = (int(v) for v in args.launch.split(","))
        ctx.set_launch(tpb, k)
    elif not args.no_autotune:
        t = time.time()
        tpb, k, ms = ctx.autotune(img, X0, Y0, Z0, SEED0, steps_per_echo=spe, bc="mirror", trial_echoes=100,
                                  delta=delta)
        log(f"autotune: threads {tpb} x walkers/lane {k} ({ms:.2f} ms per 100 echoes) in {time.time() - t:.1f}s")
    tpb, k = ctx.get_launch()

    def one_step():
        X.copy_(X0)
        Y.copy_(Y0)
        Z.copy_(Z0)
        SEED.copy_(SEED0)
        EN.fill_(1.0)
        R.cpmg_run(ctx, img, X, Y, Z, SEED, echoes=E, steps_per_echo=spe, bc="mirror", delta=delta, energy=EN,
                   echo_sums=SUMS, sync=True)

    for i in range(args.warmup):
        one_step()
        log(f"warmup {i + 1}/{args.warmup} done")
    ctx.set_timing(True)
    ctx.get_timing(reset=True)
    torch.cuda.synchronize()
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        one_step()
        log(f"step {i + 1}/{args.steps} done")
    torch.cuda.synchronize()
    ctx.synchronize()
    t1 = time.perf_counter()
    kern_ms, launches = ctx.get_timing(reset=True)
    ctx.set_timing(False)

    walker_steps = float(N) * spe * E
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = walker_steps / (ms_per_step * 1e-3)
    avg_kernel_s = kern_ms * 1e-3 / max(launches, 1)
    steps_per_launch = walker_steps * args.steps / max(launches, 1)
    achieved = steps_per_launch * BYTES_PER_STEP / avg_kernel_s
    # HBM bytes per launch come from the committed FETCH_SIZE / WRITE_SIZE profile of this
    # same command; they are only reported when that profile was taken at this launch shape.
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            prof = json.load(f)
        if prof.get("launch") == f"{tpb},{k}":
            traffic = prof.get("hbm_bytes_per_launch")

    sums = SUMS.cpu().numpy()
    decay_ok = bool(np.all(np.isfinite(sums)) and sums[0] == N and np.all(np.diff(sums) <= 0))

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        import oracle as O
        blocks = R.pack_bitblocks(solid)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        ns = min(10_000, N)
        # all host threads: first 10,000 walkers over the full 10,000 echoes
        secs, es_cpu, xo, yo, zo = O.cpmg_reference_loop(blocks, (S, S, S), x0[:ns], y0[:ns], z0[:ns], seeds[:ns],
                                                         delta, echoes=E, steps_per_echo=spe, bc="mirror",
                                                         use_digit_sum=True, threads=threads)
        cpu_rate = ns * spe * E / secs
        log(f"cpu baseline: {threads} threads {cpu_rate:.3e} walker-steps/s ({secs:.1f}s)")
        gx, gy, gz = (T[:ns].cpu().numpy() for T in (X, Y, Z))
        parity = bool(np.array_equal(gx, xo) and np.array_equal(gy, yo) and np.array_equal(gz, zo))
        # one thread: first 1,000 walkers over 2,000 echoes
        n1, e1 = min(1000, N), min(2000, E)
        secs1, *_ = O.cpmg_reference_loop(blocks, (S, S, S), x0[:n1], y0[:n1], z0[:n1], seeds[:n1], delta,
                                          echoes=e1, steps_per_echo=spe, bc="mirror", use_digit_sum=True,
                                          threads=1)
        rate1 = n1 * spe * e1 / secs1
        log(f"cpu baseline: 1 thread {rate1:.3e} walker-steps/s ({secs1:.1f}s)")
        cpu = {"value": cpu_rate, "unit": "walker-steps/s", "cores": threads, "kind": "port",
               "sample": f"first {ns} walkers x {E} echoes x {spe} steps (reference loop order, digit-sum mod6, "
                         f"AoS walkers), OpenMP static over walkers",
               "single_thread_value": rate1,
               "single_thread_sample": f"first {n1} walkers x {e1} echoes x {spe} steps, 1 thread",
               "gpu_over_cpu": value / cpu_rate}

    out = {
        "metric": "walker-steps/s, CPMG 512^3 1M walkers",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"CPMG {S}^3 sphere pack (r 6-16, porosity {phi:.3f}), {N} walkers, {E} echoes x "
                               f"{spe} steps, mirror BC, voxel {voxel} um, rho {rho} um/s",
                   "walker_steps_per_step": walker_steps, "launch": {"threads_per_block": tpb,
                                                                    "walkers_per_lane": k}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "walk_kernel<mirror, per-walker delta>",
                     "avg_kernel_ms": avg_kernel_s * 1e3, "launches": launches,
                     "algorithmic_bytes_per_launch": steps_per_launch * BYTES_PER_STEP},
        "cpu_baseline": cpu,
        "checks": {"decay_monotone": decay_ok, "subsample_positions_bit_exact": parity},
    }
    out["roofline"]["achieved"] = achieved / 1e9
    out["roofline"]["peak"] = HBM_PEAK / 1e9
    print(json.dumps(out), flush=True)
    img.close()
    ctx.close()


if __name__ == "__main__":
    main()
