"""Collision-histogram CPMG method, restated in numpy (TEST INFRASTRUCTURE
ONLY: used by tests/ as the checker; never part of the product path).

Follows, under src/backend/ of RandomWalk_NMR:
  histogram list        B/NMR_Simulation/NMR_Simulation.cpp:1456-1479
  bins and amplitudes   B/NMR_Simulation/CollisionHistogram.cpp:57-250
  penalties             B/NMR_Simulation/NMR_cpmg.cpp:235-279
  decay                 B/NMR_Simulation/NMR_cpmg.cpp:281-344
Pinned in tests/test_host.py by closed forms (one-bin decay = penalty^e,
uniform-rate walkers, probability conservation) and brute-force binning.
"""
import math

import numpy as np


def histogram_list(echoes, n):
    if echoes < n:
        n = 1
    per = echoes // n
    segs = [(i * per, i * per + per) for i in range(n)]
    f, l = segs[-1]
    segs[-1] = (f, l + echoes - n * per)
    return segs


def collision_histogram(collisions, steps, size, log_scale=False):
    xi_all = [c / float(steps) for c in collisions]
    n = len(collisions)
    if not log_scale:
        gap = 1.0 / size
        bins = [i * gap + 0.5 * gap for i in range(size)]
        amps = [0.0] * size
        for xi in xi_all:
            idx = int(math.floor(xi / gap))
            if idx >= size:
                idx -= 1
            amps[idx] += 1.0
        return np.array(bins), np.array([a / n for a in amps])
    nb = size + 1
    nz = [xi for xi, c in zip(xi_all, collisions) if c != 0]
    if nz:
        e0 = round(math.log10(min(nz)))
        step = (0.0 - e0) / (size - 1.0)
        bins = [e0 - 1.0] + [10.0 ** (e0 + step * (i - 1)) for i in range(1, nb)]
    else:
        bins = [0.0] * nb
    amps = [0.0] * nb
    if nz:
        gap = math.log10(bins[2]) - math.log10(bins[1])
        lo = math.log10(bins[1])
    for xi, c in zip(xi_all, collisions):
        if c == 0:
            amps[0] += 1.0
        else:
            idx = int(math.floor(round((math.log10(xi) - lo) / gap))) + 1
            amps[idx if 0 < idx < nb else 1] += 1.0
    return np.array(bins), np.array([a / n for a in amps])


def sigmoid_rho(p, xi):
    return (p[1] + (p[0] - p[1]) / (1 + math.exp(-p[3] * (xi - p[2])))) + \
        (p[5] + (p[4] - p[5]) / (1 + math.exp(-p[7] * (xi - p[6]))))


def penalties(bins, rho, step_length, D0, steps_per_echo, sigmoid=None):
    out = []
    for xi in bins:
        r = rho if sigmoid is None else sigmoid_rho(sigmoid, xi)
        d = 1.0 - (2.0 / 3.0) * ((r * step_length) / (D0 * 1e3))
        out.append(d ** (xi * steps_per_echo))
    return np.array(out)


def decay(amps, segments, pen):
    sig = [1.0]
    for h, (f, l) in enumerate(segments):
        dist = [sig[-1] * a for a in amps[h]]
        for _ in range(f, l):
            dist = [d * p for d, p in zip(dist, pen)]
            s = 0.0
            for d in dist:
                s += d
            sig.append(s)
    return np.array(sig)
