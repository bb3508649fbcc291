"""CPU oracle for the random-walk NMR hot path.

TEST INFRASTRUCTURE ONLY: this package is the checker the HIP path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it; the product package
(``randomwalk_nmr_amd``) never does.

``rwnmr_oracle.c`` restates the walker step, folds, bitBlock packing, the
CPMG/map walk, the PFGSE measurement and the reference seeding; ``laplace.py``
restates the Laplace inversion in numpy/scipy.  Each routine cites the
reference file:line it follows (paths relative to src/backend/ of
RandomWalk_NMR).

Parity status of each oracle routine (how it is pinned, see the tests):
    xorshift64 / mod6 / direction  pinned (numpy known answers, brute force)
    folds / bitBlock               pinned (brute force, hand-packed PNG stacks)
    walk (CPMG, map, labels)       pinned (closed forms + invariants)
    pfgse                          pinned (free diffusion closed form, k=0 column)
    mt19937 / Lemire placement     pinned (C++ standard value, numpy, libstdc++)
    setup_walkers (point / cubic,  pinned (brute-force reachable sets on tiny
      voxel division)                images, draw counts, seed_and_place)
    xorshift64, accumulated        pinned against the reference's own code
      linspace/logspace,             (oracle/_ref/libref.so, oracle/build_ref.py;
      Stejskal-Tanner fit            tests/test_ref.py)
    laplace                        pinned (scipy.optimize.nnls, closed forms;
                                     the L-curve curvature against a finite-
                                     difference geometric curvature, with the
                                     reference's one denominator substitution,
                                     TikhonovSolver.cpp:111, checked as such)
"""
from __future__ import annotations

import ctypes as _ct
import os as _os
import subprocess as _sp

import numpy as _np

_HERE = _os.path.dirname(_os.path.abspath(__file__))
_LIB_PATH = _os.path.join(_HERE, "liborc.so")
_lib = None

BC_NOFLUX, BC_PERIODIC, BC_MIRROR = 0, 1, 2
BC_CODES = {"noflux": 0, "periodic": 1, "mirror": 2}

_P = _ct.c_void_p
_I = _ct.c_int
_I64 = _ct.c_int64
_U32 = _ct.c_uint32
_U64 = _ct.c_uint64
_D = _ct.c_double


def build(force: bool = False) -> str:
    """Compile liborc.so with the committed Makefile (gcc, no GPU needed)."""
    src = _os.path.join(_HERE, "rwnmr_oracle.c")
    if force or not _os.path.exists(_LIB_PATH) or _os.path.getmtime(_LIB_PATH) < _os.path.getmtime(src):
        _sp.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _os.path.exists(_LIB_PATH):
            build()
        L = _ct.CDLL(_LIB_PATH)
        L.orc_xorshift64.restype = _U64
        L.orc_xorshift64.argtypes = [_U64]
        L.orc_mod6_digit_sum.restype = _U64
        L.orc_mod6_digit_sum.argtypes = [_U64]
        L.orc_fold_periodic.restype = _I
        L.orc_fold_periodic.argtypes = [_I, _I, _I]
        L.orc_fold_mirror.restype = _I
        L.orc_fold_mirror.argtypes = [_I, _I, _I]
        L.orc_bitblock_pack.restype = None
        L.orc_bitblock_pack.argtypes = [_P, _I, _I, _I, _P]
        L.orc_walk.restype = _I
        L.orc_walk.argtypes = [_P, _I, _I, _I, _P, _I, _I, _I, _I, _I64, _P, _P, _P, _P, _P, _P, _P,
                               _P, _P, _I64, _I64, _P, _I]
        L.orc_walk_field.restype = _I
        L.orc_walk_field.argtypes = [_P, _I, _I, _I, _I, _I, _I, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D,
                                     _I64, _I64, _P, _I]
        L.orc_cpmg_reference_loop.restype = _D
        L.orc_cpmg_reference_loop.argtypes = [_P, _I, _I, _I, _I, _I, _I, _I64, _P, _P, _P, _P, _P,
                                              _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _I]
        L.orc_pfgse.restype = _I
        L.orc_pfgse.argtypes = [_P, _I, _I, _I, _I, _I, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                _I, _P, _I, _P, _I, _I, _P, _P, _P, _D, _P, _P, _I]
        L.orc_mt_state_size.restype = _ct.c_size_t
        L.orc_mt_seed.restype = None
        L.orc_mt_seed.argtypes = [_P, _U32]
        L.orc_mt_next.restype = _U32
        L.orc_mt_next.argtypes = [_P]
        L.orc_lemire_draw.restype = _U32
        L.orc_lemire_draw.argtypes = [_P, _U32]
        L.orc_seed_and_place.restype = None
        L.orc_seed_and_place.argtypes = [_U64, _I64, _U32, _P, _I64, _P, _P, _P, _P]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_ct.c_void_p)


def _c(a, dtype):
    return _np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- scalars

def xorshift64(s: int) -> int:
    return int(lib().orc_xorshift64(s))


def mod6_digit_sum(a: int) -> int:
    return int(lib().orc_mod6_digit_sum(a))


def fold_periodic(p: int, shift: int, d: int) -> int:
    return int(lib().orc_fold_periodic(p, shift, d))


def fold_mirror(p: int, shift: int, d: int) -> int:
    return int(lib().orc_fold_mirror(p, shift, d))


# --------------------------------------------------------------- bitBlock

def bitblock_dims(W, H, D):
    return (W + 3) // 4, (H + 3) // 4, (D + 3) // 4


def bitblock_pack(solid: _np.ndarray) -> _np.ndarray:
    """solid: uint8/bool volume indexed [z, y, x]; returns uint64 blocks."""
    solid = _c(solid, _np.uint8)
    D, H, W = solid.shape
    bC, bR, bD = bitblock_dims(W, H, D)
    out = _np.zeros(bC * bR * bD, dtype=_np.uint64)
    lib().orc_bitblock_pack(_ptr(solid), W, H, D, _ptr(out))
    return out


# ------------------------------------------------------------------- walk

def walk(blocks, dims, x, y, z, seed, *, echoes, steps_per_echo, bc="mirror", shift=0,
         penalty=None, energy=None, hits=None, labels=None, label_delta=None, label_hits=False,
         echo_sums=True, use_digit_sum=False, threads=0):
    """CPMG / map walk (see orc_walk).  Arrays are copied; returns a dict with
    x, y, z, seed, energy, hits, echo_sums (echoes+1) and label_hits."""
    W, H, D = dims
    n = len(x)
    x = _c(x, _np.int32).copy()
    y = _c(y, _np.int32).copy()
    z = _c(z, _np.int32).copy()
    seed = _c(seed, _np.uint64).copy()
    energy = _np.ones(n) if energy is None else _c(energy, _np.float64).copy()
    hits = _np.zeros(n, _np.uint32) if hits is None else _c(hits, _np.uint32).copy()
    blocks = _c(blocks, _np.uint64)
    pen = None if penalty is None else _c(_np.broadcast_to(penalty, (n,)), _np.float64)
    lab = None if labels is None else _c(labels, _np.uint8)
    ld = None if label_delta is None else _c(label_delta, _np.float64)
    nl = 0 if ld is None else len(ld)
    lh = _np.zeros((n, nl), _np.uint32) if (label_hits and nl) else None
    es = _np.zeros(echoes + 1) if echo_sums else None
    if pen is None and lab is None:
        pen = _np.ones(n)
    rc = lib().orc_walk(_ptr(blocks), W, H, D, _ptr(lab), nl, BC_CODES.get(bc, bc), shift,
                        int(use_digit_sum), n, _ptr(x), _ptr(y), _ptr(z), _ptr(seed), _ptr(energy),
                        _ptr(hits), _ptr(pen), _ptr(ld), _ptr(lh), echoes, steps_per_echo, _ptr(es),
                        threads)
    if rc:
        raise ValueError(f"orc_walk failed ({rc})")
    return dict(x=x, y=y, z=z, seed=seed, energy=energy, hits=hits, echo_sums=es, label_hits=lh)


def walk_field(blocks, dims, x, y, z, seed, *, echoes, steps_per_echo, field, gammatau, bc="mirror", shift=0,
               penalty=None, p_alive=None, phase=None, hits=None, use_digit_sum=False, threads=0):
    """Internal-field CPMG (see orc_walk_field).  Returns a dict with x, y,
    z, seed, p_alive, phase, hits and echo_sums (sums of p_alive cos(phase))."""
    W, H, D = dims
    n = len(x)
    x = _c(x, _np.int32).copy()
    y = _c(y, _np.int32).copy()
    z = _c(z, _np.int32).copy()
    seed = _c(seed, _np.uint64).copy()
    pa = _np.ones(n) if p_alive is None else _c(p_alive, _np.float64).copy()
    ph = _np.zeros(n) if phase is None else _c(phase, _np.float64).copy()
    hits = _np.zeros(n, _np.uint32) if hits is None else _c(hits, _np.uint32).copy()
    pen = _np.ones(n) if penalty is None else _c(_np.broadcast_to(penalty, (n,)), _np.float64)
    fld = _c(field, _np.float64)
    if fld.size != W * H * D:
        raise ValueError("field must hold W*H*D values")
    es = _np.zeros(echoes + 1)
    blocks = _c(blocks, _np.uint64)
    rc = lib().orc_walk_field(_ptr(blocks), W, H, D, BC_CODES.get(bc, bc), shift, int(use_digit_sum), n, _ptr(x),
                              _ptr(y), _ptr(z), _ptr(seed), _ptr(pa), _ptr(ph), _ptr(hits), _ptr(pen), _ptr(fld),
                              float(gammatau), echoes, steps_per_echo, _ptr(es), threads)
    if rc:
        raise ValueError(f"orc_walk_field failed ({rc})")
    return dict(x=x, y=y, z=z, seed=seed, p_alive=pa, phase=ph, hits=hits, echo_sums=es)


def cpmg_reference_loop(blocks, dims, x, y, z, seed, penalty, *, echoes, steps_per_echo,
                        bc="mirror", shift=0, use_digit_sum=True, threads=1):
    """Reference-order CPU CPMG (timing baseline).  Returns a dict: seconds
    (the step loop only), echo_sums and the final x, y, z, seed, energy, hits."""
    W, H, D = dims
    n = len(x)
    xs, ys, zs = (_c(a, _np.int32) for a in (x, y, z))
    sd = _c(seed, _np.uint64)
    pen = _c(_np.broadcast_to(penalty, (n,)), _np.float64)
    out = dict(echo_sums=_np.zeros(echoes + 1), x=_np.zeros(n, _np.int32), y=_np.zeros(n, _np.int32),
               z=_np.zeros(n, _np.int32), seed=_np.zeros(n, _np.uint64), energy=_np.zeros(n),
               hits=_np.zeros(n, _np.uint32))
    blocks = _c(blocks, _np.uint64)
    secs = lib().orc_cpmg_reference_loop(_ptr(blocks), W, H, D, BC_CODES.get(bc, bc), shift,
                                         int(use_digit_sum), n, _ptr(xs), _ptr(ys), _ptr(zs),
                                         _ptr(sd), _ptr(pen), echoes, steps_per_echo, _ptr(out["echo_sums"]),
                                         *(_ptr(out[k]) for k in ("x", "y", "z", "seed", "energy", "hits")),
                                         threads)
    if secs < 0:
        raise MemoryError("orc_cpmg_reference_loop")
    out["seconds"] = secs
    return out


def pfgse(blocks, dims, x, y, z, seed, *, steps_inc, kx, ky, kz, voxel, bc="mirror", shift=0,
          penalty=None, set_delta=None, energy=None, hits=None, x0=None, y0=None, z0=None,
          threads=0):
    """PFGSE walk + all-k measurement (see orc_pfgse).  Returns dict with M
    shaped [n_sets or 1, n_delta, n_k] and the final walker state."""
    W, H, D = dims
    n = len(x)
    x0 = _c(x if x0 is None else x0, _np.int32).copy()
    y0 = _c(y if y0 is None else y0, _np.int32).copy()
    z0 = _c(z if z0 is None else z0, _np.int32).copy()
    x = _c(x, _np.int32).copy()
    y = _c(y, _np.int32).copy()
    z = _c(z, _np.int32).copy()
    seed = _c(seed, _np.uint64).copy()
    energy = _np.ones(n) if energy is None else _c(energy, _np.float64).copy()
    hits = _np.zeros(n, _np.uint32) if hits is None else _c(hits, _np.uint32).copy()
    steps_inc = _c(steps_inc, _np.int64)
    kshape = _np.shape(kx)
    per_delta = len(kshape) == 2
    kx, ky, kz = (_c(_np.broadcast_to(_np.asarray(a, _np.float64), kshape), _np.float64) for a in (kx, ky, kz))
    nk = kshape[-1]
    nsets = 0 if set_delta is None else len(set_delta)
    sd = None if set_delta is None else _c(set_delta, _np.float64)
    pen = None if penalty is None else _c(_np.broadcast_to(penalty, (n,)), _np.float64)
    if nsets == 0 and pen is None:
        pen = _np.ones(n)
    M = _np.zeros((max(nsets, 1), len(steps_inc), nk))
    msd = _np.zeros((len(steps_inc), 3))
    blocks = _c(blocks, _np.uint64)
    rc = lib().orc_pfgse(_ptr(blocks), W, H, D, BC_CODES.get(bc, bc), shift, n, _ptr(x0), _ptr(y0),
                         _ptr(z0), _ptr(x), _ptr(y), _ptr(z), _ptr(seed), _ptr(energy), _ptr(hits),
                         _ptr(pen), nsets, _ptr(sd), len(steps_inc), _ptr(steps_inc), nk, int(per_delta),
                         _ptr(kx), _ptr(ky), _ptr(kz), float(voxel), _ptr(M), _ptr(msd), threads)
    if rc:
        raise ValueError(f"orc_pfgse failed ({rc})")
    return dict(M=M, msd=msd, x=x, y=y, z=z, seed=seed, energy=energy, hits=hits)


# ---------------------------------------------------------------- seeding

class MT19937:
    """The oracle's mt19937 (standard linear seeding)."""

    def __init__(self, seed: int):
        self._buf = _ct.create_string_buffer(int(lib().orc_mt_state_size()))
        lib().orc_mt_seed(self._buf, seed & 0xFFFFFFFF)

    def next(self) -> int:
        return int(lib().orc_mt_next(self._buf))

    def lemire(self, rng: int) -> int:
        return int(lib().orc_lemire_draw(self._buf, rng & 0xFFFFFFFF))


def pore_list(solid: _np.ndarray) -> _np.ndarray:
    """Pore voxels (solid == 0) in z, y, x raster order as (x, y, z) rows
    (B/NMR_Simulation/NMR_Simulation.cpp:930-980)."""
    zz, yy, xx = _np.nonzero(_np.asarray(solid) == 0)
    return _np.stack([xx, yy, zz], axis=1).astype(_np.int32)


def reference_pore_cap(n_pores: int, volume: int) -> int:
    """numberOfPores as the reference recomputes it from the porosity
    (B/NMR_Simulation/NMR_Simulation.cpp:915-928)."""
    porosity = float(n_pores) / float(volume)
    return int(porosity * float(volume))


def setup_walkers(solid, master_seed: int, n: int, mode: str = "random", deviation: int = 0, shift: int = 0):
    """NMR_Simulation's walker set-up, one draw at a time in Python on the
    oracle's mt19937 (B/NMR_Simulation/NMR_Simulation.cpp: generator :86-95,
    setWalkers :324-410, createWalkers :1139-1204, createWalkersIDList
    :1086-1136, placeWalkersInCubicSpace :1395-1454 over the cube of
    countPoresInCubicSpace / createPoreList(v1, v2) :728-816, :982-1069 with
    v1 = c - deviation, v2 = c + deviation, c = (dim - 1) / 2, upper bound
    exclusive and clipped to the image; applyVoxelDivision :192-268).
    Returns (seeds, x, y, z, draws) or None when the reference places no
    walker (point on a solid voxel, empty cube)."""
    solid = _np.asarray(solid)
    D, H, W = solid.shape
    g = MT19937(master_seed)
    for _ in range(4096):
        g.next()
    draws = 4096
    seeds = _np.zeros(n, _np.uint64)
    for i in range(n):
        hi, lo = g.next(), g.next()
        seeds[i] = (hi << 32) | lo
        draws += 2
    x = _np.zeros(n, _np.int64)
    y = _np.zeros(n, _np.int64)
    z = _np.zeros(n, _np.int64)

    if mode == "random":
        pores = pore_list(solid)
        cap = reference_pore_cap(len(pores), solid.size)
        picks = []
        for i in range(n):
            picks.append(g.lemire(cap + 1))
        draws += _lemire_outputs(master_seed, 4096 + 2 * n, [cap + 1] * n)
        for i, idx in enumerate(picks):
            idx = min(idx, len(pores) - 1)
            x[i], y[i], z[i] = pores[idx]
    elif mode == "point":
        cx, cy, cz = W // 2, H // 2, D // 2
        if solid[cz, cy, cx]:
            return None
        x[:], y[:], z[:] = cx, cy, cz
    else:
        lo = [(W - 1) // 2 - deviation, (H - 1) // 2 - deviation, (D - 1) // 2 - deviation]
        hi = [min((W - 1) // 2 + deviation, W), min((H - 1) // 2 + deviation, H), min((D - 1) // 2 + deviation, D)]
        cube = []
        if min(lo) >= 0:
            for zz in range(lo[2], hi[2]):
                for yy in range(lo[1], hi[1]):
                    for xx in range(lo[0], hi[0]):
                        if not solid[zz, yy, xx]:
                            cube.append((xx, yy, zz))
        if not cube:
            return None
        picks = [g.lemire(len(cube) + 1) for _ in range(n)]
        draws += _lemire_outputs(master_seed, 4096 + 2 * n, [len(cube) + 1] * n)
        for i, idx in enumerate(picks):
            x[i], y[i], z[i] = cube[min(idx, len(cube) - 1)]
    if shift:
        f = 1 << shift
        start = draws
        for i in range(n):
            x[i] = x[i] * f + g.lemire(f)
            y[i] = y[i] * f + g.lemire(f)
            z[i] = z[i] * f + g.lemire(f)
        draws += _lemire_outputs(master_seed, start, [f] * (3 * n))
    return seeds, x.astype(_np.int32), y.astype(_np.int32), z.astype(_np.int32), draws


def _lemire_outputs(master_seed: int, skip: int, ranges) -> int:
    """32-bit outputs Lemire's method takes for the given ranges, starting
    `skip` outputs into the stream: one per draw, plus one per rejection
    (low word of output * range under (2^32 - range) mod range)."""
    g = MT19937(master_seed)
    for _ in range(skip):
        g.next()
    used = 0
    for r in ranges:
        thr = ((1 << 32) - r) % r
        while True:
            used += 1
            if ((g.next() * r) & 0xFFFFFFFF) >= thr:
                break
    return used


def seed_and_place(master_seed: int, n: int, pores: _np.ndarray, volume: int):
    """Seeds and start voxels drawn the reference's way (see orc_seed_and_place)."""
    pores = _c(pores, _np.int32)
    cap = reference_pore_cap(len(pores), volume)
    seeds = _np.zeros(n, _np.uint64)
    x = _np.zeros(n, _np.int32)
    y = _np.zeros(n, _np.int32)
    z = _np.zeros(n, _np.int32)
    lib().orc_seed_and_place(master_seed, n, cap, _ptr(pores), len(pores), _ptr(seeds), _ptr(x),
                             _ptr(y), _ptr(z))
    return seeds, x, y, z
