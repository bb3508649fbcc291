"""Python host API over the C ABI: contexts, images, walkers and the
CPMG / map / PFGSE runs.

Arrays may be numpy arrays (host memory: the call copies in and out and
returns when done) or torch CUDA tensors (device memory already resident in
HBM: the call enqueues on the context's stream; ``sync=True`` waits).
Reference paths are under src/backend/ of RandomWalk_NMR.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from ._lib import Walkers, check, load

BC = {"noflux": 0, "periodic": 1, "mirror": 2}


def bc_code(bc) -> int:
    if isinstance(bc, str):
        if bc not in BC:
            raise ValueError(f"unknown boundary condition {bc!r}")
        return BC[bc]
    return int(bc)


def _is_dev(a) -> bool:
    return a is not None and hasattr(a, "data_ptr") and getattr(a, "is_cuda", False)


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ct.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ct.c_void_p)


def _host(a, dtype, n=None, copy=True):
    if a is None:
        return None
    out = np.array(a, dtype=dtype, copy=copy, order="C") if copy else np.ascontiguousarray(a, dtype=dtype)
    if n is not None and out.shape != (n,):
        out = np.ascontiguousarray(np.broadcast_to(out, (n,)), dtype=dtype)
    return out


class Context:
    """rwnmr_ctx: one device, one stream, grow-only device scratch."""

    def __init__(self, device: int = 0, threads_per_block: int = 0, walkers_per_lane: int = 0):
        L = load()
        h = ct.c_void_p()
        code = L.rwnmr_ctx_create(ct.byref(h), device)
        if code != 0:
            raise _lib.RwnmrError(code, f"rwnmr_ctx_create(device={device}) failed (no GPU?)")
        self.handle = h
        self.device = device
        if threads_per_block or walkers_per_lane:
            self.set_launch(threads_per_block, walkers_per_lane)

    def set_launch(self, threads_per_block: int = 0, walkers_per_lane: int = 0):
        check(load().rwnmr_ctx_set_launch(self.handle, threads_per_block, walkers_per_lane), self.handle)

    def get_launch(self):
        t, k = ct.c_uint32(), ct.c_uint32()
        check(load().rwnmr_ctx_get_launch(self.handle, ct.byref(t), ct.byref(k)), self.handle)
        return t.value, k.value

    def set_image_layout(self, sx: int = 1, sy: int = 1, sz: int = 1):
        """Brick layout (log2 blocks per axis) of images uploaded afterwards."""
        check(load().rwnmr_ctx_set_image_layout(self.handle, sx, sy, sz), self.handle)

    def set_locality(self, enable: bool = True, resort_every: int = 1000, min_walkers: int = 65536):
        """Morton ordering of walkers + XCD-striped tiles for large runs."""
        check(load().rwnmr_ctx_set_locality(self.handle, int(enable), int(resort_every), int(min_walkers)),
              self.handle)

    def set_unfolded(self, mode: int = 3):
        """Periodic / mirror walks (and noflux with modes 3, 4 on an image
        of sides >= 2, on the periodic unfolded image): 0 folded kernels, 1 unfolded image (no
        per-step fold), 2 unfolded + per-walker 16-byte chunk cache, 3 and 4
        the chunk tested by one shift and the next direction drawn a step
        ahead, through the 4 KiB byte-key table (3, default) or the 64 KiB
        wide direction table (4); 5 as 3, with CPMG of one walker per lane
        loading both possible next chunks a step ahead (latency-bound runs
        of few walkers); 6 as 3, with mirror walks keeping an 8-byte (4 x 4
        x 4 voxel) chunk (3 VALU per step fewer, a quarter more chunk
        loads)."""
        check(load().rwnmr_ctx_set_unfolded(self.handle, int(mode)), self.handle)

    def set_timing(self, enable: bool = True):
        check(load().rwnmr_ctx_set_timing(self.handle, int(enable)), self.handle)

    def get_timing(self, reset: bool = False):
        """(accumulated walk/PFGSE kernel milliseconds, launches), HIP events."""
        ms, n = ct.c_double(), ct.c_uint64()
        check(load().rwnmr_ctx_get_timing(self.handle, ct.byref(ms), ct.byref(n), int(reset)), self.handle)
        return ms.value, n.value

    def autotune(self, img, x, y, z, seed, *, steps_per_echo, bc="mirror", trial_echoes=64, penalty=None,
                 delta=1.0):
        """Pick the fastest launch shape on copies of these walkers
        (rwnmr_ctx_autotune).  Returns (threads_per_block, walkers_per_lane, ms)."""
        dev = _is_dev(x)
        if dev:
            _torch_sync()
        else:
            x, y, z = _host(x, np.int32), _host(y, np.int32), _host(z, np.int32)
            seed = _host(seed, np.uint64)
        w = _walker_struct(x, y, z, seed, None, None, penalty if dev else None, delta, len(x), int(dev))
        t, k, ms = ct.c_uint32(), ct.c_uint32(), ct.c_double()
        check(load().rwnmr_ctx_autotune(self.handle, img.handle, ct.byref(w), int(steps_per_echo), bc_code(bc),
                                        int(trial_echoes), ct.byref(t), ct.byref(k), ct.byref(ms)), self.handle)
        return t.value, k.value, ms.value

    def get_tuning(self):
        """(threads_per_block, walkers_per_lane, unfolded_mode, resort_every) now in use
        (rwnmr_ctx_get_tuning)."""
        t, k, r = ct.c_uint32(), ct.c_uint32(), ct.c_uint32()
        u = ct.c_int32()
        check(load().rwnmr_ctx_get_tuning(self.handle, ct.byref(t), ct.byref(k), ct.byref(u), ct.byref(r)),
              self.handle)
        return t.value, k.value, u.value, r.value

    def probe_gather(self, table_bytes: int, loads_per_thread: int = 256):
        """Random 16-byte loads per second from a device table of
        table_bytes (rwnmr_probe_gather).  Returns (loads_per_s, kernel_ms)."""
        rate, ms = ct.c_double(), ct.c_double()
        code = load().rwnmr_probe_gather(self.handle, int(table_bytes), int(loads_per_thread), ct.byref(rate),
                                         ct.byref(ms))
        check(code, self.handle)
        return rate.value, ms.value

    @property
    def stream(self) -> int:
        return load().rwnmr_ctx_stream(self.handle) or 0

    def synchronize(self):
        check(load().rwnmr_ctx_synchronize(self.handle), self.handle)

    def close(self):
        if getattr(self, "handle", None):
            load().rwnmr_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ------------------------------------------------------------------ images

def bitblock_dims(W, H, D):
    """ceil(dim / 4) per axis (B/BitBlock/bitBlock.cpp:138-166)."""
    return (W + 3) // 4, (H + 3) // 4, (D + 3) // 4


def pack_bitblocks(solid: np.ndarray) -> np.ndarray:
    """Pack a [z, y, x] volume (nonzero = solid) into bitBlock words
    (B/BitBlock/bitBlock.cpp:168-221)."""
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    D, H, W = solid.shape
    bC, bR, bD = bitblock_dims(W, H, D)
    out = np.zeros(bC * bR * bD, np.uint64)
    check(load().rwnmr_bitblock_pack(_ptr(solid), W, H, D, _ptr(out)))
    return out


class Image:
    """rwnmr_image: a bitBlock image (and optional solid labels) in HBM."""

    def __init__(self, ctx: Context, blocks: np.ndarray, dims, labels: np.ndarray | None = None,
                 n_labels: int = 0):
        W, H, D = (int(v) for v in dims)
        blocks = np.ascontiguousarray(blocks, dtype=np.uint64)
        bC, bR, bD = bitblock_dims(W, H, D)
        if blocks.size != bC * bR * bD:
            raise ValueError(f"expected {bC * bR * bD} bitBlock words, got {blocks.size}")
        lab = None
        if labels is not None:
            lab = np.ascontiguousarray(labels, dtype=np.uint8)
            if lab.size != W * H * D:
                raise ValueError("labels must have W*H*D entries")
            if not n_labels:
                n_labels = int(lab.max()) + 1
        h = ct.c_void_p()
        check(load().rwnmr_image_upload(ctx.handle, _ptr(blocks), W, H, D, _ptr(lab), int(n_labels),
                                        ct.byref(h)), ctx.handle)
        self.ctx = ctx
        self.handle = h
        self.dims = (W, H, D)
        self.n_labels = int(n_labels) if lab is not None else 0

    @classmethod
    def from_solid(cls, ctx: Context, solid: np.ndarray, labels=None, n_labels: int = 0):
        D, H, W = solid.shape
        return cls(ctx, pack_bitblocks(solid), (W, H, D), labels, n_labels)

    def close(self):
        if getattr(self, "handle", None):
            load().rwnmr_image_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- synthetic inputs

def sphere_pack(W, H, D, r_min, r_max, porosity, seed):
    """Seeded overlapping solid spheres until the pore fraction <= porosity.
    Returns (solid [D, H, W] uint8, achieved porosity)."""
    solid = np.zeros((D, H, W), np.uint8)
    phi = ct.c_double()
    check(load().rwnmr_synth_sphere_pack(W, H, D, float(r_min), float(r_max), float(porosity), int(seed),
                                         _ptr(solid), ct.byref(phi)))
    return solid, phi.value


def voronoi_labels(solid, n_sites, n_labels, seed):
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    D, H, W = solid.shape
    lab = np.zeros_like(solid)
    check(load().rwnmr_synth_voronoi_labels(W, H, D, n_sites, n_labels, int(seed), _ptr(solid), _ptr(lab)))
    return lab


def seed_and_place(solid, master_seed, n):
    """Reference seeding and uniform placement (include/rwnmr/rwnmr_host.h)."""
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    D, H, W = solid.shape
    seeds = np.zeros(n, np.uint64)
    x = np.zeros(n, np.int32)
    y = np.zeros(n, np.int32)
    z = np.zeros(n, np.int32)
    check(load().rwnmr_seed_and_place(_ptr(solid), W, H, D, int(master_seed), int(n), _ptr(seeds), _ptr(x),
                                      _ptr(y), _ptr(z)))
    return seeds, x, y, z


PLACEMENTS = {"random": 0, "point": 1, "cubic": 2}


def setup_walkers(solid, master_seed, n, *, placement="random", deviation=0, shift=0):
    """The reference's whole walker set-up: seeds, placement mode
    ("random" / "point" / "cubic") and voxel division (include/rwnmr/rwnmr_host.h).
    Returns dict(seed, x, y, z, occupancy, draws)."""
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    D, H, W = solid.shape
    seeds = np.zeros(n, np.uint64)
    x, y, z = (np.zeros(n, np.int32) for _ in range(3))
    occ, draws = ct.c_double(), ct.c_uint64()
    check(load().rwnmr_setup_walkers(_ptr(solid), W, H, D, int(master_seed), int(n), PLACEMENTS[placement],
                                     int(deviation), int(shift), _ptr(seeds), _ptr(x), _ptr(y), _ptr(z),
                                     ct.byref(occ), ct.byref(draws)))
    return dict(seed=seeds, x=x, y=y, z=z, occupancy=occ.value, draws=draws.value)


def time_step(voxel, D0):
    return load().rwnmr_time_step(float(voxel), float(D0))


def time_framework(time_ms, voxel, D0, steps_per_echo):
    e, s = ct.c_uint32(), ct.c_uint64()
    check(load().rwnmr_time_framework(float(time_ms), float(voxel), float(D0), int(steps_per_echo),
                                      ct.byref(e), ct.byref(s)))
    return e.value, s.value


def penalty(rho, voxel, D0):
    return load().rwnmr_penalty(float(rho), float(voxel), float(D0))


# ---------------------------------------------------------------- walkers

def _walker_struct(x, y, z, seed, energy, hits, pen, delta, n, on_device):
    w = Walkers()
    w.x, w.y, w.z, w.seed = _ptr(x), _ptr(y), _ptr(z), _ptr(seed)
    w.energy, w.hits, w.penalty = _ptr(energy), _ptr(hits), _ptr(pen)
    w.delta_uniform = float(delta)
    w.n = int(n)
    w.on_device = int(on_device)
    return w


def _torch_sync():
    import torch
    torch.cuda.current_stream().synchronize()


def cpmg_run(ctx: Context, img: Image, x, y, z, seed, *, echoes, steps_per_echo, bc="mirror", shift=0,
             penalty=None, delta=1.0, energy=None, hits=None, echo_sums=True, label_delta=None,
             label_hits=None, sync=True):
    """CPMG walk (rwnmr_cpmg_run / rwnmr_cpmg_run_labels).

    Host (numpy) inputs are copied and the results returned in a dict
    (x, y, z, seed, energy, hits, echo_sums, label_hits).  Device (torch CUDA)
    inputs are updated in place; echo_sums must then be a device tensor or
    None."""
    dev = _is_dev(x)
    n = len(x)
    L = load()
    if dev:
        _torch_sync()
        es = echo_sums if (echo_sums is not None and echo_sums is not True and echo_sums is not False) else None
        w = _walker_struct(x, y, z, seed, energy, hits, penalty, delta, n, 1)
        if label_delta is not None:
            code = L.rwnmr_cpmg_run_labels(ctx.handle, img.handle, ct.byref(w), echoes, steps_per_echo,
                                           bc_code(bc), shift, _ptr(label_delta), _ptr(label_hits), _ptr(es))
        else:
            code = L.rwnmr_cpmg_run(ctx.handle, img.handle, ct.byref(w), echoes, steps_per_echo, bc_code(bc),
                                    shift, _ptr(es))
        check(code, ctx.handle)
        if sync:
            ctx.synchronize()
        return dict(x=x, y=y, z=z, seed=seed, energy=energy, hits=hits, echo_sums=es, label_hits=label_hits)
    x, y, z = _host(x, np.int32), _host(y, np.int32), _host(z, np.int32)
    seed = _host(seed, np.uint64)
    energy = np.ones(n) if energy is None else _host(energy, np.float64, n)
    hits = np.zeros(n, np.uint32) if hits is None else _host(hits, np.uint32, n)
    pen = None if penalty is None else _host(penalty, np.float64, n)
    es = np.zeros(echoes + 1) if echo_sums else None
    w = _walker_struct(x, y, z, seed, energy, hits, pen, delta, n, 0)
    lh = None
    if label_delta is not None:
        ld = _host(label_delta, np.float64)
        if label_hits is not None:
            lh = _host(label_hits, np.uint32).reshape(-1, n) if not isinstance(label_hits, bool) else \
                np.zeros((img.n_labels, n), np.uint32)
        code = L.rwnmr_cpmg_run_labels(ctx.handle, img.handle, ct.byref(w), echoes, steps_per_echo, bc_code(bc),
                                       shift, _ptr(ld), _ptr(lh), _ptr(es))
    else:
        code = L.rwnmr_cpmg_run(ctx.handle, img.handle, ct.byref(w), echoes, steps_per_echo, bc_code(bc), shift,
                                _ptr(es))
    check(code, ctx.handle)
    return dict(x=x, y=y, z=z, seed=seed, energy=energy, hits=hits, echo_sums=es, label_hits=lh)


def cpmg_forward_jacobian(ctx: Context, img: Image, x, y, z, seed, *, echoes, steps_per_echo, label_delta,
                          bc="mirror", energy=None, label_hits=None, echo_sums=None, jac=None, sync=True):
    """Per-label CPMG that also returns the Jacobian of the echo sums with
    respect to the log deltas (rwnmr_cpmg_forward_jacobian):
    jac[l, n] = dM(n)/d log delta_l = sum_w h_wl(n) E_w(n).
    Host inputs are copied and a dict (x, y, z, seed, energy, label_hits,
    echo_sums [echoes+1], jac [n_labels, echoes+1]) returned; device inputs
    are updated in place, echo_sums and jac must then be device tensors of
    echoes+1 and n_labels * (echoes+1) doubles."""
    dev = _is_dev(x)
    n = len(x)
    L = load()
    if dev:
        _torch_sync()
        w = _walker_struct(x, y, z, seed, energy, None, None, 1.0, n, 1)
        check(L.rwnmr_cpmg_forward_jacobian(ctx.handle, img.handle, ct.byref(w), echoes, steps_per_echo,
                                            bc_code(bc), _ptr(label_delta), _ptr(label_hits), _ptr(echo_sums),
                                            _ptr(jac)), ctx.handle)
        if sync:
            ctx.synchronize()
        return dict(x=x, y=y, z=z, seed=seed, energy=energy, label_hits=label_hits, echo_sums=echo_sums, jac=jac)
    x, y, z = _host(x, np.int32), _host(y, np.int32), _host(z, np.int32)
    seed = _host(seed, np.uint64)
    energy = np.ones(n) if energy is None else _host(energy, np.float64, n)
    lh = None if label_hits is None else _host(label_hits, np.uint32).reshape(-1, n)
    ld = _host(label_delta, np.float64)
    es = np.zeros(echoes + 1)
    jc = np.zeros((img.n_labels, echoes + 1))
    w = _walker_struct(x, y, z, seed, energy, None, None, 1.0, n, 0)
    check(L.rwnmr_cpmg_forward_jacobian(ctx.handle, img.handle, ct.byref(w), echoes, steps_per_echo, bc_code(bc),
                                        _ptr(ld), _ptr(lh), _ptr(es), _ptr(jc)), ctx.handle)
    return dict(x=x, y=y, z=z, seed=seed, energy=energy, label_hits=lh, echo_sums=es, jac=jc)


def cpmg_run_label_sets(ctx: Context, img: Image, x, y, z, seed, *, echoes, steps_per_echo, set_label_delta,
                        bc="mirror", energy=None, echo_sums=None, sync=True):
    """Several per-label relaxivity sets on one trajectory
    (rwnmr_cpmg_run_label_sets): set_label_delta is [n_sets, n_labels].
    Host inputs are copied and a dict (x, y, z, seed, energy, echo_sums
    [n_sets, echoes+1]) returned; device inputs are updated in place and
    echo_sums must be a device tensor of n_sets * (echoes+1) doubles."""
    dev = _is_dev(x)
    n = len(x)
    L = load()
    if dev:
        _torch_sync()
        n_sets = set_label_delta.shape[0]
        w = _walker_struct(x, y, z, seed, energy, None, None, 1.0, n, 1)
        check(L.rwnmr_cpmg_run_label_sets(ctx.handle, img.handle, ct.byref(w), echoes, steps_per_echo, bc_code(bc),
                                          n_sets, _ptr(set_label_delta), _ptr(echo_sums)), ctx.handle)
        if sync:
            ctx.synchronize()
        return dict(x=x, y=y, z=z, seed=seed, energy=energy, echo_sums=echo_sums)
    sld = np.ascontiguousarray(set_label_delta, np.float64)
    n_sets = sld.shape[0]
    x, y, z = _host(x, np.int32), _host(y, np.int32), _host(z, np.int32)
    seed = _host(seed, np.uint64)
    energy = np.ones(n) if energy is None else _host(energy, np.float64, n)
    es = np.zeros((n_sets, echoes + 1))
    w = _walker_struct(x, y, z, seed, energy, None, None, 1.0, n, 0)
    check(L.rwnmr_cpmg_run_label_sets(ctx.handle, img.handle, ct.byref(w), echoes, steps_per_echo, bc_code(bc),
                                      n_sets, _ptr(sld), _ptr(es)), ctx.handle)
    return dict(x=x, y=y, z=z, seed=seed, energy=energy, echo_sums=es)


class CpmgGraph:
    """A CPMG run captured as a HIP graph (rwnmr_cpmg_graph_create) on
    device walkers x, y, z, seed (torch CUDA tensors, walked in place; energy
    and hits optional).  With start=(x0, y0, z0, seed0[, energy0]) every
    launch() is one full run from that start state; without it launches
    continue the walk.  Keep the tensors alive while the graph exists."""

    def __init__(self, ctx: Context, img: Image, x, y, z, seed, *, echoes, steps_per_echo, bc="mirror",
                 penalty=None, delta=1.0, energy=None, hits=None, echo_sums=None, start=None):
        self.ctx = ctx
        self._keep = (img, x, y, z, seed, penalty, energy, hits, echo_sums, start)
        n = len(x)
        self._w = _walker_struct(x, y, z, seed, energy, hits, penalty, delta, n, 1)
        st = None
        if start is not None:
            e0 = start[4] if len(start) > 4 else None
            self._s = _walker_struct(start[0], start[1], start[2], start[3], e0, None, None, 1.0, n, 1)
            st = ct.byref(self._s)
        h = ct.c_void_p()
        _torch_sync()
        check(load().rwnmr_cpmg_graph_create(ctx.handle, img.handle, ct.byref(self._w), st, int(echoes),
                                             int(steps_per_echo), bc_code(bc), _ptr(echo_sums), ct.byref(h)),
              ctx.handle)
        self.handle = h

    def launch(self, sync=False):
        check(load().rwnmr_graph_launch(self.handle), self.ctx.handle)
        if sync:
            self.ctx.synchronize()

    def close(self):
        if getattr(self, "handle", None):
            load().rwnmr_graph_destroy(self.handle)
            self.handle = None


def map_run(ctx: Context, img: Image, x, y, z, seed, *, steps, bc="mirror", shift=0, hits=None, sync=True):
    """Hit-count walk (rwnmr_map_run)."""
    dev = _is_dev(x)
    n = len(x)
    L = load()
    if dev:
        _torch_sync()
        w = _walker_struct(x, y, z, seed, None, hits, None, 1.0, n, 1)
        check(L.rwnmr_map_run(ctx.handle, img.handle, ct.byref(w), int(steps), bc_code(bc), shift), ctx.handle)
        if sync:
            ctx.synchronize()
        return dict(x=x, y=y, z=z, seed=seed, hits=hits)
    x, y, z = _host(x, np.int32), _host(y, np.int32), _host(z, np.int32)
    seed = _host(seed, np.uint64)
    hits = np.zeros(n, np.uint32) if hits is None else _host(hits, np.uint32, n)
    w = _walker_struct(x, y, z, seed, None, hits, None, 1.0, n, 0)
    check(L.rwnmr_map_run(ctx.handle, img.handle, ct.byref(w), int(steps), bc_code(bc), shift), ctx.handle)
    return dict(x=x, y=y, z=z, seed=seed, hits=hits)


def pfgse_run(ctx: Context, img: Image, x, y, z, seed, *, steps_inc, kx, ky, kz, voxel, bc="mirror",
              x0=None, y0=None, z0=None, penalty=None, delta=1.0, set_delta=None, energy=None, hits=None,
              M=None, msd=None, sync=True):
    """PFGSE walk + all-k measurement (rwnmr_pfgse_run).  kx/ky/kz are 1-D
    (one k set) or 2-D [n_delta, n_k] (a k set per diffusion time).  Returns
    dict with M shaped [n_sets or 1, n_delta, n_k], msd [n_delta, 3] (sums of
    squared displacements in voxel^2) plus the walker state."""
    dev = _is_dev(x)
    n = len(x)
    L = load()
    nd = len(steps_inc)
    per_delta = len(kx.shape) == 2
    nk = kx.shape[-1]
    ns = 0 if set_delta is None else len(set_delta)
    if dev:
        _torch_sync()
        w = _walker_struct(x, y, z, seed, energy, hits, penalty, delta, n, 1)
        code = L.rwnmr_pfgse_run(ctx.handle, img.handle, ct.byref(w), _ptr(x0), _ptr(y0), _ptr(z0),
                                 _ptr(steps_inc), nd, _ptr(kx), _ptr(ky), _ptr(kz), nk, int(per_delta),
                                 float(voxel), bc_code(bc), 0, ns, _ptr(set_delta), _ptr(M), _ptr(msd))
        check(code, ctx.handle)
        if sync:
            ctx.synchronize()
        return dict(M=M, msd=msd, x=x, y=y, z=z, seed=seed, energy=energy, hits=hits)
    x0 = _host(x if x0 is None else x0, np.int32)
    y0 = _host(y if y0 is None else y0, np.int32)
    z0 = _host(z if z0 is None else z0, np.int32)
    x, y, z = _host(x, np.int32), _host(y, np.int32), _host(z, np.int32)
    seed = _host(seed, np.uint64)
    energy = np.ones(n) if energy is None else _host(energy, np.float64, n)
    hits = np.zeros(n, np.uint32) if hits is None else _host(hits, np.uint32, n)
    pen = None if penalty is None else _host(penalty, np.float64, n)
    si = _host(steps_inc, np.uint32)
    kx, ky, kz = (np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), np.shape(kx)))
                  for a in (kx, ky, kz))
    sd = None if set_delta is None else _host(set_delta, np.float64)
    M = np.zeros((max(ns, 1), nd, nk))
    msd = np.zeros((nd, 3))
    w = _walker_struct(x, y, z, seed, energy, hits, pen, delta, n, 0)
    code = L.rwnmr_pfgse_run(ctx.handle, img.handle, ct.byref(w), _ptr(x0), _ptr(y0), _ptr(z0), _ptr(si), nd,
                             _ptr(kx), _ptr(ky), _ptr(kz), nk, int(per_delta), float(voxel), bc_code(bc), 0, ns,
                             _ptr(sd), _ptr(M), _ptr(msd))
    check(code, ctx.handle)
    return dict(M=M, msd=msd, x=x, y=y, z=z, seed=seed, energy=energy, hits=hits)


def cpmg_field_run(ctx: Context, img: Image, x, y, z, seed, *, echoes, steps_per_echo, field, gyromagnetic_ratio,
                   time_interval_ms, bc="mirror", penalty=None, delta=1.0, p_alive=None, phase=None, energy=None,
                   hits=None, echo_sums=True, sync=True):
    """Internal-field CPMG (rwnmr_cpmg_field_run): echo sums of
    p_alive cos(phase).  Host inputs are copied and results returned (x, y,
    z, seed, p_alive, phase, energy, hits, echo_sums); device inputs are
    updated in place."""
    dev = _is_dev(x)
    n = len(x)
    L = load()
    if dev:
        _torch_sync()
        es = echo_sums if (echo_sums is not None and echo_sums is not True and echo_sums is not False) else None
        w = _walker_struct(x, y, z, seed, energy, hits, penalty, delta, n, 1)
        check(L.rwnmr_cpmg_field_run(ctx.handle, img.handle, ct.byref(w), echoes, steps_per_echo, bc_code(bc),
                                     _ptr(field), float(gyromagnetic_ratio), float(time_interval_ms), _ptr(p_alive),
                                     _ptr(phase), _ptr(es)), ctx.handle)
        if sync:
            ctx.synchronize()
        return dict(x=x, y=y, z=z, seed=seed, p_alive=p_alive, phase=phase, energy=energy, hits=hits, echo_sums=es)
    x, y, z = _host(x, np.int32), _host(y, np.int32), _host(z, np.int32)
    seed = _host(seed, np.uint64)
    energy = np.ones(n) if energy is None else _host(energy, np.float64, n)
    hits = np.zeros(n, np.uint32) if hits is None else _host(hits, np.uint32, n)
    pen = None if penalty is None else _host(penalty, np.float64, n)
    pa = np.ones(n) if p_alive is None else _host(p_alive, np.float64, n)
    ph = np.zeros(n) if phase is None else _host(phase, np.float64, n)
    fld = np.ascontiguousarray(field, np.float64)
    if fld.size != img.dims[0] * img.dims[1] * img.dims[2]:
        raise ValueError("field must hold W*H*D values")
    es = np.zeros(echoes + 1) if echo_sums else None
    w = _walker_struct(x, y, z, seed, energy, hits, pen, delta, n, 0)
    check(L.rwnmr_cpmg_field_run(ctx.handle, img.handle, ct.byref(w), echoes, steps_per_echo, bc_code(bc),
                                 _ptr(fld), float(gyromagnetic_ratio), float(time_interval_ms), _ptr(pa), _ptr(ph),
                                 _ptr(es)), ctx.handle)
    return dict(x=x, y=y, z=z, seed=seed, p_alive=pa, phase=ph, energy=energy, hits=hits, echo_sums=es)


def pfgse_run_sampled(ctx: Context, img: Image, x, y, z, seed, *, n_samples, steps_inc, kx, ky, kz, voxel,
                      bc="mirror", penalty=None, delta=1.0, set_delta=None):
    """PFGSE with walker samples (rwnmr_pfgse_run_sampled), host arrays.
    Returns dict(M_samples [n_samples, sets|1, n_delta, n_k], M, msd and
    the final walker state)."""
    n = len(x)
    x0, y0, z0 = _host(x, np.int32), _host(y, np.int32), _host(z, np.int32)
    x, y, z = x0.copy(), y0.copy(), z0.copy()
    seed = _host(seed, np.uint64)
    energy, hits = np.ones(n), np.zeros(n, np.uint32)
    pen = None if penalty is None else _host(penalty, np.float64, n)
    si = np.ascontiguousarray(steps_inc, np.uint32)
    kx, ky, kz = (np.ascontiguousarray(k, np.float64) for k in (kx, ky, kz))
    per_delta = kx.ndim == 2
    nd, nk = len(si), kx.shape[-1]
    sd = None if set_delta is None else np.ascontiguousarray(set_delta, np.float64)
    ns = 0 if sd is None else len(sd)
    Ms = np.zeros((n_samples, max(ns, 1), nd, nk))
    M = np.zeros((max(ns, 1), nd, nk))
    msd = np.zeros((nd, 3))
    w = _walker_struct(x, y, z, seed, energy, hits, pen, delta, n, 0)
    check(load().rwnmr_pfgse_run_sampled(ctx.handle, img.handle, ct.byref(w), _ptr(x0), _ptr(y0), _ptr(z0), _ptr(si),
                                         nd, _ptr(kx), _ptr(ky), _ptr(kz), nk, int(per_delta), float(voxel),
                                         bc_code(bc), ns, _ptr(sd), int(n_samples), _ptr(Ms), _ptr(M), _ptr(msd)),
          ctx.handle)
    return dict(M_samples=Ms, M=M, msd=msd, x=x, y=y, z=z, seed=seed, energy=energy, hits=hits)


def pfgse_sample_stats(M_samples, rhs, n_points):
    """Walker-sampling statistics of one diffusion time (rwnmr_pfgse_sample_stats):
    M_samples [n_samples, n_k]."""
    Ms = np.ascontiguousarray(M_samples, np.float64)
    S, nk = Ms.shape
    r = np.ascontiguousarray(rhs, np.float64)
    out = {k: np.zeros(nk) for k in ("mean_M", "std_M", "mean_lhs", "std_lhs")}
    Ds = np.zeros(S)
    dm, dsd, de = ct.c_double(), ct.c_double(), ct.c_double()
    check(load().rwnmr_pfgse_sample_stats(_ptr(Ms), S, nk, _ptr(r), int(n_points), _ptr(out["mean_M"]),
                                          _ptr(out["std_M"]), _ptr(out["mean_lhs"]), _ptr(out["std_lhs"]), _ptr(Ds),
                                          ct.byref(dm), ct.byref(dsd), ct.byref(de)))
    out.update(D_samples=Ds, D_mean=dm.value, D_std=dsd.value, D_err_mean=de.value)
    return out


def collision_histogram(collisions, steps, *, size, log_scale=False):
    """(bins, amps) of walker collision rates (rwnmr_collision_histogram)."""
    c = np.ascontiguousarray(collisions, np.uint32)
    nb = load().rwnmr_histogram_bins(int(size), int(log_scale))
    bins, amps = np.zeros(nb), np.zeros(nb)
    check(load().rwnmr_collision_histogram(_ptr(c), len(c), int(steps), int(size), int(log_scale), _ptr(bins),
                                           _ptr(amps)))
    return bins, amps


def histogram_list(echoes, n_histograms):
    """[(first_echo, last_echo)] of the histogram segments."""
    f = np.zeros(n_histograms, np.uint32)
    l = np.zeros(n_histograms, np.uint32)
    n = ct.c_int32()
    check(load().rwnmr_histogram_list(int(echoes), int(n_histograms), ct.byref(n), _ptr(f), _ptr(l)))
    return [(int(f[i]), int(l[i])) for i in range(n.value)]


def histogram_penalties(bins, *, rho, step_length, D0, steps_per_echo, sigmoid=None):
    b = np.ascontiguousarray(bins, np.float64)
    out = np.zeros(len(b))
    sg = None if sigmoid is None else np.ascontiguousarray(sigmoid, np.float64)
    check(load().rwnmr_histogram_penalties(_ptr(b), len(b), float(rho), _ptr(sg), float(step_length), float(D0),
                                           int(steps_per_echo), _ptr(out)))
    return out


def histogram_decay(amps, segments, penalties):
    a = np.ascontiguousarray(amps, np.float64)
    nh, nb = a.shape
    f = np.array([s[0] for s in segments], np.uint32)
    l = np.array([s[1] for s in segments], np.uint32)
    p = np.ascontiguousarray(penalties, np.float64)
    sig = np.zeros(1 + int((l - f).sum()))
    check(load().rwnmr_histogram_decay(_ptr(a), nh, nb, _ptr(f), _ptr(l), _ptr(p), _ptr(sig)))
    return sig


def cpmg_histogram_simulate(ctx: Context, img: Image, x, y, z, seed, *, echoes, steps_per_echo, bc="mirror",
                            n_histograms=1, size=100, log_scale=False, rho=0.0, sigmoid=None, step_length=1.0,
                            D0=2.5):
    """Collision-histogram CPMG on the GPU map walk
    (rwnmr_cpmg_histogram_simulate).  Host walkers are copied and returned;
    device walkers are updated in place.  Returns dict(signal, bins, amps,
    x, y, z, seed)."""
    dev = _is_dev(x)
    n = len(x)
    if dev:
        _torch_sync()
        w = _walker_struct(x, y, z, seed, None, None, None, 1.0, n, 1)
    else:
        x, y, z = _host(x, np.int32), _host(y, np.int32), _host(z, np.int32)
        seed = _host(seed, np.uint64)
        w = _walker_struct(x, y, z, seed, None, None, None, 1.0, n, 0)
    nb = load().rwnmr_histogram_bins(int(size), int(log_scale))
    sig = np.zeros(echoes + 1)
    bins = np.zeros(nb)
    amps = np.zeros((n_histograms, nb))
    used = ct.c_int32()
    sg = None if sigmoid is None else np.ascontiguousarray(sigmoid, np.float64)
    check(load().rwnmr_cpmg_histogram_simulate(ctx.handle, img.handle, ct.byref(w), int(echoes), int(steps_per_echo),
                                               bc_code(bc), 0, int(n_histograms), int(size), int(log_scale),
                                               float(rho), _ptr(sg), float(step_length), float(D0), _ptr(sig),
                                               _ptr(bins), _ptr(amps), ct.byref(used)), ctx.handle)
    return dict(signal=sig, bins=bins, amps=amps[:used.value], x=x, y=y, z=z, seed=seed)


def cpmg_batch_run(ctx: Context, instances, *, sequential=False, sync=True):
    """Many CPMG instances in one persistent launch (rwnmr_cpmg_batch_run).

    instances: list of dicts with keys img, x, y, z, seed (torch CUDA
    tensors, updated in place), echo_sums (CUDA float64, echoes+1), echoes,
    steps_per_echo, bc, and optionally energy, hits, penalty, delta.
    sequential=True runs one launch per instance (the comparison path)."""
    _torch_sync()
    arr = (_lib.CpmgInstance * len(instances))()
    for i, q in enumerate(instances):
        arr[i].image = q["img"].handle
        arr[i].walkers = _walker_struct(q["x"], q["y"], q["z"], q["seed"], q.get("energy"), q.get("hits"),
                                        q.get("penalty"), q.get("delta", 1.0), len(q["x"]), 1)
        arr[i].echoes = int(q["echoes"])
        arr[i].steps_per_echo = int(q["steps_per_echo"])
        arr[i].bc = bc_code(q.get("bc", "mirror"))
        arr[i].echo_sums = _ptr(q["echo_sums"])
    check(load().rwnmr_cpmg_batch_run(ctx.handle, arr, len(instances), 1 if sequential else 0), ctx.handle)
    if sync:
        ctx.synchronize()
