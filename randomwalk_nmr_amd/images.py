"""Image ingestion and the reference's text outputs (include/rwnmr/rwnmr_io.h):
PNG slice stacks binarised as the reference does (any colour channel
nonzero = solid, B/NMR_Simulation/NMR_Simulation.cpp:536-617), the micro-CT
config file (B/ConfigFiles/uct_config.cpp), pore / interface statistics,
ImageInfo.txt and walkers.csv (NMR_Simulation.cpp:1644-1712)."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from ._lib import UctConfig, check, load


def _ptr(a):
    return ct.c_void_p(a.ctypes.data)


def read_png(path):
    """Binarised PNG slice: uint8 [height, width], 1 = solid."""
    L = load()
    w, h = ct.c_int32(), ct.c_int32()
    check(L.rwnmr_png_size(str(path).encode(), ct.byref(w), ct.byref(h)))
    out = np.zeros((h.value, w.value), np.uint8)
    check(L.rwnmr_png_binarize(str(path).encode(), w.value, h.value, _ptr(out)))
    return out


class UctImage:
    """A uct.config: defaults of config/default/uct.config, then the keys of
    `path` (and of `more` files, in order) applied on top."""

    def __init__(self, path=None, *more):
        self.cfg = UctConfig()
        load().rwnmr_uct_config_defaults(ct.byref(self.cfg))
        for p in ([path] if path else []) + list(more):
            check(load().rwnmr_uct_config_read(str(p).encode(), ct.byref(self.cfg)))

    def __getattr__(self, name):
        cfg = self.__dict__["cfg"]
        v = getattr(cfg, name)
        return v.decode() if isinstance(v, bytes) else v

    def set(self, **kw):
        for k, v in kw.items():
            setattr(self.cfg, k, v.encode() if isinstance(v, str) else v)
        return self

    def slice_path(self, i):
        buf = ct.create_string_buffer(4096)
        check(load().rwnmr_uct_slice_path(ct.byref(self.cfg), int(i), buf, 4096))
        return buf.value.decode()

    def load(self):
        """Binary volume uint8 [depth, height, width], 1 = solid."""
        L = load()
        w, h, d = ct.c_int32(), ct.c_int32(), ct.c_int32()
        check(L.rwnmr_uct_dims(ct.byref(self.cfg), ct.byref(w), ct.byref(h), ct.byref(d)))
        vol = np.zeros((d.value, h.value, w.value), np.uint8)
        check(L.rwnmr_uct_load(ct.byref(self.cfg), w.value, h.value, d.value, _ptr(vol)))
        return vol


def image_stats(solid, voxel=1.0):
    """dict(pores, porosity, interfaces, svp) of a binary volume [D, H, W]."""
    solid = np.ascontiguousarray(solid, np.uint8)
    D, H, W = solid.shape
    pores, ifc = ct.c_uint64(), ct.c_uint64()
    por, svp = ct.c_double(), ct.c_double()
    check(load().rwnmr_image_stats(_ptr(solid), W, H, D, float(voxel), ct.byref(pores), ct.byref(por),
                                   ct.byref(ifc), ct.byref(svp)))
    return dict(pores=pores.value, porosity=por.value, interfaces=ifc.value, svp=svp.value)


def write_image_info(path, *, image_path, shape, resolution, voxel_shift, step_length, porosity, svp):
    D, H, W = shape
    check(load().rwnmr_write_image_info(str(path).encode(), str(image_path).encode(), W, H, D, float(resolution),
                                        int(voxel_shift), float(step_length), float(porosity), float(svp)))


def write_walkers_csv(path, x0, y0, z0, x, y, z, collisions, xi_rate, energy, seed):
    arrs = [np.ascontiguousarray(a, t) for a, t in ((x0, np.int32), (y0, np.int32), (z0, np.int32),
                                                    (x, np.int32), (y, np.int32), (z, np.int32),
                                                    (collisions, np.uint32), (xi_rate, np.float64),
                                                    (energy, np.float64), (seed, np.uint64))]
    check(load().rwnmr_write_walkers_csv(str(path).encode(), len(arrs[0]), *[_ptr(a) for a in arrs]))


def internal_field_uniform(solid, *, resolution, gradient, direction):
    """Field of a uniform gradient (InternalField::fillDataArray): float64
    [D, H, W], coordinate * gradient * 1e-6 * resolution in pores, 0 in solid."""
    solid = np.ascontiguousarray(solid, np.uint8)
    D, H, W = solid.shape
    out = np.zeros((D, H, W))
    check(load().rwnmr_internal_field_uniform(_ptr(solid), W, H, D, float(resolution), float(gradient),
                                              int(direction), _ptr(out)))
    return out


def read_internal_field(path, shape):
    """Raw W*H*D doubles (InternalField::readDataFromFile) as [D, H, W]."""
    D, H, W = shape
    out = np.zeros((D, H, W))
    check(load().rwnmr_internal_field_read(str(path).encode(), W, H, D, _ptr(out)))
    return out
