"""The reference's output tree (include/rwnmr/rwnmr_output.h): SimInfo.txt,
NMR_binImage.txt and the CPMG / multitau / PFGSE result files, written by
the library in the reference's exact formats, plus the path of the
command-line program built next to the library (``rwnmr``).
"""
from __future__ import annotations

import os

import numpy as np

from ._lib import HERE, check, load
from .core import _ptr

APP_PATH = os.path.join(HERE, "rwnmr")


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _s(p):
    return os.fspath(p).encode()


def make_dirs(path):
    check(load().rwnmr_make_dirs(_s(path)))


def write_sim_info(path, *, name, data_path, image_path, voxel_resolution, D0, n_images, occupancy, seed, gpu=True,
                   bc="noflux"):
    check(load().rwnmr_write_sim_info(_s(path), name.encode(), _s(data_path), _s(image_path),
                                      float(voxel_resolution), float(D0), int(n_images), float(occupancy), int(seed),
                                      int(bool(gpu)), bc.encode()))


def write_bin_image(path, blocks, shape):
    """shape = (depth, rows, cols) of the image the bitBlock words describe."""
    b = np.ascontiguousarray(blocks, dtype=np.uint64)
    D, H, W = shape
    check(load().rwnmr_write_bin_image(_s(path), _ptr(b), W, H, D))


def write_cpmg_decay(path, times, amps, noise=None):
    t, a = _d(times), _d(amps)
    nz = None if noise is None else _d(noise)
    check(load().rwnmr_write_cpmg_decay(_s(path), len(t), _ptr(t), _ptr(a), None if nz is None else _ptr(nz)))


def write_two_columns(path, header, a, b):
    a, b = _d(a), _d(b)
    check(load().rwnmr_write_two_columns(_s(path), header.encode(), len(a), _ptr(a), _ptr(b)))


def write_histogram_list(path, bins, amps):
    """bins / amps: [n_hist][n_bins]."""
    b, a = _d(bins), _d(amps)
    check(load().rwnmr_write_histogram_list(_s(path), b.shape[0], b.shape[1], _ptr(b), _ptr(a)))


def write_pfgse_parameters(path, *, D0, pulse_width, gamma, gmax, gradients, times):
    t = _d(times)
    check(load().rwnmr_write_pfgse_parameters(_s(path), float(D0), float(pulse_width), float(gamma), float(gmax[0]),
                                              float(gmax[1]), float(gmax[2]), int(gradients), _ptr(t), len(t)))


def write_pfgse_gradient(path, g, k):
    """g, k: [n][3] gradient and wave vectors."""
    g, k = _d(g), _d(k)
    cols = [np.ascontiguousarray(c) for c in (g[:, 0], g[:, 1], g[:, 2], k[:, 0], k[:, 1], k[:, 2])]
    check(load().rwnmr_write_pfgse_gradient(_s(path), len(g), *[_ptr(c) for c in cols]))


def write_pfgse_echoes(path, *, gradient, Mkt, noise, Mkt_std, lhs, lhs_std, rhs):
    cols = [_d(c) for c in (gradient, Mkt, noise, Mkt_std, lhs, lhs_std, rhs)]
    check(load().rwnmr_write_pfgse_echoes(_s(path), len(cols[0]), *[_ptr(c) for c in cols]))


def write_pfgse_msd(path, msd, msd_std, dmsd, dmsd_std):
    cols = [_d(c) for c in (msd, msd_std, dmsd, dmsd_std)]
    check(load().rwnmr_write_pfgse_msd(_s(path), *[_ptr(c) for c in cols]))


def write_pfgse_results_header(path):
    check(load().rwnmr_write_pfgse_results_header(_s(path)))


def append_pfgse_results(path, time, row, points):
    """row: D_sat, error, std, D_msd, std, DmsdX, std, DmsdY, std, DmsdZ, std."""
    r = _d(row)
    if r.size != 11:
        raise ValueError("a PFGSE results row has 11 values")
    check(load().rwnmr_append_pfgse_results(_s(path), float(time), _ptr(r), int(points)))


def threshold_lhs_window(lhs, value, window):
    a = _d(lhs)
    return int(load().rwnmr_pfgse_threshold_lhs_window(_ptr(a), len(a), float(value), int(window)))
