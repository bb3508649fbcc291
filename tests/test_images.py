"""Image ingestion and text outputs (CPU): PNG decoding and binarisation,
uct.config reading, slice naming, pore / interface statistics, ImageInfo.txt
and walkers.csv formats.

Pins: the reference's three sample PNG stacks (tests/golden/png, data copied
from RandomWalk_NMR input/images) against their binarised volumes decoded by
PIL (tests/golden/images.npz); a PNG encoder written here from the PNG
specification (every row filter, colour type, bit depth and Adam7) against
the binarisation rule applied to the raw samples; statistics against numpy
with periodic neighbours.
"""
import os
import struct
import zlib

import numpy as np
import pytest

import randomwalk_nmr_amd as R
from randomwalk_nmr_amd import images as I

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
STACKS = {
    "tiny_3D": ("tiny_3D", "tiny_3D_", 10),
    "sphere_r2_5": ("sphere_2.5voxels", "sphere_r=2.5voxels_", 7),
    "antisphere_r2_5": ("antisphere_r=2.5voxels", "antisphere_r=2.5voxels_", 5),
}


@pytest.mark.parametrize("key", sorted(STACKS))
def test_reference_png_stacks_via_uct_config(tmp_path, key):
    folder, prefix, slices = STACKS[key]
    cfg = tmp_path / "uct.config"
    cfg.write_text("--- MICRO-CT IMAGE PARAMETERS\n"
                   "-- 'IMG_FILE_LIST' == alternative mode: absolute path to a file\n"
                   f"DIR_PATH: {os.path.join(GOLD, 'png', folder)}/\n"
                   f"FILENAME: {prefix}\nFIRST_IDX: 0\nDIGITS: 1\nEXTENSION: .png\nSLICES: {slices}\n"
                   "RESOLUTION: 2.5\nVOXEL_DIVISION: 1\n")
    u = I.UctImage(str(cfg))
    assert (u.slices, u.resolution, u.voxel_division, u.img_files_list) == (slices, 2.5, 1, "Empty")
    vol = u.load()
    gold = np.load(os.path.join(GOLD, "images.npz"))
    assert np.array_equal(vol, gold[f"{key}_solid"])
    assert np.array_equal(R.pack_bitblocks(vol), gold[f"{key}_blocks"])
    # the same stack through an explicit file list
    lst = tmp_path / "list.txt"
    lst.write_text("".join(u.slice_path(i) + "\n" for i in range(slices)))
    u2 = I.UctImage(str(cfg)).set(img_files_list=str(lst))
    assert np.array_equal(u2.load(), vol)


def test_uct_config_parsing_and_slice_names(tmp_path):
    cfg = tmp_path / "u.config"
    cfg.write_text("DIR_PATH: /data/rock/\nFILENAME: slice_\nFIRST_IDX: 7\nDIGITS: 4\nEXTENSION: .png\n"
                   "SLICES: 5\nRESOLUTION: 0.75\nVOXEL_DIVISION: 2\nUNKNOWN_KEY: 3\n")
    u = I.UctImage(str(cfg))
    assert u.dir_path == "/data/rock/" and u.filename == "slice_" and u.first_idx == 7 and u.digits == 4
    assert (u.slices, u.resolution, u.voxel_division) == (5, 0.75, 2)
    assert [u.slice_path(i) for i in (0, 4)] == ["/data/rock/slice_0007.png", "/data/rock/slice_0011.png"]
    u.set(digits=1, first_idx=98)
    assert u.slice_path(3) == "/data/rock/slice_101.png"   # setw pads, never truncates
    with pytest.raises(R.RwnmrError):
        u.slice_path(5)
    bad = tmp_path / "bad.config"
    bad.write_text("SLICES: many\n")
    with pytest.raises(R.RwnmrError):
        I.UctImage(str(bad))


# ---------------------------------------------------------------- PNG writer

def _chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def _filter_rows(rows, bpp, rng):
    """Apply a random PNG filter (0-4) to each row of raw bytes."""
    out, prev = [], bytes(len(rows[0]))
    for row in rows:
        ft = int(rng.integers(0, 5))
        f = bytearray(len(row))
        for i in range(len(row)):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ft]
            f[i] = (row[i] - pred) & 0xFF
        out.append(bytes([ft]) + bytes(f))
        prev = row
    return b"".join(out)


def _pack_samples(samples, depth):
    """Row of integer samples -> bytes (big-endian, MSB-first sub-byte)."""
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in samples)
    if depth == 8:
        return bytes(int(v) for v in samples)
    bits = "".join(format(int(v), f"0{depth}b") for v in samples)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


def _write_png(path, img, ctype, depth, rng, palette=None, interlace=False):
    """img: [H, W, channels] integer samples."""
    H, W, ch = img.shape
    bpp = max(1, ch * depth // 8)
    passes = ([(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
              if interlace else [(0, 0, 1, 1)])
    raw = b""
    for x0, y0, dx, dy in passes:
        sub = img[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        rows = [_pack_samples(sub[y].reshape(-1), depth) for y in range(sub.shape[0])]
        raw += _filter_rows(rows, bpp, rng)
    data = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, depth, ctype, 0, 0, int(interlace)))
    if palette is not None:
        data += _chunk(b"PLTE", bytes(palette.reshape(-1).tolist()))
    comp = zlib.compress(raw)
    data += _chunk(b"IDAT", comp[: len(comp) // 2]) + _chunk(b"IDAT", comp[len(comp) // 2:])
    data += _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


CASES = [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 4), (3, 8), (4, 8), (4, 16),
         (6, 8), (6, 16)]


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("ctype,depth", CASES)
def test_png_decoder_every_format(tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth + 7 * interlace)
    H, W = 37, 29
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    top = (1 << depth) - 1
    img = rng.integers(0, top + 1, (H, W, ch))
    img[rng.random((H, W)) < 0.4] = 0            # many all-zero pixels
    palette = None
    if ctype == 3:
        n = 1 << depth
        palette = rng.integers(0, 256, (n, 3))
        palette[0] = 0                            # index 0 is black
        palette[1 % n] = [0, 0, 7]
        img = rng.integers(0, n, (H, W, 1))
    if depth == 16:
        img[rng.random((H, W)) < 0.2] = rng.integers(1, 256)   # nonzero low byte only: pore after 8-bit reduction
    path = str(tmp_path / "t.png")
    _write_png(path, img, ctype, depth, rng, palette, interlace)
    # expected: imread colour rule
    if ctype == 3:
        colour = palette[img[..., 0]]
    elif ctype in (0, 4):
        colour = np.repeat(img[..., :1], 3, axis=2)
    else:
        colour = img[..., :3]
    if depth == 16:
        colour = colour >> 8
    want = (colour != 0).any(axis=2).astype(np.uint8)
    assert np.array_equal(I.read_png(path), want)


def test_png_rejects_garbage(tmp_path):
    p = tmp_path / "x.png"
    p.write_bytes(b"not a png")
    with pytest.raises(R.RwnmrError):
        I.read_png(str(p))


@pytest.mark.parametrize("shape", [(1, 7, 9), (6, 5, 8), (9, 9, 9)])
def test_image_stats_match_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    solid = (rng.random(shape) < 0.35).astype(np.uint8)
    got = I.image_stats(solid, voxel=0.5)
    pores = int((solid == 0).sum())
    ifc = int((solid != np.roll(solid, -1, axis=2)).sum() + (solid != np.roll(solid, -1, axis=1)).sum())
    if shape[0] > 1:
        ifc += int((solid != np.roll(solid, -1, axis=0)).sum())
    assert got["pores"] == pores and got["interfaces"] == ifc
    assert got["porosity"] == pores / solid.size
    assert abs(got["svp"] - ifc * 0.25 / (pores * 0.125)) <= 1e-12 * got["svp"]


def test_image_info_and_walkers_csv_formats(tmp_path):
    p = tmp_path / "ImageInfo.txt"
    I.write_image_info(p, image_path="/img/rock_.png", shape=(5, 6, 7), resolution=1.25, voxel_shift=1,
                       step_length=0.625, porosity=0.2345678912, svp=1.0 / 3.0)
    assert p.read_text() == ("image path: /img/rock_.png\nwidth(x): 7\nheight(y): 6\ndepth(z): 5\n"
                             "resolution: 1.25\nvoxel shift: 1\nstep length: 0.625\nporosity: 0.234568\n"
                             "SVp: 0.333333\n")
    c = tmp_path / "walkers.csv"
    I.write_walkers_csv(c, [1, -2], [3, 4], [5, 6], [7, 8], [9, -10], [11, 12], [13, 0], [0.5, 1e-7],
                        [0.123456789, 1.0], np.array([2**63 + 5, 17], np.uint64))
    lines = c.read_text().splitlines()
    assert lines[0] == "PositionXi,PositionYi,PositionZi,PositionXf,PositionYf,PositionZf,Collisions,XIRate,Energy,RNGSeed"
    assert lines[1] == "1,3,5,7,9,11,13,0.5,0.123457,9223372036854775813"
    assert lines[2] == "-2,4,6,8,-10,12,0,1e-07,1,17"


def test_internal_field_uniform_and_file_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    solid = (rng.random((4, 5, 6)) < 0.3).astype(np.uint8)
    for direction, axis in ((0, 2), (1, 1), (2, 0)):
        f = I.internal_field_uniform(solid, resolution=2.0, gradient=30.0, direction=direction)
        coord = np.indices(solid.shape)[axis].astype(float)
        want = np.where(solid == 1, 0.0, coord * (30.0 * 1e-6 * 2.0))
        assert np.array_equal(f, want)
    p = tmp_path / "field.bin"
    f.tofile(p)
    assert np.array_equal(I.read_internal_field(p, solid.shape), f)
    with pytest.raises(R.RwnmrError):
        I.read_internal_field(p, (4, 5, 7))
