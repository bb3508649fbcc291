"""Convert the reference's three tiny PNG stacks (input/images/ of
RandomWalk_NMR) into voxel fixtures: tests/golden/images.npz holds, per
stack, the binarised volume [z, y, x] (any nonzero channel = solid, as
B/NMR_Simulation/NMR_Simulation.cpp:577-617 binarises) and its bitBlock words
packed by an independent numpy packer written here from the layout of
B/BitBlock/bitBlock.h:113-134.

Run in a checkout where /root/reference exists; GPU test runs only read the
committed .npz.
"""
import os

import numpy as np
from PIL import Image

REF = "/root/reference/input/images"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "images.npz")
STACKS = {
    "tiny_3D": ("tiny_3D", "tiny_3D_", 10),
    "sphere_r2_5": ("sphere_2.5voxels", "sphere_r=2.5voxels_", 7),
    "antisphere_r2_5": ("antisphere_r=2.5voxels", "antisphere_r=2.5voxels_", 5),
}


def numpy_pack(solid):
    D, H, W = solid.shape
    bC, bR, bD = (W + 3) // 4, (H + 3) // 4, (D + 3) // 4
    out = np.zeros(bC * bR * bD, np.uint64)
    z, y, x = np.nonzero(solid)
    blk = (x // 4) + (y // 4) * bC + (z // 4) * bC * bR
    bit = (x % 4) + 4 * (y % 4) + 16 * (z % 4)
    np.bitwise_or.at(out, blk, np.left_shift(np.uint64(1), bit.astype(np.uint64)))
    return out


def main():
    arrays = {}
    for key, (folder, prefix, slices) in STACKS.items():
        vol = []
        for i in range(slices):
            img = np.asarray(Image.open(os.path.join(REF, folder, f"{prefix}{i}.png")).convert("RGB"))
            vol.append((img != 0).any(axis=2).astype(np.uint8))
        solid = np.stack(vol)
        arrays[f"{key}_solid"] = solid
        arrays[f"{key}_blocks"] = numpy_pack(solid)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
