// walk_common.h -- step arithmetic shared by the HIP kernels and the host
// (the same inline functions compile for gfx950 and for x86 so the host can
// check them exhaustively on small domains).
//
// Semantics restated from RandomWalk_NMR (paths under src/backend/):
//   xorshift64 13/7/17            B/RNG/xorshift.cpp:17-24
//   direction = state % 6 + 1     B/Walker/walker.h:346-358, 566-582 (the
//                                 reference's digit-sum mod6 equals % 6)
//   unit moves                    B/Walker/walker.h:360-391, walker_defs.h:6-23
//   noflux flip                   B/NMR_Simulation/NMR_cpmg_cuda.cpp:1847-1902
//   periodic / mirror image fold  B/NMR_Simulation/NMR_cpmg_cuda.cpp:1991-2015
//   bitBlock addressing           B/BitBlock/bitBlock.h:113-139
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define RW_HD __host__ __device__ __forceinline__
#else
#define RW_HD static inline
#endif

namespace rwnmr {

enum Boundary : int { kNoflux = 0, kPeriodic = 1, kMirror = 2 };

RW_HD uint64_t xorshift64(uint64_t s)
{
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}

// s % 6 from 32-bit operations only.  s mod 2 is the low bit.  s mod 3 uses
// 2^k == (-1)^k (mod 3): with p = popcount(s) and e = popcount of the even
// bits, s == e + 2 (p - e) = 2p - e (mod 3), a number in [0, 128]; its
// residue comes from a 24-bit multiply (q / 3 == (q * 171) >> 9 for q < 256).
// The two residues combine by CRT: r6 = r3 + 3 * ((r3 ^ r2) & 1).
RW_HD uint32_t mod6(uint64_t s)
{
    const uint32_t lo = (uint32_t)s;
    const uint32_t hi = (uint32_t)(s >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t p = __builtin_popcount(lo) + __builtin_popcount(hi);
    const uint32_t e = __builtin_popcount(lo & 0x55555555u) + __builtin_popcount(hi & 0x55555555u);
#else
    const uint32_t p = (uint32_t)__builtin_popcount(lo) + (uint32_t)__builtin_popcount(hi);
    const uint32_t e = (uint32_t)__builtin_popcount(lo & 0x55555555u) +
                       (uint32_t)__builtin_popcount(hi & 0x55555555u);
#endif
    const uint32_t q = 2u * p - e;
    const uint32_t r3 = q - 3u * ((q * 171u) >> 9);
    const uint32_t r2 = lo & 1u;
    return r3 + 3u * ((r3 ^ r2) & 1u);
}

// Residue r = state % 6 -> move.  r: 0 North (y-1), 1 West (x-1),
// 2 South (y+1), 3 East (x+1), 4 Up (z+1), 5 Down (z-1).
// axis: 0 = x, 1 = y, 2 = z.  sign: +1 for r in {2,3,4}, else -1.
RW_HD int move_axis(uint32_t r) { return r >= 4u ? 2 : (int)((r & 1u) ^ 1u); }
RW_HD int move_sign(uint32_t r) { return ((0x1Cu >> r) & 1u) ? 1 : -1; }

// Mirror fold of an image-voxel coordinate g (C truncating division as in
// the reference formula).
RW_HD int fold_mirror(int g, int d)
{
    int m = g % d;
    if (m < 0)
        m += d;
    const int reflected = (g > 0) ? (g / d) % 2 : ((-g - 1 + d) / d) % 2;
    return reflected ? d - 1 - m : m;
}

RW_HD int fold_periodic(int g, int d)
{
    int m = g % d;
    return m < 0 ? m + d : m;
}

// Reflection flag of the mirror tiling at g (1 when fold_mirror(g) runs
// backwards as g increases).
RW_HD int mirror_flag(int g, int d)
{
    return (g > 0) ? (g / d) % 2 : ((-g - 1 + d) / d) % 2;
}

// One unit move of a folded coordinate i with reflection flag m inside a
// mirror tiling of period 2d: the image coordinate moves by +-1, or stays at
// the face while the flag flips.  Equals fold_mirror(g + sign) for
// i == fold_mirror(g), m == mirror_flag(g).
RW_HD void mirror_step(int i, int m, int sign, int d, int &ni, int &nm)
{
    const int t = m ? -sign : sign;
    const int c = i + t;
    const int out = (c < 0) | (c >= d);
    ni = out ? i : c;
    nm = m ^ out;
}

RW_HD int periodic_step(int i, int sign, int d)
{
    const int c = i + sign;
    return c < 0 ? d - 1 : (c >= d ? 0 : c);
}

RW_HD uint32_t block_index(int ix, int iy, int iz, int bC, int bCR)
{
    return (uint32_t)(ix >> 2) + (uint32_t)(iy >> 2) * (uint32_t)bC + (uint32_t)(iz >> 2) * (uint32_t)bCR;
}

RW_HD uint32_t bit_index(int ix, int iy, int iz)
{
    return (uint32_t)((ix & 3) | ((iy & 3) << 2) | ((iz & 3) << 4));
}

// Device layout of the bitBlock words.  Blocks are grouped in bricks of
// 2^sx x 2^sy x 2^sz blocks stored contiguously (bricks in x-fastest order
// over a grid of nbx x nby x nbz bricks), so one cache line holds a compact
// 3-D neighbourhood instead of a thin row along x.  sx = sy = sz = 0 is the
// reference layout (block index bx + by*bC + bz*bC*bR).
struct BrickLayout {
    int sx, sy, sz;      // log2 of the brick extent in blocks per axis
    int nbx, nbxy;       // bricks along x, bricks per z-layer
};

RW_HD uint32_t brick_block_index(int ix, int iy, int iz, const BrickLayout &L)
{
    const uint32_t bx = (uint32_t)ix >> 2, by = (uint32_t)iy >> 2, bz = (uint32_t)iz >> 2;
    const uint32_t brick = (bx >> L.sx) + (by >> L.sy) * (uint32_t)L.nbx + (bz >> L.sz) * (uint32_t)L.nbxy;
    const uint32_t inner = (bx & ((1u << L.sx) - 1u)) | ((by & ((1u << L.sy) - 1u)) << L.sx) |
                           ((bz & ((1u << L.sz) - 1u)) << (L.sx + L.sy));
    return (brick << (L.sx + L.sy + L.sz)) | inner;
}

// 30-bit Morton (Z-order) code of block coordinates (10 bits per axis),
// used to order walkers by position.
RW_HD uint32_t spread3(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

RW_HD uint32_t morton3(uint32_t bx, uint32_t by, uint32_t bz)
{
    return spread3(bx) | (spread3(by) << 1) | (spread3(bz) << 2);
}

}  // namespace rwnmr
