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

// The state before one xorshift64 step: each y = x ^ (x << a) is undone by
// x = y ^ (y << a) ^ (y << 2a) ^ ... (and likewise for >>), in reverse order.
// Lets a walk that draws its next direction one step ahead return the state
// of the steps it actually took.
RW_HD uint64_t xorshift64_inverse(uint64_t s)
{
    s ^= (s << 17) ^ (s << 34) ^ (s << 51);
    s ^= (s >> 7) ^ (s >> 14) ^ (s >> 21) ^ (s >> 28) ^ (s >> 35) ^ (s >> 42) ^ (s >> 49) ^ (s >> 56) ^ (s >> 63);
    s ^= (s << 13) ^ (s << 26) ^ (s << 39) ^ (s << 52);
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

// ------------------------------------------------------------------------
// Unfolded image with dilated voxel addresses (periodic and mirror walks).
//
// Mirror boundaries reflect the image into a tiling of period 2d per axis,
// so fold_mirror(g) == fold of (g mod 2d) inside one 2W x 2H x 2D "unfolded"
// image (u < d -> u, else 2d - 1 - u); periodic boundaries have period d.
// Storing that unfolded image once turns both boundaries into one periodic
// walk with no per-step fold.  Voxel (ux, uy, uz) of the unfolded image has
// the bit address A = dep(ux, Mx) | dep(uy, My) | dep(uz, Mz), where dep
// deposits the coordinate's bits into the positions of a mask and the three
// masks interleave round-robin from bit 0 (x0 y0 z0 x1 y1 z1 ...: one 32-bit
// word is 4 x 4 x 2 voxels, one 128-byte line 16 x 8 x 8, a Z-order curve
// above that).  Axes whose bits run out are skipped.  A unit move along one
// axis is then three 32-bit operations on A (a dilated add):
//     +1: A' = (A & ~M) | (((A | ~M) + lowbit(M)) & M)
//     -1: A' = (A & ~M) | (((A | ~M) + M) & M)
// both wrapping at 2^k.  A period that is not a power of two needs one fix
// after the add (field == dep(P) -> 0 going up; field == all ones ->
// dep(P - 1) going down).
// Bit order: each aligned run of 64 addresses is one 64-bit value (word
// 2i the low half, 2i + 1 the high half) holding address a at bit
// 63 - (a & 63), so a left shift by a & 63 (the hardware count of a 64-bit
// shift) brings the voxel's bit to the sign bit: the wall test is a shift
// and a sign compare.

struct ZLayout {
    uint32_t m[3];       // address bits of x, y, z
    uint32_t period[3];  // W, H, D (periodic) or 2W, 2H, 2D (mirror)
    uint32_t bits;       // address bits in use (the volume holds 2^bits voxels)
    uint32_t wrapfix;    // 1 when some period is not a power of two
};

RW_HD uint32_t zceil_log2(uint32_t v)
{
    uint32_t k = 0;
    while (k < 32 && (1ull << k) < (uint64_t)v)
        ++k;
    return k;
}

// Layout for field periods px, py, pz (each >= 1).  Returns 0 when the
// address would need more than 32 bits.
RW_HD int zlayout_make(uint32_t px, uint32_t py, uint32_t pz, ZLayout &L)
{
    const uint32_t k[3] = {zceil_log2(px), zceil_log2(py), zceil_log2(pz)};
    const uint32_t total = k[0] + k[1] + k[2];
    if (total > 32)
        return 0;
    uint32_t used[3] = {0, 0, 0};
    L.m[0] = L.m[1] = L.m[2] = 0;
    L.period[0] = px;
    L.period[1] = py;
    L.period[2] = pz;
    uint32_t pos = 0;
    while (pos < total)
        for (int a = 0; a < 3 && pos < total; ++a)
            if (used[a] < k[a]) {
                L.m[a] |= 1u << pos;
                ++pos;
                ++used[a];
            }
    L.bits = total;
    L.wrapfix = 0;
    for (int a = 0; a < 3; ++a)
        if ((1ull << k[a]) != (uint64_t)L.period[a])
            L.wrapfix = 1;
    return 1;
}

// Deposit the low bits of v into the set bits of m (lowest first).
RW_HD uint32_t zdeposit(uint32_t v, uint32_t m)
{
    uint32_t r = 0;
    for (uint32_t b = 1; m; m &= m - 1u, b <<= 1)
        if (v & b)
            r |= m & (~m + 1u);
    return r;
}

// Inverse of zdeposit: gather the bits of a at the set bits of m.
RW_HD uint32_t zextract(uint32_t a, uint32_t m)
{
    uint32_t r = 0;
    for (uint32_t b = 1; m; m &= m - 1u, b <<= 1)
        if (a & m & (~m + 1u))
            r |= b;
    return r;
}

RW_HD uint32_t floor_mod(int32_t g, uint32_t p)
{
    const int64_t m = (int64_t)g % (int64_t)p;
    return (uint32_t)(m < 0 ? m + (int64_t)p : m);
}

// Image voxel of unfolded coordinate u on an axis of image extent d.
RW_HD int zunfold(uint32_t u, int d, int mirror)
{
    return (mirror && u >= (uint32_t)d) ? 2 * d - 1 - (int)u : (int)u;
}

// One direction of the step table, indexed by idx = 2 * (s % 3) + (s & 1)
// (a bijection onto the residue s % 6 by the Chinese remainder theorem).
struct ZDir {
    uint32_t nm;      // ~mask of the moving axis
    uint32_t add;     // lowbit(mask) going up, mask going down
    int32_t dp;       // packed unbounded displacement: +-1 (x), +-2^10 (y), +-2^20 (z)
    uint32_t m;       // mask of the moving axis
    uint32_t wrapv;   // field value that needs the wrap fix (never matches if none)
    uint32_t fixv;    // replacement field value
};

// Noflux move on the periodic unfolded image (period = image size, every
// size >= 2): the forward move (m, add, dp) crosses the face when
// ((c & mp) | (A & mm)) == tgt (c: the candidate, A: the current address;
// going up the field wraps to 0 or, for a period that is not a power of
// two, reaches the period; going down the walker is at 0); the reference
// then turns around (B/NMR_Simulation/NMR_cpmg_cuda.cpp:1847-1902): addend
// add2, displacement dp2.  Walkers never leave the image, so no wrap
// fix-up is needed.
struct ZFlip {
    uint32_t m, add;
    int32_t dp;
    uint32_t tgt, add2;
    int32_t dp2;
    uint32_t mp, mm;
};

// s % 6 residue for table index idx (inverse of the CRT pairing).
RW_HD uint32_t zdir_residue(uint32_t idx)
{
    return (0x521430u >> (4u * idx)) & 0xFu;   // idx 0..5 -> 0, 3, 4, 1, 2, 5
}

// Direction key of a 64-bit state s with bytes b0 .. b7 (b0 lowest): 256 = 4
// (mod 6) and 4^j = 4 (mod 6) for j >= 1, so s = b0 + 4 (b1 + ... + b7) (mod
// 6).  The kernels form key = b0 + 4 (b1 + ... + b7) + acc with two byte dot
// products (v_dot4_u32_u8) of the state's words with these weights; key - acc
// <= kZKeyMax, and the table index of a key k is zkey_index(k).
constexpr uint32_t kZKeyWeightsLo = 0x04040401u, kZKeyWeightsHi = 0x04040404u;
constexpr uint32_t kZKeyMax = 255u * 29u;   // b0 + 4 x 7 bytes

// Byte dot product of v_dot4_u32_u8 (the host restatement the tests use).
RW_HD uint32_t zdot4(uint32_t a, uint32_t w, uint32_t acc)
{
    for (int i = 0; i < 4; ++i)
        acc += ((a >> (8 * i)) & 0xFFu) * ((w >> (8 * i)) & 0xFFu);
    return acc;
}

RW_HD uint32_t zkey_index(uint32_t k) { return 2u * (k % 3u) + (k & 1u); }

RW_HD ZFlip zdir_noflux(const ZLayout &L, uint32_t idx)
{
    const uint32_t r = zdir_residue(idx);
    const int ax = move_axis(r), sg = move_sign(r);
    const uint32_t m = L.m[ax], low = m & (~m + 1u);
    const int32_t unit = ax == 0 ? 1 : (ax == 1 ? 1 << 10 : 1 << 20);
    const uint32_t k = (uint32_t)__builtin_popcount(m);
    const uint32_t p = L.period[ax];
    ZFlip f;
    f.m = m;
    f.add = sg > 0 ? low : m;
    f.dp = sg * unit;
    f.add2 = sg > 0 ? m : low;
    f.dp2 = -sg * unit;
    f.mp = sg > 0 ? m : 0u;
    f.mm = sg > 0 ? 0u : m;
    f.tgt = (sg > 0 && (1ull << k) != (uint64_t)p) ? zdeposit(p, m) : 0u;
    return f;
}

// flags: bit a reverses the address move along axis a (the mirror image of
// the mirror tiling, u -> 2d - 1 - u, is the tiling itself, so a mirror walker
// may sit at the reflected address with its moves reversed along that axis);
// dp keeps the drawn direction's unbounded displacement.
RW_HD ZDir zdir_make(const ZLayout &L, uint32_t idx, uint32_t flags = 0u)
{
    const uint32_t r = zdir_residue(idx);
    const int ax = move_axis(r), sg0 = move_sign(r);
    const int sg = ((flags >> ax) & 1u) ? -sg0 : sg0;
    const uint32_t m = L.m[ax];
    ZDir d;
    d.m = m;
    d.nm = ~m;
    d.add = sg > 0 ? (m & (~m + 1u)) : m;
    d.dp = sg0 * (ax == 0 ? 1 : (ax == 1 ? 1 << 10 : 1 << 20));
    const uint32_t k = (uint32_t)__builtin_popcount(m);
    const uint32_t p = L.period[ax];
    if ((1ull << k) == (uint64_t)p) {
        d.wrapv = 0xFFFFFFFFu;   // no axis owns every bit, so this never matches
        d.fixv = 0;
    } else if (sg > 0) {
        d.wrapv = zdeposit(p, m);
        d.fixv = 0;
    } else {
        d.wrapv = m;
        d.fixv = zdeposit(p - 1u, m);
    }
    return d;
}

// Candidate address after one move.
RW_HD uint32_t zmove(uint32_t A, const ZDir &d, int wrapfix)
{
    uint32_t c = (A & d.nm) | (((A | d.nm) + d.add) & ~d.nm);
    if (wrapfix && (c & d.m) == d.wrapv)
        c = (c & d.nm) | d.fixv;
    return c;
}

// Byte sum of the 64-bit state (== state mod 3, since 256 == 1 mod 3) and
// the table index; sb <= 2040 and (sb * 2731) >> 13 == sb / 3 there.
RW_HD uint32_t zbyte_sum(uint64_t s)
{
    uint32_t t = 0;
    for (int i = 0; i < 8; ++i)
        t += (uint32_t)(s >> (8 * i)) & 0xFFu;
    return t;
}

RW_HD uint32_t zdir_index_from_sum(uint32_t sb, uint32_t lo)
{
    const uint32_t q = (sb * 2731u) >> 13;
    return 2u * (sb - 3u * q) + (lo & 1u);
}

// Packed displacement -> three int32 (|dx|, |dy|, |dz| <= 511).
RW_HD void zunpack(int32_t p, int32_t &dx, int32_t &dy, int32_t &dz)
{
    dx = (int32_t)((uint32_t)p << 22) >> 22;
    const int32_t p1 = (p - dx) >> 10;
    dy = (int32_t)((uint32_t)p1 << 22) >> 22;
    dz = (p1 - dy) >> 10;
}

}  // namespace rwnmr
