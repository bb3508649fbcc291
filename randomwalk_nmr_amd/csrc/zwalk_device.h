// zwalk_device.h -- the unfolded-image walk of one tile of walkers, shared
// by the single-instance kernel (zwalk_kernels.hip) and the batched
// persistent kernel.  Layout and move arithmetic: walk_common.h, "Unfolded
// image with dilated voxel addresses".
//
// Per step and walker, all in VGPRs: xorshift64; a key congruent to the
// state mod 6 from two byte dot products (zdot_key) indexes a byte table in
// LDS that points at the direction entry (axis mask, dilated addend, packed
// displacement) of s % 6; the dilated add gives the candidate voxel
// address with the periodic wrap included; then the wall bit: with the
// register chunk cache (CACHE != 0) each walker keeps the 16-byte chunk
// (8 x 4 x 4 voxels) of its last lookup and reads memory only when the
// candidate leaves it (about one step in five), else one 4-byte gather per
// step; then either the move (address and packed displacement) or a wall hit
// (hit count + 1, energy *= delta, in that order as the reference).  The
// unbounded coordinates the reference reports are rebuilt from the packed
// displacement every kZFlush steps.  Nothing in the loop folds, divides or
// branches on the boundary.
//
// Reference semantics (paths under src/backend/ of RandomWalk_NMR):
//   CPMG_walk_{periodic,mirror}   B/NMR_Simulation/NMR_cpmg_cuda.cpp:300-402, 574-676
//   map_3D_{periodic,mirror}      B/NMR_Simulation/NMR_Simulation_cuda.cpp:418-721
//   image fold                    B/NMR_Simulation/NMR_cpmg_cuda.cpp:1991-2015
//   step / direction              B/Walker/walker.h:346-391, 498-507
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_common.h"
#include "walk_device.h"
#include "zwalk_kernels.h"

namespace rwnmr {

__device__ __forceinline__ uint32_t zaddress(int32_t x, int32_t y, int32_t z, const ZLayout &zl)
{
    return zdeposit(floor_mod(x, zl.period[0]), zl.m[0]) | zdeposit(floor_mod(y, zl.period[1]), zl.m[1]) |
           zdeposit(floor_mod(z, zl.period[2]), zl.m[2]);
}

// Direction table of layout zl in LDS (threads 0..5 write; the caller syncs).
__device__ __forceinline__ void zbuild_table(const ZLayout &zl, uint4 *tab, uint2 *fix)
{
    if (threadIdx.x < 6) {
        const ZDir d = zdir_make(zl, threadIdx.x);
        tab[threadIdx.x] = make_uint4(d.m, d.add, (uint32_t)d.dp, d.nm);
        fix[threadIdx.x] = make_uint2(d.wrapv, d.fixv);
    }
}

// Mirror walks on the unfolded image start every walker in the first octant
// (its folded position; flag bit a set when its unbounded coordinate is in an
// odd mirror image along axis a) and read the direction table of its flags:
// 8 copies of the 6-entry table (and of the wrap fix-up table) at
// kZFlagStride-byte steps (fix-up: half that), so the walker's table offset
// is flags x kZFlagStride, constant over the launch.  Threads 0..47 write;
// the caller syncs.
constexpr uint32_t kZFlagStride = 128u;

__device__ __forceinline__ void zbuild_table_flags(const ZLayout &zl, uint4 *tab, uint2 *fix)
{
    if (threadIdx.x < 48) {
        const uint32_t flags = threadIdx.x / 6u, idx = threadIdx.x % 6u;
        const ZDir d = zdir_make(zl, idx, flags);
        tab[flags * (kZFlagStride / 16u) + idx] = make_uint4(d.m, d.add, (uint32_t)d.dp, d.nm);
        fix[flags * (kZFlagStride / 16u) + idx] = make_uint2(d.wrapv, d.fixv);
    }
}

// Wide direction table: one 16-byte entry per (state & 1, byte sum of the
// state), index (state & 1) * 2048 + sb, so the step reads its entry without
// reducing sb mod 3 (sb <= 2040; 64 KiB of LDS).  Periods that are powers of
// two only (no wrap fix-up table).  Copied from the 6-entry table `small`
// (built and synced before); all threads of the block write; the caller
// syncs.
constexpr uint32_t kZWideEntries = 4096;

__device__ __forceinline__ void zbuild_table_wide(const uint4 *small, uint4 *tab)
{
    for (uint32_t e = threadIdx.x; e < kZWideEntries; e += blockDim.x) {
        const uint32_t sb = e & 2047u;
        tab[e] = small[2u * (sb % 3u) + (e >> 11)];
    }
}

// Direction key of the state (walk_common.h: key = b0 + 4 (b1 + ... + b7) +
// acc, congruent to s + acc mod 6): 2 VALU, v_dot4_u32_u8 on the high word
// with weights 4 4 4 4, then on the low word with 4 4 4 1.
__device__ __forceinline__ uint32_t zdot_key(uint32_t lo, uint32_t hi, uint32_t acc)
{
    return __builtin_amdgcn_udot4(lo, kZKeyWeightsLo, __builtin_amdgcn_udot4(hi, kZKeyWeightsHi, acc, false),
                                  false);
}

// Byte-key table: for each key k <= kZKeyMax the byte offset
// zkey_index(k) x bytes of the direction entry of s % 6 = k % 6 in the
// 6-entry table (the entry index 2 (s % 3) + (s % 2) of zdir_make); 7424
// bytes of LDS.  All threads of the block write; the caller syncs.
constexpr uint32_t kZKeyEntries = 7424;   // kZKeyMax + 1 rounded up to 64 bytes
static_assert(kZKeyEntries > kZKeyMax && kZKeyEntries % 64u == 0u, "key table size");

// bytes: the entry size of the direction table the offsets point into (16,
// or 32 for the noflux table).
__device__ __forceinline__ void zbuild_key_table(uint8_t *keys, uint32_t bytes = 16u)
{
    for (uint32_t e = threadIdx.x; e < kZKeyEntries; e += blockDim.x)
        keys[e] = (uint8_t)(bytes * zkey_index(e));
}

// Flag-keyed byte tables (TAB 3; mirror walks with reflection flags): the
// walker's flags select one of 8 byte-key tables (key + flags x
// kZKeyEntries, 58 KiB of LDS), whose bytes point straight at the direction
// entry for those flags, so the step adds no per-walker table offset (the
// flags' table offset enters as the accumulator of the key).  Twelve entries serve
// every flag combination: reflecting an axis flips the dilated addend and
// the wrap fix-up of its two directions but not their packed displacement
// (the unbounded move), so entry 2 idx + f holds direction idx with its axis
// reflected when f is 1 (16 bytes each; wrap fix-up entries 8 bytes).
constexpr uint32_t kZFlagKeyEntries = 8u * kZKeyEntries;
constexpr uint32_t kZFlippedEntries = 12u;

// Threads 0..11 write; the caller syncs.
__device__ __forceinline__ void zbuild_table_flipped(const ZLayout &zl, uint4 *tab, uint2 *fix)
{
    if (threadIdx.x < kZFlippedEntries) {
        const uint32_t idx = threadIdx.x >> 1, f = threadIdx.x & 1u;
        const uint32_t ax = (uint32_t)move_axis(zdir_residue(idx));
        const ZDir d = zdir_make(zl, idx, f << ax);
        tab[threadIdx.x] = make_uint4(d.m, d.add, (uint32_t)d.dp, d.nm);
        fix[threadIdx.x] = make_uint2(d.wrapv, d.fixv);
    }
}

// The byte of entry e (flags e / kZKeyEntries, key k = e % kZKeyEntries) is
// the offset 16 (2 idx + flip) of its entry in zbuild_table_flipped, with
// idx = 2 (k % 3) + (k % 2) as in zbuild_key_table.  All threads of the
// block write 4 keys per word (a word never straddles two flag tables);
// the caller syncs.
__device__ __forceinline__ void zbuild_flag_key_table(uint32_t *keys)
{
    for (uint32_t w = threadIdx.x; w < kZFlagKeyEntries / 4u; w += blockDim.x) {
        uint32_t v = 0u;
        for (uint32_t b = 0; b < 4u; ++b) {
            const uint32_t e = 4u * w + b;
            const uint32_t flags = e / kZKeyEntries, key = e % kZKeyEntries;
            const uint32_t idx = zkey_index(key);
            const uint32_t ax = (uint32_t)move_axis(zdir_residue(idx));
            v |= (16u * (2u * idx + ((flags >> ax) & 1u))) << (8u * b);
        }
        keys[w] = v;
    }
}

// Noflux direction table of layout zl (periodic image): two 16-byte halves
// per direction, {m, add, dp, tgt} and {add2, dp2, mp, mm} (ZFlip).
// Threads 0..5 write; the caller syncs.
__device__ __forceinline__ void zbuild_table_noflux(const ZLayout &zl, uint4 *tab)
{
    if (threadIdx.x < 6) {
        const ZFlip f = zdir_noflux(zl, threadIdx.x);
        tab[2 * threadIdx.x] = make_uint4(f.m, f.add, (uint32_t)f.dp, f.tgt);
        tab[2 * threadIdx.x + 1] = make_uint4(f.add2, (uint32_t)f.dp2, f.mp, f.mm);
    }
}

// The direction entry of the next move: advances the RNG state s and reads
// the table entry (16 bytes per entry in tab_bytes, 8 in fix_bytes) of
// state % 6.  TAB 0: the 6-entry table, s % 3 by arithmetic; 1: tab_bytes is
// the wide table (zbuild_table_wide); 2: the 6-entry table reached through
// the byte-key table `keys` (zbuild_key_table); 3: the 12-entry table of
// zbuild_table_flipped reached through the flag-keyed tables `keys`
// (zbuild_flag_key_table), fo = flags x 4096.
struct ZEntry {
    uint32_t m, add, dp, fx, fy;      // axis mask, dilated addend, packed displacement, wrap fix-up
    uint32_t tgt, add2, dp2, mp, mm;  // noflux: face test and the turned-around move (ZFlip)
};

// fo: the walker's table offset (mirror walks with reflection flags), else 0;
// TAB 3: its key offset.
template <bool WRAPFIX, int TAB = 0, bool NOFLUX = false>
__device__ __forceinline__ void zdirection(uint64_t &s, const char *tab_bytes, const char *fix_bytes, ZEntry &e,
                                           const uint8_t *keys = nullptr, uint32_t fo = 0u)
{
    static_assert(!NOFLUX || (TAB == 2 && !WRAPFIX), "noflux: byte-key table, no wrap fix-up");
    static_assert(!(TAB == 1 && WRAPFIX), "the wide table has no wrap fix-up");
    s = xorshift64(s);
    const uint32_t lo = (uint32_t)s, hi = (uint32_t)(s >> 32);
    uint32_t off;
    if (TAB == 3) {
        // the flags' key table: its offset is the key's accumulator
        off = keys[zdot_key(lo, hi, fo)];
    } else if (TAB == 2) {
        off = keys[zdot_key(lo, hi, 0u)];
    } else if (TAB == 1) {
        // byte sum sb <= 2040 == state (mod 3)
        const uint32_t sb = __builtin_amdgcn_sad_u8(hi, 0u, __builtin_amdgcn_sad_u8(lo, 0u, 0u));
        off = (sb << 4) | ((lo << 15) & 0x8000u);
    } else {
        const uint32_t sb = __builtin_amdgcn_sad_u8(hi, 0u, __builtin_amdgcn_sad_u8(lo, 0u, 0u));
        // q = sb / 3 by a 24-bit multiply; entry 2 * (sb - 3q) + (state & 1)
        // at 16 bytes per entry
        const uint32_t q = __umul24(sb, 2731u) >> 13;
        off = (uint32_t)__mul24((int)q, -96) + ((sb << 5) | ((lo & 1u) << 4));
    }
    if (TAB != 3)
        off += fo;
    const uint4 t = *reinterpret_cast<const uint4 *>(tab_bytes + off);
    e.m = t.x;
    e.add = t.y;
    e.dp = t.z;
    if (NOFLUX) {
        const uint4 u = *reinterpret_cast<const uint4 *>(tab_bytes + off + 16u);
        e.tgt = t.w;
        e.add2 = u.x;
        e.dp2 = u.y;
        e.mp = u.z;
        e.mm = u.w;
    }
    if (WRAPFIX) {
        const uint2 f = *reinterpret_cast<const uint2 *>(fix_bytes + (off >> 1));
        e.fx = f.x;
        e.fy = f.y;
    }
}

// Candidate address of the move of entry e from A (dilated add, periodic
// wrap included; WRAPFIX: non-power-of-two period fix-up).
template <bool WRAPFIX>
__device__ __forceinline__ uint32_t zmove_entry(uint32_t A, const ZEntry &e)
{
    uint32_t c = (A & ~e.m) | (((A | ~e.m) + e.add) & e.m);
    if (WRAPFIX)
        c = (c & e.m) == e.fx ? ((c & ~e.m) | e.fy) : c;
    return c;
}

// Noflux candidate of entry e from A: the forward move, or the turned-around
// one when the forward move crosses a face of the image; dp receives the
// packed displacement of the move taken.
__device__ __forceinline__ uint32_t zmove_noflux(uint32_t A, const ZEntry &e, int32_t &dp)
{
    const uint32_t c = (A & ~e.m) | (((A | ~e.m) + e.add) & e.m);
    const bool crossed = ((c & e.mp) | (A & e.mm)) == e.tgt;
    const uint32_t cf = (A & ~e.m) | (((A | ~e.m) + e.add2) & e.m);
    dp = (int32_t)(crossed ? e.dp2 : e.dp);
    return crossed ? cf : c;
}

// Candidate address of one move from A: advances the RNG state s, reads the
// direction entry and returns the dilated-add result; dp receives the packed
// displacement.
template <bool WRAPFIX, bool WIDE = false>
__device__ __forceinline__ uint32_t zpropose(uint64_t &s, uint32_t A, const char *tab_bytes, const char *fix_bytes,
                                             int32_t &dp)
{
    static_assert(!(WIDE && WRAPFIX), "the wide table has no wrap fix-up");
    s = xorshift64(s);
    const uint32_t lo = (uint32_t)s, hi = (uint32_t)(s >> 32);
    // byte sum sb <= 2040 == state (mod 3)
    const uint32_t sb = __builtin_amdgcn_sad_u8(hi, 0u, __builtin_amdgcn_sad_u8(lo, 0u, 0u));
    uint32_t off;
    if (WIDE) {
        off = (sb << 4) | ((lo << 15) & 0x8000u);
    } else {
        const uint32_t q = __umul24(sb, 2731u) >> 13;
        off = (uint32_t)__mul24((int)q, -96) + ((sb << 5) | ((lo & 1u) << 4));
    }
    const uint4 t = *reinterpret_cast<const uint4 *>(tab_bytes + off);
    const uint32_t m = t.x;
    uint32_t c = (A & ~m) | (((A | ~m) + t.y) & m);
    if (WRAPFIX) {
        const uint2 f = *reinterpret_cast<const uint2 *>(fix_bytes + (off >> 1));
        c = (c & m) == f.x ? ((c & ~m) | f.y) : c;
    }
    dp = (int32_t)t.z;
    // keep the displacement in this one table read (the compiler would
    // otherwise re-read it inside the move branch)
    asm volatile("" : "+v"(dp));
    return c;
}

// Register chunk cache: make QV hold the 16-byte chunk of address c (read
// from words only when c left the chunk QB).  c < 2^bits and the volume holds
// at least 4 words, so the read is in bounds.
__device__ __forceinline__ void zchunk_fill(const char *words, uint32_t c, uint32_t &QB, uint4 &QV)
{
    const uint32_t cid = c >> 7;
    if (cid != QB) {
        QV = *reinterpret_cast<const uint4 *>(words + ((c >> 3) & ~15u));
        QB = cid;
    }
}

// The 32-bit word of QV that holds address c's bit (descending order in
// each 64-bit half: addresses 0-31 of the half in its high word).
__device__ __forceinline__ uint32_t zchunk_word(uint32_t c, const uint4 &QV)
{
    const bool up = (c & 64u) != 0u;
    const uint32_t lo = up ? QV.z : QV.x, hi = up ? QV.w : QV.y;
    return (c & 32u) ? lo : hi;
}

// Address c's bit from its word: bit 31 - (c & 31), brought to the sign bit.
__device__ __forceinline__ uint32_t zchunk_bit(uint32_t c, const uint4 &QV)
{
    return (zchunk_word(c, QV) << (c & 31u)) >> 31;
}

// Buffer descriptor over the unfolded image's 16-byte chunks (base address,
// stride 16, num_records = chunk count, built by the compiler's
// make_buffer_rsrc builtin) for the chunk reloads; ids at or past the chunk
// count read zeros instead of memory.  The last descriptor word (32-bit data
// format, no swizzle) is the GFX9 layout (gfx90a / gfx942 / gfx950): other
// families lay it out differently, so they are refused at compile time.
#if defined(__HIP_DEVICE_COMPILE__) && !defined(__GFX9__)
#error "the chunk buffer descriptor is laid out for GFX9 targets (gfx90a, gfx942, gfx950)"
#endif
constexpr int32_t kZChunkRsrcWord3 = 0x00020000;
typedef __amdgpu_buffer_rsrc_t zrsrc_t;
typedef uint32_t zv4u __attribute__((ext_vector_type(4)));
typedef uint32_t zv2u __attribute__((ext_vector_type(2)));

#if defined(RWNMR_RAW_BUFFER_LOAD)
// Raw buffer load (a documented clang builtin) at byte offset cid * 16: one
// shift more per reload than the indexed form below.
__device__ __forceinline__ zv4u zchunk_load(zrsrc_t rsrc, uint32_t cid)
{
    return __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int32_t)(cid << 4), 0, 0);
}
// 8-byte chunks (64 addresses) at byte offset cid * 8
__device__ __forceinline__ zv2u zchunk_load8(zrsrc_t rsrc, uint32_t cid)
{
    return __builtin_amdgcn_raw_buffer_load_b64(rsrc, (int32_t)(cid << 3), 0, 0);
}
#else
// Indexed (struct) buffer load: the chunk id itself is the index, so no
// address shift.  Clang has no builtin for the indexed form, so the LLVM
// intrinsic is bound by name; the build probes that this compiles and
// otherwise builds the raw form above (RWNMR_RAW_BUFFER_LOAD).
__device__ zv4u zbuffer_load_v4(zrsrc_t rsrc, int32_t vindex, int32_t voffset, int32_t soffset, int32_t aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v4i32");
__device__ __forceinline__ zv4u zchunk_load(zrsrc_t rsrc, uint32_t cid)
{
    return zbuffer_load_v4(rsrc, (int32_t)cid, 0, 0, 0);
}
// 8-byte chunks (64 addresses): a descriptor of stride 8 (zchunk_rsrc8)
__device__ zv2u zbuffer_load_v2(zrsrc_t rsrc, int32_t vindex, int32_t voffset, int32_t soffset, int32_t aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v2i32");
__device__ __forceinline__ zv2u zchunk_load8(zrsrc_t rsrc, uint32_t cid)
{
    return zbuffer_load_v2(rsrc, (int32_t)cid, 0, 0, 0);
}
#endif

__device__ __forceinline__ zrsrc_t zchunk_rsrc(const uint32_t *words, const ZLayout &zl)
{
    const int32_t chunks = (int32_t)(zl.bits > 7 ? (1u << (zl.bits - 7)) : 1u);   // 128 addresses each
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(words), (short)16, chunks, kZChunkRsrcWord3);
}

// The same over the image's 8-byte chunks (4 x 4 x 4 voxels, 64 addresses;
// stride 8, num_records = that chunk count).
__device__ __forceinline__ zrsrc_t zchunk_rsrc8(const uint32_t *words, const ZLayout &zl)
{
    const int32_t chunks = (int32_t)(zl.bits > 6 ? (1u << (zl.bits - 6)) : 1u);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(words), (short)8, chunks, kZChunkRsrcWord3);
}

// 8-byte chunk cache: QV.x / QV.y hold the chunk of address c (QB its id,
// c >> 6), read when c left the chunk held.
__device__ __forceinline__ void zchunk_fill_buf8(const zrsrc_t &rsrc, uint32_t c, uint32_t &QB, uint4 &QV)
{
    const uint32_t cid = c >> 6;
    if (cid != QB) {
        const zv2u v = zchunk_load8(rsrc, cid);
        QV.x = v.x;
        QV.y = v.y;
    }
    QB = cid;
}

__device__ __forceinline__ void zchunk_fill_buf(const zrsrc_t &rsrc, uint32_t c, uint32_t &QB, uint4 &QV)
{
    const uint32_t cid = c >> 7;
    if (cid != QB) {
        const zv4u v = zchunk_load(rsrc, cid);
        QV = make_uint4(v.x, v.y, v.z, v.w);
    }
    // unconditional (equal when no load was needed): the chunk id then lives
    // in the register it was computed in instead of being copied on reloads
    QB = cid;
}

template <int CW>
__device__ __forceinline__ uint32_t zchunk_wall(uint32_t c, const uint4 &QV)
{
    uint64_t half = ((uint64_t)QV.y << 32) | QV.x;
    if (CW == 128) {
        // the half that bit 6 of c selects, by two bitfield inserts under the
        // sign-extended bit (3 VALU; a compare and two selects take 4)
        // (as inline VALU: the compiler turns the mask form back into selects)
        const uint32_t sel = (uint32_t)__builtin_amdgcn_sbfe((int)c, 6u, 1u);
        uint32_t lo, hi;
        asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(lo) : "v"(sel), "v"(QV.z), "v"(QV.x));
        asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(hi) : "v"(sel), "v"(QV.w), "v"(QV.y));
        half = ((uint64_t)hi << 32) | lo;
    }
    // descending bit order: address c's bit is bit 63 - (c & 63); the empty
    // asm keeps the sign test on the high word (the compiler would widen it
    // to a 64-bit compare)
    uint32_t top = (uint32_t)((half << (c & 63u)) >> 32);
    asm("" : "+v"(top));
    return top >> 31;
}

// Label-volume address of address c: the first-octant address when
// a.zlfold, else c itself (t = 0, a.zlf_mask all ones).  The mirror image of
// equal power-of-two periods 2^k is its own reflection u -> 2^k - 1 - u
// along every axis, which complements the axis' bits, and the voxel there
// carries the same label.  Address bit 3i + a holds bit i of axis a, so the
// axes to reflect are the bits of t = c >> (3k - 3), and their complement
// masks below the top three bits add up to t x (1 + 8 + ... + 8^(k-2))
// (disjoint bit positions, no carries); zoctant_constants splits that
// multiplier into a 24-bit part (a.zlf_mul) and a shifted top term.
__device__ __forceinline__ uint32_t zoctant(uint32_t c, const WalkArgs &a)
{
    const uint32_t t = __builtin_amdgcn_ubfe(c, a.zlf_shift, 3u);
    return (c ^ ((t << a.zlf_shift2) + __umul24(t, a.zlf_mul))) & a.zlf_mask;
}

// Label of the voxel at address c of a labelled walk.
__device__ __forceinline__ uint32_t zlabel_at(const WalkArgs &a, uint32_t c)
{
    return a.zlabels[zoctant(c, a)];
}

// The coordinate held in the bits m of address v (the axis' dilated bits
// gathered to the low end).  m is wave-uniform, so the loop is scalar: 2
// VALU per bit.
__device__ __forceinline__ uint32_t zcompact(uint32_t v, uint32_t m)
{
    uint32_t u = 0u, j = 0u;
    for (uint32_t mm = m; mm; mm &= mm - 1u, ++j)
        u |= ((v >> __builtin_ctz(mm)) & 1u) << j;
    return u;
}

// One tile (blockDim.x * K consecutive walkers from tile * blockDim.x * K,
// or a.tile_walkers from tile * a.tile_walkers with balanced tiles) walked
// for a.echoes x a.steps_per_echo steps on the unfolded image a.zwords /
// a.zl, with the direction table already in LDS.  EMODE 0 = hit
// counting (map), 1 = CPMG with per-walker or uniform delta, 2 = per-label
// delta (ldelta: the label deltas in LDS; labels from a.zlabels; per-label
// hit counts in a.label_hits).  Echo partials go to part[row * part_stride]
// (nullable).  All threads of the block call it.
// CACHE 0: no chunk cache; 1: 128-bit chunk, word select (the original
// form); 64 / 128: chunk of that many bits with the shift test.  TAB: the
// direction table form of zdirection (1: tab is the wide table; 2: keys is
// the byte-key table).
// NOFLUX: noflux walk on the periodic unfolded image (tab: the noflux table,
// keys built for 32-byte entries); walkers must start inside the image.
// EMODE 2 with K == 1 keeps the per-label hit counts in LDS (lh: blockDim
// rows of kLhStride entries; one LDS add per hit instead of a compare and
// add per label), else in registers.
// MFLAGS: mirror walk with reflection flags (tab / fix: zbuild_table_flags,
// or zbuild_table_flipped with TAB 3).
// NS == 8 (EMODE 2, K == 1, pipelined chunk cache): label-set walk, each
// walker carrying 8 energies, one per set of label deltas (ldelta: 8 rows of
// kMaxLabels, set-major), all starting from the walker's energy; a hit
// multiplies each by its set's delta of the hit label, in the order of the
// single-set walk, so each set's energies equal that walk's bit for bit.
// Echo partials of set q go to part[q * (a.echoes + 1) * part_stride + row
// * part_stride] (lds8: the staging of stage_echo_sets8); set 0's energy is
// stored.
// JAC (EMODE 2, K == 1, NS == 1): the Jacobian forward; besides the echo
// sums, each echo's block sums of h_l E for the 8 labels (the walker's hits
// on label l up to the echo, from its LDS row, times its energy: dE/d log
// delta_l, since E is the sequential product of the hit labels' deltas) go
// to a.jac_partials + tile, row stride part_stride, label stride (a.echoes +
// 1) part_stride, through the 8-value reduce-scatter (lds8: the staging of
// stage_echo_sets8).
// ADISP (mirror walks with reflection flags, every period above 2 kZFlush:
// zaddress_displacement_ok): no packed displacement per move; each flush
// recovers the walker's unbounded moves from its address change since the
// last flush (per axis, the coordinate change taken modulo the period into
// (-period / 2, period / 2), negated along reflected axes), one VALU per
// move fewer.
// HITS false (CPMG, a.hits null): no wall-hit count, as the reference's
// CPMG kernels (NMR_cpmg_cuda.cpp:380-386: a hit only multiplies the
// energy), one VALU per step fewer.
template <int EMODE, int K, bool WRAPFIX, int CACHE, int TAB = 0, bool NOFLUX = false, bool MFLAGS = false,
          int NS = 1, bool JAC = false, bool ADISP = false, bool HITS = true>
__device__ __forceinline__ void zwalk_tile(const WalkArgs &a, uint64_t tile, double *part, uint64_t part_stride,
                                           double (*lds)[kEchoCols], const uint4 *tab, const uint2 *fix,
                                           const double *ldelta = nullptr, const uint8_t *keys = nullptr,
                                           uint32_t *lh = nullptr, double *lds8 = nullptr)
{
    constexpr bool WIDE = TAB == 1;
    static_assert(!JAC || (EMODE == 2 && K == 1 && NS == 1 && CACHE >= 64), "Jacobian forward: labelled, K == 1");
    static_assert(!NOFLUX || CACHE >= 64, "noflux: pipelined chunk-cache walk only");
    static_assert(TAB != 3 || (MFLAGS && CACHE >= 64), "flag-keyed tables: pipelined mirror walk only");
    static_assert(NS == 1 || (NS == 8 && EMODE == 2 && K == 1 && CACHE >= 64), "label sets: 8 sets, K == 1");
    static_assert(!ADISP || (MFLAGS && !NOFLUX && CACHE >= 64), "address displacements: pipelined mirror walk");
    static_assert(HITS || (EMODE == 1 && CACHE >= 64), "uncounted hits: pipelined CPMG walk");
    // label-set walks count no per-label hits (a.label_hits is null)
    constexpr bool LH_LDS = EMODE == 2 && K == 1 && NS == 1;
    constexpr int NL = (EMODE == 2 && !LH_LDS && NS == 1) ? kMaxLabels : 1;
    const char *words = reinterpret_cast<const char *>(a.zwords);
    const zrsrc_t zrsrc = CACHE == 64 ? zchunk_rsrc8(a.zwords, a.zl) : zchunk_rsrc(a.zwords, a.zl);
    // the tile's walkers: blockDim x K from tile x blockDim x K, or with
    // balanced tiles a.tile_walkers (at most blockDim, K == 1) from tile x
    // a.tile_walkers
    const uint64_t tw = a.tile_walkers ? (uint64_t)a.tile_walkers : (uint64_t)blockDim.x * K;
    const uint64_t base = tile * tw + threadIdx.x;
    const uint64_t n = a.n;
    const uint64_t end = (tile + 1) * tw < n ? (tile + 1) * tw : n;
    const uint32_t echoes = a.echoes, spe = a.steps_per_echo;
    const int n_labels = a.n_labels;

    uint32_t A[K], C[K], Hc[K];
    uint32_t A0[K];   // ADISP: the address at the last flush
    uint32_t FO[K];   // MFLAGS: the walker's direction table offset (flags x kZFlagStride)
    uint32_t QB[K];   // chunk cache: chunk id (address >> 6 or >> 7) held in QV
    uint4 QV[K];
    int32_t P[K], DP[K], X[K], Y[K], Z[K];
    uint64_t S[K];
    double E[K], DL[K];
    double ES[NS];   // label-set walks: one energy per set
    uint32_t LH[K][NL];
    // labelled walks: a hit's label is loaded at the hit and applied one step
    // later (or at the echo's end), so its latency overlaps the next step's
    // chunk load instead of following the hit's own
    uint32_t PW[K];
    bool PEND[K];
    uint32_t *my_lh = LH_LDS ? lh + threadIdx.x * (uint32_t)kLhStride : nullptr;
    bool act[K];
    static_for<K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const uint64_t id = base + (uint64_t)k * blockDim.x;
        act[k] = id < end;
        const uint64_t j = act[k] ? id : 0;
        X[k] = act[k] ? a.x[j] : 0;
        Y[k] = act[k] ? a.y[j] : 0;
        Z[k] = act[k] ? a.z[j] : 0;
        if (NOFLUX && act[k] &&
            !((unsigned)X[k] < (unsigned)a.W && (unsigned)Y[k] < (unsigned)a.H && (unsigned)Z[k] < (unsigned)a.D)) {
            // a noflux walker must start inside the image (as the folded walk)
            act[k] = false;
            atomicOr(a.error_flag, 1u);
            X[k] = Y[k] = Z[k] = 0;
        }
        S[k] = act[k] ? a.seed[j] : 1ull;
        E[k] = act[k] ? (a.energy ? a.energy[j] : 1.0) : 0.0;   // idle lanes add 0 to echo sums
        Hc[k] = (act[k] && a.hits) ? a.hits[j] : 0u;
        DL[k] = EMODE == 1 ? ((act[k] && a.penalty) ? a.penalty[j] : a.delta_uniform) : 1.0;
        if (MFLAGS) {
            // start at the folded position (first octant of the mirror image)
            // with the reflection flags of the unbounded one
            const int32_t w = (int32_t)(a.zl.period[0] >> 1), h = (int32_t)(a.zl.period[1] >> 1),
                          d = (int32_t)(a.zl.period[2] >> 1);
            A[k] = zaddress(fold_mirror(X[k], w), fold_mirror(Y[k], h), fold_mirror(Z[k], d), a.zl);
            FO[k] = (TAB == 3 ? kZKeyEntries : kZFlagStride) *
                    ((uint32_t)mirror_flag(X[k], w) | ((uint32_t)mirror_flag(Y[k], h) << 1) |
                     ((uint32_t)mirror_flag(Z[k], d) << 2));
        } else {
            A[k] = zaddress(X[k], Y[k], Z[k], a.zl);
            FO[k] = 0u;
        }
        P[k] = 0;
        A0[k] = A[k];
        PW[k] = 0u;
        PEND[k] = false;
        if (NS > 1)
            static_for<NS>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                ES[q] = (act[k] && a.set_energy) ? a.set_energy[(uint64_t)q * n + j] : E[k];
            });
        QB[k] = 0xFFFFFFFFu;   // no chunk: addresses >> 7 stay below 2^25
        QV[k] = make_uint4(0u, 0u, 0u, 0u);
        static_for<NL>([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            LH[k][l] = (EMODE == 2 && a.label_hits && act[k] && l < n_labels) ? a.label_hits[(uint64_t)l * n + j]
                                                                               : 0u;
        });
        if (LH_LDS)
            for (int l = 0; l < kMaxLabels; ++l)
                my_lh[l] = (a.label_hits && act[k] && l < n_labels) ? a.label_hits[(uint64_t)l * n + j] : 0u;
    });

    auto echo_value = [&]() {
        double v = E[0];
        static_for<K - 1>([&](auto kc) { v += E[decltype(kc)::value + 1]; });
        return v;
    };
    auto flush = [&]() {
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            int32_t dx, dy, dz;
            if constexpr (ADISP) {
                const uint32_t flags = FO[k] / (TAB == 3 ? kZKeyEntries : kZFlagStride);
                int32_t d[3];
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    const int32_t per = (int32_t)a.zl.period[ax];
                    int32_t s = (int32_t)zcompact(A[k], a.zl.m[ax]) - (int32_t)zcompact(A0[k], a.zl.m[ax]);
                    if (s > per / 2)
                        s -= per;
                    else if (s < -(per / 2))
                        s += per;
                    d[ax] = ((flags >> ax) & 1u) ? -s : s;
                }
                dx = d[0];
                dy = d[1];
                dz = d[2];
                A0[k] = A[k];
            } else {
                zunpack(P[k], dx, dy, dz);
                P[k] = 0;
            }
            X[k] += dx;
            Y[k] += dy;
            Z[k] += dz;
        });
    };
    const char *tab_bytes = reinterpret_cast<const char *>(tab);
    const char *fix_bytes = reinterpret_cast<const char *>(fix);
    // CACHE 64 / 128 draw each walker's next direction one step ahead (T),
    // between the chunk load of a step and its use, so the RNG, the table
    // read and the load overlap; the state is rewound by one step at the end.
    constexpr bool PIPE = CACHE >= 64;
    ZEntry T[K];
    if (PIPE)
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            zdirection<WRAPFIX, TAB, NOFLUX>(S[k], tab_bytes, fix_bytes, T[k], keys, FO[k]);
        });
    // apply walker k's pending hit: energy *= its label's delta (per set) and
    // the per-label hit count, in hit order
    auto apply_label = [&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const uint32_t lab = PW[k];
        if constexpr (NS > 1) {
            static_for<NS>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                ES[q] *= ldelta[q * kMaxLabels + lab];
            });
        } else {
            E[k] *= ldelta[lab];
        }
        if (LH_LDS) {
            __hip_atomic_fetch_add(my_lh + lab, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            static_for<NL>([&](auto lc) {
                constexpr int l = decltype(lc)::value;
                LH[k][l] += lab == (uint32_t)l ? 1u : 0u;
            });
        }
    };
    constexpr bool DEFER = CACHE >= 64 && EMODE == 2;
    auto settle_labels = [&]() {
        if constexpr (DEFER)
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                if (PEND[k]) {
                    apply_label(kc);
                    PEND[k] = false;
                }
            });
    };
    auto step = [&]() {
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if (NOFLUX) {
                C[k] = zmove_noflux(A[k], T[k], DP[k]);
            } else if (PIPE) {
                C[k] = zmove_entry<WRAPFIX>(A[k], T[k]);
                DP[k] = (int32_t)T[k].dp;
            } else {
                C[k] = zpropose<WRAPFIX, WIDE>(S[k], A[k], tab_bytes, fix_bytes, DP[k]);
            }
        });
        uint32_t wallv[K];
        if (CACHE == 1) {
            // each walker keeps the 16-byte chunk (8 x 4 x 4 voxels) of its last
            // lookup; only candidates outside it read memory
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                zchunk_fill(words, C[k], QB[k], QV[k]);
            });
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                wallv[k] = zchunk_bit(C[k], QV[k]);
            });
        } else if (CACHE == 64 || CACHE == 128) {
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                if (CACHE == 128)
                    zchunk_fill_buf(zrsrc, C[k], QB[k], QV[k]);
                else
                    zchunk_fill_buf8(zrsrc, C[k], QB[k], QV[k]);
            });
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                zdirection<WRAPFIX, TAB, NOFLUX>(S[k], tab_bytes, fix_bytes, T[k], keys, FO[k]);
            });
            // keep the direction's table reads ahead of the wall test, so
            // their latency overlaps the chunk load's
            __builtin_amdgcn_sched_barrier(0);
            // the previous step's hit (its label load overlapped this step's chunk load)
            settle_labels();
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                wallv[k] = zchunk_wall<CACHE == 64 ? 64 : 128>(C[k], QV[k]);
            });
        } else {
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                // c < 2^bits, so the byte offset of its 64-bit pair of words is
                // inside the volume (at least 4 words)
                const uint64_t pair = *reinterpret_cast<const uint64_t *>(words + ((C[k] >> 3) & ~7u));
                wallv[k] = (uint32_t)((pair << (C[k] & 63u)) >> 63);
            });
        }
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const uint32_t wall = wallv[k];
            if (CACHE >= 64 && EMODE == 1) {
                // two short divergent paths (4 VALU in all) instead of the
                // selects and unconditional multiply the compiler would make
                if (__builtin_expect(wall != 0u, 0)) {
                    if (HITS)
                        Hc[k] += 1u;
                    E[k] *= DL[k];
                } else {
                    A[k] = C[k];
                    if (!ADISP)
                        P[k] += DP[k];
                    else   // a move of one copy: kept a branch, not turned into selects
                        asm volatile("" : "+v"(A[k]));
                }
                return;
            }
            if (CACHE >= 64 && EMODE == 2) {
                if (__builtin_expect(wall != 0u, 0)) {
                    Hc[k] += 1u;
                    PW[k] = zlabel_at(a, C[k]);
                    PEND[k] = true;
                } else {
                    A[k] = C[k];
                    if (!ADISP)
                        P[k] += DP[k];
                }
                return;
            }
            A[k] = wall ? A[k] : C[k];
            if (!ADISP)
                P[k] += wall ? 0 : DP[k];
            Hc[k] += wall;
            if (EMODE == 1 && wall)
                E[k] *= DL[k];
            if (EMODE == 2 && wall) {
                const uint32_t lab = zlabel_at(a, C[k]);
                E[k] *= ldelta[lab];
                if (LH_LDS) {
                    __hip_atomic_fetch_add(my_lh + lab, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    static_for<NL>([&](auto lc) {
                        constexpr int l = decltype(lc)::value;
                        LH[k][l] += lab == (uint32_t)l ? 1u : 0u;
                    });
                }
            }
        });
    };

    const uint64_t set_stride = (uint64_t)(echoes + 1) * part_stride;
    uint32_t since = 0;
    auto walk_echo = [&]() {
        uint32_t left = spe;
        while (left) {
            const uint32_t ns = left < kZFlush - since ? left : kZFlush - since;
            // two steps per trip: the pipelined direction entry then lives in
            // alternating registers instead of being copied every step
            uint32_t i = 0;
            for (; i + 2 <= ns; i += 2) {
                step();
                step();
            }
            if (i < ns)
                step();
            since += ns;
            left -= ns;
            if (since == kZFlush) {
                flush();
                since = 0;
            }
        }
        settle_labels();
    };
    if constexpr (NS > 1) {
        if (part)
            stage_echo_sets8(lds8, 0, echoes, ES, part, part_stride, set_stride);
        for (uint32_t echo = 1; echo <= echoes; ++echo) {
            walk_echo();
            if (part)
                stage_echo_sets8(lds8, echo, echoes, ES, part, part_stride, set_stride);
        }
    } else if constexpr (JAC) {
        // the labels' h_l E at echo 0 and after each echo (the pending hit
        // settled), staged beside the grouped echo sums
        double *jpart = part ? a.jac_partials + tile : nullptr;
        uint32_t jecho = 0;
        auto stage_jac = [&]() {
            double v[8];
#pragma unroll
            for (int l = 0; l < 8; ++l)
                v[l] = (double)my_lh[l] * E[0];
            stage_echo_sets8(lds8, jecho, echoes, v, jpart, part_stride, set_stride);
        };
        static_assert(kMaxLabels == 8, "one reduce-scatter lane group per label");
        if (jpart)
            stage_jac();
        auto walk_echo_jac = [&]() {
            walk_echo();
            ++jecho;
            if (jpart)
                stage_jac();
        };
        run_echoes<true>(echoes, lds, part, part_stride, walk_echo_jac, echo_value);
    } else {
        // one walker per lane: echoes reduced in groups of 4; several walkers
        // per lane stage each echo (the group buffer would push those
        // register-heavy forms into scratch)
        run_echoes<K == 1>(echoes, lds, part, part_stride, walk_echo, echo_value);
    }
    flush();
    if (NS > 1)
        E[0] = ES[0];
    if (PIPE)
        static_for<K>([&](auto kc) { S[decltype(kc)::value] = xorshift64_inverse(S[decltype(kc)::value]); });

    static_for<K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (act[k]) {
            const uint64_t id = base + (uint64_t)k * blockDim.x;
            a.x[id] = X[k];
            a.y[id] = Y[k];
            a.z[id] = Z[k];
            a.seed[id] = S[k];
            if (a.energy)
                a.energy[id] = E[k];
            if (NS > 1 && a.set_energy)
                static_for<NS>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    a.set_energy[(uint64_t)q * n + id] = ES[q];
                });
            if (a.hits)
                a.hits[id] = Hc[k];
            if (EMODE == 2 && a.label_hits) {
                if (LH_LDS) {
                    for (int l = 0; l < n_labels; ++l)
                        a.label_hits[(uint64_t)l * n + id] = my_lh[l];
                } else {
                    static_for<NL>([&](auto lc) {
                        constexpr int l = decltype(lc)::value;
                        if (l < n_labels)
                            a.label_hits[(uint64_t)l * n + id] = LH[k][l];
                    });
                }
            }
        }
    });
}

// Speculative CPMG tile for latency-bound runs (few waves per SIMD, where a
// step waits for its chunk load with nothing else to issue): one walker per
// lane, no chunk cache, directions drawn two steps ahead.  At the start of
// step t both possible candidates of step t + 1 are formed, from the current
// position (if step t hits a wall) and from step t's candidate (if it
// moves), and both of their 16-byte chunks are loaded; then the direction of
// step t + 2 is drawn, and only then does step t wait for its own chunk,
// loaded at the start of step t - 1 (the empty asm holds that wait after the
// draw).  Two chunk loads per walker-step (more VALU and L2 requests than
// the chunk-cache walk), so it pays only when the walk is latency-bound.
// Same walk, bit for bit, as zwalk_tile (EMODE 1).
template <bool WRAPFIX, bool NOFLUX, bool MFLAGS>
__device__ __forceinline__ void zwalk_tile_spec(const WalkArgs &a, uint64_t tile, double *part, uint64_t part_stride,
                                                double (*lds)[kEchoCols], const uint4 *tab, const uint2 *fix,
                                                const uint8_t *keys)
{
    const zrsrc_t zrsrc = zchunk_rsrc(a.zwords, a.zl);
    const uint64_t id = tile * blockDim.x + threadIdx.x;
    const uint32_t echoes = a.echoes, spe = a.steps_per_echo;
    bool act = id < a.n;
    const uint64_t j = act ? id : 0;
    int32_t X = act ? a.x[j] : 0, Y = act ? a.y[j] : 0, Z = act ? a.z[j] : 0;
    if (NOFLUX && act && !((unsigned)X < (unsigned)a.W && (unsigned)Y < (unsigned)a.H && (unsigned)Z < (unsigned)a.D)) {
        act = false;   // a noflux walker must start inside the image
        atomicOr(a.error_flag, 1u);
        X = Y = Z = 0;
    }
    uint64_t S = act ? a.seed[j] : 1ull;
    double E = act ? (a.energy ? a.energy[j] : 1.0) : 0.0;   // idle lanes add 0 to echo sums
    uint32_t Hc = (act && a.hits) ? a.hits[j] : 0u;
    const double DL = (act && a.penalty) ? a.penalty[j] : a.delta_uniform;
    uint32_t A, FO = 0u;
    if (MFLAGS) {
        const int32_t w = (int32_t)(a.zl.period[0] >> 1), h = (int32_t)(a.zl.period[1] >> 1),
                      d = (int32_t)(a.zl.period[2] >> 1);
        A = zaddress(fold_mirror(X, w), fold_mirror(Y, h), fold_mirror(Z, d), a.zl);
        FO = kZFlagStride * ((uint32_t)mirror_flag(X, w) | ((uint32_t)mirror_flag(Y, h) << 1) |
                             ((uint32_t)mirror_flag(Z, d) << 2));
    } else {
        A = zaddress(X, Y, Z, a.zl);
    }
    int32_t P = 0;
    const char *tab_bytes = reinterpret_cast<const char *>(tab);
    const char *fix_bytes = reinterpret_cast<const char *>(fix);
    auto move = [&](uint32_t from, const ZEntry &t, int32_t &dp) -> uint32_t {
        if (NOFLUX)
            return zmove_noflux(from, t, dp);
        dp = (int32_t)t.dp;
        return zmove_entry<WRAPFIX>(from, t);
    };
    auto chunk = [&](uint32_t c) -> uint4 {
        const zv4u v = zchunk_load(zrsrc, c >> 7);
        return make_uint4(v.x, v.y, v.z, v.w);
    };
    // the first step's candidate C and its chunk (in both slots), and the
    // second step's direction T
    ZEntry T;
    zdirection<WRAPFIX, 2, NOFLUX>(S, tab_bytes, fix_bytes, T, keys, FO);
    int32_t DP;
    uint32_t C = move(A, T, DP);
    uint4 QS = chunk(C), QM = chunk(C);
    bool stay = true;   // which of QS / QM holds the chunk of C
    zdirection<WRAPFIX, 2, NOFLUX>(S, tab_bytes, fix_bytes, T, keys, FO);
    auto step = [&]() {
        int32_t dps, dpm;
        const uint32_t cs = move(A, T, dps);   // next candidate if this step hits
        const uint32_t cm = move(C, T, dpm);   // next candidate if this step moves
        const uint4 qs = chunk(cs);
        const uint4 qm = chunk(cm);
        ZEntry Tn;
        zdirection<WRAPFIX, 2, NOFLUX>(S, tab_bytes, fix_bytes, Tn, keys, FO);
        // this step's chunks were loaded a step ago: wait for them here,
        // after the loads above and the draw, not before
        asm volatile("" : "+v"(QS.x), "+v"(QS.y), "+v"(QS.z), "+v"(QS.w), "+v"(QM.x), "+v"(QM.y), "+v"(QM.z),
                     "+v"(QM.w)::"memory");
        const uint4 q = stay ? QS : QM;
        const bool wall = zchunk_wall<128>(C, q) != 0u;
        if (__builtin_expect(wall, 0)) {
            Hc += 1u;
            E *= DL;
        } else {
            A = C;
            P += DP;
        }
        C = wall ? cs : cm;
        DP = wall ? dps : dpm;
        QS = qs;
        QM = qm;
        stay = wall;
        T = Tn;
    };
    uint32_t since = 0;
    auto flush = [&]() {
        int32_t dx, dy, dz;
        zunpack(P, dx, dy, dz);
        X += dx;
        Y += dy;
        Z += dz;
        P = 0;
    };
    auto walk_echo = [&]() {
        uint32_t left = spe;
        while (left) {
            const uint32_t ns = left < kZFlush - since ? left : kZFlush - since;
            uint32_t i = 0;
            for (; i + 2 <= ns; i += 2) {
                step();
                step();
            }
            if (i < ns)
                step();
            since += ns;
            left -= ns;
            if (since == kZFlush) {
                flush();
                since = 0;
            }
        }
    };
    run_echoes<true>(echoes, lds, part, part_stride, walk_echo, [&]() { return E; });
    flush();
    // the directions were drawn two steps ahead
    S = xorshift64_inverse(xorshift64_inverse(S));
    if (act) {
        a.x[id] = X;
        a.y[id] = Y;
        a.z[id] = Z;
        a.seed[id] = S;
        if (a.energy)
            a.energy[id] = E;
        if (a.hits)
            a.hits[id] = Hc;
    }
}

}  // namespace rwnmr
