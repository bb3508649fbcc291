// zwalk_kernels.h -- host-side view of the unfolded-image walk kernels
// (periodic and mirror CPMG / map walks; see zwalk_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_common.h"
#include "walk_kernels.h"

namespace rwnmr {

// Steps between flushes of the packed displacement into x, y, z (each packed
// component must stay within +-511).
static constexpr uint32_t kZFlush = 511;

// Address displacements recover the unbounded moves when every period of
// the layout exceeds twice the flush interval: a walker moves at most
// kZFlush voxels per axis between flushes, less than half a period, so the
// coordinate change modulo the period is the move itself.
__host__ __device__ inline bool zaddress_displacement_ok(const ZLayout &zl)
{
    return zl.period[0] > 2u * kZFlush && zl.period[1] > 2u * kZFlush && zl.period[2] > 2u * kZFlush;
}

// True when a CPMG launch of shape ls runs the mirror walk with flag-keyed
// tables and address displacements (zwalk_kernel<1, 1, ..., TAB 3, ...,
// ADISP>): one walker per lane, 1024-thread blocks (two per CU).
inline bool zadisp_walk(const WalkArgs &a, int emode, const LaunchShape &ls)
{
    return a.zwords && a.zmflags && !a.znoflux && (a.zcache == 2 || a.zcache == 3 || a.zcache == 5) &&
           emode == 1 && ls.k == 1 && ls.threads >= 1024 && a.shift == 0 && zaddress_displacement_ok(a.zl);
}

// Balanced tiles of that walk: with more walkers than one round of `slots`
// resident blocks holds, tiles of equal walker counts (at most blockDim)
// that fill whole rounds, so no CU idles through a part-empty last round
// (1M walkers on 512 slots: 1024 tiles of 977 walkers, two full rounds,
// instead of 977 tiles of 1024, the second round 465 of 512).  Returns the
// walkers per tile, or 0 for the usual tiles of blockDim x K walkers.
inline uint64_t zbalanced_tile_walkers(const WalkArgs &a, int emode, const LaunchShape &ls, uint32_t slots)
{
    const uint64_t round = (uint64_t)slots * ls.threads;
    if (!zadisp_walk(a, emode, ls) || slots == 0 || a.n <= round)
        return 0;
    const uint64_t rounds = (a.n + round - 1) / round;
    return (a.n + rounds * slots - 1) / (rounds * slots);
}

// Words of the unfolded bit volume for layout zl (at least one 16-byte chunk).
inline uint64_t zimage_words(const ZLayout &zl) { return zl.bits > 7 ? (1ull << (zl.bits - 5)) : 4ull; }

// Build the unfolded image: bit A of out is voxel (zunfold(ux), zunfold(uy),
// zunfold(uz)) of the device bitBlock image `blocks` (brick layout L), 1 for
// addresses outside the periods.
hipError_t launch_zimage_build(const uint64_t *blocks, const BrickLayout &L, int W, int H, int D, int mirror,
                               const ZLayout &zl, uint32_t *out, hipStream_t st);

// Mirror image of three equal power-of-two periods 2^k, address bit 3i + a
// holding bit i of axis a: the top three address bits are the axes' top
// bits, so the first octant is the first eighth of the addresses and the
// labels can be kept for it alone (read at zoctant(c)).
inline bool zoctant_layout(const ZLayout &zl)
{
    if (zl.wrapfix || zl.bits < 3 || zl.bits % 3 != 0)
        return false;
    uint32_t p = 0;
    for (uint32_t i = 0; i < zl.bits / 3; ++i)
        p |= 1u << (3 * i);
    return zl.m[0] == p && zl.m[1] == p << 1 && zl.m[2] == p << 2;
}

// Constants of zoctant in a: the fold when `fold` (zoctant_layout(a.zl)
// holds), else the identity (t = c >> 31 = 0 below 2^31 addresses).  With k
// = bits / 3 the fold's multiplier is P = 1 + 8 + ... + 8^(k-2); it fits the
// 24-bit multiply up to k = 9, and for k = 10 its top term 2^24 is added as
// t << 24 instead (a shift of 31 otherwise leaves only bits above the mask).
inline void zoctant_constants(WalkArgs &a, bool fold)
{
    a.zlfold = fold ? 1u : 0u;
    a.zlf_shift = fold ? a.zl.bits - 3 : 31u;
    uint32_t p = 0;
    for (uint32_t i = 0; fold && i + 1 < a.zl.bits / 3; ++i)
        p |= 1u << (3 * i);
    a.zlf_shift2 = (p >> 24) ? 24u : 31u;
    a.zlf_mul = p & 0xFFFFFFu;
    a.zlf_mask = fold ? (1u << (a.zl.bits - 3)) - 1u : 0xFFFFFFFFu;
}

// Bytes of the label volume: one per address, or one per first-octant
// address (octant); a multiple of 4.
inline uint64_t zlabel_bytes(const ZLayout &zl, bool octant)
{
    const uint64_t b = 1ull << (octant ? zl.bits - 3 : zl.bits);
    return b < 4 ? 4 : b;
}

// Build the label volume (one byte per voxel, the unfolded image's address
// order; octant: the first octant's addresses only) from the device labels
// (64 bytes per block of `blocks`' brick layout L); 0 outside the periods.
hipError_t launch_zlabels_build(const uint8_t *labels_blocked, const BrickLayout &L, int W, int H, int D, int mirror,
                                const ZLayout &zl, bool octant, uint8_t *out, hipStream_t st);

// keys[i] = unfolded address of walker i >> 5 (its 32-bit word), iota[i] = i.
hipError_t launch_zkeys(const int32_t *x, const int32_t *y, const int32_t *z, uint64_t n, const ZLayout &zl, int mflags,
                        uint32_t *keys, uint32_t *iota, hipStream_t st);

// PFGSE walk + measurement on the unfolded image (a.zwords / a.zl).
hipError_t launch_zpfgse(const WalkArgs &a, const PfgseArgs &p, hipStream_t st);

// CPMG (emode 1), map (emode 0) or per-label CPMG (emode 2, needs a.zlabels
// and a.label_delta) walk on a.zwords / a.zl; a.shift must be 0.
hipError_t launch_zwalk(const WalkArgs &a, int emode, const LaunchShape &ls, hipStream_t st);

// Label-set CPMG walk (a.n_label_sets <= 8 rows of a.n_labels deltas in
// a.label_delta; a.zlabels needed; one walker per lane, `threads` per
// block); partials [set][echo rows][a.partial_stride] of 8 sets.
hipError_t launch_zwalk_sets(const WalkArgs &a, uint32_t threads, hipStream_t st);
// Jacobian forward (per-label CPMG that also returns the block sums of h_l E
// per label and echo in a.jac_partials): a.zlabels, a.label_delta,
// a.partials and a.jac_partials needed; one walker per lane, `threads` per
// block.
hipError_t launch_zwalk_jac(const WalkArgs &a, uint32_t threads, hipStream_t st);

}  // namespace rwnmr
