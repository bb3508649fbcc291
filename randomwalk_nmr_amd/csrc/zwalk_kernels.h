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

// Words of the unfolded bit volume for layout zl (at least one 16-byte chunk).
inline uint64_t zimage_words(const ZLayout &zl) { return zl.bits > 7 ? (1ull << (zl.bits - 5)) : 4ull; }

// Build the unfolded image: bit A of out is voxel (zunfold(ux), zunfold(uy),
// zunfold(uz)) of the device bitBlock image `blocks` (brick layout L), 1 for
// addresses outside the periods.
hipError_t launch_zimage_build(const uint64_t *blocks, const BrickLayout &L, int W, int H, int D, int mirror,
                               const ZLayout &zl, uint32_t *out, hipStream_t st);

// Words of the unfolded 4-bit label volume (8 voxels per word, at least one).
inline uint64_t zlabel_words(const ZLayout &zl) { return zl.bits > 3 ? (1ull << (zl.bits - 3)) : 1ull; }

// Build the unfolded label volume from the device labels (64 bytes per block
// of `blocks`' brick layout L); 0 outside the periods.
hipError_t launch_zlabels_build(const uint8_t *labels_blocked, const BrickLayout &L, int W, int H, int D, int mirror,
                                const ZLayout &zl, uint32_t *out, hipStream_t st);

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

}  // namespace rwnmr
