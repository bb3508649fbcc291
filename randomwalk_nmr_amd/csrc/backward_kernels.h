// backward_kernels.h -- checkpoint layout and the replay backward kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_kernels.h"

namespace rwnmr {

// One checkpoint segment, structure of arrays over n walkers.
struct CkptView {
    int32_t *x, *y, *z;
    uint64_t *seed;
    double *energy;
    uint32_t *hits;   // [n_labels][n]
};

// Bytes per segment: 3 int32 + u64 seed + f64 energy + n_labels u32, 8-aligned.
__host__ __device__ inline uint64_t ckpt_segment_bytes(uint64_t n, uint32_t n_labels)
{
    const uint64_t words = (3ull * n * 4 + 7) / 8 * 8;
    return words + 16ull * n + ((4ull * n_labels * n + 7) / 8) * 8;
}

__host__ __device__ inline CkptView ckpt_view(void *base, uint64_t n, uint32_t n_segments, uint32_t n_labels,
                                              uint32_t seg)
{
    (void)n_segments;
    uint8_t *p = (uint8_t *)base + (uint64_t)seg * ckpt_segment_bytes(n, n_labels);
    CkptView v;
    v.x = (int32_t *)p;
    v.y = v.x + n;
    v.z = v.y + n;
    p += (3ull * n * 4 + 7) / 8 * 8;
    v.seed = (uint64_t *)p;
    v.energy = (double *)(p + 8ull * n);
    v.hits = (uint32_t *)(p + 16ull * n);
    return v;
}

// The per-label hits at the end of the run, after the n_segments segments.
__host__ __device__ inline uint32_t *ckpt_final_hits(void *base, uint64_t n, uint32_t n_segments, uint32_t n_labels)
{
    return (uint32_t *)((uint8_t *)base + (uint64_t)n_segments * ckpt_segment_bytes(n, n_labels));
}

// Whole checkpoint buffer: the segments, then the final per-label hits.
__host__ __device__ inline uint64_t ckpt_total_bytes(uint64_t n, uint32_t n_segments, uint32_t n_labels)
{
    return ckpt_segment_bytes(n, n_labels) * n_segments + ((4ull * n_labels * n + 7) / 8) * 8;
}

// Threads per block of the unfolded replay (its LDS holds kMaxLabels doubles
// per thread).
constexpr uint32_t kZBackThreads = 512;

struct BackwardArgs {
    void *ckpt;
    uint64_t n;
    uint32_t n_segments;
    int32_t n_labels;
    uint32_t every;            // echoes per segment
    uint32_t echoes;
    uint32_t steps_per_echo;
    const double *grad_echo;   // [echoes + 1] dL/dM(n)
    double *partials;          // [n_labels][partial_stride]
    uint64_t partial_stride;   // = segments * blocks
    uint32_t blocks;           // walker tiles (blocks) per segment
};

// Segment and walker tile of this workgroup of a replay launch (a 1-D grid
// of 8 * ceil(blocks * segments / 8) workgroups).  Workgroup g runs on XCD
// g % 8, and XCD j replays the j-th contiguous eighth of the tile-major,
// segment-minor list, so with position-ordered walkers each XCD's L2 holds
// one region of the image in every segment.  False for a spare workgroup.
__device__ __forceinline__ bool backward_block(const BackwardArgs &b, uint32_t &seg, uint64_t &tile)
{
    const uint64_t total = (uint64_t)b.blocks * b.n_segments;
    const uint64_t per_xcd = (total + kXcds - 1) / kXcds;
    const uint64_t t = (uint64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
    if (t >= total)
        return false;
    tile = t / b.n_segments;
    seg = (uint32_t)(t % b.n_segments);
    return true;
}

// Workgroups of a replay launch of `blocks` tiles per segment.
inline uint64_t backward_grid(uint64_t blocks, uint32_t n_segments)
{
    const uint64_t total = blocks * n_segments;
    return ((total + kXcds - 1) / kXcds) * kXcds;
}

hipError_t launch_backward(const WalkArgs &a, const BackwardArgs &b, int bc, uint32_t threads, hipStream_t st);

}  // namespace rwnmr
