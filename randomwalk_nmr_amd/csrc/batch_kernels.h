// batch_kernels.h -- host-side view of the batched persistent CPMG kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_kernels.h"

namespace rwnmr {

static constexpr uint32_t kBatchThreads = 256;

// Per-instance descriptor (device copy).  Walker arrays are device pointers.
struct BatchInstance {
    const uint64_t *blocks;
    int32_t W, H, D, bC, bCR;
    BrickLayout layout;
    int32_t bc;
    int32_t *x, *y, *z;
    uint64_t *seed;
    double *energy;
    uint32_t *hits;
    const double *penalty;
    double delta_uniform;
    uint64_t n;
    uint32_t echoes, steps_per_echo;
    double *partials;    // [(echoes + 1)][n_tiles]
    uint32_t n_tiles;
    const uint32_t *zwords;   // unfolded image (periodic / mirror), else null: folded walk
    ZLayout zl;
};

struct BatchTile {
    uint32_t inst;
    uint32_t tile;       // tile index inside the instance
};

struct BatchArgs {
    const BatchInstance *inst;
    const BatchTile *tiles;
    uint32_t n_tiles;
    uint32_t *counter;   // work-queue head, zero before launch
    uint32_t *error_flag;
};

// One fixed-order sum: *out = in[0] + ... + in[cols - 1].
struct ReduceRow {
    const double *in;
    uint32_t cols;
    double *out;
};

hipError_t launch_batch(const BatchArgs &b, uint32_t k, uint32_t grid, hipStream_t st);
hipError_t batch_resident_blocks(uint32_t k, int *per_cu);
hipError_t launch_reduce_segments(const ReduceRow *rows, uint64_t n_rows, hipStream_t st);

}  // namespace rwnmr
