// locality_kernels.h -- position ordering of walkers (see locality_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rwnmr {

// Walker state in sorted order plus each walker's index in the caller's order.
struct SortedState {
    int32_t *x, *y, *z;
    uint64_t *seed;
    double *energy;
    uint32_t *hits;      // nullable
    double *penalty;     // nullable (uniform delta)
    uint32_t *orig;
};

hipError_t launch_morton_keys(int bc, const int32_t *x, const int32_t *y, const int32_t *z, uint64_t n, int W, int H,
                              int D, uint32_t *keys, uint32_t *iota, hipStream_t st);
hipError_t sort_pairs_temp_bytes(uint64_t n, int end_bit, size_t *bytes);
hipError_t sort_pairs(void *temp, size_t bytes, const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in,
                      uint32_t *vals_out, uint64_t n, int end_bit, hipStream_t st);
hipError_t launch_gather_walkers(const SortedState &dst, const SortedState &src, const uint32_t *perm,
                                 const uint32_t *orig_src, const double *penalty, uint64_t n, hipStream_t st);
hipError_t launch_scatter_walkers(const SortedState &out, const SortedState &s, uint64_t n, hipStream_t st);

}  // namespace rwnmr
