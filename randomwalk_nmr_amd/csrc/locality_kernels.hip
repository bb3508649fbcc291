// locality_kernels.hip -- keep walkers that are close in the image close in
// memory and on the same XCD.
//
// A walk step reads one bitBlock word near the walker, so the cache
// footprint of a workgroup is the union of its walkers' neighbourhoods.
// Ordering walkers by the Morton code of their image block (radix sort of
// 30-bit keys; the unfolded-image walk sorts by its own Z-order word address) makes every tile a compact cluster, and striping the sorted
// tiles per XCD (tile_of_block) gives each XCD's L2 one region of the image.
// Per-walker results do not depend on the order (each walker owns its RNG
// stream), so the permutation is undone exactly at the end.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <hipcub/hipcub.hpp>

#include "locality_kernels.h"
#include "walk_common.h"

namespace rwnmr {

template <int BC>
__global__ __launch_bounds__(256) void morton_keys_kernel(const int32_t *x, const int32_t *y, const int32_t *z,
                                                         uint64_t n, int W, int H, int D, uint32_t *keys,
                                                         uint32_t *iota)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int ix = x[i], iy = y[i], iz = z[i];
    if (BC == kMirror) {
        ix = fold_mirror(ix, W);
        iy = fold_mirror(iy, H);
        iz = fold_mirror(iz, D);
    } else if (BC == kPeriodic) {
        ix = fold_periodic(ix, W);
        iy = fold_periodic(iy, H);
        iz = fold_periodic(iz, D);
    } else {
        // noflux positions are inside the image; clamp defensively for the key
        ix = ix < 0 ? 0 : (ix >= W ? W - 1 : ix);
        iy = iy < 0 ? 0 : (iy >= H ? H - 1 : iy);
        iz = iz < 0 ? 0 : (iz >= D ? D - 1 : iz);
    }
    keys[i] = morton3((uint32_t)ix >> 2, (uint32_t)iy >> 2, (uint32_t)iz >> 2);
    iota[i] = (uint32_t)i;
}

// dst[i] = src[perm[i]] for the walker state; orig_dst[i] = orig_src[perm[i]]
// (orig_src null: identity).  Penalties are read through the original
// index from the caller's array.
__global__ __launch_bounds__(256) void gather_walkers_kernel(SortedState dst, SortedState src, const uint32_t *perm,
                                                            const uint32_t *orig_src, const double *penalty,
                                                            uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t p = perm[i];
    const uint32_t o = orig_src ? orig_src[p] : p;
    dst.x[i] = src.x[p];
    dst.y[i] = src.y[p];
    dst.z[i] = src.z[p];
    dst.seed[i] = src.seed[p];
    dst.energy[i] = src.energy ? src.energy[p] : 1.0;
    if (dst.hits)
        dst.hits[i] = src.hits ? src.hits[p] : 0u;
    if (dst.penalty)
        dst.penalty[i] = penalty[o];
    dst.orig[i] = o;
}

// Caller's arrays back in the caller's order: out[orig[i]] = s[i].
__global__ __launch_bounds__(256) void scatter_walkers_kernel(SortedState out, SortedState s, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t o = s.orig[i];
    out.x[o] = s.x[i];
    out.y[o] = s.y[i];
    out.z[o] = s.z[i];
    out.seed[o] = s.seed[i];
    if (out.energy)
        out.energy[o] = s.energy[i];
    if (out.hits)
        out.hits[o] = s.hits[i];
}

hipError_t launch_morton_keys(int bc, const int32_t *x, const int32_t *y, const int32_t *z, uint64_t n, int W, int H,
                              int D, uint32_t *keys, uint32_t *iota, hipStream_t st)
{
    const uint32_t grid = (uint32_t)((n + 255) / 256);
    switch (bc) {
    case kNoflux: hipLaunchKernelGGL(morton_keys_kernel<kNoflux>, dim3(grid), dim3(256), 0, st, x, y, z, n, W, H, D, keys, iota); break;
    case kPeriodic: hipLaunchKernelGGL(morton_keys_kernel<kPeriodic>, dim3(grid), dim3(256), 0, st, x, y, z, n, W, H, D, keys, iota); break;
    default: hipLaunchKernelGGL(morton_keys_kernel<kMirror>, dim3(grid), dim3(256), 0, st, x, y, z, n, W, H, D, keys, iota); break;
    }
    return hipGetLastError();
}

hipError_t sort_pairs_temp_bytes(uint64_t n, int end_bit, size_t *bytes)
{
    return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, end_bit);
}

hipError_t sort_pairs(void *temp, size_t bytes, const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in,
                      uint32_t *vals_out, uint64_t n, int end_bit, hipStream_t st)
{
    return hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys_in, keys_out, vals_in, vals_out, (int)n, 0, end_bit,
                                              st);
}

hipError_t launch_gather_walkers(const SortedState &dst, const SortedState &src, const uint32_t *perm,
                                 const uint32_t *orig_src, const double *penalty, uint64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(gather_walkers_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, dst, src, perm,
                       orig_src, penalty, n);
    return hipGetLastError();
}

hipError_t launch_scatter_walkers(const SortedState &out, const SortedState &s, uint64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(scatter_walkers_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, out, s, n);
    return hipGetLastError();
}

}  // namespace rwnmr
