// batch_kernels.hip -- many CPMG instances (uneven images, walker counts,
// relaxivities, echo counts) in one persistent launch.
//
// The host cuts every instance into tiles of blockDim.x * K walkers and
// orders the tile list by cost (steps x walkers, longest first).  A grid
// sized to the device's resident capacity then loops: one workgroup pops the
// next tile index from a global atomic counter, loads that instance's
// descriptor (image pointer, dimensions, boundary, relaxivity, echo count)
// and walks the tile with the same tile walk as the single-instance kernels
// (zwalk_tile on the instance's unfolded image for periodic and mirror
// boundaries, walk_tile otherwise), writing one FP64 partial per (echo, tile).  Every workgroup leaves
// the loop once the counter passes the tile count, so the grid always drains.
// The reference runs one sample per process invocation
// (B/NMR_Simulation/NMR_cpmg_cuda.cpp:929-1774); this has no counterpart.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "batch_kernels.h"
#include "walk_device.h"
#include "zwalk_device.h"

namespace rwnmr {

template <int K>
__global__ __launch_bounds__(256) void batch_kernel(const BatchArgs b)
{
    __shared__ double lds[kEchoChunk][kEchoCols];
    __shared__ uint4 tab[6];
    __shared__ uint2 fix[6];
    __shared__ uint8_t keys[kZKeyEntries];   // the byte-key table: the same for every instance
    __shared__ uint32_t next;
    zbuild_key_table(keys);
    for (;;) {
        if (threadIdx.x == 0)
            next = atomicAdd(b.counter, 1u);
        __syncthreads();
        // uniform across the block: scalar loads of the tile and instance
        // descriptors (kept out of VGPRs and scratch)
        const uint32_t t = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();  // every thread has read `next` before it is reused
        if (t >= b.n_tiles)
            break;
        const BatchTile bt = b.tiles[t];
        // fields read through a reference (a copy of the whole descriptor
        // went through scratch)
        const BatchInstance &I = b.inst[__builtin_amdgcn_readfirstlane(bt.inst)];
        WalkArgs a;
        a.blocks = I.blocks;
        a.labels_blocked = nullptr;
        a.W = I.W;
        a.H = I.H;
        a.D = I.D;
        a.bC = I.bC;
        a.bCR = I.bCR;
        a.layout = I.layout;
        a.tile_map = 0;
        a.n_tiles = 0;
        a.tile_walkers = 0;   // tiles of blockDim x K walkers
        a.shift = 0;
        a.x = I.x;
        a.y = I.y;
        a.z = I.z;
        a.seed = I.seed;
        a.energy = I.energy;
        a.hits = I.hits;
        a.penalty = I.penalty;
        a.delta_uniform = I.delta_uniform;
        a.label_delta = nullptr;
        a.n_labels = 0;
        a.label_hits = nullptr;
        a.n = I.n;
        a.echoes = I.echoes;
        a.steps_per_echo = I.steps_per_echo;
        a.partials = nullptr;
        a.partial_stride = 0;
        a.error_flag = b.error_flag;
        a.zwords = I.zwords;
        a.zl = I.zl;
        a.zcache = 2;
        a.znoflux = 0;   // noflux instances run the folded walk here
        a.zmflags = 0;   // mirror instances keep their unfolded start positions
        double *part = I.partials + bt.tile;
        if (I.zwords) {
            // the previous tile's table reads ended at its last block barrier
            // (the key table was synced by the barriers of the first tile)
            zbuild_table(I.zl, tab, fix);
            __syncthreads();
            if (I.zl.wrapfix)
                zwalk_tile<1, K, true, 128, 2>(a, bt.tile, part, I.n_tiles, lds, tab, fix, nullptr, keys);
            else
                zwalk_tile<1, K, false, 128, 2>(a, bt.tile, part, I.n_tiles, lds, tab, fix, nullptr, keys);
            continue;
        }
        switch (I.bc) {
        case kNoflux: walk_tile<kNoflux, 1, K>(a, bt.tile, part, I.n_tiles, lds); break;
        case kPeriodic: walk_tile<kPeriodic, 1, K>(a, bt.tile, part, I.n_tiles, lds); break;
        default: walk_tile<kMirror, 1, K>(a, bt.tile, part, I.n_tiles, lds); break;
        }
    }
}

// Segmented fixed-order sums: out[r] = sum_c in[r].base[c * 0 + c] over
// rows of different lengths (one block per row).
__global__ __launch_bounds__(256) void reduce_segments_kernel(const ReduceRow *rows, uint64_t n_rows)
{
    __shared__ double red[256];
    for (uint64_t r = blockIdx.x; r < n_rows; r += gridDim.x) {
        const ReduceRow row = rows[r];
        double s = 0.0;
        for (uint32_t c = threadIdx.x; c < row.cols; c += blockDim.x)
            s += row.in[c];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int off = blockDim.x / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off)
                red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            *row.out = red[0];
        __syncthreads();
    }
}

hipError_t launch_batch(const BatchArgs &b, uint32_t k, uint32_t grid, hipStream_t st)
{
    switch (k) {
    case 1: hipLaunchKernelGGL(batch_kernel<1>, dim3(grid), dim3(kBatchThreads), 0, st, b); break;
    case 2: hipLaunchKernelGGL(batch_kernel<2>, dim3(grid), dim3(kBatchThreads), 0, st, b); break;
    case 4: hipLaunchKernelGGL(batch_kernel<4>, dim3(grid), dim3(kBatchThreads), 0, st, b); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t batch_resident_blocks(uint32_t k, int *per_cu)
{
    switch (k) {
    case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, batch_kernel<1>, kBatchThreads, 0);
    case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, batch_kernel<2>, kBatchThreads, 0);
    case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, batch_kernel<4>, kBatchThreads, 0);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reduce_segments(const ReduceRow *rows, uint64_t n_rows, hipStream_t st)
{
    if (n_rows == 0)
        return hipSuccess;
    const uint32_t grid = (uint32_t)(n_rows < 65536 ? n_rows : 65536);
    hipLaunchKernelGGL(reduce_segments_kernel, dim3(grid), dim3(256), 0, st, rows, n_rows);
    return hipGetLastError();
}

}  // namespace rwnmr
