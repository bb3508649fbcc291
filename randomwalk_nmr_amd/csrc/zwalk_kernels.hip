// zwalk_kernels.hip -- gfx950 CPMG / map walk for periodic and mirror
// boundaries on the unfolded image (step described in zwalk_device.h), the
// unfolded-image build and the Z-order sort keys.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_common.h"
#include "walk_device.h"
#include "zwalk_device.h"
#include "zwalk_kernels.h"

namespace rwnmr {

template <int EMODE, int K, bool WRAPFIX, bool RCACHE>
__global__ __launch_bounds__(max_threads_for_k(K)) void zwalk_kernel(const WalkArgs a)
{
    __shared__ double lds[kEchoChunk][kMaxWaves];
    __shared__ uint4 tab[6];
    __shared__ uint2 fix[6];
    __shared__ double ldelta[kMaxLabels];
    const uint64_t tile = tile_of_block(a);
    if (a.tile_map && tile >= a.n_tiles)
        return;  // spare workgroup of the XCD-striped grid (whole block leaves together)
    zbuild_table(a.zl, tab, fix);
    if (EMODE == 2 && threadIdx.x < kMaxLabels)
        ldelta[threadIdx.x] = (int)threadIdx.x < a.n_labels ? a.label_delta[threadIdx.x] : 1.0;
    __syncthreads();
    zwalk_tile<EMODE, K, WRAPFIX, RCACHE>(a, tile, a.partials ? a.partials + tile : nullptr, a.partial_stride, lds,
                                          tab, fix, ldelta);
}

// One thread per output word of the unfolded image.
__global__ __launch_bounds__(256) void zimage_build_kernel(const uint64_t *blocks, BrickLayout L, int W, int H, int D,
                                                          int mirror, ZLayout zl, uint32_t *out, uint64_t nwords)
{
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords)
        return;
    const uint64_t nbits = 1ull << zl.bits;
    uint32_t word = 0;
    for (uint32_t b = 0; b < 32; ++b) {
        const uint64_t A64 = w * 32u + b;
        uint32_t bit = 1;
        if (A64 < nbits) {
            const uint32_t A = (uint32_t)A64;
            const uint32_t ux = zextract(A, zl.m[0]), uy = zextract(A, zl.m[1]), uz = zextract(A, zl.m[2]);
            if (ux < zl.period[0] && uy < zl.period[1] && uz < zl.period[2]) {
                const int ix = zunfold(ux, W, mirror), iy = zunfold(uy, H, mirror), iz = zunfold(uz, D, mirror);
                bit = (uint32_t)(blocks[brick_block_index(ix, iy, iz, L)] >> bit_index(ix, iy, iz)) & 1u;
            }
        }
        word |= bit << b;
    }
    out[w] = word;
}

// One thread per output word (8 labels of 4 bits).
__global__ __launch_bounds__(256) void zlabels_build_kernel(const uint8_t *labels_blocked, BrickLayout L, int W, int H,
                                                           int D, int mirror, ZLayout zl, uint32_t *out,
                                                           uint64_t nwords)
{
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords)
        return;
    const uint64_t nbits = 1ull << zl.bits;
    uint32_t word = 0;
    for (uint32_t j = 0; j < 8; ++j) {
        const uint64_t A64 = w * 8u + j;
        if (A64 >= nbits)
            break;
        const uint32_t A = (uint32_t)A64;
        const uint32_t ux = zextract(A, zl.m[0]), uy = zextract(A, zl.m[1]), uz = zextract(A, zl.m[2]);
        if (ux < zl.period[0] && uy < zl.period[1] && uz < zl.period[2]) {
            const int ix = zunfold(ux, W, mirror), iy = zunfold(uy, H, mirror), iz = zunfold(uz, D, mirror);
            const uint32_t lab = labels_blocked[(uint64_t)brick_block_index(ix, iy, iz, L) * 64u + bit_index(ix, iy, iz)];
            word |= (lab & 0xFu) << (4u * j);
        }
    }
    out[w] = word;
}

__global__ __launch_bounds__(256) void zkeys_kernel(const int32_t *x, const int32_t *y, const int32_t *z, uint64_t n,
                                                   ZLayout zl, uint32_t *keys, uint32_t *iota)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    keys[i] = zaddress(x[i], y[i], z[i], zl) >> 5;
    iota[i] = (uint32_t)i;
}

hipError_t launch_zimage_build(const uint64_t *blocks, const BrickLayout &L, int W, int H, int D, int mirror,
                               const ZLayout &zl, uint32_t *out, hipStream_t st)
{
    const uint64_t nw = zimage_words(zl);
    hipLaunchKernelGGL(zimage_build_kernel, dim3((uint32_t)((nw + 255) / 256)), dim3(256), 0, st, blocks, L, W, H, D,
                       mirror, zl, out, nw);
    return hipGetLastError();
}

hipError_t launch_zlabels_build(const uint8_t *labels_blocked, const BrickLayout &L, int W, int H, int D, int mirror,
                                const ZLayout &zl, uint32_t *out, hipStream_t st)
{
    const uint64_t nw = zlabel_words(zl);
    hipLaunchKernelGGL(zlabels_build_kernel, dim3((uint32_t)((nw + 255) / 256)), dim3(256), 0, st, labels_blocked, L, W,
                       H, D, mirror, zl, out, nw);
    return hipGetLastError();
}

hipError_t launch_zkeys(const int32_t *x, const int32_t *y, const int32_t *z, uint64_t n, const ZLayout &zl,
                        uint32_t *keys, uint32_t *iota, hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(zkeys_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, x, y, z, n, zl, keys, iota);
    return hipGetLastError();
}

template <int EMODE, bool WRAPFIX, bool RCACHE>
static hipError_t launch_z(const WalkArgs &a, const LaunchShape &ls, uint32_t blocks, hipStream_t st)
{
    switch (ls.k) {
    case 1: hipLaunchKernelGGL((zwalk_kernel<EMODE, 1, WRAPFIX, RCACHE>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
    case 2: hipLaunchKernelGGL((zwalk_kernel<EMODE, 2, WRAPFIX, RCACHE>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
    case 4: hipLaunchKernelGGL((zwalk_kernel<EMODE, 4, WRAPFIX, RCACHE>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int EMODE>
static hipError_t launch_z_mode(const WalkArgs &a, const LaunchShape &ls, uint32_t blocks, hipStream_t st)
{
    if (a.zl.wrapfix)
        return a.zcache ? launch_z<EMODE, true, true>(a, ls, blocks, st) : launch_z<EMODE, true, false>(a, ls, blocks, st);
    return a.zcache ? launch_z<EMODE, false, true>(a, ls, blocks, st) : launch_z<EMODE, false, false>(a, ls, blocks, st);
}

hipError_t launch_zwalk(const WalkArgs &a, int emode, const LaunchShape &ls, hipStream_t st)
{
    if (a.n == 0)
        return hipSuccess;
    if (a.shift != 0 || a.zwords == nullptr || ls.threads > max_threads_for_k((int)ls.k) || ls.threads % 64 != 0 ||
        ls.threads == 0)
        return hipErrorInvalidValue;
    const uint64_t per_block = (uint64_t)ls.threads * ls.k;
    const uint64_t tiles = (a.n + per_block - 1) / per_block;
    if (tiles > 0x7FFFFFF0ull || (a.tile_map && a.n_tiles != tiles))
        return hipErrorInvalidValue;
    const uint32_t blocks = a.tile_map ? (uint32_t)(((tiles + kXcds - 1) / kXcds) * kXcds) : (uint32_t)tiles;
    if (emode == 0)
        return launch_z_mode<0>(a, ls, blocks, st);
    if (emode == 1)
        return launch_z_mode<1>(a, ls, blocks, st);
    if (emode == 2 && a.zlabels && a.label_delta && a.n_labels > 0 && a.n_labels <= kMaxLabels)
        return launch_z_mode<2>(a, ls, blocks, st);
    return hipErrorInvalidValue;
}

}  // namespace rwnmr
