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

template <int EMODE, int K, bool WRAPFIX, int CACHE, int TAB, bool NOFLUX = false, bool MFLAGS = false,
          bool ADISP = false, bool HITS = true>
__global__ __launch_bounds__(max_threads_for_k(K)) void zwalk_kernel(const WalkArgs a)
{
    __shared__ double lds[kEchoChunk][kEchoCols];
    __shared__ uint4 tab[TAB == 1 ? kZWideEntries
                                  : (TAB == 3 ? kZFlippedEntries : (MFLAGS ? 8 * kZFlagStride / 16 : (NOFLUX ? 12 : 6)))];
    __shared__ uint4 small[6];
    __shared__ uint2 fix[TAB == 3 ? kZFlippedEntries : (MFLAGS ? 8 * kZFlagStride / 16 : 6)];
    __shared__ uint32_t keys[TAB == 3 ? kZFlagKeyEntries / 4 : (TAB == 2 ? kZKeyEntries / 4 : 1)];
    __shared__ double ldelta[kMaxLabels];
    extern __shared__ uint32_t zlh[];   // EMODE 2, K == 1: per-label hit counts (blockDim x kLhStride)
    const uint64_t tile = tile_of_block(a);
    if (a.tile_map && tile >= a.n_tiles)
        return;  // spare workgroup of the XCD-striped grid (whole block leaves together)
    uint8_t *key_bytes = reinterpret_cast<uint8_t *>(keys);
    if (NOFLUX)
        zbuild_table_noflux(a.zl, tab);
    else if (TAB == 3)
        zbuild_table_flipped(a.zl, tab, fix);
    else if (MFLAGS)
        zbuild_table_flags(a.zl, tab, fix);
    else
        zbuild_table(a.zl, TAB == 1 ? small : tab, fix);
    if (TAB == 3)
        zbuild_flag_key_table(keys);
    if (TAB == 2)
        zbuild_key_table(key_bytes, NOFLUX ? 32u : 16u);
    if (EMODE == 2 && threadIdx.x < kMaxLabels)
        ldelta[threadIdx.x] = (int)threadIdx.x < a.n_labels ? a.label_delta[threadIdx.x] : 1.0;
    __syncthreads();
    if (TAB == 1) {
        zbuild_table_wide(small, tab);
        __syncthreads();
    }
    zwalk_tile<EMODE, K, WRAPFIX, CACHE, TAB, NOFLUX, MFLAGS, 1, false, ADISP, HITS>(
        a, tile, a.partials ? a.partials + tile : nullptr, a.partial_stride, lds, tab, fix, ldelta, key_bytes,
        EMODE == 2 && K == 1 ? zlh : nullptr);
}

// Speculative CPMG walk (zwalk_tile_spec): one walker per lane, for runs
// too small to hide the chunk-load latency of a step.
template <bool WRAPFIX, bool NOFLUX, bool MFLAGS>
__global__ __launch_bounds__(1024) void zwalk_spec_kernel(const WalkArgs a)
{
    __shared__ double lds[kEchoChunk][kEchoCols];
    __shared__ uint4 tab[MFLAGS ? 8 * kZFlagStride / 16 : (NOFLUX ? 12 : 6)];
    __shared__ uint2 fix[MFLAGS ? 8 * kZFlagStride / 16 : 6];
    __shared__ uint8_t keys[kZKeyEntries];
    const uint64_t tile = tile_of_block(a);
    if (a.tile_map && tile >= a.n_tiles)
        return;  // spare workgroup of the XCD-striped grid (whole block leaves together)
    if (NOFLUX)
        zbuild_table_noflux(a.zl, tab);
    else if (MFLAGS)
        zbuild_table_flags(a.zl, tab, fix);
    else
        zbuild_table(a.zl, tab, fix);
    zbuild_key_table(keys, NOFLUX ? 32u : 16u);
    __syncthreads();
    zwalk_tile_spec<WRAPFIX, NOFLUX, MFLAGS>(a, tile, a.partials ? a.partials + tile : nullptr, a.partial_stride,
                                             lds, tab, fix, keys);
}

// Label-set CPMG: 8 sets of label deltas evaluated on one trajectory (one
// walker per lane; zwalk_tile with NS = 8).  Unused sets (a.n_label_sets <
// 8) carry deltas of 1.
template <bool WRAPFIX, bool NOFLUX, bool MFLAGS>
__global__ __launch_bounds__(1024) void zwalk_sets_kernel(const WalkArgs a)
{
    extern __shared__ double lds8[];   // kEchoChunk x 8 sets x waves of the block
    __shared__ uint4 tab[MFLAGS ? 8 * kZFlagStride / 16 : (NOFLUX ? 12 : 6)];
    __shared__ uint2 fix[MFLAGS ? 8 * kZFlagStride / 16 : 6];
    __shared__ uint8_t keys[kZKeyEntries];
    __shared__ double ldelta[8 * kMaxLabels];
    const uint64_t tile = tile_of_block(a);
    if (a.tile_map && tile >= a.n_tiles)
        return;
    if (NOFLUX)
        zbuild_table_noflux(a.zl, tab);
    else if (MFLAGS)
        zbuild_table_flags(a.zl, tab, fix);
    else
        zbuild_table(a.zl, tab, fix);
    zbuild_key_table(keys, NOFLUX ? 32u : 16u);
    for (uint32_t t = threadIdx.x; t < 8u * kMaxLabels; t += blockDim.x) {
        const uint32_t q = t / kMaxLabels, l = t % kMaxLabels;
        ldelta[t] = (q < a.n_label_sets && (int)l < a.n_labels) ? a.label_delta[q * (uint32_t)a.n_labels + l] : 1.0;
    }
    __syncthreads();
    zwalk_tile<2, 1, WRAPFIX, 128, 2, NOFLUX, MFLAGS, 8>(a, tile, a.partials ? a.partials + tile : nullptr,
                                                         a.partial_stride, nullptr, tab, fix, ldelta, keys, nullptr,
                                                         lds8);
}

// Jacobian forward: the per-label CPMG walk (zwalk_tile EMODE 2, one walker
// per lane, per-label hit counts in LDS) that also sums h_l E per label and
// echo (zwalk_tile JAC).  Dynamic LDS: the 8-label echo staging (kEchoChunk
// x 8 x waves doubles), then the hit-count rows (blockDim x kLhStride).
template <bool WRAPFIX, bool NOFLUX, bool MFLAGS>
__global__ __launch_bounds__(1024) void zwalk_jac_kernel(const WalkArgs a)
{
    extern __shared__ double zjdyn[];
    __shared__ double lds[kEchoChunk][kEchoCols];
    __shared__ uint4 tab[MFLAGS ? 8 * kZFlagStride / 16 : (NOFLUX ? 12 : 6)];
    __shared__ uint2 fix[MFLAGS ? 8 * kZFlagStride / 16 : 6];
    __shared__ uint8_t keys[kZKeyEntries];
    __shared__ double ldelta[kMaxLabels];
    const uint64_t tile = tile_of_block(a);
    if (a.tile_map && tile >= a.n_tiles)
        return;
    double *lds8 = zjdyn;
    uint32_t *zlh = reinterpret_cast<uint32_t *>(zjdyn + (size_t)kEchoChunk * 8 * (blockDim.x / kWave));
    if (NOFLUX)
        zbuild_table_noflux(a.zl, tab);
    else if (MFLAGS)
        zbuild_table_flags(a.zl, tab, fix);
    else
        zbuild_table(a.zl, tab, fix);
    zbuild_key_table(keys, NOFLUX ? 32u : 16u);
    if (threadIdx.x < kMaxLabels)
        ldelta[threadIdx.x] = (int)threadIdx.x < a.n_labels ? a.label_delta[threadIdx.x] : 1.0;
    __syncthreads();
    zwalk_tile<2, 1, WRAPFIX, 128, 2, NOFLUX, MFLAGS, 1, true>(a, tile, a.partials ? a.partials + tile : nullptr,
                                                               a.partial_stride, lds, tab, fix, ldelta, keys, zlh,
                                                               lds8);
}

// PFGSE on the unfolded image: the walk between diffusion times counts hits
// (register chunk cache, packed displacement flushed every kZFlush steps and
// at every diffusion time), then pow(delta, hits) per energy set and the
// shared measurement (pfgse_measure_k).  K walkers per lane (walkers
// threadIdx.x + k blockDim.x of the block's blockDim.x K): their steps
// interleave, so each wave keeps K chunk reloads in flight, and their
// measurement terms are added in the lane before the cross-lane sums.
// NOFLUX: noflux walk on the periodic unfolded image (face test and
// turn-around, zmove_noflux); walkers must start inside the image.  MAXSETS:
// the most energy sets the instance handles (1, 2 or kMaxSets; registers).
// Two walkers per lane fit 168 VGPRs (3 waves per SIMD) for rotation k rows,
// up to two sets and the periodic / mirror walk; the other instances take
// more registers and 2 waves.
template <bool WRAPFIX, bool ROT, bool NOFLUX, int MAXSETS, int K>
__global__ __launch_bounds__(kPfgseThreads, (K > 1 && ROT && MAXSETS <= 2 && !NOFLUX) ? 3 : 1) void zpfgse_kernel(
    const WalkArgs a, const PfgseArgs p)
{
    __shared__ double lds[kMaxSets * kPfgseKChunk][kPfgseWaves];
    __shared__ uint4 tab[NOFLUX ? 12 : 6];
    __shared__ uint2 fix[6];
    __shared__ uint8_t keys[kZKeyEntries];
    if (NOFLUX)
        zbuild_table_noflux(a.zl, tab);
    else
        zbuild_table(a.zl, tab, fix);
    zbuild_key_table(keys, NOFLUX ? 32u : 16u);
    __syncthreads();
    const int nsets = p.n_sets > 0 ? (int)p.n_sets : 1;
    bool active[K];
    uint64_t id[K];
    int32_t X[K], Y[K], Z[K], x0[K], y0[K], z0[K], P[K];
    uint64_t S[K];
    uint32_t hits[K], A[K], QB[K], H[K];
    uint4 QV[K];
    double own_delta[K], e[K][MAXSETS];
    ZEntry T[K];
    const zrsrc_t zrsrc = zchunk_rsrc(a.zwords, a.zl);
    const char *tab_bytes = reinterpret_cast<const char *>(tab);
    const char *fix_bytes = reinterpret_cast<const char *>(fix);
    static_for<K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        bool inside;
        id[k] = pfgse_walker(p, blockIdx.x, threadIdx.x + (uint32_t)k * blockDim.x, blockDim.x * K, a.n, inside);
        active[k] = inside;
        uint64_t j = active[k] ? id[k] : 0;
        X[k] = active[k] ? a.x[j] : 0;
        Y[k] = active[k] ? a.y[j] : 0;
        Z[k] = active[k] ? a.z[j] : 0;
        if (NOFLUX && active[k] &&
            !((unsigned)X[k] < (unsigned)a.W && (unsigned)Y[k] < (unsigned)a.H && (unsigned)Z[k] < (unsigned)a.D)) {
            active[k] = false;   // a noflux walker must start inside the image
            atomicOr(a.error_flag, 1u);
            j = 0;
            X[k] = Y[k] = Z[k] = 0;
        }
        S[k] = active[k] ? a.seed[j] : 1ull;
        hits[k] = (active[k] && a.hits) ? a.hits[j] : 0u;
        x0[k] = active[k] ? p.x0[j] : 0;
        y0[k] = active[k] ? p.y0[j] : 0;
        z0[k] = active[k] ? p.z0[j] : 0;
        own_delta[k] = (active[k] && a.penalty) ? a.penalty[j] : a.delta_uniform;
        static_for<MAXSETS>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            e[k][q] = !active[k] ? 0.0 : ((q == 0 && a.energy) ? a.energy[j] : 1.0);
        });
        A[k] = zaddress(X[k], Y[k], Z[k], a.zl);
        QB[k] = 0xFFFFFFFFu;
        QV[k] = make_uint4(0u, 0u, 0u, 0u);
        P[k] = 0;
        // the next direction is drawn a step ahead, during the chunk load
        // (the state is rewound by one step at the end)
        zdirection<WRAPFIX, 2, NOFLUX>(S[k], tab_bytes, fix_bytes, T[k], keys);
    });
    uint32_t since = 0;
    auto flush = [&]() {
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            int32_t dx, dy, dz;
            zunpack(P[k], dx, dy, dz);
            X[k] += dx;
            Y[k] += dy;
            Z[k] += dz;
            P[k] = 0;
        });
        since = 0;
    };
    for (uint32_t i = 0; i < p.n_delta; ++i) {
        static_for<K>([&](auto kc) { H[decltype(kc)::value] = 0u; });
        uint32_t left = p.steps_inc[i];
        auto step = [&]() {
            uint32_t c[K];
            int32_t dp[K];
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                if (NOFLUX) {
                    c[k] = zmove_noflux(A[k], T[k], dp[k]);
                } else {
                    c[k] = zmove_entry<WRAPFIX>(A[k], T[k]);
                    dp[k] = (int32_t)T[k].dp;
                }
                zchunk_fill_buf(zrsrc, c[k], QB[k], QV[k]);
            });
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                zdirection<WRAPFIX, 2, NOFLUX>(S[k], tab_bytes, fix_bytes, T[k], keys);
            });
            // the table reads stay ahead of the wall tests, so their latency
            // overlaps the chunk loads'
            __builtin_amdgcn_sched_barrier(0);
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                // two short divergent paths instead of selects
                if (__builtin_expect(zchunk_wall<128>(c[k], QV[k]) != 0u, 0)) {
                    H[k] += 1u;
                } else {
                    A[k] = c[k];
                    P[k] += dp[k];
                }
            });
        };
        while (left) {
            const uint32_t ns = left < kZFlush - since ? left : kZFlush - since;
            // two steps per trip (the pipelined entry alternates registers)
            uint32_t st = 0;
            for (; st + 2 <= ns; st += 2) {
                step();
                step();
            }
            if (st < ns)
                step();
            since += ns;
            left -= ns;
            if (since == kZFlush)
                flush();
        }
        flush();
        double dX[K], dY[K], dZ[K];
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            hits[k] += H[k];
            static_for<MAXSETS>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if (q < nsets)
                    e[k][q] *= pow(p.n_sets > 0 ? p.set_delta[q] : own_delta[k], (double)H[k]);
            });
            dX[k] = (double)(X[k] - x0[k]);
            dY[k] = (double)(Y[k] - y0[k]);
            dZ[k] = (double)(Z[k] - z0[k]);
        });
        pfgse_measure_k<ROT, MAXSETS, K>(p, i, dX, dY, dZ, e, nsets, lds);
        if (p.msd_rows)
            pfgse_msd_k<K>(p, i, dX, dY, dZ, active, lds);
    }
    static_for<K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        S[k] = xorshift64_inverse(S[k]);
        if (active[k]) {
            a.x[id[k]] = X[k];
            a.y[id[k]] = Y[k];
            a.z[id[k]] = Z[k];
            a.seed[id[k]] = S[k];
            if (a.energy)
                a.energy[id[k]] = e[k][0];
            if (a.hits)
                a.hits[id[k]] = hits[k];
        }
    });
}

template <bool WRAPFIX, bool ROT, bool NOFLUX>
static void launch_zpfgse_sets(const WalkArgs &a, const PfgseArgs &p, dim3 g, dim3 t, hipStream_t st)
{
    if (p.n_sets <= 1)
        hipLaunchKernelGGL((zpfgse_kernel<WRAPFIX, ROT, NOFLUX, 1, kPfgseK>), g, t, 0, st, a, p);
    else if (p.n_sets <= 2)
        hipLaunchKernelGGL((zpfgse_kernel<WRAPFIX, ROT, NOFLUX, 2, kPfgseK>), g, t, 0, st, a, p);
    else
        hipLaunchKernelGGL((zpfgse_kernel<WRAPFIX, ROT, NOFLUX, kMaxSets, kPfgseK>), g, t, 0, st, a, p);
}

hipError_t launch_zpfgse(const WalkArgs &a, const PfgseArgs &p, hipStream_t st)
{
    if (a.n == 0)
        return hipSuccess;
    if (a.zwords == nullptr || a.shift != 0 || p.n_sets > (uint32_t)kMaxSets)
        return hipErrorInvalidValue;
    const dim3 g(pfgse_grid_blocks(p, a.n)), t(kPfgseThreads);
    if (a.znoflux) {
        if (p.k_rot)
            launch_zpfgse_sets<false, true, true>(a, p, g, t, st);
        else
            launch_zpfgse_sets<false, false, true>(a, p, g, t, st);
    } else if (a.zl.wrapfix) {
        if (p.k_rot)
            launch_zpfgse_sets<true, true, false>(a, p, g, t, st);
        else
            launch_zpfgse_sets<true, false, false>(a, p, g, t, st);
    } else {
        if (p.k_rot)
            launch_zpfgse_sets<false, true, false>(a, p, g, t, st);
        else
            launch_zpfgse_sets<false, false, false>(a, p, g, t, st);
    }
    return hipGetLastError();
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
        // descending order in each 64-bit pair of words: bit p of the pair
        // holds address 64 * pair + 63 - p
        const uint64_t A64 = (w >> 1) * 64u + 63u - 32u * (uint32_t)(w & 1u) - b;
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

// One thread per output word (4 labels of one byte, addresses 4w .. 4w + 3;
// with octant these are first-octant addresses, whose axis coordinates
// stay below the half periods).
__global__ __launch_bounds__(256) void zlabels_build_kernel(const uint8_t *labels_blocked, BrickLayout L, int W, int H,
                                                           int D, int mirror, ZLayout zl, uint32_t *out,
                                                           uint64_t nwords, uint64_t naddr)
{
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords)
        return;
    uint32_t word = 0;
    for (uint32_t j = 0; j < 4; ++j) {
        const uint64_t A64 = w * 4u + j;
        if (A64 >= naddr)
            break;
        const uint32_t A = (uint32_t)A64;
        const uint32_t ux = zextract(A, zl.m[0]), uy = zextract(A, zl.m[1]), uz = zextract(A, zl.m[2]);
        if (ux < zl.period[0] && uy < zl.period[1] && uz < zl.period[2]) {
            const int ix = zunfold(ux, W, mirror), iy = zunfold(uy, H, mirror), iz = zunfold(uz, D, mirror);
            const uint32_t lab = labels_blocked[(uint64_t)brick_block_index(ix, iy, iz, L) * 64u + bit_index(ix, iy, iz)];
            word |= (lab & 0xFFu) << (8u * j);
        }
    }
    out[w] = word;
}

// mflags: the key of the folded position (where a mirror walk with
// reflection flags starts the walker)
__global__ __launch_bounds__(256) void zkeys_kernel(const int32_t *x, const int32_t *y, const int32_t *z, uint64_t n,
                                                   ZLayout zl, int mflags, uint32_t *keys, uint32_t *iota)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int32_t xi = x[i], yi = y[i], zi = z[i];
    if (mflags) {
        xi = fold_mirror(xi, (int32_t)(zl.period[0] >> 1));
        yi = fold_mirror(yi, (int32_t)(zl.period[1] >> 1));
        zi = fold_mirror(zi, (int32_t)(zl.period[2] >> 1));
    }
    keys[i] = zaddress(xi, yi, zi, zl) >> 5;
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
                                const ZLayout &zl, bool octant, uint8_t *out, hipStream_t st)
{
    const uint64_t nw = zlabel_bytes(zl, octant) / 4;
    const uint64_t naddr = 1ull << (octant ? zl.bits - 3 : zl.bits);
    hipLaunchKernelGGL(zlabels_build_kernel, dim3((uint32_t)((nw + 255) / 256)), dim3(256), 0, st, labels_blocked, L, W,
                       H, D, mirror, zl, reinterpret_cast<uint32_t *>(out), nw, naddr);
    return hipGetLastError();
}

hipError_t launch_zkeys(const int32_t *x, const int32_t *y, const int32_t *z, uint64_t n, const ZLayout &zl,
                        int mflags, uint32_t *keys, uint32_t *iota, hipStream_t st)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(zkeys_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, x, y, z, n, zl, mflags, keys,
                       iota);
    return hipGetLastError();
}

template <int EMODE, bool WRAPFIX, int CACHE, int TAB, bool NOFLUX = false, bool MFLAGS = false>
static hipError_t launch_z(const WalkArgs &a, const LaunchShape &ls, uint32_t blocks, hipStream_t st)
{
    // mirror CPMG of one walker per lane on images whose periods all exceed
    // twice the flush interval: unbounded moves from address displacements
    if constexpr (EMODE == 1 && MFLAGS && TAB == 3 && !NOFLUX) {
        if (ls.k == 1 && zaddress_displacement_ok(a.zl)) {
            // hit counts only when the caller asked for them
            if (a.hits)
                hipLaunchKernelGGL((zwalk_kernel<EMODE, 1, WRAPFIX, CACHE, TAB, NOFLUX, MFLAGS, true>), dim3(blocks),
                                   dim3(ls.threads), 0, st, a);
            else
                hipLaunchKernelGGL((zwalk_kernel<EMODE, 1, WRAPFIX, CACHE, TAB, NOFLUX, MFLAGS, true, false>),
                                   dim3(blocks), dim3(ls.threads), 0, st, a);
            return hipGetLastError();
        }
    }
    // EMODE 2 with one walker per lane keeps the per-label hit counts in LDS
    const size_t lh_bytes = EMODE == 2 ? (size_t)kLhStride * ls.threads * sizeof(uint32_t) : 0;
    switch (ls.k) {
    case 1: hipLaunchKernelGGL((zwalk_kernel<EMODE, 1, WRAPFIX, CACHE, TAB, NOFLUX, MFLAGS>), dim3(blocks), dim3(ls.threads), lh_bytes, st, a); break;
    case 2: hipLaunchKernelGGL((zwalk_kernel<EMODE, 2, WRAPFIX, CACHE, TAB, NOFLUX, MFLAGS>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
    case 4: hipLaunchKernelGGL((zwalk_kernel<EMODE, 4, WRAPFIX, CACHE, TAB, NOFLUX, MFLAGS>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// a.zcache: 0 no chunk cache; 1 128-bit chunk with word select; 2 and 3 a
// 128-bit chunk with the shift test and the direction drawn one step ahead,
// through the byte-key table (2) or, for power-of-two periods, the wide
// direction table (3; the 6-entry table otherwise); 4 the speculative CPMG
// walk; 5 as 2 with a 64-bit (4 x 4 x 4 voxel) chunk, mirror walks with
// reflection flags (others run 2).
template <bool WRAPFIX, bool NOFLUX, bool MFLAGS>
static hipError_t launch_z_spec(const WalkArgs &a, const LaunchShape &ls, uint32_t blocks, hipStream_t st)
{
    hipLaunchKernelGGL((zwalk_spec_kernel<WRAPFIX, NOFLUX, MFLAGS>), dim3(blocks), dim3(ls.threads), 0, st, a);
    return hipGetLastError();
}

template <int EMODE>
static hipError_t launch_z_mode(const WalkArgs &a, const LaunchShape &ls, uint32_t blocks, hipStream_t st)
{
    // a.zcache 4: the speculative CPMG walk (one walker per lane); other
    // modes and lane shapes take the pipelined chunk-cache walk
    if (a.zcache == 4 && EMODE == 1 && ls.k == 1) {
        if (a.znoflux)
            return launch_z_spec<false, true, false>(a, ls, blocks, st);
        if (a.zmflags)
            return a.zl.wrapfix ? launch_z_spec<true, false, true>(a, ls, blocks, st)
                                : launch_z_spec<false, false, true>(a, ls, blocks, st);
        return a.zl.wrapfix ? launch_z_spec<true, false, false>(a, ls, blocks, st)
                            : launch_z_spec<false, false, false>(a, ls, blocks, st);
    }
    if (a.zcache == 4) {
        WalkArgs b = a;
        b.zcache = 2;
        return launch_z_mode<EMODE>(b, ls, blocks, st);
    }
    if (a.znoflux)   // noflux on the periodic image: pipelined byte-key walk only
        return a.zcache >= 2 ? launch_z<EMODE, false, 128, 2, true>(a, ls, blocks, st) : hipErrorInvalidValue;
    if (a.zmflags && a.zcache == 5) {   // mirror, 8-byte chunk cache
        if (EMODE < 2 && ls.threads >= 1024)
            return a.zl.wrapfix ? launch_z<EMODE, true, 64, 3, false, true>(a, ls, blocks, st)
                                : launch_z<EMODE, false, 64, 3, false, true>(a, ls, blocks, st);
        return a.zl.wrapfix ? launch_z<EMODE, true, 64, 2, false, true>(a, ls, blocks, st)
                            : launch_z<EMODE, false, 64, 2, false, true>(a, ls, blocks, st);
    }
    if (a.zcache == 5) {
        WalkArgs b = a;
        b.zcache = 2;
        return launch_z_mode<EMODE>(b, ls, blocks, st);
    }
    if (a.zmflags && a.zcache >= 2) {   // mirror, walkers folded to the first octant with reflection flags
        // CPMG and map walks of 1024-thread blocks read the flag-keyed
        // tables (58 KiB of LDS: two blocks per CU, 8 waves per SIMD)
        if (EMODE < 2 && ls.threads >= 1024)
            return a.zl.wrapfix ? launch_z<EMODE, true, 128, 3, false, true>(a, ls, blocks, st)
                                : launch_z<EMODE, false, 128, 3, false, true>(a, ls, blocks, st);
        return a.zl.wrapfix ? launch_z<EMODE, true, 128, 2, false, true>(a, ls, blocks, st)
                            : launch_z<EMODE, false, 128, 2, false, true>(a, ls, blocks, st);
    }
    if (a.zl.wrapfix) {
        switch (a.zcache) {
        case 0: return launch_z<EMODE, true, 0, 0>(a, ls, blocks, st);
        case 1: return launch_z<EMODE, true, 1, 0>(a, ls, blocks, st);
        case 2: return launch_z<EMODE, true, 128, 2>(a, ls, blocks, st);
        case 3: return launch_z<EMODE, true, 128, 0>(a, ls, blocks, st);
        default: return hipErrorInvalidValue;
        }
    }
    switch (a.zcache) {
    case 0: return launch_z<EMODE, false, 0, 0>(a, ls, blocks, st);
    case 1: return launch_z<EMODE, false, 1, 0>(a, ls, blocks, st);
    case 2: return launch_z<EMODE, false, 128, 2>(a, ls, blocks, st);
    case 3: return launch_z<EMODE, false, 128, 1>(a, ls, blocks, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_zwalk_sets(const WalkArgs &a, uint32_t threads, hipStream_t st)
{
    if (a.n == 0)
        return hipSuccess;
    if (a.shift != 0 || !a.zwords || !a.zlabels || !a.label_delta || a.n_labels <= 0 || a.n_labels > kMaxLabels ||
        a.n_label_sets == 0 || a.n_label_sets > 8 || threads == 0 || threads > 1024 || threads % 64 != 0)
        return hipErrorInvalidValue;
    const uint64_t tiles = (a.n + threads - 1) / threads;
    if (tiles > 0x7FFFFFF0ull || (a.tile_map && a.n_tiles != tiles))
        return hipErrorInvalidValue;
    const uint32_t blocks = a.tile_map ? (uint32_t)(((tiles + kXcds - 1) / kXcds) * kXcds) : (uint32_t)tiles;
    // the per-set echo staging, sized for this block's waves
    const size_t lds8_bytes = (size_t)kEchoChunk * 8 * (threads / kWave) * sizeof(double);
    if (a.znoflux) {
        hipLaunchKernelGGL((zwalk_sets_kernel<false, true, false>), dim3(blocks), dim3(threads), lds8_bytes, st, a);
    } else if (a.zmflags) {
        if (a.zl.wrapfix)
            hipLaunchKernelGGL((zwalk_sets_kernel<true, false, true>), dim3(blocks), dim3(threads), lds8_bytes, st, a);
        else
            hipLaunchKernelGGL((zwalk_sets_kernel<false, false, true>), dim3(blocks), dim3(threads), lds8_bytes, st,
                               a);
    } else if (a.zl.wrapfix) {
        hipLaunchKernelGGL((zwalk_sets_kernel<true, false, false>), dim3(blocks), dim3(threads), lds8_bytes, st, a);
    } else {
        hipLaunchKernelGGL((zwalk_sets_kernel<false, false, false>), dim3(blocks), dim3(threads), lds8_bytes, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_zwalk_jac(const WalkArgs &a, uint32_t threads, hipStream_t st)
{
    if (a.n == 0)
        return hipSuccess;
    if (a.shift != 0 || !a.zwords || !a.zlabels || !a.label_delta || !a.partials || !a.jac_partials ||
        a.n_labels <= 0 || a.n_labels > kMaxLabels || threads == 0 || threads > 1024 || threads % 64 != 0)
        return hipErrorInvalidValue;
    const uint64_t tiles = (a.n + threads - 1) / threads;
    if (tiles > 0x7FFFFFF0ull || (a.tile_map && a.n_tiles != tiles))
        return hipErrorInvalidValue;
    const uint32_t blocks = a.tile_map ? (uint32_t)(((tiles + kXcds - 1) / kXcds) * kXcds) : (uint32_t)tiles;
    const size_t dyn = (size_t)kEchoChunk * 8 * (threads / kWave) * sizeof(double) +
                       (size_t)kLhStride * threads * sizeof(uint32_t);
    if (a.znoflux) {
        hipLaunchKernelGGL((zwalk_jac_kernel<false, true, false>), dim3(blocks), dim3(threads), dyn, st, a);
    } else if (a.zmflags) {
        if (a.zl.wrapfix)
            hipLaunchKernelGGL((zwalk_jac_kernel<true, false, true>), dim3(blocks), dim3(threads), dyn, st, a);
        else
            hipLaunchKernelGGL((zwalk_jac_kernel<false, false, true>), dim3(blocks), dim3(threads), dyn, st, a);
    } else if (a.zl.wrapfix) {
        hipLaunchKernelGGL((zwalk_jac_kernel<true, false, false>), dim3(blocks), dim3(threads), dyn, st, a);
    } else {
        hipLaunchKernelGGL((zwalk_jac_kernel<false, false, false>), dim3(blocks), dim3(threads), dyn, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_zwalk(const WalkArgs &a, int emode, const LaunchShape &ls, hipStream_t st)
{
    if (a.n == 0)
        return hipSuccess;
    if (a.shift != 0 || a.zwords == nullptr || ls.threads > max_threads_for_k((int)ls.k) || ls.threads % 64 != 0 ||
        ls.threads == 0)
        return hipErrorInvalidValue;
    const uint64_t per_block = (uint64_t)ls.threads * ls.k;
    uint64_t tiles = (a.n + per_block - 1) / per_block;
    if (a.tile_walkers) {
        // balanced tiles: only for the address-displacement mirror walk, one
        // lane per walker of the tile
        if (!zadisp_walk(a, emode, ls) || a.tile_walkers > ls.threads)
            return hipErrorInvalidValue;
        tiles = (a.n + a.tile_walkers - 1) / a.tile_walkers;
    }
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
