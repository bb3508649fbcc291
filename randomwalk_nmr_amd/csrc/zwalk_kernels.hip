// zwalk_kernels.hip -- gfx950 CPMG / map walk for periodic and mirror
// boundaries on the unfolded image (layout and move arithmetic in
// walk_common.h, "Unfolded image with dilated voxel addresses").
//
// Per step and walker, all in VGPRs: xorshift64; s % 3 from the byte sum of
// the state (two v_sad_u8) and a 24-bit multiply; s % 2 from the low bit;
// the pair indexes a 6-entry direction table in LDS (axis mask, dilated
// addend, packed displacement); the dilated add gives the candidate voxel
// address with the periodic wrap included; one 4-byte gather of the word
// holding the voxel's bit (or, with the register chunk cache, a 16-byte
// gather only when the candidate leaves the walker's cached 8 x 4 x 4-voxel
// chunk, about one step in five); then either the move (address and packed
// displacement) or a wall hit (hit count + 1, energy *= delta, in that
// order as the reference).  The unbounded coordinates the reference reports
// are rebuilt from the packed displacement every kZFlush steps.  Nothing in
// the loop folds, divides or branches on the boundary.
//
// Reference semantics (paths under src/backend/ of RandomWalk_NMR):
//   CPMG_walk_{periodic,mirror}   B/NMR_Simulation/NMR_cpmg_cuda.cpp:300-402, 574-676
//   map_3D_{periodic,mirror}      B/NMR_Simulation/NMR_Simulation_cuda.cpp:418-721
//   image fold                    B/NMR_Simulation/NMR_cpmg_cuda.cpp:1991-2015
//   step / direction              B/Walker/walker.h:346-391, 498-507
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

// One tile of blockDim.x * K walkers; EMODE 0 = hit counting (map), 1 = CPMG
// with per-walker or uniform delta.
template <int EMODE, int K, bool WRAPFIX, bool RCACHE>
__global__ __launch_bounds__(max_threads_for_k(K)) void zwalk_kernel(const WalkArgs a)
{
    __shared__ double lds[kEchoChunk][kMaxWaves];
    __shared__ uint4 tab[6];
    __shared__ uint2 fix[6];
    const uint64_t tile = tile_of_block(a);
    if (a.tile_map && tile >= a.n_tiles)
        return;  // spare workgroup of the XCD-striped grid (whole block leaves together)
    if (threadIdx.x < 6) {
        const ZDir d = zdir_make(a.zl, threadIdx.x);
        tab[threadIdx.x] = make_uint4(d.m, d.add, (uint32_t)d.dp, d.nm);
        fix[threadIdx.x] = make_uint2(d.wrapv, d.fixv);
    }
    __syncthreads();

    const char *words = reinterpret_cast<const char *>(a.zwords);
    double *part = a.partials ? a.partials + tile : nullptr;
    const uint64_t base = tile * blockDim.x * K + threadIdx.x;
    const uint64_t n = a.n;
    const uint32_t echoes = a.echoes, spe = a.steps_per_echo;

    uint32_t A[K], C[K], Hc[K];
    uint32_t QB[K];   // RCACHE: 128-voxel chunk (address >> 7) held in QV
    uint4 QV[K];
    int32_t P[K], DP[K], X[K], Y[K], Z[K];
    uint64_t S[K];
    double E[K], DL[K];
    bool act[K];
    static_for<K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const uint64_t id = base + (uint64_t)k * blockDim.x;
        act[k] = id < n;
        const uint64_t j = act[k] ? id : 0;
        X[k] = act[k] ? a.x[j] : 0;
        Y[k] = act[k] ? a.y[j] : 0;
        Z[k] = act[k] ? a.z[j] : 0;
        S[k] = act[k] ? a.seed[j] : 1ull;
        E[k] = act[k] ? (a.energy ? a.energy[j] : 1.0) : 0.0;   // idle lanes add 0 to echo sums
        Hc[k] = (act[k] && a.hits) ? a.hits[j] : 0u;
        DL[k] = EMODE == 1 ? ((act[k] && a.penalty) ? a.penalty[j] : a.delta_uniform) : 1.0;
        A[k] = zaddress(X[k], Y[k], Z[k], a.zl);
        P[k] = 0;
        QB[k] = 0xFFFFFFFFu;   // no chunk: addresses >> 7 stay below 2^25
        QV[k] = make_uint4(0u, 0u, 0u, 0u);
    });

    auto echo_value = [&]() {
        double v = 0.0;
        static_for<K>([&](auto kc) { v += E[decltype(kc)::value]; });
        return v;
    };
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
    };
    const char *tab_bytes = reinterpret_cast<const char *>(tab);
    const char *fix_bytes = reinterpret_cast<const char *>(fix);
    auto step = [&]() {
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            S[k] = xorshift64(S[k]);
            const uint32_t lo = (uint32_t)S[k], hi = (uint32_t)(S[k] >> 32);
            // byte sum sb <= 2040 == state (mod 3); q = sb / 3 by a 24-bit multiply;
            // table entry 2 * (sb - 3q) + (state & 1) at 16 bytes per entry
            const uint32_t sb = __builtin_amdgcn_sad_u8(hi, 0u, __builtin_amdgcn_sad_u8(lo, 0u, 0u));
            const uint32_t q = __umul24(sb, 2731u) >> 13;
            const uint32_t off = (uint32_t)__mul24((int)q, -96) + ((sb << 5) | ((lo & 1u) << 4));
            const uint4 t = *reinterpret_cast<const uint4 *>(tab_bytes + off);
            const uint32_t m = t.x;
            uint32_t c = (A[k] & ~m) | (((A[k] | ~m) + t.y) & m);
            if (WRAPFIX) {
                const uint2 f = *reinterpret_cast<const uint2 *>(fix_bytes + (off >> 1));
                c = (c & m) == f.x ? ((c & ~m) | f.y) : c;
            }
            C[k] = c;
            DP[k] = (int32_t)t.z;
            // keep the displacement in this one table read (the compiler would
            // otherwise re-read it inside the move branch)
            asm volatile("" : "+v"(DP[k]));
        });
        uint32_t word[K];
        if (RCACHE) {
            // each walker keeps the 16-byte chunk (8 x 4 x 4 voxels) of its last
            // lookup; only candidates outside it read memory
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                const uint32_t cid = C[k] >> 7;
                if (cid != QB[k]) {
                    // c < 2^bits and the volume holds at least 4 words: in bounds
                    QV[k] = *reinterpret_cast<const uint4 *>(words + ((C[k] >> 3) & ~15u));
                    QB[k] = cid;
                }
            });
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                const bool up = (C[k] & 64u) != 0u;
                const uint32_t lo = up ? QV[k].z : QV[k].x, hi = up ? QV[k].w : QV[k].y;
                word[k] = (C[k] & 32u) ? hi : lo;
            });
        } else {
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                // c < 2^bits, so the byte offset of its word is inside the volume
                word[k] = *reinterpret_cast<const uint32_t *>(words + ((C[k] >> 3) & ~3u));
            });
        }
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const uint32_t wall = __builtin_amdgcn_ubfe(word[k], C[k], 1u);   // offset = C & 31
            A[k] = wall ? A[k] : C[k];
            P[k] += wall ? 0 : DP[k];
            Hc[k] += wall;
            if (EMODE == 1 && wall)
                E[k] *= DL[k];
        });
    };

    if (part)
        stage_echo(lds, 0, echoes, echo_value(), part, a.partial_stride);
    uint32_t since = 0;
    for (uint32_t echo = 1; echo <= echoes; ++echo) {
        uint32_t left = spe;
        while (left) {
            const uint32_t ns = left < kZFlush - since ? left : kZFlush - since;
            for (uint32_t i = 0; i < ns; ++i)
                step();
            since += ns;
            left -= ns;
            if (since == kZFlush) {
                flush();
                since = 0;
            }
        }
        if (part)
            stage_echo(lds, echo, echoes, echo_value(), part, a.partial_stride);
    }
    flush();

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
            if (a.hits)
                a.hits[id] = Hc[k];
        }
    });
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
    return hipErrorInvalidValue;
}

}  // namespace rwnmr
