// zwalk_device.h -- the unfolded-image walk of one tile of walkers, shared
// by the single-instance kernel (zwalk_kernels.hip) and the batched
// persistent kernel.  Layout and move arithmetic: walk_common.h, "Unfolded
// image with dilated voxel addresses".
//
// Per step and walker, all in VGPRs: xorshift64; s % 3 from the byte sum of
// the state (two v_sad_u8) and a 24-bit multiply; s % 2 from the low bit;
// the pair indexes a 6-entry direction table in LDS (axis mask, dilated
// addend, packed displacement); the dilated add gives the candidate voxel
// address with the periodic wrap included; then the wall bit: with the
// register chunk cache (RCACHE) each walker keeps the 16-byte chunk
// (8 x 4 x 4 voxels) of its last lookup and reads memory only when the
// candidate leaves it (about one step in five), else one 4-byte gather per
// step; then either the move (address and packed displacement) or a wall hit
// (hit count + 1, energy *= delta, in that order as the reference).  The
// unbounded coordinates the reference reports are rebuilt from the packed
// displacement every kZFlush steps.  Nothing in the loop folds, divides or
// branches on the boundary.
//
// Reference semantics (paths under src/backend/ of RandomWalk_NMR):
//   CPMG_walk_{periodic,mirror}   B/NMR_Simulation/NMR_cpmg_cuda.cpp:300-402, 574-676
//   map_3D_{periodic,mirror}      B/NMR_Simulation/NMR_Simulation_cuda.cpp:418-721
//   image fold                    B/NMR_Simulation/NMR_cpmg_cuda.cpp:1991-2015
//   step / direction              B/Walker/walker.h:346-391, 498-507
#pragma once
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

// Direction table of layout zl in LDS (threads 0..5 write; the caller syncs).
__device__ __forceinline__ void zbuild_table(const ZLayout &zl, uint4 *tab, uint2 *fix)
{
    if (threadIdx.x < 6) {
        const ZDir d = zdir_make(zl, threadIdx.x);
        tab[threadIdx.x] = make_uint4(d.m, d.add, (uint32_t)d.dp, d.nm);
        fix[threadIdx.x] = make_uint2(d.wrapv, d.fixv);
    }
}

// Candidate address of one move from A: advances the RNG state s, reads the
// direction entry (16 bytes per entry in tab_bytes, 8 in fix_bytes) and
// returns the dilated-add result; dp receives the packed displacement.
template <bool WRAPFIX>
__device__ __forceinline__ uint32_t zpropose(uint64_t &s, uint32_t A, const char *tab_bytes, const char *fix_bytes,
                                             int32_t &dp)
{
    s = xorshift64(s);
    const uint32_t lo = (uint32_t)s, hi = (uint32_t)(s >> 32);
    // byte sum sb <= 2040 == state (mod 3); q = sb / 3 by a 24-bit multiply;
    // table entry 2 * (sb - 3q) + (state & 1) at 16 bytes per entry
    const uint32_t sb = __builtin_amdgcn_sad_u8(hi, 0u, __builtin_amdgcn_sad_u8(lo, 0u, 0u));
    const uint32_t q = __umul24(sb, 2731u) >> 13;
    const uint32_t off = (uint32_t)__mul24((int)q, -96) + ((sb << 5) | ((lo & 1u) << 4));
    const uint4 t = *reinterpret_cast<const uint4 *>(tab_bytes + off);
    const uint32_t m = t.x;
    uint32_t c = (A & ~m) | (((A | ~m) + t.y) & m);
    if (WRAPFIX) {
        const uint2 f = *reinterpret_cast<const uint2 *>(fix_bytes + (off >> 1));
        c = (c & m) == f.x ? ((c & ~m) | f.y) : c;
    }
    dp = (int32_t)t.z;
    // keep the displacement in this one table read (the compiler would
    // otherwise re-read it inside the move branch)
    asm volatile("" : "+v"(dp));
    return c;
}

// Register chunk cache: make QV hold the 16-byte chunk of address c (read
// from words only when c left the chunk QB).  c < 2^bits and the volume holds
// at least 4 words, so the read is in bounds.
__device__ __forceinline__ void zchunk_fill(const char *words, uint32_t c, uint32_t &QB, uint4 &QV)
{
    const uint32_t cid = c >> 7;
    if (cid != QB) {
        QV = *reinterpret_cast<const uint4 *>(words + ((c >> 3) & ~15u));
        QB = cid;
    }
}

// The 32-bit word of QV that holds address c's bit.
__device__ __forceinline__ uint32_t zchunk_word(uint32_t c, const uint4 &QV)
{
    const bool up = (c & 64u) != 0u;
    const uint32_t lo = up ? QV.z : QV.x, hi = up ? QV.w : QV.y;
    return (c & 32u) ? hi : lo;
}

// Solid label (4 bits per voxel, same address order) of address c.
__device__ __forceinline__ uint32_t zlabel(const uint32_t *zlabels, uint32_t c)
{
    return (zlabels[c >> 3] >> ((c & 7u) * 4u)) & 0xFu;
}

// One tile (blockDim.x * K consecutive walkers from tile * blockDim.x * K)
// walked for a.echoes x a.steps_per_echo steps on the unfolded image
// a.zwords / a.zl, with the direction table already in LDS.  EMODE 0 = hit
// counting (map), 1 = CPMG with per-walker or uniform delta, 2 = per-label
// delta (ldelta: the label deltas in LDS; labels from a.zlabels; per-label
// hit counts in a.label_hits).  Echo partials go to part[row * part_stride]
// (nullable).  All threads of the block call it.
template <int EMODE, int K, bool WRAPFIX, bool RCACHE>
__device__ __forceinline__ void zwalk_tile(const WalkArgs &a, uint64_t tile, double *part, uint64_t part_stride,
                                           double (*lds)[kMaxWaves], const uint4 *tab, const uint2 *fix,
                                           const double *ldelta = nullptr)
{
    constexpr int NL = EMODE == 2 ? kMaxLabels : 1;
    const char *words = reinterpret_cast<const char *>(a.zwords);
    const uint64_t base = tile * blockDim.x * K + threadIdx.x;
    const uint64_t n = a.n;
    const uint32_t echoes = a.echoes, spe = a.steps_per_echo;
    const int n_labels = a.n_labels;

    uint32_t A[K], C[K], Hc[K];
    uint32_t QB[K];   // RCACHE: 128-voxel chunk (address >> 7) held in QV
    uint4 QV[K];
    int32_t P[K], DP[K], X[K], Y[K], Z[K];
    uint64_t S[K];
    double E[K], DL[K];
    uint32_t LH[K][NL];
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
        static_for<NL>([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            LH[k][l] = (EMODE == 2 && a.label_hits && act[k] && l < n_labels) ? a.label_hits[(uint64_t)l * n + j]
                                                                               : 0u;
        });
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
            C[k] = zpropose<WRAPFIX>(S[k], A[k], tab_bytes, fix_bytes, DP[k]);
        });
        uint32_t word[K];
        if (RCACHE) {
            // each walker keeps the 16-byte chunk (8 x 4 x 4 voxels) of its last
            // lookup; only candidates outside it read memory
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                zchunk_fill(words, C[k], QB[k], QV[k]);
            });
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                word[k] = zchunk_word(C[k], QV[k]);
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
            if (EMODE == 2 && wall) {
                const uint32_t lab = zlabel(a.zlabels, C[k]);
                E[k] *= ldelta[lab];
                static_for<NL>([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    LH[k][l] += lab == (uint32_t)l ? 1u : 0u;
                });
            }
        });
    };

    if (part)
        stage_echo(lds, 0, echoes, echo_value(), part, part_stride);
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
            stage_echo(lds, echo, echoes, echo_value(), part, part_stride);
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
            if (EMODE == 2 && a.label_hits)
                static_for<NL>([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    if (l < n_labels)
                        a.label_hits[(uint64_t)l * n + id] = LH[k][l];
                });
        }
    });
}

}  // namespace rwnmr
