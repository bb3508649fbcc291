// walk_device.h -- device-side pieces shared by the gfx950 walk kernels
// (plain CPMG/map/PFGSE launches and the batched persistent kernel): the
// walker registers, one step split into propose (RNG, direction, fold,
// block address) and commit, start validation and the FP64 wave / block
// echo reductions.  Step semantics are documented in walk_common.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_common.h"
#include "walk_kernels.h"

namespace rwnmr {

static constexpr int kWave = 64;
static constexpr int kEchoChunk = 32;   // echoes staged in LDS between block syncs
static constexpr int kMaxWaves = 16;    // up to 1024 threads per block

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
        v += __shfl_xor(v, off, kWave);
    return v;
}

// Registers of one walker.  ix/iy/iz are the image-voxel (folded)
// coordinates, mx/my/mz the mirror reflection flags.
struct Walker {
    int x, y, z;
    int ix, iy, iz;
    int mx, my, mz;
    uint64_t s;
    double e;
    double delta;
    uint32_t h;
    int active;
};

template <int BC>
__device__ __forceinline__ void init_fold(const WalkArgs &a, Walker &w)
{
    if (BC == kMirror) {
        w.ix = fold_mirror(w.x, a.W);
        w.iy = fold_mirror(w.y, a.H);
        w.iz = fold_mirror(w.z, a.D);
        w.mx = mirror_flag(w.x, a.W);
        w.my = mirror_flag(w.y, a.H);
        w.mz = mirror_flag(w.z, a.D);
    } else if (BC == kPeriodic) {
        w.ix = fold_periodic(w.x, a.W);
        w.iy = fold_periodic(w.y, a.H);
        w.iz = fold_periodic(w.z, a.D);
        w.mx = w.my = w.mz = 0;
    } else {
        w.ix = w.x;
        w.iy = w.y;
        w.iz = w.z;
        w.mx = w.my = w.mz = 0;
    }
}

// Candidate image voxel of the next move plus the move itself.
struct Move {
    int ax;     // axis 0/1/2
    int sg;     // +-1
    int nix, niy, niz;
    int nm;     // new reflection flag of the moving axis
    uint32_t blk, bit;
};

template <int BC>
__device__ __forceinline__ Move propose(const WalkArgs &a, Walker &w)
{
    Move mv;
    w.s = xorshift64(w.s);
    const uint32_t r = mod6(w.s);
    mv.ax = move_axis(r);
    int sg = move_sign(r);
    const int i = mv.ax == 0 ? w.ix : (mv.ax == 1 ? w.iy : w.iz);
    const int d = mv.ax == 0 ? a.W : (mv.ax == 1 ? a.H : a.D);
    int ni, nm = 0;
    if (BC == kMirror) {
        const int m = mv.ax == 0 ? w.mx : (mv.ax == 1 ? w.my : w.mz);
        mirror_step(i, m, sg, d, ni, nm);
    } else if (BC == kPeriodic) {
        ni = periodic_step(i, sg, d);
    } else {
        // noflux: an outward move at a face turns around before the wall test
        if ((sg < 0 && i == 0) || (sg > 0 && i == d - 1))
            sg = -sg;
        ni = i + sg;
    }
    mv.sg = sg;
    mv.nm = nm;
    mv.nix = mv.ax == 0 ? ni : w.ix;
    mv.niy = mv.ax == 1 ? ni : w.iy;
    mv.niz = mv.ax == 2 ? ni : w.iz;
    mv.blk = block_index(mv.nix, mv.niy, mv.niz, a.bC, a.bCR);
    mv.bit = bit_index(mv.nix, mv.niy, mv.niz);
    return mv;
}

template <int BC>
__device__ __forceinline__ void commit(Walker &w, const Move &mv)
{
    w.x += mv.ax == 0 ? mv.sg : 0;
    w.y += mv.ax == 1 ? mv.sg : 0;
    w.z += mv.ax == 2 ? mv.sg : 0;
    w.ix = mv.nix;
    w.iy = mv.niy;
    w.iz = mv.niz;
    if (BC == kMirror) {
        w.mx = mv.ax == 0 ? mv.nm : w.mx;
        w.my = mv.ax == 1 ? mv.nm : w.my;
        w.mz = mv.ax == 2 ? mv.nm : w.mz;
    }
}

// Validity of a walker's start position: noflux walkers must start inside
// the image (the fold keeps periodic/mirror lookups in range for any start).
template <int BC>
__device__ __forceinline__ int start_ok(const WalkArgs &a, const Walker &w)
{
    if (BC != kNoflux)
        return 1;
    return (unsigned)w.x < (unsigned)a.W && (unsigned)w.y < (unsigned)a.H && (unsigned)w.z < (unsigned)a.D;
}

// Stage one echo's block sum; every kEchoChunk echoes (and at the last echo)
// the block writes its partials out[row * row_stride] for the staged rows.
// All threads of the block must call this.
__device__ __forceinline__ void stage_echo(double (*lds)[kMaxWaves], uint32_t echo, uint32_t last_echo, double v,
                                           double *out, uint64_t row_stride)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    v = wave_sum(v);
    const int slot = echo % kEchoChunk;
    if (lane == 0)
        lds[slot][wave] = v;
    if (slot == kEchoChunk - 1 || echo == last_echo) {
        __syncthreads();
        const int cnt = slot + 1;
        if ((int)threadIdx.x < cnt) {
            double s = 0.0;
            for (int q = 0; q < nw; ++q)
                s += lds[threadIdx.x][q];
            const uint64_t row = (uint64_t)(echo - slot) + threadIdx.x;
            out[row * row_stride] = s;
        }
        __syncthreads();
    }
}

// One tile of walkers (blockDim.x * K consecutive walkers starting at
// tile * blockDim.x * K) walked for a.echoes x a.steps_per_echo steps.
// EMODE: 0 = hit counting only (map), 1 = per-walker delta, 2 = per-label
// delta.  Echo partials go to part[row * part_stride] (nullable).
template <int BC, int EMODE, int K>
__device__ __forceinline__ void walk_tile(const WalkArgs &a, uint64_t tile, double *part, uint64_t part_stride,
                                          double (*lds)[kMaxWaves])
{
    const uint64_t base = tile * blockDim.x * K + threadIdx.x;
    Walker w[K];
    uint32_t lhits[K][kMaxLabels];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint64_t id = base + (uint64_t)k * blockDim.x;
        w[k].active = id < a.n;
        const uint64_t j = w[k].active ? id : 0;
        w[k].x = w[k].active ? a.x[j] : 0;
        w[k].y = w[k].active ? a.y[j] : 0;
        w[k].z = w[k].active ? a.z[j] : 0;
        w[k].s = w[k].active ? a.seed[j] : 1;
        w[k].e = (w[k].active && a.energy) ? a.energy[j] : 1.0;
        w[k].h = (w[k].active && a.hits) ? a.hits[j] : 0u;
        w[k].delta = EMODE == 1 ? ((w[k].active && a.penalty) ? a.penalty[j] : a.delta_uniform) : 1.0;
        if (w[k].active && !start_ok<BC>(a, w[k])) {
            w[k].active = 0;
            atomicOr(a.error_flag, 1u);
        }
        if (!w[k].active) {
            w[k].x = w[k].y = w[k].z = 0;
            w[k].e = 0.0;
        }
        init_fold<BC>(a, w[k]);
#pragma unroll
        for (int l = 0; l < kMaxLabels; ++l)
            lhits[k][l] = (EMODE == 2 && a.label_hits && w[k].active && l < a.n_labels)
                              ? a.label_hits[(uint64_t)l * a.n + j]
                              : 0u;
    }

    if (part) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k)
            v += w[k].active ? w[k].e : 0.0;
        stage_echo(lds, 0, a.echoes, v, part, part_stride);
    }

    for (uint32_t echo = 1; echo <= a.echoes; ++echo) {
        for (uint32_t st = 0; st < a.steps_per_echo; ++st) {
            Move mv[K];
            uint64_t word[K];
#pragma unroll
            for (int k = 0; k < K; ++k)
                mv[k] = propose<BC>(a, w[k]);
#pragma unroll
            for (int k = 0; k < K; ++k)
                word[k] = w[k].active ? a.blocks[mv[k].blk] : 0ull;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int wall = (int)((word[k] >> mv[k].bit) & 1ull);
                if (!wall) {
                    commit<BC>(w[k], mv[k]);
                } else {
                    w[k].h += 1u;
                    if (EMODE == 1) {
                        w[k].e *= w[k].delta;
                    } else if (EMODE == 2) {
                        const uint32_t lab = a.labels_blocked[(uint64_t)mv[k].blk * 64u + mv[k].bit];
                        w[k].e *= a.label_delta[lab];
#pragma unroll
                        for (int l = 0; l < kMaxLabels; ++l)
                            lhits[k][l] += (lab == (uint32_t)l) ? 1u : 0u;
                    }
                }
            }
        }
        if (part) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < K; ++k)
                v += w[k].active ? w[k].e : 0.0;
            stage_echo(lds, echo, a.echoes, v, part, part_stride);
        }
    }

#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (!w[k].active)
            continue;
        const uint64_t id = base + (uint64_t)k * blockDim.x;
        a.x[id] = w[k].x;
        a.y[id] = w[k].y;
        a.z[id] = w[k].z;
        a.seed[id] = w[k].s;
        if (a.energy)
            a.energy[id] = w[k].e;
        if (a.hits)
            a.hits[id] = w[k].h;
        if (EMODE == 2 && a.label_hits)
#pragma unroll
            for (int l = 0; l < kMaxLabels; ++l)
                if (l < a.n_labels)
                    a.label_hits[(uint64_t)l * a.n + id] = lhits[k][l];
    }
}

}  // namespace rwnmr
