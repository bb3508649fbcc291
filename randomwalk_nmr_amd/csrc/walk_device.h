// walk_device.h -- device-side pieces shared by the gfx950 walk kernels
// (plain CPMG/map/PFGSE launches, the batched persistent kernel and the
// backward replay): the walker registers, one step split into propose (RNG,
// direction, fold, block address) and commit, start validation and the FP64
// wave / block echo reductions.  Step semantics are documented in
// walk_common.h.
//
// Register discipline: a lane's K walkers live in arrays indexed only by
// compile-time constants (static_for), and the image geometry is copied into
// plain values (Geom) once, so every walker field stays in a VGPR and the
// dimensions in SGPRs.  (Conditional reads of struct fields selected by a
// runtime value are written on values, never on addresses; otherwise the
// compiler folds them into address selects and demotes the walker to
// LDS/scratch.)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "walk_common.h"
#include "walk_kernels.h"

namespace rwnmr {

static constexpr int kWave = 64;
static constexpr int kEchoChunk = 16;   // echoes staged in LDS between block syncs
static constexpr int kMaxWaves = 16;    // up to 1024 threads per block
// one staged echo row: a partial per half-wave, plus one pad column so the
// rows of consecutive echoes start on different LDS banks
static constexpr int kEchoCols = 2 * kMaxWaves + 1;

// f(integral_constant<0>), ..., f(integral_constant<N-1>), fully inlined.
template <int N, class F> __device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (N > 0) {
        static_for<N - 1>(f);
        f(std::integral_constant<int, N - 1>{});
    }
}

// v + v of lane ^ 16 (ROW16) or lane ^ 32, by the gfx950 permlane swaps:
// with both operands v, the swap returns v with rows 1 and 3 (lanes 32-63)
// replaced by rows 0 and 2 (lanes 0-31), and v with rows 0 and 2 (lanes
// 0-31) replaced by rows 1 and 3 (lanes 32-63); their sum is the butterfly
// sum, with the lower lane's value first in both lanes of a pair.
template <bool ROW16>
__device__ __forceinline__ double xor_rows_sum(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const auto l = ROW16 ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                         : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = ROW16 ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                         : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}

template <int CTRL, int ROWS> __device__ __forceinline__ double dpp_f64(double v);

// Sum of v over the 64 lanes, the same in every lane: lane ^ 1, lane ^ 2,
// half-row mirror and row mirror by DPP, then lane ^ 16 and lane ^ 32 by the
// permlane swaps (18 VALU, no LDS).  Fixed order, so deterministic.  The
// whole wave must be active.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0xB1, 0xF>(v);
    v += dpp_f64<0x4E, 0xF>(v);
    v += dpp_f64<0x141, 0xF>(v);
    v += dpp_f64<0x140, 0xF>(v);
    v = xor_rows_sum<true>(v);
    return xor_rows_sum<false>(v);
}

// v from the lane given by the DPP control CTRL (row_mirror 0x140: lane
// 15 - i of the row; row_half_mirror 0x141: lane 7 - i of the half row;
// quad_perm 0x4E: lane ^ 2; quad_perm 0xB1: lane ^ 1; row_bcast:15 0x142:
// lane 15 of the previous row, written only in the rows ROWS selects), both
// 32-bit halves.  Every source lane of these controls exists, so no old value
// is needed (the unwritten rows of row_bcast:15 hold junk).  Callers run
// with the whole wave active.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, CTRL, ROWS, 0xF, true),
                            __builtin_amdgcn_mov_dpp(lo, CTRL, ROWS, 0xF, true));
}

// Sums of v over the half-waves, lanes 0-31 in lane 31 and lanes 32-63 in
// lane 63 (other lanes hold partial sums): the 16 lanes of each row by four
// DPP steps (lane ^ 1, lane ^ 2, half-row mirror, row mirror), then rows 0
// and 2 into rows 1 and 3 (row_bcast:15).  Fixed order, so deterministic;
// 15 VALU and no LDS round trips.
__device__ __forceinline__ double half_wave_sum(double v)
{
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    return v + dpp_f64<0x142, 0xA>(v);
}

// One butterfly step of the reduce-scatter: items [0, n) and [n, 2n) of v;
// the lane keeps the half its bit selects and adds the other half of the
// partner lane (which has the opposite bit and the same earlier bits).
template <int CTRL, int N>
__device__ __forceinline__ void wave_halve(double *v, bool bit)
{
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double keep = bit ? v[i + N] : v[i];
        const double send = bit ? v[i] : v[i + N];
        v[i] = keep + dpp_f64<CTRL>(send);
    }
}

// Lanes l and l ^ 32 (ROW16 false) or l ^ 16 (ROW16 true) both hold x and
// y: the lane with that bit clear returns x summed over the pair, the lane
// with it set y summed over the pair, the lower lane's value first in both.
// One gfx950 permlane swap per 32-bit half moves both values (the swap
// exchanges the upper lanes of its first operand with the lower lanes of
// its second), so a halving step costs 3 VALU and no select.
template <bool ROW16>
__device__ __forceinline__ double swap_halve(double x, double y)
{
    const int xl = __double2loint(x), xh = __double2hiint(x), yl = __double2loint(y), yh = __double2hiint(y);
    const auto l = ROW16 ? __builtin_amdgcn_permlane16_swap(xl, yl, false, false)
                         : __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
    const auto h = ROW16 ? __builtin_amdgcn_permlane16_swap(xh, yh, false, false)
                         : __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
    return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}

// Sums over the 64 lanes of N = 4, 8 or 16 values per lane at once (a
// reduce-scatter): halvings across lane ^ 32 and lane ^ 16 by swap_halve,
// then (N >= 8) across bit 3 by the row mirror and (N == 16) bit 2 by the
// half-row mirror with selects, then plain DPP sums over the remaining lane
// bits.  Lane l returns the wave total of value l >> (6 - log2 N): lanes
// 16q (N = 4), 8q (N = 8) or 4q (N = 16) hold value q.  About 5.3, 4.3 and
// 3.9 VALU per value.  Fixed order, so deterministic.  The whole wave must
// be active; v is overwritten.
template <int N>
__device__ __forceinline__ double wave_reduce_scatter(double (&v)[N])
{
    static_assert(N == 4 || N == 8 || N == 16, "4, 8 or 16 values");
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int i = 0; i < N / 2; ++i)
        v[i] = swap_halve<false>(v[i], v[i + N / 2]);   // bit 5
#pragma unroll
    for (int i = 0; i < N / 4; ++i)
        v[i] = swap_halve<true>(v[i], v[i + N / 4]);    // bit 4
    if constexpr (N >= 8)
        wave_halve<0x140, N / 8>(v, (lane & 8) != 0);   // row mirror: bit 3
    if constexpr (N >= 16)
        wave_halve<0x141, N / 16>(v, (lane & 4) != 0);  // half-row mirror: bit 2
    double r = v[0];
    r += dpp_f64<0xB1>(r);                              // lane ^ 1
    r += dpp_f64<0x4E>(r);                              // lane ^ 2
    if constexpr (N < 16)
        r += dpp_f64<0x141>(r);                         // half-row mirror
    if constexpr (N < 8)
        r += dpp_f64<0x140>(r);                         // row mirror
    return r;
}

// Image geometry as plain values (uniform across the grid).
struct Geom {
    const uint64_t *blocks;
    int W, H, D;
    BrickLayout L;
};

__device__ __forceinline__ Geom make_geom(const WalkArgs &a)
{
    Geom g;
    g.blocks = a.blocks;
    g.W = a.W;
    g.H = a.H;
    g.D = a.D;
    g.L = a.layout;
    return g;
}

// Tile walked by this workgroup.  Workgroups are dealt round-robin to the
// XCDs (workgroup b runs on XCD b % 8), so with tile_map == 1 XCD j walks the
// j-th contiguous eighth of the (position-sorted) tiles and its L2 holds one
// compact region of the image.  Returns n_tiles or more for a spare block.
__device__ __forceinline__ uint64_t tile_of_block(const WalkArgs &a)
{
    if (!a.tile_map)
        return blockIdx.x;
    const uint32_t per_xcd = (a.n_tiles + kXcds - 1) / kXcds;
    return (uint64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
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

__device__ __forceinline__ int pick3(int ax, int a0, int a1, int a2)
{
    const int t = ax == 1 ? a1 : a2;
    return ax == 0 ? a0 : t;
}

template <int BC>
__device__ __forceinline__ void init_fold(const Geom &g, Walker &w)
{
    if (BC == kMirror) {
        w.ix = fold_mirror(w.x, g.W);
        w.iy = fold_mirror(w.y, g.H);
        w.iz = fold_mirror(w.z, g.D);
        w.mx = mirror_flag(w.x, g.W);
        w.my = mirror_flag(w.y, g.H);
        w.mz = mirror_flag(w.z, g.D);
    } else if (BC == kPeriodic) {
        w.ix = fold_periodic(w.x, g.W);
        w.iy = fold_periodic(w.y, g.H);
        w.iz = fold_periodic(w.z, g.D);
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
    int ni;     // new folded coordinate of the moving axis
    int nm;     // new reflection flag of the moving axis
    uint32_t blk, bit;
};

template <int BC>
__device__ __forceinline__ Move propose(const Geom &g, Walker &w)
{
    Move mv;
    w.s = xorshift64(w.s);
    const uint32_t r = mod6(w.s);
    mv.ax = move_axis(r);
    int sg = move_sign(r);
    const int ix = w.ix, iy = w.iy, iz = w.iz;
    const int i = pick3(mv.ax, ix, iy, iz);
    const int d = pick3(mv.ax, g.W, g.H, g.D);
    int ni, nm = 0;
    if (BC == kMirror) {
        const int m = pick3(mv.ax, w.mx, w.my, w.mz);
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
    mv.ni = ni;
    mv.nm = nm;
    const int nix = mv.ax == 0 ? ni : ix;
    const int niy = mv.ax == 1 ? ni : iy;
    const int niz = mv.ax == 2 ? ni : iz;
    mv.blk = brick_block_index(nix, niy, niz, g.L);
    mv.bit = bit_index(nix, niy, niz);
    return mv;
}

template <int BC>
__device__ __forceinline__ void commit(Walker &w, const Move &mv)
{
    const int dx = mv.ax == 0 ? mv.sg : 0, dy = mv.ax == 1 ? mv.sg : 0, dz = mv.ax == 2 ? mv.sg : 0;
    w.x += dx;
    w.y += dy;
    w.z += dz;
    w.ix = mv.ax == 0 ? mv.ni : w.ix;
    w.iy = mv.ax == 1 ? mv.ni : w.iy;
    w.iz = mv.ax == 2 ? mv.ni : w.iz;
    if (BC == kMirror) {
        w.mx = mv.ax == 0 ? mv.nm : w.mx;
        w.my = mv.ax == 1 ? mv.nm : w.my;
        w.mz = mv.ax == 2 ? mv.nm : w.mz;
    }
}

// Validity of a walker's start position: noflux walkers must start inside
// the image (the fold keeps periodic/mirror lookups in range for any start).
template <int BC>
__device__ __forceinline__ int start_ok(const Geom &g, const Walker &w)
{
    if (BC != kNoflux)
        return 1;
    return (unsigned)w.x < (unsigned)g.W && (unsigned)w.y < (unsigned)g.H && (unsigned)w.z < (unsigned)g.D;
}

// Stage one echo's block sum; every kEchoChunk echoes (and at the last echo)
// the block writes its partials out[row * row_stride] for the staged rows.
// Each half-wave stages its own partial, so a wave spends 15 VALU per echo.
// All threads of the block must call this.
__device__ __forceinline__ void stage_echo(double (*lds)[kEchoCols], uint32_t echo, uint32_t last_echo, double v,
                                           double *out, uint64_t row_stride)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    v = half_wave_sum(v);
    const int slot = echo % kEchoChunk;
    if ((lane & 31) == 31)
        lds[slot][2 * wave + (lane >> 5)] = v;
    if (slot == kEchoChunk - 1 || echo == last_echo) {
        __syncthreads();
        const int cnt = slot + 1;
        if ((int)threadIdx.x < cnt) {
            double s = 0.0;
            for (int q = 0; q < 2 * nw; ++q)
                s += lds[threadIdx.x][q];
            const uint64_t row = (uint64_t)(echo - slot) + threadIdx.x;
            out[row * row_stride] = s;
        }
        __syncthreads();
    }
}

// Stage the block sums of the echoes echo & ~3 .. echo, whose per-lane
// values are v[0 .. (echo & 3)] (entries past echo & 3 are ignored; v is
// overwritten, so the reduction works in place): one 4-value reduce-scatter
// (21 VALU, about 5 per echo instead of a 15-VALU half-wave sum each)
// leaves the wave total of v[r] in row r, whose first lane stages it.
// Called when echo & 3 == 3 or echo == last_echo; every kEchoChunk echoes
// (a multiple of 4) and at the last echo the block writes its partials
// out[row * row_stride] for the staged rows.  All threads of the block must
// call this, with the whole wave active.
__device__ __forceinline__ void stage_echo4(double (*lds)[kEchoCols], uint32_t echo, uint32_t last_echo,
                                            double (&v)[4], double *out, uint64_t row_stride)
{
    static_assert(kEchoChunk % 4 == 0, "echo groups of 4 inside a chunk");
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    const double r = wave_reduce_scatter<4>(v);
    const uint32_t g0 = echo & ~3u;
    const uint32_t row = (uint32_t)lane >> 4;
    if ((lane & 15) == 0 && g0 + row <= echo)
        lds[(g0 + row) % kEchoChunk][wave] = r;
    const int slot = echo % kEchoChunk;
    if (slot == kEchoChunk - 1 || echo == last_echo) {
        __syncthreads();
        const int cnt = slot + 1;
        if ((int)threadIdx.x < cnt) {
            double s = 0.0;
            for (int q = 0; q < nw; ++q)
                s += lds[threadIdx.x][q];
            const uint64_t orow = (uint64_t)(echo - slot) + threadIdx.x;
            out[orow * row_stride] = s;
        }
        __syncthreads();
    }
}

// Walk echoes 1 .. echoes (walk_echo() advances every lane's walkers by one
// echo; echo_value() is the lane's value for the echo sums) and stage the
// echo sums into part (nullable; staging as stage_echo / stage_echo4).
// GROUP4 reduces the values of 4 echoes at once (echoes 4g .. 4g + 3, echo
// 0 the start state), with the loop unrolled over a group so the values
// stay in registers; otherwise one half-wave sum per echo.  The image walks
// of one walker per lane (folded, unfolded, batched, internal field) use
// GROUP4, so the echo sums of the same energies agree bit for bit between
// their kernels.
// All threads of the block call it.
template <bool GROUP4, class WalkEcho, class EchoValue>
__device__ __forceinline__ void run_echoes(uint32_t echoes, double (*lds)[kEchoCols], double *part,
                                           uint64_t part_stride, WalkEcho &&walk_echo, EchoValue &&echo_value)
{
    if constexpr (!GROUP4) {
        if (part)
            stage_echo(lds, 0, echoes, echo_value(), part, part_stride);
        for (uint32_t echo = 1; echo <= echoes; ++echo) {
            walk_echo();
            if (part)
                stage_echo(lds, echo, echoes, echo_value(), part, part_stride);
        }
    } else {
        double EB[4] = {echo_value(), 0.0, 0.0, 0.0};
        for (uint32_t g0 = 0;; g0 += 4) {
            static_for<4>([&](auto jc) {
                constexpr uint32_t j = decltype(jc)::value;
                if ((j > 0 || g0 > 0) && g0 + j <= echoes) {   // uniform
                    walk_echo();
                    EB[j] = echo_value();
                }
            });
            const bool end = g0 + 3 >= echoes;
            if (part)
                stage_echo4(lds, end ? echoes : g0 + 3, echoes, EB, part, part_stride);
            if (end)
                break;
        }
    }
}

// Echo sums of 8 energy sets per walker: one reduce-scatter gives lanes 8q
// .. 8q + 7 the wave total of set q, lanes 0, 8, ..., 56 stage their set's
// wave total in LDS for kEchoChunk echoes (lds8: kEchoChunk x 8 x waves of
// the block, [slot][set][wave]), and the block then writes one partial per
// (set, echo): out[set * set_stride + row * row_stride].  About 4.3 VALU per
// set and echo instead of 15 for separate half-wave sums.  Fixed order, so
// deterministic.  All threads of the block call it, with the whole wave
// active.
__device__ __forceinline__ void stage_echo_sets8(double *lds8, uint32_t echo, uint32_t last_echo,
                                                 const double (&v)[8], double *out, uint64_t row_stride,
                                                 uint64_t set_stride)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    double t[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
        t[q] = v[q];
    const double r = wave_reduce_scatter<8>(t);
    const int slot = echo % kEchoChunk;
    if ((lane & 7) == 0)
        lds8[(slot * 8 + (lane >> 3)) * nw + wave] = r;
    if (slot == kEchoChunk - 1 || echo == last_echo) {
        __syncthreads();
        const int cnt = slot + 1;
        for (int q = (int)threadIdx.x; q < 8 * cnt; q += (int)blockDim.x) {
            const int row = q >> 3, set = q & 7;
            double s = 0.0;
            for (int w = 0; w < nw; ++w)
                s += lds8[q * nw + w];
            out[(uint64_t)set * set_stride + ((uint64_t)(echo - slot) + (uint64_t)row) * row_stride] = s;
        }
        __syncthreads();
    }
}

// PFGSE measurement at diffusion time i for the K walkers of each lane, with
// displacements (dX, dY, dZ)[k] voxels and energies e[k][0..nsets): block
// sums of e cos((k . dr) voxel) for every k of the time's k row and every set
// q, staged in LDS (lds rows q * kPfgseKChunk + k) and written as partials
// [(q * n_delta + i) * n_k + k][blockIdx.x].  cos is evaluated once per k and
// walker and shared by the sets; with ROT (k row k_0 + j d) the cosines
// follow the three-term recurrence cos(ph0 + (j+1) phd) = 2 cos(phd)
// cos(ph0 + j phd) - cos(ph0 + (j-1) phd), one FMA per k instead of a cos
// (over 64 k the error stays below 5e-13 of a term, far inside the 1e-11
// M(0) tolerance); the recurrence runs through whole groups of 8 k, the
// values past the last k unused.  A lane adds its K walkers' terms before
// the cross-lane reduce-scatter, so that cost is paid once per K walkers.
// Idle walkers carry zero energies, so their terms are zero without a
// select.  All threads of the block call it.
template <bool ROT, int MAXSETS, int K>
__device__ __forceinline__ void pfgse_measure_k(const PfgseArgs &p, uint32_t i, const double (&dX)[K],
                                                const double (&dY)[K], const double (&dZ)[K],
                                                const double (&e)[K][MAXSETS], int nsets,
                                                double (*lds)[kPfgseWaves])
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const uint64_t krow = (uint64_t)i * p.k_delta_stride;
    double c[K], cprev[K], twoc1[K];
    static_for<K>([&](auto kc_) {
        constexpr int w = decltype(kc_)::value;
        c[w] = 0.0;
        cprev[w] = 0.0;
        twoc1[w] = 2.0;
        if (ROT) {
            const double ph0 = (p.kx[krow] * dX[w] + p.ky[krow] * dY[w] + p.kz[krow] * dZ[w]) * p.voxel;
            const double phd = ((p.kx[krow + 1] - p.kx[krow]) * dX[w] + (p.ky[krow + 1] - p.ky[krow]) * dY[w] +
                                (p.kz[krow + 1] - p.kz[krow]) * dZ[w]) * p.voxel;
            double sn = 0.0, s1 = 0.0, c1 = 1.0;
            sincos(ph0, &sn, &c[w]);
            sincos(phd, &s1, &c1);
            cprev[w] = c[w] * c1 + sn * s1;   // cos(ph0 - phd)
            twoc1[w] = 2.0 * c1;
        }
    });
    for (uint32_t k0 = 0; k0 < p.n_k; k0 += kPfgseKChunk) {
        const uint32_t kc = p.n_k - k0 < (uint32_t)kPfgseKChunk ? p.n_k - k0 : (uint32_t)kPfgseKChunk;
        // groups of 8 k: the cosines of the group, then the reduce-scatters
        // of their terms
        for (uint32_t g0 = 0; g0 < kc; g0 += 8) {
            // the 8 cosines of walker w (ROT: the recurrence advanced 8 k)
            // handed to use(j, cos) in k order
            double cs[ROT ? 1 : K][8];
            if constexpr (!ROT)
                static_for<K>([&](auto kc_) {
                    constexpr int w = decltype(kc_)::value;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const uint32_t kk = g0 + (uint32_t)j;
                        cs[w][j] = 0.0;
                        if (kk < kc) {   // uniform
                            const uint64_t kk_i = krow + k0 + kk;
                            cs[w][j] = cos((p.kx[kk_i] * dX[w] + p.ky[kk_i] * dY[w] + p.kz[kk_i] * dZ[w]) * p.voxel);
                        }
                    }
                });
            double c0[K], cp0[K];   // ROT: the recurrence state at the group's first k
            static_for<K>([&](auto kc_) {
                constexpr int w = decltype(kc_)::value;
                c0[w] = c[w];
                cp0[w] = cprev[w];
            });
            auto cosines = [&](auto wc, auto &&use) {
                constexpr int w = decltype(wc)::value;
                if constexpr (ROT) {
                    // every set pair runs the 8 recurrence steps from the
                    // group's start (the cosines are not kept across pairs)
                    double cc = c0[w], cp = cp0[w];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        use(j, cc);
                        const double cn = fma(twoc1[w], cc, -cp);
                        cp = cc;
                        cc = cn;
                    }
                    c[w] = cc;
                    cprev[w] = cp;
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        use(j, cs[w][j]);
                }
            };
            // the lane's terms: its walkers' products, added in walker order
            // sets in pairs: one 16-value reduce-scatter (lanes 4(8p + j) ..
            // + 3 hold k j of set q0 + p), a lone last set by an 8-value one
            // (lanes 8j .. 8j + 7)
            static_for<(MAXSETS + 1) / 2>([&](auto pc) {
                constexpr int q0 = 2 * decltype(pc)::value;
                // the pair's second set (in bounds also where MAXSETS is odd
                // and the pair branch cannot run)
                constexpr int q1 = q0 + 1 < MAXSETS ? q0 + 1 : q0;
                if (q0 + 1 < MAXSETS && q0 + 1 < nsets) {
                    double t[16];
                    static_for<K>([&](auto wc) {
                        constexpr int w = decltype(wc)::value;
                        cosines(wc, [&](int j, double cv) {
                            t[j] = w == 0 ? cv * e[w][q0] : fma(cv, e[w][q0], t[j]);
                            t[8 + j] = w == 0 ? cv * e[w][q1] : fma(cv, e[w][q1], t[8 + j]);
                        });
                    });
                    const double v = wave_reduce_scatter<16>(t);
                    const uint32_t q = (uint32_t)q0 + ((uint32_t)lane >> 5);
                    const uint32_t kk = g0 + (((uint32_t)lane >> 2) & 7u);
                    if ((lane & 3) == 0 && kk < kc)
                        lds[q * kPfgseKChunk + kk][wave] = v;
                } else if (q0 < nsets) {
                    double t[8];
                    static_for<K>([&](auto wc) {
                        constexpr int w = decltype(wc)::value;
                        cosines(wc, [&](int j, double cv) {
                            t[j] = w == 0 ? cv * e[w][q0] : fma(cv, e[w][q0], t[j]);
                        });
                    });
                    const double v = wave_reduce_scatter<8>(t);
                    const uint32_t kk = g0 + ((uint32_t)lane >> 3);
                    if ((lane & 7) == 0 && kk < kc)
                        lds[q0 * kPfgseKChunk + kk][wave] = v;
                }
            });
        }
        __syncthreads();
        for (uint32_t r = threadIdx.x; r < (uint32_t)nsets * kc; r += blockDim.x) {
            const uint32_t q = r / kc, kk = r % kc;
            double sum = 0.0;
            for (int w = 0; w < kPfgseWaves; ++w)
                sum += lds[q * kPfgseKChunk + kk][w];
            const uint64_t row = ((uint64_t)q * p.n_delta + i) * p.n_k + k0 + kk;
            p.partials[row * p.partial_stride + blockIdx.x] = sum;
        }
        __syncthreads();
    }
}

// One walker per lane (the folded PFGSE kernel).
template <bool ROT, int MAXSETS = kMaxSets>
__device__ __forceinline__ void pfgse_measure(const PfgseArgs &p, uint32_t i, double dX, double dY, double dZ,
                                              const double *e, int nsets, bool active,
                                              double (*lds)[kPfgseWaves])
{
    (void)active;
    const double x[1] = {dX}, y[1] = {dY}, z[1] = {dZ};
    double ee[1][MAXSETS];
    for (int q = 0; q < MAXSETS; ++q)
        ee[0][q] = e[q];
    pfgse_measure_k<ROT, MAXSETS, 1>(p, i, x, y, z, ee, nsets, lds);
}

// Sums of squared displacements per axis at diffusion time i over the K
// walkers of each lane (voxel^2; integer-valued, so the FP64 sums are exact
// in any order).
template <int K>
__device__ __forceinline__ void pfgse_msd_k(const PfgseArgs &p, uint32_t i, const double (&dX)[K],
                                            const double (&dY)[K], const double (&dZ)[K], const bool (&active)[K],
                                            double (*lds)[kPfgseWaves])
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    double xx = 0.0, yy = 0.0, zz = 0.0;
    static_for<K>([&](auto kc_) {
        constexpr int w = decltype(kc_)::value;
        xx += active[w] ? dX[w] * dX[w] : 0.0;
        yy += active[w] ? dY[w] * dY[w] : 0.0;
        zz += active[w] ? dZ[w] * dZ[w] : 0.0;
    });
    const double sx = wave_sum(xx);
    const double sy = wave_sum(yy);
    const double sz = wave_sum(zz);
    if (lane == 0) {
        lds[0][wave] = sx;
        lds[1][wave] = sy;
        lds[2][wave] = sz;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double sum = 0.0;
        for (int w = 0; w < kPfgseWaves; ++w)
            sum += lds[threadIdx.x][w];
        const uint64_t row = p.msd_rows + (uint64_t)i * 3 + threadIdx.x;
        p.partials[row * p.partial_stride + blockIdx.x] = sum;
    }
    __syncthreads();
}

__device__ __forceinline__ void pfgse_msd(const PfgseArgs &p, uint32_t i, double dX, double dY, double dZ,
                                          bool active, double (*lds)[kPfgseWaves])
{
    const double x[1] = {dX}, y[1] = {dY}, z[1] = {dZ};
    const bool act[1] = {active};
    pfgse_msd_k<1>(p, i, x, y, z, act, lds);
}

// One tile of walkers (blockDim.x * K consecutive walkers starting at
// tile * blockDim.x * K) walked for a.echoes x a.steps_per_echo steps.
// EMODE: 0 = hit counting only (map), 1 = per-walker delta, 2 = per-label
// delta.  Echo partials go to part[row * part_stride] (nullable).
template <int BC, int EMODE, int K>
__device__ __forceinline__ void walk_tile(const WalkArgs &a, uint64_t tile, double *part, uint64_t part_stride,
                                          double (*lds)[kEchoCols])
{
    constexpr int NL = EMODE == 2 ? kMaxLabels : 1;
    const Geom g = make_geom(a);
    const uint64_t base = tile * blockDim.x * K + threadIdx.x;
    const uint64_t n = a.n;
    const uint32_t echoes = a.echoes, spe = a.steps_per_echo;
    const int n_labels = a.n_labels;
    Walker w[K];
    uint32_t lhits[K][NL];
    static_for<K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const uint64_t id = base + (uint64_t)k * blockDim.x;
        w[k].active = id < n;
        const uint64_t j = w[k].active ? id : 0;
        w[k].x = w[k].active ? a.x[j] : 0;
        w[k].y = w[k].active ? a.y[j] : 0;
        w[k].z = w[k].active ? a.z[j] : 0;
        w[k].s = w[k].active ? a.seed[j] : 1;
        w[k].e = (w[k].active && a.energy) ? a.energy[j] : 1.0;
        w[k].h = (w[k].active && a.hits) ? a.hits[j] : 0u;
        w[k].delta = EMODE == 1 ? ((w[k].active && a.penalty) ? a.penalty[j] : a.delta_uniform) : 1.0;
        if (w[k].active && !start_ok<BC>(g, w[k])) {
            w[k].active = 0;
            atomicOr(a.error_flag, 1u);
        }
        if (!w[k].active) {
            w[k].x = w[k].y = w[k].z = 0;
            w[k].e = 0.0;
        }
        init_fold<BC>(g, w[k]);
        static_for<NL>([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            lhits[k][l] = (EMODE == 2 && a.label_hits && w[k].active && l < n_labels)
                              ? a.label_hits[(uint64_t)l * n + j]
                              : 0u;
        });
    });

    auto echo_value = [&]() {
        double v = 0.0;
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            v += w[k].active ? w[k].e : 0.0;
        });
        return v;
    };
    run_echoes<K == 1>(echoes, lds, part, part_stride, [&]() {
        for (uint32_t st = 0; st < spe; ++st) {
            Move mv[K];
            uint64_t word[K];
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                mv[k] = propose<BC>(g, w[k]);
            });
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                word[k] = w[k].active ? g.blocks[mv[k].blk] : 0ull;
            });
            static_for<K>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
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
                        static_for<NL>([&](auto lc) {
                            constexpr int l = decltype(lc)::value;
                            lhits[k][l] += (lab == (uint32_t)l) ? 1u : 0u;
                        });
                    }
                }
            });
        }
    }, echo_value);

    static_for<K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (w[k].active) {
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
                static_for<NL>([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    if (l < n_labels)
                        a.label_hits[(uint64_t)l * n + id] = lhits[k][l];
                });
        }
    });
}

// Internal-field CPMG on one tile (CPMG_walk_{noflux,periodic,mirror}_field,
// B/NMR_Simulation/NMR_cpmg_cuda.cpp:146-293, 409-567, 683-841): per echo two
// phase inversions (phase = -phase) of steps_per_echo / 2 steps each; before
// and after every step the phase gains gammatau * field at the walker's image
// voxel (one fused multiply-add, as CUDA contracts phase + gammatau * field);
// a wall hit multiplies p_alive by delta.  Echo value p_alive cos(phase).
// Runs on the folded geometry (the field is indexed by the image voxel).
template <int BC, int K>
__device__ __forceinline__ void walk_tile_field(const WalkArgs &a, uint64_t tile, double *part, uint64_t part_stride,
                                                double (*lds)[kEchoCols])
{
    const Geom g = make_geom(a);
    const uint64_t base = tile * blockDim.x * K + threadIdx.x;
    const uint64_t n = a.n;
    const uint32_t echoes = a.echoes, half = a.steps_per_echo / 2;
    const uint64_t row = (uint64_t)a.W, slab = (uint64_t)a.W * (uint64_t)a.H;
    const double gt = a.gammatau;
    Walker w[K];
    double PA[K], PH[K];
    static_for<K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const uint64_t id = base + (uint64_t)k * blockDim.x;
        w[k].active = id < n;
        const uint64_t j = w[k].active ? id : 0;
        w[k].x = w[k].active ? a.x[j] : 0;
        w[k].y = w[k].active ? a.y[j] : 0;
        w[k].z = w[k].active ? a.z[j] : 0;
        w[k].s = w[k].active ? a.seed[j] : 1;
        w[k].h = (w[k].active && a.hits) ? a.hits[j] : 0u;
        w[k].delta = (w[k].active && a.penalty) ? a.penalty[j] : a.delta_uniform;
        w[k].e = 0.0;
        PA[k] = w[k].active ? a.p_alive[j] : 0.0;
        PH[k] = w[k].active ? a.phase[j] : 0.0;
        if (w[k].active && !start_ok<BC>(g, w[k])) {
            w[k].active = 0;
            atomicOr(a.error_flag, 1u);
        }
        if (!w[k].active) {
            w[k].x = w[k].y = w[k].z = 0;
            PA[k] = 0.0;
            PH[k] = 0.0;
        }
        init_fold<BC>(g, w[k]);
    });
    auto fval = [&](const Walker &v) {
        return a.field[(uint64_t)v.ix + (uint64_t)v.iy * row + (uint64_t)v.iz * slab];
    };
    auto echo_value = [&]() {
        double v = 0.0;
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            v += PA[k] * cos(PH[k]);
        });
        return v;
    };
    run_echoes<K == 1>(echoes, lds, part, part_stride, [&]() {
        for (int inv = 0; inv < 2; ++inv) {
            static_for<K>([&](auto kc) { PH[decltype(kc)::value] = -PH[decltype(kc)::value]; });
            for (uint32_t st = 0; st < half; ++st) {
                Move mv[K];
                uint64_t word[K];
                static_for<K>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    PH[k] = fma(gt, fval(w[k]), PH[k]);
                    mv[k] = propose<BC>(g, w[k]);
                });
                static_for<K>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    word[k] = w[k].active ? g.blocks[mv[k].blk] : 0ull;
                });
                static_for<K>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    if (!((word[k] >> mv[k].bit) & 1ull)) {
                        commit<BC>(w[k], mv[k]);
                    } else {
                        w[k].h += 1u;
                        PA[k] *= w[k].delta;
                    }
                    PH[k] = fma(gt, fval(w[k]), PH[k]);
                });
            }
        }
    }, echo_value);
    static_for<K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (w[k].active) {
            const uint64_t id = base + (uint64_t)k * blockDim.x;
            a.x[id] = w[k].x;
            a.y[id] = w[k].y;
            a.z[id] = w[k].z;
            a.seed[id] = w[k].s;
            a.p_alive[id] = PA[k];
            a.phase[id] = PH[k];
            if (a.energy)
                a.energy[id] = PA[k] * cos(PH[k]);
            if (a.hits)
                a.hits[id] = w[k].h;
        }
    });
}

}  // namespace rwnmr
