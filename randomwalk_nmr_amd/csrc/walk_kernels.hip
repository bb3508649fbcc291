// walk_kernels.hip -- gfx950 kernels of the random-walk NMR hot path.
//
// One lane owns K walkers (structure-of-arrays in HBM, walker k of a lane at
// index base + k * blockDim.x so every load/store is coalesced).  Each step:
// xorshift64 in registers, direction = state % 6 (32-bit popcount form),
// image fold kept incrementally in registers (no integer division in the
// loop), one 8-byte bitBlock gather, then move or wall hit (hit count +1 and
// energy *= delta).  Per-echo magnetisation is reduced in-wave (cross-lane
// shuffles), staged per block in LDS for 32 echoes at a time, and written as
// one FP64 partial per (echo, block); a second kernel sums the partials of
// each echo in a fixed order, so echo sums are deterministic run to run.
//
// Reference semantics (paths under src/backend/ of RandomWalk_NMR):
//   CPMG_walk_{noflux,periodic,mirror}  B/NMR_Simulation/NMR_cpmg_cuda.cpp:37-139, 300-402, 574-676
//   map_3D_{noflux,periodic,mirror}     B/NMR_Simulation/NMR_Simulation_cuda.cpp:418-721
//   CPMG_energyReduce + host sum        B/NMR_Simulation/NMR_cpmg_cuda.cpp:888-925, 1780-1801
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_common.h"
#include "walk_device.h"
#include "zwalk_kernels.h"

namespace rwnmr {

// One block per tile of blockDim.x * K walkers.
template <int BC, int EMODE, int K>
__global__ __launch_bounds__(max_threads_for_k(K)) void walk_kernel(const WalkArgs a)
{
    __shared__ double lds[kEchoChunk][kEchoCols];
    const uint64_t tile = tile_of_block(a);
    if (a.tile_map && tile >= a.n_tiles)
        return;  // spare workgroup of the XCD-striped grid (whole block leaves together)
    walk_tile<BC, EMODE, K>(a, tile, a.partials ? a.partials + tile : nullptr, a.partial_stride, lds);
}

template <int BC, int K>
__global__ __launch_bounds__(max_threads_for_k(K)) void walk_field_kernel(const WalkArgs a)
{
    __shared__ double lds[kEchoChunk][kEchoCols];
    const uint64_t tile = tile_of_block(a);
    if (a.tile_map && tile >= a.n_tiles)
        return;
    walk_tile_field<BC, K>(a, tile, a.partials ? a.partials + tile : nullptr, a.partial_stride, lds);
}

// Voxel-division path (shift > 0): walker coordinates are sub-voxel units
// and the image voxel is (p >> shift); folds recomputed each step.
template <int BC, int EMODE>
__global__ __launch_bounds__(1024) void walk_kernel_shift(const WalkArgs a)
{
    __shared__ double lds[kEchoChunk][kEchoCols];
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int active = id < a.n;
    const uint64_t j = active ? id : 0;
    int x = active ? a.x[j] : 0, y = active ? a.y[j] : 0, z = active ? a.z[j] : 0;
    uint64_t s = active ? a.seed[j] : 1;
    double e = (active && a.energy) ? a.energy[j] : 1.0;
    uint32_t h = (active && a.hits) ? a.hits[j] : 0u;
    const double delta = EMODE == 1 ? ((active && a.penalty) ? a.penalty[j] : a.delta_uniform) : 1.0;
    const int sh = a.shift;
    if (active && BC == kNoflux &&
        !((unsigned)(x >> sh) < (unsigned)a.W && (unsigned)(y >> sh) < (unsigned)a.H &&
          (unsigned)(z >> sh) < (unsigned)a.D)) {
        active = 0;
        atomicOr(a.error_flag, 1u);
    }
    if (!active)
        e = 0.0;
    if (a.partials)
        stage_echo(lds, 0, a.echoes, active ? e : 0.0, a.partials + blockIdx.x, a.partial_stride);
    for (uint32_t echo = 1; echo <= a.echoes; ++echo) {
        for (uint32_t st = 0; st < a.steps_per_echo; ++st) {
            s = xorshift64(s);
            const uint32_t r = mod6(s);
            const int ax = move_axis(r);
            int sg = move_sign(r);
            if (BC == kNoflux) {
                const int g = (ax == 0 ? x : (ax == 1 ? y : z)) >> sh;
                const int d = ax == 0 ? a.W : (ax == 1 ? a.H : a.D);
                if ((sg < 0 && g == 0) || (sg > 0 && g == d - 1))
                    sg = -sg;
            }
            const int nx = x + (ax == 0 ? sg : 0);
            const int ny = y + (ax == 1 ? sg : 0);
            const int nz = z + (ax == 2 ? sg : 0);
            int ix, iy, iz;
            if (BC == kMirror) {
                ix = fold_mirror(nx >> sh, a.W);
                iy = fold_mirror(ny >> sh, a.H);
                iz = fold_mirror(nz >> sh, a.D);
            } else if (BC == kPeriodic) {
                ix = fold_periodic(nx >> sh, a.W);
                iy = fold_periodic(ny >> sh, a.H);
                iz = fold_periodic(nz >> sh, a.D);
            } else {
                ix = nx >> sh;
                iy = ny >> sh;
                iz = nz >> sh;
            }
            const uint32_t blk = brick_block_index(ix, iy, iz, a.layout);
            const uint32_t bit = bit_index(ix, iy, iz);
            const uint64_t word = active ? a.blocks[blk] : 0ull;
            if (!((word >> bit) & 1ull)) {
                x = nx;
                y = ny;
                z = nz;
            } else {
                h += 1u;
                if (EMODE == 1)
                    e *= delta;
                else if (EMODE == 2)
                    e *= a.label_delta[a.labels_blocked[(uint64_t)blk * 64u + bit]];
            }
        }
        if (a.partials)
            stage_echo(lds, echo, a.echoes, active ? e : 0.0, a.partials + blockIdx.x, a.partial_stride);
    }
    if (active) {
        a.x[id] = x;
        a.y[id] = y;
        a.z[id] = z;
        a.seed[id] = s;
        if (a.energy)
            a.energy[id] = e;
        if (a.hits)
            a.hits[id] = h;
    }
}

// PFGSE: the walk between diffusion times counts hits only (relaxation never
// changes a trajectory), then every energy set is updated once with
// pow(delta, hits of the increment) and the block sums of E cos(phase) for all
// k are staged through LDS (pfgse_measure).  Partials:
// [(set * n_delta + i) * n_k + k][block].  Used for noflux; periodic and
// mirror run zpfgse_kernel on the unfolded image.
template <int BC, bool ROT>
__global__ __launch_bounds__(kPfgseThreads) void pfgse_kernel(const WalkArgs a, const PfgseArgs p)
{
    __shared__ double lds[kMaxSets * kPfgseKChunk][kPfgseWaves];
    bool inside;
    const uint64_t id = pfgse_walker(p, blockIdx.x, threadIdx.x, blockDim.x, a.n, inside);
    const Geom g = make_geom(a);
    Walker w;
    w.active = inside;
    const uint64_t j = w.active ? id : 0;
    w.x = w.active ? a.x[j] : 0;
    w.y = w.active ? a.y[j] : 0;
    w.z = w.active ? a.z[j] : 0;
    w.s = w.active ? a.seed[j] : 1;
    w.h = (w.active && a.hits) ? a.hits[j] : 0u;
    const int x0 = w.active ? p.x0[j] : 0;
    const int y0 = w.active ? p.y0[j] : 0;
    const int z0 = w.active ? p.z0[j] : 0;
    const double own_delta = (w.active && a.penalty) ? a.penalty[j] : a.delta_uniform;
    double e[kMaxSets];
    const int nsets = p.n_sets > 0 ? (int)p.n_sets : 1;
    const bool ok = w.active && start_ok<BC>(g, w);
    if (w.active && !ok)
        atomicOr(a.error_flag, 1u);
    w.active = ok;
    static_for<kMaxSets>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        e[q] = !w.active ? 0.0 : ((q == 0 && a.energy) ? a.energy[j] : 1.0);
    });
    if (!w.active)
        w.x = w.y = w.z = 0;
    init_fold<BC>(g, w);

    for (uint32_t i = 0; i < p.n_delta; ++i) {
        uint32_t h = 0;
        const uint32_t steps = p.steps_inc[i];
        for (uint32_t st = 0; st < steps; ++st) {
            Move mv = propose<BC>(g, w);
            const uint64_t word = w.active ? g.blocks[mv.blk] : 0ull;
            if (!((word >> mv.bit) & 1ull))
                commit<BC>(w, mv);
            else
                h += 1u;
        }
        w.h += h;
        static_for<kMaxSets>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if (q < nsets)
                e[q] *= pow(p.n_sets > 0 ? p.set_delta[q] : own_delta, (double)h);
        });
        const double dX = (double)(w.x - x0);
        const double dY = (double)(w.y - y0);
        const double dZ = (double)(w.z - z0);
        pfgse_measure<ROT>(p, i, dX, dY, dZ, e, nsets, w.active, lds);
        if (p.msd_rows)
            pfgse_msd(p, i, dX, dY, dZ, w.active, lds);
    }
    if (w.active) {
        a.x[id] = w.x;
        a.y[id] = w.y;
        a.z[id] = w.z;
        a.seed[id] = w.s;
        if (a.energy)
            a.energy[id] = e[0];
        if (a.hits)
            a.hits[id] = w.h;
    }
}

hipError_t launch_pfgse(const WalkArgs &a, const PfgseArgs &p, int bc, uint32_t threads, hipStream_t st)
{
    if (a.n == 0)
        return hipSuccess;
    if (threads != kPfgseThreads)
        return hipErrorInvalidValue;
    // the grid and the sample grouping were sized for the kernel that runs
    if (p.walkers_per_block != (a.zwords ? kPfgseThreads * (uint32_t)kPfgseK : kPfgseThreads))
        return hipErrorInvalidValue;
    if (a.zwords)
        return launch_zpfgse(a, p, st);
    const uint32_t blocks = pfgse_grid_blocks(p, a.n);
#define RW_PF(B)                                                                                             \
    if (bc == B) {                                                                                           \
        if (p.k_rot)                                                                                         \
            hipLaunchKernelGGL((pfgse_kernel<B, true>), dim3(blocks), dim3(threads), 0, st, a, p);           \
        else                                                                                                 \
            hipLaunchKernelGGL((pfgse_kernel<B, false>), dim3(blocks), dim3(threads), 0, st, a, p);          \
        return hipGetLastError();                                                                            \
    }
    RW_PF(kNoflux) RW_PF(kPeriodic) RW_PF(kMirror)
#undef RW_PF
    return hipErrorInvalidValue;
}

// out[r] = sum_c partials[r * cols + c], fixed summation order per row.
__global__ __launch_bounds__(256) void reduce_rows_kernel(const double *partials, uint64_t rows, uint32_t cols,
                                                          double *out)
{
    __shared__ double red[256];
    for (uint64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        double s = 0.0;
        for (uint32_t c = threadIdx.x; c < cols; c += blockDim.x)
            s += partials[r * cols + c];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int off = blockDim.x / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off)
                red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            out[r] = red[0];
        __syncthreads();
    }
}

// ---------------------------------------------------------------- launch

template <int BC, int EMODE>
static hipError_t launch_bc_emode(const WalkArgs &a, const LaunchShape &ls, hipStream_t st)
{
    const uint64_t per_block = (uint64_t)ls.threads * ls.k;
    const uint32_t tiles = (uint32_t)((a.n + per_block - 1) / per_block);
    if (a.tile_map && a.n_tiles != tiles)
        return hipErrorInvalidValue;
    // XCD-striped tiles: a grid of 8 * ceil(tiles / 8) workgroups (spares exit)
    const uint32_t blocks = a.tile_map ? ((tiles + kXcds - 1) / kXcds) * kXcds : tiles;
    if constexpr (EMODE == 3) {
        // the field walk runs at voxel resolution only
        if (a.shift != 0 || !a.field || !a.p_alive || !a.phase)
            return hipErrorInvalidValue;
        switch (ls.k) {
        case 1: hipLaunchKernelGGL((walk_field_kernel<BC, 1>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
        case 2: hipLaunchKernelGGL((walk_field_kernel<BC, 2>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
        case 4: hipLaunchKernelGGL((walk_field_kernel<BC, 4>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    } else {
        if (a.shift != 0) {
            const uint32_t tpb = ls.threads;
            const uint64_t sblocks = (a.n + tpb - 1) / tpb;
            hipLaunchKernelGGL((walk_kernel_shift<BC, EMODE>), dim3((uint32_t)sblocks), dim3(tpb), 0, st, a);
            return hipGetLastError();
        }
        switch (ls.k) {
        case 1: hipLaunchKernelGGL((walk_kernel<BC, EMODE, 1>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
        case 2: hipLaunchKernelGGL((walk_kernel<BC, EMODE, 2>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
        case 4: hipLaunchKernelGGL((walk_kernel<BC, EMODE, 4>), dim3(blocks), dim3(ls.threads), 0, st, a); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
}

uint32_t walk_grid_blocks(uint64_t n, const LaunchShape &ls, int shift)
{
    const uint64_t per_block = shift ? (uint64_t)ls.threads : (uint64_t)ls.threads * ls.k;
    return (uint32_t)((n + per_block - 1) / per_block);
}

hipError_t launch_walk(const WalkArgs &a, int bc, int emode, const LaunchShape &ls, hipStream_t st)
{
    if (a.n == 0)
        return hipSuccess;
    if (walk_grid_blocks(a.n, ls, a.shift) > 0x7FFFFFFFu || ls.threads > max_threads_for_k((int)ls.k))
        return hipErrorInvalidValue;
#define RW_CASE(B, E)                                                                                   \
    if (bc == B && emode == E)                                                                            \
        return launch_bc_emode<B, E>(a, ls, st);
    RW_CASE(kNoflux, 0) RW_CASE(kNoflux, 1) RW_CASE(kNoflux, 2)
    RW_CASE(kPeriodic, 0) RW_CASE(kPeriodic, 1) RW_CASE(kPeriodic, 2)
    RW_CASE(kMirror, 0) RW_CASE(kMirror, 1) RW_CASE(kMirror, 2)
    RW_CASE(kNoflux, 3) RW_CASE(kPeriodic, 3) RW_CASE(kMirror, 3)
#undef RW_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_reduce_rows(const double *partials, uint64_t rows, uint32_t cols, double *out, hipStream_t st)
{
    if (rows == 0)
        return hipSuccess;
    const uint32_t grid = (uint32_t)(rows < 65536 ? rows : 65536);
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(grid), dim3(256), 0, st, partials, rows, cols, out);
    return hipGetLastError();
}

}  // namespace rwnmr
