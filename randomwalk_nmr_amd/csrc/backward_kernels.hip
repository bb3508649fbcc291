// backward_kernels.hip -- gradient of the CPMG echo sums with respect to the
// per-label wall penalties, by trajectory replay from checkpoints.
//
// Relaxation never changes a trajectory (a wall hit leaves the walker in
// place, B/NMR_Simulation/NMR_cpmg_cuda.cpp:113-124), so each walker's energy
// after echo n is E_w(n) = prod over its hits of delta_label(hit), and
//   dM(n)/d delta_l = sum_w E_w(n) h_wl(n) / delta_l,
// with h_wl(n) the walker's hits on label l up to echo n.  For an upstream
// gradient g(n) = dL/dM(n) the kernel accumulates, per walker and label,
//   G_wl = sum_n g(n) E_w(n) h_wl(n)
// while replaying one checkpoint segment (state saved every `every`
// echoes: position, RNG state, energy, per-label hits), then reduces G over
// walkers.  One thread per (walker, segment): segments replay in parallel
// and no hit history is stored.  Periodic and mirror images replay on the
// unfolded image (zwalk_device.h), noflux on the folded one.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "backward_kernels.h"
#include "walk_device.h"
#include "zwalk_device.h"

namespace rwnmr {

template <int BC>
__global__ __launch_bounds__(256) void backward_kernel(const WalkArgs a, const BackwardArgs b)
{
    __shared__ double lds[kMaxLabels][kMaxWaves];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    uint32_t seg;
    uint64_t tile;
    if (!backward_block(b, seg, tile))
        return;   // spare workgroup of the XCD-striped grid (the whole block leaves)
    const uint64_t id = tile * blockDim.x + threadIdx.x;
    const uint64_t n = b.n;
    const bool active = id < n;
    const uint64_t j = active ? id : 0;
    const CkptView cv = ckpt_view(b.ckpt, n, b.n_segments, b.n_labels, seg);
    const Geom g = make_geom(a);
    Walker w;
    w.active = active;
    w.x = active ? cv.x[j] : 0;
    w.y = active ? cv.y[j] : 0;
    w.z = active ? cv.z[j] : 0;
    w.s = active ? cv.seed[j] : 1;
    w.e = active ? cv.energy[j] : 0.0;
    w.h = 0;
    w.delta = 1.0;
    init_fold<BC>(g, w);
    uint32_t h[kMaxLabels];
    double G[kMaxLabels];
    const int n_labels = b.n_labels;
    static_for<kMaxLabels>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        h[l] = (active && l < n_labels) ? cv.hits[(uint64_t)l * n + j] : 0u;
        G[l] = 0.0;
    });
    const uint32_t e0 = seg * b.every;
    const uint32_t e1 = (e0 + b.every < b.echoes) ? e0 + b.every : b.echoes;
    for (uint32_t echo = e0 + 1; echo <= e1; ++echo) {
        for (uint32_t st = 0; st < b.steps_per_echo; ++st) {
            Move mv = propose<BC>(g, w);
            const uint64_t word = active ? g.blocks[mv.blk] : 0ull;
            if (!((word >> mv.bit) & 1ull)) {
                commit<BC>(w, mv);
            } else {
                const uint32_t lab = a.labels_blocked[(uint64_t)mv.blk * 64u + mv.bit];
                w.e *= a.label_delta[lab];
                static_for<kMaxLabels>([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    h[l] += (lab == (uint32_t)l) ? 1u : 0u;
                });
            }
        }
        const double ge = b.grad_echo[echo] * w.e;
        static_for<kMaxLabels>([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            G[l] += ge * (double)h[l];
        });
    }
    static_for<kMaxLabels>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        const double v = wave_sum(active ? G[l] : 0.0);
        if (lane == 0)
            lds[l][wave] = v;
    });
    __syncthreads();
    if ((int)threadIdx.x < b.n_labels) {
        double s = 0.0;
        for (int q = 0; q < nw; ++q)
            s += lds[threadIdx.x][q];
        const uint64_t col = (uint64_t)seg * b.blocks + tile;
        b.partials[(uint64_t)threadIdx.x * b.partial_stride + col] = s;
    }
}

// The same replay on the unfolded image (periodic / mirror; a.zwords,
// a.zl, a.zlabels) with the forward walk's step: 16-byte register chunk
// cache tested by one shift, byte-key direction table, next direction drawn
// while the chunk load is in flight.  The per-echo sum over labels is
// rearranged so a step touches one label only: with Q(n) = sum of g E over
// the segment's echoes up to n and h_wl(n) = h_wl(start) + hits so far,
//   G_wl = h_wl(end) Q(end) - R_wl,  R_wl = sum over the segment's label-l
// hits of Q at the hit (Q before the hit's echo),
// so a hit adds Q to R[label] (one LDS add) and an echo adds g E to Q;
// h_wl(end) is the next segment's checkpoint (or the final hits).
template <bool WRAPFIX>
__global__ __launch_bounds__(kZBackThreads) void zbackward_kernel(const WalkArgs a, const BackwardArgs b)
{
    __shared__ double lds[kMaxLabels][kMaxWaves];
    __shared__ uint4 tab[6];
    __shared__ uint2 fix[6];
    __shared__ uint8_t keys[kZKeyEntries];
    __shared__ double ldelta[kMaxLabels];
    // R_wl, label-major: the block size is fixed, so a label's offset is one
    // shift-and-add, and 32 KiB keeps 4 blocks (8 waves per SIMD) per CU
    __shared__ double racc[kMaxLabels * kZBackThreads];
    uint32_t seg;
    uint64_t tile;
    if (!backward_block(b, seg, tile))
        return;   // spare workgroup of the XCD-striped grid (the whole block leaves)
    zbuild_table(a.zl, tab, fix);
    zbuild_key_table(keys, 16u);
    if (threadIdx.x < kMaxLabels)
        ldelta[threadIdx.x] = (int)threadIdx.x < b.n_labels ? a.label_delta[threadIdx.x] : 1.0;
    double *my_r = racc + threadIdx.x;
    for (int l = 0; l < kMaxLabels; ++l)
        my_r[l * kZBackThreads] = 0.0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    const uint64_t id = tile * blockDim.x + threadIdx.x;
    const uint64_t n = b.n;
    const bool active = id < n;
    const uint64_t j = active ? id : 0;
    const CkptView cv = ckpt_view(b.ckpt, n, b.n_segments, b.n_labels, seg);
    const uint32_t *h_end = seg + 1 < b.n_segments ? ckpt_view(b.ckpt, n, b.n_segments, b.n_labels, seg + 1).hits
                                                   : ckpt_final_hits(b.ckpt, n, b.n_segments, b.n_labels);
    uint32_t A = zaddress(active ? cv.x[j] : 0, active ? cv.y[j] : 0, active ? cv.z[j] : 0, a.zl);
    uint64_t s = active ? cv.seed[j] : 1ull;
    double e = active ? cv.energy[j] : 0.0;   // idle lanes contribute 0
    double Q = 0.0;
    uint32_t QB = 0xFFFFFFFFu;
    uint4 QV = make_uint4(0u, 0u, 0u, 0u);
    const zrsrc_t zrsrc = zchunk_rsrc(a.zwords, a.zl);
    const char *tab_bytes = reinterpret_cast<const char *>(tab);
    const char *fix_bytes = reinterpret_cast<const char *>(fix);
    ZEntry T;
    zdirection<WRAPFIX, 2>(s, tab_bytes, fix_bytes, T, keys);
    // a hit's label is loaded at the hit and applied one step later (or at
    // the echo's end, before Q moves on), so its latency overlaps the next
    // step's chunk load
    uint32_t lab = 0u;
    bool pend = false;
    auto settle = [&]() {
        if (pend) {
            e *= ldelta[lab];
            __hip_atomic_fetch_add(my_r + lab * kZBackThreads, Q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            pend = false;
        }
    };
    auto step = [&]() {
        const uint32_t c = zmove_entry<WRAPFIX>(A, T);
        zchunk_fill_buf(zrsrc, c, QB, QV);
        zdirection<WRAPFIX, 2>(s, tab_bytes, fix_bytes, T, keys);
        // the table reads stay ahead of the wall test, so their latency
        // overlaps the chunk load's
        __builtin_amdgcn_sched_barrier(0);
        settle();
        if (__builtin_expect(zchunk_wall<128>(c, QV) != 0u, 0)) {
            lab = zlabel_at(a, c);
            pend = true;
        } else {
            A = c;
        }
    };
    const uint32_t e0 = seg * b.every;
    const uint32_t e1 = (e0 + b.every < b.echoes) ? e0 + b.every : b.echoes;
    const uint32_t spe = b.steps_per_echo;
    for (uint32_t echo = e0 + 1; echo <= e1; ++echo) {
        uint32_t st = 0;
        for (; st + 2 <= spe; st += 2) {
            step();
            step();
        }
        if (st < spe)
            step();
        settle();
        Q += b.grad_echo[echo] * e;
    }
    const int n_labels = b.n_labels;
    static_for<kMaxLabels>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        double G = 0.0;
        if (active && l < n_labels)
            G = (double)h_end[(uint64_t)l * n + j] * Q - my_r[l * kZBackThreads];
        const double v = wave_sum(G);
        if (lane == 0)
            lds[l][wave] = v;
    });
    __syncthreads();
    if ((int)threadIdx.x < b.n_labels) {
        double sum = 0.0;
        for (int q = 0; q < nw; ++q)
            sum += lds[threadIdx.x][q];
        const uint64_t col = (uint64_t)seg * b.blocks + tile;
        b.partials[(uint64_t)threadIdx.x * b.partial_stride + col] = sum;
    }
}

hipError_t launch_backward(const WalkArgs &a, const BackwardArgs &b, int bc, uint32_t threads, hipStream_t st)
{
    if (a.zwords) {
        if (!a.zlabels || threads != kZBackThreads)
            return hipErrorInvalidValue;
        if ((uint64_t)b.blocks != (b.n + threads - 1) / threads)
            return hipErrorInvalidValue;
        const dim3 grid((uint32_t)backward_grid(b.blocks, b.n_segments));
        if (a.zl.wrapfix)
            hipLaunchKernelGGL(zbackward_kernel<true>, grid, dim3(threads), 0, st, a, b);
        else
            hipLaunchKernelGGL(zbackward_kernel<false>, grid, dim3(threads), 0, st, a, b);
        return hipGetLastError();
    }
    const uint32_t bx = (uint32_t)((b.n + threads - 1) / threads);
    if ((uint64_t)b.blocks != bx)
        return hipErrorInvalidValue;
    const dim3 grid((uint32_t)backward_grid(bx, b.n_segments));
    switch (bc) {
    case kNoflux: hipLaunchKernelGGL(backward_kernel<kNoflux>, grid, dim3(threads), 0, st, a, b); break;
    case kPeriodic: hipLaunchKernelGGL(backward_kernel<kPeriodic>, grid, dim3(threads), 0, st, a, b); break;
    case kMirror: hipLaunchKernelGGL(backward_kernel<kMirror>, grid, dim3(threads), 0, st, a, b); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rwnmr
