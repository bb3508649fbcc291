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
    const uint32_t seg = blockIdx.y;
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
        const uint64_t col = (uint64_t)seg * gridDim.x + blockIdx.x;
        b.partials[(uint64_t)threadIdx.x * b.partial_stride + col] = s;
    }
}

// The same replay on the unfolded image (periodic / mirror; a.zwords,
// a.zl, a.zlabels), with the per-walker register chunk cache.
template <bool WRAPFIX>
__global__ __launch_bounds__(256) void zbackward_kernel(const WalkArgs a, const BackwardArgs b)
{
    __shared__ double lds[kMaxLabels][kMaxWaves];
    __shared__ uint4 tab[6];
    __shared__ uint2 fix[6];
    __shared__ double ldelta[kMaxLabels];
    zbuild_table(a.zl, tab, fix);
    if (threadIdx.x < kMaxLabels)
        ldelta[threadIdx.x] = (int)threadIdx.x < b.n_labels ? a.label_delta[threadIdx.x] : 1.0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    const uint32_t seg = blockIdx.y;
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = b.n;
    const bool active = id < n;
    const uint64_t j = active ? id : 0;
    const CkptView cv = ckpt_view(b.ckpt, n, b.n_segments, b.n_labels, seg);
    uint32_t A = zaddress(active ? cv.x[j] : 0, active ? cv.y[j] : 0, active ? cv.z[j] : 0, a.zl);
    uint64_t s = active ? cv.seed[j] : 1ull;
    double e = active ? cv.energy[j] : 0.0;
    uint32_t QB = 0xFFFFFFFFu;
    uint4 QV = make_uint4(0u, 0u, 0u, 0u);
    uint32_t h[kMaxLabels];
    double G[kMaxLabels];
    const int n_labels = b.n_labels;
    static_for<kMaxLabels>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        h[l] = (active && l < n_labels) ? cv.hits[(uint64_t)l * n + j] : 0u;
        G[l] = 0.0;
    });
    const char *words = reinterpret_cast<const char *>(a.zwords);
    const char *tab_bytes = reinterpret_cast<const char *>(tab);
    const char *fix_bytes = reinterpret_cast<const char *>(fix);
    const uint32_t e0 = seg * b.every;
    const uint32_t e1 = (e0 + b.every < b.echoes) ? e0 + b.every : b.echoes;
    for (uint32_t echo = e0 + 1; echo <= e1; ++echo) {
        for (uint32_t st = 0; st < b.steps_per_echo; ++st) {
            int32_t dp;
            const uint32_t c = zpropose<WRAPFIX>(s, A, tab_bytes, fix_bytes, dp);
            zchunk_fill(words, c, QB, QV);
            if (!zchunk_bit(c, QV)) {
                A = c;
            } else {
                const uint32_t lab = zlabel(a.zlabels, c);
                e *= ldelta[lab];
                static_for<kMaxLabels>([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    h[l] += (lab == (uint32_t)l) ? 1u : 0u;
                });
            }
        }
        const double ge = b.grad_echo[echo] * e;
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
        double sum = 0.0;
        for (int q = 0; q < nw; ++q)
            sum += lds[threadIdx.x][q];
        const uint64_t col = (uint64_t)seg * gridDim.x + blockIdx.x;
        b.partials[(uint64_t)threadIdx.x * b.partial_stride + col] = sum;
    }
}

hipError_t launch_backward(const WalkArgs &a, const BackwardArgs &b, int bc, uint32_t threads, hipStream_t st)
{
    if (a.zwords) {
        if (!a.zlabels || threads != 256)
            return hipErrorInvalidValue;
        const dim3 grid((uint32_t)((b.n + threads - 1) / threads), b.n_segments);
        if (a.zl.wrapfix)
            hipLaunchKernelGGL(zbackward_kernel<true>, grid, dim3(threads), 0, st, a, b);
        else
            hipLaunchKernelGGL(zbackward_kernel<false>, grid, dim3(threads), 0, st, a, b);
        return hipGetLastError();
    }
    const uint32_t bx = (uint32_t)((b.n + threads - 1) / threads);
    const dim3 grid(bx, b.n_segments);
    switch (bc) {
    case kNoflux: hipLaunchKernelGGL(backward_kernel<kNoflux>, grid, dim3(threads), 0, st, a, b); break;
    case kPeriodic: hipLaunchKernelGGL(backward_kernel<kPeriodic>, grid, dim3(threads), 0, st, a, b); break;
    case kMirror: hipLaunchKernelGGL(backward_kernel<kMirror>, grid, dim3(threads), 0, st, a, b); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rwnmr
