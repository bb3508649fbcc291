// probe_kernels.hip -- random-gather probe: the rate of independent random
// 16-byte loads (the shape of the walk kernels' chunk reloads) from a device
// table of a given size.  The table size decides which level serves them:
// a table inside one XCD's 4 MB L2, one inside the 256 MB Infinity Cache
// (MALL), or one far larger (HBM).  Measurement only; no reference
// counterpart (the reference publishes no gather rate).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/rwnmr/rwnmr.h"

namespace {

constexpr uint32_t kProbeThreads = 256;
constexpr uint32_t kProbeBlocks = 4096;   // 16 waves per CU on 256 CUs
constexpr int kProbeUnroll = 8;           // loads in flight per lane

// Each lane draws chunk indices with xorshift32 (no data dependence between
// loads, so kProbeUnroll are in flight) and folds the loaded words into one
// value so no load is dead.
__global__ __launch_bounds__(kProbeThreads) void gather_probe_kernel(const uint4 *__restrict__ table, uint32_t mask,
                                                                     uint32_t rounds, uint32_t *out)
{
    const uint32_t id = blockIdx.x * kProbeThreads + threadIdx.x;
    uint32_t s = id * 2654435761u + 0x9E3779B9u;
    s = s ? s : 1u;
    uint32_t acc = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
        uint4 v[kProbeUnroll];
#pragma unroll
        for (int j = 0; j < kProbeUnroll; ++j) {
            s ^= s << 13;
            s ^= s >> 17;
            s ^= s << 5;
            v[j] = table[s & mask];
        }
#pragma unroll
        for (int j = 0; j < kProbeUnroll; ++j)
            acc += v[j].x ^ v[j].y ^ v[j].z ^ v[j].w;
    }
    out[id] = acc;
}

__global__ void probe_fill_kernel(uint4 *table, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        table[i] = make_uint4((uint32_t)i, (uint32_t)(i >> 32), ~(uint32_t)i, 0x5A5A5A5Au);
}

}  // namespace

extern "C" rwnmr_status rwnmr_probe_gather(rwnmr_ctx *ctx, uint64_t table_bytes, uint32_t loads_per_thread,
                                           double *loads_per_s, double *kernel_ms)
{
    if (!ctx || !loads_per_s || table_bytes < 16 || (table_bytes & (table_bytes - 1)) != 0 ||
        table_bytes > (16ull << 32) || loads_per_thread < kProbeUnroll)
        return RWNMR_ERR_INVALID;
    hipStream_t st = static_cast<hipStream_t>(rwnmr_ctx_stream(ctx));
    const uint64_t chunks = table_bytes / 16;
    const uint32_t rounds = loads_per_thread / kProbeUnroll;
    const uint64_t threads = (uint64_t)kProbeBlocks * kProbeThreads;
    uint4 *table = nullptr;
    uint32_t *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rwnmr_status status = RWNMR_OK;
    float ms = 0.f;
    if (hipMalloc((void **)&table, table_bytes) != hipSuccess || hipMalloc((void **)&out, threads * 4) != hipSuccess) {
        status = RWNMR_ERR_ALLOC;
        goto done;
    }
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        status = RWNMR_ERR_HIP;
        goto done;
    }
    hipLaunchKernelGGL(probe_fill_kernel, dim3(2048), dim3(256), 0, st, table, chunks);
    // one untimed launch brings the table into whatever level holds it
    hipLaunchKernelGGL(gather_probe_kernel, dim3(kProbeBlocks), dim3(kProbeThreads), 0, st, table,
                       (uint32_t)(chunks - 1), rounds, out);
    (void)hipEventRecord(e0, st);
    hipLaunchKernelGGL(gather_probe_kernel, dim3(kProbeBlocks), dim3(kProbeThreads), 0, st, table,
                       (uint32_t)(chunks - 1), rounds, out);
    (void)hipEventRecord(e1, st);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f) {
        status = RWNMR_ERR_HIP;
        goto done;
    }
    *loads_per_s = (double)threads * (double)rounds * kProbeUnroll / (1e-3 * ms);
    if (kernel_ms)
        *kernel_ms = ms;
done:
    (void)hipStreamSynchronize(st);
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    if (table)
        (void)hipFree(table);
    if (out)
        (void)hipFree(out);
    return status;
}
