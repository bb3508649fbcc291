// capi.hip -- C ABI entry points of the MI355X random-walk NMR hot path
// (declared in include/rwnmr/rwnmr.h).
//
// A context owns one HIP stream and grow-only device buffers, so repeated
// calls (fit iterations, PFGSE diffusion times, benchmarks) allocate nothing
// after the first call of a given size.  Reference paths under src/backend/.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rwnmr/rwnmr.h"
#include "backward_kernels.h"
#include "batch_kernels.h"
#include "locality_kernels.h"
#include "zwalk_kernels.h"
#include "walk_kernels.h"

using rwnmr::LaunchShape;
using rwnmr::PfgseArgs;
using rwnmr::WalkArgs;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t want)
    {
        if (want <= bytes)
            return hipSuccess;
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess)
            bytes = want;
        return e;
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ void fill_f64_kernel(double *p, double v, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// Device scratch of the no-allocation CPMG entry point
// (rwnmr_cpmg_run_noalloc): the error word in the first 256 bytes, the echo
// partials after it.  One copy per GPU, held until the process ends.  Host
// code never passes the array's name to a kernel or a copy: each GPU's copy
// has its own device address, read once per device by hipGetSymbolAddress
// (scratch_address) and kept in g_scratch_ptr.
constexpr size_t kScratchBytes = 16u << 20;
constexpr size_t kScratchFlagBytes = 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices];

// Device address of `scratch` on `device`, which must be the current device;
// resolved on the first call per device.  Empty string on success.
std::string scratch_address(int device, void **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxScratchDevices)
        return "device index beyond the scratch table";
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return std::string();
}

}  // namespace

struct rwnmr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    LaunchShape shape{256, 1};
    // walker staging for host-pointer calls
    DevBuf x, y, z, seed, energy, hits, penalty, x0, y0, z0;
    // per-call small arrays and reductions
    DevBuf partials, sums, label_delta, label_hits, steps_inc, kvec, set_delta;
    DevBuf jac_partials, jac_rows;   // Jacobian forward: per-label partials and host-call rows
    // internal-field CPMG staging (host-pointer calls)
    DevBuf field, p_alive, phase;
    DevBuf msamples;   // PFGSE per-sample sums (host-pointer calls)
    // label-set CPMG: the walkers' start state (replayed per group of 8 sets)
    // and the 8 set energies carried between launches
    DevBuf lset_x, lset_y, lset_z, lset_seed, lset_energy, lset_energies;
    // batched CPMG: descriptors, tile list, reduction rows, queue head
    DevBuf batch_inst, batch_tiles, batch_rows, batch_counter;
    DevBuf flag;  // uint32 error word
    uint32_t *host_flag = nullptr;  // pinned readback: [0] of flag, [1] of the scratch error word
    // this device's copy of `scratch` (no-allocation entry point); null with
    // scratch_err set when its address could not be read
    void *scratch = nullptr;
    std::string scratch_err;
    bool scratch_flag_pending = false;  // a no-allocation run's error word is still to be checked
    // optional HIP-event timing of the walk / PFGSE kernels on this stream
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> event_pool;
    double kernel_ms = 0.0;
    uint64_t kernel_launches = 0;
    // live images: destroying the context releases their device memory and
    // detaches them, so a later rwnmr_image_destroy only frees the handle
    std::vector<rwnmr_image *> images;
    // device layout for images uploaded from now on (log2 brick extents)
    int brick_sx = 1, brick_sy = 1, brick_sz = 1;
    // locality ordering: Morton-sorted walkers, XCD-striped tiles, re-sorted
    // every resort_every echoes (only for runs of at least min_walkers)
    int locality = 1;
    uint32_t resort_every = 1000;
    uint64_t locality_min_walkers = 65536;
    DevBuf lkeys, lidx, lsort_tmp;
    DevBuf sortA[8], sortB[8];
    // periodic / mirror walks on the unfolded image (built per image on first use):
    // 0 folded kernels, 1 unfolded, 2 unfolded with the per-walker 128-bit
    // chunk cache, 3 and 4 with the chunk tested by one shift and the next
    // direction drawn a step ahead, through the 4 KiB byte-key table (3, the
    // default) or the 64 KiB wide direction table (4); 5 the speculative
    // CPMG walk; 6 as 3 with an 8-byte (4 x 4 x 4 voxel) chunk for mirror
    // walks (others run 3)
    int unfolded = 3;
    // resident 1024-thread blocks of the flag-keyed mirror walk (two per CU):
    // the tile count of balanced tiles (zbalanced_tile_walkers)
    uint32_t balanced_slots = 0;
};

// Unfolded image of one boundary condition (dilated addresses).
struct ZImage {
    uint32_t *words = nullptr;
    rwnmr::ZLayout zl{};
    int state = 0;   // 0: not built, 1: built, -1: not available (over 2^32 voxels or no memory)
    uint8_t *labels = nullptr;    // solid labels, one byte each, in the same order (per-label walks;
                                  // the first octant only when zoctant_layout(zl))
    int lstate = 0;  // as state
};

struct rwnmr_image {
    rwnmr_ctx *ctx = nullptr;
    int32_t W = 0, H = 0, D = 0;
    int32_t bC = 0, bR = 0, bD = 0;
    uint64_t nblocks = 0;        // blocks of the reference layout
    rwnmr::BrickLayout layout{0, 0, 0, 0, 0};
    uint64_t nwords = 0;         // words of the device layout (bricks padded)
    uint64_t *blocks = nullptr;        // device
    uint8_t *labels_blocked = nullptr; // device, 64 bytes per block
    int32_t n_labels = 0;
    // unfolded bit volumes: [0] periodic (W x H x D), [1] mirror (2W x 2H x 2D)
    mutable ZImage z[2];
};

namespace {

rwnmr_status fail(rwnmr_ctx *c, rwnmr_status code, const std::string &msg)
{
    if (c)
        c->err = msg;
    return code;
}

#define RW_HIP(ctx, call)                                                                                   \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            return fail((ctx), e_ == hipErrorOutOfMemory ? RWNMR_ERR_ALLOC : RWNMR_ERR_HIP,                 \
                        std::string(#call) + ": " + hipGetErrorString(e_));                                 \
    } while (0)

hipEvent_t take_event(rwnmr_ctx *c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

// Record the start of a timed kernel; returns the event (or null when off).
hipEvent_t timing_begin(rwnmr_ctx *c)
{
    if (!c->timing)
        return nullptr;
    hipEvent_t e = take_event(c);
    if (e)
        (void)hipEventRecord(e, c->stream);
    return e;
}

void timing_end(rwnmr_ctx *c, hipEvent_t start)
{
    if (!c->timing || !start)
        return;
    hipEvent_t e = take_event(c);
    if (!e)
        return;
    (void)hipEventRecord(e, c->stream);
    c->pending.emplace_back(start, e);
}

// Fold finished event pairs into the kernel-time counters (stream idle).
void timing_collect(rwnmr_ctx *c)
{
    for (auto &pr : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
            c->kernel_ms += ms;
            c->kernel_launches += 1;
        }
        c->event_pool.push_back(pr.first);
        c->event_pool.push_back(pr.second);
    }
    c->pending.clear();
}

rwnmr_status clear_flag(rwnmr_ctx *c)
{
    RW_HIP(c, c->flag.reserve(sizeof(uint32_t)));
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, c->stream, c->flag.as<uint32_t>(), 0u, 1ull);
    RW_HIP(c, hipGetLastError());
    return RWNMR_OK;
}

// Synchronise and turn a raised kernel flag into RWNMR_ERR_WALKER.
rwnmr_status sync_and_check(rwnmr_ctx *c)
{
    c->host_flag[0] = 0;
    c->host_flag[1] = 0;
    if (c->flag.p)
        RW_HIP(c, hipMemcpyAsync(c->host_flag, c->flag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    const bool scratch_flag = c->scratch_flag_pending;
    if (scratch_flag)
        RW_HIP(c, hipMemcpyAsync(c->host_flag + 1, c->scratch, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    RW_HIP(c, hipStreamSynchronize(c->stream));
    c->scratch_flag_pending = false;
    timing_collect(c);
    if (c->host_flag[0] || c->host_flag[1])
        return fail(c, RWNMR_ERR_WALKER, "a walker started outside the image (noflux boundary)");
    return RWNMR_OK;
}

rwnmr_status check_image(rwnmr_ctx *c, const rwnmr_image *img)
{
    if (!img || !img->blocks)
        return fail(c, RWNMR_ERR_INVALID, "image is null");
    if (img->ctx != c)
        return fail(c, RWNMR_ERR_INVALID, "image belongs to another context");
    return RWNMR_OK;
}

rwnmr_status check_walkers(rwnmr_ctx *c, const rwnmr_image *img, const rwnmr_walkers *w, int32_t bc,
                           uint32_t shift)
{
    if (!w)
        return fail(c, RWNMR_ERR_INVALID, "walkers is null");
    if (w->n && (!w->x || !w->y || !w->z || !w->seed))
        return fail(c, RWNMR_ERR_INVALID, "walker position/seed arrays must be non-null");
    if (bc < RWNMR_BC_NOFLUX || bc > RWNMR_BC_MIRROR)
        return fail(c, RWNMR_ERR_INVALID, "unknown boundary condition");
    if (shift > 16)
        return fail(c, RWNMR_ERR_INVALID, "voxel-division shift above 16");
    if (bc == RWNMR_BC_NOFLUX && (img->W < 2 || img->H < 2 || img->D < 2))
        return fail(c, RWNMR_ERR_INVALID, "noflux needs every image dimension >= 2");
    if (w->n > (1ull << 40))
        return fail(c, RWNMR_ERR_INVALID, "too many walkers");
    if (!w->on_device && bc == RWNMR_BC_NOFLUX) {
        for (uint64_t i = 0; i < w->n; ++i) {
            const int gx = w->x[i] >> shift, gy = w->y[i] >> shift, gz = w->z[i] >> shift;
            if ((unsigned)gx >= (unsigned)img->W || (unsigned)gy >= (unsigned)img->H ||
                (unsigned)gz >= (unsigned)img->D)
                return fail(c, RWNMR_ERR_WALKER, "walker " + std::to_string(i) + " starts outside the image");
        }
    }
    return RWNMR_OK;
}

// Device views of the walker arrays: the caller's pointers (on_device) or
// the context's staging copies (host pointers, copied in here).
struct WalkerView {
    int32_t *x, *y, *z;
    uint64_t *seed;
    double *energy;
    uint32_t *hits;
    const double *penalty;
};

rwnmr_status stage_in(rwnmr_ctx *c, const rwnmr_walkers *w, WalkerView &v)
{
    if (w->on_device) {
        v = {w->x, w->y, w->z, w->seed, w->energy, w->hits, w->penalty};
        return RWNMR_OK;
    }
    const uint64_t n = w->n;
    RW_HIP(c, c->x.reserve(n * 4 + 4));
    RW_HIP(c, c->y.reserve(n * 4 + 4));
    RW_HIP(c, c->z.reserve(n * 4 + 4));
    RW_HIP(c, c->seed.reserve(n * 8 + 8));
    v.x = c->x.as<int32_t>();
    v.y = c->y.as<int32_t>();
    v.z = c->z.as<int32_t>();
    v.seed = c->seed.as<uint64_t>();
    v.energy = nullptr;
    v.hits = nullptr;
    v.penalty = nullptr;
    RW_HIP(c, hipMemcpyAsync(v.x, w->x, n * 4, hipMemcpyHostToDevice, c->stream));
    RW_HIP(c, hipMemcpyAsync(v.y, w->y, n * 4, hipMemcpyHostToDevice, c->stream));
    RW_HIP(c, hipMemcpyAsync(v.z, w->z, n * 4, hipMemcpyHostToDevice, c->stream));
    RW_HIP(c, hipMemcpyAsync(v.seed, w->seed, n * 8, hipMemcpyHostToDevice, c->stream));
    if (w->energy) {
        RW_HIP(c, c->energy.reserve(n * 8 + 8));
        v.energy = c->energy.as<double>();
        RW_HIP(c, hipMemcpyAsync(v.energy, w->energy, n * 8, hipMemcpyHostToDevice, c->stream));
    }
    if (w->hits) {
        RW_HIP(c, c->hits.reserve(n * 4 + 4));
        v.hits = c->hits.as<uint32_t>();
        RW_HIP(c, hipMemcpyAsync(v.hits, w->hits, n * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (w->penalty) {
        RW_HIP(c, c->penalty.reserve(n * 8 + 8));
        RW_HIP(c, hipMemcpyAsync(c->penalty.p, w->penalty, n * 8, hipMemcpyHostToDevice, c->stream));
        v.penalty = c->penalty.as<double>();
    }
    return RWNMR_OK;
}

rwnmr_status stage_out(rwnmr_ctx *c, rwnmr_walkers *w, const WalkerView &v)
{
    if (w->on_device)
        return RWNMR_OK;
    const uint64_t n = w->n;
    RW_HIP(c, hipMemcpyAsync(w->x, v.x, n * 4, hipMemcpyDeviceToHost, c->stream));
    RW_HIP(c, hipMemcpyAsync(w->y, v.y, n * 4, hipMemcpyDeviceToHost, c->stream));
    RW_HIP(c, hipMemcpyAsync(w->z, v.z, n * 4, hipMemcpyDeviceToHost, c->stream));
    RW_HIP(c, hipMemcpyAsync(w->seed, v.seed, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (w->energy)
        RW_HIP(c, hipMemcpyAsync(w->energy, v.energy, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (w->hits)
        RW_HIP(c, hipMemcpyAsync(w->hits, v.hits, n * 4, hipMemcpyDeviceToHost, c->stream));
    return RWNMR_OK;
}

WalkArgs base_args(const rwnmr_image *img, const WalkerView &v, const rwnmr_walkers *w, uint32_t shift,
                   uint32_t *flag)
{
    WalkArgs a;
    memset(&a, 0, sizeof(a));
    a.blocks = img->blocks;
    a.labels_blocked = img->labels_blocked;
    a.W = img->W;
    a.H = img->H;
    a.D = img->D;
    a.bC = img->bC;
    a.bCR = img->bC * img->bR;
    a.layout = img->layout;
    a.tile_map = 0;
    a.n_tiles = 0;
    a.shift = (int32_t)shift;
    a.x = v.x;
    a.y = v.y;
    a.z = v.z;
    a.seed = v.seed;
    a.energy = v.energy;
    a.hits = v.hits;
    a.penalty = v.penalty;
    a.delta_uniform = w->delta_uniform;
    a.n = w->n;
    a.error_flag = flag;
    return a;
}

// Unfolded image for a periodic or mirror walk, built on the context stream
// on first use; null when the walk runs on the folded kernels (option off,
// noflux, more than 2^32 voxels in the unfolded volume, or no memory for it).
const ZImage *zimage_for(rwnmr_ctx *c, const rwnmr_image *img, int bc, bool build = true)
{
    if (!c->unfolded || (bc != rwnmr::kPeriodic && bc != rwnmr::kMirror))
        return nullptr;
    ZImage &z = img->z[bc == rwnmr::kMirror ? 1 : 0];
    if (z.state == 0 && build) {
        const uint32_t f = bc == rwnmr::kMirror ? 2u : 1u;
        if (!rwnmr::zlayout_make(f * (uint32_t)img->W, f * (uint32_t)img->H, f * (uint32_t)img->D, z.zl)) {
            z.state = -1;
            return nullptr;
        }
        if (hipMalloc((void **)&z.words, rwnmr::zimage_words(z.zl) * 4) != hipSuccess) {
            (void)hipGetLastError();
            z.words = nullptr;
            z.state = -1;
            return nullptr;
        }
        if (rwnmr::launch_zimage_build(img->blocks, img->layout, img->W, img->H, img->D,
                                       bc == rwnmr::kMirror ? 1 : 0, z.zl, z.words, c->stream) != hipSuccess) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(z.words);
            z.words = nullptr;
            z.state = -1;
            return nullptr;
        }
        z.state = 1;
    }
    return z.state == 1 ? &z : nullptr;
}

void free_zimages(rwnmr_image *img)
{
    for (ZImage &z : img->z) {
        if (z.words)
            (void)hipFree(z.words);
        if (z.labels)
            (void)hipFree(z.labels);
        z.words = nullptr;
        z.labels = nullptr;
        z.state = 0;
        z.lstate = 0;
    }
}

// Label volume of an image with solid labels (built on first use); null
// when unavailable.
const uint8_t *zlabels_for(rwnmr_ctx *c, const rwnmr_image *img, int bc)
{
    if (!img->labels_blocked)
        return nullptr;
    const ZImage *zi = zimage_for(c, img, bc);
    if (!zi)
        return nullptr;
    ZImage &z = img->z[bc == rwnmr::kMirror ? 1 : 0];
    if (z.lstate == 0 && z.zl.bits > 31)
        z.lstate = -1;   // labelled walks on the unfolded image address below 2^31
    if (z.lstate == 0) {
        const bool octant = bc == rwnmr::kMirror && rwnmr::zoctant_layout(z.zl);
        if (hipMalloc((void **)&z.labels, rwnmr::zlabel_bytes(z.zl, octant)) != hipSuccess) {
            (void)hipGetLastError();
            z.labels = nullptr;
            z.lstate = -1;
            return nullptr;
        }
        if (rwnmr::launch_zlabels_build(img->labels_blocked, img->layout, img->W, img->H, img->D,
                                        bc == rwnmr::kMirror ? 1 : 0, z.zl, octant, z.labels, c->stream) != hipSuccess) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(z.labels);
            z.labels = nullptr;
            z.lstate = -1;
            return nullptr;
        }
        z.lstate = 1;
    }
    return z.lstate == 1 ? z.labels : nullptr;
}

// Use the unfolded-image kernel for this walk when one applies.  noflux_ok:
// the caller's kernel also runs noflux walks on the periodic unfolded image
// (the CPMG, map and PFGSE walks, with the pipelined forms 3 and 4 and every
// image size >= 2).  mflags_ok: the caller's kernel starts mirror walkers in
// the first octant of the mirror image with reflection flags (the CPMG and
// map walks, forms 3 and 4, both then on the byte-key table), so each launch
// begins with the walkers' image lines inside one eighth of the mirror
// image.
// build: false uses an unfolded image only when an earlier call built it
// (callers that must not allocate device memory).
void attach_zimage(rwnmr_ctx *c, const rwnmr_image *img, int bc, int emode, WalkArgs &a, bool noflux_ok = false,
                   bool mflags_ok = false, bool build = true)
{
    a.zwords = nullptr;
    a.zlabels = nullptr;
    rwnmr::zoctant_constants(a, false);
    a.znoflux = 0;
    a.zmflags = 0;
    if (emode < 0 || emode > 2 || a.shift != 0)
        return;
    int zbc = bc;
    if (bc == rwnmr::kNoflux) {
        if (!noflux_ok || c->unfolded < 3 || img->W < 2 || img->H < 2 || img->D < 2)
            return;
        zbc = rwnmr::kPeriodic;
    }
    const uint8_t *zlab = nullptr;
    if (emode == 2 && (!build || !(zlab = zlabels_for(c, img, zbc))))
        return;
    if (const ZImage *z = zimage_for(c, img, zbc, build)) {
        a.zwords = z->words;
        a.zl = z->zl;
        a.zcache = c->unfolded >= 2 ? (uint32_t)(c->unfolded - 1) : 0u;
        a.zlabels = zlab;
        // mirror image of equal power-of-two periods: labels kept for the
        // first octant and read at the folded address
        rwnmr::zoctant_constants(a, zlab && bc == rwnmr::kMirror && rwnmr::zoctant_layout(z->zl));
        a.znoflux = bc == rwnmr::kNoflux ? 1u : 0u;
        a.zmflags = (bc == rwnmr::kMirror && mflags_ok && c->unfolded >= 3) ? 1u : 0u;
    }
}

hipError_t launch_any_walk(const WalkArgs &a, int bc, int emode, const LaunchShape &ls, hipStream_t st)
{
    return a.zwords ? rwnmr::launch_zwalk(a, emode, ls, st) : rwnmr::launch_walk(a, bc, emode, ls, st);
}

// True when every k row (n_rows rows of n_k vectors; kx, ky, kz blocks of nkv
// doubles) is k_0 + j (k_1 - k_0) to within a few ulps of its magnitude.
bool k_rows_arithmetic(const double *k, uint64_t nkv, uint32_t n_k, uint32_t n_rows)
{
    if (n_k < 2)
        return false;
    for (int comp = 0; comp < 3; ++comp)
        for (uint32_t r = 0; r < n_rows; ++r) {
            const double *row = k + (uint64_t)comp * nkv + (uint64_t)r * n_k;
            const double d = row[1] - row[0];
            double mag = 0.0;
            for (uint32_t j = 0; j < n_k; ++j)
                mag = std::max(mag, std::fabs(row[j]));
            for (uint32_t j = 0; j < n_k; ++j)
                if (std::fabs(row[j] - (row[0] + (double)j * d)) > 4e-15 * mag)
                    return false;
        }
    return true;
}

// Upper bound on the partial-sum scratch of one walk launch; longer CPMG runs
// are split into several launches of fewer echoes.
constexpr uint64_t kMaxPartialBytes = 1ull << 30;
// walker counts up to one wave per SIMD (1024 SIMDs x 64 lanes): the
// autotuner also times the speculative walk (form 5) for these
constexpr uint64_t kSpecMaxWalkers = 1024ull * 64ull;

// Tiles (echo-partial columns) of a walk launch over a.n walkers: balanced
// tiles (zbalanced_tile_walkers, which set a.tile_walkers) or the usual grid
// of blockDim x K walkers per tile (walk_grid_blocks).
uint32_t walk_columns(const rwnmr_ctx *c, WalkArgs &a, int emode, const LaunchShape &ls)
{
    a.tile_walkers = 0;
    const uint64_t tw = rwnmr::zbalanced_tile_walkers(a, emode, ls, c->balanced_slots);
    if (tw == 0)
        return rwnmr::walk_grid_blocks(a.n, ls, a.shift);
    a.tile_walkers = (uint32_t)tw;
    return (uint32_t)((a.n + tw - 1) / tw);
}

// Walk in position order: every `resort_every` echoes the walkers are sorted
// by the Morton code of their image block and gathered into context arrays,
// the walk runs with XCD-striped tiles, and at the end the final state is
// scattered back to the caller's order.  Per-walker results are identical to
// the unsorted walk; echo sums differ only by summation order.
// `a` holds the caller's (or staged) device arrays; sums is device or null.
rwnmr_status walk_local(rwnmr_ctx *c, const rwnmr_image *img, const WalkArgs &a, int bc, int emode, uint32_t echoes,
                        uint32_t steps_per_echo, double *sums, uint64_t max_chunk)
{
    const uint64_t n = a.n;
    if (n > 0xFFFFFFFFull)
        return fail(c, RWNMR_ERR_INVALID, "locality ordering supports up to 2^32 walkers");
    const LaunchShape ls = c->shape;
    // a's tiles (balanced tiles when walk_columns set them)
    const uint64_t per_tile = a.tile_walkers ? (uint64_t)a.tile_walkers : (uint64_t)ls.threads * ls.k;
    const uint32_t tiles = (uint32_t)((n + per_tile - 1) / per_tile);
    // keys: the unfolded word address (zl.bits - 5 bits) or the 30-bit Morton code of the block
    const int key_bits = a.zwords ? (int)std::max<uint32_t>(1u, a.zl.bits > 5 ? a.zl.bits - 5 : 1u) : 30;
    size_t tmp_bytes = 0;
    RW_HIP(c, rwnmr::sort_pairs_temp_bytes(n, key_bits, &tmp_bytes));
    RW_HIP(c, c->lsort_tmp.reserve(tmp_bytes + 256));
    RW_HIP(c, c->lkeys.reserve(8 * n));
    RW_HIP(c, c->lidx.reserve(8 * n));
    const size_t sizes[8] = {4, 4, 4, 8, 8, 4, 8, 4};
    for (int i = 0; i < 8; ++i) {
        RW_HIP(c, c->sortA[i].reserve(sizes[i] * n));
        RW_HIP(c, c->sortB[i].reserve(sizes[i] * n));
    }
    auto state_of = [&](DevBuf *b) {
        rwnmr::SortedState s;
        s.x = b[0].as<int32_t>();
        s.y = b[1].as<int32_t>();
        s.z = b[2].as<int32_t>();
        s.seed = b[3].as<uint64_t>();
        s.energy = b[4].as<double>();
        s.hits = a.hits ? b[5].as<uint32_t>() : nullptr;
        s.penalty = a.penalty ? b[6].as<double>() : nullptr;
        s.orig = b[7].as<uint32_t>();
        return s;
    };
    rwnmr::SortedState caller;
    caller.x = a.x;
    caller.y = a.y;
    caller.z = a.z;
    caller.seed = a.seed;
    caller.energy = a.energy;
    caller.hits = a.hits;
    caller.penalty = nullptr;
    caller.orig = nullptr;
    DevBuf *sets[2] = {c->sortA, c->sortB};
    int which = 0;
    rwnmr::SortedState cur = caller;
    uint32_t *keys = c->lkeys.as<uint32_t>(), *keys_alt = keys + n;
    uint32_t *idx = c->lidx.as<uint32_t>(), *idx_alt = idx + n;
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(c->resort_every, max_chunk));
    uint32_t done = 0;
    bool first = true;
    do {
        // sort the current state by position and gather it into the other set
        const rwnmr::SortedState dst = state_of(sets[which]);
        which ^= 1;
        if (a.zwords)
            RW_HIP(c, rwnmr::launch_zkeys(cur.x, cur.y, cur.z, n, a.zl, (int)a.zmflags, keys, idx, c->stream));
        else
            RW_HIP(c, rwnmr::launch_morton_keys(bc, cur.x, cur.y, cur.z, n, img->W, img->H, img->D, keys, idx,
                                                c->stream));
        RW_HIP(c, rwnmr::sort_pairs(c->lsort_tmp.p, tmp_bytes, keys, keys_alt, idx, idx_alt, n, key_bits, c->stream));
        RW_HIP(c, rwnmr::launch_gather_walkers(dst, cur, idx_alt, first ? nullptr : cur.orig, a.penalty, n, c->stream));
        cur = dst;
        const uint32_t ne = (uint32_t)std::min<uint64_t>(chunk, (uint64_t)echoes - done);
        WalkArgs b = a;
        b.x = cur.x;
        b.y = cur.y;
        b.z = cur.z;
        b.seed = cur.seed;
        b.energy = cur.energy;
        b.hits = cur.hits;
        b.penalty = cur.penalty;
        b.tile_map = 1;
        b.n_tiles = tiles;
        b.echoes = ne;
        b.steps_per_echo = steps_per_echo;
        b.partials = sums ? c->partials.as<double>() : nullptr;
        b.partial_stride = tiles;
        hipEvent_t t0 = timing_begin(c);
        RW_HIP(c, launch_any_walk(b, bc, emode, ls, c->stream));
        timing_end(c, t0);
        if (sums)
            RW_HIP(c, rwnmr::launch_reduce_rows(b.partials, (uint64_t)ne + 1, tiles, sums + done, c->stream));
        done += ne;
        first = false;
    } while (done < echoes);
    RW_HIP(c, rwnmr::launch_scatter_walkers(caller, cur, n, c->stream));
    return RWNMR_OK;
}

// Internal-field inputs of a CPMG run (emode 3).
struct FieldRun {
    const double *field;   // W*H*D, host or device like the walkers
    double gammatau;
    double *p_alive, *phase;   // nullable: start at 1 / 0, not returned
};

// Shared CPMG driver: emode 1 (per-walker delta), 2 (per-label delta) or 3
// (internal field, `fr`).  jac (emode 2, nullable): also the Jacobian rows
// jac[l][0..echoes] = sum_w h_wl(n) E_w(n) for l < n_labels (unfolded image,
// one walker per lane, per-label hit counts kept across launches).
rwnmr_status cpmg_common(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                         uint32_t steps_per_echo, int32_t bc, uint32_t shift, int emode,
                         const double *label_delta, uint32_t *label_hits, double *echo_sums,
                         const FieldRun *fr = nullptr, double *jac = nullptr)
{
    rwnmr_status s;
    if (!c)
        return RWNMR_ERR_INVALID;
    if ((s = check_image(c, img)) != RWNMR_OK)
        return s;
    if ((s = check_walkers(c, img, w, bc, shift)) != RWNMR_OK)
        return s;
    if (emode == 2) {
        if (!img->labels_blocked || img->n_labels <= 0)
            return fail(c, RWNMR_ERR_INVALID, "image has no solid labels");
        if (!label_delta)
            return fail(c, RWNMR_ERR_INVALID, "label_delta is null");
        if (shift != 0 && label_hits)
            return fail(c, RWNMR_ERR_INVALID, "per-label hit counts need shift == 0");
    }
    if (jac && (emode != 2 || shift != 0 || !echo_sums))
        return fail(c, RWNMR_ERR_INVALID, "Jacobian forward: per-label run with echo sums and shift == 0");
    WalkerView v;
    if ((s = clear_flag(c)) != RWNMR_OK)
        return s;
    if ((s = stage_in(c, w, v)) != RWNMR_OK)
        return s;
    WalkArgs a = base_args(img, v, w, shift, c->flag.as<uint32_t>());
    a.n_labels = img->n_labels;
    if (emode == 3) {
        if (!fr || !fr->field)
            return fail(c, RWNMR_ERR_INVALID, "field run: null field");
        if (shift != 0)
            return fail(c, RWNMR_ERR_INVALID, "field run: voxel division (shift > 0) is not supported");
        const uint64_t vox = (uint64_t)img->W * img->H * img->D, n = w->n;
        a.gammatau = fr->gammatau;
        if (w->on_device) {
            a.field = fr->field;
        } else {
            RW_HIP(c, c->field.reserve(8 * vox));
            RW_HIP(c, hipMemcpyAsync(c->field.p, fr->field, 8 * vox, hipMemcpyHostToDevice, c->stream));
            a.field = c->field.as<double>();
        }
        if (w->on_device && (fr->p_alive != nullptr) != (fr->phase != nullptr))
            return fail(c, RWNMR_ERR_INVALID, "field run: give both p_alive and phase device arrays, or neither");
        if (w->on_device && fr->p_alive && fr->phase) {
            a.p_alive = fr->p_alive;
            a.phase = fr->phase;
        } else {
            RW_HIP(c, c->p_alive.reserve(8 * n + 8));
            RW_HIP(c, c->phase.reserve(8 * n + 8));
            a.p_alive = c->p_alive.as<double>();
            a.phase = c->phase.as<double>();
            const hipMemcpyKind kind = w->on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
            const uint32_t nb = (uint32_t)((n + 255) / 256);
            if (fr->p_alive)
                RW_HIP(c, hipMemcpyAsync(a.p_alive, fr->p_alive, 8 * n, kind, c->stream));
            else if (n)
                hipLaunchKernelGGL(fill_f64_kernel, dim3(nb), dim3(256), 0, c->stream, a.p_alive, 1.0, n);
            if (fr->phase)
                RW_HIP(c, hipMemcpyAsync(a.phase, fr->phase, 8 * n, kind, c->stream));
            else if (n)
                hipLaunchKernelGGL(fill_f64_kernel, dim3(nb), dim3(256), 0, c->stream, a.phase, 0.0, n);
            RW_HIP(c, hipGetLastError());
        }
    }
    if (emode == 2) {
        if (w->on_device) {
            a.label_delta = label_delta;
            a.label_hits = label_hits;
        } else {
            RW_HIP(c, c->label_delta.reserve(sizeof(double) * img->n_labels));
            RW_HIP(c, hipMemcpyAsync(c->label_delta.p, label_delta, sizeof(double) * img->n_labels,
                                     hipMemcpyHostToDevice, c->stream));
            a.label_delta = c->label_delta.as<double>();
            if (label_hits) {
                const size_t lb = sizeof(uint32_t) * (size_t)img->n_labels * w->n;
                RW_HIP(c, c->label_hits.reserve(lb + 4));
                RW_HIP(c, hipMemcpyAsync(c->label_hits.p, label_hits, lb, hipMemcpyHostToDevice, c->stream));
                a.label_hits = c->label_hits.as<uint32_t>();
            }
        }
        if (jac && !a.label_hits) {
            // the Jacobian needs each walker's hits since the start, kept
            // across launches: start them at zero in context memory
            const uint64_t cnt = (uint64_t)img->n_labels * w->n;
            RW_HIP(c, c->label_hits.reserve(4 * cnt + 4));
            a.label_hits = c->label_hits.as<uint32_t>();
            hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)((cnt + 255) / 256)), dim3(256), 0, c->stream,
                               a.label_hits, 0u, cnt);
            RW_HIP(c, hipGetLastError());
        }
    }
    attach_zimage(c, img, bc, emode, a, true, true);
    if (jac && (!a.zwords || !a.zlabels))
        return fail(c, RWNMR_ERR_INVALID, "Jacobian forward: needs the unfolded labelled image");
    // the Jacobian forward holds one walker per lane
    const LaunchShape shape = jac ? LaunchShape{c->shape.threads, 1u} : c->shape;
    // tiles, one echo-partial column each (balanced tiles where they apply)
    const uint32_t blocks = jac ? rwnmr::walk_grid_blocks(w->n, shape, (int)shift) : walk_columns(c, a, emode, shape);
    a.partial_stride = blocks;
    double *sums = nullptr;
    if (echo_sums) {
        if (w->on_device) {
            sums = echo_sums;
        } else {
            RW_HIP(c, c->sums.reserve(sizeof(double) * ((size_t)echoes + 1)));
            sums = c->sums.as<double>();
        }
    }
    // echoes per launch so that the partial buffer stays bounded
    uint64_t chunk = echoes ? echoes : 1;
    if (echo_sums && blocks) {
        // the Jacobian forward stages kMaxLabels more rows per echo
        const uint64_t per_row = 8ull * blocks * (jac ? 1 + rwnmr::kMaxLabels : 1);
        const uint64_t max_rows = std::max<uint64_t>(2, kMaxPartialBytes / per_row);
        chunk = std::min<uint64_t>(chunk, max_rows - 1);
    }
    if (echo_sums)
        RW_HIP(c, c->partials.reserve(sizeof(double) * (size_t)blocks * (size_t)(chunk + 1) + 8));
    double *jrows = nullptr;
    if (jac) {
        RW_HIP(c, c->jac_partials.reserve(sizeof(double) * rwnmr::kMaxLabels * (size_t)blocks * (size_t)(chunk + 1) +
                                          8));
        if (w->on_device) {
            jrows = jac;
        } else {
            RW_HIP(c, c->jac_rows.reserve(sizeof(double) * (size_t)img->n_labels * ((size_t)echoes + 1)));
            jrows = c->jac_rows.as<double>();
        }
    }
    const bool local = emode < 2 && shift == 0 && c->locality && w->n >= c->locality_min_walkers;
    if (local) {
        a.steps_per_echo = steps_per_echo;
        if ((s = walk_local(c, img, a, bc, emode, echoes, steps_per_echo, echo_sums ? sums : nullptr, chunk)) !=
            RWNMR_OK)
            return s;
    } else {
    if (a.energy == nullptr && chunk < echoes && w->n) {
        // energies must persist between launches: keep them in context memory
        RW_HIP(c, c->energy.reserve(w->n * 8 + 8));
        a.energy = c->energy.as<double>();
        const uint64_t nb = (w->n + 255) / 256;
        hipLaunchKernelGGL(fill_f64_kernel, dim3((uint32_t)nb), dim3(256), 0, c->stream, a.energy, 1.0, w->n);
        RW_HIP(c, hipGetLastError());
    }
    a.steps_per_echo = steps_per_echo;
    if (shift == 0) {
        // XCD-striped tiles: XCD j walks the j-th contiguous eighth of the
        // walkers, so position-ordered walkers keep one region of the image
        // per XCD L2
        a.tile_map = 1;
        a.n_tiles = blocks;
    }
    uint32_t done = 0;
    do {
        const uint32_t ne = (uint32_t)std::min<uint64_t>(chunk, (uint64_t)echoes - done);
        a.echoes = ne;
        a.partials = echo_sums ? c->partials.as<double>() : nullptr;
        a.jac_partials = jac ? c->jac_partials.as<double>() : nullptr;
        hipEvent_t t0 = timing_begin(c);
        if (jac)
            RW_HIP(c, rwnmr::launch_zwalk_jac(a, shape.threads, c->stream));
        else
            RW_HIP(c, launch_any_walk(a, bc, emode, c->shape, c->stream));
        timing_end(c, t0);
        if (echo_sums) {
            // rows 0..ne of this launch are echoes done..done+ne (row 0 repeats
            // the previous launch's last echo with identical bits)
            RW_HIP(c, rwnmr::launch_reduce_rows(a.partials, (uint64_t)ne + 1, blocks, sums + done, c->stream));
        }
        if (jac)
            for (int l = 0; l < img->n_labels; ++l)
                RW_HIP(c, rwnmr::launch_reduce_rows(a.jac_partials + (uint64_t)l * (ne + 1) * blocks, (uint64_t)ne + 1,
                                                    blocks, jrows + (uint64_t)l * (echoes + 1) + done, c->stream));
        done += ne;
    } while (done < echoes);
    }
    if (!w->on_device) {
        if (echo_sums)
            RW_HIP(c, hipMemcpyAsync(echo_sums, sums, sizeof(double) * ((size_t)echoes + 1), hipMemcpyDeviceToHost,
                                     c->stream));
        if (jac)
            RW_HIP(c, hipMemcpyAsync(jac, jrows, sizeof(double) * (size_t)img->n_labels * ((size_t)echoes + 1),
                                     hipMemcpyDeviceToHost, c->stream));
        if (emode == 2 && label_hits)
            RW_HIP(c, hipMemcpyAsync(label_hits, a.label_hits, sizeof(uint32_t) * (size_t)img->n_labels * w->n,
                                     hipMemcpyDeviceToHost, c->stream));
        if (emode == 3 && fr->p_alive)
            RW_HIP(c, hipMemcpyAsync(fr->p_alive, a.p_alive, 8 * w->n, hipMemcpyDeviceToHost, c->stream));
        if (emode == 3 && fr->phase)
            RW_HIP(c, hipMemcpyAsync(fr->phase, a.phase, 8 * w->n, hipMemcpyDeviceToHost, c->stream));
        if ((s = stage_out(c, w, v)) != RWNMR_OK)
            return s;
        return sync_and_check(c);
    }
    return RWNMR_OK;
}

rwnmr_status pfgse_common(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, const int32_t *x0,
                          const int32_t *y0, const int32_t *z0, const uint32_t *steps_inc, uint32_t n_delta,
                          const double *kx, const double *ky, const double *kz, uint32_t n_k, int32_t k_per_delta,
                          double voxel, int32_t bc, uint32_t shift, uint32_t n_sets, const double *set_delta,
                          double *M, double *msd, uint32_t n_samples, double *M_samples);

}  // namespace

extern "C" {

const char *rwnmr_version(void) { return "rwnmr-mi355x 0.1 (gfx950)"; }

rwnmr_status rwnmr_ctx_create(rwnmr_ctx **out, int32_t device)
{
    if (!out)
        return RWNMR_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return RWNMR_ERR_HIP;
    if (device < 0 || device >= count)
        return RWNMR_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess)
        return RWNMR_ERR_HIP;
    rwnmr_ctx *c = new (std::nothrow) rwnmr_ctx();
    if (!c)
        return RWNMR_ERR_ALLOC;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return RWNMR_ERR_HIP;
    }
    if (hipHostMalloc((void **)&c->host_flag, 2 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return RWNMR_ERR_ALLOC;
    }
    c->host_flag[0] = 0;
    c->host_flag[1] = 0;
    // the no-allocation entry point's scratch on this device (the device is current)
    c->scratch_err = scratch_address(device, &c->scratch);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->balanced_slots = 2u * (uint32_t)cus;
    *out = c;
    return RWNMR_OK;
}

void rwnmr_ctx_destroy(rwnmr_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (rwnmr_image *img : c->images) {
        free_zimages(img);
        if (img->blocks)
            (void)hipFree(img->blocks);
        if (img->labels_blocked)
            (void)hipFree(img->labels_blocked);
        img->blocks = nullptr;
        img->labels_blocked = nullptr;
        img->ctx = nullptr;
    }
    c->images.clear();
    for (DevBuf *b : {&c->x, &c->y, &c->z, &c->seed, &c->energy, &c->hits, &c->penalty, &c->x0, &c->y0, &c->z0,
                      &c->partials, &c->sums, &c->label_delta, &c->label_hits, &c->steps_inc, &c->kvec,
                      &c->set_delta, &c->flag, &c->batch_inst, &c->batch_tiles, &c->batch_rows,
                      &c->batch_counter, &c->lkeys, &c->lidx, &c->lsort_tmp, &c->field, &c->p_alive, &c->phase, &c->msamples,
                      &c->lset_x, &c->lset_y, &c->lset_z, &c->lset_seed, &c->lset_energy, &c->lset_energies,
                      &c->jac_partials, &c->jac_rows})
        b->release();
    for (int i = 0; i < 8; ++i) {
        c->sortA[i].release();
        c->sortB[i].release();
    }
    if (c->host_flag)
        (void)hipHostFree(c->host_flag);
    timing_collect(c);
    for (hipEvent_t e : c->event_pool)
        (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

rwnmr_status rwnmr_ctx_set_image_layout(rwnmr_ctx *c, int32_t sx, int32_t sy, int32_t sz)
{
    if (!c)
        return RWNMR_ERR_INVALID;
    if (sx < 0 || sy < 0 || sz < 0 || sx > 4 || sy > 4 || sz > 4)
        return fail(c, RWNMR_ERR_INVALID, "brick shifts must be in [0, 4]");
    c->brick_sx = sx;
    c->brick_sy = sy;
    c->brick_sz = sz;
    return RWNMR_OK;
}

rwnmr_status rwnmr_ctx_set_locality(rwnmr_ctx *c, int32_t enable, uint32_t resort_every, uint64_t min_walkers)
{
    if (!c)
        return RWNMR_ERR_INVALID;
    c->locality = enable != 0;
    c->resort_every = resort_every ? resort_every : 1000;
    c->locality_min_walkers = min_walkers;
    return RWNMR_OK;
}

rwnmr_status rwnmr_ctx_set_unfolded(rwnmr_ctx *c, int32_t enable)
{
    if (!c)
        return RWNMR_ERR_INVALID;
    if (enable < 0 || enable > 6)
        return fail(c, RWNMR_ERR_INVALID, "set_unfolded: mode must be 0 .. 6");
    c->unfolded = enable;
    return RWNMR_OK;
}

rwnmr_status rwnmr_ctx_set_timing(rwnmr_ctx *c, int32_t enable)
{
    if (!c)
        return RWNMR_ERR_INVALID;
    c->timing = enable != 0;
    return RWNMR_OK;
}

rwnmr_status rwnmr_ctx_get_timing(rwnmr_ctx *c, double *kernel_ms, uint64_t *launches, int32_t reset)
{
    if (!c || !kernel_ms || !launches)
        return RWNMR_ERR_INVALID;
    RW_HIP(c, hipStreamSynchronize(c->stream));
    timing_collect(c);
    *kernel_ms = c->kernel_ms;
    *launches = c->kernel_launches;
    if (reset) {
        c->kernel_ms = 0.0;
        c->kernel_launches = 0;
    }
    return RWNMR_OK;
}

const char *rwnmr_last_error(const rwnmr_ctx *c) { return c ? c->err.c_str() : "null context"; }

rwnmr_status rwnmr_ctx_synchronize(rwnmr_ctx *c)
{
    if (!c)
        return RWNMR_ERR_INVALID;
    return sync_and_check(c);
}

void *rwnmr_ctx_stream(rwnmr_ctx *c) { return c ? (void *)c->stream : nullptr; }

rwnmr_status rwnmr_ctx_set_launch(rwnmr_ctx *c, uint32_t tpb, uint32_t k)
{
    if (!c)
        return RWNMR_ERR_INVALID;
    if (tpb == 0)
        tpb = 256;
    if (k == 0)
        k = 1;
    if (tpb % 64 != 0 || !(k == 1 || k == 2 || k == 4) || tpb > rwnmr::max_threads_for_k((int)k))
        return fail(c, RWNMR_ERR_INVALID,
                    "threads_per_block must be a multiple of 64 and at most 1024 (k=1), 512 (k=2) or 256 (k=4)");
    c->shape.threads = tpb;
    c->shape.k = k;
    return RWNMR_OK;
}

rwnmr_status rwnmr_ctx_get_launch(const rwnmr_ctx *c, uint32_t *tpb, uint32_t *k)
{
    if (!c || !tpb || !k)
        return RWNMR_ERR_INVALID;
    *tpb = c->shape.threads;
    *k = c->shape.k;
    return RWNMR_OK;
}

rwnmr_status rwnmr_bitblock_pack(const uint8_t *solid, int32_t W, int32_t H, int32_t D, uint64_t *blocks)
{
    if (!solid || !blocks || W <= 0 || H <= 0 || D <= 0)
        return RWNMR_ERR_INVALID;
    const int32_t bC = (W + 3) / 4, bR = (H + 3) / 4, bD = (D + 3) / 4;
    const size_t nb = (size_t)bC * bR * bD;
    memset(blocks, 0, nb * sizeof(uint64_t));
    for (int32_t z = 0; z < D; ++z)
        for (int32_t y = 0; y < H; ++y) {
            const uint8_t *row = solid + ((size_t)z * H + y) * W;
            const size_t brow = (size_t)(y >> 2) * bC + (size_t)(z >> 2) * bC * bR;
            const int bitrow = ((y & 3) << 2) | ((z & 3) << 4);
            for (int32_t x = 0; x < W; ++x)
                if (row[x])
                    blocks[brow + (x >> 2)] |= 1ull << (bitrow | (x & 3));
        }
    return RWNMR_OK;
}

rwnmr_status rwnmr_image_upload(rwnmr_ctx *c, const uint64_t *blocks, int32_t W, int32_t H, int32_t D,
                                const uint8_t *labels, int32_t n_labels, rwnmr_image **out)
{
    if (!c)
        return RWNMR_ERR_INVALID;
    if (!out || !blocks || W <= 0 || H <= 0 || D <= 0)
        return fail(c, RWNMR_ERR_INVALID, "image_upload: null pointer or non-positive dimension");
    if (labels && (n_labels <= 0 || n_labels > rwnmr::kMaxLabels))
        return fail(c, RWNMR_ERR_INVALID, "image_upload: n_labels must be in [1, 8]");
    *out = nullptr;
    rwnmr_image *img = new (std::nothrow) rwnmr_image();
    if (!img)
        return fail(c, RWNMR_ERR_ALLOC, "image_upload: host allocation");
    img->ctx = c;
    img->W = W;
    img->H = H;
    img->D = D;
    img->bC = (W + 3) / 4;
    img->bR = (H + 3) / 4;
    img->bD = (D + 3) / 4;
    img->nblocks = (uint64_t)img->bC * img->bR * img->bD;
    if (img->nblocks > 0xFFFFFFFFull) {
        delete img;
        return fail(c, RWNMR_ERR_INVALID, "image_upload: more than 2^32 blocks");
    }
    // device layout: bricks of 2^sx x 2^sy x 2^sz blocks, brick grid padded
    rwnmr::BrickLayout &L = img->layout;
    L.sx = c->brick_sx;
    L.sy = c->brick_sy;
    L.sz = c->brick_sz;
    const int nbx = (img->bC + (1 << L.sx) - 1) >> L.sx;
    const int nby = (img->bR + (1 << L.sy) - 1) >> L.sy;
    const int nbz = (img->bD + (1 << L.sz) - 1) >> L.sz;
    L.nbx = nbx;
    L.nbxy = nbx * nby;
    img->nwords = ((uint64_t)nbx * nby * nbz) << (L.sx + L.sy + L.sz);
    if (img->nwords > 0xFFFFFFFFull) {
        delete img;
        return fail(c, RWNMR_ERR_INVALID, "image_upload: more than 2^32 blocks after brick padding");
    }
    // device word of every reference block
    std::vector<uint32_t> where(img->nblocks);
    for (int32_t bz = 0; bz < img->bD; ++bz)
        for (int32_t by = 0; by < img->bR; ++by)
            for (int32_t bx = 0; bx < img->bC; ++bx)
                where[(size_t)bx + (size_t)by * img->bC + (size_t)bz * img->bC * img->bR] =
                    rwnmr::brick_block_index(bx << 2, by << 2, bz << 2, L);
    std::vector<uint64_t> dev_words(img->nwords, 0ull);
    for (uint64_t b = 0; b < img->nblocks; ++b)
        dev_words[where[b]] = blocks[b];
    hipError_t e = hipMalloc((void **)&img->blocks, img->nwords * 8);
    if (e == hipSuccess)
        e = hipMemcpyAsync(img->blocks, dev_words.data(), img->nwords * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && labels) {
        for (size_t i = 0; i < (size_t)W * H * D; ++i)
            if (labels[i] >= (uint8_t)n_labels) {
                (void)hipStreamSynchronize(c->stream);
                (void)hipFree(img->blocks);
                delete img;
                return fail(c, RWNMR_ERR_INVALID, "image_upload: label out of range");
            }
        std::vector<uint8_t> lb(img->nwords * 64, 0);
        for (int32_t z = 0; z < D; ++z)
            for (int32_t y = 0; y < H; ++y)
                for (int32_t x = 0; x < W; ++x) {
                    const size_t b = where[(size_t)(x >> 2) + (size_t)(y >> 2) * img->bC +
                                           (size_t)(z >> 2) * img->bC * img->bR];
                    const int bit = (x & 3) | ((y & 3) << 2) | ((z & 3) << 4);
                    lb[b * 64 + bit] = labels[((size_t)z * H + y) * W + x];
                }
        e = hipMalloc((void **)&img->labels_blocked, lb.size());
        if (e == hipSuccess)
            e = hipMemcpyAsync(img->labels_blocked, lb.data(), lb.size(), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        img->n_labels = n_labels;
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        if (img->blocks)
            (void)hipFree(img->blocks);
        if (img->labels_blocked)
            (void)hipFree(img->labels_blocked);
        delete img;
        return fail(c, e == hipErrorOutOfMemory ? RWNMR_ERR_ALLOC : RWNMR_ERR_HIP,
                    std::string("image_upload: ") + hipGetErrorString(e));
    }
    c->images.push_back(img);
    *out = img;
    return RWNMR_OK;
}

void rwnmr_image_destroy(rwnmr_image *img)
{
    if (!img)
        return;
    if (img->ctx) {
        rwnmr_ctx *c = img->ctx;
        (void)hipStreamSynchronize(c->stream);
        c->images.erase(std::remove(c->images.begin(), c->images.end(), img), c->images.end());
        free_zimages(img);
        if (img->blocks)
            (void)hipFree(img->blocks);
        if (img->labels_blocked)
            (void)hipFree(img->labels_blocked);
    }
    // an image whose context is gone was already released by rwnmr_ctx_destroy
    delete img;
}

rwnmr_status rwnmr_map_run(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, uint64_t steps, int32_t bc,
                           uint32_t shift)
{
    rwnmr_status s;
    if (!c)
        return RWNMR_ERR_INVALID;
    if ((s = check_image(c, img)) != RWNMR_OK)
        return s;
    if ((s = check_walkers(c, img, w, bc, shift)) != RWNMR_OK)
        return s;
    if (!w->hits)
        return fail(c, RWNMR_ERR_INVALID, "map_run needs a hits array");
    WalkerView v;
    if ((s = clear_flag(c)) != RWNMR_OK)
        return s;
    if ((s = stage_in(c, w, v)) != RWNMR_OK)
        return s;
    WalkArgs a = base_args(img, v, w, shift, c->flag.as<uint32_t>());
    a.energy = nullptr;  // the map walk never touches energies
    a.echoes = 1;
    attach_zimage(c, img, bc, 0, a, true, true);
    uint64_t left = steps;
    while (left > 0) {
        const uint32_t now = (uint32_t)std::min<uint64_t>(left, 1ull << 30);
        a.steps_per_echo = now;
        hipEvent_t t0 = timing_begin(c);
        RW_HIP(c, launch_any_walk(a, bc, 0, c->shape, c->stream));
        timing_end(c, t0);
        left -= now;
    }
    if (!w->on_device) {
        double *saved = w->energy;
        w->energy = nullptr;
        s = stage_out(c, w, v);
        w->energy = saved;
        if (s != RWNMR_OK)
            return s;
        return sync_and_check(c);
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_ctx_autotune(rwnmr_ctx *c, const rwnmr_image *img, const rwnmr_walkers *w,
                                uint32_t steps_per_echo, int32_t bc, uint32_t trial_echoes, uint32_t *best_threads,
                                uint32_t *best_k, double *best_ms)
{
    rwnmr_status s;
    if (!c)
        return RWNMR_ERR_INVALID;
    if ((s = check_image(c, img)) != RWNMR_OK)
        return s;
    if ((s = check_walkers(c, img, w, bc, 0)) != RWNMR_OK)
        return s;
    if (w->n == 0 || trial_echoes == 0 || steps_per_echo == 0)
        return fail(c, RWNMR_ERR_INVALID, "autotune needs walkers, echoes and steps");
    // trial copies of the walkers, so the caller's state is untouched
    const uint64_t n = w->n;
    DevBuf tx, ty, tz, ts, te, tsum;
    auto cleanup = [&]() {
        for (DevBuf *b : {&tx, &ty, &tz, &ts, &te, &tsum})
            b->release();
    };
    const hipMemcpyKind kind = w->on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (tx.reserve(n * 4) != hipSuccess || ty.reserve(n * 4) != hipSuccess || tz.reserve(n * 4) != hipSuccess ||
        ts.reserve(n * 8) != hipSuccess || te.reserve(n * 8) != hipSuccess ||
        tsum.reserve(8ull * (trial_echoes + 1)) != hipSuccess) {
        cleanup();
        return fail(c, RWNMR_ERR_ALLOC, "autotune: trial buffers");
    }
    const bool was_timing = c->timing;
    const double saved_ms = c->kernel_ms;
    const uint64_t saved_launches = c->kernel_launches;
    const LaunchShape saved = c->shape;
    const int saved_unfolded = c->unfolded;
    const uint32_t saved_resort = c->resort_every;
    auto restore = [&]() {
        c->timing = was_timing;
        c->kernel_ms = saved_ms;
        c->kernel_launches = saved_launches;
        c->shape = saved;
        c->unfolded = saved_unfolded;
        c->resort_every = saved_resort;
        cleanup();
    };
    c->timing = true;
    (void)hipStreamSynchronize(c->stream);
    timing_collect(c);
    // One trial run from the caller's start state; returns the walk kernels'
    // milliseconds (whole == false) or the whole run's, re-sorts included.
    auto trial = [&](double &ms, bool whole) -> rwnmr_status {
        rwnmr_walkers tw = *w;
        tw.x = tx.as<int32_t>();
        tw.y = ty.as<int32_t>();
        tw.z = tz.as<int32_t>();
        tw.seed = ts.as<uint64_t>();
        tw.energy = te.as<double>();
        tw.hits = nullptr;
        tw.on_device = 1;
        if (tw.penalty && !w->on_device)
            tw.penalty = nullptr;  // host penalties: time with the uniform delta instead
        hipError_t e = hipMemcpyAsync(tw.x, w->x, n * 4, kind, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(tw.y, w->y, n * 4, kind, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(tw.z, w->z, n * 4, kind, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(tw.seed, w->seed, n * 8, kind, c->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(fill_f64_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, c->stream,
                               tw.energy, 1.0, n);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess)
            return fail(c, RWNMR_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
        c->kernel_ms = 0.0;
        c->kernel_launches = 0;
        hipEvent_t t0 = whole ? take_event(c) : nullptr, t1 = whole ? take_event(c) : nullptr;
        if (whole && (!t0 || !t1))
            return fail(c, RWNMR_ERR_HIP, "autotune: no events");
        if (whole)
            (void)hipEventRecord(t0, c->stream);
        rwnmr_status st = cpmg_common(c, img, &tw, trial_echoes, steps_per_echo, bc, 0, 1, nullptr, nullptr,
                                      tsum.as<double>());
        if (whole)
            (void)hipEventRecord(t1, c->stream);
        if (st == RWNMR_OK)
            st = sync_and_check(c);
        ms = c->kernel_ms;
        if (whole) {
            float f = 0.f;
            if (st == RWNMR_OK && hipEventElapsedTime(&f, t0, t1) == hipSuccess)
                ms = f;
            c->event_pool.push_back(t0);
            c->event_pool.push_back(t1);
        }
        return st;
    };
    // 1. launch shape: threads per block x walkers per lane
    LaunchShape best = saved;
    double best_t = 1e300;
    const uint32_t threads[] = {64, 128, 256, 512, 1024};
    const uint32_t ks[] = {1, 2, 4};
    for (uint32_t t : threads)
        for (uint32_t k : ks) {
            if (t > rwnmr::max_threads_for_k((int)k))
                continue;
            c->shape = LaunchShape{t, k};
            double ms = 0.0;
            if ((s = trial(ms, false)) != RWNMR_OK) {
                restore();
                return s;
            }
            // more walkers per lane must win by more than 3%: short trials on a
            // cool device favour them, sustained runs do not
            const double margin = k > best.k ? 0.97 : 1.0;
            if (ms < best_t * margin) {
                best_t = ms;
                best = LaunchShape{t, k};
            }
        }
    c->shape = best;
    // 2. image form (how the bitBlock is staged for the walk): the unfolded
    // Z-order image with the register chunk cache (the default) against the
    // folded brick-ordered bitBlock words
    int best_form = saved_unfolded;
    if (saved_unfolded != 0) {
        c->unfolded = 0;
        double ms = 0.0;
        if ((s = trial(ms, false)) != RWNMR_OK) {
            restore();
            return s;
        }
        if (ms < best_t) {
            best_t = ms;
            best_form = 0;
        }
        c->unfolded = best_form;
    }
    // mirror walks: the 8-byte (4 x 4 x 4 voxel) chunk cache (form 6; 3 VALU
    // per step fewer, a quarter more chunk loads) against the 16-byte one.
    // They differ by a few percent, so each is timed twice, in mirrored order,
    // and keeps its better time
    if (bc == rwnmr::kMirror && best_form == 3 && saved_unfolded >= 3) {
        double form_ms[2] = {1e300, 1e300};
        const int forms[2] = {3, 6};
        for (int t = 0; t < 4; ++t) {
            const int fi = (t == 1 || t == 2) ? 1 : 0;
            c->unfolded = forms[fi];
            double ms = 0.0;
            if ((s = trial(ms, false)) != RWNMR_OK) {
                restore();
                return s;
            }
            form_ms[fi] = std::min(form_ms[fi], ms);
        }
        if (form_ms[1] < form_ms[0]) {
            best_form = 6;
            best_t = form_ms[1];
        } else {
            best_t = std::min(best_t, form_ms[0]);
        }
        c->unfolded = best_form;
    }
    // runs of at most one wave per SIMD are latency-bound: the speculative
    // two-candidate walk (form 5, one walker per lane) against the rest
    if (saved_unfolded >= 3 && c->shape.k == 1 && n <= kSpecMaxWalkers) {
        c->unfolded = 5;
        double ms = 0.0;
        if ((s = trial(ms, false)) != RWNMR_OK) {
            restore();
            return s;
        }
        if (ms < best_t) {
            best_t = ms;
            best_form = 5;
        }
        c->unfolded = best_form;
    }
    // 3. echoes per launch of locality-ordered runs (the re-sort interval),
    // timed over the whole run with its re-sorts; the candidates differ by
    // about 1%, so each is timed twice, in mirrored order (drifts of the
    // clock over the phase cancel), and keeps its better time
    uint32_t best_resort = saved_resort;
    if (c->locality && n >= c->locality_min_walkers && trial_echoes >= 4) {
        const uint32_t cands[3] = {trial_echoes / 4, trial_echoes / 2, trial_echoes};
        double cand_ms[3] = {1e300, 1e300, 1e300};
        for (int t = 0; t < 6; ++t) {
            const int ci = t < 3 ? t : 5 - t;
            c->resort_every = std::max<uint32_t>(1u, cands[ci]);
            double ms = 0.0;
            if ((s = trial(ms, true)) != RWNMR_OK) {
                restore();
                return s;
            }
            cand_ms[ci] = std::min(cand_ms[ci], ms);
        }
        double best_whole = 1e300;
        for (int ci = 0; ci < 3; ++ci)
            if (cand_ms[ci] < best_whole) {
                best_whole = cand_ms[ci];
                best_resort = std::max<uint32_t>(1u, cands[ci]);
            }
    }
    c->timing = was_timing;
    c->kernel_ms = saved_ms;
    c->kernel_launches = saved_launches;
    c->shape = best;
    c->unfolded = best_form;
    c->resort_every = best_resort;
    cleanup();
    if (best_threads)
        *best_threads = best.threads;
    if (best_k)
        *best_k = best.k;
    if (best_ms)
        *best_ms = best_t;
    return RWNMR_OK;
}

rwnmr_status rwnmr_ctx_get_tuning(const rwnmr_ctx *c, uint32_t *threads_per_block, uint32_t *walkers_per_lane,
                                  int32_t *unfolded, uint32_t *resort_every)
{
    if (!c)
        return RWNMR_ERR_INVALID;
    if (threads_per_block)
        *threads_per_block = c->shape.threads;
    if (walkers_per_lane)
        *walkers_per_lane = c->shape.k;
    if (unfolded)
        *unfolded = c->unfolded;
    if (resort_every)
        *resort_every = c->resort_every;
    return RWNMR_OK;
}

rwnmr_status rwnmr_cpmg_run(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                            uint32_t steps_per_echo, int32_t bc, uint32_t shift, double *echo_sums)
{
    return cpmg_common(c, img, w, echoes, steps_per_echo, bc, shift, 1, nullptr, nullptr, echo_sums);
}

rwnmr_status rwnmr_cpmg_run_noalloc(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                    uint32_t steps_per_echo, int32_t bc, uint32_t shift, double *echo_sums)
{
    rwnmr_status s;
    if (!c)
        return RWNMR_ERR_INVALID;
    if ((s = check_image(c, img)) != RWNMR_OK)
        return s;
    if ((s = check_walkers(c, img, w, bc, shift)) != RWNMR_OK)
        return s;
    if (!w->on_device)
        return fail(c, RWNMR_ERR_INVALID, "cpmg_run_noalloc: walker arrays must be device memory (on_device = 1)");
    if (!c->scratch)
        return fail(c, RWNMR_ERR_HIP, "cpmg_run_noalloc: no device scratch: " + c->scratch_err);
    uint32_t *flag = static_cast<uint32_t *>(c->scratch);
    double *part = reinterpret_cast<double *>(static_cast<unsigned char *>(c->scratch) + kScratchFlagBytes);
    const uint64_t part_doubles = (kScratchBytes - kScratchFlagBytes) / sizeof(double);
    const uint32_t blocks = rwnmr::walk_grid_blocks(w->n, c->shape, (int)shift);
    // echoes per launch: rows 0..chunk of [echo][block] partials must fit the scratch
    uint64_t chunk = echoes ? echoes : 1;
    if (echo_sums && blocks) {
        const uint64_t rows = part_doubles / blocks;
        if (rows < 2)
            return fail(c, RWNMR_ERR_INVALID,
                        "cpmg_run_noalloc: " + std::to_string(w->n) + " walkers need " +
                            std::to_string(16ull * blocks) + " bytes of echo partials per launch; the scratch holds " +
                            std::to_string(part_doubles * sizeof(double)) + " (at most " +
                            std::to_string(part_doubles / 2) + " blocks of walkers)");
        chunk = std::min<uint64_t>(chunk, rows - 1);
    }
    if (chunk < echoes && !w->energy && w->n)
        return fail(c, RWNMR_ERR_INVALID,
                    "cpmg_run_noalloc: a run of " + std::to_string(echoes) + " echoes takes several launches (" +
                        std::to_string(chunk) + " echoes each) and needs a device energy array to carry the "
                        "energies between them");
    // the error word starts at zero on every call (a fill kernel, no memset node)
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, c->stream, flag, 0u, 1ull);
    RW_HIP(c, hipGetLastError());
    c->scratch_flag_pending = true;
    const WalkerView v{w->x, w->y, w->z, w->seed, w->energy, w->hits, w->penalty};
    WalkArgs a = base_args(img, v, w, shift, flag);
    a.n_labels = img->n_labels;
    // an unfolded image is used only when an earlier call built it
    attach_zimage(c, img, bc, 1, a, true, true, false);
    a.partial_stride = blocks;
    a.steps_per_echo = steps_per_echo;
    uint32_t done = 0;
    do {
        const uint32_t ne = (uint32_t)std::min<uint64_t>(chunk, (uint64_t)echoes - done);
        a.echoes = ne;
        a.partials = echo_sums ? part : nullptr;
        RW_HIP(c, launch_any_walk(a, bc, 1, c->shape, c->stream));
        if (echo_sums)
            RW_HIP(c, rwnmr::launch_reduce_rows(part, (uint64_t)ne + 1, blocks, echo_sums + done, c->stream));
        done += ne;
    } while (done < echoes);
    return RWNMR_OK;
}

}  // extern "C"

// A CPMG run captured as a HIP graph: resets of the walkers from start
// arrays (optional), the error-word clear, the walk launches and the echo
// reductions, replayed by one hipGraphLaunch.  The graph owns its echo
// partials (allocated before the capture), so nothing allocates in the
// graph and no context buffer can move under it; it records fill kernels
// and copies only (no memset nodes).
struct rwnmr_graph {
    rwnmr_ctx *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double *partials = nullptr;
};

namespace {

void graph_release(rwnmr_graph *g)
{
    if (g->exec)
        (void)hipGraphExecDestroy(g->exec);
    if (g->graph)
        (void)hipGraphDestroy(g->graph);
    if (g->partials)
        (void)hipFree(g->partials);
    delete g;
}

}  // namespace

extern "C" {

rwnmr_status rwnmr_cpmg_graph_create(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w,
                                     const rwnmr_walkers *start, uint32_t echoes, uint32_t steps_per_echo, int32_t bc,
                                     double *echo_sums, rwnmr_graph **out)
{
    rwnmr_status s;
    if (!c || !out)
        return RWNMR_ERR_INVALID;
    *out = nullptr;
    if ((s = check_image(c, img)) != RWNMR_OK)
        return s;
    if ((s = check_walkers(c, img, w, bc, 0)) != RWNMR_OK)
        return s;
    if (!w->on_device || (start && (!start->on_device || start->n != w->n || !start->x || !start->y || !start->z ||
                                    !start->seed)))
        return fail(c, RWNMR_ERR_INVALID,
                    "cpmg_graph_create: device walkers, and start arrays (if given) of the same count on the device");
    const uint64_t n = w->n;
    const uint32_t blocks = rwnmr::walk_grid_blocks(n, c->shape, 0);
    uint64_t chunk = echoes ? echoes : 1;
    if (echo_sums && blocks)
        chunk = std::min<uint64_t>(chunk, std::max<uint64_t>(2, kMaxPartialBytes / (8ull * blocks)) - 1);
    if (chunk < echoes && !w->energy && n)
        return fail(c, RWNMR_ERR_INVALID, "cpmg_graph_create: a run of several launches needs a device energy array");
    // everything the graph touches exists before the capture: the error word,
    // the unfolded image and the graph's own partials
    RW_HIP(c, c->flag.reserve(sizeof(uint32_t)));
    const WalkerView v{w->x, w->y, w->z, w->seed, w->energy, w->hits, w->penalty};
    WalkArgs a = base_args(img, v, w, 0, c->flag.as<uint32_t>());
    a.n_labels = img->n_labels;
    attach_zimage(c, img, bc, 1, a, true, true, true);
    a.partial_stride = blocks;
    a.steps_per_echo = steps_per_echo;
    a.tile_map = 1;
    a.n_tiles = blocks;
    RW_HIP(c, hipStreamSynchronize(c->stream));
    rwnmr_graph *g = new (std::nothrow) rwnmr_graph();
    if (!g)
        return fail(c, RWNMR_ERR_ALLOC, "cpmg_graph_create: no host memory");
    g->ctx = c;
    if (echo_sums && blocks &&
        hipMalloc((void **)&g->partials, 8ull * blocks * (chunk + 1)) != hipSuccess) {
        (void)hipGetLastError();
        graph_release(g);
        return fail(c, RWNMR_ERR_ALLOC, "cpmg_graph_create: no device memory for the echo partials");
    }
    auto capture = [&]() -> rwnmr_status {
        if (start) {
            RW_HIP(c, hipMemcpyAsync(w->x, start->x, 4 * n, hipMemcpyDeviceToDevice, c->stream));
            RW_HIP(c, hipMemcpyAsync(w->y, start->y, 4 * n, hipMemcpyDeviceToDevice, c->stream));
            RW_HIP(c, hipMemcpyAsync(w->z, start->z, 4 * n, hipMemcpyDeviceToDevice, c->stream));
            RW_HIP(c, hipMemcpyAsync(w->seed, start->seed, 8 * n, hipMemcpyDeviceToDevice, c->stream));
            const uint32_t nb = (uint32_t)((n + 255) / 256);
            if (w->energy) {
                if (start->energy)
                    RW_HIP(c, hipMemcpyAsync(w->energy, start->energy, 8 * n, hipMemcpyDeviceToDevice, c->stream));
                else if (n)
                    hipLaunchKernelGGL(fill_f64_kernel, dim3(nb), dim3(256), 0, c->stream, w->energy, 1.0, n);
            }
            if (w->hits) {
                if (start->hits)
                    RW_HIP(c, hipMemcpyAsync(w->hits, start->hits, 4 * n, hipMemcpyDeviceToDevice, c->stream));
                else if (n)
                    hipLaunchKernelGGL(fill_u32_kernel, dim3(nb), dim3(256), 0, c->stream, w->hits, 0u, n);
            }
            RW_HIP(c, hipGetLastError());
        }
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, c->stream, c->flag.as<uint32_t>(), 0u, 1ull);
        RW_HIP(c, hipGetLastError());
        uint32_t done = 0;
        do {
            const uint32_t ne = (uint32_t)std::min<uint64_t>(chunk, (uint64_t)echoes - done);
            a.echoes = ne;
            a.partials = echo_sums ? g->partials : nullptr;
            RW_HIP(c, launch_any_walk(a, bc, 1, c->shape, c->stream));
            if (echo_sums)
                RW_HIP(c, rwnmr::launch_reduce_rows(g->partials, (uint64_t)ne + 1, blocks, echo_sums + done,
                                                    c->stream));
            done += ne;
        } while (done < echoes);
        return RWNMR_OK;
    };
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        graph_release(g);
        return fail(c, RWNMR_ERR_HIP, "cpmg_graph_create: hipStreamBeginCapture failed");
    }
    s = capture();
    const hipError_t e = hipStreamEndCapture(c->stream, &g->graph);
    if (s != RWNMR_OK || e != hipSuccess) {
        (void)hipGetLastError();
        graph_release(g);
        return s != RWNMR_OK ? s : fail(c, RWNMR_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        graph_release(g);
        return fail(c, RWNMR_ERR_HIP, "cpmg_graph_create: hipGraphInstantiate failed");
    }
    *out = g;
    return RWNMR_OK;
}

rwnmr_status rwnmr_graph_launch(rwnmr_graph *g)
{
    if (!g || !g->exec)
        return RWNMR_ERR_INVALID;
    RW_HIP(g->ctx, hipGraphLaunch(g->exec, g->ctx->stream));
    return RWNMR_OK;
}

void rwnmr_graph_destroy(rwnmr_graph *g)
{
    if (!g)
        return;
    (void)hipStreamSynchronize(g->ctx->stream);
    graph_release(g);
}

rwnmr_status rwnmr_cpmg_field_run(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                  uint32_t steps_per_echo, int32_t bc, const double *field,
                                  double gyromagnetic_ratio, double time_interval_ms, double *p_alive, double *phase,
                                  double *echo_sums)
{
    FieldRun fr;
    fr.field = field;
    // gamma = 1e6 * giromagnetic ratio, tau = 1e-3 * time interval, half of their product per
    // phase sample (NMR_cpmg_cuda.cpp:1010-1011 and the kernels' gammatau)
    fr.gammatau = 0.5 * (1.0e6 * gyromagnetic_ratio) * (1.0e-3 * time_interval_ms);
    fr.p_alive = p_alive;
    fr.phase = phase;
    return cpmg_common(c, img, w, echoes, steps_per_echo, bc, 0, 3, nullptr, nullptr, echo_sums, &fr);
}

rwnmr_status rwnmr_cpmg_run_labels(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                   uint32_t steps_per_echo, int32_t bc, uint32_t shift, const double *label_delta,
                                   uint32_t *label_hits, double *echo_sums)
{
    return cpmg_common(c, img, w, echoes, steps_per_echo, bc, shift, 2, label_delta, label_hits, echo_sums);
}

rwnmr_status rwnmr_cpmg_forward_jacobian(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                         uint32_t steps_per_echo, int32_t bc, const double *label_delta,
                                         uint32_t *label_hits, double *echo_sums, double *jac)
{
    if (!c)
        return RWNMR_ERR_INVALID;
    if (!jac || !echo_sums)
        return fail(c, RWNMR_ERR_INVALID, "forward_jacobian: null echo_sums or jac");
    return cpmg_common(c, img, w, echoes, steps_per_echo, bc, 0, 2, label_delta, label_hits, echo_sums, nullptr, jac);
}

rwnmr_status rwnmr_cpmg_run_label_sets(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                       uint32_t steps_per_echo, int32_t bc, uint32_t n_sets,
                                       const double *set_label_delta, double *echo_sums)
{
    rwnmr_status s;
    if (!c)
        return RWNMR_ERR_INVALID;
    if ((s = check_image(c, img)) != RWNMR_OK)
        return s;
    if ((s = check_walkers(c, img, w, bc, 0)) != RWNMR_OK)
        return s;
    if (!img->labels_blocked || img->n_labels <= 0)
        return fail(c, RWNMR_ERR_INVALID, "label sets: image has no solid labels");
    if (n_sets == 0 || !set_label_delta || !echo_sums)
        return fail(c, RWNMR_ERR_INVALID, "label sets: need n_sets > 0, set_label_delta and echo_sums");
    const uint64_t n = w->n;
    const size_t nl = (size_t)img->n_labels, rows = (size_t)echoes + 1;
    if (n == 0) {
        std::vector<double> zero(rows * n_sets, 0.0);
        if (w->on_device)
            RW_HIP(c, hipMemcpyAsync(echo_sums, zero.data(), 8 * zero.size(), hipMemcpyHostToDevice, c->stream));
        else
            memcpy(echo_sums, zero.data(), 8 * zero.size());
        return w->on_device ? RWNMR_OK : sync_and_check(c);
    }
    WalkerView v;
    if ((s = clear_flag(c)) != RWNMR_OK)
        return s;
    if ((s = stage_in(c, w, v)) != RWNMR_OK)
        return s;
    const hipMemcpyKind in_kind = w->on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    RW_HIP(c, c->label_delta.reserve(8 * nl * n_sets));
    RW_HIP(c, hipMemcpyAsync(c->label_delta.p, set_label_delta, 8 * nl * n_sets, in_kind, c->stream));
    double *sums = echo_sums;
    if (!w->on_device) {
        RW_HIP(c, c->sums.reserve(8 * rows * n_sets));
        sums = c->sums.as<double>();
    }
    WalkArgs a = base_args(img, v, w, 0, c->flag.as<uint32_t>());
    a.n_labels = img->n_labels;
    attach_zimage(c, img, bc, 2, a, true, true);
    // the start state, replayed for every set (fallback) or group of 8 sets
    auto save_start = [&]() -> rwnmr_status {
        RW_HIP(c, c->lset_x.reserve(4 * n));
        RW_HIP(c, c->lset_y.reserve(4 * n));
        RW_HIP(c, c->lset_z.reserve(4 * n));
        RW_HIP(c, c->lset_seed.reserve(8 * n));
        RW_HIP(c, hipMemcpyAsync(c->lset_x.p, v.x, 4 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(c->lset_y.p, v.y, 4 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(c->lset_z.p, v.z, 4 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(c->lset_seed.p, v.seed, 8 * n, hipMemcpyDeviceToDevice, c->stream));
        if (v.energy) {
            RW_HIP(c, c->lset_energy.reserve(8 * n));
            RW_HIP(c, hipMemcpyAsync(c->lset_energy.p, v.energy, 8 * n, hipMemcpyDeviceToDevice, c->stream));
        }
        return RWNMR_OK;
    };
    auto restore_start = [&]() -> rwnmr_status {
        RW_HIP(c, hipMemcpyAsync(v.x, c->lset_x.p, 4 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(v.y, c->lset_y.p, 4 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(v.z, c->lset_z.p, 4 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(v.seed, c->lset_seed.p, 8 * n, hipMemcpyDeviceToDevice, c->stream));
        if (v.energy)
            RW_HIP(c, hipMemcpyAsync(v.energy, c->lset_energy.p, 8 * n, hipMemcpyDeviceToDevice, c->stream));
        return RWNMR_OK;
    };
    if (!a.zwords || !a.zlabels) {
        // no unfolded image: one labelled walk per set from the same start state
        if (n_sets > 1 && (s = save_start()) != RWNMR_OK)
            return s;
        // the set whose energies are returned (the first of the last group of 8) walks last
        const uint32_t last = 8u * ((n_sets - 1u) / 8u);
        for (uint32_t i = 0; i < n_sets; ++i) {
            const uint32_t q = i + 1 < n_sets ? (i < last ? i : i + 1) : last;
            if (i > 0 && (s = restore_start()) != RWNMR_OK)
                return s;
            rwnmr_walkers ws = *w;
            ws.on_device = 1;
            ws.x = v.x;
            ws.y = v.y;
            ws.z = v.z;
            ws.seed = v.seed;
            ws.energy = v.energy;
            ws.hits = nullptr;
            ws.penalty = nullptr;
            if ((s = cpmg_common(c, img, &ws, echoes, steps_per_echo, bc, 0, 2,
                                 c->label_delta.as<double>() + (size_t)q * nl, nullptr, sums + (size_t)q * rows)) !=
                RWNMR_OK)
                return s;
        }
    } else {
        const uint32_t threads = c->shape.threads;
        const uint32_t blocks = (uint32_t)((n + threads - 1) / threads);
        // echoes per launch: 8 sets x (chunk + 1) rows of partials within the bound
        const uint64_t max_rows = std::max<uint64_t>(2, kMaxPartialBytes / (8ull * 8ull * blocks));
        const uint64_t chunk = std::min<uint64_t>(echoes ? echoes : 1, max_rows - 1);
        const uint32_t groups = (n_sets + 7) / 8;
        RW_HIP(c, c->partials.reserve(8ull * 8ull * (chunk + 1) * blocks + 8));
        const bool carry = chunk < echoes;
        if (carry)
            RW_HIP(c, c->lset_energies.reserve(8ull * 8ull * n));
        if (groups > 1 && (s = save_start()) != RWNMR_OK)
            return s;
        for (uint32_t g = 0; g < groups; ++g) {
            // every group walks the same trajectories from the start state
            if (g > 0 && (s = restore_start()) != RWNMR_OK)
                return s;
            if (carry) {
                // the 8 carried set energies start at the walkers' start energies
                for (int q = 0; q < 8; ++q) {
                    double *row = c->lset_energies.as<double>() + (size_t)q * n;
                    if (v.energy)
                        RW_HIP(c, hipMemcpyAsync(row, v.energy, 8 * n, hipMemcpyDeviceToDevice, c->stream));
                    else
                        hipLaunchKernelGGL(fill_f64_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0,
                                           c->stream, row, 1.0, n);
                }
                RW_HIP(c, hipGetLastError());
            }
            const uint32_t gsets = std::min<uint32_t>(8u, n_sets - 8u * g);
            WalkArgs b = a;
            b.label_delta = c->label_delta.as<double>() + (size_t)8 * g * nl;
            b.n_label_sets = gsets;
            b.partial_stride = blocks;
            b.tile_map = 1;   // XCD-striped tiles (position-ordered walkers keep an image region per L2)
            b.n_tiles = blocks;
            b.steps_per_echo = steps_per_echo;
            b.set_energy = carry ? c->lset_energies.as<double>() : nullptr;
            double *part = c->partials.as<double>();
            uint32_t done = 0;
            do {
                const uint32_t ne = (uint32_t)std::min<uint64_t>(chunk, (uint64_t)echoes - done);
                b.echoes = ne;
                b.partials = part;
                hipEvent_t t0 = timing_begin(c);
                RW_HIP(c, rwnmr::launch_zwalk_sets(b, threads, c->stream));
                timing_end(c, t0);
                for (uint32_t q = 0; q < gsets; ++q)
                    RW_HIP(c, rwnmr::launch_reduce_rows(part + (size_t)q * (ne + 1) * blocks, (uint64_t)ne + 1,
                                                        blocks, sums + (size_t)(8 * g + q) * rows + done, c->stream));
                done += ne;
            } while (done < echoes);
        }
    }
    if (!w->on_device) {
        RW_HIP(c, hipMemcpyAsync(echo_sums, sums, 8 * rows * n_sets, hipMemcpyDeviceToHost, c->stream));
        if ((s = stage_out(c, w, v)) != RWNMR_OK)
            return s;
        return sync_and_check(c);
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_pfgse_run(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, const int32_t *x0,
                             const int32_t *y0, const int32_t *z0, const uint32_t *steps_inc, uint32_t n_delta,
                             const double *kx, const double *ky, const double *kz, uint32_t n_k, int32_t k_per_delta,
                             double voxel, int32_t bc, uint32_t shift, uint32_t n_sets, const double *set_delta,
                             double *M, double *msd)
{
    return pfgse_common(c, img, w, x0, y0, z0, steps_inc, n_delta, kx, ky, kz, n_k, k_per_delta, voxel, bc, shift,
                        n_sets, set_delta, M, msd, 0, nullptr);
}

rwnmr_status rwnmr_pfgse_run_sampled(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, const int32_t *x0,
                                     const int32_t *y0, const int32_t *z0, const uint32_t *steps_inc,
                                     uint32_t n_delta, const double *kx, const double *ky, const double *kz,
                                     uint32_t n_k, int32_t k_per_delta, double voxel, int32_t bc, uint32_t n_sets,
                                     const double *set_delta, uint32_t n_samples, double *M_samples, double *M,
                                     double *msd)
{
    if (n_samples == 0 || !M_samples)
        return fail(c, RWNMR_ERR_INVALID, "pfgse_run_sampled: need n_samples > 0 and M_samples");
    return pfgse_common(c, img, w, x0, y0, z0, steps_inc, n_delta, kx, ky, kz, n_k, k_per_delta, voxel, bc, 0,
                        n_sets, set_delta, M, msd, n_samples, M_samples);
}

}  // extern "C"

namespace {

rwnmr_status pfgse_common(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w, const int32_t *x0,
                          const int32_t *y0, const int32_t *z0, const uint32_t *steps_inc, uint32_t n_delta,
                          const double *kx, const double *ky, const double *kz, uint32_t n_k, int32_t k_per_delta,
                          double voxel, int32_t bc, uint32_t shift, uint32_t n_sets, const double *set_delta,
                          double *M, double *msd, uint32_t n_samples, double *M_samples)
{
    rwnmr_status s;
    if (!c)
        return RWNMR_ERR_INVALID;
    if ((s = check_image(c, img)) != RWNMR_OK)
        return s;
    if ((s = check_walkers(c, img, w, bc, shift)) != RWNMR_OK)
        return s;
    if (shift != 0)
        return fail(c, RWNMR_ERR_INVALID, "pfgse_run supports shift == 0 only");
    if (n_sets > (uint32_t)rwnmr::kMaxSets)
        return fail(c, RWNMR_ERR_INVALID, "pfgse_run: at most 8 energy sets");
    if (!M || (n_delta && !steps_inc) || (n_k && (!kx || !ky || !kz)) || (n_sets && !set_delta) ||
        (w->n && (!x0 || !y0 || !z0)))
        return fail(c, RWNMR_ERR_INVALID, "pfgse_run: null array");
    if (w->n == 0)
        return fail(c, RWNMR_ERR_INVALID, "pfgse_run: no walkers");
    const uint32_t threads = rwnmr::kPfgseThreads;
    if (n_samples && w->n / n_samples == 0)
        return fail(c, RWNMR_ERR_INVALID, "pfgse sampling: fewer walkers than samples");
    WalkerView v;
    if ((s = clear_flag(c)) != RWNMR_OK)
        return s;
    if ((s = stage_in(c, w, v)) != RWNMR_OK)
        return s;
    WalkArgs a = base_args(img, v, w, shift, c->flag.as<uint32_t>());
    attach_zimage(c, img, bc, 1, a, true);
    PfgseArgs p;
    memset(&p, 0, sizeof(p));
    // the unfolded-image kernel walks kPfgseK walkers per lane
    p.walkers_per_block = a.zwords ? threads * (uint32_t)rwnmr::kPfgseK : threads;
    if (n_samples) {
        // NMR_PFGSE::recoverDsatWithSampling: walkersPerSample = N / samples
        p.group_walkers = w->n / n_samples;
        p.blocks_per_group = (uint32_t)((p.group_walkers + p.walkers_per_block - 1) / p.walkers_per_block);
        p.full_groups = n_samples;
    }
    const uint32_t blocks = rwnmr::pfgse_grid_blocks(p, w->n);
    const uint64_t mrows = (uint64_t)(n_sets ? n_sets : 1) * n_delta * n_k;
    const uint64_t rows = mrows + (msd ? 3ull * n_delta : 0ull);
    if (rows * blocks * 8ull > (8ull << 30))
        return fail(c, RWNMR_ERR_INVALID, "pfgse_run: partial buffer above 8 GiB; split the k or delta set");
    p.n_delta = n_delta;
    p.n_k = n_k;
    p.k_delta_stride = k_per_delta ? n_k : 0u;
    p.voxel = voxel;
    p.n_sets = n_sets;
    if (w->on_device) {
        p.x0 = x0;
        p.y0 = y0;
        p.z0 = z0;
        p.steps_inc = steps_inc;
        p.kx = kx;
        p.ky = ky;
        p.kz = kz;
        p.set_delta = set_delta;
    } else {
        const uint64_t n = w->n;
        RW_HIP(c, c->x0.reserve(n * 4 + 4));
        RW_HIP(c, c->y0.reserve(n * 4 + 4));
        RW_HIP(c, c->z0.reserve(n * 4 + 4));
        RW_HIP(c, hipMemcpyAsync(c->x0.p, x0, n * 4, hipMemcpyHostToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(c->y0.p, y0, n * 4, hipMemcpyHostToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(c->z0.p, z0, n * 4, hipMemcpyHostToDevice, c->stream));
        p.x0 = c->x0.as<int32_t>();
        p.y0 = c->y0.as<int32_t>();
        p.z0 = c->z0.as<int32_t>();
        RW_HIP(c, c->steps_inc.reserve(4ull * n_delta + 4));
        RW_HIP(c, hipMemcpyAsync(c->steps_inc.p, steps_inc, 4ull * n_delta, hipMemcpyHostToDevice, c->stream));
        p.steps_inc = c->steps_inc.as<uint32_t>();
        const uint64_t nkv = (uint64_t)n_k * (k_per_delta ? n_delta : 1u);
        RW_HIP(c, c->kvec.reserve(24ull * nkv + 24));
        double *kv = c->kvec.as<double>();
        RW_HIP(c, hipMemcpyAsync(kv, kx, 8ull * nkv, hipMemcpyHostToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(kv + nkv, ky, 8ull * nkv, hipMemcpyHostToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(kv + 2ull * nkv, kz, 8ull * nkv, hipMemcpyHostToDevice, c->stream));
        p.kx = kv;
        p.ky = kv + nkv;
        p.kz = kv + 2ull * nkv;
        if (n_sets) {
            RW_HIP(c, c->set_delta.reserve(8ull * n_sets));
            RW_HIP(c, hipMemcpyAsync(c->set_delta.p, set_delta, 8ull * n_sets, hipMemcpyHostToDevice, c->stream));
            p.set_delta = c->set_delta.as<double>();
        }
    }
    RW_HIP(c, c->partials.reserve(8ull * rows * blocks + 8));
    p.partials = c->partials.as<double>();
    p.partial_stride = blocks;
    p.msd_rows = msd ? mrows : 0;
    {
        // k rows that are arithmetic progressions (to rounding) take the
        // rotation form of the phase sum
        const uint64_t nkv = (uint64_t)n_k * (k_per_delta ? n_delta : 1u);
        std::vector<double> hk(3 * nkv);
        const hipMemcpyKind kind = w->on_device ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
        RW_HIP(c, hipMemcpyAsync(hk.data(), kx, 8 * nkv, kind, c->stream));
        RW_HIP(c, hipMemcpyAsync(hk.data() + nkv, ky, 8 * nkv, kind, c->stream));
        RW_HIP(c, hipMemcpyAsync(hk.data() + 2 * nkv, kz, 8 * nkv, kind, c->stream));
        RW_HIP(c, hipStreamSynchronize(c->stream));
        p.k_rot = k_rows_arithmetic(hk.data(), nkv, n_k, k_per_delta ? n_delta : 1u) ? 1u : 0u;
    }
    double *Mdev = M, *msd_dev = msd;
    if (!w->on_device) {
        RW_HIP(c, c->sums.reserve(8ull * rows + 8));
        Mdev = c->sums.as<double>();
        msd_dev = Mdev + mrows;
    }
    hipEvent_t t0 = timing_begin(c);
    RW_HIP(c, rwnmr::launch_pfgse(a, p, bc, threads, c->stream));
    timing_end(c, t0);
    RW_HIP(c, rwnmr::launch_reduce_rows(p.partials, mrows, blocks, Mdev, c->stream));
    if (msd)
        RW_HIP(c, rwnmr::launch_reduce_rows(p.partials + mrows * blocks, 3ull * n_delta, blocks, msd_dev, c->stream));
    double *Ms_dev = M_samples;
    if (n_samples) {
        // per-sample sums: the sample's own block columns, fixed order
        if (!w->on_device) {
            RW_HIP(c, c->msamples.reserve(8ull * mrows * n_samples + 8));
            Ms_dev = c->msamples.as<double>();
        }
        std::vector<rwnmr::ReduceRow> rows_list;
        rows_list.reserve((size_t)mrows * n_samples);
        for (uint32_t smp = 0; smp < n_samples; ++smp)
            for (uint64_t r = 0; r < mrows; ++r)
                rows_list.push_back(rwnmr::ReduceRow{p.partials + r * blocks + (uint64_t)smp * p.blocks_per_group,
                                                     p.blocks_per_group, Ms_dev + (uint64_t)smp * mrows + r});
        RW_HIP(c, hipStreamSynchronize(c->stream));   // the row table buffer may still be read by a prior launch
        RW_HIP(c, c->batch_rows.reserve(sizeof(rwnmr::ReduceRow) * rows_list.size()));
        RW_HIP(c, hipMemcpy(c->batch_rows.p, rows_list.data(), sizeof(rwnmr::ReduceRow) * rows_list.size(),
                            hipMemcpyHostToDevice));
        RW_HIP(c, rwnmr::launch_reduce_segments(c->batch_rows.as<rwnmr::ReduceRow>(), rows_list.size(), c->stream));
    }
    if (!w->on_device) {
        if (n_samples)
            RW_HIP(c, hipMemcpyAsync(M_samples, Ms_dev, 8ull * mrows * n_samples, hipMemcpyDeviceToHost, c->stream));
        RW_HIP(c, hipMemcpyAsync(M, Mdev, 8ull * mrows, hipMemcpyDeviceToHost, c->stream));
        if (msd)
            RW_HIP(c, hipMemcpyAsync(msd, msd_dev, 24ull * n_delta, hipMemcpyDeviceToHost, c->stream));
        if ((s = stage_out(c, w, v)) != RWNMR_OK)
            return s;
        return sync_and_check(c);
    }
    return RWNMR_OK;
}

}  // namespace


extern "C" uint64_t rwnmr_ckpt_bytes(uint64_t n, uint32_t n_segments, uint32_t n_labels)
{
    return rwnmr::ckpt_total_bytes(n, n_segments, n_labels);
}

extern "C" rwnmr_status rwnmr_cpmg_forward_ckpt(rwnmr_ctx *c, const rwnmr_image *img, rwnmr_walkers *w,
                                                uint32_t echoes, uint32_t steps_per_echo, int32_t bc,
                                                const double *label_delta, uint32_t every, void *ckpt,
                                                double *echo_sums)
{
    rwnmr_status s;
    if (!c)
        return RWNMR_ERR_INVALID;
    if ((s = check_image(c, img)) != RWNMR_OK)
        return s;
    if ((s = check_walkers(c, img, w, bc, 0)) != RWNMR_OK)
        return s;
    if (!w->on_device || !w->energy || !label_delta || !ckpt || !echo_sums || every == 0 || echoes == 0)
        return fail(c, RWNMR_ERR_INVALID,
                    "forward_ckpt: device walkers with energies, label deltas, checkpoint buffer, echo sums");
    if (!img->labels_blocked)
        return fail(c, RWNMR_ERR_INVALID, "forward_ckpt: image has no solid labels");
    const uint64_t n = w->n;
    const uint32_t nl = (uint32_t)img->n_labels;
    const uint32_t nseg = (echoes + every - 1) / every;
    RW_HIP(c, c->label_hits.reserve(4ull * nl * n + 4));
    uint32_t *lh = c->label_hits.as<uint32_t>();
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)((nl * n + 255) / 256)), dim3(256), 0, c->stream, lh, 0u,
                       (uint64_t)nl * n);
    RW_HIP(c, hipGetLastError());
    for (uint32_t seg = 0; seg < nseg; ++seg) {
        const rwnmr::CkptView v = rwnmr::ckpt_view(ckpt, n, nseg, nl, seg);
        RW_HIP(c, hipMemcpyAsync(v.x, w->x, 4 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(v.y, w->y, 4 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(v.z, w->z, 4 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(v.seed, w->seed, 8 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(v.energy, w->energy, 8 * n, hipMemcpyDeviceToDevice, c->stream));
        RW_HIP(c, hipMemcpyAsync(v.hits, lh, 4ull * nl * n, hipMemcpyDeviceToDevice, c->stream));
        const uint32_t e0 = seg * every;
        const uint32_t ne = std::min(every, echoes - e0);
        s = cpmg_common(c, img, w, ne, steps_per_echo, bc, 0, 2, label_delta, lh, echo_sums + e0);
        if (s != RWNMR_OK)
            return s;
    }
    // the per-label hits at the end (the replay of the last segment reads them)
    RW_HIP(c, hipMemcpyAsync(rwnmr::ckpt_final_hits(ckpt, n, nseg, nl), lh, 4ull * nl * n, hipMemcpyDeviceToDevice,
                             c->stream));
    return RWNMR_OK;
}

extern "C" rwnmr_status rwnmr_cpmg_backward(rwnmr_ctx *c, const rwnmr_image *img, const void *ckpt, uint64_t n,
                                            uint32_t echoes, uint32_t steps_per_echo, int32_t bc,
                                            const double *label_delta, uint32_t every, const double *grad_echo,
                                            double *grad_delta)
{
    rwnmr_status s;
    if (!c)
        return RWNMR_ERR_INVALID;
    if ((s = check_image(c, img)) != RWNMR_OK)
        return s;
    if (!img->labels_blocked || !ckpt || !label_delta || !grad_echo || !grad_delta || every == 0 || echoes == 0 ||
        n == 0)
        return fail(c, RWNMR_ERR_INVALID, "backward: null argument or empty run");
    if (bc < RWNMR_BC_NOFLUX || bc > RWNMR_BC_MIRROR)
        return fail(c, RWNMR_ERR_INVALID, "backward: unknown boundary condition");
    const uint32_t nl = (uint32_t)img->n_labels;
    const uint32_t nseg = (echoes + every - 1) / every;
    WalkArgs a;
    memset(&a, 0, sizeof(a));
    a.blocks = img->blocks;
    a.labels_blocked = img->labels_blocked;
    a.W = img->W;
    a.H = img->H;
    a.D = img->D;
    a.bC = img->bC;
    a.bCR = img->bC * img->bR;
    a.layout = img->layout;
    a.label_delta = label_delta;
    a.n_labels = (int32_t)nl;
    a.n = n;
    attach_zimage(c, img, bc, 2, a);
    const uint32_t threads = a.zwords ? rwnmr::kZBackThreads : 256u;
    const uint64_t bx = (n + threads - 1) / threads;
    if (rwnmr::backward_grid(bx, nseg) > 0x7FFFFFFFull)
        return fail(c, RWNMR_ERR_INVALID, "backward: grid too large");
    rwnmr::BackwardArgs b;
    b.ckpt = const_cast<void *>(ckpt);
    b.n = n;
    b.n_segments = nseg;
    b.n_labels = (int32_t)nl;
    b.every = every;
    b.echoes = echoes;
    b.steps_per_echo = steps_per_echo;
    b.grad_echo = grad_echo;
    b.partial_stride = (uint64_t)nseg * bx;
    b.blocks = (uint32_t)bx;
    RW_HIP(c, c->partials.reserve(8ull * nl * b.partial_stride + 8));
    b.partials = c->partials.as<double>();
    hipEvent_t t0 = timing_begin(c);
    RW_HIP(c, rwnmr::launch_backward(a, b, bc, threads, c->stream));
    timing_end(c, t0);
    RW_HIP(c, c->sums.reserve(8ull * nl + 8));
    RW_HIP(c, rwnmr::launch_reduce_rows(b.partials, nl, (uint32_t)b.partial_stride, c->sums.as<double>(), c->stream));
    double raw[rwnmr::kMaxLabels], dl[rwnmr::kMaxLabels];
    RW_HIP(c, hipMemcpyAsync(raw, c->sums.p, 8ull * nl, hipMemcpyDeviceToHost, c->stream));
    RW_HIP(c, hipMemcpyAsync(dl, label_delta, 8ull * nl, hipMemcpyDeviceToHost, c->stream));
    if ((s = sync_and_check(c)) != RWNMR_OK)
        return s;
    for (uint32_t l = 0; l < nl; ++l)
        grad_delta[l] = raw[l] / dl[l];
    return RWNMR_OK;
}

extern "C" rwnmr_status rwnmr_cpmg_batch_run(rwnmr_ctx *c, const rwnmr_cpmg_instance *inst, uint32_t n_inst,
                                             uint32_t flags)
{
    if (!c)
        return RWNMR_ERR_INVALID;
    if (!inst || n_inst == 0)
        return fail(c, RWNMR_ERR_INVALID, "batch_run: no instances");
    if (flags > 1)
        return fail(c, RWNMR_ERR_INVALID, "batch_run: unknown flags");
    for (uint32_t i = 0; i < n_inst; ++i) {
        const rwnmr_cpmg_instance &q = inst[i];
        rwnmr_status s = check_image(c, q.image);
        if (s != RWNMR_OK)
            return s;
        if ((s = check_walkers(c, q.image, &q.walkers, q.bc, 0)) != RWNMR_OK)
            return s;
        if (!q.walkers.on_device || !q.echo_sums)
            return fail(c, RWNMR_ERR_INVALID, "batch_run: walkers and echo_sums must be device arrays");
        if (q.walkers.n == 0)
            return fail(c, RWNMR_ERR_INVALID, "batch_run: instance without walkers");
    }
    if (flags & 1u) {
        // one launch per instance (the comparison path)
        for (uint32_t i = 0; i < n_inst; ++i) {
            rwnmr_walkers w = inst[i].walkers;
            rwnmr_status s = cpmg_common(c, inst[i].image, &w, inst[i].echoes, inst[i].steps_per_echo, inst[i].bc, 0,
                                         1, nullptr, nullptr, inst[i].echo_sums);
            if (s != RWNMR_OK)
                return s;
        }
        return RWNMR_OK;
    }
    const uint32_t K = c->shape.k;
    const uint64_t tile_walkers = (uint64_t)rwnmr::kBatchThreads * K;
    std::vector<rwnmr::BatchInstance> desc(n_inst);
    std::vector<rwnmr::BatchTile> tiles;
    std::vector<std::pair<double, uint32_t>> order;
    uint64_t part_total = 0, rows_total = 0;
    for (uint32_t i = 0; i < n_inst; ++i) {
        const rwnmr_cpmg_instance &q = inst[i];
        const uint64_t nt = (q.walkers.n + tile_walkers - 1) / tile_walkers;
        if (nt > 0xFFFFFFFFull)
            return fail(c, RWNMR_ERR_INVALID, "batch_run: too many tiles");
        part_total += (uint64_t)(q.echoes + 1) * nt;
        rows_total += q.echoes + 1ull;
    }
    RW_HIP(c, hipStreamSynchronize(c->stream));  // descriptor buffers may still be in use
    if (clear_flag(c) != RWNMR_OK)
        return RWNMR_ERR_HIP;
    RW_HIP(c, c->partials.reserve(8ull * part_total + 8));
    double *part = c->partials.as<double>();
    std::vector<rwnmr::ReduceRow> rows;
    rows.reserve(rows_total);
    uint64_t off = 0;
    for (uint32_t i = 0; i < n_inst; ++i) {
        const rwnmr_cpmg_instance &q = inst[i];
        const rwnmr_image *img = q.image;
        const uint32_t nt = (uint32_t)((q.walkers.n + tile_walkers - 1) / tile_walkers);
        rwnmr::BatchInstance &d = desc[i];
        d.blocks = img->blocks;
        d.W = img->W;
        d.H = img->H;
        d.D = img->D;
        d.bC = img->bC;
        d.bCR = img->bC * img->bR;
        d.layout = img->layout;
        d.bc = q.bc;
        d.x = q.walkers.x;
        d.y = q.walkers.y;
        d.z = q.walkers.z;
        d.seed = q.walkers.seed;
        d.energy = q.walkers.energy;
        d.hits = q.walkers.hits;
        d.penalty = q.walkers.penalty;
        d.delta_uniform = q.walkers.delta_uniform;
        d.n = q.walkers.n;
        d.echoes = q.echoes;
        d.steps_per_echo = q.steps_per_echo;
        d.partials = part + off;
        d.n_tiles = nt;
        rwnmr::WalkArgs za;
        memset(&za, 0, sizeof(za));
        attach_zimage(c, img, q.bc, 1, za);
        d.zwords = za.zwords;
        d.zl = za.zl;
        for (uint32_t t = 0; t < nt; ++t) {
            const uint64_t walkers = std::min<uint64_t>(tile_walkers, q.walkers.n - (uint64_t)t * tile_walkers);
            const double cost = (double)q.echoes * q.steps_per_echo * (double)walkers;
            order.emplace_back(-cost, (uint32_t)tiles.size());
            tiles.push_back(rwnmr::BatchTile{i, t});
        }
        for (uint32_t e = 0; e <= q.echoes; ++e)
            rows.push_back(rwnmr::ReduceRow{part + off + (uint64_t)e * nt, nt, q.echo_sums + e});
        off += (uint64_t)(q.echoes + 1) * nt;
    }
    // longest tiles first (LPT order) so the queue's tail is short work
    std::stable_sort(order.begin(), order.end());
    std::vector<rwnmr::BatchTile> sorted(tiles.size());
    for (size_t i = 0; i < order.size(); ++i)
        sorted[i] = tiles[order[i].second];
    RW_HIP(c, c->batch_inst.reserve(sizeof(rwnmr::BatchInstance) * desc.size()));
    RW_HIP(c, c->batch_tiles.reserve(sizeof(rwnmr::BatchTile) * sorted.size()));
    RW_HIP(c, c->batch_rows.reserve(sizeof(rwnmr::ReduceRow) * rows.size()));
    RW_HIP(c, c->batch_counter.reserve(sizeof(uint32_t)));
    RW_HIP(c, hipMemcpy(c->batch_inst.p, desc.data(), sizeof(rwnmr::BatchInstance) * desc.size(),
                        hipMemcpyHostToDevice));
    RW_HIP(c, hipMemcpy(c->batch_tiles.p, sorted.data(), sizeof(rwnmr::BatchTile) * sorted.size(),
                        hipMemcpyHostToDevice));
    RW_HIP(c, hipMemcpy(c->batch_rows.p, rows.data(), sizeof(rwnmr::ReduceRow) * rows.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, c->stream, c->batch_counter.as<uint32_t>(), 0u, 1ull);
    RW_HIP(c, hipGetLastError());
    int per_cu = 0, cus = 0;
    RW_HIP(c, rwnmr::batch_resident_blocks(K, &per_cu));
    RW_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const uint64_t resident = (uint64_t)std::max(per_cu, 1) * (uint64_t)std::max(cus, 1);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(resident, sorted.size());
    rwnmr::BatchArgs b;
    b.inst = c->batch_inst.as<rwnmr::BatchInstance>();
    b.tiles = c->batch_tiles.as<rwnmr::BatchTile>();
    b.n_tiles = (uint32_t)sorted.size();
    b.counter = c->batch_counter.as<uint32_t>();
    b.error_flag = c->flag.as<uint32_t>();
    hipEvent_t t0 = timing_begin(c);
    RW_HIP(c, rwnmr::launch_batch(b, K, grid, c->stream));
    timing_end(c, t0);
    RW_HIP(c, rwnmr::launch_reduce_segments(c->batch_rows.as<rwnmr::ReduceRow>(), rows.size(), c->stream));
    return RWNMR_OK;
}
