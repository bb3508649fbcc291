// walk_kernels.h -- host-side view of the gfx950 walk kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_common.h"

namespace rwnmr {

static constexpr uint32_t kXcds = 8;     // MI355X: 8 accelerator dies, each with its own L2

static constexpr int kMaxLabels = 8;   // solid labels per image (per-label relaxivity)
// per-label hit counts of a labelled walk in LDS: one row of kMaxLabels
// counts per walker, padded to an odd stride so a wave's lanes hit distinct banks
static constexpr int kLhStride = kMaxLabels + 1;

// Everything one walk launch reads or writes.  Walker arrays are device
// pointers of length n (structure of arrays).
struct WalkArgs {
    const uint64_t *blocks;          // bitBlock words in the image's device layout
    const uint8_t *labels_blocked;   // optional: 64 label bytes per block (bit order)
    int32_t W, H, D;                 // image columns, rows, depth
    int32_t bC, bCR;                 // block columns, block columns * block rows
    BrickLayout layout;              // device word order of the blocks
    uint32_t tile_map;               // 0: block b walks tile b; 1: tiles striped per XCD
    uint32_t n_tiles;                // tiles of the launch (tile_map == 1)
    uint32_t tile_walkers;           // 0: a tile is blockDim x K walkers; else the walkers of a tile, at most
                                     // blockDim (balanced tiles, zbalanced_tile_walkers)
    int32_t shift;                   // log2(voxel division)
    int32_t *x, *y, *z;
    uint64_t *seed;
    double *energy;                  // nullable: start at 1.0, not stored
    uint32_t *hits;                  // nullable
    const double *penalty;           // nullable: delta_uniform for every walker
    double delta_uniform;
    const double *label_delta;       // [n_labels] device
    int32_t n_labels;
    uint32_t *label_hits;            // nullable: [n_labels][n]
    uint64_t n;
    uint32_t echoes;
    uint32_t steps_per_echo;
    double *partials;                // nullable: [(echoes + 1)][partial_stride]
    uint32_t partial_stride;         // = grid blocks of the launch
    uint32_t *error_flag;            // device word, bit 0: walker started outside the image
    const uint32_t *zwords;          // unfolded image (periodic / mirror fast path), else null
    ZLayout zl;                      // its dilated address layout
    uint32_t zcache;                 // chunk cache in registers: 0 none, 1 16-byte (word select), 2/3 16-byte, pipelined,
                                     // 4 speculative walk, 5 8-byte pipelined (mirror)
    uint32_t znoflux;                // 1: noflux walk on the periodic unfolded image (zl of that image)
    uint32_t zmflags;                // 1: mirror walk started in the first octant with reflection flags
    const uint8_t *zlabels;          // solid labels, one byte per voxel in the unfolded image's address order
                                     // (per-label walks; with zlfold the first octant's addresses only)
    uint32_t zlfold;                 // 1: labels read at the first-octant address (zoctant)
    uint32_t zlf_shift, zlf_mul, zlf_shift2, zlf_mask;   // zoctant constants (zoctant_constants)
    uint32_t n_label_sets;           // label-set walks: label_delta holds [n_label_sets][n_labels]
    double *set_energy;              // label-set walks: nullable [8][n] energies carried between launches
    double *jac_partials;            // Jacobian forward: [kMaxLabels][(echoes + 1)][partial_stride] block sums of
                                     // h_l E (hits on label l up to the echo times the energy)
    // internal-field CPMG (emode 3): field[x + y W + z W H], phase step gammatau,
    // per-walker surviving fraction and phase (in/out)
    const double *field;
    double gammatau;
    double *p_alive, *phase;
};

struct LaunchShape {
    uint32_t threads;   // per block, multiple of 64, <= 1024
    uint32_t k;         // walkers per lane: 1, 2 or 4
};

// Largest block for K walkers per lane: the register budget per lane shrinks
// as the block grows (512 VGPRs per SIMD lane / waves per SIMD), so wider K
// runs in smaller blocks to keep every walker in registers.
__host__ __device__ constexpr uint32_t max_threads_for_k(int k) { return k >= 4 ? 512u : 1024u; }

static constexpr uint32_t kPfgseThreads = 256;
static constexpr int kPfgseWaves = (int)kPfgseThreads / 64;
static constexpr int kPfgseK = 2;   // walkers per lane of the unfolded-image PFGSE kernel
static constexpr int kMaxSets = 8;       // energy sets evaluated on one PFGSE trajectory
static constexpr int kPfgseKChunk = 64;  // k values staged in LDS per block sync

// PFGSE-specific arguments (device pointers).  The shift > 0 voxel division
// is not supported by the PFGSE kernel.
struct PfgseArgs {
    const int32_t *x0, *y0, *z0;     // initial positions (displacement origin)
    const uint32_t *steps_inc;       // [n_delta] steps between consecutive diffusion times
    uint32_t n_delta;
    const double *kx, *ky, *kz;      // wave vectors: [n_k], or [n_delta][n_k] per diffusion time
    uint32_t n_k;
    uint32_t k_delta_stride;         // 0: one k set for every diffusion time; n_k: a set per time
    double voxel;                    // voxel size (phase = k . dr * voxel)
    uint32_t n_sets;                 // 0: per-walker delta; else uniform deltas below
    const double *set_delta;         // [n_sets]
    double *partials;                // [(n_sets|1) * n_delta * n_k (+ 3 n_delta)][partial_stride]
    uint32_t partial_stride;         // = grid blocks
    uint64_t msd_rows;               // 0: no displacement sums; else first row of the
                                     // [n_delta][3] squared-displacement sums
    uint32_t k_rot;                  // 1: every k row is an arithmetic progression (n_k >= 2):
                                     // cos(k_j . dr) by complex rotation instead of cos()
    // walker samples: full_groups groups of group_walkers consecutive walkers,
    // each walked by blocks_per_group blocks of its own (so per-sample sums are
    // sums of whole partial columns); walkers after them follow in plain
    // blocks.  group_walkers == 0: plain blocks for every walker.
    uint64_t group_walkers;
    uint32_t blocks_per_group, full_groups;
    uint32_t walkers_per_block;      // kPfgseThreads x walkers per lane of the kernel that runs
};

// Walker index of thread threadIdx.x of block b under the sample grouping of
// p (see PfgseArgs); `inside` is false for padding threads.
__host__ __device__ inline uint64_t pfgse_walker(const PfgseArgs &p, uint64_t b, uint32_t t, uint32_t threads,
                                                 uint64_t n, bool &inside)
{
    if (p.group_walkers == 0) {
        const uint64_t id = b * threads + t;
        inside = id < n;
        return id;
    }
    const uint64_t grouped = (uint64_t)p.blocks_per_group * p.full_groups;
    if (b < grouped) {
        const uint64_t g = b / p.blocks_per_group, lb = b % p.blocks_per_group;
        const uint64_t local = lb * threads + t;
        inside = local < p.group_walkers;
        return g * p.group_walkers + local;
    }
    const uint64_t id = p.group_walkers * p.full_groups + (b - grouped) * threads + t;
    inside = id < n;
    return id;
}

uint32_t walk_grid_blocks(uint64_t n, const LaunchShape &ls, int shift);

// Blocks of a PFGSE launch of n walkers under the grouping of p
// (p.walkers_per_block walkers per block).
inline uint32_t pfgse_grid_blocks(const PfgseArgs &p, uint64_t n)
{
    const uint64_t per = p.walkers_per_block;
    if (p.group_walkers == 0)
        return (uint32_t)((n + per - 1) / per);
    const uint64_t rest = n - p.group_walkers * p.full_groups;
    return (uint32_t)((uint64_t)p.blocks_per_group * p.full_groups + (rest + per - 1) / per);
}
hipError_t launch_pfgse(const WalkArgs &a, const PfgseArgs &p, int bc, uint32_t threads, hipStream_t st);
hipError_t launch_walk(const WalkArgs &a, int bc, int emode, const LaunchShape &ls, hipStream_t st);
hipError_t launch_reduce_rows(const double *partials, uint64_t rows, uint32_t cols, double *out,
                              hipStream_t st);

}  // namespace rwnmr
