/*
 * rwnmr.h -- C ABI of the MI355X random-walk NMR hot path.
 *
 * Plain C types only (no HIP or torch types): an FFI (ctypes, cgo, JNI, ...)
 * binds these symbols directly.  Every call returns an rwnmr_status (0 = OK)
 * and never exits the process; rwnmr_last_error() explains a failure.
 *
 * Ownership: the caller owns every input and output array; a context owns its
 * device memory, HIP stream and launch-shape cache.  A context is not
 * thread-safe (one host thread drives it), like the reference's single
 * mutex-guarded inverter (B/Laplace/src/nmrinv.cpp:4-5).
 *
 * Host vs device pointers: rwnmr_walkers.on_device = 0 means every array of
 * the call is host memory (the call copies in, runs, copies out and returns
 * when done); 1 means every array is device memory already resident in HBM
 * (the call only enqueues work on the context stream; rwnmr_ctx_synchronize
 * waits and reports kernel-side errors).
 *
 * Reference paths below are under src/backend/ of RandomWalk_NMR.
 */
#ifndef RWNMR_H
#define RWNMR_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rwnmr_status;
#define RWNMR_OK 0
#define RWNMR_ERR_INVALID 1     /* bad argument (sizes, null pointers, ranges) */
#define RWNMR_ERR_HIP 2         /* HIP runtime error */
#define RWNMR_ERR_ALLOC 3       /* device or host allocation failed */
#define RWNMR_ERR_STATE 4       /* call out of order (e.g. lambda not set) */
#define RWNMR_ERR_WALKER 5      /* a walker started outside the image (noflux) */
#define RWNMR_ERR_NUMERIC 6     /* solver failed to converge */

/* Boundary conditions: rwnmr_config BC key (B/ConfigFiles/rwnmr_config.cpp:235-246). */
#define RWNMR_BC_NOFLUX 0
#define RWNMR_BC_PERIODIC 1
#define RWNMR_BC_MIRROR 2

typedef struct rwnmr_ctx rwnmr_ctx;
typedef struct rwnmr_image rwnmr_image;

/* Library version string. */
const char *rwnmr_version(void);

/* Context on HIP device `device`: stream, scratch, launch-shape cache.
 * Replaces the per-call cudaMalloc/cudaDeviceReset life cycle of
 * B/NMR_Simulation/NMR_cpmg_cuda.cpp:929-1774 (no device reset, ever). */
rwnmr_status rwnmr_ctx_create(rwnmr_ctx **out, int32_t device);
void rwnmr_ctx_destroy(rwnmr_ctx *ctx);
const char *rwnmr_last_error(const rwnmr_ctx *ctx);
/* Wait for the context stream; returns RWNMR_ERR_WALKER if a kernel saw a
 * walker start outside the image since the last check. */
rwnmr_status rwnmr_ctx_synchronize(rwnmr_ctx *ctx);
/* The context's hipStream_t, as an opaque pointer (for event timing). */
void *rwnmr_ctx_stream(rwnmr_ctx *ctx);
/* Launch shape for walk kernels: threads per block (64..1024, multiple of
 * 64) and walkers per lane (1, 2 or 4).  0 selects the built-in default. */
rwnmr_status rwnmr_ctx_set_launch(rwnmr_ctx *ctx, uint32_t threads_per_block, uint32_t walkers_per_lane);
rwnmr_status rwnmr_ctx_get_launch(const rwnmr_ctx *ctx, uint32_t *threads_per_block, uint32_t *walkers_per_lane);
/* HIP-event timing of every walk / PFGSE kernel launched on the context
 * stream (replaces the cudaEvent + omp_get_wtime buckets of
 * NMR_cpmg_cuda.cpp:934-989, 1756-1773).  get_timing synchronises and
 * returns the accumulated kernel milliseconds and launch count. */
rwnmr_status rwnmr_ctx_set_timing(rwnmr_ctx *ctx, int32_t enable);
rwnmr_status rwnmr_ctx_get_timing(rwnmr_ctx *ctx, double *kernel_ms, uint64_t *launches, int32_t reset);
/* Device layout of images uploaded after this call: the bitBlock words are
 * regrouped in bricks of 2^sx x 2^sy x 2^sz blocks so a cache line holds a
 * compact 3-D neighbourhood (0,0,0 keeps the reference order,
 * B/BitBlock/bitBlock.h:113-122).  Default 1,1,1.  Results do not depend on it. */
rwnmr_status rwnmr_ctx_set_image_layout(rwnmr_ctx *ctx, int32_t sx, int32_t sy, int32_t sz);
/* Locality ordering for CPMG and map runs of at least min_walkers walkers:
 * walkers are sorted by the Morton code of their image block every
 * resort_every echoes and tiles are striped so each XCD walks one compact
 * region of the image (its L2 holds that region).  Walker results are
 * identical with it on or off; echo sums differ only in summation order.
 * Default on, every 1000 echoes, from 65,536 walkers. */
rwnmr_status rwnmr_ctx_set_locality(rwnmr_ctx *ctx, int32_t enable, uint32_t resort_every, uint64_t min_walkers);
/* Periodic and mirror CPMG / map walks (shift 0, uniform or per-walker
 * delta) on an unfolded copy of the image: the mirror tiling of period 2d
 * (or the periodic one of period d) stored once with Z-order voxel
 * addresses, so a step needs no fold (replaces the per-step fold of
 * NMR_cpmg_cuda.cpp:1991-2015).  Built per image on first use: 2W x 2H x 2D
 * bits for mirror (128 MiB at 512^3), W x H x D for periodic; volumes over
 * 2^32 voxels use the folded kernels.  mode 0: folded kernels; 1: unfolded;
 * 2: unfolded, each walker also keeping the 16-byte chunk (8 x 4 x 4
 * voxels) of its last lookup in registers so most steps read no memory;
 * 3 (default) and 4: the chunk tested by one shift and the next direction
 * drawn a step ahead, through the 4 KiB byte-key table (3) or the 64 KiB
 * wide direction table (4).  Modes 3 and 4 also run noflux walks (image
 * sides >= 2) on the periodic unfolded image, and start mirror CPMG / map
 * walkers at their folded positions with reflection flags (mirrored
 * direction tables, byte-key form).  5: as 3 for every walk except CPMG with
 * one walker per lane, which loads the chunks of both possible next
 * candidates a step ahead (two loads per walker-step) and draws directions
 * two steps ahead: for latency-bound runs of few walkers.  6: as 3, with
 * mirror walks keeping an 8-byte chunk (4 x 4 x 4 voxels: 3 VALU per step
 * fewer, a quarter more chunk loads); other walks run as 3.  Results are
 * identical in every mode. */
rwnmr_status rwnmr_ctx_set_unfolded(rwnmr_ctx *ctx, int32_t mode);

/* Upload a packed image in the reference bitBlock layout: block index
 * (x>>2) + (y>>2)*bC + (z>>2)*bC*bR, bit (x&3) + 4*(y&3) + 16*(z&3),
 * 1 = solid, bC = ceil(cols/4), bR = ceil(rows/4), bD = ceil(depth/4)
 * (B/BitBlock/bitBlock.h:113-139, B/BitBlock/bitBlock.cpp:138-221; the
 * H2D copy of NMR_cpmg_cuda.cpp:1044-1048).  `blocks` is host memory of
 * bC*bR*bD words.  solid_labels (nullable) is a cols*rows*depth byte volume
 * (x fastest) of solid-voxel labels in [0, n_labels), n_labels <= 8, used by
 * the per-label relaxivity path (no reference counterpart: the reference has
 * one relaxivity per walker, B/Walker/walker.h:198-209). */
rwnmr_status rwnmr_image_upload(rwnmr_ctx *ctx, const uint64_t *blocks, int32_t cols, int32_t rows,
                                int32_t depth, const uint8_t *solid_labels, int32_t n_labels,
                                rwnmr_image **out);
void rwnmr_image_destroy(rwnmr_image *img);
/* Pack a host voxel volume (x fastest, nonzero = solid) into bitBlock words
 * (bC*bR*bD entries), as B/BitBlock/bitBlock.cpp:168-221 does. */
rwnmr_status rwnmr_bitblock_pack(const uint8_t *solid, int32_t cols, int32_t rows, int32_t depth,
                                 uint64_t *blocks_out);

/* Walker set, structure of arrays (the SoA staging of
 * B/NMR_Simulation/NMR_cpmg_cuda.cpp:1110-1170).  x/y/z are current
 * positions (unfolded for periodic/mirror), seed the current xorshift64
 * state.  energy (nullable): magnetisation, in/out (NULL: starts at 1.0).
 * hits (nullable): wall-hit counter, incremented in place.  penalty
 * (nullable): per-walker delta = 1 - (2/3) rho eps / (D0 1e3)
 * (B/Walker/walker.h:203-209); NULL with no label path uses delta_uniform. */
typedef struct {
    int32_t *x, *y, *z;
    uint64_t *seed;
    double *energy;
    uint32_t *hits;
    const double *penalty;
    double delta_uniform;
    uint64_t n;
    int32_t on_device;
} rwnmr_walkers;

/* Autotuner: times trial_echoes-long CPMG runs of copies of the walkers and
 * keeps the fastest settings in the context, in three steps: (1) the launch
 * shape, threads per block 64..1024 x walkers per lane 1/2/4 (walk kernel
 * time); (2) the image form the walk reads, the unfolded Z-order image with
 * the per-walker register chunk cache against the folded brick-ordered
 * bitBlock words (rwnmr_ctx_set_unfolded 3 / 0; walk kernel time), for
 * mirror walks the 8-byte chunk against the 16-byte one (6 / 3, each timed
 * twice in mirrored order), and for at most 65,536 walkers with one walker
 * per lane the speculative walk (5); (3) for
 * runs that use the locality ordering, the echoes per launch between walker
 * re-sorts, trial_echoes / 4, / 2 or trial_echoes (whole-run time, re-sorts
 * included).  best_* report step 1's shape and the fastest walk time; the
 * rest is read back with rwnmr_ctx_get_tuning.  The reference fixes all of
 * this in its config (BLOCKS/THREADSPERBLOCK/ECHOESPERKERNEL,
 * config/default/rwnmr.config:43-47). */
rwnmr_status rwnmr_ctx_autotune(rwnmr_ctx *ctx, const rwnmr_image *img, const rwnmr_walkers *w,
                                uint32_t steps_per_echo, int32_t bc, uint32_t trial_echoes,
                                uint32_t *best_threads, uint32_t *best_k, double *best_ms);
/* The context's current launch shape, image form (rwnmr_ctx_set_unfolded
 * mode) and echoes between locality re-sorts. */
rwnmr_status rwnmr_ctx_get_tuning(const rwnmr_ctx *ctx, uint32_t *threads_per_block, uint32_t *walkers_per_lane,
                                  int32_t *unfolded, uint32_t *resort_every);

/* Map simulation: advance every walker `steps` steps counting wall hits
 * (map_3D_{noflux,periodic,mirror}, B/NMR_Simulation/NMR_Simulation_cuda.cpp:418-721,
 * driven by mapSimulation_CUDA_3D_histograms :726-1270).  shift = log2 of
 * the voxel division (0 for none). */
rwnmr_status rwnmr_map_run(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, uint64_t steps,
                           int32_t bc, uint32_t shift);

/* CPMG: `echoes` x `steps_per_echo` steps; every wall hit multiplies the
 * walker's energy by its delta (CPMG_walk_*, NMR_cpmg_cuda.cpp:37-139,
 * 300-402, 574-676).  echo_sums (echoes+1 doubles): [0] = sum of initial
 * energies, [e] = sum after echo e (image_simulation_cuda,
 * NMR_cpmg_cuda.cpp:929-1774, before bulk and normalisation).  Sums are
 * FP64 with a fixed reduction order (deterministic). */
rwnmr_status rwnmr_cpmg_run(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                            uint32_t steps_per_echo, int32_t bc, uint32_t shift, double *echo_sums);

/* A CPMG run captured once as a HIP graph and replayed with one graph launch
 * (no reference counterpart: the reference launches and synchronises every
 * 16 echoes, B/NMR_Simulation/NMR_cpmg_cuda.cpp:1174-1334).  The graph holds,
 * in order: when `start` is given, copies of its device arrays into w's
 * (x, y, z, seed; energy from start->energy or 1.0 when w->energy is set;
 * hits from start->hits or 0 when w->hits is set); the error-word clear; the
 * walk launches (no locality re-sort) and the echo reductions into the device
 * echo_sums (echoes+1, nullable).  w and start must be device memory
 * (on_device = 1).  Every replay (rwnmr_graph_launch, asynchronous on the
 * context stream; rwnmr_ctx_synchronize waits and reports a noflux walker
 * started outside the image) repeats that work on the same arrays: with
 * `start` each replay is one full run from the start state, without it each
 * replay continues the walk.  Results equal rwnmr_cpmg_run's (walkers bit for
 * bit, echo sums to rounding).  The graph owns its echo partials and keeps
 * the pointers it was built with: destroy it before the image, the arrays or
 * the context. */
typedef struct rwnmr_graph rwnmr_graph;
rwnmr_status rwnmr_cpmg_graph_create(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w,
                                     const rwnmr_walkers *start, uint32_t echoes, uint32_t steps_per_echo, int32_t bc,
                                     double *echo_sums, rwnmr_graph **out);
rwnmr_status rwnmr_graph_launch(rwnmr_graph *graph);
void rwnmr_graph_destroy(rwnmr_graph *graph);

/* CPMG that allocates nothing: rwnmr_cpmg_run for device-resident walkers
 * (on_device must be 1) and a device echo_sums (echoes+1, nullable) that
 * calls no hipMalloc / hipFree and no allocating library routine.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  Its echo partials and error word live
 * in a fixed 16 MiB device array of the library, one per GPU, held until the
 * process ends: a run needs at most 16 MiB / 16 B = 1,048,560 blocks of
 * walkers (threads per block x walkers per lane each) and is split into
 * launches of as many echoes as the array holds, in which case w->energy
 * must be non-null.  It never builds an unfolded image (it uses one an
 * earlier call built, else the folded kernels) and never re-sorts walkers;
 * walker results equal rwnmr_cpmg_run's bit for bit, echo sums to rounding.
 * Asynchronous; rwnmr_ctx_synchronize reports a noflux walker that started
 * outside the image.  All calls on one GPU share the array: calls that
 * overlap on one GPU (two contexts, two threads) race. */
rwnmr_status rwnmr_cpmg_run_noalloc(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                    uint32_t steps_per_echo, int32_t bc, uint32_t shift, double *echo_sums);

/* Internal-field CPMG (NMR_cpmg::image_simulation_cuda with an internal
 * field, kernels CPMG_walk_{noflux,periodic,mirror}_field,
 * B/NMR_Simulation/NMR_cpmg_cuda.cpp:146-293, 409-567, 683-841, driver
 * :1006-1011, 1080-1257): per echo two phase inversions of
 * steps_per_echo / 2 steps; the phase gains gammatau * field[x + y W + z W H]
 * at the walker's image voxel before and after every step, with
 * gammatau = 0.5 * (1e6 * gyromagnetic_ratio) * (1e-3 * time_interval_ms);
 * a wall hit multiplies p_alive by the walker's delta.  echo_sums[e] =
 * sum_w p_alive cos(phase) after echo e (echoes + 1 values).  field holds
 * W*H*D doubles (host or device, as the walkers).  p_alive / phase are
 * in/out per walker (NULL: start at 1 / 0 and are not returned; on device
 * give both or neither).  shift (voxel division) is not supported here. */
rwnmr_status rwnmr_cpmg_field_run(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                  uint32_t steps_per_echo, int32_t bc, const double *field,
                                  double gyromagnetic_ratio, double time_interval_ms, double *p_alive, double *phase,
                                  double *echo_sums);

/* CPMG with a relaxivity per solid label: a hit on a voxel of label l
 * multiplies by label_delta[l] (host or device like the walkers, n_labels of
 * the image).  label_hits (nullable, [n_labels][n] uint32, in/out) counts hits
 * per label.  No reference counterpart (new data for the relaxivity fit). */
rwnmr_status rwnmr_cpmg_run_labels(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w,
                                   uint32_t echoes, uint32_t steps_per_echo, int32_t bc, uint32_t shift,
                                   const double *label_delta, uint32_t *label_hits, double *echo_sums);

/* CPMG with n_sets sets of per-label relaxivity evaluated on one trajectory
 * (relaxation never changes a trajectory): set_label_delta holds
 * [n_sets][n_labels] deltas and echo_sums receives [n_sets][echoes+1] sums,
 * each set's equal to an rwnmr_cpmg_run_labels call with that set's deltas
 * from the same start state (per-walker energies bit for bit, sums to
 * rounding).  Sets are walked 8 at a time, each walker carrying 8 energies;
 * every group of 8 replays the trajectories from the start state.  The
 * walkers end as after one run; w->energy (nullable) returns the last
 * group's first set's energies; w->hits is not updated.  Host or device
 * placement of every array follows w->on_device.  No reference counterpart:
 * it scores a whole population of the relaxivity fit's genetic algorithm
 * (B/GA/ga_core.cpp:463-490 evaluates one individual per full simulation). */
rwnmr_status rwnmr_cpmg_run_label_sets(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                       uint32_t steps_per_echo, int32_t bc, uint32_t n_sets,
                                       const double *set_label_delta, double *echo_sums);

/* One CPMG instance of a batch: its image, device-resident walkers
 * (on_device must be 1; per-walker penalty or delta_uniform), echo count,
 * steps per echo, boundary, and a device echo_sums array of echoes+1. */
typedef struct {
    const rwnmr_image *image;
    rwnmr_walkers walkers;
    uint32_t echoes;
    uint32_t steps_per_echo;
    int32_t bc;
    double *echo_sums;
} rwnmr_cpmg_instance;

/* Many CPMG instances (uneven images, walker counts, relaxivities, echo
 * counts) in one persistent launch fed by an atomic tile queue, longest
 * tiles first.  flags = 1 runs the instances one launch each instead (the
 * comparison path).  Results equal n_inst separate rwnmr_cpmg_run calls
 * bit for bit in the walkers; echo sums agree to rounding.  Asynchronous on
 * the context stream.  No reference counterpart: the reference simulates
 * one sample per process (B/NMR_Simulation/NMR_cpmg_cuda.cpp:929-1774). */
rwnmr_status rwnmr_cpmg_batch_run(rwnmr_ctx *ctx, const rwnmr_cpmg_instance *inst, uint32_t n_inst, uint32_t flags);

/* Differentiable CPMG (no reference counterpart: the reference fits
 * relaxivity with a genetic algorithm, B/GA/ga_core.cpp:463-490, over one
 * relaxivity per walker).  Relaxation never changes a trajectory, so
 * dM(n)/d delta_l = sum_w E_w(n) h_wl(n) / delta_l with h_wl(n) the
 * walker's hits on solid label l up to echo n.
 *
 * rwnmr_ckpt_bytes: size of the checkpoint buffer for n walkers, n_segments
 * = ceil(echoes / every) segments and n_labels labels.
 * rwnmr_cpmg_forward_ckpt: the labelled CPMG run (as rwnmr_cpmg_run_labels,
 * device walkers with an energy array, device label_delta and echo_sums of
 * echoes+1) that also stores, every `every` echoes, each walker's position,
 * RNG state, energy and per-label hits into ckpt (device memory of
 * rwnmr_ckpt_bytes), and the per-label hits at the end.  Asynchronous.
 * rwnmr_cpmg_backward: given dL/dM(n) (device, echoes+1) returns
 * dL/d delta_l (host, n_labels) by replaying every segment from its
 * checkpoint in parallel (one thread per walker and segment) instead of
 * storing hit histories.  Synchronous. */
uint64_t rwnmr_ckpt_bytes(uint64_t n, uint32_t n_segments, uint32_t n_labels);
rwnmr_status rwnmr_cpmg_forward_ckpt(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                     uint32_t steps_per_echo, int32_t bc, const double *label_delta, uint32_t every,
                                     void *ckpt, double *echo_sums);
rwnmr_status rwnmr_cpmg_backward(rwnmr_ctx *ctx, const rwnmr_image *img, const void *ckpt, uint64_t n,
                                 uint32_t echoes, uint32_t steps_per_echo, int32_t bc, const double *label_delta,
                                 uint32_t every, const double *grad_echo, double *grad_delta);

/* rwnmr_cpmg_forward_jacobian (no reference counterpart; for least-squares
 * fits of the per-label relaxivities): the labelled CPMG run of
 * rwnmr_cpmg_run_labels (shift 0) that also returns, in one pass, the
 * Jacobian of the echo sums with respect to the log deltas,
 * jac[l][n] = dM(n)/d log delta_l = sum_w h_wl(n) E_w(n) for l < n_labels
 * and n = 0..echoes ([n_labels][echoes+1], host or device like the
 * walkers).  label_hits (nullable, [n_labels][n] uint32, in/out) are the
 * hits before the run, counted in h (NULL: zero).  Needs the unfolded
 * labelled image (the default); one walker per lane with the context's
 * threads per block. */
rwnmr_status rwnmr_cpmg_forward_jacobian(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                         uint32_t steps_per_echo, int32_t bc, const double *label_delta,
                                         uint32_t *label_hits, double *echo_sums, double *jac);

/* PFGSE, reference GPU semantics (NMR_PFGSE::simulation_cuda,
 * B/NMR_Simulation/NMR_pfgse_cuda.cpp:536-1348): for each diffusion time i
 * in order, walk steps_inc[i] more steps (PFG_map_*, :34-339), multiply the
 * energy by pow(delta, hits of this increment) (PFG_evaluate_energy,
 * :342-354) and record M[i][k] = sum_w E_w cos((k . (r_w - r0_w)) * voxel)
 * for every k (PFG_measure + PFG_reduce, :358-389, 504-532, one pass for
 * all k instead of a launch per k).  x0/y0/z0 are the walkers' initial
 * positions.  n_sets == 0: per-walker delta (w->penalty / delta_uniform),
 * M is [n_delta][n_k].  n_sets > 0: set_delta[n_sets] uniform deltas are
 * evaluated on the same trajectories (relaxation does not change them) and
 * M is [n_sets][n_delta][n_k]; w->energy then returns set 0's energy.
 * k_per_delta = 0: kx/ky/kz hold one set of n_k wave vectors (the
 * reference's fixed gradient list, NMR_pfgse.cpp:272-313); 1: they hold
 * n_delta x n_k values, a k set per diffusion time.
 * msd (nullable, [n_delta][3]): sums over walkers of the squared x, y, z
 * displacements in voxel^2 at each diffusion time, unweighted (the inputs of
 * recoverDmsdWithoutSampling, NMR_pfgse.cpp:1098-1160; exact integers).
 * Host/device placement of every array follows w->on_device. */
rwnmr_status rwnmr_pfgse_run(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, const int32_t *x0,
                             const int32_t *y0, const int32_t *z0, const uint32_t *steps_inc,
                             uint32_t n_delta, const double *kx, const double *ky, const double *kz,
                             uint32_t n_k, int32_t k_per_delta, double voxel, int32_t bc, uint32_t shift,
                             uint32_t n_sets, const double *set_delta, double *M, double *msd);
/* PFGSE with walker-sampling statistics (NMR_PFGSE::recoverDsatWithSampling,
 * B/NMR_Simulation/NMR_pfgse.cpp:876-1060; samples as
 * NMR_pfgse_cuda.cpp:1350-2906 form them): walkers split into n_samples
 * consecutive groups of N / n_samples (the remaining walkers walk but
 * belong to no sample); M_samples gets [n_samples][sets|1][n_delta][n_k]
 * sums per sample, M and msd the all-walker results of rwnmr_pfgse_run.
 * Each sample is summed over whole blocks of its own, so its sums do not
 * depend on the other samples. */
rwnmr_status rwnmr_pfgse_run_sampled(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, const int32_t *x0,
                                     const int32_t *y0, const int32_t *z0, const uint32_t *steps_inc,
                                     uint32_t n_delta, const double *kx, const double *ky, const double *kz,
                                     uint32_t n_k, int32_t k_per_delta, double voxel, int32_t bc, uint32_t n_sets,
                                     const double *set_delta, uint32_t n_samples, double *M_samples, double *M,
                                     double *msd);

/* Random-gather probe (measurement; no reference counterpart): times
 * loads_per_thread independent random 16-byte loads per lane (the walk
 * kernels' chunk reloads), over 4096 x 256 lanes, from a device table of
 * table_bytes (a power of two, 16 B .. 64 GiB), which the probe allocates,
 * fills, reads once untimed and frees.  The table size decides what serves
 * the loads: L2 (a few MiB), the Infinity Cache / MALL (up to 256 MiB) or
 * HBM.  Returns random loads per second (each a separate cache line) and
 * the timed kernel's milliseconds.  Synchronous on the context stream. */
rwnmr_status rwnmr_probe_gather(rwnmr_ctx *ctx, uint64_t table_bytes, uint32_t loads_per_thread, double *loads_per_s,
                                double *kernel_ms);

/* Laplace inversion, same names and semantics as the reference's C API
 * (B/Laplace/include/nmrinv.h:13-39, B/Laplace/src/nmrinv.cpp): one global
 * inverter guarded by a mutex; FP64 throughout. */
void initialize_inverter(int num_raw_bins, double *raw_bins, double min_t2, double max_t2, int num_t2_bins,
                         int num_lambdas, int prune_num, double noise_amp);
void initialize_inverter_with_t2_bins(int num_raw_bins, double *raw_bins, int num_t2_bins, double *t2_bins,
                                      int num_lambdas, int prune_num, double noise_amp);
double find_best_lambda(int num_raw_amps, double *raw_amps);
void get_processed_raw_amps(int data_size, double *raw_amps, double *processed_raw_amps);
bool invert(int data_size, double *raw_amps, double *out_result);
bool invert_with_multiple_noises(const int raw_amps_size, double *raw_amps_ptr, const int num_inversions,
                                  double *raw_noise_ptr, double *t2_amps_ptr);
bool get_lcurve_values(double *solution_norms_ptr, double *residual_norms_ptr, double *lambdas_ptr,
                        double *curvature_ptr);
double get_tikhonov_lambda(void);
void set_lambda(double tik_lambda);
bool get_noise(double *out_noise_data);
void set_noise(double noise_amp, double *in_noise_data);
void get_used_raw_bins(double *out_used_raw_bins);
void get_used_t2_bins(double *out_used_t2_bins);
/* Sizes of the arrays the getters above fill (the reference leaves these to
 * the caller's knowledge of its config). */
int rwnmr_inverter_num_used_raw_bins(void);
/* log10 range of the L-curve lambdas (cpmg.config MIN_LAMBDA / MAX_LAMBDA,
 * NMRInverterConfig of B/Laplace/include/nmrinv_core.h; the reference's C
 * API has no setter and keeps the default -4 .. 2).  Applies to the next
 * initialize_inverter*. */
void rwnmr_inverter_set_lambda_range(double min_log10, double max_log10);
int rwnmr_inverter_num_t2_bins(void);
int rwnmr_inverter_num_lambdas(void);

#ifdef __cplusplus
}
#endif

#endif /* RWNMR_H */
