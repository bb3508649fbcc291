/*
 * rwnmr_drivers.h -- C ABI of the CPMG and PFGSE sequence drivers (the
 * host logic of NMR_cpmg::run_simulation and NMR_PFGSE::run around the GPU
 * kernels).  Reference paths under src/backend/ of RandomWalk_NMR.
 */
#ifndef RWNMR_DRIVERS_H
#define RWNMR_DRIVERS_H

#include "rwnmr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* CPMG sequence parameters (cpmg.config OBS_TIME / APPLY_BULK and
 * rwnmr.config STEPS_PER_ECHO / D0 / BULK_TIME / RHO / BC; uct.config
 * RESOLUTION / VOXEL_DIVISION give voxel_size and shift). */
typedef struct {
    double observation_time;   /* ms */
    uint32_t steps_per_echo;
    double voxel_size;         /* um (resolution / voxel division) */
    double D0;                 /* um^2/ms */
    double bulk_time;          /* ms */
    int32_t apply_bulk;
    double rho;                /* um/s, uniform relaxivity (RHO_TYPE uniform) */
    const double *sigmoid;     /* nullable: 8 sigmoid parameters (RHO_TYPE sigmoid) */
    int32_t bc;
    uint32_t shift;
} rwnmr_cpmg_params;

/* NMR_cpmg::run_simulation (B/NMR_Simulation/NMR_cpmg.cpp:112-175, 346-385):
 * time framework (echoes = ceil(round(T 6 D0 / eps^2) / steps_per_echo)),
 * optional sigmoid relaxivity from a map pass, the GPU CPMG run from the
 * walkers' current state, bulk factor exp(-t/T2b), normalisation by the
 * t = 0 value.  times/signal receive echoes+1 points (capacity must be
 * enough); uniform rho is used when w->penalty is NULL. */
rwnmr_status rwnmr_cpmg_simulate(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w,
                                 const rwnmr_cpmg_params *p, uint32_t capacity, double *times, double *signal,
                                 uint32_t *n_points);

/* PFGSE parameters (pfgse.config PULSE_WIDTH / APPLY_BULK / THRESHOLD_*,
 * rwnmr.config D0 / BULK_TIME / RHO / STEPS_PER_ECHO / BC). */
typedef struct {
    double voxel_size;         /* um */
    double D0;                 /* um^2/ms */
    double bulk_time;          /* ms */
    int32_t apply_bulk;
    uint32_t steps_per_echo;
    int32_t bc;
    double pulse_width;        /* ms; enters RHS = -k^2 (Delta - pulse/3) */
    double rho;                /* um/s (uniform), used when n_rho_sets == 0 */
    const double *sigmoid;     /* nullable: 8 sigmoid parameters */
    uint32_t n_rho_sets;       /* > 0: evaluate these relaxivities on one trajectory set */
    const double *rho_sets;    /* [n_rho_sets] um/s */
    int32_t k_per_delta;       /* k arrays are [n_times][n_k] instead of [n_k] */
    uint32_t threshold_samples;/* THRESHOLD_TYPE samples: points in the fit (0 = all) */
    double threshold_lhs;      /* THRESHOLD_TYPE lhs: value in (0,1), else unused */
    uint32_t threshold_window;
} rwnmr_pfgse_params;

/* Caller-owned outputs, S = n_rho_sets (or 1), T = n_times, K = n_k. */
typedef struct {
    double *Mkt;        /* [S][T][K] M(k, Delta) after bulk */
    double *rhs;        /* [S][T][K] -k^2 (Delta - pulse/3) */
    double *lhs;        /* [S][T][K] log(max(M_k/M_0, 1e-6)) */
    double *D_sat;      /* [S][T] Stejskal-Tanner D */
    double *D_sat_err;  /* [S][T] its standard error */
    uint32_t *points;   /* [S][T] points used in the fit */
    double *msd;        /* [T][3] sums of squared displacements, voxel^2 */
    double *D_msd;      /* [T] msd / (6 Delta) */
    uint64_t *steps;    /* [T] total steps at each Delta */
} rwnmr_pfgse_result;

/* NMR_PFGSE::run (B/NMR_Simulation/NMR_pfgse.cpp:98-131, 272-361, 441-692,
 * 1098-1160) on host walkers: steps per Delta from the time framework,
 * one GPU pass over all Delta and k (rwnmr_pfgse_run), then bulk, RHS,
 * LHS, threshold, Stejskal-Tanner D and msd D per Delta.  Exposure times
 * must be ascending. */
rwnmr_status rwnmr_pfgse_simulate(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w,
                                  const rwnmr_pfgse_params *p, const double *exposure_times, uint32_t n_times,
                                  const double *kx, const double *ky, const double *kz, uint32_t n_k,
                                  rwnmr_pfgse_result *out);

/* Gradient list 0..Gmax with accumulated steps (setGradientVector,
 * NMR_pfgse.cpp:272-297). */
void rwnmr_pfgse_gradients(double gmax_x, double gmax_y, double gmax_z, uint32_t points, double *gx, double *gy,
                           double *gz);
/* k = (pulse 1e-3)(gamma 1e6)(G 1e-8) (computeWaveVectorK, NMR_pfgse.cpp:483-486). */
double rwnmr_pfgse_wave_vector(double gradient, double pulse_width, double gamma);
/* No-intercept least squares slope and its standard error
 * (LeastSquareAdjust.cpp:77-170; NMR_pfgse.cpp:678-682). */
int32_t rwnmr_stejskal_tanner(const double *rhs, const double *lhs, uint32_t n, double *D, double *D_err);
/* Walker-sampling statistics of one diffusion time (NMR_PFGSE::
 * recoverDsatWithSampling, B/NMR_Simulation/NMR_pfgse.cpp:876-1060, after
 * its optional bulk and noise terms, which the caller applies to M_samples):
 * M_samples is [n_samples][n_k] (one energy set of rwnmr_pfgse_run_sampled);
 * per sample LHS_k = log(max(M_k / M_0, 1e-6)); mean and population standard
 * deviation over samples of M and LHS per k (any output may be NULL); one
 * no-intercept Stejskal-Tanner fit per sample over the first n_points points
 * (rhs shared), D_mean / D_std / mean fit error over samples. */
rwnmr_status rwnmr_pfgse_sample_stats(const double *M_samples, uint32_t n_samples, uint32_t n_k, const double *rhs,
                                      uint32_t n_points, double *mean_M, double *std_M, double *mean_lhs,
                                      double *std_lhs, double *D_samples, double *D_mean, double *D_std,
                                      double *D_err_mean);
/* Points kept by the LHS-value threshold (setThresholdFromLHSValue, NMR_pfgse.cpp:399-419). */
uint32_t rwnmr_pfgse_threshold_lhs(const double *lhs, uint32_t n, double value, uint32_t window);
/* The same with noise (setThresholdFromLHSWindow, NMR_pfgse.cpp:405-439):
 * the first point where the mean of the last `window` LHS values falls
 * below log(value); n when n < window. */
uint32_t rwnmr_pfgse_threshold_lhs_window(const double *lhs, uint32_t n, double value, uint32_t window);

/* ---- multitau (one CPMG per echo time) ---- */

/* multitau.config TAU_MIN / TAU_MAX / TAU_POINTS / TAU_SCALE / TAU_VALUES /
 * COMPLETE_DECAY. */
typedef struct {
    double tau_min, tau_max;   /* ms */
    uint32_t tau_points;
    int32_t tau_scale;         /* 0 linear, 1 log, 2 manual (tau_values) */
    const double *tau_values;  /* manual: [tau_points] ms */
    int32_t complete_decay;    /* 1: each CPMG runs the cpmg observation time */
} rwnmr_multitau_params;

/* Echo-time sequence (NMR_multitau::setTauSequence,
 * B/NMR_Simulation/NMR_multitau.cpp:73-110): tau from linspace / logspace
 * (start + step i, base 10) / the manual list; steps = ceil(tau / dt), made
 * even, and forced strictly increasing (else previous + 2); echo time =
 * steps dt.  steps / echo_times hold tau_points values. */
rwnmr_status rwnmr_multitau_sequence(const rwnmr_multitau_params *p, double time_interval, uint32_t *steps,
                                     double *echo_times);
/* NMR_multitau::run (NMR_multitau.cpp:119-204): for every echo time one
 * CPMG (rwnmr_cpmg_simulate with steps_per_echo = that time's steps and the
 * exposure = that echo time, or the cpmg observation time with
 * complete_decay), each from the initial walker state w0 (left unchanged);
 * signal[i] = the normalised first-echo amplitude of run i
 * (multitau_decay.csv).  echo_times / signal hold tau_points values. */
rwnmr_status rwnmr_multitau_simulate(rwnmr_ctx *ctx, const rwnmr_image *img, const rwnmr_walkers *w0,
                                     const rwnmr_multitau_params *mp, const rwnmr_cpmg_params *cp,
                                     double *echo_times, double *signal);

/* ---- collision-histogram ("fast") CPMG method ---- */

/* Bins of a histogram of `size`: size (linear) or size + 1 (log scale; bin 0
 * is the no-collision control entry), CollisionHistogram::createBlankHistogram
 * (B/NMR_Simulation/CollisionHistogram.cpp:57-72). */
int32_t rwnmr_histogram_bins(int32_t size, int32_t log_scale);
/* Echo segments of the histogram list (NMR_Simulation::initHistogramList,
 * B/NMR_Simulation/NMR_Simulation.cpp:1456-1479): n_histograms segments of
 * echoes / n echoes (one if echoes < n), the remainder added to the last.
 * first_echo / last_echo hold n_histograms entries; *n_used gets the count. */
rwnmr_status rwnmr_histogram_list(uint32_t echoes, int32_t n_histograms, int32_t *n_used, uint32_t *first_echo,
                                  uint32_t *last_echo);
/* Histogram of collision rates xi = collisions / steps over n walkers
 * (CollisionHistogram::fillHistogram, CollisionHistogram.cpp:74-250):
 * linear: bins (i + 1/2) / size, index floor(xi size) (xi = 1 in the last
 * bin); log: bin 0 = round(log10 min rate) - 1 counts walkers without
 * collisions, bins 1.. = 10^(e0 + (i - 1)(0 - e0)/(size - 1)) with
 * e0 = round(log10 min rate), index round((log10 xi - log10 bin1) / gap) + 1
 * (non-positive -> bin 1).  Amplitudes are fractions of walkers. */
rwnmr_status rwnmr_collision_histogram(const uint32_t *collisions, uint64_t n, uint64_t steps, int32_t size,
                                       int32_t log_scale, double *bins, double *amps);
/* Per-bin penalties delta(rho(xi))^(xi steps_per_echo) with xi = the bin value
 * (NMR_cpmg::createPenaltiesVector, B/NMR_Simulation/NMR_cpmg.cpp:235-279);
 * rho uniform, or the 8-parameter two-sigmoid of xi when sigmoid != NULL
 * (B/Walker/walker.h:174-196); delta = 1 - (2/3) rho step_length / (D0 1e3). */
rwnmr_status rwnmr_histogram_penalties(const double *bins, int32_t n_bins, double rho, const double *sigmoid,
                                       double step_length, double D0, uint32_t steps_per_echo, double *penalties);
/* Histogram-propagated decay (NMR_cpmg::histogram_simulation,
 * NMR_cpmg.cpp:281-344): signal[0] = 1; for each segment h the bin energies
 * start at (last signal) * amps[h][i] and are multiplied by the penalties
 * once per echo of the segment, the signal being their sum.  amps is
 * [n_hist][n_bins]; signal gets 1 + sum of segment lengths values. */
rwnmr_status rwnmr_histogram_decay(const double *amps, int32_t n_hist, int32_t n_bins, const uint32_t *first_echo,
                                   const uint32_t *last_echo, const double *penalties, double *signal);
/* The whole histogram method on the GPU map walk (mapSimulation with
 * histograms, B/NMR_Simulation/NMR_Simulation_cuda.cpp:726-1270, then
 * NMR_cpmg::run_simulation's "histogram" branch, NMR_cpmg.cpp:158-165): per
 * segment a map walk of steps_per_echo x (segment echoes) steps with fresh
 * collision counts (walkers keep moving from segment to segment), a segment
 * histogram; a global histogram of the summed collisions over all steps
 * gives the bins for the penalties.  signal: echoes + 1 values (before bulk
 * relaxation and normalisation).  bins_out ([n_bins]) and amps_out
 * ([n_used][n_bins]) are nullable. */
rwnmr_status rwnmr_cpmg_histogram_simulate(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                           uint32_t steps_per_echo, int32_t bc, uint32_t shift, int32_t n_histograms,
                                           int32_t size, int32_t log_scale, double rho, const double *sigmoid,
                                           double step_length, double D0, double *signal, double *bins_out,
                                           double *amps_out, int32_t *n_used);

#ifdef __cplusplus
}
#endif

#endif /* RWNMR_DRIVERS_H */
