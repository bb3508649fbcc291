/*
 * rwnmr_host.h -- C ABI of the host-side set-up around the hot path:
 * synthetic images, walker seeding/placement, time framework, penalty.
 * Reference paths are under src/backend/ of RandomWalk_NMR.
 */
#ifndef RWNMR_HOST_H
#define RWNMR_HOST_H

#include "rwnmr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Seeded overlapping solid spheres (centres uniform in the box, radius
 * uniform in [r_min, r_max], voxel solid when its centre lies in a sphere),
 * added until the pore fraction is <= target_porosity.  solid: W*H*D bytes,
 * x fastest, 1 = solid, 0 = pore (the reference's binarisation,
 * B/NMR_Simulation/NMR_Simulation.cpp:577-617).  No reference counterpart:
 * the reference reads micro-CT images; this builds the benchmark inputs. */
int32_t rwnmr_synth_sphere_pack(int32_t W, int32_t H, int32_t D, double r_min, double r_max,
                                double target_porosity, uint64_t seed, uint8_t *solid, double *porosity_out);

/* Label every solid voxel by its nearest of n_sites seeded Voronoi sites,
 * label = site % n_labels (pore voxels get 0).  solid may be NULL (label
 * everything).  No reference counterpart (per-label relaxivity input). */
int32_t rwnmr_synth_voronoi_labels(int32_t W, int32_t H, int32_t D, int32_t n_sites, int32_t n_labels,
                                   uint64_t seed, const uint8_t *solid, uint8_t *labels);

/* Number of pore voxels (countPoresInBitBlock,
 * B/NMR_Simulation/NMR_Simulation.cpp:818-910). */
int64_t rwnmr_count_pores(const uint8_t *solid, int32_t W, int32_t H, int32_t D);

/* Walker seeds and start voxels drawn the reference's way:
 * mt19937(master_seed mod 2^32) with 4096 outputs discarded
 * (NMR_Simulation.cpp:86-95); per walker seed = (out << 32) | next out
 * (createWalkers :1139-1204); then n indices uniform over
 * [0, numberOfPores] inclusive with numberOfPores = (uint)(porosity * volume)
 * (:915-928, :1086-1136) by Lemire's method (libstdc++ >= 11), into the pore
 * list in z, y, x raster order (:930-980, placeWalkersUniformly :1271-1333).
 * An index equal to the pore count (out of range in the reference) selects
 * the last pore. */
int32_t rwnmr_seed_and_place(const uint8_t *solid, int32_t W, int32_t H, int32_t D, uint64_t master_seed,
                             uint64_t n, uint64_t *seeds, int32_t *x, int32_t *y, int32_t *z);

/* The whole walker set-up of an NMR_Simulation (RNG seeding :86-95,
 * setWalkers :324-410, applyVoxelDivision :192-268 of
 * B/NMR_Simulation/NMR_Simulation.cpp) from rwnmr.config WALKERS_PLACEMENT:
 * mode 0 "random" = rwnmr_seed_and_place; 1 "point": every walker at
 * (W/2, H/2, D/2), an error if that voxel is solid; 2 "cubic": walkers
 * uniform over the pore voxels of [c - deviation, c + deviation) per axis,
 * c = (dim - 1) / 2 (an error if there are none).  shift > 0
 * (uct.config VOXEL_DIVISION) then moves every walker to
 * pos * 2^shift + a draw in [0, 2^shift - 1] per axis.  occupancy gets
 * walkers / pore voxels; draws (nullable) the number of 32-bit generator
 * outputs consumed, discard included, so a caller can continue the stream
 * (std::mt19937(seed mod 2^32), discard(draws)) as the reference's later
 * draws (PFGSE noise) do. */
int32_t rwnmr_setup_walkers(const uint8_t *solid, int32_t W, int32_t H, int32_t D, uint64_t master_seed, uint64_t n,
                            int32_t mode, int32_t deviation, uint32_t shift, uint64_t *seeds, int32_t *x,
                            int32_t *y, int32_t *z, double *occupancy, uint64_t *draws);

/* dt = voxel^2 / (6 D0) in ms (setTimeInterval, NMR_Simulation.cpp:179-183). */
double rwnmr_time_step(double voxel_size, double D0);

/* steps = round(time * 6 D0 / voxel^2) (at least steps_per_echo),
 * echoes = ceil(steps / steps_per_echo), steps = echoes * steps_per_echo
 * (setNumberOfStepsFromTime / setTimeFramework, NMR_Simulation.cpp:276-298). */
int32_t rwnmr_time_framework(double time_ms, double voxel_size, double D0, uint32_t steps_per_echo,
                             uint32_t *echoes_out, uint64_t *steps_out);

/* delta = 1 - (2/3) rho voxel / (D0 1e3)  (computeDecreaseFactor,
 * B/Walker/walker.h:203-209; rho in um/s, voxel in um, D0 in um^2/ms). */
double rwnmr_penalty(double rho, double voxel_size, double D0);

#ifdef __cplusplus
}
#endif

#endif /* RWNMR_HOST_H */
