/*
 * rwnmr_output.h -- C ABI of the reference's output tree: the text and CSV
 * files a simulation leaves under <root>/db/<NAME>/ (SimInfo.txt,
 * NMR_binImage.txt, the CPMG, multitau and PFGSE result files), written
 * byte for byte in the reference's formats.  Reference paths are under
 * src/backend/ of RandomWalk_NMR (B/).
 *
 * Number formats are those of a C++ ostream: "%.17g" where the reference
 * sets std::numeric_limits<double>::max_digits10, "%.6g" for its default
 * precision; integers in decimal.  Every writer returns RWNMR_ERR_INVALID
 * when the file cannot be opened.
 */
#ifndef RWNMR_OUTPUT_H
#define RWNMR_OUTPUT_H

#include "rwnmr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Create a directory, mode 0700, and any missing parents (createDirectory,
 * B/FileHandler/smallFunctions.cpp:83-96, makes one level); an existing
 * one is fine. */
rwnmr_status rwnmr_make_dirs(const char *path);

/* SimInfo.txt (NMR_Simulation::saveInfo, B/NMR_Simulation/NMR_Simulation.cpp:1588-1642).
 * occupancy is walkers / pores (printed x 100 with a "%"). */
rwnmr_status rwnmr_write_sim_info(const char *path, const char *name, const char *data_path, const char *image_path,
                                  double voxel_resolution, double D0, int32_t n_images, double occupancy,
                                  uint64_t initial_seed, int32_t gpu, const char *bc);

/* NMR_binImage.txt of a 3D image (BitBlock::saveBitBlockArray_3D,
 * B/BitBlock/bitBlock.cpp:223-234, B/FileHandler/fileHandler.cpp:192-235):
 * header, dimensions, then "index, word" per block (4x4x4 bitBlock words as
 * rwnmr_bitblock_pack makes them). */
rwnmr_status rwnmr_write_bin_image(const char *path, const uint64_t *blocks, int32_t cols, int32_t rows,
                                   int32_t depth);

/* cpmg_decay.csv (NMR_cpmg::writeT2decay, B/NMR_Simulation/NMR_cpmg.cpp:487-512):
 * "time,signal,noise,noiseless", signal = amps + noise; noise nullable (zeros). */
rwnmr_status rwnmr_write_cpmg_decay(const char *path, uint32_t n, const double *times, const double *amps,
                                    const double *noise);
/* Two columns "a,b" with a header (cpmg_T2.csv "NMRT2_bins,NMRT2_amps",
 * NMR_cpmg.cpp:514-536; histogram files "Bins,Amps", :578-601,
 * NMR_pfgse.cpp:1537-1560, NMR_multitau.cpp:288-311; multitau_decay.csv
 * "echo_time,signal", NMR_multitau.cpp:221-246). */
rwnmr_status rwnmr_write_two_columns(const char *path, const char *header, uint32_t n, const double *a,
                                     const double *b);
/* Histogram list (writeHistogramList, NMR_cpmg.cpp:603-636): header
 * "Bins0,Amps0,Bins1,Amps1,...," then one row per bin, every value followed
 * by a comma.  bins/amps are [n_hist][n_bins]. */
rwnmr_status rwnmr_write_histogram_list(const char *path, int32_t n_hist, int32_t n_bins, const double *bins,
                                        const double *amps);

/* PFGSE_parameters.txt (NMR_PFGSE::writeParameters, NMR_pfgse.cpp:1359-1391). */
rwnmr_status rwnmr_write_pfgse_parameters(const char *path, double D0, double pulse_width, double gamma,
                                          double gmax_x, double gmax_y, double gmax_z, uint32_t gradients,
                                          const double *times, uint32_t n_times);
/* PFGSE_gradient.csv (writeGvector, :1393-1427): Id,Gx,Gy,Gz,Kx,Ky,Kz. */
rwnmr_status rwnmr_write_pfgse_gradient(const char *path, uint32_t n, const double *gx, const double *gy,
                                        const double *gz, const double *kx, const double *ky, const double *kz);
/* timesamples/echoes_<i>.csv (writeEchoes, :1429-1462). */
rwnmr_status rwnmr_write_pfgse_echoes(const char *path, uint32_t n, const double *gradient, const double *Mkt,
                                      const double *noise, const double *Mkt_std, const double *lhs,
                                      const double *lhs_std, const double *rhs);
/* timesamples/msd_<i>.csv (writeMsd, :1464-1495): msd[3], msd_std[3],
 * dmsd[3], dmsd_std[3] in x, y, z order; no trailing newline. */
rwnmr_status rwnmr_write_pfgse_msd(const char *path, const double *msd, const double *msd_std, const double *dmsd,
                                   const double *dmsd_std);
/* PFGSE_results.csv: header (createResultsFile, :1598-1625) ... */
rwnmr_status rwnmr_write_pfgse_results_header(const char *path);
/* ... and one appended row per diffusion time (writeResults, :1627-1656):
 * d[11] = D_sat, D_sat error, D_sat std, (points), D_msd, D_msd std,
 * DmsdX, its std, DmsdY, its std, DmsdZ, its std; points is printed as an
 * integer between d[2] and d[3]. */
rwnmr_status rwnmr_append_pfgse_results(const char *path, double time, const double *d, uint32_t points);

#ifdef __cplusplus
}
#endif

#endif /* RWNMR_OUTPUT_H */
