/*
 * rwnmr_io.h -- C ABI of image ingestion and the reference's text outputs:
 * PNG slice stacks binarised as the reference does, the micro-CT config
 * (uct.config), pore / interface statistics, ImageInfo.txt, walkers.csv.
 * Reference paths are under src/backend/ of RandomWalk_NMR (B/).
 */
#ifndef RWNMR_IO_H
#define RWNMR_IO_H

#include "rwnmr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The uct.config keys (B/ConfigFiles/uct_config.h; defaults of
 * config/default/uct.config).  img_files_list "Empty" means: build the slice
 * names from dir_path + filename + zero-padded index + extension. */
typedef struct rwnmr_uct_config {
    char dir_path[1024];       /* DIR_PATH */
    char filename[256];        /* FILENAME */
    int32_t first_idx;         /* FIRST_IDX */
    int32_t digits;            /* DIGITS */
    char extension[32];        /* EXTENSION */
    int32_t slices;            /* SLICES */
    double resolution;         /* RESOLUTION (um per voxel) */
    int32_t voxel_division;    /* VOXEL_DIVISION */
    char img_files_list[1024]; /* IMG_FILES_LIST */
} rwnmr_uct_config;

/* Defaults of config/default/uct.config (uct_config.cpp:15-25 reads that
 * file first, then the user's). */
void rwnmr_uct_config_defaults(rwnmr_uct_config *cfg);
/* Apply the "KEY: value" lines of a uct.config file on top of cfg
 * (uct_config::readConfigFile, B/ConfigFiles/uct_config.cpp:47-115: every
 * ": " of a line splits a key candidate from the rest of the line; values
 * are not trimmed).  Unknown keys are ignored; a malformed number fails. */
rwnmr_status rwnmr_uct_config_read(const char *path, rwnmr_uct_config *cfg);
/* Path of slice `slice` (0-based): line `slice` of IMG_FILES_LIST, else
 * DIR_PATH + FILENAME + setfill('0') setw(DIGITS) (FIRST_IDX + slice) +
 * EXTENSION (uct_config.cpp:165-196, uct_config.h:73-78). */
rwnmr_status rwnmr_uct_slice_path(const rwnmr_uct_config *cfg, int32_t slice, char *buf, uint32_t buflen);
/* Width and height of the first slice, depth = SLICES. */
rwnmr_status rwnmr_uct_dims(const rwnmr_uct_config *cfg, int32_t *width, int32_t *height, int32_t *depth);
/* All slices binarised into solid[z][y][x] (1 = solid), every slice must
 * be width x height (NMR_Simulation::loadRockImage / loadRockImageFromList,
 * B/NMR_Simulation/NMR_Simulation.cpp:484-575). */
rwnmr_status rwnmr_uct_load(const rwnmr_uct_config *cfg, int32_t width, int32_t height, int32_t depth,
                            uint8_t *solid);

/* PNG header size. */
rwnmr_status rwnmr_png_size(const char *path, int32_t *width, int32_t *height);
/* Decode one PNG and binarise it as imread(path, IMREAD_COLOR) followed by
 * createBinaryMap (B/NMR_Simulation/NMR_Simulation.cpp:536-617): a pixel is
 * solid (1) when any colour channel is nonzero after conversion to 8-bit
 * colour (alpha dropped, 16-bit samples reduced to the high byte, grey and
 * palette expanded).  All colour types, bit depths and Adam7 interlacing. */
rwnmr_status rwnmr_png_binarize(const char *path, int32_t width, int32_t height, uint8_t *solid);

/* Pore voxels, porosity, pore-matrix interfaces and S/Vp of a binary volume
 * (countPoresInBitBlock, countInterfacePoreMatrix, updateSVp, updatePorosity:
 * B/NMR_Simulation/NMR_Simulation.cpp:675-727, 818-923): an interface is a
 * pair of face neighbours (periodic wrap, +x, +y, and +z when depth > 1) of
 * different phase; S/Vp = interfaces * voxel^2 / (pores * voxel^3). */
rwnmr_status rwnmr_image_stats(const uint8_t *solid, int32_t width, int32_t height, int32_t depth, double voxel,
                               uint64_t *pores, double *porosity, uint64_t *interfaces, double *svp);

/* ImageInfo.txt in the reference's format (NMR_Simulation::saveImageInfo,
 * B/NMR_Simulation/NMR_Simulation.cpp:1644-1672). */
rwnmr_status rwnmr_write_image_info(const char *path, const char *image_path, int32_t width, int32_t height,
                                    int32_t depth, double resolution, int32_t voxel_shift, double step_length,
                                    double porosity, double svp);

/* walkers.csv in the reference's format (NMR_Simulation::saveWalkers,
 * B/NMR_Simulation/NMR_Simulation.cpp:1674-1712): initial and final
 * positions, collisions, XI rate, energy, initial RNG seed, precision 6. */
rwnmr_status rwnmr_write_walkers_csv(const char *path, uint64_t n, const int32_t *x0, const int32_t *y0,
                                     const int32_t *z0, const int32_t *x, const int32_t *y, const int32_t *z,
                                     const uint32_t *collisions, const double *xi_rate, const double *energy,
                                     const uint64_t *seed);

/* Internal field of a uniform gradient (InternalField::fillDataArray,
 * B/NMR_Simulation/InternalField.cpp:95-134): value = coordinate along
 * `direction` (0 x, 1 y, 2 z) * gradient * 1e-6 * resolution in pore voxels,
 * 0 in solid voxels; field[x + y W + z W H]. */
rwnmr_status rwnmr_internal_field_uniform(const uint8_t *solid, int32_t width, int32_t height, int32_t depth,
                                          double resolution, double gradient, int32_t direction, double *field);
/* Internal field from a raw file of W*H*D native doubles
 * (InternalField::readDataFromFile, InternalField.cpp:167-200; the file size
 * must match). */
rwnmr_status rwnmr_internal_field_read(const char *path, int32_t width, int32_t height, int32_t depth,
                                       double *field);

#ifdef __cplusplus
}
#endif

#endif /* RWNMR_IO_H */
