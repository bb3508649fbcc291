// output.cpp -- the reference's result files (SimInfo.txt, NMR_binImage.txt,
// CPMG / multitau / PFGSE CSV and text files) in its exact formats.
//
// Reference behaviour restated (paths under src/backend/ of RandomWalk_NMR):
//   SimInfo.txt                         B/NMR_Simulation/NMR_Simulation.cpp:1588-1642
//   NMR_binImage.txt (3D)               B/FileHandler/fileHandler.cpp:192-235
//   cpmg_decay.csv, cpmg_T2.csv,        B/NMR_Simulation/NMR_cpmg.cpp:487-536, 578-636
//   cpmg_histogram.csv, cpmg_histList.csv
//   multitau_decay.csv                  B/NMR_Simulation/NMR_multitau.cpp:221-246
//   PFGSE_parameters.txt, _gradient.csv, B/NMR_Simulation/NMR_pfgse.cpp:1359-1495, 1598-1656
//   echoes_i.csv, msd_i.csv, PFGSE_results.csv
//
// An ostream with setprecision(p) and the default float field prints a
// double as printf("%.*g", p); the reference uses p = max_digits10 (17) for
// its results and the default 6 elsewhere.
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <string>

#include "../../include/rwnmr/rwnmr_output.h"

namespace {

struct File {
    FILE *f;
    File(const char *path, const char *mode) : f(path ? fopen(path, mode) : nullptr) {}
    ~File()
    {
        if (f)
            fclose(f);
    }
    bool ok() const { return f != nullptr; }
};

inline void g17(FILE *f, double v) { fprintf(f, "%.17g", v); }

}  // namespace

extern "C" {

rwnmr_status rwnmr_make_dirs(const char *path)
{
    if (!path || !*path)
        return RWNMR_ERR_INVALID;
    std::string p(path);
    for (size_t i = 1; i <= p.size(); ++i) {
        if (i == p.size() || p[i] == '/') {
            const std::string part = p.substr(0, i);
            struct stat st;
            if (stat(part.c_str(), &st) == 0) {
                if (!S_ISDIR(st.st_mode))
                    return RWNMR_ERR_INVALID;
                continue;
            }
            if (mkdir(part.c_str(), 0700) != 0 && errno != EEXIST)
                return RWNMR_ERR_INVALID;
        }
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_sim_info(const char *path, const char *name, const char *data_path, const char *image_path,
                                  double voxel_resolution, double D0, int32_t n_images, double occupancy,
                                  uint64_t initial_seed, int32_t gpu, const char *bc)
{
    File o(path, "w");
    if (!o.ok() || !name || !data_path || !image_path || !bc)
        return RWNMR_ERR_INVALID;
    const char *rule = "------------------------------------------------------\n";
    fputs(rule, o.f);
    fprintf(o.f, ">>> NMR SIMULATION 3D PARAMETERS: %s\n", name);
    fputs(rule, o.f);
    fprintf(o.f, "Data path: %s\n", data_path);
    fprintf(o.f, "Image path: %s\n", image_path);
    fprintf(o.f, "Image resolution (um/voxel): %g\n", voxel_resolution);
    fprintf(o.f, "Diffusion coefficient (um^2/ms): %g\n", D0);
    fprintf(o.f, "Number of images: %d\n", n_images);
    fprintf(o.f, "Walkers pore occupancy in simulation: %g%%\n", occupancy * 100.0);
    // the reference always marks the seed as set by the program (both calls
    // of setInitialSeed pass true), so "(defined by user)" never appears
    fprintf(o.f, "Initial seed: %llu\n", (unsigned long long)initial_seed);
    fprintf(o.f, "GPU usage: %s\n", gpu ? "ON" : "OFF");
    fprintf(o.f, "BC: %s\n", bc);
    fputs(rule, o.f);
    fputs("\n", o.f);
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_bin_image(const char *path, const uint64_t *blocks, int32_t cols, int32_t rows,
                                   int32_t depth)
{
    if (!blocks || cols <= 0 || rows <= 0 || depth <= 0)
        return RWNMR_ERR_INVALID;
    File o(path, "w");
    if (!o.ok())
        return RWNMR_ERR_INVALID;
    const int32_t bR = (rows + 3) / 4, bC = (cols + 3) / 4, bD = (depth + 3) / 4;
    const int64_t nb = (int64_t)bR * bC * bD;
    fputs("blocks, bRows, bColumns, bDepth, imgRows, imgColumns, imgDepth, \n", o.f);
    fprintf(o.f, "%d, %d, %d, %d, %d, %d, %d\n", (int)nb, bR, bC, bD, rows, cols, depth);
    fputs("\nblockID, blockData\n", o.f);
    for (int64_t i = 0; i < nb; ++i)
        fprintf(o.f, "%d, %llu\n", (int)i, (unsigned long long)blocks[i]);
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_cpmg_decay(const char *path, uint32_t n, const double *times, const double *amps,
                                    const double *noise)
{
    if ((n && (!times || !amps)))
        return RWNMR_ERR_INVALID;
    File o(path, "w");
    if (!o.ok())
        return RWNMR_ERR_INVALID;
    fputs("time,signal,noise,noiseless\n", o.f);
    for (uint32_t i = 0; i < n; ++i) {
        const double nz = noise ? noise[i] : 0.0;
        g17(o.f, times[i]);
        fputc(',', o.f);
        g17(o.f, amps[i] + nz);
        fputc(',', o.f);
        g17(o.f, nz);
        fputc(',', o.f);
        g17(o.f, amps[i]);
        fputc('\n', o.f);
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_two_columns(const char *path, const char *header, uint32_t n, const double *a,
                                     const double *b)
{
    if (!header || (n && (!a || !b)))
        return RWNMR_ERR_INVALID;
    File o(path, "w");
    if (!o.ok())
        return RWNMR_ERR_INVALID;
    fprintf(o.f, "%s\n", header);
    for (uint32_t i = 0; i < n; ++i) {
        g17(o.f, a[i]);
        fputc(',', o.f);
        g17(o.f, b[i]);
        fputc('\n', o.f);
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_histogram_list(const char *path, int32_t n_hist, int32_t n_bins, const double *bins,
                                        const double *amps)
{
    if (n_hist < 0 || n_bins < 0 || ((int64_t)n_hist * n_bins && (!bins || !amps)))
        return RWNMR_ERR_INVALID;
    File o(path, "w");
    if (!o.ok())
        return RWNMR_ERR_INVALID;
    for (int32_t h = 0; h < n_hist; ++h)
        fprintf(o.f, "Bins%d,Amps%d,", h, h);
    fputc('\n', o.f);
    for (int32_t i = 0; i < n_bins; ++i) {
        for (int32_t h = 0; h < n_hist; ++h) {
            g17(o.f, bins[(size_t)h * n_bins + i]);
            fputc(',', o.f);
            g17(o.f, amps[(size_t)h * n_bins + i]);
            fputc(',', o.f);
        }
        fputc('\n', o.f);
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_pfgse_parameters(const char *path, double D0, double pulse_width, double gamma,
                                          double gmax_x, double gmax_y, double gmax_z, uint32_t gradients,
                                          const double *times, uint32_t n_times)
{
    if (n_times && !times)
        return RWNMR_ERR_INVALID;
    File o(path, "w");
    if (!o.ok())
        return RWNMR_ERR_INVALID;
    // Vector3D norm: sqrt(x x + y y + z z) (B/Math/Vector3D.h:43-48)
    const double norm = sqrt(gmax_x * gmax_x + gmax_y * gmax_y + gmax_z * gmax_z);
    fputs("RWNMR-PFGSE Parameters\n", o.f);
    fputs("D_0: ", o.f);
    g17(o.f, D0);
    fputs("\nPulse width: ", o.f);
    g17(o.f, pulse_width);
    fputs("\nGiromagnetic Ratio: ", o.f);
    g17(o.f, gamma);
    fputs("\nGradient direction: {", o.f);
    g17(o.f, gmax_x / norm);
    fputs(", ", o.f);
    g17(o.f, gmax_y / norm);
    fputs(", ", o.f);
    g17(o.f, gmax_z / norm);
    fprintf(o.f, "}\nGradients: %u\nTimes: [", gradients);
    for (uint32_t t = 0; t < n_times; ++t) {
        if (t > 0)
            fputs(", ", o.f);
        g17(o.f, times[t]);
    }
    fputs("]\n", o.f);
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_pfgse_gradient(const char *path, uint32_t n, const double *gx, const double *gy,
                                        const double *gz, const double *kx, const double *ky, const double *kz)
{
    if (n && (!gx || !gy || !gz || !kx || !ky || !kz))
        return RWNMR_ERR_INVALID;
    File o(path, "w");
    if (!o.ok())
        return RWNMR_ERR_INVALID;
    fputs("Id,Gx,Gy,Gz,Kx,Ky,Kz\n", o.f);
    for (uint32_t i = 0; i < n; ++i) {
        fprintf(o.f, "%u", i);
        const double v[6] = {gx[i], gy[i], gz[i], kx[i], ky[i], kz[i]};
        for (double x : v) {
            fputc(',', o.f);
            g17(o.f, x);
        }
        fputc('\n', o.f);
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_pfgse_echoes(const char *path, uint32_t n, const double *gradient, const double *Mkt,
                                      const double *noise, const double *Mkt_std, const double *lhs,
                                      const double *lhs_std, const double *rhs)
{
    if (n && (!gradient || !Mkt || !noise || !Mkt_std || !lhs || !lhs_std || !rhs))
        return RWNMR_ERR_INVALID;
    File o(path, "w");
    if (!o.ok())
        return RWNMR_ERR_INVALID;
    fputs("Idx,Gradient,NMR_signal(mean),NMR_signal(noise),NMR_signal(std),SAT_lhs(mean),SAT_lhs(std),SAT_rhs\n",
          o.f);
    for (uint32_t i = 0; i < n; ++i) {
        fprintf(o.f, "%u", i);
        const double v[7] = {gradient[i], Mkt[i], noise[i], Mkt_std[i], lhs[i], lhs_std[i], rhs[i]};
        for (double x : v) {
            fputc(',', o.f);
            g17(o.f, x);
        }
        fputc('\n', o.f);
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_pfgse_msd(const char *path, const double *msd, const double *msd_std, const double *dmsd,
                                   const double *dmsd_std)
{
    if (!msd || !msd_std || !dmsd || !dmsd_std)
        return RWNMR_ERR_INVALID;
    File o(path, "w");
    if (!o.ok())
        return RWNMR_ERR_INVALID;
    fputs("msdX(mean),msdX(std),msdY(mean),msdY(std),msdZ(mean),msdZ(std),"
          "DmsdX(mean),DmsdX(std),DmsdY(mean),DmsdY(std),DmsdZ(mean),DmsdZ(std)\n",
          o.f);
    for (int a = 0; a < 3; ++a) {
        g17(o.f, msd[a]);
        fputc(',', o.f);
        g17(o.f, msd_std[a]);
        fputc(',', o.f);
    }
    for (int a = 0; a < 3; ++a) {
        g17(o.f, dmsd[a]);
        fputc(',', o.f);
        g17(o.f, dmsd_std[a]);
        if (a < 2)
            fputc(',', o.f);
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_write_pfgse_results_header(const char *path)
{
    File o(path, "w");
    if (!o.ok())
        return RWNMR_ERR_INVALID;
    fputs("Time,D_sat,D_sat(error),D_sat(std),D_sat(pts),D_msd,D_msd(std),D_msdX,D_msdX(std),D_msdY,"
          "D_msdY(std),D_msdZ,D_msdZ(std)\n",
          o.f);
    return RWNMR_OK;
}

rwnmr_status rwnmr_append_pfgse_results(const char *path, double time, const double *d, uint32_t points)
{
    if (!d)
        return RWNMR_ERR_INVALID;
    File o(path, "a");
    if (!o.ok())
        return RWNMR_ERR_INVALID;
    g17(o.f, time);
    for (int i = 0; i < 3; ++i) {
        fputc(',', o.f);
        g17(o.f, d[i]);
    }
    fprintf(o.f, ",%u", points);
    for (int i = 3; i < 11; ++i) {
        fputc(',', o.f);
        g17(o.f, d[i]);
    }
    fputc('\n', o.f);
    return RWNMR_OK;
}

}  // extern "C"
