// histogram.cpp -- the collision-histogram ("fast") CPMG method: collision
// histograms of walker collision rates, per-bin penalties and the
// histogram-propagated decay.  Host arithmetic only; the collision counts
// come from the GPU map walk (rwnmr_map_run, driven by
// rwnmr_cpmg_histogram_simulate in drivers.hip).
//
// Reference behaviour restated (paths under src/backend/ of RandomWalk_NMR):
//   histogram list (echo segments)     B/NMR_Simulation/NMR_Simulation.cpp:1456-1479
//   linear / log bins and amplitudes   B/NMR_Simulation/CollisionHistogram.cpp:57-250
//   penalties per bin                  B/NMR_Simulation/NMR_cpmg.cpp:235-279
//   histogram simulation               B/NMR_Simulation/NMR_cpmg.cpp:281-344
//
// One quirk is kept as the reference has it: on the log scale bin 0 holds
// round(log10(min rate)) - 1 (a control value for walkers without
// collisions), and the penalty of that bin is computed from it like any
// other bin, i.e. delta^(bin0 * steps_per_echo) with a negative exponent.
#include <math.h>
#include <stdint.h>

#include <vector>

#include "../../include/rwnmr/rwnmr_drivers.h"

extern "C" {

int32_t rwnmr_histogram_bins(int32_t size, int32_t log_scale) { return log_scale ? size + 1 : size; }

rwnmr_status rwnmr_histogram_list(uint32_t echoes, int32_t n_histograms, int32_t *n_used, uint32_t *first_echo,
                                  uint32_t *last_echo)
{
    if (!n_used || !first_echo || !last_echo || n_histograms <= 0 || echoes == 0)
        return RWNMR_ERR_INVALID;
    int32_t n = n_histograms;
    if (echoes < (uint32_t)n)
        n = 1;
    const uint32_t per = echoes / (uint32_t)n;
    for (int32_t i = 0; i < n; ++i) {
        first_echo[i] = (uint32_t)i * per;
        last_echo[i] = first_echo[i] + per;
    }
    last_echo[n - 1] += echoes - (uint32_t)n * per;
    *n_used = n;
    return RWNMR_OK;
}

rwnmr_status rwnmr_collision_histogram(const uint32_t *collisions, uint64_t n, uint64_t steps, int32_t size,
                                       int32_t log_scale, double *bins, double *amps)
{
    if (!collisions || !bins || !amps || n == 0 || steps == 0 || size <= 0 || (log_scale && size < 2))
        return RWNMR_ERR_INVALID;
    const double st = (double)steps;
    const double walkers = (double)n;
    if (!log_scale) {
        const double gap = 1.0 / (double)size;
        for (int32_t i = 0; i < size; ++i) {
            bins[i] = i * gap + 0.5 * gap;
            amps[i] = 0.0;
        }
        for (uint64_t w = 0; w < n; ++w) {
            const double xi = collisions[w] / st;
            int32_t idx = (int32_t)floor(xi / gap);
            if (idx >= size)
                idx--;
            amps[idx] += 1.0;
        }
        for (int32_t i = 0; i < size; ++i)
            amps[i] = amps[i] / walkers;
        return RWNMR_OK;
    }
    const int32_t nb = size + 1;
    // bins: control entry, then size log-spaced rates from 10^round(log10 min rate) to 1
    uint64_t first = 0;
    while (first < n && collisions[first] == 0)
        ++first;
    if (first < n) {
        double min_rate = collisions[first] / st;
        for (uint64_t w = first; w < n; ++w)
            if (collisions[w] != 0 && collisions[w] / st < min_rate)
                min_rate = collisions[w] / st;
        const double e0 = round(log10(min_rate));
        const double step = (0.0 - e0) / ((double)size - 1.0);
        bins[0] = e0 - 1.0;
        for (int32_t i = 1; i < nb; ++i)
            bins[i] = pow(10.0, e0 + step * (i - 1));
    } else {
        for (int32_t i = 0; i < nb; ++i)
            bins[i] = 0.0;
    }
    for (int32_t i = 0; i < nb; ++i)
        amps[i] = 0.0;
    const double log_gap = log10(bins[2]) - log10(bins[1]);
    const double min_val = log10(bins[1]);
    for (uint64_t w = 0; w < n; ++w) {
        if (collisions[w] == 0) {
            amps[0] += 1.0;
        } else {
            const double xi = collisions[w] / st;
            int32_t idx = (int32_t)floor(round((log10(xi) - min_val) / log_gap)) + 1;
            if (idx > 0 && idx < nb)
                amps[idx] += 1.0;
            else
                amps[1] += 1.0;
        }
    }
    for (int32_t i = 0; i < nb; ++i)
        amps[i] = amps[i] / walkers;
    return RWNMR_OK;
}

rwnmr_status rwnmr_histogram_penalties(const double *bins, int32_t n_bins, double rho, const double *sigmoid,
                                       double step_length, double D0, uint32_t steps_per_echo, double *penalties)
{
    if (!bins || !penalties || n_bins <= 0 || D0 <= 0.0)
        return RWNMR_ERR_INVALID;
    for (int32_t i = 0; i < n_bins; ++i) {
        const double xi = bins[i];
        double r = rho;
        if (sigmoid) {
            const double *p = sigmoid;
            r = (p[1] + (p[0] - p[1]) / (1 + exp((-1) * p[3] * (xi - p[2])))) +
                (p[5] + (p[4] - p[5]) / (1 + exp((-1) * p[7] * (xi - p[6]))));
        }
        const double delta = 1.0 - (2.0 / 3.0) * ((r * step_length) / (D0 * 1e03));
        penalties[i] = pow(delta, xi * (double)steps_per_echo);
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_histogram_decay(const double *amps, int32_t n_hist, int32_t n_bins, const uint32_t *first_echo,
                                   const uint32_t *last_echo, const double *penalties, double *signal)
{
    if (!amps || !first_echo || !last_echo || !penalties || !signal || n_hist <= 0 || n_bins <= 0)
        return RWNMR_ERR_INVALID;
    std::vector<double> dist(n_bins);
    uint64_t k = 0;
    signal[k++] = 1.0;
    for (int32_t h = 0; h < n_hist; ++h) {
        const double level = signal[k - 1];
        for (int32_t i = 0; i < n_bins; ++i)
            dist[i] = level * amps[(size_t)h * n_bins + i];
        for (uint32_t e = first_echo[h]; e < last_echo[h]; ++e) {
            for (int32_t i = 0; i < n_bins; ++i)
                dist[i] *= penalties[i];
            double s = 0.0;
            for (int32_t i = 0; i < n_bins; ++i)
                s += dist[i];
            signal[k++] = s;
        }
    }
    return RWNMR_OK;
}

}  // extern "C"
