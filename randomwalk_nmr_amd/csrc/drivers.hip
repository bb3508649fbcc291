// drivers.hip -- host drivers of the CPMG and PFGSE sequences over the C
// ABI: time framework, relaxivity, the GPU run, bulk relaxation,
// normalisation, wave vectors, Stejskal-Tanner and msd diffusion
// coefficients.  Reference paths under src/backend/ of RandomWalk_NMR.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/rwnmr/rwnmr.h"
#include "../../include/rwnmr/rwnmr_drivers.h"
#include "../../include/rwnmr/rwnmr_host.h"

namespace {

// Two-sigmoid relaxivity of a walker's collision rate xi
// (Walker::setSurfaceRelaxivity, B/Walker/walker.h:174-196).
double sigmoid_rho(const double *p, double xi)
{
    const double s1 = p[1] + (p[0] - p[1]) / (1.0 + exp(-p[3] * (xi - p[2])));
    const double s2 = p[5] + (p[4] - p[5]) / (1.0 + exp(-p[7] * (xi - p[6])));
    return s1 + s2;
}

// Per-walker deltas from a sigmoid relaxivity: map the walkers over `steps`
// steps (NMR_Simulation::mapSimulation), xi = hits / steps
// (Walker::updateXIrate, walker.h:162-172), then delta
// (updateWalkersRelaxativity, NMR_Simulation.cpp:1500-1509).  The walkers
// themselves are not moved (the map runs on copies).
rwnmr_status sigmoid_penalties(rwnmr_ctx *ctx, const rwnmr_image *img, const rwnmr_walkers *w, uint64_t steps,
                               int32_t bc, uint32_t shift, const double *sig, double voxel, double D0,
                               std::vector<double> &pen)
{
    const uint64_t n = w->n;
    std::vector<int32_t> x(n), y(n), z(n);
    std::vector<uint64_t> s(n);
    std::vector<uint32_t> h(n, 0u);
    const hipMemcpyKind k = w->on_device ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
    if (hipMemcpy(x.data(), w->x, n * 4, k) != hipSuccess || hipMemcpy(y.data(), w->y, n * 4, k) != hipSuccess ||
        hipMemcpy(z.data(), w->z, n * 4, k) != hipSuccess || hipMemcpy(s.data(), w->seed, n * 8, k) != hipSuccess)
        return RWNMR_ERR_HIP;
    rwnmr_walkers m{};
    m.x = x.data();
    m.y = y.data();
    m.z = z.data();
    m.seed = s.data();
    m.hits = h.data();
    m.n = n;
    m.on_device = 0;
    rwnmr_status st = rwnmr_map_run(ctx, img, &m, steps, bc, shift);
    if (st != RWNMR_OK)
        return st;
    pen.resize(n);
    for (uint64_t i = 0; i < n; ++i) {
        const double xi = (double)h[i] / (double)steps;
        pen[i] = rwnmr_penalty(sigmoid_rho(sig, xi), voxel, D0);
    }
    return RWNMR_OK;
}

}  // namespace

extern "C" {

rwnmr_status rwnmr_cpmg_simulate(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w,
                                 const rwnmr_cpmg_params *p, uint32_t capacity, double *times, double *signal,
                                 uint32_t *n_points)
{
    if (!ctx || !img || !w || !p || !times || !signal || !n_points)
        return RWNMR_ERR_INVALID;
    uint32_t echoes = 0;
    uint64_t steps = 0;
    rwnmr_status st = rwnmr_time_framework(p->observation_time, p->voxel_size, p->D0, p->steps_per_echo, &echoes,
                                           &steps);
    if (st != RWNMR_OK)
        return st;
    if (capacity < echoes + 1)
        return RWNMR_ERR_INVALID;
    rwnmr_walkers run = *w;
    std::vector<double> pen;
    if (p->sigmoid) {
        st = sigmoid_penalties(ctx, img, w, steps, p->bc, p->shift, p->sigmoid, p->voxel_size, p->D0, pen);
        if (st != RWNMR_OK)
            return st;
        if (w->on_device)
            return RWNMR_ERR_INVALID;  // per-walker sigmoid deltas: host walkers only
        run.penalty = pen.data();
    } else if (!w->penalty) {
        run.delta_uniform = rwnmr_penalty(p->rho, p->voxel_size, p->D0);
    }
    // signal_times: t accumulates += steps_per_echo * dt (NMR_cpmg.cpp:112-130)
    const double dt = rwnmr_time_step(p->voxel_size, p->D0);
    const double step_t = (double)p->steps_per_echo * dt;
    double t = 0.0;
    times[0] = 0.0;
    for (uint32_t e = 1; e <= echoes; ++e) {
        t += step_t;
        times[e] = t;
    }
    if (w->on_device) {
        double *dsums = nullptr;
        if (hipMalloc((void **)&dsums, 8ull * (echoes + 1)) != hipSuccess)
            return RWNMR_ERR_ALLOC;
        st = rwnmr_cpmg_run(ctx, img, &run, echoes, p->steps_per_echo, p->bc, p->shift, dsums);
        if (st == RWNMR_OK)
            st = rwnmr_ctx_synchronize(ctx);
        if (st == RWNMR_OK && hipMemcpy(signal, dsums, 8ull * (echoes + 1), hipMemcpyDeviceToHost) != hipSuccess)
            st = RWNMR_ERR_HIP;
        (void)hipFree(dsums);
    } else {
        st = rwnmr_cpmg_run(ctx, img, &run, echoes, p->steps_per_echo, p->bc, p->shift, signal);
    }
    if (st != RWNMR_OK)
        return st;
    // applyBulk (NMR_cpmg.cpp:346-359) and normalizeSignal (:369-385)
    if (p->apply_bulk) {
        const double bulk = -1.0 / p->bulk_time;
        for (uint32_t e = 0; e <= echoes; ++e)
            signal[e] = exp(bulk * times[e]) * signal[e];
    }
    const double norm = 1.0 / signal[0];
    for (uint32_t e = 0; e <= echoes; ++e)
        signal[e] = norm * signal[e];
    *n_points = echoes + 1;
    return RWNMR_OK;
}

void rwnmr_pfgse_gradients(double gmax_x, double gmax_y, double gmax_z, uint32_t points, double *gx, double *gy,
                           double *gz)
{
    // setGradientVector (NMR_pfgse.cpp:272-297): 0 .. Gmax, value += gap
    const double ax = gmax_x / (double)(points - 1), ay = gmax_y / (double)(points - 1),
                 az = gmax_z / (double)(points - 1);
    double vx = 0, vy = 0, vz = 0;
    for (uint32_t i = 0; i < points; ++i) {
        gx[i] = vx;
        gy[i] = vy;
        gz[i] = vz;
        vx += ax;
        vy += ay;
        vz += az;
    }
}

double rwnmr_pfgse_wave_vector(double gradient, double pulse_width, double gamma)
{
    // computeWaveVectorK (NMR_pfgse.cpp:483-486)
    return (pulse_width * 1.0e-03) * (gamma * 1.0e+06) * (gradient * 1.0e-08);
}

int32_t rwnmr_stejskal_tanner(const double *rhs, const double *lhs, uint32_t n, double *D, double *D_err)
{
    // no-intercept least squares (LeastSquareAdjust.cpp:77-170) and
    // err = sqrt(MSE n / (n - 1)) (NMR_pfgse.cpp:678-682)
    if (n < 2 || !rhs || !lhs || !D || !D_err)
        return RWNMR_ERR_INVALID;
    double num = 0, den = 0;
    for (uint32_t i = 0; i < n; ++i) {
        num += rhs[i] * lhs[i];
        den += rhs[i] * rhs[i];
    }
    const double B = num / den;
    double res = 0;
    for (uint32_t i = 0; i < n; ++i)
        res += pow(lhs[i] - B * rhs[i], 2.0);
    res /= (double)n;
    *D = B;
    *D_err = sqrt(res * ((double)n / ((double)n - 1.0)));
    return RWNMR_OK;
}

rwnmr_status rwnmr_pfgse_sample_stats(const double *M_samples, uint32_t n_samples, uint32_t n_k, const double *rhs,
                                      uint32_t n_points, double *mean_M, double *std_M, double *mean_lhs,
                                      double *std_lhs, double *D_samples, double *D_mean, double *D_std,
                                      double *D_err_mean)
{
    // recoverDsatWithSampling (NMR_pfgse.cpp:876-1060) after its optional bulk
    // and noise terms: per-sample normalisation by M(k=0), LHS, mean and
    // population deviation over samples, one no-intercept fit per sample
    if (!M_samples || !rhs || n_samples == 0 || n_k == 0 || n_points < 2 || n_points > n_k)
        return RWNMR_ERR_INVALID;
    const uint32_t S = n_samples;
    std::vector<double> lhs((size_t)S * n_k), col(S), Ds(S), Derr(S);
    auto mean = [](const double *v, uint32_t n) {
        double sum = 0;
        for (uint32_t i = 0; i < n; ++i)
            sum += v[i];
        return sum / (double)n;
    };
    auto stdev = [](const double *v, uint32_t n, double m) {
        double sum = 0.0;
        for (uint32_t i = 0; i < n; ++i)
            sum += (v[i] - m) * (v[i] - m);
        return sqrt(sum / (double)n);
    };
    for (uint32_t s = 0; s < S; ++s) {
        const double *Ms = M_samples + (size_t)s * n_k;
        const double n0 = Ms[0] / Ms[0];
        for (uint32_t k = 0; k < n_k; ++k) {
            const double ratio = (Ms[k] / Ms[0]) / n0;
            lhs[(size_t)s * n_k + k] = ratio > 1.0e-06 ? log(ratio) : log(1.0e-06);
        }
    }
    for (uint32_t k = 0; k < n_k; ++k) {
        for (uint32_t s = 0; s < S; ++s)
            col[s] = M_samples[(size_t)s * n_k + k];
        const double m = mean(col.data(), S);
        if (mean_M)
            mean_M[k] = m;
        if (std_M)
            std_M[k] = stdev(col.data(), S, m);
        for (uint32_t s = 0; s < S; ++s)
            col[s] = lhs[(size_t)s * n_k + k];
        const double ml = mean(col.data(), S);
        if (mean_lhs)
            mean_lhs[k] = ml;
        if (std_lhs)
            std_lhs[k] = stdev(col.data(), S, ml);
    }
    for (uint32_t s = 0; s < S; ++s) {
        rwnmr_status st = rwnmr_stejskal_tanner(rhs, lhs.data() + (size_t)s * n_k, n_points, &Ds[s], &Derr[s]);
        if (st != RWNMR_OK)
            return st;
    }
    const double dm = mean(Ds.data(), S);
    if (D_samples)
        std::copy(Ds.begin(), Ds.end(), D_samples);
    if (D_mean)
        *D_mean = dm;
    if (D_std)
        *D_std = stdev(Ds.data(), S, dm);
    if (D_err_mean)
        *D_err_mean = mean(Derr.data(), S);
    return RWNMR_OK;
}

uint32_t rwnmr_pfgse_threshold_lhs(const double *lhs, uint32_t n, double value, uint32_t window)
{
    // setThresholdFromLHSValue (NMR_pfgse.cpp:399-419): the first point
    // after `window` whose LHS falls below log(value)
    if (n < window || window == 0)
        return n;
    int idx = (int)window - 1;
    const double lv = log(value);
    bool greater = true;
    while (idx < (int)n && greater) {
        if (lhs[idx] < lv)
            greater = false;
        else
            ++idx;
    }
    if (greater)
        --idx;
    return (uint32_t)idx;
}

uint32_t rwnmr_pfgse_threshold_lhs_window(const double *lhs, uint32_t n, double value, uint32_t window)
{
    // setThresholdFromLHSWindow (NMR_pfgse.cpp:405-439): a moving mean of
    // `window` LHS values, the first point where it falls below log(value)
    if (n < window || window == 0)
        return n;
    std::vector<double> win(lhs, lhs + window);
    auto wmean = [&]() {
        double s = 0.0;
        for (double v : win)
            s += v;
        return s / (double)win.size();
    };
    const double lv = log(value);
    int idx = (int)window;
    bool greater = !(wmean() < lv);
    while (idx < (int)n && greater) {
        win[(uint32_t)idx % window] = lhs[idx];
        if (wmean() < lv)
            greater = false;
        else
            ++idx;
    }
    if (greater)
        --idx;
    return (uint32_t)idx;
}

rwnmr_status rwnmr_pfgse_simulate(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w,
                                  const rwnmr_pfgse_params *p, const double *exposure_times, uint32_t n_times,
                                  const double *kx, const double *ky, const double *kz, uint32_t n_k,
                                  rwnmr_pfgse_result *out)
{
    if (!ctx || !img || !w || !p || !exposure_times || !kx || !ky || !kz || !out || n_k < 2 || n_times == 0)
        return RWNMR_ERR_INVALID;
    if (w->on_device)
        return RWNMR_ERR_INVALID;  // host walkers (the driver post-processes on the host)
    const uint64_t n = w->n;
    // steps per diffusion time: setTimeFramework(Delta) (NMR_Simulation.cpp:293-298),
    // advanced incrementally (NMR_pfgse_cuda.cpp:563-565)
    std::vector<uint32_t> inc(n_times);
    uint64_t taken = 0;
    for (uint32_t i = 0; i < n_times; ++i) {
        uint32_t echoes;
        uint64_t steps;
        rwnmr_status st = rwnmr_time_framework(exposure_times[i], p->voxel_size, p->D0, p->steps_per_echo, &echoes,
                                               &steps);
        if (st != RWNMR_OK)
            return st;
        if (steps < taken)
            return RWNMR_ERR_INVALID;  // exposure times must be ascending
        inc[i] = (uint32_t)(steps - taken);
        taken = steps;
        out->steps[i] = steps;
    }
    std::vector<int32_t> x0(w->x, w->x + n), y0(w->y, w->y + n), z0(w->z, w->z + n);
    rwnmr_walkers run = *w;
    std::vector<double> pen;
    if (p->sigmoid) {
        // initial map over the longest time (runInitialMapSimulation, NMR_pfgse.cpp:322-348)
        rwnmr_status st = sigmoid_penalties(ctx, img, w, taken, p->bc, 0, p->sigmoid, p->voxel_size, p->D0, pen);
        if (st != RWNMR_OK)
            return st;
        run.penalty = pen.data();
    } else if (!w->penalty) {
        run.delta_uniform = rwnmr_penalty(p->rho, p->voxel_size, p->D0);
    }
    const uint32_t ns = p->n_rho_sets;
    std::vector<double> sets(ns);
    for (uint32_t q = 0; q < ns; ++q)
        sets[q] = rwnmr_penalty(p->rho_sets[q], p->voxel_size, p->D0);
    const uint32_t nsets = ns ? ns : 1;
    rwnmr_status st = rwnmr_pfgse_run(ctx, img, &run, x0.data(), y0.data(), z0.data(), inc.data(), n_times, kx, ky,
                                      kz, n_k, p->k_per_delta, p->voxel_size, p->bc, 0, ns, ns ? sets.data() : nullptr,
                                      out->Mkt, out->msd);
    if (st != RWNMR_OK)
        return st;
    for (uint32_t q = 0; q < nsets; ++q)
        for (uint32_t i = 0; i < n_times; ++i) {
            const double delta_t = exposure_times[i];
            double *M = out->Mkt + ((size_t)q * n_times + i) * n_k;
            // applyBulk (NMR_pfgse.cpp:519-529)
            if (p->apply_bulk) {
                const double bm = exp(-1.0 / p->bulk_time * delta_t);
                for (uint32_t k = 0; k < n_k; ++k)
                    M[k] *= bm;
            }
            double *rhs = out->rhs + ((size_t)q * n_times + i) * n_k;
            double *lhs = out->lhs + ((size_t)q * n_times + i) * n_k;
            const double *kxi = kx + (p->k_per_delta ? (size_t)i * n_k : 0);
            const double *kyi = ky + (p->k_per_delta ? (size_t)i * n_k : 0);
            const double *kzi = kz + (p->k_per_delta ? (size_t)i * n_k : 0);
            // RHS = -|k|^2 (Delta - pulse/3) (computeRHS :450-453), LHS =
            // log(max(M_k/M_0, 1e-6)) of the normalised signal (computeLHS
            // :474-481, recoverDsatWithoutSampling :641-692)
            const double M0 = M[0];
            for (uint32_t k = 0; k < n_k; ++k) {
                const double kn = sqrt(kxi[k] * kxi[k] + kyi[k] * kyi[k] + kzi[k] * kzi[k]);
                rhs[k] = -1.0 * kn * kn * (delta_t - (p->pulse_width / 3.0));
                const double ratio = (M[k] / M0) / (M[0] / M0);
                lhs[k] = ratio > 1.0e-06 ? log(ratio) : log(1.0e-06);
            }
            uint32_t used = n_k;
            // the window is first clamped to the gradient points (applyThreshold, NMR_pfgse.cpp:139)
            if (p->threshold_lhs > 0.0 && p->threshold_lhs < 1.0)
                used = rwnmr_pfgse_threshold_lhs(lhs, n_k, p->threshold_lhs, std::min(p->threshold_window, n_k));
            else if (p->threshold_samples > 1 && p->threshold_samples <= n_k)
                used = p->threshold_samples;
            out->points[(size_t)q * n_times + i] = used;
            st = rwnmr_stejskal_tanner(rhs, lhs, used, &out->D_sat[(size_t)q * n_times + i],
                                       &out->D_sat_err[(size_t)q * n_times + i]);
            if (st != RWNMR_OK)
                return st;
        }
    // D_msd = eps^2 sum |dr|^2 / N / (6 Delta) (recoverDmsdWithoutSampling :1098-1160)
    for (uint32_t i = 0; i < n_times; ++i) {
        const double v2 = p->voxel_size * p->voxel_size;
        const double s2 = (out->msd[3 * i] + out->msd[3 * i + 1] + out->msd[3 * i + 2]) * v2 / (double)n;
        out->D_msd[i] = s2 / (6.0 * exposure_times[i]);
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_cpmg_histogram_simulate(rwnmr_ctx *ctx, const rwnmr_image *img, rwnmr_walkers *w, uint32_t echoes,
                                           uint32_t steps_per_echo, int32_t bc, uint32_t shift, int32_t n_histograms,
                                           int32_t size, int32_t log_scale, double rho, const double *sigmoid,
                                           double step_length, double D0, double *signal, double *bins_out,
                                           double *amps_out, int32_t *n_used)
{
    if (!ctx || !img || !w || !signal || echoes == 0 || steps_per_echo == 0 || n_histograms <= 0 || size <= 0 ||
        w->n == 0)
        return RWNMR_ERR_INVALID;
    const uint64_t n = w->n;
    std::vector<uint32_t> first((size_t)n_histograms), last((size_t)n_histograms);
    int32_t nh = 0;
    rwnmr_status st = rwnmr_histogram_list(echoes, n_histograms, &nh, first.data(), last.data());
    if (st != RWNMR_OK)
        return st;
    const int32_t nb = rwnmr_histogram_bins(size, log_scale);
    std::vector<double> amps((size_t)nh * nb), seg_bins(nb), bins(nb), pen(nb);
    std::vector<uint32_t> hits(n, 0u), total(n, 0u);
    // the map walk counts into a hits array of the walkers' memory kind
    uint32_t *dev_hits = nullptr;
    if (w->on_device && hipMalloc((void **)&dev_hits, n * 4) != hipSuccess)
        return RWNMR_ERR_ALLOC;
    rwnmr_walkers m = *w;
    m.energy = nullptr;
    m.hits = w->on_device ? dev_hits : hits.data();
    for (int32_t h = 0; h < nh && st == RWNMR_OK; ++h) {
        const uint64_t steps = (uint64_t)steps_per_echo * (last[h] - first[h]);
        if (w->on_device) {
            if (hipMemset(dev_hits, 0, n * 4) != hipSuccess)
                st = RWNMR_ERR_HIP;
        } else {
            std::fill(hits.begin(), hits.end(), 0u);
        }
        if (st == RWNMR_OK)
            st = rwnmr_map_run(ctx, img, &m, steps, bc, shift);
        if (st == RWNMR_OK && w->on_device) {
            if (rwnmr_ctx_synchronize(ctx) != RWNMR_OK || hipMemcpy(hits.data(), dev_hits, n * 4,
                                                                    hipMemcpyDeviceToHost) != hipSuccess)
                st = RWNMR_ERR_HIP;
        }
        if (st == RWNMR_OK)
            st = rwnmr_collision_histogram(hits.data(), n, steps, size, log_scale, seg_bins.data(),
                                           amps.data() + (size_t)h * nb);
        for (uint64_t i = 0; i < n; ++i)
            total[i] += hits[i];
    }
    if (dev_hits)
        (void)hipFree(dev_hits);
    if (st != RWNMR_OK)
        return st;
    std::vector<double> gamps(nb);
    if ((st = rwnmr_collision_histogram(total.data(), n, (uint64_t)last[nh - 1] * steps_per_echo, size, log_scale,
                                        bins.data(), gamps.data())) != RWNMR_OK)
        return st;
    if ((st = rwnmr_histogram_penalties(bins.data(), nb, rho, sigmoid, step_length, D0, steps_per_echo,
                                        pen.data())) != RWNMR_OK)
        return st;
    if ((st = rwnmr_histogram_decay(amps.data(), nh, nb, first.data(), last.data(), pen.data(), signal)) !=
        RWNMR_OK)
        return st;
    if (bins_out)
        std::copy(bins.begin(), bins.end(), bins_out);
    if (amps_out)
        std::copy(amps.begin(), amps.end(), amps_out);
    if (n_used)
        *n_used = nh;
    return RWNMR_OK;
}

rwnmr_status rwnmr_multitau_sequence(const rwnmr_multitau_params *p, double time_interval, uint32_t *steps,
                                     double *echo_times)
{
    if (!p || !steps || !echo_times || p->tau_points == 0 || time_interval <= 0.0 ||
        (p->tau_scale == 2 && !p->tau_values) || (p->tau_scale != 2 && p->tau_points < 2))
        return RWNMR_ERR_INVALID;
    const uint32_t n = p->tau_points;
    std::vector<double> tau(n);
    if (p->tau_scale == 2) {
        std::copy(p->tau_values, p->tau_values + n, tau.begin());
    } else if (p->tau_scale == 1) {
        const double a = log10(p->tau_min), b = log10(p->tau_max);
        const double step = (b - a) / ((double)n - 1.0);
        for (uint32_t i = 0; i < n; ++i)
            tau[i] = pow(10.0, a + step * i);
    } else {
        const double step = (p->tau_max - p->tau_min) / ((double)n - 1.0);
        for (uint32_t i = 0; i < n; ++i)
            tau[i] = p->tau_min + step * i;
    }
    uint32_t min_steps = 0;
    for (uint32_t i = 0; i < n; ++i) {
        int s = (int)std::ceil(tau[i] / time_interval);
        if (s % 2 != 0)
            s++;
        if ((uint32_t)s <= min_steps)
            s = (int)min_steps + 2;
        if (s <= 0)
            return RWNMR_ERR_INVALID;
        steps[i] = (uint32_t)s;
        min_steps = (uint32_t)s;
        echo_times[i] = s * time_interval;
    }
    return RWNMR_OK;
}

rwnmr_status rwnmr_multitau_simulate(rwnmr_ctx *ctx, const rwnmr_image *img, const rwnmr_walkers *w0,
                                     const rwnmr_multitau_params *mp, const rwnmr_cpmg_params *cp,
                                     double *echo_times, double *signal)
{
    if (!ctx || !img || !w0 || !mp || !cp || !echo_times || !signal || w0->n == 0)
        return RWNMR_ERR_INVALID;
    const uint32_t n_tau = mp->tau_points;
    std::vector<uint32_t> steps(n_tau);
    const double dt = rwnmr_time_step(cp->voxel_size, cp->D0);
    rwnmr_status st = rwnmr_multitau_sequence(mp, dt, steps.data(), echo_times);
    if (st != RWNMR_OK)
        return st;
    const uint64_t n = w0->n;
    // every CPMG starts from the initial walker state (NMR_cpmg resets the walkers)
    std::vector<int32_t> hx, hy, hz;
    std::vector<uint64_t> hs;
    int32_t *dx = nullptr, *dy = nullptr, *dz = nullptr;
    uint64_t *ds = nullptr;
    if (w0->on_device) {
        if (hipMalloc((void **)&dx, 4 * n) != hipSuccess || hipMalloc((void **)&dy, 4 * n) != hipSuccess ||
            hipMalloc((void **)&dz, 4 * n) != hipSuccess || hipMalloc((void **)&ds, 8 * n) != hipSuccess) {
            (void)hipFree(dx), (void)hipFree(dy), (void)hipFree(dz), (void)hipFree(ds);
            return RWNMR_ERR_ALLOC;
        }
    }
    for (uint32_t i = 0; i < n_tau && st == RWNMR_OK; ++i) {
        rwnmr_cpmg_params p = *cp;
        p.steps_per_echo = steps[i];
        p.observation_time = mp->complete_decay ? cp->observation_time : echo_times[i];
        rwnmr_walkers w = *w0;
        w.energy = nullptr;
        w.hits = nullptr;
        if (w0->on_device) {
            if (hipMemcpy(dx, w0->x, 4 * n, hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(dy, w0->y, 4 * n, hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(dz, w0->z, 4 * n, hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(ds, w0->seed, 8 * n, hipMemcpyDeviceToDevice) != hipSuccess) {
                st = RWNMR_ERR_HIP;
                break;
            }
            w.x = dx, w.y = dy, w.z = dz, w.seed = ds;
        } else {
            hx.assign(w0->x, w0->x + n);
            hy.assign(w0->y, w0->y + n);
            hz.assign(w0->z, w0->z + n);
            hs.assign(w0->seed, w0->seed + n);
            w.x = hx.data(), w.y = hy.data(), w.z = hz.data(), w.seed = hs.data();
        }
        uint32_t echoes = 0;
        uint64_t total = 0;
        if ((st = rwnmr_time_framework(p.observation_time, p.voxel_size, p.D0, p.steps_per_echo, &echoes, &total)) !=
            RWNMR_OK)
            break;
        std::vector<double> times(echoes + 1), sig(echoes + 1);
        uint32_t npts = 0;
        st = rwnmr_cpmg_simulate(ctx, img, &w, &p, echoes + 1, times.data(), sig.data(), &npts);
        if (st == RWNMR_OK)
            signal[i] = npts > 1 ? sig[1] : sig[0];
    }
    if (w0->on_device) {
        (void)hipFree(dx), (void)hipFree(dy), (void)hipFree(dz), (void)hipFree(ds);
    }
    return st;
}

}  // extern "C"
