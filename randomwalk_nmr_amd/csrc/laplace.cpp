// laplace.cpp -- FP64 Laplace (T2) inversion with the reference's C API
// (B/Laplace/include/nmrinv.h:13-39, B/Laplace/src/nmrinv.cpp) over this
// build's own numerics:
//   * thin SVD by one-sided (Hestenes) Jacobi rotations, replacing Eigen's
//     JacobiSVD in the L-curve (B/Laplace/src/TikhonovSolver.cpp:14-80);
//   * Hansen's L-curve curvature over log-spaced lambdas
//     (TikhonovSolver.cpp:82-116);
//   * Lawson & Hanson (1974, "Solving Least Squares Problems", ch. 23)
//     active-set NNLS on [K; lambda I] x = [b; 0], each passive-set
//     subproblem solved by Householder QR (the reference vendors an Eigen
//     NNLS module, B/Laplace/include/NNLS.hpp, whose code is not used);
//   * kernel matrix, T2 bins, log pruning and noise exactly as
//     B/Laplace/src/nmrinv_core.cpp:99-480 and vector_funcs.cpp:85-107.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <random>
#include <tuple>
#include <vector>

#include "rwnmr/rwnmr.h"

namespace {

using Vec = std::vector<double>;

// value += step accumulation (vector_funcs.cpp:85-100)
Vec linspace_acc(double start, double end, int size, bool include_final = true)
{
    const double step = include_final ? (end - start) / (size - 1) : (end - start) / size;
    Vec out;
    out.reserve(size);
    double v = start;
    for (int i = 0; i < size; ++i) {
        out.push_back(v);
        v += step;
    }
    return out;
}

Vec logspace_acc(double start, double end, int size)
{
    Vec v = linspace_acc(start, end, size);
    for (double &x : v)
        x = pow(10.0, x);
    return v;
}

// Row-major dense matrix.
struct Mat {
    int rows = 0, cols = 0;
    Vec a;
    Mat() = default;
    Mat(int r, int c) : rows(r), cols(c), a((size_t)r * c, 0.0) {}
    double &operator()(int i, int j) { return a[(size_t)i * cols + j]; }
    double operator()(int i, int j) const { return a[(size_t)i * cols + j]; }
};

Mat kernel_matrix(const Vec &t, const Vec &t2)
{
    Mat K((int)t.size(), (int)t2.size());
    for (int i = 0; i < K.rows; ++i)
        for (int j = 0; j < K.cols; ++j)
            K(i, j) = exp(-t[i] / t2[j]);
    return K;
}

// Thin SVD by one-sided Jacobi on the columns of A (m x n, m >= n):
// rotate column pairs until mutually orthogonal; singular values are the
// column norms (sorted descending) and U's columns the normalised columns.
void jacobi_svd(const Mat &A, Mat &U, Vec &s)
{
    const int m = A.rows, n = A.cols;
    // work column-major for cache-friendly column operations
    std::vector<Vec> col(n, Vec(m));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i)
            col[j][i] = A(i, j);
    const double tol = 1e-15;
    for (int sweep = 0; sweep < 80; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                const double *cp = col[p].data(), *cq = col[q].data();
                for (int i = 0; i < m; ++i) {
                    alpha += cp[i] * cp[i];
                    beta += cq[i] * cq[i];
                    gamma += cp[i] * cq[i];
                }
                if (gamma == 0.0)
                    continue;
                const double c0 = fabs(gamma) / sqrt(alpha * beta);
                off = std::max(off, c0);
                if (c0 <= tol)
                    continue;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                double *wp = col[p].data(), *wq = col[q].data();
                for (int i = 0; i < m; ++i) {
                    const double xp = wp[i], xq = wq[i];
                    wp[i] = c * xp - sn * xq;
                    wq[i] = sn * xp + c * xq;
                }
            }
        if (off <= tol)
            break;
    }
    std::vector<std::pair<double, int>> norms(n);
    for (int j = 0; j < n; ++j) {
        double ss = 0;
        for (int i = 0; i < m; ++i)
            ss += col[j][i] * col[j][i];
        norms[j] = {sqrt(ss), j};
    }
    std::stable_sort(norms.begin(), norms.end(), [](auto &a, auto &b) { return a.first > b.first; });
    U = Mat(m, n);
    s.assign(n, 0.0);
    for (int k = 0; k < n; ++k) {
        s[k] = norms[k].first;
        const int j = norms[k].second;
        if (s[k] > 0)
            for (int i = 0; i < m; ++i)
                U(i, k) = col[j][i] / s[k];
    }
}

// Least squares min |A_P z - b| over the passive columns P by Householder
// QR of the (m x |P|) submatrix.  Returns false if it is rank deficient.
bool passive_ls(const Mat &A, const Vec &b, const std::vector<int> &P, Vec &z)
{
    const int m = A.rows, k = (int)P.size();
    std::vector<Vec> Q(k, Vec(m));
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < m; ++i)
            Q[j][i] = A(i, P[j]);
    Vec rhs = b;
    Vec diag(k);
    for (int j = 0; j < k; ++j) {
        double norm = 0;
        for (int i = j; i < m; ++i)
            norm += Q[j][i] * Q[j][i];
        norm = sqrt(norm);
        if (norm == 0.0)
            return false;
        const double alpha = Q[j][j] > 0 ? -norm : norm;
        Vec v(m - j);
        for (int i = j; i < m; ++i)
            v[i - j] = Q[j][i];
        v[0] -= alpha;
        double vn = 0;
        for (double x : v)
            vn += x * x;
        if (vn == 0.0) {
            diag[j] = alpha;
            continue;
        }
        for (int c = j; c < k; ++c) {
            double d = 0;
            for (int i = j; i < m; ++i)
                d += v[i - j] * Q[c][i];
            d = 2.0 * d / vn;
            for (int i = j; i < m; ++i)
                Q[c][i] -= d * v[i - j];
        }
        double d = 0;
        for (int i = j; i < m; ++i)
            d += v[i - j] * rhs[i];
        d = 2.0 * d / vn;
        for (int i = j; i < m; ++i)
            rhs[i] -= d * v[i - j];
        diag[j] = Q[j][j];
    }
    z.assign(k, 0.0);
    for (int j = k - 1; j >= 0; --j) {
        double sacc = rhs[j];
        for (int c = j + 1; c < k; ++c)
            sacc -= Q[c][j] * z[c];
        if (diag[j] == 0.0)
            return false;
        z[j] = sacc / diag[j];
    }
    return true;
}

// Lawson-Hanson NNLS: min |A x - b| subject to x >= 0.
bool nnls(const Mat &A, const Vec &b, Vec &x, double tol, int max_iter)
{
    const int m = A.rows, n = A.cols;
    x.assign(n, 0.0);
    std::vector<char> passive(n, 0);
    Vec w(n), r(m), z;
    auto gradient = [&]() {
        for (int i = 0; i < m; ++i) {
            double acc = b[i];
            for (int j = 0; j < n; ++j)
                acc -= A(i, j) * x[j];
            r[i] = acc;
        }
        for (int j = 0; j < n; ++j) {
            double acc = 0;
            for (int i = 0; i < m; ++i)
                acc += A(i, j) * r[i];
            w[j] = acc;
        }
    };
    gradient();
    std::vector<char> blocked(n, 0);  // candidates whose entry was not positive
    int iter = 0;
    while (iter++ < max_iter) {
        // step 2-5: the most positive gradient component of the zero set
        int jmax = -1;
        double wmax = tol;
        for (int j = 0; j < n; ++j)
            if (!passive[j] && !blocked[j] && w[j] > wmax) {
                wmax = w[j];
                jmax = j;
            }
        if (jmax < 0)
            return true;  // Kuhn-Tucker conditions hold
        passive[jmax] = 1;
        bool first = true;
        for (;;) {
            std::vector<int> P;
            for (int j = 0; j < n; ++j)
                if (passive[j])
                    P.push_back(j);
            if (P.empty())
                break;
            if (!passive_ls(A, b, P, z))
                return false;
            if (first) {
                first = false;
                // rounding can make the entering variable non-positive: keep
                // it out and try the next candidate (Lawson-Hanson step 6)
                const size_t qj = std::find(P.begin(), P.end(), jmax) - P.begin();
                if (z[qj] <= 0.0) {
                    passive[jmax] = 0;
                    blocked[jmax] = 1;
                    break;
                }
            }
            size_t qmin = P.size();
            double alpha = 2.0;
            for (size_t q = 0; q < P.size(); ++q)
                if (z[q] <= 0.0) {
                    const double xi = x[P[q]];
                    const double a = xi / (xi - z[q]);
                    if (a < alpha) {
                        alpha = a;
                        qmin = q;
                    }
                }
            if (qmin == P.size()) {  // step 7: all positive, accept
                for (size_t q = 0; q < P.size(); ++q)
                    x[P[q]] = z[q];
                std::fill(blocked.begin(), blocked.end(), 0);
                break;
            }
            // steps 8-11: move toward z until a variable hits zero, drop it
            for (size_t q = 0; q < P.size(); ++q)
                x[P[q]] += alpha * (z[q] - x[P[q]]);
            x[P[qmin]] = 0.0;
            passive[P[qmin]] = 0;
            for (size_t q = 0; q < P.size(); ++q)
                if (x[P[q]] <= 0.0) {
                    x[P[q]] = 0.0;
                    passive[P[q]] = 0;
                }
        }
        gradient();
    }
    return false;
}

// L-curve curvature at lambda from the SVD (B/Laplace/src/TikhonovSolver.cpp:82-116):
// eta = |x|^2, rho = |r|^2 restricted to range(U), deta = d eta / d lambda.  The
// denominator is the reference's (lambda^2 eta^2 + rho^2)^1.5 (TikhonovSolver.cpp:111);
// Hansen's closed form has lambda^4 eta^2 there and is then twice the geometric
// curvature of (ln rho, ln eta).  The reference's form is kept so that the same
// lambda is chosen.
double hansen_curvature(const Vec &s, const Vec &uatb, const Vec &q, const Vec &q2, double lambda, double &sol,
                        double &res)
{
    const double sql = lambda * lambda;
    double eta = 0, rho = 0, deta = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const double s2 = s[i] * s[i];
        const double f = s2 / (s2 + sql);
        const double omf = 1.0 - f;
        const double fq = f * q[i];
        const double ou = omf * uatb[i];
        eta += fq * fq;
        rho += ou * ou;
        deta += omf * f * f * q2[i];
    }
    deta *= -4.0 / lambda;
    const double mult = -2.0 * eta * rho / deta;
    const double num = sql * deta * rho + 2.0 * lambda * eta * rho + sql * sql * eta * deta;
    const double den = pow(sql * eta * eta + rho * rho, 1.5);
    sol = sqrt(eta);
    res = sqrt(rho);
    return mult * num / den;
}

Vec normal_noise(double std, int n)
{
    std::random_device rd;
    unsigned seed = rd.entropy() != 0 ? rd()
                                      : (unsigned)std::chrono::system_clock::now().time_since_epoch().count();
    std::default_random_engine gen(seed);
    std::normal_distribution<double> dist(0.0, std);
    Vec v;
    v.reserve(n);
    for (int i = 0; i < n; ++i)
        v.push_back(dist(gen));
    return v;
}

// The inverter state of B/Laplace/include/nmrinv_core.h:104-151.
struct Inverter {
    double min_t2 = 0.1, max_t2 = 1e4;
    bool use_log = true;
    int num_t2 = 128;
    double min_lambda = -4.0, max_lambda = 2.0;
    int num_lambdas = 512;
    int prune_num = 512;
    double noise_amp = 0.0;

    int orig_num_echos = 0;
    Vec orig_raw_bins, used_raw_bins, used_t2_bins, used_raw_noise, used_raw_amps, used_t2_amps;
    int prune_start = 0;
    std::vector<std::tuple<int, int, int>> slices;
    bool pruned = false;
    Mat pure;
    double used_lambda = 0.0;
    bool lambda_set = false;
    // L-curve
    bool lc_set = false;
    Vec lambdas, solution_norms, residual_norms, curvature;

    void set_pruning()
    {
        slices.clear();
        pruned = false;
        const int n = (int)orig_raw_bins.size();
        // The reference prunes whenever prune_num != 0; with prune_num >= the
        // number of raw points its start-index search runs past the arrays,
        // so that case keeps the raw bins here.
        if (prune_num > 1 && prune_num < n) {
            Vec lb(n);
            for (int i = 0; i < n; ++i)
                lb[i] = log10(orig_raw_bins[i]);
            const double lo = lb[0];
            const double diff = (lb[n - 1] - lo) / (prune_num - 1);
            Vec nb(prune_num);
            for (int i = 0; i < prune_num; ++i)
                nb[i] = lo + diff * i;
            int p = 1;
            while (p < prune_num - 1 && (pow(10.0, nb[p]) - orig_raw_bins[p]) <= 0.0)
                ++p;
            --p;
            int ib = 0;
            while (ib < n && (lb[ib] - nb[p]) <= 0.0)
                ++ib;
            for (int pi = p; pi < prune_num - 1; ++pi) {
                const int start = ib;
                const double bi = nb[pi], bf = nb[pi + 1];
                while (ib < n && lb[ib] >= bi && lb[ib] < bf)
                    ++ib;
                slices.emplace_back(pi, start, ib);
            }
            prune_start = p;
            pruned = true;
            used_raw_bins = prune_curve(orig_raw_bins.data(), n, true);
        } else {
            used_raw_bins = orig_raw_bins;
        }
    }

    Vec prune_curve(const double *data, int n, bool use_logmean) const
    {
        Vec out(prune_num, 0.0);
        std::copy(data, data + prune_start, out.begin());
        for (auto &t : slices) {
            const int pi = std::get<0>(t), b = std::get<1>(t), e = std::get<2>(t);
            double sacc = 0.0;
            if (use_logmean) {
                for (int i = b; i < e; ++i)
                    sacc += log10(data[i]);
                out[pi] = pow(10.0, sacc / (e - b));
            } else {
                for (int i = b; i < e; ++i)
                    sacc += data[i];
                out[pi] = sacc / (e - b);
            }
        }
        out[prune_num - 1] = data[n - 1];
        return out;
    }

    void set_config(int n, const double *raw_bins, const double *forced_t2, int n_forced)
    {
        lambda_set = false;
        lc_set = false;
        orig_num_echos = n;
        orig_raw_bins.assign(raw_bins, raw_bins + n);
        if (forced_t2)
            used_t2_bins.assign(forced_t2, forced_t2 + n_forced);
        else
            used_t2_bins = use_log ? logspace_acc(log10(min_t2), log10(max_t2), num_t2)
                                   : linspace_acc(min_t2, max_t2, num_t2);
        if (noise_amp != 0.0)
            used_raw_noise = normal_noise(noise_amp, n);
        else
            used_raw_noise.clear();
        set_pruning();
        pure = kernel_matrix(used_raw_bins, used_t2_bins);
    }

    // process_raw_amps (nmrinv_core.cpp:357-391) for num_decays curves.
    void process(int size, const double *amps)
    {
        const int nd = orig_num_echos ? size / orig_num_echos : 0;
        const int per = (int)used_raw_bins.size();
        used_raw_amps.assign((size_t)nd * per, 0.0);
        Vec noisy;
        const double *src = amps;
        if (noise_amp != 0.0 && (int)used_raw_noise.size() == orig_num_echos) {
            noisy.assign(amps, amps + size);
            for (int d = 0; d < nd; ++d)
                for (int i = 0; i < orig_num_echos; ++i)
                    noisy[(size_t)d * orig_num_echos + i] += used_raw_noise[i];
            src = noisy.data();
        }
        for (int d = 0; d < nd; ++d) {
            const double *curve = src + (size_t)d * orig_num_echos;
            if (pruned) {
                Vec p = prune_curve(curve, orig_num_echos, false);
                std::copy(p.begin(), p.end(), used_raw_amps.begin() + (size_t)d * per);
            } else {
                std::copy(curve, curve + orig_num_echos, used_raw_amps.begin() + (size_t)d * per);
            }
        }
    }

    void find_lambda(int size, const double *amps)
    {
        process(size, amps);
        Mat U;
        Vec s;
        jacobi_svd(pure, U, s);
        int keep = 0;
        while (keep < (int)s.size() && s[keep] != 0.0)
            ++keep;
        s.resize(keep);
        Vec uatb(keep, 0.0);
        const int m = pure.rows;
        for (int j = 0; j < keep; ++j) {
            double acc = 0;
            for (int i = 0; i < m; ++i)
                acc += U(i, j) * used_raw_amps[i];
            uatb[j] = acc;
        }
        Vec q(keep), q2(keep);
        for (int j = 0; j < keep; ++j) {
            q[j] = uatb[j] / s[j];
            q2[j] = q[j] * q[j];
        }
        lambdas = logspace_acc(min_lambda, max_lambda, num_lambdas);
        solution_norms.assign(num_lambdas, 0.0);
        residual_norms.assign(num_lambdas, 0.0);
        curvature.assign(num_lambdas, 0.0);
        for (int il = 0; il < num_lambdas; ++il)
            curvature[il] = hansen_curvature(s, uatb, q, q2, lambdas[il], solution_norms[il], residual_norms[il]);
        const int best = (int)(std::max_element(curvature.begin(), curvature.end()) - curvature.begin());
        lc_set = true;
        set_lambda(lambdas[best]);
    }

    void set_lambda(double lam)
    {
        used_lambda = lam;
        lambda_set = true;
    }

    bool invert(int size, const double *amps)
    {
        if (!lambda_set)
            return false;
        process(size, amps);
        const int nd = orig_num_echos ? size / orig_num_echos : 0;
        const int per = (int)used_raw_bins.size(), nb = (int)used_t2_bins.size();
        const int rows = per + (used_lambda != 0.0 ? nb : 0);
        Mat A(rows, nb);
        for (int i = 0; i < per; ++i)
            for (int j = 0; j < nb; ++j)
                A(i, j) = pure(i, j);
        if (used_lambda != 0.0)
            for (int j = 0; j < nb; ++j)
                A(per + j, j) = used_lambda;
        used_t2_amps.assign((size_t)nd * nb, 0.0);
        bool ok = true;
#pragma omp parallel for schedule(dynamic) reduction(&& : ok)
        for (int d = 0; d < nd; ++d) {
            Vec b(rows, 0.0), x;
            std::copy(used_raw_amps.begin() + (size_t)d * per, used_raw_amps.begin() + (size_t)(d + 1) * per,
                      b.begin());
            if (!nnls(A, b, x, 1e-10, 30 * nb + 100))
                ok = false;
            else
                std::copy(x.begin(), x.end(), used_t2_amps.begin() + (size_t)d * nb);
        }
        return ok;
    }
};

std::mutex g_mutex;
Inverter g_inv;

}  // namespace

extern "C" {

void initialize_inverter(int num_raw_bins, double *raw_bins, double min_t2, double max_t2, int num_t2_bins,
                         int num_lambdas, int prune_num, double noise_amp)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    g_inv.min_t2 = min_t2;
    g_inv.max_t2 = max_t2;
    g_inv.use_log = true;
    g_inv.num_t2 = num_t2_bins;
    g_inv.num_lambdas = num_lambdas;
    g_inv.prune_num = prune_num;
    g_inv.noise_amp = noise_amp;
    g_inv.set_config(num_raw_bins, raw_bins, nullptr, 0);
}

void initialize_inverter_with_t2_bins(int num_raw_bins, double *raw_bins, int num_t2_bins, double *t2_bins,
                                      int num_lambdas, int prune_num, double noise_amp)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    g_inv.num_t2 = num_t2_bins;
    g_inv.num_lambdas = num_lambdas;
    g_inv.prune_num = prune_num;
    g_inv.noise_amp = noise_amp;
    g_inv.set_config(num_raw_bins, raw_bins, t2_bins, num_t2_bins);
}

void rwnmr_inverter_set_lambda_range(double min_log10, double max_log10)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    g_inv.min_lambda = min_log10;
    g_inv.max_lambda = max_log10;
}

double find_best_lambda(int num_raw_amps, double *raw_amps)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    g_inv.find_lambda(num_raw_amps, raw_amps);
    return g_inv.used_lambda;
}

void get_processed_raw_amps(int data_size, double *raw_amps, double *processed)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    g_inv.process(data_size, raw_amps);
    std::copy(g_inv.used_raw_amps.begin(), g_inv.used_raw_amps.end(), processed);
}

bool invert(int data_size, double *raw_amps, double *out_result)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_inv.lambda_set)
        return false;
    if (!g_inv.invert(data_size, raw_amps))
        return false;
    std::copy(g_inv.used_t2_amps.begin(), g_inv.used_t2_amps.end(), out_result);
    return true;
}

bool invert_with_multiple_noises(const int raw_amps_size, double *raw_amps_ptr, const int num_inversions,
                                 double *raw_noise_ptr, double *t2_amps_ptr)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_inv.lambda_set)
        return false;
    const int nd = g_inv.orig_num_echos ? raw_amps_size / g_inv.orig_num_echos : 0;
    const int total = nd * (int)g_inv.used_t2_bins.size();
    const double prev_amp = g_inv.noise_amp;
    const Vec prev_noise = g_inv.used_raw_noise;
    g_inv.noise_amp = 1.0;
    g_inv.used_raw_noise.assign(g_inv.orig_num_echos, 0.0);
    std::fill_n(t2_amps_ptr, total, 0.0);
    bool ok = true;
    for (int k = 0; k < num_inversions && ok; ++k) {
        std::copy(raw_noise_ptr + (size_t)k * g_inv.orig_num_echos,
                  raw_noise_ptr + (size_t)(k + 1) * g_inv.orig_num_echos, g_inv.used_raw_noise.begin());
        ok = g_inv.invert(raw_amps_size, raw_amps_ptr);
        for (int i = 0; i < total && ok; ++i)
            t2_amps_ptr[i] += g_inv.used_t2_amps[i];
    }
    for (int i = 0; i < total; ++i)
        t2_amps_ptr[i] /= num_inversions;
    g_inv.noise_amp = prev_amp;
    g_inv.used_raw_noise = prev_noise;
    return ok;
}

bool get_lcurve_values(double *solution_norms_ptr, double *residual_norms_ptr, double *lambdas_ptr,
                       double *curvature_ptr)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_inv.lc_set)
        return false;
    std::copy(g_inv.solution_norms.begin(), g_inv.solution_norms.end(), solution_norms_ptr);
    std::copy(g_inv.residual_norms.begin(), g_inv.residual_norms.end(), residual_norms_ptr);
    std::copy(g_inv.lambdas.begin(), g_inv.lambdas.end(), lambdas_ptr);
    std::copy(g_inv.curvature.begin(), g_inv.curvature.end(), curvature_ptr);
    return true;
}

double get_tikhonov_lambda(void)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    return g_inv.lambda_set ? g_inv.used_lambda : 0.0;
}

void set_lambda(double tik_lambda)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    g_inv.set_lambda(tik_lambda);
}

bool get_noise(double *out)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_inv.noise_amp == 0.0)
        return false;
    std::copy(g_inv.used_raw_noise.begin(), g_inv.used_raw_noise.end(), out);
    return true;
}

void set_noise(double noise_amp, double *in)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    g_inv.noise_amp = noise_amp;
    if (noise_amp != 0.0)
        g_inv.used_raw_noise.assign(in, in + g_inv.orig_num_echos);
}

void get_used_raw_bins(double *out)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    std::copy(g_inv.used_raw_bins.begin(), g_inv.used_raw_bins.end(), out);
}

void get_used_t2_bins(double *out)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    std::copy(g_inv.used_t2_bins.begin(), g_inv.used_t2_bins.end(), out);
}

int rwnmr_inverter_num_used_raw_bins(void)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    return (int)g_inv.used_raw_bins.size();
}

int rwnmr_inverter_num_t2_bins(void)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    return (int)g_inv.used_t2_bins.size();
}

int rwnmr_inverter_num_lambdas(void)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    return (int)g_inv.lambdas.size();
}

int32_t rwnmr_laplace_nnls(int m, int n, const double *A, const double *b, double *x_out, double lambda,
                           double tol, int max_iter)
{
    if (m <= 0 || n <= 0 || !A || !b || !x_out)
        return RWNMR_ERR_INVALID;
    const int rows = m + (lambda != 0.0 ? n : 0);
    Mat M(rows, n);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j)
            M(i, j) = A[(size_t)i * n + j];
    if (lambda != 0.0)
        for (int j = 0; j < n; ++j)
            M(m + j, j) = lambda;
    Vec bb(rows, 0.0), x;
    std::copy(b, b + m, bb.begin());
    if (!nnls(M, bb, x, tol > 0 ? tol : 1e-10, max_iter > 0 ? max_iter : 30 * n + 100))
        return RWNMR_ERR_NUMERIC;
    std::copy(x.begin(), x.end(), x_out);
    return RWNMR_OK;
}

}  // extern "C"
