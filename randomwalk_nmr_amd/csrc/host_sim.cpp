// host_sim.cpp -- host-side simulation set-up for the MI355X hot path:
// seeded synthetic rock images, walker seeding and placement the reference's
// way, and the reference's time framework / penalty formulas.
// Reference paths under src/backend/ of RandomWalk_NMR.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "rwnmr/rwnmr_host.h"

namespace {

// splitmix64 (Steele, Lea, Flood 2014): the generator of the synthetic
// images (image content is generated here; the reference reads images).
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform double in [0, 1) from the top 53 bits
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Lemire (2019), "Fast Random Integer Generation in an Interval": unbiased
// draw in [0, range) from a 32-bit generator with one multiply in the common
// case.  This is the method libstdc++ >= 11 uses for
// uniform_int_distribution over a 32-bit engine, so placement matches a
// reference build with a current compiler.
template <class G> uint32_t lemire_draw(G &g, uint32_t range)
{
    uint64_t m = (uint64_t)(uint32_t)g() * (uint64_t)range;
    uint32_t lo = (uint32_t)m;
    if (lo < range) {
        const uint32_t t = (uint32_t)(0u - range) % range;
        while (lo < t) {
            m = (uint64_t)(uint32_t)g() * (uint64_t)range;
            lo = (uint32_t)m;
        }
    }
    return (uint32_t)(m >> 32);
}

}  // namespace

extern "C" {

int32_t rwnmr_synth_sphere_pack(int32_t W, int32_t H, int32_t D, double r_min, double r_max,
                                double target_porosity, uint64_t seed, uint8_t *solid, double *porosity_out)
{
    if (W <= 0 || H <= 0 || D <= 0 || !solid || r_min <= 0 || r_max < r_min || target_porosity < 0 ||
        target_porosity > 1)
        return RWNMR_ERR_INVALID;
    const uint64_t V = (uint64_t)W * H * D;
    memset(solid, 0, V);
    uint64_t pores = V;
    const uint64_t goal = (uint64_t)floor(target_porosity * (double)V);
    SplitMix64 rng(seed);
    uint64_t guard = 0;
    while (pores > goal && guard < 100000000ull) {
        ++guard;
        const double cx = rng.uniform() * W, cy = rng.uniform() * H, cz = rng.uniform() * D;
        const double r = r_min + (r_max - r_min) * rng.uniform();
        const double r2 = r * r;
        const int x0 = std::max(0, (int)ceil(cx - r)), x1 = std::min(W - 1, (int)floor(cx + r));
        const int y0 = std::max(0, (int)ceil(cy - r)), y1 = std::min(H - 1, (int)floor(cy + r));
        const int z0 = std::max(0, (int)ceil(cz - r)), z1 = std::min(D - 1, (int)floor(cz + r));
        for (int z = z0; z <= z1; ++z) {
            const double dz2 = (z - cz) * (z - cz);
            for (int y = y0; y <= y1; ++y) {
                const double dyz = dz2 + (y - cy) * (y - cy);
                if (dyz > r2)
                    continue;
                uint8_t *row = solid + ((size_t)z * H + y) * W;
                for (int x = x0; x <= x1; ++x) {
                    const double d2 = dyz + (x - cx) * (x - cx);
                    if (d2 <= r2 && !row[x]) {
                        row[x] = 1;
                        --pores;
                    }
                }
            }
        }
    }
    if (porosity_out)
        *porosity_out = (double)pores / (double)V;
    return RWNMR_OK;
}

int32_t rwnmr_synth_voronoi_labels(int32_t W, int32_t H, int32_t D, int32_t n_sites, int32_t n_labels,
                                   uint64_t seed, const uint8_t *solid, uint8_t *labels)
{
    if (W <= 0 || H <= 0 || D <= 0 || n_sites <= 0 || n_labels <= 0 || n_labels > 255 || !labels)
        return RWNMR_ERR_INVALID;
    SplitMix64 rng(seed);
    std::vector<double> sx(n_sites), sy(n_sites), sz(n_sites);
    for (int i = 0; i < n_sites; ++i) {
        sx[i] = rng.uniform() * W;
        sy[i] = rng.uniform() * H;
        sz[i] = rng.uniform() * D;
    }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t v = ((size_t)z * H + y) * W + x;
                if (solid && !solid[v]) {
                    labels[v] = 0;
                    continue;
                }
                int best = 0;
                double bd = 1e300;
                for (int i = 0; i < n_sites; ++i) {
                    const double d = (x - sx[i]) * (x - sx[i]) + (y - sy[i]) * (y - sy[i]) + (z - sz[i]) * (z - sz[i]);
                    if (d < bd) {
                        bd = d;
                        best = i;
                    }
                }
                labels[v] = (uint8_t)(best % n_labels);
            }
    return RWNMR_OK;
}

int64_t rwnmr_count_pores(const uint8_t *solid, int32_t W, int32_t H, int32_t D)
{
    if (!solid)
        return -1;
    int64_t n = 0;
    const size_t V = (size_t)W * H * D;
    for (size_t i = 0; i < V; ++i)
        n += solid[i] ? 0 : 1;
    return n;
}

int32_t rwnmr_seed_and_place(const uint8_t *solid, int32_t W, int32_t H, int32_t D, uint64_t master_seed,
                             uint64_t n, uint64_t *seeds, int32_t *x, int32_t *y, int32_t *z)
{
    if (!solid || W <= 0 || H <= 0 || D <= 0 || (n && (!seeds || !x || !y || !z)))
        return RWNMR_ERR_INVALID;
    // B/NMR_Simulation/NMR_Simulation.cpp:86-95: mt19937 seeded with the
    // master seed (mod 2^32), 4096 outputs discarded.
    std::mt19937 g((uint32_t)master_seed);
    g.discard(4096);
    // :1139-1204 createWalkers: per walker, in order, a 64-bit seed made of two
    // consecutive 32-bit outputs, high word first.
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t hi = (uint32_t)g();
        const uint64_t lo = (uint32_t)g();
        seeds[i] = (hi << 32) | lo;
    }
    // :930-980 pore list in z, y, x raster order (linear voxel indices here).
    const uint64_t V = (uint64_t)W * H * D;
    std::vector<uint32_t> pores;
    pores.reserve(V / 4);
    for (uint64_t v = 0; v < V; ++v)
        if (!solid[v])
            pores.push_back((uint32_t)v);
    if (pores.empty())
        return n ? RWNMR_ERR_INVALID : RWNMR_OK;
    // :915-928 numberOfPores recomputed from the porosity; :1086-1136 draws in
    // [0, numberOfPores] inclusive; :1271-1333 places each walker.  An index
    // past the list end (a read past the vector in the reference) selects
    // the last pore here.
    const double porosity = (double)pores.size() / (double)V;
    const uint32_t cap = (uint32_t)(porosity * (double)V);
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t idx = lemire_draw(g, cap + 1u);
        if (idx >= pores.size())
            idx = (uint32_t)pores.size() - 1;
        const uint32_t v = pores[idx];
        x[i] = (int32_t)(v % (uint32_t)W);
        y[i] = (int32_t)((v / (uint32_t)W) % (uint32_t)H);
        z[i] = (int32_t)(v / ((uint32_t)W * (uint32_t)H));
    }
    return RWNMR_OK;
}

int32_t rwnmr_setup_walkers(const uint8_t *solid, int32_t W, int32_t H, int32_t D, uint64_t master_seed, uint64_t n,
                            int32_t mode, int32_t deviation, uint32_t shift, uint64_t *seeds, int32_t *x,
                            int32_t *y, int32_t *z, double *occupancy, uint64_t *draws)
{
    if (!solid || W <= 0 || H <= 0 || D <= 0 || mode < 0 || mode > 2 || shift > 8 ||
        (n && (!seeds || !x || !y || !z)))
        return RWNMR_ERR_INVALID;
    const uint64_t V = (uint64_t)W * H * D;
    // Counting generator: every 32-bit output is counted so the caller can
    // resume the same stream (mt19937 + discard) for later draws.
    struct Counted {
        std::mt19937 g;
        uint64_t used;
        typedef uint32_t result_type;
        uint32_t operator()()
        {
            ++used;
            return (uint32_t)g();
        }
    } g{std::mt19937((uint32_t)master_seed), 4096};
    g.g.discard(4096);
    uint64_t pores_total = 0;
    for (uint64_t v = 0; v < V; ++v)
        pores_total += solid[v] ? 0 : 1;
    if (occupancy)
        *occupancy = (double)n / (double)pores_total;   // updateWalkerOccupancy, :1081-1084
    if (mode == 0) {
        // setWalkers(n) (:365-384): createWalkers, pore list, ID list, placement
        int32_t st = rwnmr_seed_and_place(solid, W, H, D, master_seed, n, seeds, x, y, z);
        if (st != RWNMR_OK)
            return st;
        // replay the draws to count them (seed words, then the index draws)
        g.used = 4096 + 2 * n;
        for (uint64_t i = 0; i < 2 * n; ++i)
            g.g();
        const double porosity = (double)pores_total / (double)V;
        const uint32_t cap = (uint32_t)(porosity * (double)V);
        for (uint64_t i = 0; i < n; ++i)
            lemire_draw(g, cap + 1u);
    } else {
        // createWalkers (:1139-1204): two outputs per seed, high word first
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t hi = g();
            const uint64_t lo = g();
            seeds[i] = (hi << 32) | lo;
        }
        if (mode == 1) {
            // "point" (:328-337, placeWalkersInSamePoint :1335-1393): every
            // walker at the image centre, which must be a pore voxel
            const int32_t cx = W / 2, cy = H / 2, cz = D / 2;
            if (solid[(uint64_t)cz * W * H + (uint64_t)cy * W + cx])
                return RWNMR_ERR_INVALID;
            for (uint64_t i = 0; i < n; ++i) {
                x[i] = cx;
                y[i] = cy;
                z[i] = cz;
            }
        } else {
            // "cubic" (:338-349, countPoresInCubicSpace :728-816, createPoreList
            // :982-1069, placeWalkersInCubicSpace :1395-1454): pores of
            // [c - dev, c + dev) per axis (upper end clipped to the image,
            // a negative lower end leaves the cube empty), c = (dim - 1) / 2,
            // in z, y, x order; draws uniform over [0, count] inclusive (the
            // index equal to the count selects the last pore here)
            const int64_t c[3] = {(W - 1) / 2, (H - 1) / 2, (D - 1) / 2};
            const int64_t dim[3] = {W, H, D};
            int64_t lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                lo[a] = c[a] - deviation;
                hi[a] = std::min<int64_t>(c[a] + deviation, dim[a]);
                if (lo[a] < 0)
                    hi[a] = lo[a];
            }
            std::vector<uint32_t> cube;
            for (int64_t zz = lo[2]; zz < hi[2]; ++zz)
                for (int64_t yy = lo[1]; yy < hi[1]; ++yy)
                    for (int64_t xx = lo[0]; xx < hi[0]; ++xx) {
                        const uint64_t v = (uint64_t)zz * W * H + (uint64_t)yy * W + (uint64_t)xx;
                        if (!solid[v])
                            cube.push_back((uint32_t)v);
                    }
            if (cube.empty())
                return n ? RWNMR_ERR_INVALID : RWNMR_OK;
            for (uint64_t i = 0; i < n; ++i) {
                uint32_t idx = lemire_draw(g, (uint32_t)cube.size() + 1u);
                if (idx >= cube.size())
                    idx = (uint32_t)cube.size() - 1;
                const uint32_t v = cube[idx];
                x[i] = (int32_t)(v % (uint32_t)W);
                y[i] = (int32_t)((v / (uint32_t)W) % (uint32_t)H);
                z[i] = (int32_t)(v / ((uint32_t)W * (uint32_t)H));
            }
        }
    }
    if (shift > 0) {
        // applyVoxelDivision (:192-268): pos * 2^shift + a draw in
        // [0, 2^shift - 1] per axis, x then y then z, walker by walker
        const uint32_t f = 1u << shift;
        for (uint64_t i = 0; i < n; ++i) {
            x[i] = (int32_t)(x[i] * f + lemire_draw(g, f));
            y[i] = (int32_t)(y[i] * f + lemire_draw(g, f));
            z[i] = (int32_t)(z[i] * f + lemire_draw(g, f));
        }
    }
    if (draws)
        *draws = g.used;
    return RWNMR_OK;
}

double rwnmr_time_step(double voxel_size, double D0)
{
    // B/NMR_Simulation/NMR_Simulation.cpp:179-183: dt = eps^2 / (6 D0)
    return voxel_size * voxel_size / (6.0 * D0);
}

int32_t rwnmr_time_framework(double time_ms, double voxel_size, double D0, uint32_t steps_per_echo,
                             uint32_t *echoes_out, uint64_t *steps_out)
{
    if (steps_per_echo == 0 || voxel_size <= 0 || D0 <= 0 || !echoes_out || !steps_out)
        return RWNMR_ERR_INVALID;
    // :276-298 setNumberOfStepsFromTime + setTimeFramework(double)
    double steps = round(time_ms * (6.0 * D0 / (voxel_size * voxel_size)));
    if (steps < steps_per_echo)
        steps = steps_per_echo;
    const uint64_t st = (uint64_t)steps;
    const uint64_t echoes = (uint64_t)ceil((double)st / (double)steps_per_echo);
    *echoes_out = (uint32_t)echoes;
    *steps_out = echoes * steps_per_echo;
    return RWNMR_OK;
}

double rwnmr_penalty(double rho, double voxel_size, double D0)
{
    // B/Walker/walker.h:203-209
    return 1.0 - (2.0 / 3.0) * ((rho * voxel_size) / (D0 * 1e03));
}

}  // extern "C"
