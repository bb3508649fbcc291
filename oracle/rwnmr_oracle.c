/*
 * rwnmr_oracle.c -- CPU oracle for the random-walk NMR hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP path is
 * compared against: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product library never links or calls it.
 *
 * Every routine restates the behaviour of RandomWalk_NMR (paths relative to
 * src/backend/ of that repository); none of its code is copied.  Where the
 * reference's GPU and CPU paths differ, the routine says which one it follows.
 *
 * Pinning (what fixes each routine independently of this file):
 *   xorshift64           Marsaglia (2003) 13/7/17 triple; known-answer vectors
 *                        computed independently in numpy (tests/test_oracle_*.py)
 *   mod6 digit sum       brute force against a % 6 (proof: digit sum in base 8
 *                        with the (a>>2)&~1 trick preserves a mod 6)
 *   bitBlock packing     brute-force pack/unpack round trip on ragged volumes,
 *                        and the reference's 3 PNG stacks packed by hand in numpy
 *   folds                brute-force reflective / periodic tilings
 *   walk                 closed forms: all-pore image keeps E == 1 and the
 *                        mean-square displacement equals the step count;
 *                        energy == sequential product of delta over the hit
 *                        count; hits(map) == wall events of the CPMG replay
 *   mt19937              the C++ standard's 10000th-output requirement
 *                        (4123659995) and numpy's RandomState stream
 *   Lemire sampler       std::uniform_int_distribution of libstdc++ 11
 *   pfgse                free-diffusion closed form exp(-k^2 D0 Delta) and the
 *                        k = 0 column equal to the plain energy sum
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_BC_NOFLUX 0
#define ORC_BC_PERIODIC 1
#define ORC_BC_MIRROR 2

/* ------------------------------------------------------------------ RNG */

/* Marsaglia xorshift64 with shifts 13, 7, 17; the new state is also the
 * output.  B/RNG/xorshift.cpp:17-24, B/Walker/walker.h:557-564,
 * B/NMR_Simulation/NMR_cpmg_cuda.cpp:1820-1827. */
uint64_t orc_xorshift64(uint64_t s)
{
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}

/* The reference's residue-mod-6 routine: repeated base-8 digit sums with the
 * "(a >> 2) & ~1" step, then one conditional subtraction.
 * B/Walker/walker.h:566-582, B/NMR_Simulation/NMR_cpmg_cuda.cpp:1829-1845.
 * It equals a % 6 for every a (tests check this by brute force). */
uint64_t orc_mod6_digit_sum(uint64_t a)
{
    while (a > 11) {
        uint64_t acc = 0;
        while (a != 0) {
            acc += a & 7u;
            a = (a >> 2) & ~(uint64_t)1;
        }
        a = (uint64_t)(int)acc;
    }
    return a > 5 ? a - 6 : a;
}

/* Advance the walker's state and draw the direction 1..6
 * (1 North, 2 West, 3 South, 4 East, 5 Up, 6 Down; B/Walker/walker_defs.h:6-23).
 * B/Walker/walker.h:346-358.  use_digit_sum selects the reference's loop
 * (for timing the reference CPU path); otherwise the C remainder is used. */
static inline int orc_draw_direction(uint64_t *seed, int use_digit_sum)
{
    uint64_t s = orc_xorshift64(*seed);
    *seed = s;
    return (int)(use_digit_sum ? orc_mod6_digit_sum(s) : s % 6u) + 1;
}

/* Unit displacement of a direction.  B/Walker/walker.h:360-391. */
static inline void orc_direction_delta(int dir, int *dx, int *dy, int *dz)
{
    *dx = *dy = *dz = 0;
    switch (dir) {
    case 1: *dy = -1; break; /* North */
    case 2: *dx = -1; break; /* West  */
    case 3: *dy = +1; break; /* South */
    case 4: *dx = +1; break; /* East  */
    case 5: *dz = +1; break; /* Up    */
    case 6: *dz = -1; break; /* Down  */
    default: break;
    }
}

/* noflux: an outward move at an image face is replaced by the opposite
 * direction before the wall test.  B/Walker/walker.h:444-491,
 * B/NMR_Simulation/NMR_cpmg_cuda.cpp:1847-1902 (positions already shifted
 * to image voxels, NMR_cpmg_cuda.cpp:88-94). */
static inline int orc_noflux_flip(int dir, int gx, int gy, int gz, int W, int H, int D)
{
    switch (dir) {
    case 1: return gy == 0 ? 3 : dir;
    case 3: return gy == H - 1 ? 1 : dir;
    case 2: return gx == 0 ? 4 : dir;
    case 4: return gx == W - 1 ? 2 : dir;
    case 5: return gz == D - 1 ? 6 : dir;
    case 6: return gz == 0 ? 5 : dir;
    default: return dir;
    }
}

/* ---------------------------------------------------------------- folds */

/* Periodic image coordinate of an unbounded walker coordinate, C truncating
 * remainder then a correction.  B/NMR_Simulation/NMR_cpmg_cuda.cpp:1991-1999. */
int orc_fold_periodic(int p, int shift, int d)
{
    int g = p >> shift;
    int m = g % d;
    return m < 0 ? m + d : m;
}

/* Mirror image coordinate: the tile index parity (computed with C truncating
 * division, with the reference's separate formula for g <= 0) selects a
 * reflected copy.  B/NMR_Simulation/NMR_cpmg_cuda.cpp:2001-2015. */
int orc_fold_mirror(int p, int shift, int d)
{
    int g = p >> shift;
    int m = g % d;
    if (m < 0)
        m += d;
    int reflected = (g > 0) ? (g / d) % 2 : ((-g - 1 + d) / d) % 2;
    return reflected ? d - 1 - m : m;
}

/* ------------------------------------------------------------- bitBlock */

/* Block-grid sizes: ceil(dim / 4) per axis.  B/BitBlock/bitBlock.cpp:138-166. */
void orc_bitblock_dims(int W, int H, int D, int *bC, int *bR, int *bD)
{
    *bC = (W + 3) / 4;
    *bR = (H + 3) / 4;
    *bD = (D + 3) / 4;
}

/* Pack a voxel volume (x fastest, nonzero = solid) into 4x4x4 uint64 blocks:
 * block index bx + by*bC + bz*bC*bR, bit x%4 + 4*(y%4) + 16*(z%4), padding
 * bits 0.  B/BitBlock/bitBlock.cpp:168-221, B/BitBlock/bitBlock.h:113-134. */
void orc_bitblock_pack(const uint8_t *solid, int W, int H, int D, uint64_t *blocks)
{
    int bC, bR, bD;
    orc_bitblock_dims(W, H, D, &bC, &bR, &bD);
    memset(blocks, 0, sizeof(uint64_t) * (size_t)bC * bR * bD);
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
                if (solid[((size_t)z * H + y) * W + x]) {
                    size_t b = (size_t)(x >> 2) + (size_t)(y >> 2) * bC + (size_t)(z >> 2) * bC * bR;
                    int bit = (x & 3) + ((y & 3) << 2) + ((z & 3) << 4);
                    blocks[b] |= 1ull << bit;
                }
}

/* Wall test of an in-image voxel.  B/BitBlock/bitBlock.h:113-139. */
static inline int orc_is_wall(const uint64_t *blocks, int bC, int bR, int x, int y, int z)
{
    size_t b = (size_t)(x >> 2) + (size_t)(y >> 2) * bC + (size_t)(z >> 2) * (size_t)bC * bR;
    int bit = (x & 3) + ((y & 3) << 2) + ((z & 3) << 4);
    return (int)((blocks[b] >> bit) & 1u);
}

int orc_bitblock_is_wall(const uint64_t *blocks, int bC, int bR, int x, int y, int z)
{
    return orc_is_wall(blocks, bC, bR, x, y, z);
}

/* --------------------------------------------------------------- a step */

typedef struct {
    const uint64_t *blocks;
    int W, H, D;       /* image columns, rows, depth */
    int bC, bR;        /* block columns, block rows */
    const uint8_t *labels; /* optional W*H*D solid-label volume (x fastest) */
    int bc, shift;
} orc_geom;

/* One reference step (B/Walker/walker.cpp:215-231 for the noflux CPU path,
 * B/NMR_Simulation/NMR_cpmg_cuda.cpp:82-125 / 351-387 / 625-662 for the
 * GPU noflux / periodic / mirror kernels).  Returns -1 when the walker moved,
 * otherwise the label of the solid voxel it hit (0 without a label volume). */
static inline int orc_step(const orc_geom *g, int *x, int *y, int *z, uint64_t *seed,
                           int use_digit_sum)
{
    int dir = orc_draw_direction(seed, use_digit_sum);
    if (g->bc == ORC_BC_NOFLUX)
        dir = orc_noflux_flip(dir, *x >> g->shift, *y >> g->shift, *z >> g->shift, g->W, g->H, g->D);
    int dx, dy, dz;
    orc_direction_delta(dir, &dx, &dy, &dz);
    int nx = *x + dx, ny = *y + dy, nz = *z + dz;
    int ix, iy, iz;
    if (g->bc == ORC_BC_PERIODIC) {
        ix = orc_fold_periodic(nx, g->shift, g->W);
        iy = orc_fold_periodic(ny, g->shift, g->H);
        iz = orc_fold_periodic(nz, g->shift, g->D);
    } else if (g->bc == ORC_BC_MIRROR) {
        ix = orc_fold_mirror(nx, g->shift, g->W);
        iy = orc_fold_mirror(ny, g->shift, g->H);
        iz = orc_fold_mirror(nz, g->shift, g->D);
    } else {
        ix = nx >> g->shift;
        iy = ny >> g->shift;
        iz = nz >> g->shift;
    }
    if (!orc_is_wall(g->blocks, g->bC, g->bR, ix, iy, iz)) {
        *x = nx;
        *y = ny;
        *z = nz;
        return -1;
    }
    return g->labels ? (int)g->labels[((size_t)iz * g->H + iy) * g->W + ix] : 0;
}

/* Neumaier-compensated long double accumulator: the exact-sum reference for
 * the per-echo magnetisation (reference sums: NMR_cpmg_cuda.cpp:888-925,
 * 1780-1801, NMR_cpmg.cpp:218-227). */
typedef struct {
    long double s, c;
} orc_acc;

static inline void orc_acc_add(orc_acc *a, long double v)
{
    long double t = a->s + v;
    if (fabsl(a->s) >= fabsl(v))
        a->c += (a->s - t) + v;
    else
        a->c += (v - t) + a->s;
    a->s = t;
}

static int orc_threads(int requested)
{
#ifdef _OPENMP
    return requested > 0 ? requested : omp_get_max_threads();
#else
    (void)requested;
    return 1;
#endif
}

/* ------------------------------------------------------------ CPMG walk */

/*
 * CPMG / map walk of n walkers over `echoes` x `steps_per_echo` steps.
 *
 * Follows the GPU kernels CPMG_walk_{noflux,periodic,mirror}
 * (B/NMR_Simulation/NMR_cpmg_cuda.cpp:37-139, 300-402, 574-676) and the map
 * kernels (B/NMR_Simulation/NMR_Simulation_cuda.cpp:418-721): every wall hit
 * leaves the walker in place, adds one hit and multiplies its energy by its
 * penalty (per-walker `penalty`, or label_delta[label of the hit voxel] when
 * the geometry carries labels).
 *
 * echo_sums (optional, echoes+1 doubles): entry 0 is the sum of the initial
 * energies, entry e the exact (compensated long double) sum after echo e.
 * label_hits (optional, n*n_labels): hits per solid label, walker-major.
 * energy may be NULL (no energy tracking); hits may be NULL.
 */
int orc_walk(const uint64_t *blocks, int W, int H, int D, const uint8_t *labels, int n_labels,
             int bc, int shift, int use_digit_sum, int64_t n, int32_t *x, int32_t *y, int32_t *z,
             uint64_t *seed, double *energy, uint32_t *hits, const double *penalty,
             const double *label_delta, uint32_t *label_hits, int64_t echoes, int64_t steps_per_echo,
             double *echo_sums, int threads)
{
    if (n < 0 || echoes < 0 || steps_per_echo < 0 || W <= 0 || H <= 0 || D <= 0)
        return 1;
    if (labels && (n_labels <= 0 || !label_delta))
        return 1;
    orc_geom g;
    g.blocks = blocks;
    g.W = W;
    g.H = H;
    g.D = D;
    g.bC = (W + 3) / 4;
    g.bR = (H + 3) / 4;
    g.labels = labels;
    g.bc = bc;
    g.shift = shift;

    int nt = orc_threads(threads);
    orc_acc *acc = NULL;
    if (echo_sums) {
        acc = (orc_acc *)calloc((size_t)nt * (size_t)(echoes + 1), sizeof(orc_acc));
        if (!acc)
            return 2;
    }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        orc_acc *my = acc ? acc + (size_t)tid * (size_t)(echoes + 1) : NULL;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t w = 0; w < n; ++w) {
            int px = x[w], py = y[w], pz = z[w];
            uint64_t s = seed[w];
            double e = energy ? energy[w] : 1.0;
            uint32_t h = hits ? hits[w] : 0u;
            uint32_t *lh = label_hits ? label_hits + (size_t)w * n_labels : NULL;
            double dw = penalty ? penalty[w] : 1.0;
            if (my)
                orc_acc_add(&my[0], e);
            for (int64_t ec = 1; ec <= echoes; ++ec) {
                for (int64_t st = 0; st < steps_per_echo; ++st) {
                    int lab = orc_step(&g, &px, &py, &pz, &s, use_digit_sum);
                    if (lab >= 0) {
                        ++h;
                        if (labels) {
                            e = e * label_delta[lab];
                            if (lh)
                                ++lh[lab];
                        } else {
                            e = e * dw;
                        }
                    }
                }
                if (my)
                    orc_acc_add(&my[ec], e);
            }
            x[w] = px;
            y[w] = py;
            z[w] = pz;
            seed[w] = s;
            if (energy)
                energy[w] = e;
            if (hits)
                hits[w] = h;
        }
    }
    if (echo_sums) {
        for (int64_t ec = 0; ec <= echoes; ++ec) {
            orc_acc tot = {0.0L, 0.0L};
            for (int t = 0; t < nt; ++t) {
                orc_acc_add(&tot, acc[(size_t)t * (echoes + 1) + ec].s);
                orc_acc_add(&tot, acc[(size_t)t * (echoes + 1) + ec].c);
            }
            echo_sums[ec] = (double)(tot.s + tot.c);
        }
        free(acc);
    }
    return 0;
}

/* Image voxel of a walker position under the geometry's boundary (the
 * fold the GPU kernels apply before a bitBlock or field lookup). */
static inline void orc_image_voxel(const orc_geom *g, int x, int y, int z, int *ix, int *iy, int *iz)
{
    if (g->bc == ORC_BC_PERIODIC) {
        *ix = orc_fold_periodic(x, g->shift, g->W);
        *iy = orc_fold_periodic(y, g->shift, g->H);
        *iz = orc_fold_periodic(z, g->shift, g->D);
    } else if (g->bc == ORC_BC_MIRROR) {
        *ix = orc_fold_mirror(x, g->shift, g->W);
        *iy = orc_fold_mirror(y, g->shift, g->H);
        *iz = orc_fold_mirror(z, g->shift, g->D);
    } else {
        *ix = x >> g->shift;
        *iy = y >> g->shift;
        *iz = z >> g->shift;
    }
}

/*
 * Internal-field CPMG (B/NMR_Simulation/NMR_cpmg_cuda.cpp:146-293 noflux,
 * 409-567 periodic, 683-841 mirror).  Per echo two phase inversions
 * (phase = -phase), each of steps_per_echo / 2 (integer division) steps;
 * before and after every step the phase gains gammatau * field[voxel] with
 * voxel = ix + iy W + iz W H of the walker's image position, written as a
 * fused multiply-add (the CUDA build contracts phase + gammatau * field);
 * a wall hit multiplies p_alive by the walker's penalty (no energy change
 * otherwise).  Echo value: p_alive * cos(phase); echo 0 from the initial
 * state.  p_alive / phase are in/out (NULL: start at 1 / 0).
 */
int orc_walk_field(const uint64_t *blocks, int W, int H, int D, int bc, int shift, int use_digit_sum, int64_t n,
                   int32_t *x, int32_t *y, int32_t *z, uint64_t *seed, double *p_alive, double *phase,
                   uint32_t *hits, const double *penalty, const double *field, double gammatau, int64_t echoes,
                   int64_t steps_per_echo, double *echo_sums, int threads)
{
    if (n < 0 || echoes < 0 || steps_per_echo < 0 || W <= 0 || H <= 0 || D <= 0 || !field)
        return 1;
    orc_geom g;
    g.blocks = blocks;
    g.W = W;
    g.H = H;
    g.D = D;
    g.bC = (W + 3) / 4;
    g.bR = (H + 3) / 4;
    g.labels = NULL;
    g.bc = bc;
    g.shift = shift;
    const int64_t half = steps_per_echo / 2;
    int nt = orc_threads(threads);
    orc_acc *acc = NULL;
    if (echo_sums) {
        acc = (orc_acc *)calloc((size_t)nt * (size_t)(echoes + 1), sizeof(orc_acc));
        if (!acc)
            return 2;
    }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        orc_acc *my = acc ? acc + (size_t)tid * (size_t)(echoes + 1) : NULL;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t w = 0; w < n; ++w) {
            int px = x[w], py = y[w], pz = z[w];
            uint64_t s = seed[w];
            double pa = p_alive ? p_alive[w] : 1.0;
            double ph = phase ? phase[w] : 0.0;
            uint32_t h = hits ? hits[w] : 0u;
            const double dw = penalty ? penalty[w] : 1.0;
            if (my)
                orc_acc_add(&my[0], pa * cos(ph));
            for (int64_t ec = 1; ec <= echoes; ++ec) {
                for (int inv = 0; inv < 2; ++inv) {
                    ph = -ph;
                    for (int64_t st = 0; st < half; ++st) {
                        int ix, iy, iz;
                        orc_image_voxel(&g, px, py, pz, &ix, &iy, &iz);
                        ph = fma(gammatau, field[(size_t)ix + (size_t)iy * W + (size_t)iz * W * H], ph);
                        if (orc_step(&g, &px, &py, &pz, &s, use_digit_sum) >= 0) {
                            ++h;
                            pa = pa * dw;
                        }
                        orc_image_voxel(&g, px, py, pz, &ix, &iy, &iz);
                        ph = fma(gammatau, field[(size_t)ix + (size_t)iy * W + (size_t)iz * W * H], ph);
                    }
                }
                if (my)
                    orc_acc_add(&my[ec], pa * cos(ph));
            }
            x[w] = px;
            y[w] = py;
            z[w] = pz;
            seed[w] = s;
            if (p_alive)
                p_alive[w] = pa;
            if (phase)
                phase[w] = ph;
            if (hits)
                hits[w] = h;
        }
    }
    if (echo_sums) {
        for (int64_t ec = 0; ec <= echoes; ++ec) {
            orc_acc tot = {0.0L, 0.0L};
            for (int t = 0; t < nt; ++t) {
                orc_acc_add(&tot, acc[(size_t)t * (echoes + 1) + ec].s);
                orc_acc_add(&tot, acc[(size_t)t * (echoes + 1) + ec].c);
            }
            echo_sums[ec] = (double)(tot.s + tot.c);
        }
        free(acc);
    }
    return 0;
}

/*
 * Reference CPU CPMG loop order, kept for timing only (the CPU baseline):
 * echo-major, walker AoS records, the digit-sum mod6, a sequential
 * E *= delta on each hit and a serial per-echo sum.
 * B/NMR_Simulation/NMR_cpmg.cpp:177-233 (single-threaded there; its OpenMP
 * pragmas are commented out, :209,:220), B/Walker/walker.cpp:215-231.
 * threads > 1 partitions the walkers statically inside each echo, as the
 * reference's OMPLoopEnabler would.  Returns the seconds spent in the step
 * loop (steady clock), or a negative value on error.
 */
typedef struct {
    int32_t x0, y0, z0;
    int32_t x, y, z;
    uint64_t seed0, seed;
    double delta, energy;
    uint32_t hits;
} orc_walker_rec;

double orc_cpmg_reference_loop(const uint64_t *blocks, int W, int H, int D, int bc, int shift,
                               int use_digit_sum, int64_t n, const int32_t *x, const int32_t *y,
                               const int32_t *z, const uint64_t *seed, const double *penalty,
                               int64_t echoes, int64_t steps_per_echo, double *echo_sums,
                               int32_t *x_out, int32_t *y_out, int32_t *z_out, uint64_t *seed_out,
                               double *energy_out, uint32_t *hits_out, int threads)
{
    orc_walker_rec *wk = (orc_walker_rec *)malloc(sizeof(orc_walker_rec) * (size_t)(n > 0 ? n : 1));
    if (!wk)
        return -1.0;
    for (int64_t w = 0; w < n; ++w) {
        wk[w].x0 = wk[w].x = x[w];
        wk[w].y0 = wk[w].y = y[w];
        wk[w].z0 = wk[w].z = z[w];
        wk[w].seed0 = wk[w].seed = seed[w];
        wk[w].delta = penalty[w];
        wk[w].energy = 1.0;
        wk[w].hits = 0;
    }
    orc_geom g;
    g.blocks = blocks;
    g.W = W;
    g.H = H;
    g.D = D;
    g.bC = (W + 3) / 4;
    g.bR = (H + 3) / 4;
    g.labels = NULL;
    g.bc = bc;
    g.shift = shift;
    int nt = orc_threads(threads);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    double sum = 0.0;
    for (int64_t w = 0; w < n; ++w)
        sum += wk[w].energy;
    if (echo_sums)
        echo_sums[0] = sum;
    for (int64_t ec = 1; ec <= echoes; ++ec) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
        for (int64_t w = 0; w < n; ++w) {
            orc_walker_rec *r = &wk[w];
            for (int64_t st = 0; st < steps_per_echo; ++st) {
                if (orc_step(&g, &r->x, &r->y, &r->z, &r->seed, use_digit_sum) >= 0) {
                    r->energy = r->energy * r->delta;
                    ++r->hits;
                }
            }
        }
        sum = 0.0;
        for (int64_t w = 0; w < n; ++w)
            sum += wk[w].energy;
        if (echo_sums)
            echo_sums[ec] = sum;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    (void)nt;
    if (x_out)
        for (int64_t w = 0; w < n; ++w) {
            x_out[w] = wk[w].x;
            y_out[w] = wk[w].y;
            z_out[w] = wk[w].z;
        }
    for (int64_t w = 0; w < n; ++w) {
        if (seed_out)
            seed_out[w] = wk[w].seed;
        if (energy_out)
            energy_out[w] = wk[w].energy;
        if (hits_out)
            hits_out[w] = wk[w].hits;
    }
    free(wk);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ---------------------------------------------------------------- PFGSE */

/*
 * PFGSE with the reference GPU semantics (B/NMR_Simulation/NMR_pfgse_cuda.cpp):
 * for each diffusion time i, walk steps_inc[i] more steps counting the hits of
 * this increment (PFG_map_*, :34-339, hits restarted at 0 per call, :730),
 * multiply the energy once by pow(delta, hits) (PFG_evaluate_energy, :342-354),
 * then M[i][k] = sum_w E_w cos((k . (r_w - r0_w)) * voxel)
 * (PFG_measure :358-389 with dotProduct :3097-3100; sum over k at :892-944).
 *
 * penalty: per-walker delta (n), or, when n_sets > 0, set_delta[n_sets] gives
 * one uniform delta per set and M is laid out [set][i][k].  energy (n) holds
 * the walkers' energy for set 0 / per-walker mode on return (may be NULL).
 * k_per_delta != 0: kx/ky/kz hold n_delta x n_k values (a k set per time).
 * Sums are compensated long double; cos is libm's double cos.
 */
int orc_pfgse(const uint64_t *blocks, int W, int H, int D, int bc, int shift, int64_t n,
              const int32_t *x0, const int32_t *y0, const int32_t *z0, int32_t *x, int32_t *y,
              int32_t *z, uint64_t *seed, double *energy, uint32_t *hits, const double *penalty,
              int n_sets, const double *set_delta, int n_delta, const int64_t *steps_inc, int n_k,
              int k_per_delta, const double *kx, const double *ky, const double *kz, double voxel,
              double *M, double *msd, int threads)
{
    /* msd (optional, n_delta x 3): per-axis sums of squared displacements
     * (voxel^2) at each diffusion time, the inputs of
     * recoverDmsdWithoutSampling (B/NMR_Simulation/NMR_pfgse.cpp:1098-1160). */
    if (msd)
        memset(msd, 0, sizeof(double) * 3 * (size_t)(n_delta > 0 ? n_delta : 0));
    if (n < 0 || n_delta < 0 || n_k < 0)
        return 1;
    int ns = n_sets > 0 ? n_sets : 1;
    orc_geom g;
    g.blocks = blocks;
    g.W = W;
    g.H = H;
    g.D = D;
    g.bC = (W + 3) / 4;
    g.bR = (H + 3) / 4;
    g.labels = NULL;
    g.bc = bc;
    g.shift = shift;
    int nt = orc_threads(threads);
    size_t per = (size_t)ns * n_delta * n_k;
    orc_acc *acc = (orc_acc *)calloc((size_t)nt * (per ? per : 1), sizeof(orc_acc));
    if (!acc)
        return 2;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        orc_acc *my = acc + (size_t)tid * per;
        double e[64];
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t w = 0; w < n; ++w) {
            int px = x[w], py = y[w], pz = z[w];
            uint64_t s = seed[w];
            uint32_t htot = hits ? hits[w] : 0u;
            for (int q = 0; q < ns; ++q)
                e[q] = (q == 0 && energy) ? energy[w] : 1.0;
            for (int i = 0; i < n_delta; ++i) {
                uint32_t h = 0;
                for (int64_t st = 0; st < steps_inc[i]; ++st)
                    if (orc_step(&g, &px, &py, &pz, &s, 0) >= 0)
                        ++h;
                htot += h;
                for (int q = 0; q < ns; ++q) {
                    double d = n_sets > 0 ? set_delta[q] : penalty[w];
                    e[q] = e[q] * pow(d, (double)h);
                }
                double dX = (double)(px - x0[w]);
                double dY = (double)(py - y0[w]);
                double dZ = (double)(pz - z0[w]);
                if (msd) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                    msd[3 * i + 0] += dX * dX;
#ifdef _OPENMP
#pragma omp atomic
#endif
                    msd[3 * i + 1] += dY * dY;
#ifdef _OPENMP
#pragma omp atomic
#endif
                    msd[3 * i + 2] += dZ * dZ;
                }
                for (int k = 0; k < n_k; ++k) {
                    const size_t kk = (k_per_delta ? (size_t)i * n_k : 0) + (size_t)k;
                    double c = cos((kx[kk] * dX + ky[kk] * dY + kz[kk] * dZ) * voxel);
                    for (int q = 0; q < ns; ++q)
                        orc_acc_add(&my[((size_t)q * n_delta + i) * n_k + k], c * e[q]);
                }
            }
            x[w] = px;
            y[w] = py;
            z[w] = pz;
            seed[w] = s;
            if (hits)
                hits[w] = htot;
            if (energy)
                energy[w] = e[0];
        }
    }
    for (size_t j = 0; j < per; ++j) {
        orc_acc tot = {0.0L, 0.0L};
        for (int t = 0; t < nt; ++t) {
            orc_acc_add(&tot, acc[(size_t)t * per + j].s);
            orc_acc_add(&tot, acc[(size_t)t * per + j].c);
        }
        M[j] = (double)(tot.s + tot.c);
    }
    free(acc);
    return 0;
}

/* ------------------------------------------------------------ mt19937 */

/* 32-bit Mersenne Twister (Matsumoto & Nishimura 1998) with the standard
 * linear seeding; the reference seeds one with the master seed and discards
 * 4096 outputs (B/NMR_Simulation/NMR_Simulation.cpp:86-95). */
typedef struct {
    uint32_t mt[624];
    int idx;
} orc_mt19937;

void orc_mt_seed(orc_mt19937 *st, uint32_t seed)
{
    st->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        st->mt[i] = 1812433253u * (st->mt[i - 1] ^ (st->mt[i - 1] >> 30)) + (uint32_t)i;
    st->idx = 624;
}

uint32_t orc_mt_next(orc_mt19937 *st)
{
    if (st->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t yv = (st->mt[i] & 0x80000000u) | (st->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = st->mt[(i + 397) % 624] ^ (yv >> 1);
            if (yv & 1u)
                v ^= 0x9908b0dfu;
            st->mt[i] = v;
        }
        st->idx = 0;
    }
    uint32_t v = st->mt[st->idx++];
    v ^= v >> 11;
    v ^= (v << 7) & 0x9d2c5680u;
    v ^= (v << 15) & 0xefc60000u;
    v ^= v >> 18;
    return v;
}

size_t orc_mt_state_size(void) { return sizeof(orc_mt19937); }

/* Lemire (2019) nearly-divisionless bounded draw in [0, range) from a 32-bit
 * generator, the method libstdc++ >= 11 uses for uniform_int_distribution
 * over a 32-bit engine.  range 0 means the full 2^32. */
uint32_t orc_lemire_draw(orc_mt19937 *st, uint32_t range)
{
    if (range == 0)
        return orc_mt_next(st);
    uint64_t m = (uint64_t)orc_mt_next(st) * (uint64_t)range;
    uint32_t lo = (uint32_t)m;
    if (lo < range) {
        uint32_t t = (uint32_t)(0u - range) % range;
        while (lo < t) {
            m = (uint64_t)orc_mt_next(st) * (uint64_t)range;
            lo = (uint32_t)m;
        }
    }
    return (uint32_t)(m >> 32);
}

/*
 * Walker seeds and start voxels the way the reference draws them
 * (B/NMR_Simulation/NMR_Simulation.cpp:86-95, 1139-1204, 1086-1136, 1271-1333):
 * one mt19937 seeded with master_seed mod 2^32, 4096 outputs discarded; then,
 * for each walker in order, seed = (first output << 32) | second output; then
 * n pore indices drawn uniformly from [0, n_pores_cap] inclusive (Lemire);
 * an index past the end of the pore list selects the last pore.
 * pores_xyz holds the pore list (x,y,z triples, raster order z,y,x).
 */
void orc_seed_and_place(uint64_t master_seed, int64_t n, uint32_t n_pores_cap, const int32_t *pores_xyz,
                        int64_t n_pores, uint64_t *seeds, int32_t *x, int32_t *y, int32_t *z)
{
    orc_mt19937 st;
    orc_mt_seed(&st, (uint32_t)master_seed);
    for (int i = 0; i < 4096; ++i)
        orc_mt_next(&st);
    for (int64_t w = 0; w < n; ++w) {
        uint64_t hi = orc_mt_next(&st);
        uint64_t lo = orc_mt_next(&st);
        seeds[w] = (hi << 32) | lo;
    }
    for (int64_t w = 0; w < n; ++w) {
        uint32_t idx = orc_lemire_draw(&st, n_pores_cap + 1u);
        if ((int64_t)idx >= n_pores)
            idx = (uint32_t)(n_pores - 1);
        x[w] = pores_xyz[3 * (size_t)idx + 0];
        y[w] = pores_xyz[3 * (size_t)idx + 1];
        z[w] = pores_xyz[3 * (size_t)idx + 2];
    }
}
