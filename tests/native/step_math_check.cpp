// Host build of the kernels' step arithmetic (randomwalk_nmr_amd/csrc/walk_common.h)
// checked exhaustively on small domains against plain definitions.
// Prints "OK" and exits 0, or prints the first failures and exits 1.
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "../../randomwalk_nmr_amd/csrc/walk_common.h"

using namespace rwnmr;

static int fails = 0;
#define EXPECT(c, ...)                          \
    do {                                        \
        if (!(c)) {                             \
            if (fails < 20)                     \
                printf(__VA_ARGS__);            \
            ++fails;                            \
        }                                       \
    } while (0)

static int textbook_reflect(int g, int d)
{
    // reflective tiling of period 2d: 0..d-1, d-1..0, 0..d-1, ...
    int t = g % (2 * d);
    if (t < 0)
        t += 2 * d;
    return t < d ? t : 2 * d - 1 - t;
}

// Unfolded-image arithmetic (walk_common.h, "Unfolded image with dilated
// voxel addresses") against direct coordinate arithmetic.
static void check_unfolded()
{
    // s % 3 by byte sum and a 24-bit multiply; the table index pairs with s % 6
    for (uint32_t sb = 0; sb <= 2040; ++sb)
        EXPECT(((sb * 2731u) >> 13) == sb / 3u, "q sb=%u\n", sb);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < 5000000; ++i) {
        s = xorshift64(s);
        EXPECT(zbyte_sum(s) % 3u == s % 3u, "byte sum %llu\n", (unsigned long long)s);
        EXPECT(zdir_residue(zdir_index_from_sum(zbyte_sum(s), (uint32_t)s)) == s % 6u, "residue %llu\n",
               (unsigned long long)s);
    }
    for (uint64_t v = 0; v < 100000; ++v)
        EXPECT(zdir_residue(zdir_index_from_sum(zbyte_sum(v), (uint32_t)v)) == v % 6u, "residue small\n");
    for (uint64_t v : {~0ull, 1ull << 63, 0xFFFFFFFFull, 0x8080808080808080ull})
        EXPECT(zdir_residue(zdir_index_from_sum(zbyte_sum(v), (uint32_t)v)) == v % 6u, "residue edge\n");
    // the dot-product key: congruent to s mod 6, at most kZKeyMax, and its
    // table index names the entry of residue s % 6
    auto dot_key = [](uint64_t v) {
        return zdot4((uint32_t)v, kZKeyWeightsLo, zdot4((uint32_t)(v >> 32), kZKeyWeightsHi, 0u));
    };
    s = 0x2545F4914F6CDD1Dull;
    uint32_t kmax = 0;
    for (int i = 0; i < 5000000; ++i) {
        s = xorshift64(s);
        const uint32_t k = dot_key(s);
        kmax = k > kmax ? k : kmax;
        EXPECT(k % 6u == s % 6u, "dot key %llu\n", (unsigned long long)s);
    }
    for (uint64_t v = 0; v < 100000; ++v)
        EXPECT(dot_key(v) % 6u == v % 6u, "dot key small\n");
    for (uint64_t v : {~0ull, ~0ull - 1, 1ull << 63, 0xFFFFFFFFull, 0xFF00000000000000ull, 0x8080808080808080ull,
                       0x0101010101010101ull, 0xFFull}) {
        EXPECT(dot_key(v) % 6u == v % 6u, "dot key edge %llx\n", (unsigned long long)v);
        EXPECT(dot_key(v) <= kZKeyMax, "dot key range %llx\n", (unsigned long long)v);
    }
    EXPECT(dot_key(~0ull) == kZKeyMax, "dot key max\n");
    EXPECT(kmax <= kZKeyMax, "dot key stream max %u\n", kmax);
    for (uint32_t k = 0; k <= kZKeyMax; ++k)
        EXPECT(zdir_residue(zkey_index(k)) == k % 6u, "key index %u\n", k);
    // zdot4 is the byte dot product
    for (uint32_t a : {0u, 0xFFFFFFFFu, 0x01020304u, 0x80FF7F01u})
        for (uint32_t w : {0u, 0xFFFFFFFFu, kZKeyWeightsLo, kZKeyWeightsHi}) {
            uint32_t ref = 7u;
            for (int i = 0; i < 4; ++i)
                ref += ((a >> (8 * i)) & 255u) * ((w >> (8 * i)) & 255u);
            EXPECT(zdot4(a, w, 7u) == ref, "zdot4\n");
        }
    // unfold == the reference folds, for any unbounded coordinate
    for (int d = 1; d <= 13; ++d)
        for (int g = -100; g <= 100; ++g) {
            EXPECT(zunfold(floor_mod(g, 2u * d), d, 1) == fold_mirror(g, d), "unfold mirror g=%d d=%d\n", g, d);
            EXPECT(zunfold(floor_mod(g, (uint32_t)d), d, 0) == fold_periodic(g, d), "unfold periodic\n");
        }
    // layouts: disjoint masks covering the address bits; every move of every
    // voxel lands on the wrapped neighbour
    const uint32_t periods[][3] = {{1, 1, 1}, {2, 2, 2}, {8, 8, 8}, {16, 4, 2}, {5, 7, 3}, {10, 14, 6},
                                   {12, 2, 9}, {1, 6, 4}, {26, 18, 22}, {32, 32, 1}};
    for (const auto &p : periods) {
        ZLayout L{};
        EXPECT(zlayout_make(p[0], p[1], p[2], L) == 1, "layout\n");
        EXPECT((L.m[0] & L.m[1]) == 0 && (L.m[0] & L.m[2]) == 0 && (L.m[1] & L.m[2]) == 0, "disjoint\n");
        const uint64_t all = (uint64_t)L.m[0] | L.m[1] | L.m[2];
        EXPECT(all == (L.bits == 32 ? 0xFFFFFFFFull : ((1ull << L.bits) - 1)), "cover\n");
        for (uint32_t z = 0; z < p[2]; ++z)
            for (uint32_t y = 0; y < p[1]; ++y)
                for (uint32_t x = 0; x < p[0]; ++x) {
                    const uint32_t u[3] = {x, y, z};
                    const uint32_t A = zdeposit(x, L.m[0]) | zdeposit(y, L.m[1]) | zdeposit(z, L.m[2]);
                    for (int a = 0; a < 3; ++a)
                        EXPECT(zextract(A, L.m[a]) == u[a], "extract\n");
                    for (uint32_t idx = 0; idx < 6; ++idx) {
                        const ZDir dir = zdir_make(L, idx);
                        const uint32_t r = zdir_residue(idx);
                        const int ax = move_axis(r), sg = move_sign(r);
                        uint32_t v[3] = {x, y, z};
                        v[ax] = (uint32_t)(((int)v[ax] + sg + (int)p[ax]) % (int)p[ax]);
                        const uint32_t want = zdeposit(v[0], L.m[0]) | zdeposit(v[1], L.m[1]) | zdeposit(v[2], L.m[2]);
                        EXPECT(zmove(A, dir, (int)L.wrapfix) == want, "zmove p=%u,%u,%u u=%u,%u,%u idx=%u\n", p[0],
                               p[1], p[2], x, y, z, idx);
                        EXPECT(dir.dp == sg * (ax == 0 ? 1 : ax == 1 ? 1024 : 1 << 20), "dp\n");
                    }
                }
    }
    // packed displacement round trip
    uint64_t t = 777;
    for (int i = 0; i < 200000; ++i) {
        t = xorshift64(t);
        const int n = (int)(t % 512u);
        int dx = 0, dy = 0, dz = 0, pk = 0;
        for (int j = 0; j < n; ++j) {
            t = xorshift64(t);
            const uint32_t r = (uint32_t)(t % 6u);
            const int ax = move_axis(r), sg = move_sign(r);
            (ax == 0 ? dx : ax == 1 ? dy : dz) += sg;
            pk += sg * (ax == 0 ? 1 : ax == 1 ? 1024 : 1 << 20);
        }
        int ux, uy, uz;
        zunpack(pk, ux, uy, uz);
        EXPECT(ux == dx && uy == dy && uz == dz, "unpack %d %d %d\n", dx, dy, dz);
    }
    // host walk on the unfolded volume == walk with the reference fold per step
    for (int mirror = 0; mirror <= 1; ++mirror) {
        const int W = 7, H = 5, D = 6;
        uint8_t solid[D][H][W];
        uint64_t g = 99 + mirror;
        for (int z = 0; z < D; ++z)
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    g = xorshift64(g);
                    solid[z][y][x] = (g % 10u) < 3u;
                }
        solid[2][2][3] = 0;
        const uint32_t f = mirror ? 2u : 1u;
        ZLayout L{};
        zlayout_make(f * W, f * H, f * D, L);
        std::vector<uint32_t> vol((L.bits > 5 ? (1u << (L.bits - 5)) : 1u), 0u);
        for (uint64_t A = 0; A < (1ull << L.bits); ++A) {
            const uint32_t ux = zextract((uint32_t)A, L.m[0]), uy = zextract((uint32_t)A, L.m[1]),
                           uz = zextract((uint32_t)A, L.m[2]);
            uint32_t bit = 1;
            if (ux < L.period[0] && uy < L.period[1] && uz < L.period[2])
                bit = solid[zunfold(uz, D, mirror)][zunfold(uy, H, mirror)][zunfold(ux, W, mirror)];
            vol[A >> 5] |= bit << (A & 31);
        }
        for (int w = 0; w < 50; ++w) {
            uint64_t sa = 1000 + w, sb = sa;
            int X = 3, Y = 2, Z = 2, hx = 0, hy = 0, hz = 0;   // folded walker: unbounded X, Y, Z
            uint32_t A = zdeposit(floor_mod(3, L.period[0]), L.m[0]) | zdeposit(2, L.m[1]) | zdeposit(2, L.m[2]);
            int pk = 0, hitsA = 0, hitsB = 0;
            for (int st = 0; st < 3000; ++st) {
                sa = xorshift64(sa);
                const uint32_t r = mod6(sa);
                const int ax = move_axis(r), sg = move_sign(r);
                const int nx = X + (ax == 0 ? sg : 0), ny = Y + (ax == 1 ? sg : 0), nz = Z + (ax == 2 ? sg : 0);
                const int fx = mirror ? fold_mirror(nx, W) : fold_periodic(nx, W);
                const int fy = mirror ? fold_mirror(ny, H) : fold_periodic(ny, H);
                const int fz = mirror ? fold_mirror(nz, D) : fold_periodic(nz, D);
                if (solid[fz][fy][fx])
                    ++hitsA;
                else
                    X = nx, Y = ny, Z = nz;
                sb = xorshift64(sb);
                const uint32_t idx = zdir_index_from_sum(zbyte_sum(sb), (uint32_t)sb);
                const ZDir dir = zdir_make(L, idx);
                const uint32_t c = zmove(A, dir, (int)L.wrapfix);
                if ((vol[c >> 5] >> (c & 31)) & 1u)
                    ++hitsB;
                else
                    A = c, pk += dir.dp;
                if (st % 500 == 499) {
                    int dx, dy, dz;
                    zunpack(pk, dx, dy, dz);
                    hx += dx, hy += dy, hz += dz;
                    pk = 0;
                }
            }
            EXPECT(hitsA == hitsB && X == 3 + hx && Y == 2 + hy && Z == 2 + hz, "walk mirror=%d w=%d\n", mirror, w);
        }
    }
}

int main()
{
    // mod6 against the C remainder: xorshift stream, edges, small values
    uint64_t s = 88172645463325252ull;
    for (int i = 0; i < 20000000; ++i) {
        s = xorshift64(s);
        EXPECT(mod6(s) == s % 6u, "mod6 %llu\n", (unsigned long long)s);
    }
    for (uint64_t v = 0; v < 1000000; ++v)
        EXPECT(mod6(v) == v % 6u, "mod6 small %llu\n", (unsigned long long)v);
    // xorshift64_inverse undoes one step: along the stream, on small values
    // and single bits
    s = 1234567891011ull;
    for (int i = 0; i < 2000000; ++i) {
        const uint64_t t = xorshift64(s);
        EXPECT(xorshift64_inverse(t) == s, "inverse %llu\n", (unsigned long long)s);
        s = t;
    }
    for (uint64_t v = 0; v < 100000; ++v)
        EXPECT(xorshift64_inverse(xorshift64(v)) == v, "inverse small %llu\n", (unsigned long long)v);
    for (int b = 0; b < 64; ++b)
        EXPECT(xorshift64(xorshift64_inverse(1ull << b)) == (1ull << b), "inverse bit %d\n", b);
    const uint64_t edges[] = {~0ull, ~0ull - 1, 1ull << 63, (1ull << 63) - 1, 0xAAAAAAAAAAAAAAAAull,
                              0x5555555555555555ull, 0xFFFFFFFFull, 0x100000000ull};
    for (uint64_t v : edges)
        EXPECT(mod6(v) == v % 6u, "mod6 edge %llu\n", (unsigned long long)v);
    // residue -> move (B/Walker/walker_defs.h:6-23, walker.h:360-391)
    const int ax[6] = {1, 0, 1, 0, 2, 2}, sg[6] = {-1, -1, +1, +1, +1, -1};
    for (uint32_t r = 0; r < 6; ++r) {
        EXPECT(move_axis(r) == ax[r], "axis %u\n", r);
        EXPECT(move_sign(r) == sg[r], "sign %u\n", r);
    }
    // folds and the incremental mirror/periodic steps
    for (int d = 1; d <= 13; ++d)
        for (int g = -60; g <= 60; ++g) {
            EXPECT(fold_mirror(g, d) == textbook_reflect(g, d), "fold_mirror g=%d d=%d\n", g, d);
            int p = g % d;
            if (p < 0)
                p += d;
            EXPECT(fold_periodic(g, d) == p, "fold_periodic g=%d d=%d\n", g, d);
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                int ni, nm;
                mirror_step(fold_mirror(g, d), mirror_flag(g, d), sgn, d, ni, nm);
                EXPECT(ni == fold_mirror(g + sgn, d), "mirror_step i g=%d d=%d s=%d\n", g, d, sgn);
                EXPECT(nm == mirror_flag(g + sgn, d) || d == 1, "mirror_step m g=%d d=%d s=%d\n", g, d, sgn);
                EXPECT(periodic_step(fold_periodic(g, d), sgn, d) == fold_periodic(g + sgn, d),
                       "periodic_step g=%d d=%d\n", g, d);
            }
        }
    // a long mirror walk tracked incrementally equals the direct fold
    for (int d = 1; d <= 9; ++d) {
        int g = 3, i = fold_mirror(g, d), m = mirror_flag(g, d);
        uint64_t st = 12345 + d;
        for (int t = 0; t < 200000; ++t) {
            st = xorshift64(st);
            const int sgn = (st >> 17) & 1 ? 1 : -1;
            int ni, nm;
            mirror_step(i, m, sgn, d, ni, nm);
            g += sgn;
            i = ni;
            m = nm;
            EXPECT(i == fold_mirror(g, d), "walk d=%d t=%d\n", d, t);
        }
    }
    // bitBlock addressing (B/BitBlock/bitBlock.h:113-134)
    for (int z = 0; z < 9; ++z)
        for (int y = 0; y < 9; ++y)
            for (int x = 0; x < 9; ++x) {
                EXPECT(block_index(x, y, z, 3, 9) == (uint32_t)(x / 4 + (y / 4) * 3 + (z / 4) * 9), "blk\n");
                EXPECT(bit_index(x, y, z) == (uint32_t)(x % 4 + 4 * (y % 4) + 16 * (z % 4)), "bit\n");
            }
    check_unfolded();
    if (fails) {
        printf("%d failures\n", fails);
        return 1;
    }
    printf("OK\n");
    return 0;
}
