// Host build of the kernels' step arithmetic (randomwalk_nmr_amd/csrc/walk_common.h)
// checked exhaustively on small domains against plain definitions.
// Prints "OK" and exits 0, or prints the first failures and exits 1.
#include <stdint.h>
#include <stdio.h>

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
    if (fails) {
        printf("%d failures\n", fails);
        return 1;
    }
    printf("OK\n");
    return 0;
}
