// Draws of libstdc++'s std::uniform_int_distribution over std::mt19937, the
// way the reference draws walker seeds and pore indices
// (B/NMR_Simulation/NMR_Simulation.cpp:86-95, 1109, 1167): prints
// argv-driven draws one per line.
// usage: lemire_check <seed> <n_seeds> <cap> <n_draws>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <random>

int main(int argc, char **argv)
{
    if (argc != 5)
        return 2;
    const unsigned long seed = strtoul(argv[1], nullptr, 10);
    const long n_seeds = strtol(argv[2], nullptr, 10);
    const unsigned long cap = strtoul(argv[3], nullptr, 10);
    const long n_draws = strtol(argv[4], nullptr, 10);
    std::mt19937 g(seed);
    g.discard(4096);
    std::uniform_int_distribution<uint64_t> u64;
    for (long i = 0; i < n_seeds; ++i)
        printf("%llu\n", (unsigned long long)u64(g));
    std::uniform_int_distribution<std::mt19937::result_type> dist(0, cap);
    for (long i = 0; i < n_draws; ++i)
        printf("%lu\n", (unsigned long)dist(g));
    return 0;
}
