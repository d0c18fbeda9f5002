// CPU check of the engine's integer emulation of the reference's x87
// long-double rounding (chaincompressor_amd/csrc/chain_int.hpp) against
// real long double arithmetic, as cpp/chain/lcps_cpp.cpp:596-641 computes it.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

#include "../../chaincompressor_amd/csrc/chain_int.hpp"

int main(int argc, char** argv) {
    const uint64_t iters = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 2000000;
    std::mt19937_64 rng(12345);
    uint64_t bad = 0;
    for (uint64_t it = 0; it < iters; ++it) {
        const int m = 12 + (int)(rng() % 9);
        const uint64_t M = 1ull << m;
        // S > M (the reduction only runs when the sum exceeds the precision)
        const uint64_t S = M + 1 + rng() % (it % 3 == 0 ? 64 : (1ull << 18));
        const uint64_t p = 1 + rng() % M;
        const long double relative = (long double)M / (long double)S;
        const uint64_t want = (uint64_t)std::floor(relative * (uint32_t)p);
        const uint64_t got = chain::rvn1_reduced_prob(m, S, p);
        if (want != got && bad++ < 10) std::printf("reduce m=%d S=%llu p=%llu want=%llu got=%llu\n", m,
                                                 (unsigned long long)S, (unsigned long long)p,
                                                 (unsigned long long)want, (unsigned long long)got);
        // first rounding
        const uint64_t n = 1 + rng() % (it % 2 ? 5000 : (1ull << 34));
        const uint64_t f = rng() % (n + 1);
        const uint32_t w1 = (uint32_t)std::floor((long double)f * (uint32_t)M / (long double)n + 0.5L);
        const uint32_t g1 = chain::rvn1_first_prob(f, m, n);
        if (w1 != g1 && bad++ < 10) std::printf("first f=%llu n=%llu m=%d want=%u got=%u\n", (unsigned long long)f,
                                              (unsigned long long)n, m, w1, g1);
    }
    // exhaustive small corner: S just above M, every p
    for (int m = 12; m <= 20; m += 4)
        for (uint64_t S = (1ull << m) + 1; S < (1ull << m) + 40; ++S)
            for (uint64_t p = 1; p <= (1ull << m); p += 1 + (p >> 4)) {
                const long double relative = (long double)(1ull << m) / (long double)S;
                const uint64_t want = (uint64_t)std::floor(relative * (uint32_t)p);
                if (want != chain::rvn1_reduced_prob(m, S, p) && bad++ < 10) std::printf("corner m=%d S=%llu p=%llu\n", m, (unsigned long long)S, (unsigned long long)p);
            }
    std::printf("mismatches %llu\n", (unsigned long long)bad);
    return bad != 0;
}
