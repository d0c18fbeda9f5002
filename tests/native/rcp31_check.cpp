// CPU check of the rANS encoder's division by multiplication
// (rcp31_shift / rcp31_mult / rcp31_div in chaincompressor_amd/csrc/chain_int.hpp).
// (x * m) >> s is nondecreasing in x, so it equals floor(x / f) on all of
// [0, 1024 f) exactly when it does at every quotient boundary k f - 1 and k f;
// those are checked for every frequency 1 <= f <= 2^20.
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "../../chaincompressor_amd/csrc/chain_int.hpp"

int main(int argc, char** argv) {
    const uint32_t fmax = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : (1u << 20);
    uint64_t bad = 0, checks = 0;
    for (uint32_t f = 1; f <= fmax; ++f) {
        const uint32_t l = chain::rcp31_shift(f);
        const uint64_t m = chain::rcp31_mult(f, l);
        const uint64_t xmax = 1024ull * f;
        for (uint64_t k = 0; k <= 1024; ++k) {
            const uint64_t xs[2] = {k * f ? k * f - 1 : 0, k * f};
            for (uint64_t x : xs) {
                if (x >= xmax) continue;
                ++checks;
                if (chain::rcp31_div((uint32_t)x, m, l) != x / f && bad++ < 10)
                    std::printf("f=%u x=%llu\n", f, (unsigned long long)x);
            }
        }
    }
    std::printf("checks=%llu bad=%llu\n", (unsigned long long)checks, (unsigned long long)bad);
    return bad ? 1 : 0;
}
