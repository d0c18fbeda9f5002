// Integer arithmetic of the Chain engine shared by device kernels, host code
// and host-only tools (no HIP headers needed): bit lengths, division by p,
// the RVN1 precision and normalisation rules of cpp/chain/lcps_cpp.cpp:539-684
// in exact integer form, and exact division by multiplication.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define CH_HD __host__ __device__
#else
#define CH_HD
#endif

namespace chain {

// q / p for q < 2^63 (Granlund & Montgomery 1994, Thm 4.2 with N = 63):
// m = ceil(2^(63+l)/p) fits 64 bits and floor(q/p) = mulhi(q, m) >> (l-1).
CH_HD inline uint64_t div_p(uint64_t q, uint64_t p, uint64_t magic, uint32_t l) {
    if (l == 0) return q;
#ifdef __HIP_DEVICE_COMPILE__
    return __umul64hi(q, magic) >> (l - 1);
#else
    return (uint64_t)(((unsigned __int128)q * magic) >> 64) >> (l - 1);
#endif
}

CH_HD inline int bitlen64(uint64_t v) {
#ifdef __HIP_DEVICE_COMPILE__
    return v ? 64 - __clzll((long long)v) : 0;
#else
    return v ? 64 - __builtin_clzll(v) : 0;
#endif
}



#define RANS_VALUE_LIMIT 4096u   // RVC1 value sections: symbols < 4096
#define RANS_TAG_LIMIT 64u       // Tagged widths are 1..63

// Exact floor(x / f) for 0 <= x < 2^31 by one multiplication
// (Granlund & Montgomery 1994, Thm 4.2 with N = 31): l = ceil(log2 f),
// m = ceil(2^(31+l) / f) < 2^33, floor(x / f) = (x * m) >> (31 + l).
// The rANS encoder divides its state (< 1024 f <= 2^30) by frequencies f <= 2^20.
CH_HD inline uint32_t rcp31_shift(uint32_t f) {
#ifdef __HIP_DEVICE_COMPILE__
    return f > 1 ? 32u - (uint32_t)__clz(f - 1) : 0u;
#else
    return f > 1 ? 32u - (uint32_t)__builtin_clz(f - 1) : 0u;
#endif
}
CH_HD inline uint64_t rcp31_mult(uint32_t f, uint32_t l) {
    return ((1ull << (31 + l)) + f - 1) / f;
}
CH_HD inline uint32_t rcp31_div(uint32_t x, uint64_t m, uint32_t l) {
    return (uint32_t)(((uint64_t)x * m) >> (31 + l));
}

// lcps_cpp.cpp:541-552
CH_HD inline int rvn1_precision_bits(uint64_t unique) {
    int ub = bitlen64(unique) + 2;
    if (ub > 18) ub = 18;
    int pb = (3 * ub) / 2;
    if (pb < 12) pb = 12;
    if (pb > 20) pb = 20;
    return pb;
}

// floor((long double)f * M / n + 0.5L) of lcps_cpp.cpp:596-601. With n < 2^42
// the two roundings cannot cross an integer (the gap to the next integer of
// (2fM + n) / 2n is at least 1/2n > 2^-43, the rounding error below 2^-43),
// so the value equals the exact integer expression.
CH_HD inline uint32_t rvn1_first_prob(uint64_t f, uint32_t pb, uint64_t n) {
    const uint64_t num = 2 * (f << pb) + n;
    return (uint32_t)(num / (2 * n));
}

// 128-by-32 division helper: (hi:lo) / d with d < 2^32 and quotient < 2^64
CH_HD inline uint64_t div128_small(uint64_t hi, uint64_t lo, uint64_t d, uint64_t* rem) {
    uint64_t r = hi % d;
    uint64_t t = (r << 32) | (lo >> 32);
    const uint64_t q1 = t / d;
    r = t % d;
    t = (r << 32) | (lo & 0xffffffffull);
    const uint64_t q0 = t / d;
    *rem = t % d;
    return (q1 << 32) | q0;
}

CH_HD inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// floor(RN64(RN64(2^m / S) * p)): the reduction step of lcps_cpp.cpp:624-641,
// where x87 long double rounds each operation to a 64-bit significand,
// ties to even. Requires 0 < S < 2^32, p < 2^32.
CH_HD inline uint64_t rvn1_reduced_prob(int m, uint64_t S, uint64_t p) {
    // relative = sig * 2^e, 64-bit significand
    const int bs = bitlen64(S);
    uint64_t sig;
    int e;
    if ((S & (S - 1)) == 0) {
        sig = 1ull << 63;
        e = m - (bs - 1) - 63;
    } else {
        // 2^(63+bs) / S lies in (2^63, 2^64); bs <= 32 so the numerator is hi:lo
        const int k = 63 + bs;
        const uint64_t hi = k >= 64 ? (1ull << (k - 64)) : 0;
        const uint64_t lo = k >= 64 ? 0 : (1ull << k);
        uint64_t r;
        uint64_t q = div128_small(hi, lo, S, &r);
        e = m - 63 - bs;
        if (2 * r > S || (2 * r == S && (q & 1))) {
            q += 1;
            if (q == 0) { q = 1ull << 63; e += 1; }
        }
        sig = q;
    }
    // product sig * p, rounded to 64 significant bits
    const uint64_t lo = sig * p, hi = mulhi64(sig, p);
    if (hi != 0) {
        const int s = bitlen64(hi);   // 1..32
        uint64_t q = (lo >> s) | (hi << (64 - s));
        const uint64_t rem = lo & ((1ull << s) - 1);
        const uint64_t half = 1ull << (s - 1);
        int ss = s;
        if (rem > half || (rem == half && (q & 1))) {
            q += 1;
            if (q == 0) { q = 1ull << 63; ss += 1; }
        }
        sig = q;
        e += ss;
    } else {
        sig = lo;
    }
    if (e >= 0) return e >= 64 ? ~0ull : (sig << e);
    if (e <= -64) return 0;
    return sig >> (-e);
}

// ------------------------------------------------------------------------
// Size estimates of the LCH1 chooser (this project's container, so the rule
// is its own): a RVC1 section is estimated from its symbol counts and
// normalised probabilities as sum(count * (pb - log2 f)) bits plus the
// chunks' state and size words; the candidates whose estimate lies within
// 1/256 (+ 64 bytes) of the smallest are encoded exactly and the smallest of
// those wins. Everything is integer arithmetic, so the device and the oracle
// take the same decisions.

// log2(1 + k/256) in 1/65536, rounded to nearest
static constexpr uint16_t kLog2Frac[256] = {
        0,   369,   736,  1102,  1466,  1829,  2190,  2551,  2909,  3267,  3623,  3978,
     4331,  4683,  5034,  5384,  5732,  6079,  6425,  6769,  7112,  7454,  7795,  8134,
     8473,  8810,  9146,  9480,  9814, 10146, 10477, 10807, 11136, 11464, 11791, 12116,
    12440, 12764, 13086, 13407, 13727, 14046, 14363, 14680, 14996, 15310, 15624, 15937,
    16248, 16559, 16868, 17177, 17484, 17791, 18096, 18401, 18704, 19007, 19308, 19609,
    19909, 20207, 20505, 20802, 21098, 21393, 21687, 21980, 22272, 22564, 22854, 23144,
    23433, 23720, 24007, 24293, 24579, 24863, 25146, 25429, 25711, 25992, 26272, 26551,
    26830, 27108, 27384, 27660, 27936, 28210, 28484, 28757, 29029, 29300, 29571, 29840,
    30109, 30378, 30645, 30912, 31178, 31443, 31707, 31971, 32234, 32496, 32758, 33019,
    33279, 33538, 33797, 34055, 34312, 34569, 34825, 35080, 35334, 35588, 35841, 36094,
    36346, 36597, 36847, 37097, 37346, 37595, 37842, 38090, 38336, 38582, 38827, 39072,
    39316, 39559, 39802, 40044, 40286, 40527, 40767, 41006, 41246, 41484, 41722, 41959,
    42196, 42432, 42667, 42902, 43137, 43370, 43603, 43836, 44068, 44300, 44530, 44761,
    44990, 45220, 45448, 45676, 45904, 46131, 46357, 46583, 46809, 47034, 47258, 47482,
    47705, 47928, 48150, 48372, 48593, 48813, 49034, 49253, 49472, 49691, 49909, 50127,
    50344, 50560, 50776, 50992, 51207, 51422, 51636, 51850, 52063, 52276, 52488, 52700,
    52911, 53122, 53332, 53542, 53751, 53960, 54169, 54377, 54584, 54791, 54998, 55204,
    55410, 55615, 55820, 56025, 56229, 56432, 56635, 56838, 57040, 57242, 57443, 57644,
    57845, 58045, 58245, 58444, 58643, 58841, 59039, 59237, 59434, 59631, 59827, 60023,
    60219, 60414, 60609, 60803, 60997, 61190, 61384, 61576, 61769, 61961, 62152, 62343,
    62534, 62725, 62915, 63104, 63294, 63483, 63671, 63859, 64047, 64234, 64421, 64608,
    64794, 64980, 65166, 65351,
};

// log2(f) in 1/65536 for f >= 1: the exponent plus the table entry of the 8
// bits below the leading one (lower bits dropped)
CH_HD inline uint32_t log2_q16(uint32_t f) {
    const int e = bitlen64(f) - 1;
    const uint32_t k = e >= 8 ? (f >> (e - 8)) & 255u : (f << (8 - e)) & 255u;
    return ((uint32_t)e << 16) + kLog2Frac[k];
}

// estimated cost of one symbol of probability f / 2^pb, in 1/65536 bits
CH_HD inline uint32_t rvc1_symbol_cost_q16(uint32_t f, uint32_t pb) {
    return (pb << 16) - log2_q16(f);
}

// estimated bytes of a RVC1 section: the exact header, table and chunk size
// words, the chunks' lane states and the estimated stream bits rounded up
CH_HD inline uint64_t rvc1_section_estimate(uint64_t count, uint64_t table_len, uint32_t lanes,
                                            uint32_t chunk_symbols, uint64_t bits_q16) {
    if (count == 0) return 30;
    const uint64_t nchunks = (count + chunk_symbols - 1) / chunk_symbols;
    return 30 + table_len + 4 * nchunks + 4ull * lanes * nchunks + ((bits_q16 + (8ull << 16) - 1) >> 19);
}

// contender mask of the LCH1 chooser: bit 0 Tagged, bit 1 Fixed, bit 2
// Raw-rANS (rans_est = ~0 when ineligible); Fixed sizes are exact
CH_HD inline uint32_t lch1_contenders(uint64_t tagr_est, uint64_t fixp, uint64_t rans_est) {
    uint64_t best = tagr_est < fixp ? tagr_est : fixp;
    best = rans_est < best ? rans_est : best;
    const uint64_t lim = best + (best >> 8) + 64;
    return (tagr_est <= lim ? 1u : 0u) | (fixp <= lim ? 2u : 0u) | (rans_est <= lim ? 4u : 0u);
}

// lanes of an RVC1 section: the largest power of two <= lanes_max that
// leaves at least 256 symbols per lane (1 for short sections)
CH_HD inline uint32_t rvc1_lanes_for(uint64_t n, uint32_t lanes_max) {
    uint32_t L = lanes_max ? lanes_max : 1;
    while (L > 1 && n < (uint64_t)L * 256u) L >>= 1;
    return L;
}


}  // namespace chain
