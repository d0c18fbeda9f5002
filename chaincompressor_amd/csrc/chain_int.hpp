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

// lanes of an RVC1 section: the largest power of two <= lanes_max that
// leaves at least 256 symbols per lane (1 for short sections)
CH_HD inline uint32_t rvc1_lanes_for(uint64_t n, uint32_t lanes_max) {
    uint32_t L = lanes_max ? lanes_max : 1;
    while (L > 1 && n < (uint64_t)L * 256u) L >>= 1;
    return L;
}


}  // namespace chain
