// Host-side LCP2 export of Chain streams (Zstd-22 + RVN1 chooser).
//
// The GPU engine writes the reference's raw container LCP3 byte for byte
// (cpp/chain/lcps_cpp.cpp:1184-1211, 2306-2329). This library turns such a
// stream into the reference's entropy-coded container LCP2, as
// lcps_compress2(..., ec=True) writes it (cpp/chain/lcps_cpp.cpp:2017-2338):
//   every component becomes the smallest of
//     ZTG4  RVN1-coded widths + Zstd of the MSB-first point-major values (:1093-1138)
//     ZFMX  per-dimension minimum / width + Zstd of the fixed payload     (:856-933)
//     RVMX  RVN1 of the point-major values, when all are < 2^18          (:935-995)
//   ties in that order (:1140-1174); the corrections are Zstd-coded.
// With the same libzstd (the reference was built against zstd 1.4.9) the
// bytes are identical to lcps_compress2's; any libzstd gives a stream the
// reference decoder (lcps_cpu_decompress_serial) reads.
//
// The RVN1 table uses the integer forms of the reference's long-double
// normalisation (chain_int.hpp), the same code the GPU kernels run.
#include "../../include/chain_lcp2.h"

#include <zstd.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "chain_int.hpp"

namespace {

using chain::bitlen64;

constexpr uint64_t kAlphabetLimit = 1u << 18;   // RVN1 symbols, lcps_cpp.cpp:541-552

struct Fail : std::runtime_error {
    int code;
    Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void need(bool ok, const char* msg, int code = CHAIN_LCP2_EFORMAT) {
    if (!ok) throw Fail(code, msg);
}

// ---------------------------------------------------------------------------
// little-endian byte writer / reader

struct Out {
    std::vector<uint8_t> b;
    void raw(const void* p, size_t n) {
        const uint8_t* q = static_cast<const uint8_t*>(p);
        b.insert(b.end(), q, q + n);
    }
    void u8(uint8_t v) { b.push_back(v); }
    void u32(uint32_t v) { for (int i = 0; i < 4; ++i) b.push_back((uint8_t)(v >> (8 * i))); }
    void u64(uint64_t v) { for (int i = 0; i < 8; ++i) b.push_back((uint8_t)(v >> (8 * i))); }
    void blob(const std::vector<uint8_t>& v) {   // u64 length + bytes
        u64(v.size());
        raw(v.data(), v.size());
    }
    void varuint(uint64_t v) {
        while (v >= 0x80) {
            b.push_back((uint8_t)(v | 0x80));
            v >>= 7;
        }
        b.push_back((uint8_t)v);
    }
};

struct In {
    const uint8_t* p;
    size_t pos, end;
    const uint8_t* take(size_t n) {
        need(n <= end - pos, "truncated stream");
        const uint8_t* q = p + pos;
        pos += n;
        return q;
    }
    uint32_t u32() { const uint8_t* q = take(4); uint32_t v = 0; for (int i = 0; i < 4; ++i) v |= (uint32_t)q[i] << (8 * i); return v; }
    uint64_t u64() { const uint8_t* q = take(8); uint64_t v = 0; for (int i = 0; i < 8; ++i) v |= (uint64_t)q[i] << (8 * i); return v; }
};

// ---------------------------------------------------------------------------
// Zstd level 22 (lcps_cpp.cpp:112-132)

std::vector<uint8_t> zstd22(const std::vector<uint8_t>& src, int level) {
    std::vector<uint8_t> out(ZSTD_compressBound(src.size()));
    const size_t k = ZSTD_compress(out.data(), out.size(), src.data(), src.size(), level);
    need(!ZSTD_isError(k), "ZSTD_compress failed", CHAIN_LCP2_EZSTD);
    out.resize(k);
    return out;
}

// ---------------------------------------------------------------------------
// RVN1 value-rANS (lcps_cpp.cpp:539-821)

struct Table {
    int pb = 0;
    std::vector<uint32_t> prob, cum;
};

Table build_table(const std::vector<uint64_t>& v) {
    uint64_t maxv = 0;
    for (uint64_t x : v) maxv = std::max(maxv, x);
    need(maxv < kAlphabetLimit, "RVN1 alphabet exceeds 2^18");
    const uint32_t A = (uint32_t)maxv + 1;
    std::vector<uint64_t> freq(A, 0);
    for (uint64_t x : v) ++freq[x];
    uint64_t unique = 0;
    for (uint64_t f : freq) unique += f != 0;
    Table t;
    t.pb = chain::rvn1_precision_bits(unique);
    const uint64_t M = 1ull << t.pb;
    t.prob.assign(A, 0);
    uint64_t sum = 0;
    for (uint32_t s = 0; s < A; ++s) {
        uint32_t pr = chain::rvn1_first_prob(freq[s], (uint32_t)t.pb, v.size());
        if (pr == 0 && freq[s] != 0) pr = 1;
        t.prob[s] = pr;
        sum += pr;
    }
    if (sum != M) {
        // std::stable_sort by probability: ascending (probability, index)
        std::vector<uint32_t> order(A);
        for (uint32_t s = 0; s < A; ++s) order[s] = s;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return t.prob[a] < t.prob[b]; });
        if (sum < M) {
            t.prob[order[A - 1]] += (uint32_t)(M - sum);
        } else {
            int64_t error = (int64_t)(sum - M);
            while (error > 0) {
                const uint64_t at_start = sum;
                bool changed = false;
                for (uint32_t j = A; j-- > 1 && error > 0;) {
                    uint32_t& pr = t.prob[order[j]];
                    if (pr <= 1) {
                        need(j != A - 1, "rANS normalisation cannot retain every symbol");
                        break;
                    }
                    const uint64_t np = chain::rvn1_reduced_prob(t.pb, at_start, pr);
                    int64_t red = (int64_t)pr - (int64_t)np;
                    if (red == 0) red = 1;
                    if (red >= (int64_t)pr) red = (int64_t)pr - 1;
                    if (red > error) red = error;
                    pr -= (uint32_t)red;
                    sum -= (uint64_t)red;
                    error -= red;
                    changed = true;
                }
                need(changed, "rANS normalisation did not converge");
            }
        }
    }
    t.cum.assign(A, 0);
    uint64_t c = 0;
    for (uint32_t s = 0; s < A; ++s) {
        t.cum[s] = (uint32_t)c;
        c += t.prob[s];
    }
    need(c == M, "rANS probability total mismatch");
    return t;
}

// compact table: varuint(A), then per symbol a zero run or a 6/14/22-bit
// probability with its 2-bit length token (lcps_cpp.cpp:686-729)
std::vector<uint8_t> table_bytes(const Table& t) {
    Out o;
    const uint32_t A = (uint32_t)t.prob.size();
    o.varuint(A);
    for (uint32_t i = 0; i < A; ++i) {
        const uint32_t pr = t.prob[i];
        if (pr == 0) {
            uint32_t run = 0;
            while (run < 63 && i + run + 1 < A && t.prob[i + run + 1] == 0) ++run;
            o.u8((uint8_t)((run << 2) | 3u));
            i += run;
            continue;
        }
        const uint32_t extra = pr >= (1u << 14) ? 2 : pr >= (1u << 6) ? 1 : 0;
        o.u8((uint8_t)((pr << 2) | extra));
        for (uint32_t k = 0; k < extra; ++k) o.u8((uint8_t)(pr >> (8 * (k + 1) - 2)));
    }
    return o.b;
}

// "RVN1" u8 precision u64 count u32 final_state blob(table) u64 n bytes
std::vector<uint8_t> rvn1(const std::vector<uint64_t>& v) {
    Out o;
    o.raw("RVN1", 4);
    if (v.empty()) {
        o.u8(0);
        o.u64(0);
        o.u32(1);
        o.u64(0);
        o.u64(0);
        return o.b;
    }
    const Table t = build_table(v);
    const uint64_t M = 1ull << t.pb, lower = 4 * M;
    uint64_t x = lower;
    std::vector<uint8_t> emitted;   // in encoding order (last symbol first)
    emitted.reserve(v.size());
    for (size_t k = v.size(); k-- > 0;) {
        const uint64_t f = t.prob[v[k]];
        const uint64_t xmax = 1024 * f;   // (lower >> pb) * 256 * f
        while (x >= xmax) {
            emitted.push_back((uint8_t)x);
            x >>= 8;
        }
        x = (x / f) * M + (x % f) + t.cum[v[k]];
    }
    o.u8((uint8_t)t.pb);
    o.u64(v.size());
    o.u32((uint32_t)x);
    o.blob(table_bytes(t));
    o.u64(emitted.size());
    for (size_t k = emitted.size(); k-- > 0;) o.u8(emitted[k]);   // decoder order
    return o.b;
}

// ---------------------------------------------------------------------------
// component candidates

struct Mat {
    std::vector<int64_t> v;   // n x dims, point-major
    uint64_t n = 0;
    uint32_t dims = 0;
};

// MSB-first writer
struct Bits {
    std::vector<uint8_t> b;
    uint64_t nbits = 0;
    void put(uint64_t value, uint32_t width) {
        for (uint32_t k = width; k-- > 0;) {
            if ((nbits & 7) == 0) b.push_back(0);
            if ((value >> k) & 1u) b.back() |= (uint8_t)(0x80u >> (nbits & 7));
            ++nbits;
        }
    }
};

std::vector<uint8_t> ztg4(const Mat& m, int level) {
    std::vector<uint64_t> tags(m.n);
    Bits packed;
    uint32_t mw = 0;
    for (uint64_t r = 0; r < m.n; ++r) {
        uint32_t t = 1;
        for (uint32_t d = 0; d < m.dims; ++d) t = std::max<uint32_t>(t, (uint32_t)bitlen64((uint64_t)m.v[r * m.dims + d]));
        tags[r] = t;
        mw = std::max(mw, t);
        for (uint32_t d = 0; d < m.dims; ++d) packed.put((uint64_t)m.v[r * m.dims + d], t);
    }
    Out o;
    o.raw("ZTG4", 4);
    o.u32(m.dims);
    o.u64(m.n);
    o.u8((uint8_t)mw);
    o.blob(rvn1(tags));
    o.u64(packed.b.size());
    o.blob(zstd22(packed.b, level));
    return o.b;
}

std::vector<uint8_t> zfmx(const Mat& m, int level) {
    std::vector<uint8_t> payload;
    std::vector<int64_t> mins(m.dims, 0);
    std::vector<uint32_t> widths(m.dims, 1);
    std::vector<uint64_t> sizes(m.dims, 0);
    for (uint32_t d = 0; d < m.dims; ++d) {
        int64_t mn = 0, mx = 0;
        for (uint64_t r = 0; r < m.n; ++r) {
            const int64_t x = m.v[r * m.dims + d];
            if (r == 0 || x < mn) mn = x;
            if (r == 0 || x > mx) mx = x;
        }
        const uint32_t w = m.n ? std::max<uint32_t>(1, (uint32_t)bitlen64((uint64_t)(mx - mn))) : 1;
        Bits bits;
        for (uint64_t r = 0; r < m.n; ++r) bits.put((uint64_t)(m.v[r * m.dims + d] - mn), w);
        mins[d] = mn;
        widths[d] = w;
        sizes[d] = bits.b.size();
        payload.insert(payload.end(), bits.b.begin(), bits.b.end());
    }
    Out o;
    o.raw("ZFMX", 4);
    o.u32(m.dims);
    o.u64(m.n);
    for (uint32_t d = 0; d < m.dims; ++d) {
        o.u64((uint64_t)mins[d]);
        o.u8((uint8_t)widths[d]);
        o.u64(sizes[d]);
    }
    o.u64(payload.size());
    o.blob(zstd22(payload, level));
    return o.b;
}

bool rvmx(const Mat& m, std::vector<uint8_t>& out) {
    std::vector<uint64_t> v(m.v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        need(m.v[i] >= 0, "Raw rANS expects non-negative integer values");
        if ((uint64_t)m.v[i] >= kAlphabetLimit) return false;
        v[i] = (uint64_t)m.v[i];
    }
    Out o;
    o.raw("RVMX", 4);
    o.u32(m.dims);
    o.u64(m.n);
    o.blob(rvn1(v));
    out = std::move(o.b);
    return true;
}

// smallest wins; ties: ZTG4, ZFMX, RVMX (lcps_cpp.cpp:1140-1174)
std::vector<uint8_t> choose(const Mat& m, int level) {
    std::vector<uint8_t> best = ztg4(m, level);
    std::vector<uint8_t> cand = zfmx(m, level);
    if (cand.size() < best.size()) best.swap(cand);
    if (rvmx(m, cand) && cand.size() < best.size()) best.swap(cand);
    return best;
}

Mat read_raw1(In r, uint32_t expect_dims) {
    need(std::memcmp(r.take(4), "RAW1", 4) == 0, "LCP3 components must be RAW1 matrices");
    Mat m;
    m.dims = r.u32();
    m.n = r.u64();
    need(m.dims == expect_dims, "RAW1 dimension mismatch");
    need(m.n <= (r.end - r.pos) / 8 / std::max<uint32_t>(1, m.dims), "RAW1 matrix size mismatch");
    m.v.resize(m.n * m.dims);
    for (auto& x : m.v) x = (int64_t)r.u64();
    need(r.pos == r.end, "unexpected trailing RAW1 bytes");
    return m;
}

// ---------------------------------------------------------------------------
// LCP2 -> LCP3 (the decoder side, cpp/decompress/lcps_cpu_decompress_serial.cpp:981-1174)

std::vector<uint8_t> unzstd(const uint8_t* src, size_t n, uint64_t size) {
    need(size <= (1ull << 40), "implausible Zstd output size");
    std::vector<uint8_t> out(size);
    const size_t k = ZSTD_decompress(out.data(), out.size(), src, n);
    need(!ZSTD_isError(k), "ZSTD_decompress failed", CHAIN_LCP2_EZSTD);
    need(k == size, "Zstd output size mismatch");
    return out;
}

In sub_blob(In& r) {
    const uint64_t ln = r.u64();
    need(ln <= r.end - r.pos, "truncated LCP2 section");
    In s{r.p, r.pos, r.pos + ln};
    r.pos += ln;
    return s;
}

uint64_t read_varuint(In& r) {
    uint64_t v = 0;
    for (int shift = 0; shift <= 63; shift += 7) {
        const uint8_t b = *r.take(1);
        need(shift != 63 || (b & 0xfeu) == 0, "LCPS varuint overflow");
        v |= (uint64_t)(b & 0x7fu) << shift;
        if ((b & 0x80u) == 0) return v;
    }
    throw Fail(CHAIN_LCP2_EFORMAT, "invalid LCPS varuint");
}

// RVN1 decoder: table, then one state run forward over the bytes
std::vector<uint64_t> unrvn1(In r, uint64_t expected) {
    need(std::memcmp(r.take(4), "RVN1", 4) == 0, "invalid RVN1 header");
    const uint32_t pb = *r.take(1);
    const uint64_t n = r.u64();
    need(n == expected, "RVN1 symbol count mismatch");
    uint64_t x = r.u32();
    In tr = sub_blob(r);
    const uint64_t nb = r.u64();
    need(nb == r.end - r.pos, "RVN1 byte count mismatch");
    const uint8_t* bytes = r.take(nb);
    std::vector<uint64_t> out(n);
    if (n == 0) return out;
    need(pb >= 12 && pb <= 20, "invalid RVN1 precision");
    const uint64_t A = read_varuint(tr);
    need(A > 0 && A <= kAlphabetLimit, "invalid rANS alphabet");
    std::vector<uint32_t> prob(A, 0);
    for (uint64_t i = 0; i < A;) {
        const uint8_t first = *tr.take(1);
        const uint32_t token = first & 3u;
        if (token == 3) {
            const uint64_t zeros = (first >> 2) + 1u;
            need(zeros <= A - i, "invalid rANS zero run");
            i += zeros;
        } else {
            uint32_t pr = first >> 2;
            for (uint32_t b = 0; b < token; ++b) pr |= (uint32_t)*tr.take(1) << (6 + 8 * b);
            need(pr > 0, "invalid rANS probability");
            prob[i++] = pr;
        }
    }
    need(tr.pos == tr.end, "unexpected trailing RVN1 table bytes");
    const uint64_t M = 1ull << pb, lower = 4 * M;
    std::vector<uint32_t> cum(A + 1, 0), slot(M);
    for (uint64_t s = 0; s < A; ++s) {
        need(prob[s] <= M - cum[s], "invalid rANS probability sum");
        cum[s + 1] = cum[s] + prob[s];
        for (uint32_t k = cum[s]; k < cum[s + 1]; ++k) slot[k] = (uint32_t)s;
    }
    need(cum[A] == M, "incomplete rANS probability table");
    need(x >= lower && x < 256 * lower, "invalid RVN1 state");
    uint64_t pos = 0;
    for (uint64_t k = 0; k < n; ++k) {
        const uint32_t sl = (uint32_t)(x & (M - 1));
        const uint32_t s = slot[sl];
        out[k] = s;
        x = (uint64_t)prob[s] * (x >> pb) + sl - cum[s];
        while (x < lower) {
            need(pos < nb, "RVN1 stream exhausted");
            x = (x << 8) | bytes[pos++];
        }
    }
    need(x == lower && pos == nb, "RVN1 stream does not end in its initial state");
    return out;
}

uint64_t read_msb(const std::vector<uint8_t>& b, uint64_t bit, uint32_t w) {
    uint64_t v = 0;
    for (uint32_t k = 0; k < w; ++k, ++bit) v = (v << 1) | ((b[bit >> 3] >> (7 - (bit & 7))) & 1u);
    return v;
}

Mat read_component(In r, uint32_t expect_dims) {
    const uint8_t* magic = r.take(4);
    Mat m;
    m.dims = r.u32();
    m.n = r.u64();
    need(m.dims == expect_dims, "component dimension mismatch");
    need(m.n <= (1ull << 40), "implausible component size");
    const uint64_t nv = m.n * m.dims;
    m.v.assign(nv, 0);
    if (std::memcmp(magic, "RAW1", 4) == 0) {
        need(r.end - r.pos == 8 * nv, "RAW1 matrix size mismatch");
        for (auto& x : m.v) x = (int64_t)r.u64();
    } else if (std::memcmp(magic, "RVMX", 4) == 0) {
        const std::vector<uint64_t> v = unrvn1(sub_blob(r), nv);
        for (uint64_t i = 0; i < nv; ++i) m.v[i] = (int64_t)v[i];
    } else if (std::memcmp(magic, "ZFMX", 4) == 0) {
        std::vector<int64_t> mn(m.dims);
        std::vector<uint32_t> w(m.dims);
        std::vector<uint64_t> sz(m.dims);
        uint64_t total = 0;
        for (uint32_t d = 0; d < m.dims; ++d) {
            mn[d] = (int64_t)r.u64();
            w[d] = *r.take(1);
            sz[d] = r.u64();
            need(w[d] >= 1 && w[d] <= 64 && sz[d] == (m.n * w[d] + 7) / 8, "invalid fixed-width header");
            total += sz[d];
        }
        need(r.u64() == total, "fixed payload size mismatch");
        In z = sub_blob(r);
        const std::vector<uint8_t> payload = unzstd(z.p + z.pos, z.end - z.pos, total);
        uint64_t off = 0;
        for (uint32_t d = 0; d < m.dims; ++d) {
            const std::vector<uint8_t> part(payload.begin() + off, payload.begin() + off + sz[d]);
            for (uint64_t i = 0; i < m.n; ++i)
                m.v[i * m.dims + d] = (int64_t)((uint64_t)mn[d] + read_msb(part, i * w[d], w[d]));
            off += sz[d];
        }
    } else if (std::memcmp(magic, "ZTG4", 4) == 0) {
        const uint32_t mw = *r.take(1);
        const std::vector<uint64_t> tags = unrvn1(sub_blob(r), m.n);
        const uint64_t psize = r.u64();
        In z = sub_blob(r);
        const std::vector<uint8_t> packed = unzstd(z.p + z.pos, z.end - z.pos, psize);
        uint64_t bit = 0;
        for (uint64_t i = 0; i < m.n; ++i) {
            need(tags[i] >= 1 && tags[i] <= mw && tags[i] <= 64, "invalid Tagged width");
            need(bit + m.dims * tags[i] <= 8 * psize, "Tagged payload exhausted");
            for (uint32_t d = 0; d < m.dims; ++d, bit += tags[i])
                m.v[i * m.dims + d] = (int64_t)read_msb(packed, bit, (uint32_t)tags[i]);
        }
        need((bit + 7) / 8 == psize, "Tagged payload size mismatch");
    } else {
        throw Fail(CHAIN_LCP2_EFORMAT, "unknown LCP2 component");
    }
    need(r.pos == r.end, "unexpected trailing component bytes");
    return m;
}

// ---------------------------------------------------------------------------
// LCG2, the reference GPU compressor's container (cu/chain/lcps_gpu_cpp.cu:5-35;
// read by lcg2::parse, cu/decompress/lcps_gpu_decompress_ans.cu:2345-2452),
// with every check of that reader. Its Zstd sections become stored sections
// (codec 0: u8 0, u64 raw bytes, the bytes), which the GPU engine reads in
// place; the rest is copied. nvCOMP ANS sections (the tags of Tagged, the
// values of Raw) are refused: their format is nvCOMP's own.

uint32_t le32(const uint8_t* q) { uint32_t v = 0; for (int i = 0; i < 4; ++i) v |= (uint32_t)q[i] << (8 * i); return v; }
uint64_t le64(const uint8_t* q) { uint64_t v = 0; for (int i = 0; i < 8; ++i) v |= (uint64_t)q[i] << (8 * i); return v; }

// one section (lcps_gpu_decompress_ans.cu:2345-2375): codec, raw length,
// chunk table (raw chunks of 65536 bytes for Zstd), chunks
void lcg2_section(In& r, Out& o, uint8_t expected_codec, uint64_t expected_raw) {
    const uint8_t codec = *r.take(1);
    if (codec == 1)
        throw Fail(CHAIN_LCP2_EUNSUPPORTED, "LCG2 section coded with nvCOMP ANS (Tagged tags or Raw values) is not "
                                            "supported: its format is nvCOMP's own");
    need(codec == expected_codec || codec == 0, "unexpected LCG2 section codec");
    const uint64_t raw = r.u64();
    need(raw == expected_raw && raw > 0, "LCG2 section raw length mismatch");
    o.u8(0);
    o.u64(raw);
    if (codec == 0) {
        o.raw(r.take(raw), raw);
        return;
    }
    const uint64_t limit = 65536;
    const uint32_t nchunks = r.u32();
    need(nchunks == raw / limit + (raw % limit != 0), "LCG2 section chunk count mismatch");
    need(nchunks <= (r.end - r.pos) / 16, "truncated LCG2 chunk table");
    std::vector<uint64_t> ub(nchunks), cb(nchunks);
    uint64_t remaining = raw, comp_total = 0;
    for (uint32_t k = 0; k < nchunks; ++k) {
        ub[k] = r.u64();
        cb[k] = r.u64();
        need(ub[k] == std::min(limit, remaining) && cb[k] > 0, "invalid LCG2 chunk lengths");
        remaining -= ub[k];
        need(cb[k] <= r.end - r.pos - comp_total, "truncated LCG2 section");
        comp_total += cb[k];
    }
    need(remaining == 0, "incomplete LCG2 section");
    for (uint32_t k = 0; k < nchunks; ++k) {
        const std::vector<uint8_t> part = unzstd(r.take(cb[k]), cb[k], ub[k]);
        o.raw(part.data(), part.size());
    }
}

// one component (lcps_gpu_decompress_ans.cu:2377-2424)
uint64_t lcg2_component(In& top, Out& o, const char* name, uint32_t dims, uint64_t max_n) {
    need(std::memcmp(top.take(4), name, 4) == 0, "unexpected LCG2 component name");
    const uint64_t length = top.u64();
    need(length <= top.end - top.pos, "truncated LCG2 component");
    In r{top.p, top.pos, top.pos + length};
    top.pos += length;
    Out c;
    const uint8_t kind = *r.take(1);
    need(kind <= 3, "unknown LCG2 encoding");
    const uint32_t cd = r.u32();
    const uint64_t n = r.u64();
    need(cd == dims && n <= max_n, "LCG2 component shape mismatch");
    c.u8(kind);
    c.u32(cd);
    c.u64(n);
    uint64_t fixed_bytes = 0, tagged_bits = 0;
    if (kind == 2) {
        for (uint32_t d = 0; d < dims; ++d) {
            const uint8_t* f = r.take(13);
            const int32_t mn = (int32_t)le32(f);
            const uint32_t width = f[4];
            const uint64_t bytes = le64(f + 5);
            need(mn >= 0 && width >= 1 && width <= 31, "invalid LCG2 fixed field");
            need(bytes == (n * width + 7) / 8, "LCG2 fixed length mismatch");
            fixed_bytes += bytes;
            c.raw(f, 13);
        }
    } else if (kind == 1) {
        tagged_bits = r.u64();
        need(tagged_bits >= n * dims && tagged_bits <= n * dims * 31, "invalid LCG2 packed bit count");
        c.u64(tagged_bits);
    }
    const uint8_t sections = *r.take(1);
    c.u8(sections);
    if (kind == 0) {
        need(n == 0 && sections == 0, "invalid empty LCG2 component");
    } else {
        need(n > 0, "nonempty LCG2 encoding with no records");
        if (kind == 1) {
            need(sections == 2, "LCG2 Tagged needs ANS tags and Zstd values");
            lcg2_section(r, c, 1, n);
            lcg2_section(r, c, 2, (tagged_bits + 7) / 8);
        } else if (kind == 2) {
            need(sections == 1, "LCG2 Fixed needs one Zstd section");
            lcg2_section(r, c, 2, fixed_bytes);
        } else {
            need(sections == 1, "LCG2 Raw needs one ANS section");
            lcg2_section(r, c, 1, n * dims * 4);
        }
    }
    need(r.pos == r.end, "trailing LCG2 component bytes");
    o.raw(name, 4);
    o.blob(c.b);
    return n;
}

// the container (lcps_gpu_decompress_ans.cu:2426-2452)
std::vector<uint8_t> lcg2_inflate(const uint8_t* data, size_t len) {
    In r{data, 0, len};
    need(std::memcmp(r.take(4), "LCG2", 4) == 0, "not an LCG2 stream");
    need(r.u32() == 1, "unsupported LCG2 version");
    const uint32_t D = r.u32();
    const uint64_t n = r.u64();
    const uint8_t* hdr = r.take(16);   // i32 p, f32 eb, i32 wrap_min, i32 wrap_max
    const int32_t p = (int32_t)le32(hdr), wmin = (int32_t)le32(hdr + 8), wmax = (int32_t)le32(hdr + 12);
    float eb;
    const uint32_t ebits = le32(hdr + 4);
    std::memcpy(&eb, &ebits, 4);
    need(D >= 1 && D <= 64 && n > 0 && n <= 0x7fffffffull && p > 0, "invalid LCG2 shape or block size");
    need(eb > 0 && std::isfinite(2.0f * eb), "invalid LCG2 error bound");
    need(wmin >= 0 && wmax >= wmin, "invalid LCG2 wrap range");
    const uint8_t* mins = r.take(4ull * D);
    for (uint32_t d = 0; d < D; ++d) {
        float m;
        const uint32_t b = le32(mins + 4 * d);
        std::memcpy(&m, &b, 4);
        need(std::isfinite(m), "nonfinite LCG2 minimum");
    }
    need(r.u32() == 5, "expected five LCG2 components");
    Out o;
    o.raw("LCG2", 4);
    o.u32(1);
    o.u32(D);
    o.u64(n);
    o.raw(hdr, 16);
    o.raw(mins, 4ull * D);
    o.u32(5);
    const char* names[5] = {"IDR0", "DIF0", "KEP0", "CH20", "REL0"};
    uint64_t cn[5];
    for (int c = 0; c < 5; ++c) cn[c] = lcg2_component(r, o, names[c], (c == 0 || c == 3) ? 1u : D, n);
    need(r.pos == r.end, "trailing LCG2 bytes");
    need(cn[0] == n && cn[4] == n, "LCG2 coordinate stream count mismatch");
    need(cn[3] > 0 && cn[2] > 0 && cn[1] + cn[2] == cn[3], "LCG2 diff/keep count mismatch");
    return o.b;
}

}  // namespace

extern "C" {

int chain_lcp3_from_lcp2(const uint8_t* lcp2, size_t len, uint8_t** out, size_t* out_len, char* err,
                         size_t err_len) {
    try {
        need(lcp2 && out && out_len, "NULL argument", CHAIN_LCP2_EINVAL);
        In r{lcp2, 0, len};
        need(std::memcmp(r.take(4), "LCP2", 4) == 0, "not an LCP2 stream");
        const uint32_t D = r.u32();
        const uint64_t n = r.u64();
        need(D >= 1 && D <= 64, "invalid LCP2 dimensions");
        Out o;
        o.raw("LCP3", 4);
        o.u32(D);
        o.u64(n);
        o.raw(r.take(32 + 8ull * D), 32 + 8ull * D);
        const uint32_t cdims[5] = {1, D, D, 1, D};
        for (int c = 0; c < 5; ++c) {
            const Mat m = read_component(sub_blob(r), cdims[c]);
            Out comp;
            comp.raw("RAW1", 4);
            comp.u32(m.dims);
            comp.u64(m.n);
            for (int64_t x : m.v) comp.u64((uint64_t)x);
            o.blob(comp.b);
        }
        const uint64_t ncorr = r.u64();
        need(ncorr <= (1ull << 40), "implausible correction count");
        In z = sub_blob(r);
        need(r.pos == r.end, "unexpected trailing LCP2 bytes");
        o.u64(ncorr);
        o.blob(ncorr ? unzstd(z.p + z.pos, z.end - z.pos, 16 * ncorr) : std::vector<uint8_t>());
        *out = (uint8_t*)std::malloc(o.b.size() ? o.b.size() : 1);
        need(*out != nullptr, "out of host memory", CHAIN_LCP2_ENOMEM);
        std::memcpy(*out, o.b.data(), o.b.size());
        *out_len = o.b.size();
        return CHAIN_LCP2_OK;
    } catch (const Fail& f) {
        if (err && err_len) {
            std::strncpy(err, f.what(), err_len - 1);
            err[err_len - 1] = 0;
        }
        return f.code;
    } catch (const std::bad_alloc&) {
        if (err && err_len) {
            std::strncpy(err, "out of host memory", err_len - 1);
            err[err_len - 1] = 0;
        }
        return CHAIN_LCP2_ENOMEM;
    }
}


int chain_lcp2_from_lcp3(const uint8_t* lcp3, size_t len, int zstd_level, uint8_t** out, size_t* out_len,
                         uint64_t stats[4], char* err, size_t err_len) {
    try {
        need(lcp3 && out && out_len, "NULL argument", CHAIN_LCP2_EINVAL);
        need(zstd_level >= 1 && zstd_level <= ZSTD_maxCLevel(), "invalid Zstd level", CHAIN_LCP2_EINVAL);
        In r{lcp3, 0, len};
        need(std::memcmp(r.take(4), "LCP3", 4) == 0, "not an LCP3 stream");
        const uint32_t D = r.u32();
        const uint64_t n = r.u64();
        need(D >= 1 && D <= 64, "invalid LCP3 dimensions");
        Out o;
        o.raw("LCP2", 4);
        o.u32(D);
        o.u64(n);
        o.raw(r.take(8 + 8 + 8 + 8 + 8ull * D), 32 + 8ull * D);   // p, eb, wrap_min, wrap_max, mins
        const uint32_t cdims[5] = {1, D, D, 1, D};
        uint64_t sizes[5];
        uint64_t nrec = 0;
        for (int c = 0; c < 5; ++c) {
            const uint64_t ln = r.u64();
            In sub{r.p, r.pos, r.pos + ln};
            need(ln <= r.end - r.pos, "truncated LCP3 component");
            r.pos += ln;
            const Mat m = read_raw1(sub, cdims[c]);
            if (c == 3) nrec = m.n;
            const std::vector<uint8_t> enc = choose(m, zstd_level);
            sizes[c] = enc.size();
            o.blob(enc);
        }
        const uint64_t ncorr = r.u64();
        const uint64_t clen = r.u64();
        need(clen == 16 * ncorr && clen <= r.end - r.pos, "invalid LCP3 corrections");
        const uint8_t* craw = r.take(clen);
        const std::vector<uint8_t> corr(craw, craw + clen);
        need(r.pos == r.end, "unexpected trailing LCP3 bytes");
        o.u64(ncorr);
        o.blob(ncorr ? zstd22(corr, zstd_level) : std::vector<uint8_t>());
        if (stats) {   // lcps_compress2's return list tail, lcps_cpp.cpp:2331-2337
            stats[0] = sizes[1] + sizes[2] + sizes[3];
            stats[1] = sizes[0];
            stats[2] = sizes[4];
            stats[3] = n - nrec;
        }
        *out = (uint8_t*)std::malloc(o.b.size() ? o.b.size() : 1);
        need(*out != nullptr, "out of host memory", CHAIN_LCP2_ENOMEM);
        std::memcpy(*out, o.b.data(), o.b.size());
        *out_len = o.b.size();
        return CHAIN_LCP2_OK;
    } catch (const Fail& f) {
        if (err && err_len) {
            std::strncpy(err, f.what(), err_len - 1);
            err[err_len - 1] = 0;
        }
        return f.code;
    } catch (const std::bad_alloc&) {
        if (err && err_len) {
            std::strncpy(err, "out of host memory", err_len - 1);
            err[err_len - 1] = 0;
        }
        return CHAIN_LCP2_ENOMEM;
    }
}

int chain_lcg2_inflate(const uint8_t* lcg2, size_t len, uint8_t** out, size_t* out_len, char* err,
                       size_t err_len) {
    try {
        need(lcg2 && out && out_len, "NULL argument", CHAIN_LCP2_EINVAL);
        const std::vector<uint8_t> o = lcg2_inflate(lcg2, len);
        *out = (uint8_t*)std::malloc(o.size() ? o.size() : 1);
        need(*out != nullptr, "out of host memory", CHAIN_LCP2_ENOMEM);
        std::memcpy(*out, o.data(), o.size());
        *out_len = o.size();
        return CHAIN_LCP2_OK;
    } catch (const Fail& f) {
        if (err && err_len) {
            std::strncpy(err, f.what(), err_len - 1);
            err[err_len - 1] = 0;
        }
        return f.code;
    } catch (const std::bad_alloc&) {
        if (err && err_len) {
            std::strncpy(err, "out of host memory", err_len - 1);
            err[err_len - 1] = 0;
        }
        return CHAIN_LCP2_ENOMEM;
    }
}

void chain_lcp2_free(uint8_t* p) { std::free(p); }

const char* chain_lcp2_zstd_version(void) { return ZSTD_versionString(); }

}  // extern "C"
