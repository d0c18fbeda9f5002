// Host-side container writers of the Chain engine.
//
// LCP3 is the reference's raw container (ec=False), byte for byte:
//   cpp/chain/lcps_cpp.cpp:2306-2329 (header, five length-prefixed components,
//   raw exact corrections) with RAW1 components (:1184-1211).
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/chain_engine.h"

namespace chain {

struct ByteWriter {
    std::vector<uint8_t>& b;
    explicit ByteWriter(std::vector<uint8_t>& out) : b(out) {}
    void raw(const void* p, size_t n) {
        const uint8_t* q = (const uint8_t*)p;
        b.insert(b.end(), q, q + n);
    }
    void u8(uint8_t v) { b.push_back(v); }
    void u32(uint32_t v) { for (int i = 0; i < 4; ++i) b.push_back((uint8_t)(v >> (8 * i))); }
    void u64(uint64_t v) { for (int i = 0; i < 8; ++i) b.push_back((uint8_t)(v >> (8 * i))); }
    void i64(int64_t v) { u64((uint64_t)v); }
    void f64(double v) { uint64_t u; std::memcpy(&u, &v, 8); u64(u); }
    size_t size() const { return b.size(); }
    void patch_u64(size_t at, uint64_t v) { for (int i = 0; i < 8; ++i) b[at + i] = (uint8_t)(v >> (8 * i)); }
};

// host view of one frame's integer streams
struct StreamsView {
    uint32_t dims;
    uint64_t n;
    uint64_t pe;
    double eb;
    const double* mins;
    int64_t wrap_min, wrap_max;
    const uint32_t* idrel;
    uint64_t nrec;
    const uint8_t* choice2;
    uint64_t ndiff;
    const uint64_t* diff;
    uint64_t nkeep;
    const uint64_t* keep;
    const uint8_t* rel8;     // one of rel8 / rel32
    const uint32_t* rel32;
    uint64_t ncorr;
    const uint64_t* corr_idx;
    const double* corr_val;
};

inline void write_header(ByteWriter& w, const char* magic, const StreamsView& v) {
    w.raw(magic, 4);
    w.u32(v.dims);
    w.u64(v.n);
    w.i64((int64_t)v.pe);
    w.f64(v.eb);
    w.i64(v.wrap_min);
    w.i64(v.wrap_max);
    for (uint32_t d = 0; d < v.dims; ++d) w.f64(v.mins[d]);
}

template <class Get>
inline uint64_t write_raw1(ByteWriter& w, uint32_t dims, uint64_t n, Get get) {
    const size_t len_at = w.size();
    w.u64(0);
    const size_t start = w.size();
    w.raw("RAW1", 4);
    w.u32(dims);
    w.u64(n);
    for (uint64_t i = 0; i < n * dims; ++i) w.i64((int64_t)get(i));
    const uint64_t len = w.size() - start;
    w.patch_u64(len_at, len);
    return len;
}

inline void write_lcp3(const StreamsView& v, std::vector<uint8_t>& out, chain_stats& st) {
    out.clear();
    out.reserve(64 + 8 * (v.n * (1 + v.dims) + v.nrec * (1 + v.dims)) + 16 * v.ncorr);
    ByteWriter w(out);
    write_header(w, "LCP3", v);
    const uint32_t D = v.dims;
    const uint64_t s_idrel = write_raw1(w, 1, v.n, [&](uint64_t i) { return (uint64_t)v.idrel[i]; });
    const uint64_t s_diff = write_raw1(w, D, v.ndiff, [&](uint64_t i) { return v.diff[i]; });
    const uint64_t s_keep = write_raw1(w, D, v.nkeep, [&](uint64_t i) { return v.keep[i]; });
    const uint64_t s_choice = write_raw1(w, 1, v.nrec, [&](uint64_t i) { return (uint64_t)v.choice2[i]; });
    const uint64_t s_rel = write_raw1(w, D, v.n, [&](uint64_t i) {
        return v.rel8 ? (uint64_t)v.rel8[i] : (uint64_t)v.rel32[i];
    });
    w.u64(v.ncorr);
    w.u64(v.ncorr * 16);
    for (uint64_t k = 0; k < v.ncorr; ++k) {
        w.u64(v.corr_idx[k]);
        w.f64(v.corr_val[k]);
    }
    st.bytes = out.size();
    st.block_replacement_size = s_diff + s_keep + s_choice;
    st.idrel_size = s_idrel;
    st.rel_size = s_rel;
    st.num_chained = v.n - v.nrec;
    st.num_corrections = v.ncorr;
}

}  // namespace chain
