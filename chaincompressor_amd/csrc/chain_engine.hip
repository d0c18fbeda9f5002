// Host orchestration and C ABI of the Chain engine (gfx950).
//
// One call = one batch of frames through one launch per stage:
//   bbox -> [sync: plan keys/widths] -> quantise -> radix sort -> link ->
//   explicit compaction -> L2 stats/split -> corrections -> container.
// The only mid-pipeline host synchronisation is the one after the bounding
// box, which fixes the key width, radix pass count and storage types of the
// whole batch (cf. the six metadata round-trips of cu/chain/lcps_gpu_cpp.cu:842-1023).
#include "chain_common.hpp"
#include "../../include/chain_engine.h"
#include "chain_kernels_pre.hip"
#include "chain_format.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace chain {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) throw Error(CHAIN_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static inline void require(bool ok, int code, const char* msg) {
    if (!ok) throw Error(code, msg);
}

static inline uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// ------------------------------------------------------------------------
// device arena: one allocation grown on demand, carved per call

struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    void ensure(size_t bytes, size_t budget) {
        if (bytes <= cap) return;
        size_t want = bytes + bytes / 8 + (32u << 20);
        if (budget && want > budget) want = bytes;
        if (budget && want > budget)
            throw Error(CHAIN_ENOMEM, "device memory budget exceeded: need " + std::to_string(bytes) + " bytes");
        if (base) HIPCHK(hipFree(base));
        base = nullptr;
        cap = 0;
        HIPCHK(hipMalloc(&base, want));
        cap = want;
    }
    ~Arena() {
        if (base) (void)hipFree(base);
    }
};

// two-pass carving: first pass sums sizes, second hands out pointers
struct Carver {
    char* base = nullptr;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
        off += align256(count * sizeof(T) + 1);
        return p;
    }
};

struct StageTimer {
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<std::string> names;
    std::vector<hipEvent_t> ev;   // 2 per stage
    size_t used = 0;
    void begin(const char* name) {
        if (!on) return;
        if (used + 2 > ev.size()) {
            for (int k = 0; k < 2; ++k) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                ev.push_back(e);
            }
        }
        if (names.size() < used / 2 + 1) names.resize(used / 2 + 1);
        names[used / 2] = name;
        HIPCHK(hipEventRecord(ev[used], s));
    }
    void end() {
        if (!on) return;
        HIPCHK(hipEventRecord(ev[used + 1], s));
        used += 2;
    }
    void reset() { used = 0; }
    ~StageTimer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

}  // namespace chain

struct chain_engine {
    int device = 0;
    uint64_t budget = 0;
    hipStream_t stream = nullptr;
    chain::Arena arena;
    std::mutex mu;
    std::string err;
    chain::StageTimer timer;
};

namespace chain {

// ------------------------------------------------------------------------
// template dispatch helpers

template <int V> using IC = std::integral_constant<int, V>;

template <class Fn>
static void with_dims(uint32_t D, Fn&& fn) {
    switch (D) {
        case 1: fn(IC<1>{}); break;
        case 2: fn(IC<2>{}); break;
        case 3: fn(IC<3>{}); break;
        case 4: fn(IC<4>{}); break;
        default: throw Error(CHAIN_EINVAL, "the GPU engine supports 1 to 4 dimensions");
    }
}

template <class T> struct TypeTag { using type = T; };

template <class Fn>
static void with_dtype(int dtype, Fn&& fn) {
    if (dtype == CHAIN_F32) fn(TypeTag<float>{});
    else if (dtype == CHAIN_F64) fn(TypeTag<double>{});
    else throw Error(CHAIN_EINVAL, "dtype must be float32 or float64");
}

static inline uint64_t okey_inv(uint64_t k) {
    return (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
}
static inline double bits_double(uint64_t b) {
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}

static inline unsigned grid_for(uint64_t n, unsigned per = CH_WG) {
    uint64_t g = cdiv(n ? n : 1, per);
    return (unsigned)std::min<uint64_t>(g, 1u << 20);
}

// ------------------------------------------------------------------------
// compression of one batch

struct FramePlan {
    uint64_t n;
    double eb;
    int64_t p_arg, threshold;
    uint32_t flags;
    uint64_t pe;
    double mins[CH_MAXD];
    uint64_t qmax[CH_MAXD];
    uint32_t key_bits;
    uint64_t relmax;
};

struct CompressResult {
    // host copies of the streams of every frame (LCP3 path) or containers
    std::vector<std::vector<uint8_t>> containers;
    std::vector<chain_stats> stats;
};

template <class T, int D, bool SOA>
static void launch_bbox(hipStream_t st, const void* pts, const FrameDev* dfr, const uint32_t* dtb, uint32_t F,
                        uint32_t ntiles, uint64_t* part, uint32_t* bad, uint64_t* box) {
    hipLaunchKernelGGL((k_bbox<T, D, SOA>), dim3(ntiles), dim3(CH_WG), 0, st, (const T*)pts, dfr, dtb, F, part, bad);
    hipLaunchKernelGGL((k_bbox_finish<D>), dim3(F), dim3(CH_WG), 0, st, part, dtb, box);
}

static void fill64(hipStream_t st, uint64_t* p, uint64_t n, uint64_t v) {
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(CH_WG), 0, st, p, n, v);
}
static void fill32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(CH_WG), 0, st, p, n, v);
}

static void scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* status,
                     uint32_t* ticket, uint32_t* err, uint64_t* total) {
    const uint64_t tiles = cdiv(n, CH_TILE);
    fill64(st, status, tiles, 0);
    fill32(st, ticket, 1, 0);
    hipLaunchKernelGGL(k_scan_u32, dim3((unsigned)tiles), dim3(CH_WG), 0, st, in, out, n, status, ticket, err, total);
}

static void compress_batch_impl(chain_engine* e, const chain_frame* frames, uint32_t F, int dtype, int layout,
                                uint32_t D, bool on_device, int format, CompressResult& res) {
    require(F >= 1, CHAIN_EINVAL, "empty batch");
    require(D >= 1 && D <= CH_MAXD, CHAIN_EINVAL, "the GPU engine supports 1 to 4 dimensions");
    require(layout == CHAIN_AOS || layout == CHAIN_SOA, CHAIN_EINVAL, "layout must be AoS or SoA");
    require(format == CHAIN_FMT_LCP3 || format == CHAIN_FMT_LCH1, CHAIN_EINVAL, "unknown container format");
    const size_t es = dtype == CHAIN_F32 ? 4 : 8;
    require(dtype == CHAIN_F32 || dtype == CHAIN_F64, CHAIN_EINVAL, "dtype must be float32 or float64");
    hipStream_t st = e->stream;
    StageTimer& tm = e->timer;
    tm.reset();

    std::vector<FramePlan> plan(F);
    std::vector<uint64_t> pt_off(F + 1, 0);
    std::vector<uint32_t> tile_begin(F + 1, 0);
    for (uint32_t f = 0; f < F; ++f) {
        const chain_frame& fr = frames[f];
        // argument checks of cpp/chain/lcps_cpp.cpp:2028-2065
        require(fr.eb > 0.0 && std::isfinite(2.0 * fr.eb), CHAIN_EINVAL, "eb must be positive with finite 2*eb");
        require(fr.p > 0, CHAIN_EINVAL, "p must be positive");
        require(fr.n > 0, CHAIN_ERUNTIME, "zero-size array to reduction operation minimum which has no identity");
        plan[f].n = fr.n;
        plan[f].eb = fr.eb;
        plan[f].p_arg = fr.p;
        plan[f].threshold = fr.threshold;
        plan[f].flags = fr.flags;
        plan[f].pe = (fr.flags & CHAIN_NO_BL) ? 1 : (uint64_t)fr.p;
        pt_off[f + 1] = pt_off[f] + fr.n;
        tile_begin[f + 1] = tile_begin[f] + (uint32_t)cdiv(fr.n, CH_TILE);
    }
    const uint64_t T = pt_off[F];
    require(T < (1ull << 32), CHAIN_EINVAL, "a GPU batch holds fewer than 2^32 particles");
    const uint32_t ntiles = tile_begin[F];

    // ---- phase A: bounding boxes
    bool contiguous = on_device;
    if (on_device)
        for (uint32_t f = 0; f < F && contiguous; ++f)
            contiguous = (const char*)frames[f].points == (const char*)frames[0].points + pt_off[f] * D * es;

    Carver ca;
    auto carve_a = [&](Carver& c) {
        struct {
            void* input;
            uint64_t* part;
            uint64_t* box;
            uint32_t* bad;
            FrameDev* fdev;
            uint32_t* tb;
            uint32_t* ptb;
            uint32_t* err;
        } a;
        a.input = contiguous ? nullptr : c.take<char>(T * D * es);
        a.part = c.take<uint64_t>((uint64_t)ntiles * D * 4);
        a.box = c.take<uint64_t>((uint64_t)F * D * 4);
        a.bad = c.take<uint32_t>(F);
        a.fdev = c.take<FrameDev>(F);
        a.tb = c.take<uint32_t>(F + 1);
        a.ptb = c.take<uint32_t>(F + 1);
        a.err = c.take<uint32_t>(4);
        return a;
    };
    (void)carve_a(ca);
    const size_t size_a = ca.off;

    // phase B sizes depend on the plan; compute an upper bound now so the
    // arena is allocated once (keys as u64, rel as u32 worst case)
    const uint64_t sort_tiles = cdiv(T, CH_TILE);
    auto carve_b = [&](Carver& c, size_t key_bytes, size_t rel_bytes) {
        struct {
            void* keys_in;
            void* ka;
            void* kb;
            uint32_t* va;
            uint32_t* vb;
            void* rel;
            uint32_t* idrel;
            uint64_t* corr_tmp_idx;
            double* corr_tmp_val;
            uint32_t* tile_ncorr;
            uint32_t* tile_corr_off;
            uint64_t* corr_idx;
            double* corr_val;
            uint32_t* counts;
            uint32_t* offsets;
            uint64_t* status;
            uint32_t* ticket;
            uint64_t* blocks;
            uint8_t* choice2;
            uint64_t* diff;
            uint64_t* keep;
            FrameOut* fo;
            uint64_t* rec_base;
            uint64_t* totals;   // [0] rec total, [1] diff total, [2] corr total
        } b;
        b.keys_in = c.take<char>(T * key_bytes);
        b.ka = c.take<char>(T * key_bytes);
        b.kb = c.take<char>(T * key_bytes);
        b.va = c.take<uint32_t>(T);
        b.vb = c.take<uint32_t>(T);
        b.rel = c.take<char>(T * D * rel_bytes);
        b.idrel = c.take<uint32_t>(T);
        b.corr_tmp_idx = c.take<uint64_t>((uint64_t)ntiles * CH_TILE * D);
        b.corr_tmp_val = c.take<double>((uint64_t)ntiles * CH_TILE * D);
        b.tile_ncorr = c.take<uint32_t>(ntiles);
        b.tile_corr_off = c.take<uint32_t>(ntiles);
        b.corr_idx = c.take<uint64_t>(T * D);
        b.corr_val = c.take<double>(T * D);
        b.counts = c.take<uint32_t>(256 * sort_tiles);
        b.offsets = c.take<uint32_t>(256 * sort_tiles);
        b.status = c.take<uint64_t>(std::max<uint64_t>(cdiv(256 * sort_tiles, CH_TILE), std::max<uint64_t>(ntiles, sort_tiles)) + 1);
        b.ticket = c.take<uint32_t>(4);
        b.blocks = c.take<uint64_t>(T * D);
        b.choice2 = c.take<uint8_t>(T);
        b.diff = c.take<uint64_t>(T * D);
        b.keep = c.take<uint64_t>(T * D);
        b.fo = c.take<FrameOut>(F);
        b.rec_base = c.take<uint64_t>(F);
        b.totals = c.take<uint64_t>(4);
        return b;
    };
    Carver cb0;
    (void)carve_b(cb0, 8, 4);
    const size_t size_b = cb0.off;
    e->arena.ensure(size_a + size_b, e->budget);

    Carver cA;
    cA.base = e->arena.base;
    auto A = carve_a(cA);
    const void* dpts = contiguous ? frames[0].points : A.input;
    if (!contiguous) {
        for (uint32_t f = 0; f < F; ++f) {
            const size_t bytes = frames[f].n * D * es;
            HIPCHK(hipMemcpyAsync((char*)A.input + pt_off[f] * D * es, frames[f].points, bytes,
                                  on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
        }
    }
    std::vector<FrameDev> fdev(F);
    for (uint32_t f = 0; f < F; ++f) {
        std::memset(&fdev[f], 0, sizeof(FrameDev));
        fdev[f].pt_off = pt_off[f];
        fdev[f].n = plan[f].n;
        fdev[f].tile0 = tile_begin[f];
    }
    std::vector<uint32_t> ptb32(F + 1);
    for (uint32_t f = 0; f <= F; ++f) ptb32[f] = (uint32_t)pt_off[f];
    HIPCHK(hipMemcpyAsync(A.fdev, fdev.data(), F * sizeof(FrameDev), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(A.tb, tile_begin.data(), (F + 1) * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(A.ptb, ptb32.data(), (F + 1) * 4, hipMemcpyHostToDevice, st));
    fill32(st, A.bad, F, 0);
    fill32(st, A.err, 4, 0);
    tm.begin("bbox");
    with_dtype(dtype, [&](auto tt) {
        using T_ = typename decltype(tt)::type;
        with_dims(D, [&](auto dd) {
            constexpr int DD = decltype(dd)::value;
            if (layout == CHAIN_SOA) launch_bbox<T_, DD, true>(st, dpts, A.fdev, A.tb, F, ntiles, A.part, A.bad, A.box);
            else launch_bbox<T_, DD, false>(st, dpts, A.fdev, A.tb, F, ntiles, A.part, A.bad, A.box);
        });
    });
    tm.end();
    std::vector<uint64_t> box((size_t)F * D * 4);
    std::vector<uint32_t> bad(F);
    HIPCHK(hipMemcpyAsync(box.data(), A.box, box.size() * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(bad.data(), A.bad, F * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));

    // ---- plan: minima, quantised range, key layout (lcps_cpp.cpp:2072-2110)
    uint32_t key_bits_max = 0;
    uint64_t relmax = 0;
    for (uint32_t f = 0; f < F; ++f) {
        require(bad[f] == 0, CHAIN_ERUNTIME, "LCPS coordinates must be finite");
        FramePlan& P = plan[f];
        FrameDev& Fd = fdev[f];
        const double denom = 2.0 * P.eb;
        uint64_t bmax[CH_MAXD];
        uint64_t max_bid = 0;
        uint32_t shift = 0;
        for (uint32_t d = 0; d < D; ++d) {
            const uint64_t* b = &box[((size_t)f * D + d) * 4];
            double mn = bits_double(okey_inv(b[0]));
            const double mx = bits_double(okey_inv(b[1]));
            if (mn == 0.0) mn = (b[3] < b[2]) ? -0.0 : 0.0;   // std::min keeps the first zero seen
            P.mins[d] = mn;
            const double qv = std::floor((mx - mn) / denom);
            require(qv >= 0.0 && qv < 0x1p63, CHAIN_ERUNTIME, "quantized coordinate does not fit int64");
            P.qmax[d] = (uint64_t)qv;
            bmax[d] = P.qmax[d] / P.pe;
            max_bid = std::max(max_bid, bmax[d]);
            const uint32_t w = (uint32_t)bitlen64(bmax[d]);
            Fd.key_width[d] = w;
            Fd.key_shift[d] = w ? shift : 0;
            shift += w;
            Fd.mins[d] = mn;
            P.relmax = std::max(P.relmax, std::min<uint64_t>(P.pe - 1, P.qmax[d]));
        }
        require(shift <= 64, CHAIN_ERUNTIME, "block key exceeds 64 bits on the GPU path");
        P.key_bits = shift;
        key_bits_max = std::max(key_bits_max, shift);
        relmax = std::max(relmax, P.relmax);
        Fd.flags = ((P.flags & CHAIN_NO_L1) ? 0u : FL_L1) | ((P.flags & CHAIN_NO_L2) ? 0u : FL_L2);
        Fd.eb = P.eb;
        Fd.denom = denom;
        Fd.p = P.pe;
        uint32_t l = 0;
        while (l < 64 && (1ull << l) < P.pe) ++l;
        Fd.p_l = l;
        Fd.p_magic = l ? (uint64_t)((((unsigned __int128)1 << (63 + l)) + P.pe - 1) / P.pe) : 0;
        Fd.max_bid_bits = (uint32_t)bitlen64(max_bid);
        Fd.threshold = P.threshold;
    }
    require(relmax <= 0xffffffffull, CHAIN_ERUNTIME, "block offsets of 2^32 or more are not supported on the GPU path");
    const bool key64 = key_bits_max > 32;
    const bool rel8 = relmax <= 0xff;
    const size_t kbytes = key64 ? 8 : 4, rbytes = rel8 ? 1 : 4;
    const uint32_t passes = std::max<uint32_t>(1, (key_bits_max + 7) / 8);

    Carver cB;
    cB.base = e->arena.base + size_a;
    auto B = carve_b(cB, kbytes, rbytes);
    HIPCHK(hipMemcpyAsync(A.fdev, fdev.data(), F * sizeof(FrameDev), hipMemcpyHostToDevice, st));

    auto run_typed = [&](auto key_tag, auto rel_tag) {
        using KeyT = typename decltype(key_tag)::type;
        using RelT = typename decltype(rel_tag)::type;
        // K2 quantise
        tm.begin("quantize");
        with_dtype(dtype, [&](auto tt) {
            using T_ = typename decltype(tt)::type;
            with_dims(D, [&](auto dd) {
                constexpr int DD = decltype(dd)::value;
                if (layout == CHAIN_SOA)
                    hipLaunchKernelGGL((k_quantize<T_, DD, true, KeyT, RelT>), dim3(ntiles), dim3(CH_WG), 0, st,
                                       (const T_*)dpts, A.fdev, A.tb, F, (KeyT*)B.keys_in, (RelT*)B.rel,
                                       B.corr_tmp_idx, B.corr_tmp_val, B.tile_ncorr);
                else
                    hipLaunchKernelGGL((k_quantize<T_, DD, false, KeyT, RelT>), dim3(ntiles), dim3(CH_WG), 0, st,
                                       (const T_*)dpts, A.fdev, A.tb, F, (KeyT*)B.keys_in, (RelT*)B.rel,
                                       B.corr_tmp_idx, B.corr_tmp_val, B.tile_ncorr);
            });
        });
        tm.end();
        // K3 sort
        tm.begin("sort");
        const KeyT* kin = (const KeyT*)B.keys_in;
        const uint32_t* vin = nullptr;
        KeyT* kout = (KeyT*)B.ka;
        uint32_t* vout = B.va;
        for (uint32_t ps = 0; ps < passes; ++ps) {
            const uint32_t shift = 8 * ps;
            hipLaunchKernelGGL((k_sort_hist<KeyT>), dim3((unsigned)sort_tiles), dim3(CH_WG), 0, st, kin, T, shift,
                               (uint32_t)sort_tiles, B.counts);
            scan_u32(st, B.counts, B.offsets, 256 * sort_tiles, B.status, B.ticket, A.err, nullptr);
            if (ps == 0)
                hipLaunchKernelGGL((k_sort_scatter<KeyT, true>), dim3((unsigned)sort_tiles), dim3(CH_WG), 0, st, kin,
                                   vin, kout, vout, T, shift, (uint32_t)sort_tiles, B.counts, B.offsets);
            else
                hipLaunchKernelGGL((k_sort_scatter<KeyT, false>), dim3((unsigned)sort_tiles), dim3(CH_WG), 0, st,
                                   kin, vin, kout, vout, T, shift, (uint32_t)sort_tiles, B.counts, B.offsets);
            kin = kout;
            vin = vout;
            kout = (kout == (KeyT*)B.ka) ? (KeyT*)B.kb : (KeyT*)B.ka;
            vout = (vout == B.va) ? B.vb : B.va;
        }
        tm.end();
        // K4 link + decision
        tm.begin("link");
        with_dims(D, [&](auto dd) {
            constexpr int DD = decltype(dd)::value;
            hipLaunchKernelGGL((k_link<KeyT, DD>), dim3(grid_for(T)), dim3(CH_WG), 0, st, kin, vin, T, A.fdev,
                               A.ptb, F, B.idrel);
        });
        tm.end();
        // K5 explicit compaction
        tm.begin("explicit");
        fill64(st, B.status, ntiles, 0);
        fill32(st, B.ticket, 1, 0);
        with_dims(D, [&](auto dd) {
            constexpr int DD = decltype(dd)::value;
            hipLaunchKernelGGL((k_explicit<KeyT, DD>), dim3(ntiles), dim3(CH_WG), 0, st, B.idrel,
                               (const KeyT*)B.keys_in, A.fdev, A.tb, F, ntiles, B.status, B.ticket, A.err, B.blocks,
                               B.fo, &B.totals[0]);
        });
        hipLaunchKernelGGL(k_frame_counts, dim3(cdiv(F, CH_WG)), dim3(CH_WG), 0, st, B.fo, F, &B.totals[0]);
        tm.end();
    };
    auto with_rel = [&](auto key_tag) {
        if (rel8) run_typed(key_tag, TypeTag<uint8_t>{});
        else run_typed(key_tag, TypeTag<uint32_t>{});
    };
    if (key64) with_rel(TypeTag<uint64_t>{});
    else with_rel(TypeTag<uint32_t>{});

    // K6/K7 L2
    tm.begin("l2");
    // frame record bases as a flat array for the record-space kernels
    hipLaunchKernelGGL(k_copy_rec_base, dim3(cdiv(F, CH_WG)), dim3(CH_WG), 0, st, B.fo, F, B.rec_base);
    const uint64_t rec_tiles_max = cdiv(T, CH_TILE);
    with_dims(D, [&](auto dd) {
        constexpr int DD = decltype(dd)::value;
        hipLaunchKernelGGL((k_l2_stats<DD>), dim3(grid_for(T, CH_WG * CH_ITEMS)), dim3(CH_WG), 0, st, B.blocks,
                           &B.totals[0], F, B.rec_base, B.fo);
        hipLaunchKernelGGL((k_l2_wrap<DD>), dim3(cdiv(F, CH_WG)), dim3(CH_WG), 0, st, B.fo, F, A.err);
        fill64(st, B.status, rec_tiles_max, 0);
        fill32(st, B.ticket, 1, 0);
        hipLaunchKernelGGL((k_l2_split<DD>), dim3((unsigned)rec_tiles_max), dim3(CH_WG), 0, st, B.blocks,
                           &B.totals[0], F, B.rec_base, A.fdev, B.fo, B.status, B.ticket, A.err, B.choice2, B.diff,
                           B.keep, &B.totals[1]);
    });
    hipLaunchKernelGGL(k_frame_split_counts, dim3(cdiv(F, CH_WG)), dim3(CH_WG), 0, st, B.fo, F, &B.totals[1],
                       &B.totals[0]);
    tm.end();
    // corrections: per-tile counts -> offsets -> ordered gather
    tm.begin("corrections");
    scan_u32(st, B.tile_ncorr, B.tile_corr_off, ntiles, B.status, B.ticket, A.err, &B.totals[2]);
    with_dims(D, [&](auto dd) {
        constexpr int DD = decltype(dd)::value;
        hipLaunchKernelGGL((k_corr_gather<DD>), dim3(ntiles), dim3(CH_WG), 0, st, B.corr_tmp_idx, B.corr_tmp_val,
                           B.tile_ncorr, B.tile_corr_off, B.corr_idx, B.corr_val);
    });
    hipLaunchKernelGGL(k_frame_corr, dim3(cdiv(F, CH_WG)), dim3(CH_WG), 0, st, B.fo, F, A.tb, B.tile_corr_off,
                       &B.totals[2]);
    tm.end();

    // ---- results to the host
    std::vector<FrameOut> fo(F);
    uint64_t totals[4];
    uint32_t errw[4];
    HIPCHK(hipMemcpyAsync(fo.data(), B.fo, F * sizeof(FrameOut), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(totals, B.totals, sizeof totals, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(errw, A.err, sizeof errw, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    require(errw[0] == 0, CHAIN_EDEVICE, "device error word set (look-back wait limit or invalid block range)");

    const uint64_t R = totals[0], ND = totals[1], NC = totals[2];
    std::vector<uint32_t> h_idrel(T);
    std::vector<uint8_t> h_choice(R);
    std::vector<uint64_t> h_diff(ND * D), h_keep((R - ND) * D);
    std::vector<uint8_t> h_rel(T * D * rbytes);
    std::vector<uint64_t> h_cidx(NC);
    std::vector<double> h_cval(NC);
    HIPCHK(hipMemcpyAsync(h_idrel.data(), B.idrel, T * 4, hipMemcpyDeviceToHost, st));
    if (R) HIPCHK(hipMemcpyAsync(h_choice.data(), B.choice2, R, hipMemcpyDeviceToHost, st));
    if (ND) HIPCHK(hipMemcpyAsync(h_diff.data(), B.diff, ND * D * 8, hipMemcpyDeviceToHost, st));
    if (R - ND) HIPCHK(hipMemcpyAsync(h_keep.data(), B.keep, (R - ND) * D * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(h_rel.data(), B.rel, T * D * rbytes, hipMemcpyDeviceToHost, st));
    if (NC) {
        HIPCHK(hipMemcpyAsync(h_cidx.data(), B.corr_idx, NC * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(h_cval.data(), B.corr_val, NC * 8, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));

    res.containers.resize(F);
    res.stats.resize(F);
    for (uint32_t f = 0; f < F; ++f) {
        const FrameOut& o = fo[f];
        StreamsView v;
        v.dims = D;
        v.n = plan[f].n;
        v.pe = plan[f].pe;
        v.eb = plan[f].eb;
        v.mins = plan[f].mins;
        v.wrap_min = o.wrap_min;
        v.wrap_max = o.wrap_max;
        v.idrel = h_idrel.data() + pt_off[f];
        v.nrec = o.nrec;
        v.choice2 = h_choice.data() + o.rec_base;
        v.ndiff = o.ndiff;
        v.diff = h_diff.data() + o.diff_base * D;
        v.nkeep = o.nkeep;
        v.keep = h_keep.data() + o.keep_base * D;
        v.rel8 = rel8 ? h_rel.data() + pt_off[f] * D : nullptr;
        v.rel32 = rel8 ? nullptr : (const uint32_t*)h_rel.data() + pt_off[f] * D;
        v.ncorr = o.ncorr;
        v.corr_idx = h_cidx.data() + o.corr_base;
        v.corr_val = h_cval.data() + o.corr_base;
        chain_stats& s = res.stats[f];
        write_lcp3(v, res.containers[f], s);
    }
}

}  // namespace chain

// ------------------------------------------------------------------------
// C ABI

using chain::Error;

static thread_local std::string g_create_err;

template <class Fn>
static int guarded(chain_engine* e, Fn&& fn) {
    if (!e) return CHAIN_EINVAL;
    std::lock_guard<std::mutex> lock(e->mu);
    try {
        (void)hipSetDevice(e->device);
        fn();
        e->err.clear();
        return CHAIN_OK;
    } catch (const Error& x) {
        e->err = x.what();
        return x.code;
    } catch (const std::bad_alloc&) {
        e->err = "out of host memory";
        return CHAIN_ENOMEM;
    } catch (const std::exception& x) {
        e->err = x.what();
        return CHAIN_ERUNTIME;
    }
}

extern "C" {

int chain_engine_create(int device, uint64_t hbm_budget_bytes, chain_engine** out) {
    if (!out) return CHAIN_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        g_create_err = "no HIP device is visible";
        return CHAIN_EDEVICE;
    }
    if (device < 0 || device >= count) {
        g_create_err = "device index out of range";
        return CHAIN_EINVAL;
    }
    auto* e = new (std::nothrow) chain_engine();
    if (!e) return CHAIN_ENOMEM;
    e->device = device;
    e->budget = hbm_budget_bytes;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_err = "could not create the engine stream";
        delete e;
        return CHAIN_EDEVICE;
    }
    e->timer.s = e->stream;
    e->timer.on = true;
    *out = e;
    return CHAIN_OK;
}

void chain_engine_destroy(chain_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    (void)hipStreamDestroy(e->stream);
    delete e;
}

const char* chain_last_error(const chain_engine* e) { return e ? e->err.c_str() : g_create_err.c_str(); }

void* chain_engine_stream(chain_engine* e) { return e ? (void*)e->stream : nullptr; }

void chain_buf_free(chain_buf* b) {
    if (!b) return;
    std::free(b->data);
    b->data = nullptr;
    b->size = 0;
}

static void to_buf(const std::vector<uint8_t>& v, chain_buf* out) {
    out->data = (uint8_t*)std::malloc(v.size() ? v.size() : 1);
    if (!out->data) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(out->data, v.data(), v.size());
    out->size = v.size();
}

int chain_compress(chain_engine* e, const void* points, int dtype, int layout, uint32_t dims, uint64_t n, double eb,
                   int64_t p, int64_t threshold, uint32_t flags, int format, chain_buf* out, chain_stats* stats) {
    return guarded(e, [&] {
        if (!out) throw Error(CHAIN_EINVAL, "out must not be NULL");
        chain_frame fr{points, n, eb, p, threshold, flags};
        chain::CompressResult res;
        chain::compress_batch_impl(e, &fr, 1, dtype, layout, dims, false, format, res);
        to_buf(res.containers[0], out);
        if (stats) *stats = res.stats[0];
    });
}

int chain_compress_batch(chain_engine* e, const chain_frame* frames, uint32_t nframes, int dtype, int layout,
                         uint32_t dims, int points_on_device, int format, chain_buf* outs, chain_stats* stats) {
    return guarded(e, [&] {
        if (!frames || !outs) throw Error(CHAIN_EINVAL, "frames and outs must not be NULL");
        chain::CompressResult res;
        chain::compress_batch_impl(e, frames, nframes, dtype, layout, dims, points_on_device != 0, format, res);
        for (uint32_t f = 0; f < nframes; ++f) {
            to_buf(res.containers[f], &outs[f]);
            if (stats) stats[f] = res.stats[f];
        }
    });
}

int chain_decompress(chain_engine* e, const uint8_t* bytes, size_t len, int out_dtype, int out_layout,
                     chain_buf* out, uint64_t* n, uint32_t* dims) {
    return guarded(e, [&] { throw Error(CHAIN_ERUNTIME, "decompression is not built yet"); });
}

int chain_decompress_batch(chain_engine* e, const uint8_t* const* bytes, const size_t* lens, uint32_t nstreams,
                           int out_dtype, int out_layout, chain_buf* outs, uint64_t* ns, uint32_t* dims) {
    return guarded(e, [&] { throw Error(CHAIN_ERUNTIME, "decompression is not built yet"); });
}

int chain_last_stage_times(chain_engine* e, float* ms, int max, char* names, size_t names_len) {
    if (!e) return 0;
    std::lock_guard<std::mutex> lock(e->mu);
    const int k = (int)(e->timer.used / 2);
    std::string all;
    for (int i = 0; i < k && i < max; ++i) {
        float t = 0;
        (void)hipEventElapsedTime(&t, e->timer.ev[2 * i], e->timer.ev[2 * i + 1]);
        if (ms) ms[i] = t;
        if (i) all += ';';
        all += e->timer.names[i];
    }
    if (names && names_len) {
        std::strncpy(names, all.c_str(), names_len - 1);
        names[names_len - 1] = 0;
    }
    return k;
}

}  // extern "C"
