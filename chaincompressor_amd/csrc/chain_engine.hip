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
#include "chain_kernels_ent.hip"
#include "chain_kernels_dec.hip"
#include "chain_format.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <unordered_map>
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
    uint64_t allocs = 0;   // device allocations made for this arena so far
    void release() {
        if (base) (void)hipFree(base);
        base = nullptr;
        cap = 0;
    }
    void ensure(size_t bytes, size_t budget) {
        if (bytes <= cap) return;
        size_t want = bytes + bytes / 8 + (32u << 20);
        if (budget && want > budget) want = bytes;
        if (budget && want > budget)
            throw Error(CHAIN_ENOMEM, "device memory budget exceeded: need " + std::to_string(bytes) + " bytes");
        release();
        HIPCHK(hipMalloc(&base, want));
        cap = want;
        ++allocs;
    }
    ~Arena() { release(); }
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
    std::vector<uint64_t> bytes;  // algorithmic HBM bytes of each timed region (0: not stated)
    std::vector<hipEvent_t> ev;   // 2 per stage
    size_t used = 0;
    // returns the index of the timed region (or -1 when timing is off)
    int begin(const char* name, uint64_t alg_bytes = 0) {
        if (!on) return -1;
        if (used + 2 > ev.size()) {
            for (int k = 0; k < 2; ++k) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                ev.push_back(e);
            }
        }
        const size_t i = used / 2;
        if (names.size() < i + 1) {
            names.resize(i + 1);
            bytes.resize(i + 1);
        }
        names[i] = name;
        bytes[i] = alg_bytes;
        HIPCHK(hipEventRecord(ev[used], s));
        return (int)i;
    }
    void end() {
        if (!on) return;
        HIPCHK(hipEventRecord(ev[used + 1], s));
        used += 2;
    }
    void set_bytes(int i, uint64_t b) {
        if (i >= 0 && (size_t)i < bytes.size()) bytes[i] = b;
    }
    void reset() { used = 0; }
    ~StageTimer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// Pinned host staging for the small copies of a call (plans, tables,
// results): one bump allocation per copy, so H2D copies are true async DMA
// and results need one synchronisation. Blocks outgrown during a call are
// released at the next reset, after the stream has drained.
struct Staging {
    char* base = nullptr;
    size_t cap = 0, off = 0;
    bool fixed = false;   // no-allocation calls: never grow, never free
    std::vector<char*> retired;
    struct Pending {
        void* dst;
        const char* src;
        size_t n;
    };
    std::vector<Pending> pend;
    void reset(hipStream_t st) {
        (void)hipStreamSynchronize(st);
        if (!fixed) {
            for (char* r : retired) (void)hipHostFree(r);
            retired.clear();
        }
        pend.clear();
        off = 0;
    }
    char* take(size_t n) {
        n = (n + 63) & ~size_t(63);
        if (off + n > cap) {
            if (fixed)
                throw Error(CHAIN_ENOMEM, "the call's host tables exceed the engine's " + std::to_string(cap) +
                                              " bytes of pinned staging (no-allocation entry point)");
            if (base) retired.push_back(base);
            base = nullptr;
            const size_t want = std::max<size_t>(n, std::max<size_t>(2 * cap, size_t(1) << 20));
            HIPCHK(hipHostMalloc((void**)&base, want, hipHostMallocDefault));
            cap = want;
            off = 0;
        }
        char* q = base + off;
        off += n;
        return q;
    }
    void h2d(hipStream_t st, void* dst, const void* src, size_t n) {
        if (!n) return;
        char* q = take(n);
        std::memcpy(q, src, n);
        HIPCHK(hipMemcpyAsync(dst, q, n, hipMemcpyHostToDevice, st));
    }
    // tables whose device buffers lie in one range: laid out as on the
    // device in one staging block and sent with one copy (the padding
    // between buffers carries no data)
    struct Piece {
        void* dst;
        const void* src;
        size_t n;
    };
    void h2d_range(hipStream_t st, const std::vector<Piece>& ps) {
        char *lo = nullptr, *hi = nullptr;
        for (const Piece& p : ps) {
            if (!p.n) continue;
            char* d = static_cast<char*>(p.dst);
            lo = (!lo || d < lo) ? d : lo;
            hi = (!hi || d + p.n > hi) ? d + p.n : hi;
        }
        if (!lo) return;
        char* q = take((size_t)(hi - lo));
        for (const Piece& p : ps)
            if (p.n) std::memcpy(q + (static_cast<char*>(p.dst) - lo), p.src, p.n);
        HIPCHK(hipMemcpyAsync(lo, q, (size_t)(hi - lo), hipMemcpyHostToDevice, st));
    }
    // device buffers in one range, read with one copy; each piece lands in
    // its host dst at the next finish() (dst and src of a Piece are host
    // and device here)
    void d2h_range(hipStream_t st, const std::vector<Piece>& ps) {
        const char *lo = nullptr, *hi = nullptr;
        for (const Piece& p : ps) {
            if (!p.n) continue;
            const char* s = static_cast<const char*>(p.src);
            lo = (!lo || s < lo) ? s : lo;
            hi = (!hi || s + p.n > hi) ? s + p.n : hi;
        }
        if (!lo) return;
        char* q = take((size_t)(hi - lo));
        HIPCHK(hipMemcpyAsync(q, lo, (size_t)(hi - lo), hipMemcpyDeviceToHost, st));
        for (const Piece& p : ps)
            if (p.n) pend.push_back({p.dst, q + (static_cast<const char*>(p.src) - lo), p.n});
    }
    // lands in dst at the next finish()
    void d2h(hipStream_t st, void* dst, const void* src, size_t n) {
        if (!n) return;
        char* q = take(n);
        HIPCHK(hipMemcpyAsync(q, src, n, hipMemcpyDeviceToHost, st));
        pend.push_back({dst, q, n});
    }
    void finish(hipStream_t st) {
        HIPCHK(hipStreamSynchronize(st));
        for (const Pending& p : pend) std::memcpy(p.dst, p.src, p.n);
        pend.clear();
    }
    ~Staging() {
        for (char* r : retired) (void)hipHostFree(r);
        if (base) (void)hipHostFree(base);
    }
};

// executable graph of the last captured launch sequence
struct GraphCache {
    std::string key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t captures = 0, replays = 0;
    ~GraphCache() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// Tunable settings of an engine: the RVC1 parameters of the containers it
// writes (recorded in every section, so any decoder reads any setting) and
// the L1 strategy (successors inside 4096-point tiles plus a sort of one
// item per (tile, key), or one sort of every point). The output does not
// depend on any of them except the RVC1 chunking, which the container records.
struct Tuning {
    uint32_t lanes_max = 64;
    uint32_t chunk_symbols = 65536;
    uint32_t l1_tiles = 1;
};

}  // namespace chain

struct chain_engine {
    int device = 0;
    uint64_t budget = 0;
    hipStream_t stream = nullptr;
    chain::Arena arena;     // compression
    chain::Arena arena_d;   // decompression (so either direction's device output survives the other)
    // both arenas share the budget: growing one releases the other when the
    // two would not fit together (engines whose outputs never stay on the
    // device, such as the server's workers)
    bool shared_budget = false;
    std::mutex mu;
    std::string err;
    chain::StageTimer timer;
    chain::Staging stage;
    chain::Tuning tuning;
    bool use_graphs = false;
    chain::GraphCache graph;
    bool autotune = false;
    std::map<uint32_t, chain::Tuning> tuned;   // size bucket -> setting
    void* scratch = nullptr;   // this device's no-allocation scratch (device address)
    uint64_t last_need_c = 0;        // arena bytes the last compression carved
    uint64_t capacity_retries = 0;   // compressions run again with worst-case bounds
};

namespace chain {

static void ensure_arena(chain_engine* e, Arena& a, Arena& other, size_t bytes) {
    if (e->shared_budget && e->budget && bytes > a.cap && other.cap && bytes + other.cap > e->budget)
        other.release();
    a.ensure(bytes, e->budget && e->shared_budget ? e->budget - other.cap : e->budget);
}

// Grow the compression arena to `bytes` in the middle of a call, keeping its
// first `keep` bytes (the call's phase A, complete at the planning sync);
// true when the arena moved.
static bool ensure_arena_keep(chain_engine* e, size_t bytes, size_t keep) {
    Arena& a = e->arena;
    if (bytes <= a.cap) return false;
    if (e->shared_budget && e->budget && e->arena_d.cap && bytes + e->arena_d.cap > e->budget) e->arena_d.release();
    const size_t budget = e->budget && e->shared_budget ? e->budget - e->arena_d.cap : e->budget;
    size_t want = bytes + bytes / 8 + (32u << 20);
    if (budget && want > budget) want = bytes;
    if (budget && want > budget)
        throw Error(CHAIN_ENOMEM, "device memory budget exceeded: need " + std::to_string(bytes) + " bytes");
    char* nb = nullptr;
    HIPCHK(hipMalloc(&nb, want));
    if (keep && a.base) {
        hipError_t rc = hipMemcpyAsync(nb, a.base, keep, hipMemcpyDeviceToDevice, e->stream);
        if (rc == hipSuccess) rc = hipStreamSynchronize(e->stream);
        if (rc != hipSuccess) {
            (void)hipFree(nb);
            HIPCHK(rc);
        }
    }
    a.release();
    a.base = nb;
    a.cap = want;
    ++a.allocs;
    return true;
}

// Correction entries per tile in a first compression (a tile with more runs
// the call again with room for every value).
#define CORR_TILE_CAP 256u

// Output bytes of one frame's LCH1 container, bounded by Fixed packing of
// every component at widths the plan fixes: idrel < n, rel <= relmax, keep
// records are block ids (<= max_bid), diff records zig-zag corrections
// within the frame's wrap range (<= max_bid + 1), choice2 one bit; diff and
// keep records together number at most n. The chooser keeps a rANS-coded or
// Tagged candidate only when its estimate is smaller, with slack for the
// estimate's error; the layout kernel checks the bound and the call runs
// again with the worst case (8 bytes per value) if it is exceeded.
// full: that worst case. Corrections: 16 bytes each, at most corr_cap per tile.
static uint64_t out_bound_frame(uint64_t n, uint32_t D, uint32_t max_bid_bits, uint64_t relmax, uint32_t tiles,
                                uint32_t corr_cap, bool full) {
    const uint64_t hdr = 48 + 8ull * D + 16 + 16 * std::min<uint64_t>(n * D, (uint64_t)tiles * corr_cap);
    uint64_t comp = 0;
    if (full) {
        for (int c = 0; c < 5; ++c) {
            const uint32_t dc = (c == 0 || c == 3) ? 1u : D;
            comp += 8 + 16 + 17ull * dc + 8ull * n * dc;
        }
        return hdr + comp;
    }
    auto bytes = [](uint64_t count, uint64_t bits) { return (count * std::max<uint64_t>(bits, 1) + 7) / 8; };
    const uint64_t fix_hdr1 = 8 + 16 + 17, fix_hdrD = 8 + 16 + 17ull * D;
    comp += fix_hdr1 + bytes(n, bitlen64(n ? n - 1 : 0));                        // idrel
    comp += 2 * fix_hdrD + D * (bytes(n, max_bid_bits + 1ull) + 1);              // diff + keep
    comp += fix_hdr1 + bytes(n, 1);                                              // choice2
    comp += fix_hdrD + D * bytes(n, bitlen64(relmax));                           // rel
    comp += comp / 32 + 5 * 4096;
    return hdr + comp;
}

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

// Pinned host buffers for large outputs, recycled: device-to-host copies run
// at DMA speed into pages that are already mapped, and chain_buf_free hands
// a buffer back to the pool instead of unmapping it. Sizes are rounded up to
// 8 classes per power of two (<= 12.5 % slack); at most `cap` bytes stay
// cached. The pool lives until the process ends.
struct PinnedPool {
    std::mutex mu;
    std::unordered_map<void*, size_t> live;
    std::multimap<size_t, void*> idle;
    size_t cached = 0, cap = size_t(8) << 30;
    static constexpr size_t kMin = size_t(1) << 20;   // smaller outputs use malloc
    static size_t klass(size_t n) {
        size_t p2 = kMin;
        while (p2 < n) p2 <<= 1;
        if (p2 == kMin) return kMin;
        const size_t step = p2 / 16;   // classes between p2/2 and p2
        return (n + step - 1) / step * step;
    }
    void* acquire(size_t n) {
        if (n < kMin) return nullptr;
        const size_t c = klass(n);
        {
            std::lock_guard<std::mutex> lock(mu);
            auto it = idle.find(c);
            if (it != idle.end()) {
                void* p = it->second;
                idle.erase(it);
                cached -= c;
                live[p] = c;
                return p;
            }
        }
        void* p = nullptr;
        if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess) return nullptr;
        std::lock_guard<std::mutex> lock(mu);
        live[p] = c;
        return p;
    }
    bool release(void* p) {
        size_t c = 0;
        {
            std::lock_guard<std::mutex> lock(mu);
            auto it = live.find(p);
            if (it == live.end()) return false;
            c = it->second;
            live.erase(it);
            if (cached + c <= cap) {
                idle.emplace(c, p);
                cached += c;
                return true;
            }
        }
        (void)hipHostFree(p);
        return true;
    }
};

static PinnedPool& pinned_pool() {
    static PinnedPool* pool = new PinnedPool();   // never destroyed: outlives every buffer
    return *pool;
}

static uint8_t* host_alloc(size_t n) {
    if (void* p = pinned_pool().acquire(n)) return static_cast<uint8_t*>(p);
    return static_cast<uint8_t*>(std::malloc(n ? n : 1));
}

// Fill caller-owned chain_bufs straight from device memory: one host buffer
// per output (pinned and recycled when large) and one copy, no intermediate
// host vectors. On failure every buffer allocated here is released again.
static void device_to_bufs(hipStream_t st, const uint8_t* dsrc, const uint64_t* offs, const uint64_t* sizes,
                           uint32_t count, chain_buf* out) {
    uint32_t done = 0;
    try {
        for (; done < count; ++done) {
            out[done].data = host_alloc(sizes[done]);
            out[done].size = sizes[done];
            if (!out[done].data) throw std::bad_alloc();
            if (sizes[done])
                HIPCHK(hipMemcpyAsync(out[done].data, dsrc + offs[done], sizes[done], hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipStreamSynchronize(st);
        for (uint32_t k = 0; k <= done && k < count; ++k) {
            if (out[k].data && !pinned_pool().release(out[k].data)) std::free(out[k].data);
            out[k].data = nullptr;
            out[k].size = 0;
        }
        throw;
    }
}

struct CompressResult {
    // host copies of the containers (host_out: caller's bufs filled directly;
    // otherwise containers), or (keep_on_device) offsets into the
    // engine-owned device buffer d_out
    bool keep_on_device = false;
    chain_buf* host_out = nullptr;
    std::vector<std::vector<uint8_t>> containers;
    std::vector<chain_stats> stats;
    uint8_t* d_out = nullptr;
    std::vector<uint64_t> offsets, sizes;
};


template <class T, int D, bool SOA>
// the tile partials of every frame are folded in two launches: BBOX_SPLIT
// workgroups per frame over parts of its tiles, then one per frame over those
#define BBOX_SPLIT 16u
static void launch_bbox(hipStream_t st, const void* pts, const FrameDev* dfr, const uint32_t* dtb, uint32_t F,
                        uint32_t ntiles, uint64_t* part, uint64_t* part2, uint32_t* bad, uint64_t* box) {
    hipLaunchKernelGGL((k_bbox<T, D, SOA>), dim3(ntiles), dim3(CH_WG), 0, st, (const T*)pts, dfr, dtb, F, part, bad);
    hipLaunchKernelGGL((k_bbox_finish<D>), dim3(F * BBOX_SPLIT), dim3(CH_WG), 0, st, part, dtb, 0u, BBOX_SPLIT, part2);
    hipLaunchKernelGGL((k_bbox_finish<D>), dim3(F), dim3(CH_WG), 0, st, part2, (const uint32_t*)nullptr, BBOX_SPLIT, 1u,
                       box);
}

static void fill64(hipStream_t st, uint64_t* p, uint64_t n, uint64_t v) {
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(CH_WG), 0, st, p, n, v);
}
static void fill32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(CH_WG), 0, st, p, n, v);
}

// status (cdiv(n, CH_TILE) words) and ticket start at zero: cleared here
// unless the caller cleared them already (clear = false)
// items_dev: scan 256 x the tiles of that device-side item count (at most n)
static void scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* status,
                     uint32_t* ticket, uint32_t* err, uint64_t* total, bool clear = true,
                     const uint32_t* items_dev = nullptr) {
    const uint64_t tiles = cdiv(n, CH_TILE);
    if (clear) {
        fill64(st, status, tiles, 0);
        fill32(st, ticket, 1, 0);
    }
    hipLaunchKernelGGL(k_scan_u32, dim3((unsigned)tiles), dim3(CH_WG), 0, st, in, out, n, status, ticket, err, total,
                       items_dev);
}

// timed regions of the entropy stage whose algorithmic bytes are known only
// once the stage has run (they depend on the chosen encodings)
struct EntropyTimes {
    int stats[5], encode, assemble, pack[5];
};

template <class X>
static void entropy_stage(hipStream_t st, StageTimer& tm, EntropyTimes& et, uint32_t F, uint32_t D, bool rel8,
                          bool dk32, const FrameDev* fdev, FrameOut* fo, const uint32_t* idrel, const void* diff,
                          const void* keep,
                          const uint8_t* choice2, const void* rel, const uint64_t* corr_idx, const double* corr_val,
                          const uint32_t* tb, const uint32_t* tb_host, uint32_t ntiles,
                          uint32_t ntiles1k, const uint64_t* comp_chunk_bound, const Tuning& rc, uint64_t stride,
                          uint32_t* err, uint64_t out_cap, uint32_t* cap_flag, X& x) {
    const uint32_t ninst = 5 * F, nsec = 10 * F;
    hipLaunchKernelGGL(k_comp_init, dim3(ninst), dim3(CH_WG), 0, st, fdev, fo, F, D, x.ci, x.vhist, x.thist);
    hipLaunchKernelGGL(k_comp_tiles, dim3(5), dim3(CH_WG), 0, st, x.ci, F, x.ctb);
    // every component: a typed view and its launches
    auto each_comp = [&](auto&& fn) {
        with_dims(D, [&](auto dd) {
            constexpr int DD = decltype(dd)::value;
            fn(0, TypeTag<uint32_t>{}, IC<1>{}, (const void*)idrel);
            if (dk32) {
                fn(1, TypeTag<uint32_t>{}, IC<DD>{}, diff);
                fn(2, TypeTag<uint32_t>{}, IC<DD>{}, keep);
            } else {
                fn(1, TypeTag<uint64_t>{}, IC<DD>{}, diff);
                fn(2, TypeTag<uint64_t>{}, IC<DD>{}, keep);
            }
            fn(3, TypeTag<uint8_t>{}, IC<1>{}, (const void*)choice2);
            if (rel8) fn(4, TypeTag<uint8_t>{}, IC<DD>{}, rel);
            else fn(4, TypeTag<uint32_t>{}, IC<DD>{}, rel);
        });
    };
    // statistics slots: a frame has at most min(G, its tiles) + 1 of them (a
    // component never has more tiles in a frame than the frame has point
    // tiles); the most any frame can have decides the reducer's grid
    const uint32_t stats_grid = std::min<uint32_t>(ntiles, CS_GRID);
    uint32_t reduce_y = 1;
    for (uint32_t f = 0; f < F; ++f) {
        const uint32_t cnt = std::min<uint32_t>(tb_host[f + 1] - tb_host[f], stats_grid) + 1;
        reduce_y = std::max(reduce_y, (uint32_t)cdiv(cnt, CS_REDUCE_SLOTS));
    }
    each_comp([&](int c, auto vt, auto dd, const void* v) {
        using V = typename decltype(vt)::type;
        constexpr int DC = decltype(dd)::value;
        et.stats[c] = tm.begin("comp_stats");
        if (ntiles) {
            hipLaunchKernelGGL((k_comp_stats<V, DC>), dim3(stats_grid), dim3(CH_WG), 0, st, (const V*)v, c, F,
                               x.ctb + (uint64_t)c * (F + 1), x.ci, x.slots + (uint64_t)c * (stats_grid + F),
                               x.vhist, x.thist);
        }
        tm.end();
    });
    if (ntiles) {
        // one launch combines every component's slots
        tm.begin("comp_stats");
        hipLaunchKernelGGL(k_comp_reduce, dim3(F, reduce_y, 5), dim3(CH_WG), 0, st, x.slots, stats_grid, x.ctb, F,
                           x.ci, x.vhist, x.thist);
        tm.end();
    }
    tm.begin("rans_table", 5ull * F * (RANS_VALUE_LIMIT + RANS_TAG_LIMIT) * 4 + 10ull * F * RANS_VALUE_LIMIT * 8);
    hipLaunchKernelGGL(k_rans_table, dim3(nsec), dim3(CH_WG), 0, st, x.ci, x.sec, x.vhist, x.thist, F, x.pc,
                       x.tbytes, err);
    hipLaunchKernelGGL(k_prechoose, dim3(cdiv(ninst, 64)), dim3(64), 0, st, x.ci, x.sec, ninst, rc.lanes_max,
                       rc.chunk_symbols);
    hipLaunchKernelGGL(k_rans_plan, dim3(1), dim3(CH_WG), 0, st, x.sec, nsec, F, rc.lanes_max, rc.chunk_symbols,
                       x.comp_chunk);
    tm.end();
    const uint64_t chunk_bound = comp_chunk_bound[0] + comp_chunk_bound[1] + comp_chunk_bound[2] +
                                 comp_chunk_bound[3] + comp_chunk_bound[4];
    et.encode = -1;
    if (chunk_bound) {
        et.encode = tm.begin("rans_encode");
        auto encode = [&](auto dd, auto rt, auto dkt) {
            constexpr int DD = decltype(dd)::value;
            using RelT = typename decltype(rt)::type;
            using DKT = typename decltype(dkt)::type;
            hipLaunchKernelGGL((k_rans_encode<DD, RelT, DKT>), dim3((unsigned)std::min<uint64_t>(chunk_bound, ENC_WGS)),
                               dim3(64), 0, st, idrel,
                               (const DKT*)diff, (const DKT*)keep, choice2, (const RelT*)rel, F, x.ci, x.sec,
                               x.comp_chunk, x.pc, x.scratch, stride, x.chunk_len, x.chunk_pos);
        };
        with_dims(D, [&](auto dd) {
            if (rel8) {
                if (dk32) encode(dd, TypeTag<uint8_t>{}, TypeTag<uint32_t>{});
                else encode(dd, TypeTag<uint8_t>{}, TypeTag<uint64_t>{});
            } else {
                if (dk32) encode(dd, TypeTag<uint32_t>{}, TypeTag<uint32_t>{});
                else encode(dd, TypeTag<uint32_t>{}, TypeTag<uint64_t>{});
            }
        });
        tm.end();
    }
    et.assemble = tm.begin("assemble");
    hipLaunchKernelGGL(k_section_sizes, dim3(nsec), dim3(CH_WG), 0, st, x.sec, x.chunk_len, x.chunk_off);
    hipLaunchKernelGGL(k_choose, dim3(cdiv(ninst, 64)), dim3(64), 0, st, x.ci, x.sec, ninst);
    hipLaunchKernelGGL(k_layout, dim3(1), dim3(CH_WG), 0, st, x.ci, x.sec, fo, F, D, x.out_off, x.out_size, x.total,
                       x.tag_tiles, out_cap, cap_flag);
    hipLaunchKernelGGL(k_write_frames, dim3(F), dim3(CH_WG), 0, st, fdev, fo, D, x.out_off, x.out_size, corr_idx,
                       corr_val, x.total, out_cap, x.out);
    hipLaunchKernelGGL(k_write_components, dim3(ninst), dim3(CH_WG), 0, st, x.ci, x.sec, F, x.tbytes, x.chunk_len,
                       x.total, out_cap, x.out);
    if (chunk_bound)
        hipLaunchKernelGGL(k_copy_chunks, dim3((unsigned)std::min<uint64_t>(chunk_bound, COPY_WGS)), dim3(CH_WG), 0,
                           st, x.ci, x.sec, nsec,
                           x.comp_chunk, x.scratch, stride, x.chunk_len, x.chunk_pos, x.chunk_off, x.total, out_cap,
                           x.out);
    tm.end();
    // every component's look-back words and ticket, cleared by two launches
    fill64(st, x.status1k, 5ull * (ntiles1k + 1), 0);
    fill32(st, x.tickets, 5, 0);
    each_comp([&](int c, auto vt, auto dd, const void* v) {
        using V = typename decltype(vt)::type;
        constexpr int DC = decltype(dd)::value;
        et.pack[c] = tm.begin("pack");
        hipLaunchKernelGGL((k_pack_fixed<V, DC>), dim3(ntiles), dim3(CH_WG), 0, st, (const V*)v, c, F, tb, x.ci,
                           x.total, out_cap, x.out);
        hipLaunchKernelGGL((k_pack_tagged<V, DC>), dim3(std::min<uint32_t>(ntiles1k, PACK_TAG_WGS)), dim3(CH_WG), 0, st,
                           (const V*)v, c, F,
                           x.tag_tiles, x.ci, x.status1k + (uint64_t)c * (ntiles1k + 1), x.tickets + c, err,
                           x.head + (uint64_t)c * ntiles1k, x.tail + (uint64_t)c * ntiles1k,
                           x.head_pos + (uint64_t)c * ntiles1k, x.total, out_cap, x.out);
        tm.end();
    });
    // every component's seams in one launch
    tm.begin("pack");
    if (ntiles1k)
        hipLaunchKernelGGL(k_tag_seams, dim3(cdiv(ntiles1k, CH_WG), 5), dim3(CH_WG), 0, st, x.head, x.tail,
                           x.head_pos, ntiles1k, x.tag_tiles, F, x.total, out_cap, x.out);
    tm.end();
}

// bytes of every value that shapes the launch sequence of a call
struct GraphKey {
    std::string bytes;
    template <class V>
    GraphKey& add(const V& v) {
        bytes.append(reinterpret_cast<const char*>(&v), sizeof v);
        return *this;
    }
};

// Replay the cached graph when the launch shape matches, otherwise capture
// the launches once (no host synchronisation, memcpy or memset inside) and
// update or re-instantiate the executable graph.
template <class Fn>
static void run_graph(chain_engine* e, const std::string& key, Fn&& issue) {
    hipStream_t st = e->stream;
    GraphCache& G = e->graph;
    if (G.exec && G.key == key) {
        HIPCHK(hipGraphLaunch(G.exec, st));
        ++G.replays;
        return;
    }
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        issue();
    } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(st, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(st, &g));
    bool updated = false;
    if (G.exec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult why;
        updated = hipGraphExecUpdate(G.exec, g, &err_node, &why) == hipSuccess && why == hipGraphExecUpdateSuccess;
        (void)hipGetLastError();
        if (!updated) {
            (void)hipGraphExecDestroy(G.exec);
            G.exec = nullptr;
        }
    }
    if (!updated) {
        const hipError_t rc = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
        if (rc != hipSuccess) {
            (void)hipGraphDestroy(g);
            G.exec = nullptr;
            G.key.clear();
            throw Error(CHAIN_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(rc));
        }
    }
    if (G.graph) (void)hipGraphDestroy(G.graph);
    G.graph = g;
    G.key = key;
    ++G.captures;
    HIPCHK(hipGraphLaunch(G.exec, st));
}

// One compression of a batch. full = false sizes the correction lists
// (CORR_TILE_CAP values per tile) and the output (the Fixed-packing bound of
// the plan, chain_out_bound) for the data; when either overflows, nothing
// beyond the buffers is written, the call returns false and is run again
// with full = true (every value corrected, worst-case widths).
static bool compress_batch_try(chain_engine* e, const chain_frame* frames, uint32_t F, int dtype, int layout,
                               uint32_t D, bool on_device, int format, bool full, CompressResult& res) {
    require(F >= 1, CHAIN_EINVAL, "empty batch");
    require(D >= 1 && D <= CH_MAXD, CHAIN_EINVAL, "the GPU engine supports 1 to 4 dimensions");
    require(layout == CHAIN_AOS || layout == CHAIN_SOA, CHAIN_EINVAL, "layout must be AoS or SoA");
    require(format == CHAIN_FMT_LCP3 || format == CHAIN_FMT_LCH1, CHAIN_EINVAL, "unknown container format");
    const size_t es = dtype == CHAIN_F32 ? 4 : 8;
    require(dtype == CHAIN_F32 || dtype == CHAIN_F64, CHAIN_EINVAL, "dtype must be float32 or float64");
    hipStream_t st = e->stream;
    StageTimer& tm = e->timer;
    tm.reset();
    e->stage.reset(st);

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
    // correction entries per tile: a few (corrections are rare), or every value
    const uint32_t corr_cap = full ? CH_TILE * D : std::min<uint32_t>(CORR_TILE_CAP, CH_TILE * D);
    const uint64_t corr_slots = (uint64_t)ntiles * corr_cap;

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
            uint64_t* part2;   // per-frame parts of the tile partials
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
        a.part2 = c.take<uint64_t>((uint64_t)F * BBOX_SPLIT * D * 4);
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
    // look-back words: one region per scan of the pipeline (each radix pass,
    // explicit compaction, L2 split, corrections), all cleared by one launch
    const uint64_t lb_sort = cdiv(256 * sort_tiles, CH_TILE), lb_corr = cdiv(ntiles, CH_TILE);
    auto carve_b = [&](Carver& c, size_t key_bytes, size_t rel_bytes, uint32_t key_words) {
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
            uint64_t* status;   // lb_words: sort passes, explicit, L2 split, corrections
            uint32_t* ticket;   // lb_tickets, in the same order
            uint64_t lb_words;
            uint32_t lb_tickets;
            uint64_t* blocks;
            uint8_t* choice2;
            uint64_t* diff;
            uint64_t* keep;
            FrameOut* fo;
            uint64_t* rec_base;
            uint64_t* totals;   // [0] rec total, [1] diff total, [2] corr total
            void* item_key;     // one item per (tile, distinct key): key, and last << 32 | first point
            uint64_t* item_pay;
            uint32_t* item_total;
        } b;
        b.keys_in = c.take<char>(T * key_bytes * key_words);
        b.ka = c.take<char>(T * key_bytes);
        b.kb = c.take<char>(T * key_bytes);
        // sort values: point indices (u32 ping-pong), or on the in-tile link
        // path the items' (last, first) point pairs (u64, with item_pay)
        b.va = c.take<uint32_t>(2 * T);
        b.vb = b.va + T;
        b.rel = c.take<char>(T * D * rel_bytes);
        b.idrel = c.take<uint32_t>(T);
        b.corr_tmp_idx = c.take<uint64_t>(corr_slots);
        b.corr_tmp_val = c.take<double>(corr_slots);
        b.tile_ncorr = c.take<uint32_t>(ntiles);
        b.tile_corr_off = c.take<uint32_t>(ntiles);
        b.corr_idx = c.take<uint64_t>(std::min<uint64_t>(T * D, corr_slots));
        b.corr_val = c.take<double>(std::min<uint64_t>(T * D, corr_slots));
        b.counts = c.take<uint32_t>(256 * sort_tiles);
        b.offsets = c.take<uint32_t>(256 * sort_tiles);
        // at most key_bytes radix passes per key word
        const uint64_t passes = key_bytes * key_words;
        b.lb_words = passes * lb_sort + ntiles + sort_tiles + lb_corr + ntiles;
        b.lb_tickets = (uint32_t)passes + 4;
        b.status = c.take<uint64_t>(b.lb_words);
        b.ticket = c.take<uint32_t>(b.lb_tickets);
        b.blocks = c.take<uint64_t>(T * D);
        b.choice2 = c.take<uint8_t>(T);
        b.diff = c.take<uint64_t>(T * D);
        b.keep = c.take<uint64_t>(T * D);
        b.fo = c.take<FrameOut>(F);
        b.rec_base = c.take<uint64_t>(F);
        b.totals = c.take<uint64_t>(4);
        // single-word keys only (the in-tile link path)
        const bool items = key_words == 1;
        b.item_key = c.take<char>(items ? T * key_bytes : 0);
        b.item_pay = c.take<uint64_t>(items ? T : 0);
        b.item_total = c.take<uint32_t>(1);
        return b;
    };

    // LCH1 entropy stage buffers, sized from the particle counts alone
    const Tuning rc = e->tuning;
    // per-chunk scratch: states, then at most 3 bytes per symbol, rounded so
    // that the encoder's aligned block stores stay aligned
    const uint64_t stride = CHUNK_STATE_BYTES + align256(3ull * rc.chunk_symbols);
    uint64_t comp_chunk_bound[5] = {0, 0, 0, 0, 0};
    std::vector<uint32_t> tile1k(F + 1, 0);
    for (uint32_t f = 0; f < F; ++f) {
        const uint64_t n = plan[f].n;
        for (int c = 0; c < 5; ++c) {
            const uint32_t dc = (c == 0 || c == 3) ? 1u : D;
            comp_chunk_bound[c] += cdiv(n * dc, rc.chunk_symbols) + cdiv(n, rc.chunk_symbols);
        }
        tile1k[f + 1] = tile1k[f] + (uint32_t)cdiv(n, TAG_TILE);
    }
    const uint64_t chunk_bound = comp_chunk_bound[0] + comp_chunk_bound[1] + comp_chunk_bound[2] +
                                 comp_chunk_bound[3] + comp_chunk_bound[4];
    const uint32_t ntiles1k = tile1k[F];
    auto carve_c = [&](Carver& c, uint64_t out_bound) {
        struct {
            CompInst* ci;
            Section* sec;
            uint32_t* vhist;
            uint32_t* thist;
            CompSlot* slots;
            uint32_t* ctb;   // per component: first statistics tile of each frame, then the total
            uint64_t* pc;
            uint8_t* tbytes;
            uint32_t* comp_chunk;
            uint8_t* scratch;
            uint32_t* chunk_len;
            uint32_t* chunk_pos;
            uint32_t* chunk_off;
            uint64_t* out_off;
            uint64_t* out_size;
            uint64_t* total;
            uint8_t* head;
            uint8_t* tail;
            uint64_t* head_pos;
            uint64_t* status1k;
            uint32_t* tickets;
            uint32_t* tag_tiles;
            uint8_t* out;
        } x;
        x.ci = c.take<CompInst>(5 * F);
        x.sec = c.take<Section>(10 * F);
        // read back with the two above as one range
        x.out_off = c.take<uint64_t>(F);
        x.out_size = c.take<uint64_t>(F);
        x.total = c.take<uint64_t>(1);
        x.vhist = c.take<uint32_t>(5ull * F * RANS_VALUE_LIMIT);
        x.thist = c.take<uint32_t>(5ull * F * RANS_TAG_LIMIT);
        x.slots = c.take<CompSlot>(5ull * (CS_GRID + F));
        x.ctb = c.take<uint32_t>(5ull * (F + 1));
        x.pc = c.take<uint64_t>(10ull * F * RANS_VALUE_LIMIT);
        x.tbytes = c.take<uint8_t>(10ull * F * TABLE_BYTES_MAX);
        x.comp_chunk = c.take<uint32_t>(8);
        x.scratch = c.take<uint8_t>(chunk_bound * stride);
        x.chunk_len = c.take<uint32_t>(chunk_bound);
        x.chunk_pos = c.take<uint32_t>(chunk_bound);
        x.chunk_off = c.take<uint32_t>(chunk_bound);
        // seam bytes of the Tagged tiles, a slice of ntiles1k per component
        x.head = c.take<uint8_t>(5ull * ntiles1k);
        x.tail = c.take<uint8_t>(5ull * ntiles1k);
        x.head_pos = c.take<uint64_t>(5ull * ntiles1k);
        x.status1k = c.take<uint64_t>(5ull * (ntiles1k + 1));
        x.tickets = c.take<uint32_t>(5);
        x.tag_tiles = c.take<uint32_t>(5ull * (F + 1));
        x.out = c.take<uint8_t>(out_bound);
        return x;
    };
    // phase A now; phases B and C once the plan has fixed the widths
    ensure_arena(e, e->arena, e->arena_d, size_a);

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
    Staging& sg = e->stage;
    sg.h2d_range(st, {{A.fdev, fdev.data(), F * sizeof(FrameDev)},
                      {A.tb, tile_begin.data(), (F + 1) * 4},
                      {A.ptb, ptb32.data(), (F + 1) * 4}});
    fill32(st, A.bad, F, 0);
    fill32(st, A.err, 4, 0);
    tm.begin("bbox", T * D * es);
    with_dtype(dtype, [&](auto tt) {
        using T_ = typename decltype(tt)::type;
        with_dims(D, [&](auto dd) {
            constexpr int DD = decltype(dd)::value;
            if (layout == CHAIN_SOA) launch_bbox<T_, DD, true>(st, dpts, A.fdev, A.tb, F, ntiles, A.part, A.part2, A.bad, A.box);
            else launch_bbox<T_, DD, false>(st, dpts, A.fdev, A.tb, F, ntiles, A.part, A.part2, A.bad, A.box);
        });
    });
    tm.end();
    std::vector<uint64_t> box((size_t)F * D * 4);
    std::vector<uint32_t> bad(F);
    sg.d2h_range(st, {{box.data(), A.box, box.size() * 8}, {bad.data(), A.bad, F * 4}});
    sg.finish(st);

    // ---- plan: minima, quantised range, key layout (lcps_cpp.cpp:2072-2110)
    uint32_t key_bits_max = 0, nwords = 1;
    uint32_t word_bits[CH_MAXD] = {0, 0, 0, 0};
    uint64_t relmax = 0;
    for (uint32_t f = 0; f < F; ++f) {
        require(bad[f] == 0, CHAIN_ERUNTIME, "LCPS coordinates must be finite");
        FramePlan& P = plan[f];
        FrameDev& Fd = fdev[f];
        const double denom = 2.0 * P.eb;
        uint64_t bmax[CH_MAXD];
        uint64_t max_bid = 0;
        uint32_t shift = 0, word = 0, total_bits = 0;
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
            if (shift + w > 64) {   // key wider than 64 bits: start the next word
                word_bits[word] = std::max(word_bits[word], shift);
                ++word;
                shift = 0;
            }
            Fd.key_width[d] = w;
            Fd.key_shift[d] = w ? shift : 0;
            Fd.key_word[d] = word;
            shift += w;
            total_bits += w;
            Fd.mins[d] = mn;
            P.relmax = std::max(P.relmax, std::min<uint64_t>(P.pe - 1, P.qmax[d]));
        }
        word_bits[word] = std::max(word_bits[word], shift);
        nwords = std::max(nwords, word + 1);
        P.key_bits = total_bits;
        key_bits_max = std::max(key_bits_max, total_bits);
        relmax = std::max(relmax, P.relmax);
        uint64_t qtop = 0;
        for (uint32_t d = 0; d < D; ++d) qtop = std::max(qtop, P.qmax[d]);
        Fd.flags = ((P.flags & CHAIN_NO_L1) ? 0u : FL_L1) | ((P.flags & CHAIN_NO_L2) ? 0u : FL_L2) |
                   (qtop <= 0xffffffffull && P.pe <= 0x80000000ull ? FL_Q32 : 0u) | ((P.pe & (P.pe - 1)) == 0 ? FL_P2 : 0u);
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
    const bool key64 = key_bits_max > 32 || nwords > 1;
    const bool rel8 = relmax <= 0xff;
    const size_t kbytes = key64 ? 8 : 4, rbytes = rel8 ? 1 : 4;
    // diff and keep values are at most the wrap range (<= largest block id + 1):
    // 32-bit streams when every frame's block ids stay below 2^31
    bool dk32 = true;
    for (uint32_t f = 0; f < F; ++f) dk32 = dk32 && fdev[f].max_bid_bits <= 31;
    const size_t dkbytes = dk32 ? 4 : 8;
    uint32_t word_passes[CH_MAXD];
    for (uint32_t w = 0; w < CH_MAXD; ++w) word_passes[w] = (word_bits[w] + 7) / 8;
    if (nwords == 1) word_passes[0] = std::max<uint32_t>(1, (key_bits_max + 7) / 8);
    // Single-word keys: successors resolved inside each tile by the quantise
    // kernel, and across tiles by sorting one item per (tile, key). Wider
    // keys: every point sorted; the first radix pass (low byte of key word 0)
    // takes its histogram from quantise.
    const bool item_link = nwords == 1 && e->tuning.l1_tiles != 0;
    const bool hist_q = !item_link && word_passes[0] >= 1;

    // the plan fixes key, offset and value widths: phases B and C are sized
    // now, and the arena grows (keeping phase A) when they do not fit
    Carver cbx;
    (void)carve_b(cbx, kbytes, rbytes, nwords);
    const size_t size_b = cbx.off;
    uint64_t out_bound = 0;
    for (uint32_t f = 0; f < F; ++f)
        out_bound += out_bound_frame(plan[f].n, D, fdev[f].max_bid_bits, plan[f].relmax,
                                     tile_begin[f + 1] - tile_begin[f], corr_cap, full);
    size_t size_c = 0;
    if (format == CHAIN_FMT_LCH1) {
        Carver cc0;
        (void)carve_c(cc0, out_bound);
        size_c = cc0.off;
    }
    e->last_need_c = size_a + size_b + size_c;
    if (ensure_arena_keep(e, size_a + size_b + size_c, size_a)) {
        Carver cA2;
        cA2.base = e->arena.base;
        A = carve_a(cA2);
        dpts = contiguous ? frames[0].points : A.input;
    }

    Carver cB;
    cB.base = e->arena.base + size_a;
    auto B = carve_b(cB, kbytes, rbytes, nwords);
    sg.h2d(st, A.fdev, fdev.data(), F * sizeof(FrameDev));

    int t_explicit = -1, t_l2p = -1, t_l2s = -1, t_corr = -1;
    EntropyTimes et{};
    Carver cC;
    cC.base = e->arena.base + size_a + size_b;
    auto X = carve_c(cC, out_bound);   // entropy-stage buffers (LCH1 only)
    // phases A, B and C are consecutive, disjoint ranges of the arena: each
    // carve hands out exactly the bytes its sizing pass counted
    require(cB.off == size_b && (format != CHAIN_FMT_LCH1 || cC.off == size_c) &&
                size_a + size_b + size_c <= e->arena.cap,
            CHAIN_ERUNTIME, "arena carve differs from its sizing pass");
    // every launch after the planning sync: issued directly, or captured
    // once into a HIP graph and replayed for batches of the same launch shape
    uint64_t* const lb_expl = B.status + (uint64_t)kbytes * nwords * lb_sort;
    uint64_t* const lb_l2 = lb_expl + ntiles;
    uint64_t* const lb_cor = lb_l2 + sort_tiles;
    uint64_t* const lb_qlink = lb_cor + lb_corr;
    uint32_t* const tk_expl = B.ticket + kbytes * nwords;
    uint32_t* const tk_qlink = tk_expl + 3;
    auto issue = [&]() {
        fill64(st, B.status, B.lb_words, 0);
        fill32(st, B.ticket, B.lb_tickets, 0);
        auto run_typed = [&](auto key_tag, auto rel_tag) {
            using KeyT = typename decltype(key_tag)::type;
            using RelT = typename decltype(rel_tag)::type;
            // K2 quantise (with the first radix pass's histogram)
            if (hist_q) fill32(st, B.counts, 256 * sort_tiles, 0);
            tm.begin("quantize", T * D * es + T * sizeof(KeyT) * nwords + T * D * sizeof(RelT));
            with_dtype(dtype, [&](auto tt) {
                using T_ = typename decltype(tt)::type;
                with_dims(D, [&](auto dd) {
                    constexpr int DD = decltype(dd)::value;
                    if (layout == CHAIN_SOA)
                        hipLaunchKernelGGL((k_quantize<T_, DD, true, KeyT, RelT>), dim3(ntiles), dim3(CH_WG), 0, st,
                                           (const T_*)dpts, A.fdev, A.tb, F, (KeyT*)B.keys_in, (RelT*)B.rel,
                                           B.corr_tmp_idx, B.corr_tmp_val, B.tile_ncorr, corr_cap, A.err + 1,
                                           nwords, T,
                                           hist_q ? B.counts : nullptr, (uint32_t)sort_tiles);
                    else
                        hipLaunchKernelGGL((k_quantize<T_, DD, false, KeyT, RelT>), dim3(ntiles), dim3(CH_WG), 0, st,
                                           (const T_*)dpts, A.fdev, A.tb, F, (KeyT*)B.keys_in, (RelT*)B.rel,
                                           B.corr_tmp_idx, B.corr_tmp_val, B.tile_ncorr, corr_cap, A.err + 1,
                                           nwords, T,
                                           hist_q ? B.counts : nullptr, (uint32_t)sort_tiles);
                });
            });
            tm.end();
            if (item_link) {
                // K4a successors inside each tile, one item per (tile, key)
                tm.begin("tile_link", T * sizeof(KeyT) + 4 * T);
                with_dims(D, [&](auto dd) {
                    constexpr int DD = decltype(dd)::value;
                    hipLaunchKernelGGL((k_tile_link<KeyT, DD>), dim3(ntiles), dim3(CH_WG), 0, st, (const KeyT*)B.keys_in,
                                       A.fdev, A.tb, F, B.idrel, tk_qlink, lb_qlink, A.err, ntiles,
                                       (KeyT*)B.item_key, B.item_pay, B.item_total);
                });
                tm.end();
            }
            // K3 sort: the items (count on the device) or every point
            const uint32_t* n_dev = item_link ? B.item_total : nullptr;
            const KeyT* kin = (const KeyT*)(item_link ? B.item_key : B.keys_in);
            const uint32_t* vin = nullptr;
            KeyT* kout = (KeyT*)B.ka;
            uint32_t* vout = B.va;
            // the items' payloads ping-pong between item_pay and the u64 view
            // of va/vb
            const uint64_t* pin = B.item_pay;
            uint64_t* pout = reinterpret_cast<uint64_t*>(B.va);
            uint32_t done_passes = 0;
            for (uint32_t w = 0; w < nwords; ++w)
            for (uint32_t ps = 0; ps < word_passes[w]; ++ps) {
                if (w > 0 && ps == 0) {
                    // values of word w in the order left by the words before it
                    KeyT* g = (KeyT*)((kin == (const KeyT*)B.ka) ? B.kb : B.ka);
                    if (kin == (const KeyT*)B.keys_in) g = (KeyT*)B.ka;
                    hipLaunchKernelGGL(k_gather_word, dim3(grid_for(T)), dim3(CH_WG), 0, st,
                                       (const uint64_t*)B.keys_in + (uint64_t)w * T, vin, T, (uint64_t*)g);
                    kin = g;
                    kout = (g == (KeyT*)B.ka) ? (KeyT*)B.kb : (KeyT*)B.ka;
                }
                const uint32_t shift = 8 * ps;
                const uint32_t ps_all = done_passes++;
                if (!(hist_q && ps_all == 0)) {   // the first pass's histogram came with quantise
                    tm.begin("sort_hist", T * sizeof(KeyT));
                    hipLaunchKernelGGL((k_sort_hist<KeyT>), dim3((unsigned)sort_tiles), dim3(CH_WG), 0, st, kin, T,
                                       shift, (uint32_t)sort_tiles, B.counts, n_dev);
                    tm.end();
                }
                tm.begin("sort_scan", 2ull * 256 * sort_tiles * 4);
                scan_u32(st, B.counts, B.offsets, 256 * sort_tiles, B.status + (uint64_t)ps_all * lb_sort,
                         B.ticket + ps_all, A.err, nullptr, false, n_dev);
                tm.end();
                tm.begin("sort_scatter", ps_all == 0 ? T * sizeof(KeyT) + T * (sizeof(KeyT) + 4)
                                                     : 2 * T * (sizeof(KeyT) + 4));
                if (item_link) {
                    hipLaunchKernelGGL((k_sort_scatter<KeyT, uint64_t, false>), dim3((unsigned)sort_tiles), dim3(CH_WG), 0,
                                       st, kin, pin, kout, pout, T, shift, (uint32_t)sort_tiles, B.counts, B.offsets,
                                       n_dev);
                    const uint64_t* t = pin;
                    pin = pout;
                    pout = const_cast<uint64_t*>(t);
                } else if (ps_all == 0) {
                    hipLaunchKernelGGL((k_sort_scatter<KeyT, uint32_t, true>), dim3((unsigned)sort_tiles), dim3(CH_WG), 0,
                                       st, kin, vin, kout, vout, T, shift, (uint32_t)sort_tiles, B.counts, B.offsets,
                                       n_dev);
                } else {
                    hipLaunchKernelGGL((k_sort_scatter<KeyT, uint32_t, false>), dim3((unsigned)sort_tiles), dim3(CH_WG), 0,
                                       st, kin, vin, kout, vout, T, shift, (uint32_t)sort_tiles, B.counts, B.offsets,
                                       n_dev);
                }
                tm.end();
                kin = kout;
                vin = vout;
                kout = (kout == (KeyT*)B.ka) ? (KeyT*)B.kb : (KeyT*)B.ka;
                vout = (vout == B.va) ? B.vb : B.va;
            }
            // K4 link + decision
            tm.begin("link", item_link ? 0 : T * (sizeof(KeyT) * nwords + 8));
            with_dims(D, [&](auto dd) {
                constexpr int DD = decltype(dd)::value;
                if (item_link)
                    hipLaunchKernelGGL((k_link_items<KeyT, DD>), dim3((unsigned)sort_tiles), dim3(CH_WG), 0, st, kin,
                                       pin, B.item_total, A.fdev, A.ptb, F, B.idrel);
                else if (nwords == 1)
                    hipLaunchKernelGGL((k_link<KeyT, DD>), dim3(LINK_XCDS * (unsigned)cdiv(cdiv(T, CH_TILE), LINK_XCDS)),
                                       dim3(CH_WG), 0, st, kin, vin, T, A.fdev, A.ptb, F, B.idrel);
                else
                    hipLaunchKernelGGL((k_link_wide<DD>), dim3(grid_for(T)), dim3(CH_WG), 0, st,
                                       (const uint64_t*)B.keys_in, vin, T, nwords, A.fdev, A.ptb, F, B.idrel);
            });
            tm.end();
            // K5 explicit compaction
            t_explicit = tm.begin("explicit");
            with_dims(D, [&](auto dd) {
                constexpr int DD = decltype(dd)::value;
                auto expl = [&](auto bt) {
                    using BT = typename decltype(bt)::type;
                    hipLaunchKernelGGL((k_explicit<KeyT, DD, expl_tpw<KeyT>(), BT>),
                                       dim3(cdiv(ntiles, expl_tpw<KeyT>())), dim3(CH_WG), 0, st, B.idrel,
                                       (const KeyT*)B.keys_in, A.fdev, A.tb, F, ntiles, lb_expl, tk_expl, A.err,
                                       (BT*)B.blocks, B.fo, &B.totals[0], nwords, T);
                };
                if (dk32) expl(TypeTag<uint32_t>{});
                else expl(TypeTag<uint64_t>{});
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
        t_l2s = tm.begin("l2_stats");
        // frame record bases as a flat array for the record-space kernels
        hipLaunchKernelGGL(k_copy_rec_base, dim3(cdiv(F, CH_WG)), dim3(CH_WG), 0, st, B.fo, F, B.rec_base);
        const uint64_t rec_tiles_max = cdiv(T, CH_TILE);
        with_dims(D, [&](auto dd) {
            constexpr int DD = decltype(dd)::value;
            if (dk32)
                hipLaunchKernelGGL((k_l2_stats<DD, uint32_t>), dim3(std::min<unsigned>(grid_for(T, CH_WG * CH_ITEMS), 1024)),
                                   dim3(CH_WG), 0, st, (const uint32_t*)B.blocks, &B.totals[0], F, B.rec_base, B.fo);
            else
                hipLaunchKernelGGL((k_l2_stats<DD, uint64_t>), dim3(std::min<unsigned>(grid_for(T, CH_WG * CH_ITEMS), 1024)),
                                   dim3(CH_WG), 0, st, B.blocks, &B.totals[0], F, B.rec_base, B.fo);
            hipLaunchKernelGGL((k_l2_wrap<DD>), dim3(cdiv(F, CH_WG)), dim3(CH_WG), 0, st, B.fo, F, A.err);
            tm.end();
            t_l2p = tm.begin("l2_split");
            auto split = [&](auto dkt) {
                using DKT = typename decltype(dkt)::type;
                hipLaunchKernelGGL((k_l2_split<DD, DKT>), dim3((unsigned)std::min<uint64_t>(rec_tiles_max, L2S_WGS)),
                                   dim3(CH_WG), 0, st, (const DKT*)B.blocks, &B.totals[0], F, B.rec_base, A.fdev, B.fo, lb_l2,
                                   tk_expl + 1, A.err, B.choice2, (DKT*)B.diff, (DKT*)B.keep, &B.totals[1]);
            };
            if (dk32) split(TypeTag<uint32_t>{});
            else split(TypeTag<uint64_t>{});
        });
        hipLaunchKernelGGL(k_frame_split_counts, dim3(cdiv(F, CH_WG)), dim3(CH_WG), 0, st, B.fo, F, &B.totals[1],
                           &B.totals[0]);
        tm.end();

        // corrections: per-tile counts -> offsets -> ordered gather
        t_corr = tm.begin("corrections");
        scan_u32(st, B.tile_ncorr, B.tile_corr_off, ntiles, lb_cor, tk_expl + 2, A.err, &B.totals[2], false);
        with_dims(D, [&](auto dd) {
            constexpr int DD = decltype(dd)::value;
            hipLaunchKernelGGL((k_corr_gather<DD>), dim3(ntiles), dim3(CH_WG), 0, st, B.corr_tmp_idx, B.corr_tmp_val,
                               B.tile_ncorr, B.tile_corr_off, corr_cap, B.corr_idx, B.corr_val);
        });
        hipLaunchKernelGGL(k_frame_corr, dim3(cdiv(F, CH_WG)), dim3(CH_WG), 0, st, B.fo, F, A.tb, B.tile_corr_off,
                           &B.totals[2]);
        tm.end();
        if (format == CHAIN_FMT_LCH1)
            entropy_stage(st, tm, et, F, D, rel8, dk32, A.fdev, B.fo, B.idrel, B.diff, B.keep, B.choice2, B.rel,
                          B.corr_idx,
                          B.corr_val, A.tb, tile_begin.data(), ntiles, ntiles1k, comp_chunk_bound, rc, stride, A.err,
                          out_bound, A.err + 1, X);
    };
    if (e->use_graphs) {
        GraphKey key;
        key.add(format).add(dtype).add(layout).add(D).add(F).add(T).add(nwords).add(key64).add(rel8).add(dk32);
        key.add(dpts).add(e->arena.base).add(ntiles).add(ntiles1k).add(rc.lanes_max).add(rc.chunk_symbols);
        key.add(item_link).add(corr_cap).add(out_bound);
        for (uint32_t w = 0; w < CH_MAXD; ++w) key.add(word_passes[w]);
        for (int c = 0; c < 5; ++c) key.add(comp_chunk_bound[c]);
        for (uint32_t f = 0; f <= F; ++f) key.add(pt_off[f]);
        const bool was = tm.on;
        tm.begin("graph");
        tm.on = false;
        try {
            run_graph(e, key.bytes, issue);
        } catch (...) {
            tm.on = was;
            throw;
        }
        tm.on = was;
        tm.end();
    } else {
        issue();
    }

    const size_t kbytes_all = kbytes * nwords;
    // on the in-tile link path the sort and link regions work on the M items
    // (key, last << 32 | first point) counted on the device
    auto set_item_bytes = [&](uint64_t M) {
        if (!item_link || !tm.on) return;
        for (size_t i = 0; i < tm.used / 2; ++i) {
            const std::string& nm = tm.names[i];
            if (nm == "sort_hist") tm.set_bytes((int)i, M * kbytes);
            else if (nm == "sort_scan") tm.set_bytes((int)i, 2ull * 256 * cdiv(M, CH_TILE) * 4);
            else if (nm == "sort_scatter") tm.set_bytes((int)i, 2 * M * (kbytes + 8));
            else if (nm == "link") tm.set_bytes((int)i, M * (kbytes + 8) + 4 * M);
        }
    };
    auto set_pre_bytes = [&](uint64_t R, uint64_t NC) {
        tm.set_bytes(t_explicit, T * (4 + kbytes_all) + dkbytes * D * R);
        tm.set_bytes(t_l2s, dkbytes * D * R);
        tm.set_bytes(t_l2p, 2ull * dkbytes * D * R + R);
        tm.set_bytes(t_corr, 32ull * NC + 8ull * ntiles);
    };
    uint32_t n_items = 0;

    if (format == CHAIN_FMT_LCH1) {
        std::vector<uint64_t> offs(F), sizes(F);
        uint64_t total = 0;
        uint32_t errw[4];
        std::vector<CompInst> ci(5 * F);
        std::vector<FrameOut> fo(F);
        std::vector<Section> secs(tm.on ? 10 * F : 0);
        sg.d2h_range(st, {{offs.data(), X.out_off, F * 8},
                          {sizes.data(), X.out_size, F * 8},
                          {&total, X.total, 8},
                          {ci.data(), X.ci, 5 * F * sizeof(CompInst)},
                          {secs.data(), X.sec, secs.size() * sizeof(Section)}});
        sg.d2h(st, errw, A.err, sizeof errw);
        sg.d2h(st, fo.data(), B.fo, F * sizeof(FrameOut));
        sg.d2h(st, &n_items, B.item_total, 4);
        sg.finish(st);
        set_item_bytes(n_items);
        if (tm.on) {
            // algorithmic bytes of the regions whose work depends on the data
            uint64_t R = 0, NC = 0;
            for (uint32_t f = 0; f < F; ++f) {
                R += fo[f].nrec;
                NC += fo[f].ncorr;
            }
            set_pre_bytes(R, NC);
            const uint64_t esz[5] = {4, dkbytes, dkbytes, 1, rbytes};
            uint64_t copied = 0, enc_all = 0;
            for (int c = 0; c < 5; ++c) {
                uint64_t in_b = 0, enc_b = 0, pack_b = 0;
                for (uint32_t f = 0; f < F; ++f) {
                    const CompInst& I = ci[c * F + f];
                    const uint64_t vb = I.n * I.dims * esz[c];
                    in_b += vb;
                    for (int sx = 0; sx < 2; ++sx) {
                        const Section& S = secs[2 * (c * F + f) + sx];
                        if (S.nchunks) {
                            enc_b += vb + S.payload;
                            copied += S.payload;
                        }
                    }
                    if (I.kind == KIND_FIXP) {
                        uint64_t w = 0;
                        for (uint32_t d = 0; d < I.dims; ++d) w += I.fix_bytes[d];
                        pack_b += vb + w;
                    } else if (I.kind == KIND_TAGR) {
                        pack_b += vb + (I.tag_bits + 7) / 8;
                    }
                }
                tm.set_bytes(et.stats[c], in_b);
                enc_all += enc_b;
                tm.set_bytes(et.pack[c], pack_b);
            }
            tm.set_bytes(et.encode, enc_all);
            tm.set_bytes(et.assemble, 2 * copied);
        }
        if (errw[1]) {   // a correction list or the output overran its bound
            require(!full, CHAIN_EDEVICE, "capacity exceeded at the worst-case bound");
            return false;
        }
        require(errw[0] == 0, CHAIN_EDEVICE, "device error word set in the entropy stage");
        require(total <= out_bound, CHAIN_EDEVICE, "container exceeds its bound");
        res.stats.resize(F);
        for (uint32_t f = 0; f < F; ++f) {
            chain_stats& s = res.stats[f];
            s.bytes = sizes[f];
            s.block_replacement_size = ci[1 * F + f].size + ci[2 * F + f].size + ci[3 * F + f].size;
            s.idrel_size = ci[0 * F + f].size;
            s.rel_size = ci[4 * F + f].size;
            s.num_chained = plan[f].n - fo[f].nrec;
            s.num_corrections = fo[f].ncorr;
        }
        res.offsets = offs;
        res.sizes = sizes;
        if (res.keep_on_device) {
            res.d_out = X.out;
            return true;
        }
        if (res.host_out) {
            device_to_bufs(st, X.out, offs.data(), sizes.data(), F, res.host_out);
            return true;
        }
        std::vector<uint8_t> all(total);
        if (total) HIPCHK(hipMemcpyAsync(all.data(), X.out, total, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        res.containers.resize(F);
        for (uint32_t f = 0; f < F; ++f)
            res.containers[f].assign(all.begin() + offs[f], all.begin() + offs[f] + sizes[f]);
        return true;
    }

    // ---- results to the host
    std::vector<FrameOut> fo(F);
    uint64_t totals[4];
    uint32_t errw[4];
    sg.d2h(st, fo.data(), B.fo, F * sizeof(FrameOut));
    sg.d2h(st, totals, B.totals, sizeof totals);
    sg.d2h(st, errw, A.err, sizeof errw);
    sg.d2h(st, &n_items, B.item_total, 4);
    sg.finish(st);
    set_item_bytes(n_items);
    if (errw[1]) {   // a tile's corrections overran the list bound
        require(!full, CHAIN_EDEVICE, "capacity exceeded at the worst-case bound");
        return false;
    }
    require(errw[0] == 0, CHAIN_EDEVICE, "device error word set (look-back wait limit or invalid block range)");

    const uint64_t R = totals[0], ND = totals[1], NC = totals[2];
    set_pre_bytes(R, NC);
    std::vector<uint32_t> h_idrel(T);
    std::vector<uint8_t> h_choice(R);
    std::vector<uint64_t> h_diff(ND * D), h_keep((R - ND) * D);
    std::vector<uint8_t> h_rel(T * D * rbytes);
    std::vector<uint64_t> h_cidx(NC);
    std::vector<double> h_cval(NC);
    HIPCHK(hipMemcpyAsync(h_idrel.data(), B.idrel, T * 4, hipMemcpyDeviceToHost, st));
    if (R) HIPCHK(hipMemcpyAsync(h_choice.data(), B.choice2, R, hipMemcpyDeviceToHost, st));
    // 32-bit streams are copied into the upper halves and widened in place
    if (ND)
        HIPCHK(hipMemcpyAsync((char*)h_diff.data() + (dk32 ? ND * D * 4 : 0), B.diff, ND * D * dkbytes,
                              hipMemcpyDeviceToHost, st));
    if (R - ND)
        HIPCHK(hipMemcpyAsync((char*)h_keep.data() + (dk32 ? (R - ND) * D * 4 : 0), B.keep, (R - ND) * D * dkbytes,
                              hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(h_rel.data(), B.rel, T * D * rbytes, hipMemcpyDeviceToHost, st));
    if (NC) {
        HIPCHK(hipMemcpyAsync(h_cidx.data(), B.corr_idx, NC * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(h_cval.data(), B.corr_val, NC * 8, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    if (dk32) {
        // 32-bit element i (bytes 4n + 4i of the n-element vector) becomes
        // element i (bytes 8i .. 8i + 7): in ascending order a write ends at or
        // before the next 32-bit element still to be read (8i + 8 <= 4n + 4i + 4)
        auto widen = [](std::vector<uint64_t>& v) {
            const uint64_t n = v.size();
            const char* base = (const char*)v.data() + n * 4;
            for (uint64_t i = 0; i < n; ++i) {
                uint32_t x;
                std::memcpy(&x, base + 4 * i, 4);
                v[i] = x;
            }
        };
        widen(h_diff);
        widen(h_keep);
    }

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
    return true;
}

static void compress_batch_impl(chain_engine* e, const chain_frame* frames, uint32_t F, int dtype, int layout,
                                uint32_t D, bool on_device, int format, CompressResult& res) {
    const bool keep = res.keep_on_device;
    chain_buf* host_out = res.host_out;
    if (compress_batch_try(e, frames, F, dtype, layout, D, on_device, format, false, res)) return;
    res = CompressResult{};
    res.keep_on_device = keep;
    res.host_out = host_out;
    ++e->capacity_retries;
    require(compress_batch_try(e, frames, F, dtype, layout, D, on_device, format, true, res), CHAIN_EDEVICE,
            "capacity exceeded at the worst-case bound");
}


// ------------------------------------------------------------------------
// decompression of a batch of LCH1 / LCP3 streams

// little-endian reader over one stream inside the pooled input
struct HostReader {
    const uint8_t* base;   // the stream's host bytes
    uint64_t pos, end;     // offsets inside the stream
    uint64_t dev;          // offset of the stream in the device input pool
    const uint8_t* bytes(uint64_t k) {
        require(k <= end - pos, CHAIN_ERUNTIME, "truncated LCPS bitstream");
        const uint8_t* q = base + pos;
        pos += k;
        return q;
    }
    uint8_t u8() { return *bytes(1); }
    uint32_t u32() { const uint8_t* q = bytes(4); uint32_t v = 0; for (int i = 0; i < 4; ++i) v |= (uint32_t)q[i] << (8 * i); return v; }
    uint64_t u64() { const uint8_t* q = bytes(8); uint64_t v = 0; for (int i = 0; i < 8; ++i) v |= (uint64_t)q[i] << (8 * i); return v; }
    int64_t i64() { return (int64_t)u64(); }
    double f64() { return bits_double(u64()); }
    uint64_t varuint() {
        uint64_t v = 0;
        for (int shift = 0; shift <= 63; shift += 7) {
            const uint8_t b = u8();
            require(shift != 63 || (b & 0xfeu) == 0, CHAIN_ERUNTIME, "LCPS varuint overflow");
            v |= (uint64_t)(b & 0x7fu) << shift;
            if ((b & 0x80u) == 0) return v;
        }
        throw Error(CHAIN_ERUNTIME, "invalid LCPS varuint");
    }
    HostReader sub(uint64_t len) {
        require(len <= end - pos, CHAIN_ERUNTIME, "truncated LCPS bitstream");
        HostReader r{base, pos, pos + len, dev};
        pos += len;
        return r;
    }
    void finish() const { require(pos == end, CHAIN_ERUNTIME, "unexpected trailing LCPS bytes"); }
};

struct DecPlan {
    std::vector<DecFrame> frames;
    std::vector<DecSection> secs;
    std::vector<DecChunk> chunks;
    std::vector<DecFixed> fixed;
    std::vector<DecRaw> raws;
    std::vector<DecTagged> tagged;
    std::vector<DecJob> jobs;
    std::vector<uint32_t> cum;  // per section: alphabet + 1 cumulative frequencies
    uint64_t T = 0, R = 0, ND = 0, NK = 0, tags = 0, tag_tiles = 0, fixed_tiles = 0, raw_total = 0;
    uint32_t D = 0;
};

// RVC1 section (see chain_kernels_ent.hip): table parsed on the host
static void parse_rvc1(HostReader r, uint64_t expected, int comp, uint64_t dst, DecPlan& P) {
    require(std::memcmp(r.bytes(4), "RVC1", 4) == 0, CHAIN_ERUNTIME, "invalid RVC1 header");
    const uint32_t pb = r.u8();
    require(r.u64() == expected, CHAIN_ERUNTIME, "RVC1 symbol count mismatch");
    HostReader tr = r.sub(r.u64());
    const uint32_t L = r.u8();
    const uint32_t S = r.u32();
    const uint32_t nchunks = r.u32();
    if (expected == 0) {
        require(pb == 0 && tr.end == tr.pos && L == 0 && S == 0 && nchunks == 0, CHAIN_ERUNTIME,
                "invalid empty RVC1 section");
        r.finish();
        return;
    }
    require(pb >= 12 && pb <= 20, CHAIN_ERUNTIME, "invalid RVC1 precision");
    require(L >= 1 && L <= 64 && (L & (L - 1)) == 0 && S >= 1, CHAIN_ERUNTIME, "invalid RVC1 lanes or chunk size");
    require(nchunks == cdiv(expected, S), CHAIN_ERUNTIME, "RVC1 chunk count mismatch");
    const uint64_t A = tr.varuint();
    require(A > 0 && A <= RANS_VALUE_LIMIT, CHAIN_ERUNTIME, "RVC1 alphabet exceeds the GPU decoder's 4096 symbols");
    std::vector<uint32_t> prob(A, 0);
    for (uint64_t i = 0; i < A;) {
        const uint8_t first = tr.u8();
        const uint32_t token = first & 3u;
        if (token == 3) {
            const uint64_t zeros = (first >> 2) + 1u;
            require(zeros <= A - i, CHAIN_ERUNTIME, "invalid rANS zero run");
            i += zeros;
        } else {
            uint32_t pr = first >> 2;
            for (uint32_t b = 0; b < token; ++b) pr |= (uint32_t)tr.u8() << (6 + 8 * b);
            require(pr > 0, CHAIN_ERUNTIME, "invalid rANS probability");
            prob[i++] = pr;
        }
    }
    tr.finish();
    const uint32_t M = 1u << pb;
    DecSection sec;
    std::memset(&sec, 0, sizeof sec);
    sec.pb = pb;
    sec.lanes = L;
    sec.alphabet = (uint32_t)A;
    sec.chunk_symbols = S;
    sec.count = expected;
    sec.cum_off = P.cum.size();
    sec.comp = comp;
    sec.dst = dst;
    uint32_t sum = 0;
    for (uint64_t k = 0; k < A; ++k) {
        require(prob[k] <= M - sum, CHAIN_ERUNTIME, "invalid rANS probability sum");
        P.cum.push_back(sum);
        sum += prob[k];
    }
    P.cum.push_back(sum);
    require(sum == M, CHAIN_ERUNTIME, "incomplete rANS probability table");
    const uint32_t si = (uint32_t)P.secs.size();
    P.secs.push_back(sec);
    const uint8_t* sizes = r.bytes(4ull * nchunks);
    for (uint32_t k = 0; k < nchunks; k += DEC_WAVES) {
        DecJob j;
        j.sec = si;
        j.chunk0 = (uint32_t)P.chunks.size() + k;
        j.nchunks = std::min<uint32_t>(DEC_WAVES, nchunks - k);
        j.pad = 0;
        P.jobs.push_back(j);
    }
    for (uint32_t k = 0; k < nchunks; ++k) {
        uint32_t cb = 0;
        for (int i = 0; i < 4; ++i) cb |= (uint32_t)sizes[4 * k + i] << (8 * i);
        require(cb >= 4u * L, CHAIN_ERUNTIME, "RVC1 chunk too small");
        DecChunk ch;
        ch.sec = si;
        ch.bytes = cb;
        ch.byte_off = r.dev + r.pos;
        ch.first = (uint64_t)k * S;
        (void)r.bytes(cb);
        P.chunks.push_back(ch);
    }
    r.finish();
}

static void parse_component(HostReader r, int comp, uint32_t expected_dims, uint64_t max_records, uint64_t dst,
                            DecPlan& P, bool raw_only, uint64_t* n_out) {
    const uint8_t* magic = r.bytes(4);
    const uint32_t dims = r.u32();
    const uint64_t n = r.u64();
    require(dims == expected_dims && n <= max_records, CHAIN_ERUNTIME, "LCPS component shape mismatch");
    *n_out = n;
    if (raw_only) require(std::memcmp(magic, "RAW1", 4) == 0, CHAIN_ERUNTIME, "LCP3 requires raw component streams");
    if (std::memcmp(magic, "RAW1", 4) == 0) {
        require(r.end - r.pos == n * dims * 8, CHAIN_ERUNTIME, "RAW1 matrix size mismatch");
        if (n) {
            DecRaw j;
            std::memset(&j, 0, sizeof j);
            j.byte_off = r.dev + r.pos;
            j.count = n * dims;
            j.dst = dst;
            j.comp = comp;
            j.first_thread = P.raw_total;
            P.raw_total += j.count;
            P.raws.push_back(j);
        }
        return;
    }
    if (std::memcmp(magic, "FIXP", 4) == 0) {
        DecFixed j;
        std::memset(&j, 0, sizeof j);
        uint64_t total = 0;
        for (uint32_t d = 0; d < dims; ++d) {
            j.minimum[d] = (uint64_t)r.i64();
            j.width[d] = r.u8();
            const uint64_t bytes = r.u64();
            require(j.width[d] >= 1 && j.width[d] <= 64, CHAIN_ERUNTIME, "invalid fixed width");
            require(bytes == (n * j.width[d] + 7) / 8, CHAIN_ERUNTIME, "invalid fixed payload size");
            require((int64_t)j.minimum[d] >= 0, CHAIN_ERUNTIME, "fixed value overflow");
            j.byte_off[d] = total;
            total += bytes;
        }
        require(r.end - r.pos == total, CHAIN_ERUNTIME, "fixed raw size mismatch");
        for (uint32_t d = 0; d < dims; ++d) {
            j.byte_off[d] += r.dev + r.pos;
            j.end[d] = j.byte_off[d] + (n * j.width[d] + 7) / 8;
        }
        j.n = n;
        j.dims = dims;
        j.dst = dst;
        j.comp = comp;
        if (n) P.fixed.push_back(j);   // tiles are assigned once the batch is parsed
        return;
    }
    if (std::memcmp(magic, "RANS", 4) == 0) {
        parse_rvc1(r.sub(r.u64()), n * dims, comp, dst, P);
        r.finish();
        return;
    }
    require(std::memcmp(magic, "TAGR", 4) == 0, CHAIN_ERUNTIME, "unknown LCPS encoding");
    const uint32_t maxw = r.u8();
    require(n == 0 ? maxw == 0 : (maxw >= 1 && maxw <= 63), CHAIN_ERUNTIME, "Tagged maximum width mismatch");
    const uint64_t tag_off = P.tags;
    parse_rvc1(r.sub(r.u64()), n, -1, tag_off, P);
    P.tags += n;
    const uint64_t bytes = r.u64();
    const uint64_t at = r.dev + r.pos;
    (void)r.bytes(bytes);
    r.finish();
    if (n) {
        DecTagged j;
        std::memset(&j, 0, sizeof j);
        j.byte_off = at;
        j.bytes = bytes;
        j.n = n;
        j.tag_off = tag_off;
        j.dst = dst;
        j.dims = dims;
        j.tile0 = (uint32_t)P.tag_tiles;
        j.comp = comp;
        j.maxw = maxw;
        // narrow: values fit the u32 staging, payload and values fit the LDS
        j.recs = maxw <= 32 && maxw * dims <= 96 && dims <= 3 ? DEC_TAG_TILE2 : DEC_TAG_TILE;
        P.tag_tiles += cdiv(n, j.recs);
        P.tagged.push_back(j);
    }
}

// One component of an inflated LCG2 stream (lcps_gpu_decompress_ans.cu:
// 2377-2424): Fixed fields (i32 minimum >= 0, u8 width 1..31, u64 bytes)
// become a Fixed unpack job over the stored section in place; Tagged and Raw
// components need nvCOMP ANS and are refused.
static void parse_lcg2_component(HostReader r, int comp, uint32_t expected_dims, uint64_t max_records, uint64_t dst,
                                 DecPlan& P, uint64_t* n_out) {
    const uint8_t kind = r.u8();
    require(kind <= 3, CHAIN_ERUNTIME, "unknown LCG2 encoding");
    const uint32_t dims = r.u32();
    const uint64_t n = r.u64();
    require(dims == expected_dims && n <= max_records, CHAIN_ERUNTIME, "LCG2 component shape mismatch");
    *n_out = n;
    require(kind != 1 && kind != 3, CHAIN_ERUNTIME,
            "LCG2 Tagged and Raw components are coded with nvCOMP ANS, whose format is not supported");
    if (kind == 0) {
        require(n == 0 && r.u8() == 0, CHAIN_ERUNTIME, "invalid empty LCG2 component");
        r.finish();
        return;
    }
    require(n > 0, CHAIN_ERUNTIME, "nonempty LCG2 encoding with no records");
    DecFixed j;
    std::memset(&j, 0, sizeof j);
    uint64_t total = 0;
    for (uint32_t d = 0; d < dims; ++d) {
        const int32_t mn = (int32_t)r.u32();
        j.width[d] = r.u8();
        const uint64_t bytes = r.u64();
        require(mn >= 0 && j.width[d] >= 1 && j.width[d] <= 31, CHAIN_ERUNTIME, "invalid LCG2 fixed field");
        require(bytes == (n * j.width[d] + 7) / 8, CHAIN_ERUNTIME, "LCG2 fixed length mismatch");
        j.minimum[d] = (uint64_t)mn;
        j.byte_off[d] = total;
        total += bytes;
    }
    require(r.u8() == 1, CHAIN_ERUNTIME, "LCG2 Fixed needs one section");
    const uint8_t codec = r.u8();
    require(codec != 2, CHAIN_ERUNTIME, "LCG2 Zstd sections are inflated on the host first (chain_lcg2_inflate)");
    require(codec == 0, CHAIN_ERUNTIME, "unexpected LCG2 section codec");
    require(r.u64() == total, CHAIN_ERUNTIME, "LCG2 section raw length mismatch");
    const uint64_t at = r.dev + r.pos;
    (void)r.bytes(total);
    r.finish();
    for (uint32_t d = 0; d < dims; ++d) {
        j.byte_off[d] += at;
        j.end[d] = j.byte_off[d] + (n * j.width[d] + 7) / 8;
    }
    j.n = n;
    j.dims = dims;
    j.dst = dst;
    j.comp = comp;
    P.fixed.push_back(j);   // tiles are assigned once the batch is parsed
}

// An LCG2 stream of the reference GPU compressor (cu/chain/lcps_gpu_cpp.cu:
// 5-35) whose Zstd sections the host has inflated, read with the checks of
// lcg2::parse (cu/decompress/lcps_gpu_decompress_ans.cu:2426-2452). Its
// integer streams have the meaning of the LCP2 ones; the frame is rebuilt
// in float32 (DecFrame.lcg2) and has no exact corrections.
static void parse_lcg2(HostReader& r, DecPlan& P) {
    require(r.u32() == 1, CHAIN_ERUNTIME, "unsupported LCG2 version");
    DecFrame F;
    std::memset(&F, 0, sizeof F);
    const uint32_t D = r.u32();
    F.n = r.u64();
    const int32_t p = (int32_t)r.u32();
    const uint32_t ebits = r.u32();
    float eb;
    std::memcpy(&eb, &ebits, 4);
    const int32_t wmin = (int32_t)r.u32(), wmax = (int32_t)r.u32();
    require(D > 0 && F.n > 0 && F.n <= (uint64_t)INT32_MAX && p > 0, CHAIN_ERUNTIME,
            "invalid LCG2 shape or block size");
    require(D <= CH_MAXD, CHAIN_EINVAL, "the GPU engine supports 1 to 4 dimensions");
    require(P.D == 0 || P.D == D, CHAIN_EINVAL, "every stream of a batch must have the same dimensions");
    P.D = D;
    require(eb > 0.0f && std::isfinite(2.0f * eb), CHAIN_ERUNTIME, "invalid LCG2 error bound");
    require(wmin >= 0 && wmax >= wmin, CHAIN_ERUNTIME, "invalid LCG2 wrap range");
    F.p = (uint64_t)p;
    F.eb = eb;
    F.denom = 2.0 * eb;
    F.eb_f = eb;
    F.lcg2 = 1;
    F.wrap_min = wmin;
    F.wrap_max = wmax;
    F.range = (uint64_t)(wmax - wmin) + 1;
    F.max_corr = (int64_t)(F.range / 2);
    F.min_corr = -F.max_corr;
    if ((F.range & 1) == 0) --F.max_corr;
    for (uint32_t d = 0; d < D; ++d) {
        const uint32_t b = r.u32();
        float m;
        std::memcpy(&m, &b, 4);
        require(std::isfinite(m), CHAIN_ERUNTIME, "nonfinite LCG2 minimum");
        F.mins_f[d] = m;
        F.mins[d] = m;
    }
    require(r.u32() == 5, CHAIN_ERUNTIME, "expected five LCG2 components");
    F.pt_off = P.T;
    F.out_off = P.T * D;
    static const char* names[5] = {"IDR0", "DIF0", "KEP0", "CH20", "REL0"};
    const uint32_t cd[5] = {1, D, D, 1, D};
    uint64_t cn[5], dst[5];
    dst[0] = F.pt_off;
    dst[4] = F.pt_off * D;
    for (int c = 0; c < 5; ++c) {
        require(std::memcmp(r.bytes(4), names[c], 4) == 0, CHAIN_ERUNTIME, "unexpected LCG2 component name");
        HostReader cr = r.sub(r.u64());
        if (c == 1) dst[1] = P.ND * D;
        if (c == 2) dst[2] = P.NK * D;
        if (c == 3) dst[3] = P.R;
        parse_lcg2_component(cr, c, cd[c], F.n, dst[c], P, &cn[c]);
        if (c == 1) { F.diff_base = P.ND; F.ndiff = cn[1]; P.ND += cn[1]; }
        if (c == 2) { F.keep_base = P.NK; F.nkeep = cn[2]; P.NK += cn[2]; }
        if (c == 3) { F.rec_base = P.R; F.nrec = cn[3]; P.R += cn[3]; }
    }
    r.finish();
    require(cn[0] == F.n && cn[4] == F.n, CHAIN_ERUNTIME, "LCG2 coordinate stream count mismatch");
    require(cn[3] > 0 && cn[2] > 0 && cn[1] + cn[2] == cn[3], CHAIN_ERUNTIME, "LCG2 diff/keep count mismatch");
    P.T += F.n;
    P.frames.push_back(F);
}

static void parse_stream(const uint8_t* bytes, uint64_t dev_off, uint64_t len, DecPlan& P) {
    HostReader r{bytes, 0, len, dev_off};
    const uint8_t* magic = r.bytes(4);
    if (std::memcmp(magic, "LCG2", 4) == 0) {
        parse_lcg2(r, P);
        return;
    }
    const bool lch1 = std::memcmp(magic, "LCH1", 4) == 0, lcp3 = std::memcmp(magic, "LCP3", 4) == 0;
    require(lch1 || lcp3 || std::memcmp(magic, "LCP2", 4) != 0, CHAIN_ERUNTIME,
            "LCP2 streams carry Zstd sections; decode them with lcps_cpu_decompress");
    require(lch1 || lcp3, CHAIN_ERUNTIME, "expected a self-contained LCH1, LCP3 or inflated LCG2 bitstream");
    DecFrame F;
    std::memset(&F, 0, sizeof F);
    const uint32_t D = r.u32();
    F.n = r.u64();
    require(D > 0 && F.n > 0, CHAIN_ERUNTIME, "invalid LCP2 shape");
    require(D <= CH_MAXD, CHAIN_EINVAL, "the GPU engine supports 1 to 4 dimensions");
    require(P.D == 0 || P.D == D, CHAIN_EINVAL, "every stream of a batch must have the same dimensions");
    P.D = D;
    const int64_t p = r.i64();
    F.eb = r.f64();
    require(p > 0 && F.eb > 0.0 && std::isfinite(2.0 * F.eb), CHAIN_ERUNTIME, "invalid LCP2 quantization parameters");
    F.p = (uint64_t)p;
    F.denom = 2.0 * F.eb;
    F.wrap_min = r.i64();
    F.wrap_max = r.i64();
    require(F.wrap_min >= 0, CHAIN_ERUNTIME, "invalid LCP2 block range");
    require(F.wrap_min <= F.wrap_max, CHAIN_ERUNTIME, "invalid block-ID wrap range");
    const uint64_t diffw = (uint64_t)(F.wrap_max - F.wrap_min);
    require(diffw < (uint64_t)INT64_MAX, CHAIN_ERUNTIME, "block-ID range is too wide for wrapped differences");
    F.range = diffw + 1;
    F.max_corr = (int64_t)(F.range / 2);
    F.min_corr = -F.max_corr;
    if ((F.range & 1) == 0) --F.max_corr;
    for (uint32_t d = 0; d < D; ++d) {
        F.mins[d] = r.f64();
        require(std::isfinite(F.mins[d]), CHAIN_ERUNTIME, "invalid LCP2 coordinate minimum");
    }
    F.pt_off = P.T;
    F.out_off = P.T * D;
    const uint32_t cd[5] = {1, D, D, 1, D};
    uint64_t cn[5];
    uint64_t dst[5];
    // decode targets: choice2/diff/keep bases are only known after their
    // counts, so parse in order and fix the record-space targets as we go
    dst[0] = F.pt_off;
    dst[4] = F.pt_off * D;
    for (int c = 0; c < 5; ++c) {
        HostReader cr = r.sub(r.u64());
        if (c == 1) dst[1] = P.ND * D;
        if (c == 2) dst[2] = P.NK * D;
        if (c == 3) dst[3] = P.R;
        parse_component(cr, c, cd[c], F.n, dst[c], P, lcp3, &cn[c]);
        if (c == 1) { F.diff_base = P.ND; F.ndiff = cn[1]; P.ND += cn[1]; }
        if (c == 2) { F.keep_base = P.NK; F.nkeep = cn[2]; P.NK += cn[2]; }
        if (c == 3) { F.rec_base = P.R; F.nrec = cn[3]; P.R += cn[3]; }
    }
    require(cn[0] == F.n && cn[4] == F.n, CHAIN_ERUNTIME, "LCP2 point count mismatch");
    require(F.ndiff + F.nkeep == F.nrec, CHAIN_ERUNTIME, "unused block differences or kept blocks");
    F.ncorr = r.u64();
    require(F.ncorr <= F.n * D, CHAIN_ERUNTIME, "invalid LCP2 correction count");
    HostReader er = r.sub(r.u64());
    require(er.end - er.pos == 16 * F.ncorr, CHAIN_ERUNTIME, "raw exact correction size mismatch");
    F.corr_off = er.dev + er.pos;
    r.finish();
    P.T += F.n;
    P.frames.push_back(F);
}

struct DecodeResult {
    bool keep_on_device = false;
    chain_buf* host_out = nullptr;     // caller's bufs, filled directly (one per stream)
    const uint8_t* d_pool = nullptr;   // device copy of the concatenated streams (skips the upload)
    // no-allocation form: intermediates carved from this fixed device region
    // and the coordinates written to the caller's device buffer
    uint8_t* ext_arena = nullptr;
    size_t ext_cap = 0;
    void* ext_out = nullptr;
    uint64_t ext_out_cap = 0;
    std::vector<uint64_t> ns, out_off;
    uint32_t dims = 0;
    void* d_out = nullptr;
};

static void decompress_batch_impl(chain_engine* e, const uint8_t* const* bytes, const size_t* lens, uint32_t nstreams,
                                  int out_dtype, int out_layout, DecodeResult& res) {
    require(nstreams >= 1, CHAIN_EINVAL, "empty batch");
    require(out_dtype == CHAIN_F32 || out_dtype == CHAIN_F64, CHAIN_EINVAL, "output dtype must be float32 or float64");
    require(out_layout == CHAIN_AOS || out_layout == CHAIN_SOA, CHAIN_EINVAL, "layout must be AoS or SoA");
    hipStream_t st = e->stream;
    StageTimer& tm = e->timer;
    tm.reset();
    e->stage.reset(st);
    // device pool = the streams concatenated; parsed in place on the host
    std::vector<uint64_t> in_off(nstreams + 1, 0);
    for (uint32_t k = 0; k < nstreams; ++k) in_off[k + 1] = in_off[k] + lens[k];
    const uint64_t pool_size = in_off[nstreams];
    DecPlan P;
    for (uint32_t k = 0; k < nstreams; ++k) parse_stream(bytes[k], in_off[k], lens[k], P);
    const uint32_t F = nstreams, D = P.D;
    require(P.T < (1ull << 31), CHAIN_EINVAL, "a GPU decode batch holds fewer than 2^31 particles");
    // in-block offsets fit a byte when every frame's p is at most 256
    uint32_t rel8 = 1;
    for (const DecFrame& fd : P.frames) rel8 &= fd.p <= 256 ? 1u : 0u;
    // Fixed jobs' tiles: byte outputs (choice flags; offsets when they are u8)
    // whose every value lands in range take the byte path in tiles of
    // DEC_FIX_BRECS records, the others tiles of DEC_FIX_RECS
    P.fixed_tiles = 0;
    for (DecFixed& j : P.fixed) {
        const uint64_t vmax = j.comp == 3 ? 1u : 0xffu;
        bool bytes = j.comp == 3 || (j.comp == 4 && rel8);
        for (uint32_t d = 0; d < j.dims; ++d)
            bytes = bytes && j.width[d] <= 8 && j.minimum[d] <= vmax &&
                    (1ull << j.width[d]) - 1ull <= vmax - j.minimum[d];
        j.recs = bytes ? DEC_FIX_BRECS : DEC_FIX_RECS;
        j.tile0 = (uint32_t)P.fixed_tiles;
        P.fixed_tiles += cdiv(j.n, j.recs);
    }
    // tiles: points (chain resolution) and records (L2 inverse), per frame
    std::vector<uint32_t> ptb(F + 1, 0), rtb(F + 1, 0), ptb32(F + 1, 0);
    for (uint32_t f = 0; f < F; ++f) {
        P.frames[f].tile0 = ptb[f];
        P.frames[f].rtile0 = rtb[f];
        ptb[f + 1] = ptb[f] + (uint32_t)cdiv(P.frames[f].n, CH_TILE);
        rtb[f + 1] = rtb[f] + (uint32_t)cdiv(P.frames[f].nrec, CH_TILE);
        ptb32[f + 1] = (uint32_t)(P.frames[f].pt_off + P.frames[f].n);
    }
    const uint32_t npt = ptb[F], nrt = rtb[F];
    const size_t es = out_dtype == CHAIN_F32 ? 4 : 8;
    const uint64_t T = P.T;
    std::vector<uint64_t> seg_base_rec(F), seg_n_rec(F), seg_base_pt(F), seg_n_pt(F);
    for (uint32_t f = 0; f < F; ++f) {
        seg_base_rec[f] = P.frames[f].rec_base;
        seg_n_rec[f] = P.frames[f].nrec;
        seg_base_pt[f] = P.frames[f].pt_off;
        seg_n_pt[f] = P.frames[f].n;
    }
    auto carve = [&](Carver& c) {
        struct {
            uint8_t* in;
            DecFrame* fr;
            DecSection* secs;
            DecChunk* chunks;
            DecJob* jobs;
            DecFixed* fixed;
            DecRaw* raws;
            DecTagged* tagged;
            uint32_t* cum;
            uint32_t* idrel;
            void* rel;
            uint64_t *diff, *keep;
            uint8_t *choice, *tags;
            uint64_t* compact;
            uint32_t* root;
            uint32_t *zc, *zoff, *eoff;
            uint64_t* cstatus;   // look-back words of the chain tiles
            uint32_t *ptb, *rtb, *ptb32;
            uint64_t *sbr, *snr, *sbp, *snp;
            uint64_t* status;
            uint32_t* flags;
            uint64_t* aggv;
            uint32_t* done;
            uint32_t* jump_open;   // per-round open-list counts
            uint32_t *list_a, *list_b;
            uint32_t* ticket;
            uint32_t* err;
            uint64_t* totals;
            uint8_t* out;
        } x;
        x.in = c.take<uint8_t>(res.d_pool ? 0 : pool_size);
        x.fr = c.take<DecFrame>(F);
        x.secs = c.take<DecSection>(P.secs.size());
        x.chunks = c.take<DecChunk>(P.chunks.size());
        x.jobs = c.take<DecJob>(P.jobs.size());
        x.fixed = c.take<DecFixed>(P.fixed.size());
        x.raws = c.take<DecRaw>(P.raws.size());
        x.tagged = c.take<DecTagged>(P.tagged.size());
        x.cum = c.take<uint32_t>(P.cum.size());
        // the tables above and these frame arrays are uploaded as one range
        x.ptb = c.take<uint32_t>(F + 1);
        x.rtb = c.take<uint32_t>(F + 1);
        x.ptb32 = c.take<uint32_t>(F + 1);
        x.sbr = c.take<uint64_t>(F);
        x.snr = c.take<uint64_t>(F);
        x.sbp = c.take<uint64_t>(F);
        x.snp = c.take<uint64_t>(F);
        x.idrel = c.take<uint32_t>(T);
        x.rel = c.take<uint8_t>(T * D * (rel8 ? 1 : 4));
        x.diff = c.take<uint64_t>((P.ND + 1) * D);   // one spare record each: unconditional loads
        x.keep = c.take<uint64_t>((P.NK + 1) * D);
        x.choice = c.take<uint8_t>(P.R);
        x.tags = c.take<uint8_t>(P.tags);
        x.compact = c.take<uint64_t>((P.R + 1) * D);   // one spare record: unconditional gathers
        x.root = c.take<uint32_t>(T);
        x.zc = c.take<uint32_t>(nrt + 1);
        x.zoff = c.take<uint32_t>(nrt + 1);
        x.cstatus = c.take<uint64_t>(npt + 1);
        x.eoff = c.take<uint32_t>(npt + 1);
        x.status = c.take<uint64_t>(std::max<uint64_t>({P.tag_tiles, cdiv(npt + 1, CH_TILE), cdiv(nrt + 1, CH_TILE)}) + 1);
        x.flags = c.take<uint32_t>(nrt + 1);
        x.aggv = c.take<uint64_t>((uint64_t)(nrt + 1) * 2 * CH_MAXD);
        x.done = c.take<uint32_t>(npt + 1);
        x.jump_open = c.take<uint32_t>(64);
        x.list_a = c.take<uint32_t>(T);
        x.list_b = c.take<uint32_t>(T);
        x.ticket = c.take<uint32_t>(4);
        x.err = c.take<uint32_t>(4);
        x.totals = c.take<uint64_t>(4);
        x.out = c.take<uint8_t>(res.ext_out ? 0 : T * D * es);
        return x;
    };
    Carver c0;
    (void)carve(c0);
    Carver c1;
    if (res.ext_arena) {
        require(c0.off <= res.ext_cap, CHAIN_ENOMEM,
                ("the batch needs " + std::to_string(c0.off) + " bytes of device scratch; the no-allocation entry "
                 "point has " + std::to_string(res.ext_cap)).c_str());
        require(T * D * es <= res.ext_out_cap, CHAIN_EINVAL, "the output buffer is smaller than the decoded coordinates");
        c1.base = reinterpret_cast<char*>(res.ext_arena);
    } else {
        ensure_arena(e, e->arena_d, e->arena, c0.off);
        c1.base = e->arena_d.base;
    }
    auto X = carve(c1);
    if (res.ext_out) X.out = static_cast<uint8_t*>(res.ext_out);
    if (res.d_pool) X.in = const_cast<uint8_t*>(res.d_pool);
    Staging& sg = e->stage;
    // algorithmic bytes of every timed region, from the plan
    // diff and keep as u32 when every frame's wrap range lies in [0, 2^31)
    bool dk32 = true;
    for (const DecFrame& fd : P.frames)
        dk32 = dk32 && fd.wrap_min >= 0 && fd.wrap_max < (int64_t)0x80000000ll && fd.range <= (1ull << 31);
    const uint64_t dkes = dk32 ? 4 : 8;
    const uint64_t comp_es[5] = {4, dkes, dkes, 1, rel8 ? 1u : 4u};
    auto es_of = [&](int32_t comp) -> uint64_t { return comp < 0 ? 1 : comp_es[comp]; };
    uint64_t b_rans = 0, b_fixed = 0, b_raw = 0, b_tagged = 0;
    for (const DecChunk& ch : P.chunks) b_rans += ch.bytes;
    for (const DecSection& sc : P.secs) b_rans += sc.count * es_of(sc.comp);
    for (const DecFixed& j : P.fixed) {
        for (uint32_t d = 0; d < j.dims; ++d) b_fixed += (j.n * j.width[d] + 7) / 8;
        b_fixed += j.n * j.dims * es_of(j.comp);
    }
    for (const DecRaw& j : P.raws) b_raw += j.count * (8 + es_of(j.comp));
    for (const DecTagged& j : P.tagged) b_tagged += j.bytes + j.n + j.n * j.dims * es_of(j.comp);
    tm.begin("h2d", res.d_pool ? 0 : pool_size);
    if (!res.d_pool)
        for (uint32_t k = 0; k < nstreams; ++k)
            if (lens[k]) HIPCHK(hipMemcpyAsync(X.in + in_off[k], bytes[k], lens[k], hipMemcpyHostToDevice, st));
    sg.h2d_range(st, {{X.fr, P.frames.data(), F * sizeof(DecFrame)},
                      {X.secs, P.secs.data(), P.secs.size() * sizeof(DecSection)},
                      {X.chunks, P.chunks.data(), P.chunks.size() * sizeof(DecChunk)},
                      {X.fixed, P.fixed.data(), P.fixed.size() * sizeof(DecFixed)},
                      {X.raws, P.raws.data(), P.raws.size() * sizeof(DecRaw)},
                      {X.tagged, P.tagged.data(), P.tagged.size() * sizeof(DecTagged)},
                      {X.jobs, P.jobs.data(), P.jobs.size() * sizeof(DecJob)},
                      {X.cum, P.cum.data(), P.cum.size() * 4},
                      {X.ptb, ptb.data(), (F + 1) * 4},
                      {X.rtb, rtb.data(), (F + 1) * 4},
                      {X.ptb32, ptb32.data(), (F + 1) * 4},
                      {X.sbr, seg_base_rec.data(), F * 8},
                      {X.snr, seg_n_rec.data(), F * 8},
                      {X.sbp, seg_base_pt.data(), F * 8},
                      {X.snp, seg_n_pt.data(), F * 8}});
    fill32(st, X.err, 4, 0);
    tm.end();

    tm.begin("rans_decode", b_rans);
    if (!P.chunks.empty())
        hipLaunchKernelGGL(k_dec_rans, dim3((unsigned)P.jobs.size()), dim3(64 * DEC_WAVES), 0, st, X.in, X.secs,
                           X.chunks, X.jobs, X.cum, X.tags, X.idrel, (void*)X.diff, (void*)X.keep, (uint32_t)dk32,
                           X.choice, X.rel, rel8, X.err);
    tm.end();
    if (P.fixed_tiles) {
        tm.begin("unpack_fixed", b_fixed);
        hipLaunchKernelGGL(k_dec_fixed, dim3((unsigned)cdiv(P.fixed_tiles, DEC_FIX_TPW)), dim3(CH_WG), 0, st, X.in,
                           X.fixed, (uint32_t)P.fixed.size(), (uint32_t)P.fixed_tiles, X.idrel, (void*)X.diff,
                           (void*)X.keep, (uint32_t)dk32, X.choice, X.rel, rel8, X.err);
        tm.end();
    }
    if (P.raw_total) {
        tm.begin("unpack_raw", b_raw);
        hipLaunchKernelGGL(k_dec_raw, dim3(grid_for(P.raw_total)), dim3(CH_WG), 0, st, X.in, X.raws,
                           (uint32_t)P.raws.size(), P.raw_total, X.idrel, (void*)X.diff, (void*)X.keep, (uint32_t)dk32,
                           X.choice, X.rel, rel8, X.err);
        tm.end();
    }
    if (P.tag_tiles) {
        tm.begin("unpack_tagged", b_tagged);
        fill64(st, X.status, P.tag_tiles, 0);
        fill32(st, X.ticket, 1, 0);
        hipLaunchKernelGGL(k_dec_tagged, dim3((unsigned)P.tag_tiles), dim3(CH_WG), 0, st, X.in, X.tagged,
                           (uint32_t)P.tagged.size(), X.tags, X.status, X.ticket, X.idrel, (void*)X.diff,
                           (void*)X.keep, (uint32_t)dk32, X.choice, X.rel, rel8, X.err);
        tm.end();
    }
    if (nrt) {
        tm.begin("l2_counts", P.R + 8ull * nrt);
        hipLaunchKernelGGL(k_dec_tile_zero_counts, dim3(nrt), dim3(CH_WG), 0, st, X.choice, (const uint32_t*)nullptr,
                           X.rtb, X.sbr, X.snr, F, X.zc);
        scan_u32(st, X.zc, X.zoff, nrt, X.status, X.ticket, X.err, &X.totals[0]);
        fill32(st, X.flags, nrt + 1, 0);
        fill32(st, X.ticket, 1, 0);
        tm.end();
        tm.begin("l2_inverse", P.R + dkes * D * (P.ND + P.NK) + dkes * D * P.R);
        // 32-bit scan arithmetic when every frame's wrap range is at most 2^31
        uint64_t max_range = 0;
        for (const DecFrame& fd : P.frames) max_range = std::max(max_range, fd.range);
        with_dims(D, [&](auto dd) {
            constexpr int DD = decltype(dd)::value;
            auto inverse = [&](auto vt, auto dkt) {
                using V = typename decltype(vt)::type;
                using DKT = typename decltype(dkt)::type;
                hipLaunchKernelGGL((k_dec_l2<DD, V, DKT>), dim3(nrt), dim3(CH_WG), 0, st, X.fr, F, X.rtb, X.choice,
                                   (const DKT*)X.diff, (const DKT*)X.keep, X.zoff, X.flags, X.aggv, X.ticket,
                                   (DKT*)X.compact, X.err);
            };
            // dk32 implies a range of at most 2^31
            if (dk32) inverse(TypeTag<uint32_t>{}, TypeTag<uint32_t>{});
            else if (max_range <= (1ull << 31)) inverse(TypeTag<uint32_t>{}, TypeTag<uint64_t>{});
            else inverse(TypeTag<uint64_t>{}, TypeTag<uint64_t>{});
        });
        tm.end();
    }
    // the O(N) algorithmic traffic of chain resolution (read successors,
    // write roots) is stated once, on the local pass; the jump rounds over
    // tile seams are extra work of this formulation
    tm.begin("chain_local", 8 * T);
    fill32(st, X.jump_open, 8, 0);
    // the local pass ranks the explicit points itself: tiles by ticket, each
    // tile's first rank from a look-back over the tiles before it
    fill64(st, X.cstatus, npt, 0);
    fill32(st, X.ticket + 1, 1, 0);
    with_dims(D, [&](auto dd) {
        constexpr int DD = decltype(dd)::value;
        hipLaunchKernelGGL((k_dec_chain_local<DD>), dim3(npt), dim3(CH_WG), 0, st, X.fr, F, X.ptb, X.idrel, X.eoff,
                           X.cstatus, X.ticket + 1, &X.totals[1], X.root, X.list_a, &X.jump_open[0], X.err);
    });
    hipLaunchKernelGGL(k_dec_check_counts, dim3(cdiv(F, CH_WG)), dim3(CH_WG), 0, st, X.fr, F, X.ptb, X.eoff,
                       &X.totals[1], X.err);
    tm.end();
    // The exits: a few lock-step rounds of pointer jumping over the open list
    // (each round halves the remaining hops of every chain and drops the
    // settled exits), then one launch walks what is left, by then short.
    tm.begin("chain_jump");
    {
        const uint32_t sync_rounds = 4;
        const unsigned g = (unsigned)std::min<uint64_t>(cdiv(T, CH_TILE), 2048);
        for (uint32_t r = 0; r < sync_rounds; ++r) {
            uint32_t* lin = (r & 1) ? X.list_b : X.list_a;
            uint32_t* lout = (r & 1) ? X.list_a : X.list_b;
            hipLaunchKernelGGL(k_dec_chain_jump, dim3(g), dim3(CH_WG), 0, st, X.root, lin, &X.jump_open[r], lout,
                               &X.jump_open[r + 1]);
        }
        hipLaunchKernelGGL(k_dec_chain_resolve, dim3(g), dim3(CH_WG), 0, st, X.root,
                           (sync_rounds & 1) ? X.list_b : X.list_a, &X.jump_open[sync_rounds], X.err);
    }
    tm.end();
    tm.begin("dequantize", 4 * T + comp_es[4] * T * D + dkes * D * P.R + T * D * es);
    with_dims(D, [&](auto dd) {
        constexpr int DD = decltype(dd)::value;
        auto go = [&](auto ot, auto rt) {
            using O = typename decltype(ot)::type;
            using RT = typename decltype(rt)::type;
            auto run = [&](auto bt) {
                using BT = typename decltype(bt)::type;
                if (out_layout == CHAIN_SOA) {
                    hipLaunchKernelGGL((k_dec_dequant<DD, O, true, RT, BT>), dim3(npt), dim3(CH_WG), 0, st, X.fr, F,
                                       X.ptb, X.root, (const BT*)X.compact, (const RT*)X.rel, (O*)X.out, X.err);
                    hipLaunchKernelGGL((k_dec_corr<O, true>), dim3(64, F), dim3(CH_WG), 0, st, X.in, X.fr, F, D,
                                       (O*)X.out, X.err);
                } else {
                    hipLaunchKernelGGL((k_dec_dequant<DD, O, false, RT, BT>), dim3(npt), dim3(CH_WG), 0, st, X.fr, F,
                                       X.ptb, X.root, (const BT*)X.compact, (const RT*)X.rel, (O*)X.out, X.err);
                    hipLaunchKernelGGL((k_dec_corr<O, false>), dim3(64, F), dim3(CH_WG), 0, st, X.in, X.fr, F, D,
                                       (O*)X.out, X.err);
                }
            };
            if (dk32) run(TypeTag<uint32_t>{});
            else run(TypeTag<uint64_t>{});
        };
        auto go_rel = [&](auto ot) {
            if (rel8) go(ot, TypeTag<uint8_t>{});
            else go(ot, TypeTag<uint32_t>{});
        };
        if (out_dtype == CHAIN_F32) go_rel(TypeTag<float>{});
        else go_rel(TypeTag<double>{});
    });
    tm.end();
    uint32_t errw[4];
    sg.d2h(st, errw, X.err, sizeof errw);
    sg.finish(st);
    if (errw[0]) {
        std::string m = "invalid LCPS bitstream (device checks:";
        if (errw[0] & DERR_RANS) m += " rANS";
        if (errw[0] & DERR_VALUE) m += " value";
        if (errw[0] & DERR_L2) m += " block";
        if (errw[0] & DERR_CHAIN) m += " chain";
        if (errw[0] & DERR_QUANT) m += " offset";
        if (errw[0] & DERR_CORR) m += " correction";
        if (errw[0] & DERR_COUNT) m += " count";
        if (errw[0] & (DERR_WAIT | 1u)) m += " wait";
        throw Error(errw[0] & (DERR_WAIT | 1u) ? CHAIN_EDEVICE : CHAIN_ERUNTIME, m + ")");
    }
    res.ns.resize(F);
    res.out_off.resize(F);
    res.dims = D;
    for (uint32_t f = 0; f < F; ++f) {
        res.ns[f] = P.frames[f].n;
        res.out_off[f] = P.frames[f].out_off;
    }
    if (res.keep_on_device) {
        res.d_out = X.out;
        return;
    }
    if (res.host_out) {
        std::vector<uint64_t> boffs(F), bsizes(F);
        for (uint32_t f = 0; f < F; ++f) {
            boffs[f] = P.frames[f].out_off * es;
            bsizes[f] = P.frames[f].n * D * es;
        }
        device_to_bufs(st, X.out, boffs.data(), bsizes.data(), F, res.host_out);
        return;
    }
    throw Error(CHAIN_EINVAL, "decoded coordinates need a host or device destination");
}
// ------------------------------------------------------------------------
// autotuning of the entropy stage's chunking

struct TuneTrial {
    Tuning t;
    double compress_ms, decompress_ms;
    uint64_t bytes;
};

// Time compression + decompression of (a sample of) the batch under
// candidate settings and keep the fastest whose output is within 2 % of the
// smallest. Frames are taken in order up to 2^24 points (at least one). The
// search is coordinate-wise: 8192 / 16384 / 32768 / 65536 symbols per chunk
// with 64 lanes and tile-local L1; then 32 lanes, and the global-sort L1, at
// the best chunk size so far. Containers are valid under every setting.
static Tuning autotune_impl(chain_engine* e, const chain_frame* frames, uint32_t F, int dtype, int layout,
                            uint32_t D, bool on_device, std::vector<TuneTrial>& trials) {
    uint32_t Fs = 0;
    uint64_t pts = 0;
    while (Fs < F && (Fs == 0 || pts + frames[Fs].n <= (1ull << 24))) pts += frames[Fs++].n;
    const Tuning saved = e->tuning;
    const bool timer_was = e->timer.on;
    const bool graphs_was = e->use_graphs;
    e->timer.on = false;
    e->use_graphs = false;
    auto run = [&](const Tuning& t) {
        e->tuning = t;
        TuneTrial tr{t, 0, 0, 0};
        for (int rep = 0; rep < 2; ++rep) {   // the first run sizes the arenas
            CompressResult cr;
            cr.keep_on_device = true;
            auto t0 = std::chrono::steady_clock::now();
            compress_batch_impl(e, frames, Fs, dtype, layout, D, on_device, CHAIN_FMT_LCH1, cr);
            auto t1 = std::chrono::steady_clock::now();
            std::vector<std::vector<uint8_t>> host(Fs);
            std::vector<const uint8_t*> ptrs(Fs);
            std::vector<size_t> lens(Fs);
            bool contiguous = true;
            uint64_t total = 0;
            for (uint32_t f = 0; f < Fs; ++f) {
                host[f].resize(cr.sizes[f]);
                if (cr.sizes[f])
                    HIPCHK(hipMemcpyAsync(host[f].data(), cr.d_out + cr.offsets[f], cr.sizes[f],
                                          hipMemcpyDeviceToHost, e->stream));
                contiguous = contiguous && cr.offsets[f] == cr.offsets[0] + total;
                total += cr.sizes[f];
            }
            HIPCHK(hipStreamSynchronize(e->stream));
            for (uint32_t f = 0; f < Fs; ++f) {
                ptrs[f] = host[f].data();
                lens[f] = host[f].size();
            }
            DecodeResult dr;
            dr.keep_on_device = true;
            dr.d_pool = contiguous ? cr.d_out + cr.offsets[0] : nullptr;
            auto t2 = std::chrono::steady_clock::now();
            decompress_batch_impl(e, ptrs.data(), lens.data(), Fs, CHAIN_F64, CHAIN_AOS, dr);
            auto t3 = std::chrono::steady_clock::now();
            tr.compress_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            tr.decompress_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
            tr.bytes = total;
        }
        trials.push_back(tr);
    };
    auto best_of = [&]() -> const TuneTrial& {
        uint64_t best_bytes = ~0ull;
        for (const TuneTrial& t : trials) best_bytes = std::min(best_bytes, t.bytes);
        const TuneTrial* best = nullptr;
        for (const TuneTrial& t : trials)
            if (t.bytes <= best_bytes + best_bytes / 50 &&
                (!best || t.compress_ms + t.decompress_ms < best->compress_ms + best->decompress_ms))
                best = &t;
        return *best;
    };
    try {
        for (uint32_t cs : {8192u, 16384u, 32768u, 65536u}) run(Tuning{64, cs, 1});
        Tuning b = best_of().t;
        run(Tuning{32, b.chunk_symbols, b.l1_tiles});
        b = best_of().t;
        run(Tuning{b.lanes_max, b.chunk_symbols, b.l1_tiles ? 0u : 1u});
    } catch (...) {
        e->tuning = saved;
        e->timer.on = timer_was;
        e->use_graphs = graphs_was;
        throw;
    }
    e->timer.on = timer_was;
    e->use_graphs = graphs_was;
    return best_of().t;
}

// automatic mode: one tuned setting per (log2 mean frame size, log2 frames)
static uint32_t tune_bucket(const chain_frame* frames, uint32_t F) {
    uint64_t T = 0;
    for (uint32_t f = 0; f < F; ++f) T += frames[f].n;
    const uint32_t a = (uint32_t)bitlen64(T / F), b = (uint32_t)bitlen64(F);
    return (a << 8) | b;
}

static void apply_autotune(chain_engine* e, const chain_frame* frames, uint32_t F, int dtype, int layout,
                           uint32_t D, bool on_device, int format) {
    if (!e->autotune || format != CHAIN_FMT_LCH1 || F == 0) return;
    const uint32_t key = tune_bucket(frames, F);
    auto it = e->tuned.find(key);
    if (it == e->tuned.end()) {
        std::vector<TuneTrial> trials;
        it = e->tuned.emplace(key, autotune_impl(e, frames, F, dtype, layout, D, on_device, trials)).first;
    }
    e->tuning = it->second;
}

}  // namespace chain

// ------------------------------------------------------------------------
// C ABI

using chain::Error;

static thread_local std::string g_create_err;

// pinned host staging reserved at engine creation; the no-allocation entry
// point's host tables must fit in it
#define CHAIN_NOALLOC_STAGING_BYTES (4ull << 20)

// Device scratch of the no-allocation entry point: a fixed region per GPU,
// held from module load to process end. In host code the array's name is only
// a stand-in for the symbol; its device address is looked up once per device
// (with that device current) when an engine is created, and only that address
// is handed to kernels.
#define CHAIN_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) uint8_t noalloc_scratch[CHAIN_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[CHAIN_NOALLOC_MAX_DEVICES];

// with `device` current: the device address of its scratch (looked up on the
// device's first engine), or nullptr with *why set
static void* scratch_for_device(int device, std::string* why) {
    if (device < 0 || device >= CHAIN_NOALLOC_MAX_DEVICES) {
        *why = "device index beyond the no-allocation scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch));
        if (rc != hipSuccess || !ptr) {
            *why = std::string("hipGetSymbolAddress(noalloc_scratch): ") + hipGetErrorString(rc);
            return nullptr;
        }
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

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
    // the no-allocation entry point's device scratch and pinned staging are
    // set up here, so that its calls allocate nothing
    std::string why;
    e->scratch = scratch_for_device(device, &why);
    if (!e->scratch) {
        g_create_err = why;
        chain_engine_destroy(e);
        return CHAIN_EDEVICE;
    }
    try {
        (void)e->stage.take(CHAIN_NOALLOC_STAGING_BYTES);
        e->stage.off = 0;
    } catch (const std::exception& x) {
        g_create_err = x.what();
        chain_engine_destroy(e);
        return CHAIN_ENOMEM;
    }
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
    if (b->data && !chain::pinned_pool().release(b->data)) std::free(b->data);
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
        chain::apply_autotune(e, &fr, 1, dtype, layout, dims, false, format);
        chain::CompressResult res;
        res.host_out = out;
        chain::compress_batch_impl(e, &fr, 1, dtype, layout, dims, false, format, res);
        if (!res.containers.empty()) to_buf(res.containers[0], out);
        if (stats) *stats = res.stats[0];
    });
}

int chain_compress_batch(chain_engine* e, const chain_frame* frames, uint32_t nframes, int dtype, int layout,
                         uint32_t dims, int points_on_device, int format, chain_buf* outs, chain_stats* stats) {
    return guarded(e, [&] {
        if (!frames || !outs) throw Error(CHAIN_EINVAL, "frames and outs must not be NULL");
        chain::apply_autotune(e, frames, nframes, dtype, layout, dims, points_on_device != 0, format);
        chain::CompressResult res;
        res.host_out = outs;
        chain::compress_batch_impl(e, frames, nframes, dtype, layout, dims, points_on_device != 0, format, res);
        for (uint32_t f = 0; f < nframes; ++f) {
            if (!res.containers.empty()) to_buf(res.containers[f], &outs[f]);
            if (stats) stats[f] = res.stats[f];
        }
    });
}

int chain_compress_batch_device(chain_engine* e, const chain_frame* frames, uint32_t nframes, int dtype, int layout,
                                uint32_t dims, void** d_out, uint64_t* offsets, uint64_t* sizes, chain_stats* stats) {
    return guarded(e, [&] {
        if (!frames || !d_out || !offsets || !sizes) throw Error(CHAIN_EINVAL, "NULL argument");
        chain::apply_autotune(e, frames, nframes, dtype, layout, dims, true, CHAIN_FMT_LCH1);
        chain::CompressResult res;
        res.keep_on_device = true;
        chain::compress_batch_impl(e, frames, nframes, dtype, layout, dims, true, CHAIN_FMT_LCH1, res);
        *d_out = res.d_out;
        for (uint32_t f = 0; f < nframes; ++f) {
            offsets[f] = res.offsets[f];
            sizes[f] = res.sizes[f];
            if (stats) stats[f] = res.stats[f];
        }
    });
}

static void check_tuning(uint32_t lanes_max, uint32_t chunk_symbols) {
    if (lanes_max < 1 || lanes_max > 64 || (lanes_max & (lanes_max - 1)))
        throw Error(CHAIN_EINVAL, "lanes_max must be a power of two in [1, 64]");
    if (chunk_symbols < 64 || chunk_symbols > (1u << 20))
        throw Error(CHAIN_EINVAL, "chunk_symbols must be in [64, 2^20]");
}

int chain_engine_set_rans(chain_engine* e, uint32_t lanes_max, uint32_t chunk_symbols) {
    return guarded(e, [&] {
        check_tuning(lanes_max, chunk_symbols);
        e->tuning.lanes_max = lanes_max;
        e->tuning.chunk_symbols = chunk_symbols;
    });
}

int chain_engine_set_tuning(chain_engine* e, const chain_tuning* t) {
    return guarded(e, [&] {
        if (!t) throw Error(CHAIN_EINVAL, "NULL argument");
        check_tuning(t->lanes_max, t->chunk_symbols);
        e->tuning = chain::Tuning{t->lanes_max, t->chunk_symbols, t->l1_tiles ? 1u : 0u};
    });
}

int chain_engine_get_tuning(chain_engine* e, chain_tuning* t) {
    return guarded(e, [&] {
        if (!t) throw Error(CHAIN_EINVAL, "NULL argument");
        *t = chain_tuning{e->tuning.lanes_max, e->tuning.chunk_symbols, e->tuning.l1_tiles, 0};
    });
}

int chain_decompress(chain_engine* e, const uint8_t* bytes, size_t len, int out_dtype, int out_layout,
                     chain_buf* out, uint64_t* n, uint32_t* dims) {
    return guarded(e, [&] {
        if (!bytes || !out) throw Error(CHAIN_EINVAL, "NULL argument");
        chain::DecodeResult res;
        res.host_out = out;
        chain::decompress_batch_impl(e, &bytes, &len, 1, out_dtype, out_layout, res);
        if (n) *n = res.ns[0];
        if (dims) *dims = res.dims;
    });
}

int chain_decompress_batch(chain_engine* e, const uint8_t* const* bytes, const size_t* lens, uint32_t nstreams,
                           int out_dtype, int out_layout, chain_buf* outs, uint64_t* ns, uint32_t* dims) {
    return guarded(e, [&] {
        if (!bytes || !lens || !outs) throw Error(CHAIN_EINVAL, "NULL argument");
        chain::DecodeResult res;
        res.host_out = outs;
        chain::decompress_batch_impl(e, bytes, lens, nstreams, out_dtype, out_layout, res);
        for (uint32_t k = 0; k < nstreams; ++k) {
            if (ns) ns[k] = res.ns[k];
            if (dims) dims[k] = res.dims;
        }
    });
}

int chain_decompress_batch_device(chain_engine* e, const uint8_t* const* bytes, const size_t* lens,
                                  const void* d_streams, uint32_t nstreams, int out_dtype, int out_layout,
                                  void** d_out, uint64_t* out_offsets, uint64_t* ns, uint32_t* dims) {
    return guarded(e, [&] {
        if (!bytes || !lens || !d_out || !out_offsets) throw Error(CHAIN_EINVAL, "NULL argument");
        chain::DecodeResult res;
        res.keep_on_device = true;
        res.d_pool = (const uint8_t*)d_streams;
        chain::decompress_batch_impl(e, bytes, lens, nstreams, out_dtype, out_layout, res);
        *d_out = res.d_out;
        for (uint32_t k = 0; k < nstreams; ++k) {
            out_offsets[k] = res.out_off[k];
            if (ns) ns[k] = res.ns[k];
            if (dims) dims[k] = res.dims;
        }
    });
}

int chain_decompress_batch_noalloc(chain_engine* e, const uint8_t* const* bytes, const size_t* lens,
                                   const void* d_streams, uint32_t nstreams, int out_dtype, int out_layout,
                                   void* d_out, uint64_t out_capacity, uint64_t* out_offsets, uint64_t* ns,
                                   uint32_t* dims) {
    return guarded(e, [&] {
        if (!bytes || !lens || !d_streams || !d_out || !out_offsets) throw Error(CHAIN_EINVAL, "NULL argument");
        if (!e->scratch) throw Error(CHAIN_EDEVICE, "the engine has no no-allocation scratch");
        chain::DecodeResult res;
        res.d_pool = (const uint8_t*)d_streams;
        res.ext_arena = static_cast<uint8_t*>(e->scratch);
        res.ext_cap = CHAIN_NOALLOC_SCRATCH_BYTES;
        res.ext_out = d_out;
        res.ext_out_cap = out_capacity;
        res.keep_on_device = true;
        // no event creation, no pinned staging growth or release in this call
        const bool timer_was = e->timer.on;
        e->timer.on = false;
        e->stage.fixed = true;
        try {
            chain::decompress_batch_impl(e, bytes, lens, nstreams, out_dtype, out_layout, res);
        } catch (...) {
            e->timer.on = timer_was;
            e->stage.fixed = false;
            throw;
        }
        e->timer.on = timer_was;
        e->stage.fixed = false;
        for (uint32_t k = 0; k < nstreams; ++k) {
            out_offsets[k] = res.out_off[k];
            if (ns) ns[k] = res.ns[k];
            if (dims) dims[k] = res.dims;
        }
    });
}

int chain_device_malloc(chain_engine* e, uint64_t bytes, void** out) {
    return guarded(e, [&] {
        if (!out) throw Error(CHAIN_EINVAL, "NULL argument");
        *out = nullptr;
        HIPCHK(hipMalloc(out, bytes ? bytes : 1));
    });
}

int chain_device_free(chain_engine* e, void* p) {
    return guarded(e, [&] { HIPCHK(hipFree(p)); });
}

int chain_copy(chain_engine* e, void* dst, const void* src, uint64_t bytes, int to_device) {
    return guarded(e, [&] {
        const hipMemcpyKind kind = to_device == 2 ? hipMemcpyDeviceToDevice
                                   : to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
        HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
    });
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

int chain_compress_search(chain_engine* e, const void* points, int dtype, int layout, uint32_t dims, uint64_t n,
                          double eb, const int64_t* ps, uint32_t nps, const int64_t* ts, uint32_t nts, uint32_t flags,
                          int points_on_device, chain_stats* stats) {
    return guarded(e, [&] {
        if (!points || !ps || !ts || !stats || nps == 0 || nts == 0) throw Error(CHAIN_EINVAL, "NULL or empty argument");
        if (n == 0) throw Error(CHAIN_ERUNTIME, "zero-size array to reduction operation minimum which has no identity");
        const size_t es = dtype == CHAIN_F32 ? 4 : 8;
        const uint64_t bytes = n * dims * es;
        // the frame once in device memory; every configuration reads it
        void* dpts = const_cast<void*>(points);
        void* owned = nullptr;
        if (!points_on_device) {
            HIPCHK(hipMalloc(&owned, bytes));
            HIPCHK(hipMemcpyAsync(owned, points, bytes, hipMemcpyHostToDevice, e->stream));
            dpts = owned;
        }
        try {
            const uint64_t total = (uint64_t)nps * nts;
            const uint64_t per_call = std::max<uint64_t>(1, (1ull << 27) / n);   // bounded device memory
            std::vector<chain_frame> fr;
            for (uint64_t k0 = 0; k0 < total; k0 += per_call) {
                const uint64_t k1 = std::min(total, k0 + per_call);
                fr.clear();
                for (uint64_t k = k0; k < k1; ++k)
                    fr.push_back(chain_frame{dpts, n, eb, ps[k / nts], ts[k % nts], flags});
                chain::CompressResult res;
                res.keep_on_device = true;
                chain::compress_batch_impl(e, fr.data(), (uint32_t)fr.size(), dtype, layout, dims, true,
                                           CHAIN_FMT_LCH1, res);
                for (uint64_t k = k0; k < k1; ++k) stats[k] = res.stats[k - k0];
            }
        } catch (...) {
            if (owned) {
                (void)hipStreamSynchronize(e->stream);
                (void)hipFree(owned);
            }
            throw;
        }
        if (owned) {
            HIPCHK(hipStreamSynchronize(e->stream));
            HIPCHK(hipFree(owned));
        }
    });
}

int chain_engine_set_autotune(chain_engine* e, int on) {
    return guarded(e, [&] { e->autotune = on != 0; });
}

int chain_engine_autotune(chain_engine* e, const chain_frame* frames, uint32_t nframes, int dtype, int layout,
                          uint32_t dims, int points_on_device, chain_tuning* best, chain_tuning* trials,
                          float* trial_ms, uint64_t* trial_bytes, uint32_t max_trials, uint32_t* ntrials) {
    return guarded(e, [&] {
        if (!frames || nframes == 0) throw Error(CHAIN_EINVAL, "empty batch");
        std::vector<chain::TuneTrial> tr;
        const chain::Tuning b = chain::autotune_impl(e, frames, nframes, dtype, layout, dims, points_on_device != 0, tr);
        e->tuning = b;
        e->tuned[chain::tune_bucket(frames, nframes)] = b;
        if (best) *best = chain_tuning{b.lanes_max, b.chunk_symbols, b.l1_tiles, 0};
        for (size_t k = 0; k < tr.size() && k < max_trials; ++k) {
            if (trials) trials[k] = chain_tuning{tr[k].t.lanes_max, tr[k].t.chunk_symbols, tr[k].t.l1_tiles, 0};
            if (trial_ms) trial_ms[k] = (float)(tr[k].compress_ms + tr[k].decompress_ms);
            if (trial_bytes) trial_bytes[k] = tr[k].bytes;
        }
        if (ntrials) *ntrials = (uint32_t)std::min<size_t>(tr.size(), max_trials);
    });
}

int chain_engine_set_graphs(chain_engine* e, int on) {
    return guarded(e, [&] { e->use_graphs = on != 0; });
}

int chain_engine_graph_stats(chain_engine* e, uint64_t* captures, uint64_t* replays) {
    return guarded(e, [&] {
        if (captures) *captures = e->graph.captures;
        if (replays) *replays = e->graph.replays;
    });
}

int chain_engine_set_timing(chain_engine* e, int on) {
    return guarded(e, [&] { e->timer.on = on != 0; });
}

int chain_last_stage_bytes(chain_engine* e, uint64_t* bytes, int max) {
    if (!e) return 0;
    std::lock_guard<std::mutex> lock(e->mu);
    const int k = (int)(e->timer.used / 2);
    for (int i = 0; i < k && i < max; ++i)
        if (bytes) bytes[i] = e->timer.bytes[i];
    return k;
}

}  // extern "C"

#include "chain_server.hpp"
