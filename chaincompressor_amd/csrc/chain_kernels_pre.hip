// Pre-entropy kernels of the Chain compressor for gfx950 (CDNA4).
//
// Stage map to the reference CPU path (cpp/chain/lcps_cpp.cpp):
//   k_bbox / k_bbox_finish   minima                         :2072-2079
//   k_quantize               quantise, split, corrections   :2091-2110, 1494-1523
//   k_sort_*                 L1 grouping by block (stable)  :2114-2138
//   k_link                   successor + chain decision     :2155-2200
//   k_explicit               explicit block records         :2196-2199
//   k_l2_stats / k_l2_split  L2 diff/keep/choice2           :1305-1454, 1456-1478
// All work is 64-wide-wavefront native: ballots for ranking/compaction,
// LDS-staged tiles of 4096 items, single-pass decoupled look-back scans.
#include <type_traits>

#include "chain_common.hpp"

#pragma clang fp contract(off)

namespace chain {

// ------------------------------------------------------------------------
// small device helpers

__device__ inline uint64_t okey(double x) {      // order-preserving key, -0 < +0
    uint64_t b = __double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ inline uint32_t lane_id() { return __lane_id(); }
__device__ inline uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// Cross-lane steps with DPP (no LDS round trip). row_shr:k gives each lane
// the value of the lane k below it in its row of 16 (the identity where that
// lane is outside the row); row_bcast:15 gives rows 1 and 3 the last lane of
// the row before, row_bcast:31 gives rows 2 and 3 lane 31.
#define DPP_ROW_SHR(k) (0x110 + (k))
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143
template <int CTRL, int ROW_MASK>
__device__ inline uint32_t dpp_u32(uint32_t ident, uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)ident, (int)x, CTRL, ROW_MASK, 0xf, false);
}
// inclusive combine over the wavefront in lane order (Hillis-Steele inside
// each row, then the rows): every lane gets op over lanes 0..itself
template <typename Op>
__device__ inline uint32_t wave_incl_dpp(uint32_t x, uint32_t ident, Op op) {
    x = op(x, dpp_u32<DPP_ROW_SHR(1), 0xf>(ident, x));
    x = op(x, dpp_u32<DPP_ROW_SHR(2), 0xf>(ident, x));
    x = op(x, dpp_u32<DPP_ROW_SHR(4), 0xf>(ident, x));
    x = op(x, dpp_u32<DPP_ROW_SHR(8), 0xf>(ident, x));
    x = op(x, dpp_u32<DPP_ROW_BCAST15, 0xa>(ident, x));
    x = op(x, dpp_u32<DPP_ROW_BCAST31, 0xc>(ident, x));
    return x;
}

// inclusive scan of a u32 inside one 64-lane wavefront
__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    return wave_incl_dpp(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}
// minimum / maximum over the wavefront, in every lane
__device__ inline uint32_t wave_min_u32(uint32_t v) {
    const uint32_t r = wave_incl_dpp(v, 0xffffffffu, [](uint32_t a, uint32_t b) { return a < b ? a : b; });
    return (uint32_t)__builtin_amdgcn_readlane((int)r, 63);
}
__device__ inline uint32_t wave_max_u32(uint32_t v) {
    const uint32_t r = wave_incl_dpp(v, 0u, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
    return (uint32_t)__builtin_amdgcn_readlane((int)r, 63);
}

// exclusive scan over a 256-thread workgroup; returns the exclusive prefix,
// *total = sum over the workgroup. s_tmp must hold 4 entries.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) s_tmp[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t k = 0; k < CH_WG / 64; ++k) {
        uint32_t t = s_tmp[k];
        before += k < w ? t : 0u;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// largest f with begin[f] <= x (begin ascending, begin[0] = 0)
__device__ inline uint32_t upper_find(const uint32_t* begin, uint32_t nframes, uint32_t x) {
    uint32_t lo = 0, hi = nframes;   // answer in [lo, hi)
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (begin[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------------------
// single-pass decoupled look-back over tiles (status word: 2-bit flag | 62-bit value)

#define LB_AGG (1ull << 62)
#define LB_INC (2ull << 62)
#define LB_MASK ((1ull << 62) - 1ull)
#define LB_SPIN_LIMIT (1u << 24)

__device__ inline void lb_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t lb_load(uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by all lanes of wave 0. Tiles are numbered by an atomic ticket, so
// every predecessor is already resident or finished and publishes its
// aggregate before it looks back itself: the wait always ends. A bounded
// spin sets *err instead of waiting forever.
__device__ inline uint64_t lookback(uint64_t* status, uint32_t tile, uint64_t aggregate, uint32_t* err) {
    const uint32_t lane = lane_id();
    if (tile == 0) {
        if (lane == 0) lb_store(&status[0], LB_INC | aggregate);
        return 0;
    }
    if (lane == 0) lb_store(&status[tile], LB_AGG | aggregate);
    uint64_t excl = 0;
    int64_t t = (int64_t)tile - 1;
    uint32_t spins = 0;
    while (t >= 0) {
        // each lane inspects one predecessor, nearest first
        const int64_t mine = t - (int64_t)lane;
        uint64_t s = mine >= 0 ? lb_load(&status[mine]) : LB_INC;
        const uint64_t not_ready = __ballot((s >> 62) == 0);
        const uint64_t inc_mask = __ballot((s >> 62) == 2);
        // lanes up to (and including) the first inclusive one, stopping at
        // the first not-ready one
        const uint64_t first_inc = inc_mask ? (inc_mask & (~inc_mask + 1)) : 0;
        const uint64_t first_nr = not_ready ? (not_ready & (~not_ready + 1)) : 0;
        if (first_nr && (!first_inc || first_nr < first_inc)) {
            // sum the ready aggregates before the first not-ready lane, then wait
            const uint64_t take = first_nr - 1;
            uint64_t v = ((take >> lane) & 1ull) ? (s & LB_MASK) : 0ull;
            for (int k = 32; k >= 1; k >>= 1) v += __shfl_xor(v, k, 64);
            excl += v;
            t -= __popcll(take);
            if (++spins > LB_SPIN_LIMIT) { if (lane == 0) atomicOr(err, 1u); break; }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        const uint64_t take = first_inc ? (first_inc | (first_inc - 1)) : ~0ull;
        uint64_t v = ((take >> lane) & 1ull) ? (s & LB_MASK) : 0ull;
        for (int k = 32; k >= 1; k >>= 1) v += __shfl_xor(v, k, 64);
        excl += v;
        if (first_inc) break;
        t -= 64;
    }
    if (lane == 0) lb_store(&status[tile], LB_INC | (excl + aggregate));
    return excl;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------------
// generic single-pass exclusive scan of u32 -> u32 (used for radix offsets
// and per-tile counts)

// items_dev: the length is 256 x the tiles of a count of items that lives on
// the device (a radix histogram's digit-major counts); tiles past it exit
__global__ __launch_bounds__(CH_WG) void k_scan_u32(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                    uint64_t n, uint64_t* status, uint32_t* ticket, uint32_t* err,
                                                    uint64_t* total_out, const uint32_t* __restrict__ items_dev) {
    __shared__ uint32_t s_tile, s_tmp[4];
    __shared__ uint64_t s_excl;
    if (items_dev) n = 256ull * (((uint64_t)*items_dev + CH_TILE - 1) / CH_TILE);
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if ((uint64_t)tile * CH_TILE >= n) return;
    const uint64_t base = (uint64_t)tile * CH_TILE + (uint64_t)threadIdx.x * CH_ITEMS;
    // a thread's 16 items in four 16-byte loads and stores when they are all
    // in range and both arrays are 16-byte aligned (the arena's are)
    const bool vec = base + CH_ITEMS <= n && ((((uintptr_t)in) | ((uintptr_t)out)) & 15u) == 0;
    uint32_t v[CH_ITEMS];
    uint32_t sum = 0;
    if (vec) {
        const uint4* src = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int q = 0; q < CH_ITEMS / 4; ++q) {
            const uint4 u = src[q];
            v[4 * q] = u.x; v[4 * q + 1] = u.y; v[4 * q + 2] = u.z; v[4 * q + 3] = u.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k) v[k] = base + k < n ? in[base + k] : 0u;
    }
#pragma unroll
    for (int k = 0; k < CH_ITEMS; ++k) sum += v[k];
    uint32_t tot;
    uint32_t pre = block_excl_scan(sum, s_tmp, &tot);
    if (threadIdx.x < 64) {
        uint64_t e = lookback(status, tile, tot, err);
        if (threadIdx.x == 0) s_excl = e;
    }
    __syncthreads();
    uint64_t run = s_excl + pre;
    if (vec) {
        uint4* dst = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int q = 0; q < CH_ITEMS / 4; ++q) {
            uint4 u;
            u.x = (uint32_t)run; run += v[4 * q];
            u.y = (uint32_t)run; run += v[4 * q + 1];
            u.z = (uint32_t)run; run += v[4 * q + 2];
            u.w = (uint32_t)run; run += v[4 * q + 3];
            dst[q] = u;
        }
    } else {
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k) {
            if (base + k < n) out[base + k] = (uint32_t)run;
            run += v[k];
        }
    }
    if (total_out && base < n && base + CH_ITEMS >= n) *total_out = run;
}

// ------------------------------------------------------------------------
// K1: per-tile minima/maxima of every dimension; a second kernel folds the
// tiles of each frame. The first +0 and -0 index are kept so the host can
// reproduce std::min's "first of equal values" choice for a zero minimum.

template <typename T, int D, bool SOA>
__device__ inline double load_coord(const T* pts, const FrameDev& F, uint64_t i, int d) {
    if (SOA) return (double)pts[F.pt_off * D + (uint64_t)d * F.n + i];
    return (double)pts[(F.pt_off + i) * D + d];
}

template <typename T, int D, bool SOA>
__global__ __launch_bounds__(CH_WG) void k_bbox(const T* __restrict__ pts, const FrameDev* __restrict__ fr,
                                                const uint32_t* __restrict__ tile_begin, uint32_t nframes,
                                                uint64_t* __restrict__ part, uint32_t* __restrict__ bad) {
    const uint32_t tile = blockIdx.x;
    const uint32_t f = upper_find(tile_begin, nframes, tile);
    const FrameDev& F = fr[f];
    const uint64_t start = (uint64_t)(tile - F.tile0) * CH_TILE;
    const uint64_t end = start + CH_TILE < F.n ? start + CH_TILE : F.n;
    // Order-preserving keys of the input width (float32 -> double is exact
    // and keeps the order, -0 < +0 included), zero positions relative to the
    // tile; both widened once per thread before the reduction.
    using BK = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
    using BU = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
    constexpr BK BK_MAX = ~BK(0);
    BK kmn[D], kmx[D];
    uint32_t rzp[D], rzn[D];
#pragma unroll
    for (int d = 0; d < D; ++d) { kmn[d] = BK_MAX; kmx[d] = 0; rzp[d] = ~0u; rzn[d] = ~0u; }
    bool nonfinite = false;
    // every coordinate of the tile is loaded before the first is used
    T xs[CH_ITEMS][D];
#pragma unroll
    for (int j = 0; j < CH_ITEMS; ++j) {
        const uint64_t i = start + (uint64_t)j * CH_WG + threadIdx.x;
        // unconditional loads at a valid point (the tile's first)
        const uint64_t ic = i < end ? i : start;
        const T* p = SOA ? pts + F.pt_off * D + ic : pts + (F.pt_off + ic) * D;
#pragma unroll
        for (int d = 0; d < D; ++d) xs[j][d] = SOA ? p[(uint64_t)d * F.n] : p[d];
    }
#pragma unroll
    for (int j = 0; j < CH_ITEMS; ++j) {
        const uint32_t rel = (uint32_t)j * CH_WG + threadIdx.x;   // point start + rel
        if (start + rel >= end) continue;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T x = xs[j][d];
            nonfinite |= !isfinite(x);
            BU b;
            __builtin_memcpy(&b, &x, sizeof(b));
            const BK k = (b >> (8 * sizeof(BU) - 1)) ? BK(~b) : BK(b | (BU(1) << (8 * sizeof(BU) - 1)));
            kmn[d] = k < kmn[d] ? k : kmn[d];
            kmx[d] = k > kmx[d] ? k : kmx[d];
            if (x == T(0)) {
                if (signbit(x)) rzn[d] = rel < rzn[d] ? rel : rzn[d];
                else rzp[d] = rel < rzp[d] ? rel : rzp[d];
            }
        }
    }
    // widen: key of the input width -> the double's key
    auto widen = [](BK k) -> uint64_t {
        const BU b = (k >> (8 * sizeof(BK) - 1)) ? BU(k & ~(BK(1) << (8 * sizeof(BK) - 1))) : BU(~k);
        T x;
        __builtin_memcpy(&x, &b, sizeof(x));
        return okey((double)x);
    };
    __shared__ uint64_t s_red[4][CH_MAXD][4];
    const uint32_t w = threadIdx.x >> 6;
    if constexpr (sizeof(BK) == 4) {
        // 32-bit keys and tile offsets reduce across the wavefront with DPP,
        // then only the results are widened
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const uint32_t a = wave_min_u32(kmn[d]), b = wave_max_u32(kmx[d]);
            const uint32_t c = wave_min_u32(rzp[d]), e = wave_min_u32(rzn[d]);
            if (lane_id() == 0) {
                s_red[w][d][0] = a == BK_MAX ? ~0ull : widen(a);
                s_red[w][d][1] = b == 0 ? 0ull : widen(b);
                s_red[w][d][2] = c == ~0u ? ~0ull : start + c;
                s_red[w][d][3] = e == ~0u ? ~0ull : start + e;
            }
        }
    } else {
        uint64_t mn[D], mx[D], zp[D], zn[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            mn[d] = kmn[d] == BK_MAX ? ~0ull : widen(kmn[d]);
            mx[d] = kmx[d] == 0 ? 0ull : widen(kmx[d]);
            zp[d] = rzp[d] == ~0u ? ~0ull : start + rzp[d];
            zn[d] = rzn[d] == ~0u ? ~0ull : start + rzn[d];
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            uint64_t a = mn[d], b = mx[d], c = zp[d], e = zn[d];
            for (int k = 32; k >= 1; k >>= 1) {
                uint64_t ta = __shfl_xor(a, k, 64), tb = __shfl_xor(b, k, 64);
                uint64_t tc = __shfl_xor(c, k, 64), te = __shfl_xor(e, k, 64);
                a = ta < a ? ta : a; b = tb > b ? tb : b; c = tc < c ? tc : c; e = te < e ? te : e;
            }
            if (lane_id() == 0) { s_red[w][d][0] = a; s_red[w][d][1] = b; s_red[w][d][2] = c; s_red[w][d][3] = e; }
        }
    }
    if (nonfinite) atomicOr(&bad[f], 1u);
    __syncthreads();
    if (threadIdx.x < (uint32_t)D) {
        const int d = threadIdx.x;
        uint64_t a = ~0ull, b = 0, c = ~0ull, e = ~0ull;
        for (int k = 0; k < 4; ++k) {
            a = s_red[k][d][0] < a ? s_red[k][d][0] : a;
            b = s_red[k][d][1] > b ? s_red[k][d][1] : b;
            c = s_red[k][d][2] < c ? s_red[k][d][2] : c;
            e = s_red[k][d][3] < e ? s_red[k][d][3] : e;
        }
        uint64_t* o = part + ((uint64_t)tile * D + d) * 4;
        o[0] = a; o[1] = b; o[2] = c; o[3] = e;
    }
}

template <int D>
__global__ __launch_bounds__(CH_WG) void k_bbox_finish(const uint64_t* __restrict__ part,
                                                       const uint32_t* __restrict__ tile_begin, uint32_t stride,
                                                       uint32_t split, uint64_t* __restrict__ frame_box) {
    // frame f's partials are [tile_begin[f], tile_begin[f + 1]), or
    // [f * stride, (f + 1) * stride) without a table; workgroup b folds part
    // b % split of frame b / split into output slot b
    const uint32_t f = blockIdx.x / split, sp = blockIdx.x % split;
    const uint32_t f0 = tile_begin ? tile_begin[f] : f * stride, f1 = tile_begin ? tile_begin[f + 1] : (f + 1) * stride;
    const uint32_t per = (f1 - f0 + split - 1) / split;
    const uint32_t t0 = f0 + sp * per < f1 ? f0 + sp * per : f1;
    const uint32_t t1 = t0 + per < f1 ? t0 + per : f1;
    __shared__ uint64_t s_red[4][CH_MAXD][4];
    const uint32_t w = threadIdx.x >> 6;
    // four tiles per thread and round, all their partials loaded first; a
    // tile past the end reads the round's first tile again, which min, max
    // and first-zero leave unchanged
    uint64_t ma[D], mb[D], mc[D], me[D];
#pragma unroll
    for (int d = 0; d < D; ++d) { ma[d] = ~0ull; mb[d] = 0; mc[d] = ~0ull; me[d] = ~0ull; }
    for (uint32_t t = t0 + threadIdx.x; t < t1; t += 4 * CH_WG) {
        uint64_t v[4][D][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t tt = t + u * CH_WG < t1 ? t + u * CH_WG : t;
#pragma unroll
            for (int d = 0; d < D; ++d)
#pragma unroll
                for (int q = 0; q < 4; ++q) v[u][d][q] = part[((uint64_t)tt * D + d) * 4 + q];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int d = 0; d < D; ++d) {
                ma[d] = v[u][d][0] < ma[d] ? v[u][d][0] : ma[d];
                mb[d] = v[u][d][1] > mb[d] ? v[u][d][1] : mb[d];
                mc[d] = v[u][d][2] < mc[d] ? v[u][d][2] : mc[d];
                me[d] = v[u][d][3] < me[d] ? v[u][d][3] : me[d];
            }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        uint64_t a = ma[d], b = mb[d], c = mc[d], e = me[d];
        for (int k = 32; k >= 1; k >>= 1) {
            uint64_t ta = __shfl_xor(a, k, 64), tb = __shfl_xor(b, k, 64);
            uint64_t tc = __shfl_xor(c, k, 64), te = __shfl_xor(e, k, 64);
            a = ta < a ? ta : a; b = tb > b ? tb : b; c = tc < c ? tc : c; e = te < e ? te : e;
        }
        if (lane_id() == 0) { s_red[w][d][0] = a; s_red[w][d][1] = b; s_red[w][d][2] = c; s_red[w][d][3] = e; }
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)D) {
        const int d = threadIdx.x;
        uint64_t a = ~0ull, b = 0, c = ~0ull, e = ~0ull;
        for (int k = 0; k < 4; ++k) {
            a = s_red[k][d][0] < a ? s_red[k][d][0] : a;
            b = s_red[k][d][1] > b ? s_red[k][d][1] : b;
            c = s_red[k][d][2] < c ? s_red[k][d][2] : c;
            e = s_red[k][d][3] < e ? s_red[k][d][3] : e;
        }
        uint64_t* o = frame_box + ((uint64_t)blockIdx.x * D + d) * 4;
        o[0] = a; o[1] = b; o[2] = c; o[3] = e;
    }
}

// ------------------------------------------------------------------------
// K2: FP64 quantisation, block split, concatenated block key and exact
// corrections (cpp/chain/lcps_cpp.cpp:2091-2110 with :1494-1523).

__device__ inline bool within_eb(double restored, double original, double eb) {
    if (!isfinite(restored)) return false;
    const double difference = restored - original;
    const double magnitude = fabs(difference);
    if (magnitude < eb) return true;
    if (!(magnitude == eb)) return false;
    const double virtual_original = difference - restored;
    const double remainder = (restored - (difference - virtual_original)) + (-original - virtual_original);
    return difference >= 0.0 ? remainder <= 0.0 : remainder >= 0.0;
}

#define QZ_GROUP 8   // points per thread whose loads are in flight together (16: 120 VGPRs, slower)

// Quantise the tile [start, end) of frame F: rel offsets and block keys to
// global memory (and, with S_KEY, every key also to s_key[point - start]);
// with s_h set, the low-byte histogram of key word 0 for the (at most two)
// sort tiles the points fall in. Returns the thread's correction flags (bit
// j*D + d for item j, dimension d).
template <typename T, int D, bool SOA, typename KeyT, typename RelT, bool S_KEY>
__device__ inline uint64_t quantize_items(const T* __restrict__ pts, const FrameDev& F, uint64_t start,
                                          uint64_t end, KeyT* __restrict__ key_out, RelT* __restrict__ rel_out,
                                          uint32_t nwords, uint64_t total_pts, uint32_t (*s_h)[256], uint64_t st0,
                                          KeyT* s_key) {
    const double denom = F.denom, eb = F.eb;
    const uint64_t p = F.p, magic = F.p_magic;
    const uint32_t l = F.p_l;
    double mins[D];
    uint32_t shift[D], word[D];
#pragma unroll
    for (int d = 0; d < D; ++d) { mins[d] = F.mins[d]; shift[d] = F.key_shift[d]; word[d] = F.key_word[d]; }
    // frame-uniform fast paths: q < 2^32 converts with one instruction, and
    // a power-of-two p splits q with a shift and a mask
    const bool q32 = (F.flags & FL_Q32) != 0, p2 = (F.flags & FL_P2) != 0;
    uint64_t cmask = 0;   // correction flags, bit j*D + d
#pragma unroll 1
    for (int j0 = 0; j0 < CH_ITEMS; j0 += QZ_GROUP) {
        // every coordinate of the group is loaded before the first is used;
        // points past the end load the tile's first point
        T xs[QZ_GROUP][D];
#pragma unroll
        for (int jj = 0; jj < QZ_GROUP; ++jj) {
            const uint64_t i = start + (uint64_t)(j0 + jj) * CH_WG + threadIdx.x;
            const uint64_t ic = i < end ? i : start;
#pragma unroll
            for (int d = 0; d < D; ++d)
                xs[jj][d] = pts[SOA ? F.pt_off * D + (uint64_t)d * F.n + ic : (F.pt_off + ic) * D + d];
        }
#pragma unroll
        for (int jj = 0; jj < QZ_GROUP; ++jj) {
        const int j = j0 + jj;
        const uint64_t i = start + (uint64_t)j * CH_WG + threadIdx.x;
        if (i < end) {
            KeyT key = 0;
            KeyT kw[D];
#pragma unroll
            for (int d = 0; d < D; ++d) kw[d] = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const double x = (double)xs[jj][d];
                // 0 <= qv < 2^63 is an integer, so (double)(uint64_t)qv == qv
                const double qv = floor((x - mins[d]) / denom);
                const double restored = fma(qv + 0.5, denom, mins[d]);
                if (!within_eb(restored, x, eb)) cmask |= 1ull << (j * D + d);
                uint64_t b, r;
                if (q32) {
                    const uint32_t q = (uint32_t)qv;
                    const uint32_t bq = p2 ? q >> l : (uint32_t)div_p(q, p, magic, l);
                    b = bq;
                    r = p2 ? (q & (uint32_t)(p - 1)) : q - bq * (uint32_t)p;
                } else {
                    const uint64_t q = (uint64_t)qv;
                    b = p2 ? q >> l : div_p(q, p, magic, l);
                    r = q - b * p;
                }
                rel_out[(F.pt_off + i) * D + d] = (RelT)r;
                if (nwords == 1) key |= (KeyT)b << shift[d];
                else kw[word[d]] |= (KeyT)b << shift[d];
            }
            if (nwords == 1) {
                key_out[F.pt_off + i] = key;
                if (S_KEY) s_key[i - start] = key;
            } else {
#pragma unroll
                for (int w = 0; w < D; ++w)
                    if ((uint32_t)w < nwords) key_out[(uint64_t)w * total_pts + F.pt_off + i] = kw[w];
            }
            if (s_h) {
                const uint32_t digit = (uint32_t)(nwords == 1 ? key : kw[0]) & 0xffu;
                atomicAdd(&s_h[(uint32_t)((F.pt_off + i) / CH_TILE - st0)][digit], 1u);
            }
        }
        }
    }
    return cmask;
}

// Exact corrections of a tile, in element order (j, thread, d), into the
// tile's slot of corr_idx / corr_val (cap entries per tile); rare, so the
// ordered write path runs only when the tile has any. A tile with more than
// cap corrections keeps its first cap, records cap as its count and raises
// bit 0 of *cap_flag: the call is then run again with cap = CH_TILE * D.
// Called by the whole workgroup.
template <typename T, int D, bool SOA>
__device__ inline void tile_corrections(const T* __restrict__ pts, const FrameDev& F, uint64_t start, uint32_t tile,
                                        uint64_t cmask, uint64_t* __restrict__ corr_idx,
                                        double* __restrict__ corr_val, uint32_t* __restrict__ tile_ncorr,
                                        uint32_t cap, uint32_t* __restrict__ cap_flag, uint32_t* s_tmp) {
    uint32_t run = 0;
    const uint64_t cbase = (uint64_t)tile * cap;
    if (__syncthreads_or(cmask != 0)) {
        for (int j = 0; j < CH_ITEMS; ++j) {
            const uint32_t bits = (uint32_t)(cmask >> (j * D)) & ((1u << D) - 1u);
            uint32_t tot;
            uint32_t pos = run + block_excl_scan(__popc(bits), s_tmp, &tot);
            const uint64_t i = start + (uint64_t)j * CH_WG + threadIdx.x;
#pragma unroll
            for (int d = 0; d < D; ++d)
                if (bits & (1u << d)) {
                    if (pos < cap) {
                        corr_idx[cbase + pos] = i * D + d;
                        corr_val[cbase + pos] = load_coord<T, D, SOA>(pts, F, i, d);
                    }
                    ++pos;
                }
            run += tot;
        }
    }
    if (threadIdx.x == 0) {
        if (run > cap) atomicOr(cap_flag, 1u);
        tile_ncorr[tile] = run < cap ? run : cap;
    }
}

template <typename T, int D, bool SOA, typename KeyT, typename RelT>
__global__ __launch_bounds__(CH_WG) void k_quantize(const T* __restrict__ pts, const FrameDev* __restrict__ fr,
                                                    const uint32_t* __restrict__ tile_begin, uint32_t nframes,
                                                    KeyT* __restrict__ key_out, RelT* __restrict__ rel_out,
                                                    uint64_t* __restrict__ corr_idx, double* __restrict__ corr_val,
                                                    uint32_t* __restrict__ tile_ncorr, uint32_t corr_cap,
                                                    uint32_t* __restrict__ cap_flag, uint32_t nwords,
                                                    uint64_t total_pts, uint32_t* __restrict__ counts,
                                                    uint32_t sort_ntiles) {
    const uint32_t tile = blockIdx.x;
    const uint32_t f = upper_find(tile_begin, nframes, tile);
    const FrameDev& F = fr[f];
    const uint64_t start = (uint64_t)(tile - F.tile0) * CH_TILE;
    const uint64_t end = start + CH_TILE < F.n ? start + CH_TILE : F.n;
    // With counts set, the first radix pass's digit histogram (low byte of
    // key word 0) is gathered here: the tile's points fall in at most two
    // sort tiles of CH_TILE global positions, counted in LDS and added to
    // counts[digit * sort_ntiles + sort tile] (cleared before the launch).
    __shared__ uint32_t s_h[2][256];
    const uint64_t st0 = (F.pt_off + start) / CH_TILE;
    if (counts) {
        s_h[0][threadIdx.x] = 0;
        s_h[1][threadIdx.x] = 0;
        __syncthreads();
    }
    __shared__ uint32_t s_tmp[4];
    const uint64_t cmask = quantize_items<T, D, SOA, KeyT, RelT, false>(
        pts, F, start, end, key_out, rel_out, nwords, total_pts, counts ? s_h : nullptr, st0, nullptr);
    if (counts) {
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const uint32_t v = s_h[s][threadIdx.x];
            if (v) atomicAdd(&counts[(uint64_t)threadIdx.x * sort_ntiles + st0 + s], v);
        }
    }
    tile_corrections<T, D, SOA>(pts, F, start, tile, cmask, corr_idx, corr_val, tile_ncorr, corr_cap, cap_flag, s_tmp);
}

// gather per-tile correction lists into one ordered list
template <int D>
__global__ __launch_bounds__(CH_WG) void k_corr_gather(const uint64_t* __restrict__ corr_idx,
                                                       const double* __restrict__ corr_val,
                                                       const uint32_t* __restrict__ tile_ncorr,
                                                       const uint32_t* __restrict__ tile_corr_off, uint32_t cap,
                                                       uint64_t* __restrict__ out_idx, double* __restrict__ out_val) {
    const uint32_t tile = blockIdx.x;
    const uint32_t c = tile_ncorr[tile];
    const uint64_t src = (uint64_t)tile * cap;
    const uint64_t dst = tile_corr_off[tile];
    for (uint32_t k = threadIdx.x; k < c; k += CH_WG) {
        out_idx[dst + k] = corr_idx[src + k];
        out_val[dst + k] = corr_val[src + k];
    }
}

// ------------------------------------------------------------------------
// K3: stable LSD radix sort of (block key, point index), 8-bit digits.
// Pass = tile histogram -> digit-major exclusive scan -> scatter. Wavefront w
// of a tile owns its positions [w * 1024, (w + 1) * 1024), item r of lane l
// at w * 1024 + r * 64 + l, so ranking the wavefronts one after another, each
// in item order, is the tile's order (the sort stays stable). The scatter
// loads every item first, ranks them within the wavefront with ballots (8
// ballots give each lane the mask of lanes holding its digit) and a per-
// wavefront digit counter in LDS, adds the counts of the earlier wavefronts,
// stages the tile in LDS in digit order and writes contiguous runs per digit.

#define SORT_WAVE_ITEMS (CH_TILE / CH_WG)   // items per lane
template <typename KeyT>
__global__ __launch_bounds__(CH_WG) void k_sort_hist(const KeyT* __restrict__ keys, uint64_t T, uint32_t shift,
                                                     uint32_t ntiles, uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ n_dev) {
    // n_dev: the item count lives on the device. The counts are digit-major
    // with a stride of the tiles the count needs (not the launch's), so the
    // scan and the scatter touch only those; tiles past it exit.
    if (n_dev) {
        T = *n_dev;
        ntiles = (uint32_t)((T + CH_TILE - 1) / CH_TILE);
        if (blockIdx.x >= ntiles) return;
    }
    // one histogram per wavefront: atomics of different wavefronts never
    // meet on a word; the four are summed at the end
    __shared__ uint32_t s_h[CH_WG / 64][256];
#pragma unroll
    for (int w = 0; w < CH_WG / 64; ++w) s_h[w][threadIdx.x] = 0;
    const uint64_t base = (uint64_t)blockIdx.x * CH_TILE;
    KeyT kx[CH_ITEMS];
#pragma unroll
    for (int j = 0; j < CH_ITEMS; ++j) {   // unconditional loads (the tile's first key past the end)
        const uint64_t i = base + (uint64_t)j * CH_WG + threadIdx.x;
        kx[j] = keys[i < T ? i : (base < T ? base : 0)];
    }
    __syncthreads();
    uint32_t* h = s_h[threadIdx.x >> 6];
#pragma unroll
    for (int j = 0; j < CH_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * CH_WG + threadIdx.x;
        if (i < T) atomicAdd(&h[(uint32_t)(kx[j] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < CH_WG / 64; ++w) c += s_h[w][threadIdx.x];
    counts[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = c;
}

template <typename KeyT, typename ValT, bool FIRST>
__global__ __launch_bounds__(CH_WG) void k_sort_scatter(const KeyT* __restrict__ kin, const ValT* __restrict__ vin,
                                                        KeyT* __restrict__ kout, ValT* __restrict__ vout,
                                                        uint64_t T, uint32_t shift, uint32_t ntiles,
                                                        const uint32_t* __restrict__ counts,
                                                        const uint32_t* __restrict__ offsets,
                                                        const uint32_t* __restrict__ n_dev) {
    if (n_dev) {   // the item count lives on the device; the counts' stride follows it
        T = *n_dev;
        ntiles = (uint32_t)((T + CH_TILE - 1) / CH_TILE);
    }
    if ((uint64_t)blockIdx.x * CH_TILE >= T) return;
    __shared__ uint32_t s_lstart[256], s_goff[256], s_tmp[4];
    // per-wavefront digit counts, then run starts: at most CH_TILE, so 16 bits
    // (2 KB less LDS: three workgroups per CU with 32-bit keys and 64-bit values)
    __shared__ uint16_t s_wcnt[4][256];
    static_assert(CH_TILE <= 65535, "run starts fit 16 bits");
    __shared__ KeyT s_key[CH_TILE];
    __shared__ ValT s_val[CH_TILE];
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63u;
    const uint64_t base = (uint64_t)blockIdx.x * CH_TILE;
    const uint32_t cnt_tile = (uint32_t)(T - base < CH_TILE ? T - base : CH_TILE);
    KeyT key[SORT_WAVE_ITEMS];
    ValT val[SORT_WAVE_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_WAVE_ITEMS; ++r) {
        const uint32_t p = w * (CH_TILE / 4) + r * 64 + lane;
        const uint64_t pc = base + (p < cnt_tile ? p : 0u);   // unconditional (the tile's first key)
        key[r] = kin[pc];
        val[r] = FIRST ? (ValT)(base + p) : vin[pc];
    }
    {
        const uint32_t c = counts[(uint64_t)tid * ntiles + blockIdx.x];
        s_goff[tid] = offsets[(uint64_t)tid * ntiles + blockIdx.x];
        uint32_t tot;
        s_lstart[tid] = block_excl_scan(c, s_tmp, &tot);
        for (int k = 0; k < 4; ++k) s_wcnt[k][tid] = 0;
    }
    __syncthreads();
    // rank of every item among its wavefront's items of the same digit
    uint32_t wr[SORT_WAVE_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_WAVE_ITEMS; ++r) {
        const uint32_t p = w * (CH_TILE / 4) + r * 64 + lane;
        const bool valid = p < cnt_tile;
        const uint32_t digit = (uint32_t)(key[r] >> shift) & 0xffu;
        // peer mask in 32-bit halves: a lane whose digit bit b is set keeps
        // the ballot, else its complement (t = 0 / ~0)
        const uint64_t v0 = __builtin_amdgcn_ballot_w64(valid);
        uint32_t plo = (uint32_t)v0, phi = (uint32_t)(v0 >> 32);
        const uint32_t nd = ~digit;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint32_t t = (uint32_t)__builtin_amdgcn_sbfe((int32_t)nd, b, 1);
            const uint64_t bb = __builtin_amdgcn_ballot_w64(t == 0u);
            plo &= (uint32_t)bb ^ t;
            phi &= (uint32_t)(bb >> 32) ^ t;
        }
        const uint64_t peers = ((uint64_t)phi << 32) | plo;
        const uint32_t prank = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
        uint32_t before = 0;
        if (valid) before = s_wcnt[w][digit];
        wr[r] = before + prank;
        // lanes of one digit read the count before its leader updates it
        __builtin_amdgcn_wave_barrier();
        if (valid && prank == 0) s_wcnt[w][digit] = (uint16_t)(before + (uint32_t)__popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        // digit tid: start of every wavefront's run in the tile
        const uint32_t a0 = s_wcnt[0][tid], a1 = s_wcnt[1][tid], a2 = s_wcnt[2][tid];
        const uint32_t st = s_lstart[tid];
        s_wcnt[0][tid] = (uint16_t)st;
        s_wcnt[1][tid] = (uint16_t)(st + a0);
        s_wcnt[2][tid] = (uint16_t)(st + a0 + a1);
        s_wcnt[3][tid] = (uint16_t)(st + a0 + a1 + a2);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_WAVE_ITEMS; ++r) {
        const uint32_t p = w * (CH_TILE / 4) + r * 64 + lane;
        if (p < cnt_tile) {
            const uint32_t digit = (uint32_t)(key[r] >> shift) & 0xffu;
            const uint32_t lp = s_wcnt[w][digit] + wr[r];
            s_key[lp] = key[r];
            s_val[lp] = val[r];
        }
    }
    __syncthreads();
    for (uint32_t p = tid; p < cnt_tile; p += CH_WG) {
        const KeyT k = s_key[p];
        const uint32_t d = (uint32_t)(k >> shift) & 0xffu;
        const uint32_t dst = s_goff[d] + (p - s_lstart[d]);
        kout[dst] = k;
        vout[dst] = s_val[p];
    }
}

// ------------------------------------------------------------------------
// The chain decision (cpp/chain/lcps_cpp.cpp:2155-2200): a point whose
// successor lies `dist` indices later stores the distance instead of its
// block when bitlen(dist) + bitlen(bitlen(dist)) is below the cheaper explicit
// cost, min(D lb + bitlen(lb), D bitlen(max block id)), minus the threshold,
// with lb the bit length of the point's largest block coordinate. `key` is
// the concatenated block ids of a single-word key.
template <int D>
__device__ inline bool chain_relative_mb(const FrameDev& F, uint64_t mb, uint64_t dist) {
    const int64_t local_bits = bitlen64(mb);
    const int64_t tag_bits = bitlen64((uint64_t)local_bits);
    const int64_t a = local_bits * D + tag_bits;
    const int64_t c = (int64_t)D * (int64_t)F.max_bid_bits;
    const int64_t smaller = a < c ? a : c;
    const int64_t bl = bitlen64(dist);
    const int64_t lhs = bl + bitlen64((uint64_t)bl);
    return lhs < smaller - F.threshold;
}
// mb: the point's largest block coordinate
template <int D>
__device__ inline bool chain_relative(const FrameDev& F, uint64_t key, uint64_t dist) {
    uint64_t mb = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const uint32_t wd = F.key_width[d];
        const uint64_t mask = wd >= 64 ? ~0ull : ((1ull << wd) - 1ull);
        const uint64_t b = (key >> F.key_shift[d]) & mask;
        mb = b > mb ? b : mb;
    }
    return chain_relative_mb<D>(F, mb, dist);
}

// The same decision as a distance limit: for a point whose largest block
// coordinate has bit length lb, the distance is stored iff dist < limit(lb).
// The left side bl + bitlen(bl), bl = bitlen(dist), never decreases with dist,
// so the distances that qualify are exactly those below 2^bl* for the largest
// qualifying bl* (none: limit 1, i.e. never, since dist >= 1).
// Only bit lengths up to max_bl are examined: the limit is exact for every
// dist < 2^max_bl (in-tile distances are below 4096 = 2^12).
template <int D>
__device__ inline uint64_t chain_limit(const FrameDev& F, int lb, int max_bl) {
    const int64_t a = (int64_t)lb * D + bitlen64((uint64_t)lb);
    const int64_t c = (int64_t)D * (int64_t)F.max_bid_bits;
    const int64_t rhs = (a < c ? a : c) - F.threshold;
    int best = 0;
    for (int bl = 1; bl <= max_bl; ++bl)
        if (bl + bitlen64((uint64_t)bl) < rhs) best = bl;
    return best ? (1ull << best) : 1ull;   // dist < 2^best, or nothing
}

// ------------------------------------------------------------------------
// K4: successor of every point inside its block and the chain decision
// (cpp/chain/lcps_cpp.cpp:2155-2200), computed in sorted order and scattered
// to index order.

// Tiles of 4096 sorted positions, all loads of a tile issued first. The
// workgroups of a launch are dispatched to the 8 XCDs in turn (workgroup b on
// XCD b % 8), so workgroup b takes tile (b % 8) * per + b / 8: every XCD walks
// a contiguous run of the sorted order, whose scattered link writes fall in a
// smaller set of index ranges and merge better in that XCD's L2.
#define LINK_XCDS 8u
template <typename KeyT, int D>
__global__ __launch_bounds__(CH_WG) void k_link(const KeyT* __restrict__ skey, const uint32_t* __restrict__ sval,
                                                uint64_t T, const FrameDev* __restrict__ fr,
                                                const uint32_t* __restrict__ frame_pt_begin, uint32_t nframes,
                                                uint32_t* __restrict__ idrel) {
    __shared__ uint32_t s_off[CH_MAXF_LDS + 1];
    const uint32_t* offs = frame_pt_begin;
    if (nframes <= CH_MAXF_LDS) {
        for (uint32_t k = threadIdx.x; k <= nframes; k += CH_WG) s_off[k] = frame_pt_begin[k];
        __syncthreads();
        offs = s_off;
    }
    const uint32_t ntiles = (uint32_t)((T + CH_TILE - 1) / CH_TILE);
    const uint32_t per = gridDim.x / LINK_XCDS;
    const uint32_t tile = (blockIdx.x % LINK_XCDS) * per + blockIdx.x / LINK_XCDS;
    if (tile >= ntiles) return;
    const uint64_t s0 = (uint64_t)tile * CH_TILE;
    uint32_t iv[CH_ITEMS], jv[CH_ITEMS];
    KeyT kv[CH_ITEMS], kn[CH_ITEMS];
#pragma unroll
    for (int q = 0; q < CH_ITEMS; ++q) {
        const uint64_t s = s0 + (uint64_t)q * CH_WG + threadIdx.x;
        // unconditional loads at valid positions (the tile's first one)
        const uint64_t sc = s < T ? s : s0, nc = s + 1 < T ? s + 1 : s0;
        iv[q] = sval[sc];
        kv[q] = skey[sc];
        jv[q] = sval[nc];
        kn[q] = skey[nc];
    }
#pragma unroll
    for (int q = 0; q < CH_ITEMS; ++q) {
        const uint64_t s = s0 + (uint64_t)q * CH_WG + threadIdx.x;
        if (s >= T) continue;
        const uint32_t i = iv[q];
        const KeyT key = kv[q];
        uint32_t out = 0;
        if (s + 1 < T && kn[q] == key) {
            const uint32_t j = jv[q];
            const uint32_t f = upper_find(offs, nframes, i);
            const FrameDev& F = fr[f];
            if ((F.flags & FL_L1) && j < offs[f + 1]) {
                const uint64_t dist = (uint64_t)(j - i);
                if (chain_relative<D>(F, (uint64_t)key, dist)) out = (uint32_t)dist;
            }
        }
        idrel[i] = out;
    }
}

// ------------------------------------------------------------------------
// K4 for single-word keys, tile by tile: the successors that stay inside a
// tile of 4096 points of one frame (the quantise tiles).
//
// next[i] (the smallest j > i of the frame with the same block) is found in
// two steps. Inside the tile, a stable LDS radix sort of (key, position)
// puts equal keys next to each other in index order, so every point but the
// last of its key in the tile has its successor right after it; those links
// and their chain decisions are final here. The last point of each key in
// the tile (its exit) links to the first point of that key in the next tile
// of the frame that holds the key: each tile emits one item (key, first
// point, last point) per distinct key, in tile order (ordered compaction by
// a look-back over ticketed tiles), the items are sorted by key (stable, so
// a key's items stay in tile order) and k_link_items links every item's last
// point to the next item's first point. Exits are written 0 here and
// overwritten there when they chain. For data with any index locality the
// items are a small fraction of the points, and the global sort and its
// scattered link writes shrink with them.
#define QL_ITEMS (CH_TILE / CH_WG)   // positions per thread (16)
#define QL_DIGIT_MAX 10u             // widest digit of a tile sort pass
//
// The sort runs on a tile key: each dimension's block id minus the tile's
// smallest, packed into the bits the tile's range of that dimension needs.
// It is one-to-one on the tile's keys, so it groups exactly like the full key,
// and a tile of a compact cluster needs a single pass of a few bits where the
// full key needs two or more passes of eight. Ranking costs one ballot per key
// bit and item, so the passes rank exactly the tile key's bits, spread evenly
// over as few passes of at most QL_DIGIT_MAX bits as that width allows.
template <typename KeyT, int D>
__global__ __launch_bounds__(CH_WG) void k_tile_link(
    const KeyT* __restrict__ keys, const FrameDev* __restrict__ fr, const uint32_t* __restrict__ tile_begin,
    uint32_t nframes, uint32_t* __restrict__ idrel, uint32_t* ticket, uint64_t* status, uint32_t* err,
    uint32_t ntiles, KeyT* __restrict__ item_key, uint64_t* __restrict__ item_pay,
    uint32_t* __restrict__ item_total) {
    __shared__ KeyT s_key[CH_TILE];   // tile keys
    __shared__ __attribute__((aligned(16))) uint16_t s_ia[CH_TILE];
    __shared__ __attribute__((aligned(16))) uint16_t s_ib[CH_TILE];
    // per digit of a sort pass: the counts of wavefronts 0 | 1 << 16 and
    // 2 | 3 << 16, then each wavefront's first position of the digit; after
    // the sort, the in-tile links by point (< 4096; exits 0)
    __shared__ __attribute__((aligned(16))) uint32_t s_cnt[2][1u << QL_DIGIT_MAX];
    static_assert(sizeof(s_cnt) >= CH_TILE * sizeof(uint16_t), "links alias the digit counters");
    uint16_t* const s_link = reinterpret_cast<uint16_t*>(&s_cnt[0][0]);
    // per point: e such that its successor's distance is stored iff it is
    // below 2^e (the chain decision of its block); by bit length of the
    // largest block coordinate, for the frame
    __shared__ uint8_t s_exp[CH_TILE];
    __shared__ uint8_t s_le[72];
    __shared__ uint32_t s_tmp[4], s_tile;
    __shared__ KeyT s_rng[CH_WG / 64][2 * CH_MAXD];
    __shared__ uint64_t s_excl;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t f = upper_find(tile_begin, nframes, tile);
    const FrameDev& F = fr[f];
    const uint64_t start = (uint64_t)(tile - F.tile0) * CH_TILE;
    const uint64_t end = start + CH_TILE < F.n ? start + CH_TILE : F.n;
    const uint32_t cnt = (uint32_t)(end - start);
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63u;
    uint32_t* const out = idrel + F.pt_off + start;
    const bool l1 = (F.flags & FL_L1) != 0;
    uint32_t nruns = 0;
    uint16_t* src = s_ia;
    uint16_t* dst = s_ib;
    // block layout of the frame's keys, and of the tile keys
    uint32_t ksh[D];
    KeyT kmask[D], bmin[D], cmask[D];
    uint32_t csh[D], cw[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        ksh[d] = F.key_shift[d];
        const uint32_t wd = F.key_width[d];
        kmask[d] = wd >= 8 * sizeof(KeyT) ? ~KeyT(0) : (KeyT)((KeyT(1) << wd) - 1);
    }
    if (l1) {
        if (tid < 65) s_le[tid] = (uint8_t)__builtin_ctzll(chain_limit<D>(F, (int)tid, 12));
        {
            // the tile's keys, coalesced, every load issued first (past the
            // end: the tile's last key again, which leaves the range as it is)
            const KeyT* kp = keys + F.pt_off + start;
            KeyT kv[QL_ITEMS];
#pragma unroll
            for (int k = 0; k < QL_ITEMS; ++k) {
                const uint32_t j = k * CH_WG + tid;
                kv[k] = kp[j < cnt - 1 ? j : cnt - 1];
            }
            // per-dimension range of the tile's block ids
            KeyT mn[D], mx[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                mn[d] = (kv[0] >> ksh[d]) & kmask[d];
                mx[d] = mn[d];
            }
            // (a position past the end loaded the tile's last key again)
#pragma unroll
            for (int k = 1; k < QL_ITEMS; ++k) {
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const KeyT b = (kv[k] >> ksh[d]) & kmask[d];
                    mn[d] = b < mn[d] ? b : mn[d];
                    mx[d] = b > mx[d] ? b : mx[d];
                }
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if constexpr (sizeof(KeyT) == 4) {
                    mn[d] = wave_min_u32(mn[d]);
                    mx[d] = wave_max_u32(mx[d]);
                } else {
                    for (int o = 32; o >= 1; o >>= 1) {
                        const KeyT a = __shfl_xor(mn[d], o, 64), c = __shfl_xor(mx[d], o, 64);
                        mn[d] = a < mn[d] ? a : mn[d];
                        mx[d] = c > mx[d] ? c : mx[d];
                    }
                }
                if (lane == 0) {
                    s_rng[w][2 * d] = mn[d];
                    s_rng[w][2 * d + 1] = mx[d];
                }
            }
            __syncthreads();
            uint32_t acc = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                KeyT a = s_rng[0][2 * d], c = s_rng[0][2 * d + 1];
#pragma unroll
                for (int q = 1; q < CH_WG / 64; ++q) {
                    a = s_rng[q][2 * d] < a ? s_rng[q][2 * d] : a;
                    c = s_rng[q][2 * d + 1] > c ? s_rng[q][2 * d + 1] : c;
                }
                bmin[d] = a;
                cw[d] = (uint32_t)bitlen64((uint64_t)(c - a));
                cmask[d] = cw[d] >= 8 * sizeof(KeyT) ? ~KeyT(0) : (KeyT)((KeyT(1) << cw[d]) - 1);
                csh[d] = acc;
                acc += cw[d];
            }
            // tile key and chain limit of every point
#pragma unroll
            for (int k = 0; k < QL_ITEMS; ++k) {
                const uint32_t j = k * CH_WG + tid;
                KeyT ck = 0, mb = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const KeyT b = (kv[k] >> ksh[d]) & kmask[d];
                    ck |= (b - bmin[d]) << csh[d];
                    mb = b > mb ? b : mb;
                }
                const uint32_t lb = sizeof(KeyT) == 4 ? 32u - (uint32_t)__clz((uint32_t)mb)
                                                      : (uint32_t)bitlen64((uint64_t)mb);
                // positions past the end are never read: no bound test
                s_key[j] = ck;
                s_exp[j] = s_le[lb];
                s_ia[j] = (uint16_t)j;
            }
        }
        // tile keys and the identity order are read by every thread below
        // (also when no pass runs: every key of the tile equal)
        __syncthreads();
        const uint32_t cbits = csh[D - 1] + cw[D - 1];
        // Stable LSD passes. Wavefront w ranks the positions [w * 1024,
        // (w + 1) * 1024) of the current order, item r of lane l at
        // w * 1024 + r * 64 + l, round by round: one ballot per digit bit
        // narrows each lane's mask to the lanes of its digit (four rounds at
        // a time, for independent work between the ballots), and the lowest
        // lane of each group adds the group's size to the wavefront's counter
        // of that digit with one LDS atomic that returns the count before it.
        // A wavefront's LDS operations complete in order, so consecutive
        // rounds' atomics are issued back to back and their results read
        // afterwards.
        const uint32_t npass = (cbits + QL_DIGIT_MAX - 1u) / QL_DIGIT_MAX;
        const uint32_t pbits = npass ? (cbits + npass - 1u) / npass : 1u;
        const uint32_t hsh = 16u * (w & 1u);
        uint32_t* const wcnt = s_cnt[w >> 1];
        for (uint32_t shift = 0; shift < cbits; shift += pbits) {
            const uint32_t nb = cbits - shift < pbits ? cbits - shift : pbits;
            const uint32_t ndig = 1u << nb;
            for (uint32_t k = tid; k < ndig; k += CH_WG) {
                s_cnt[0][k] = 0;
                s_cnt[1][k] = 0;
            }
            __syncthreads();
            // per item: its point (12 bits) | digit << 12; then its rank among
            // the group's lanes | the group's leader lane << 6, and finally its
            // rank among the wavefront's items of the digit
            uint32_t it[QL_ITEMS], rk[QL_ITEMS];
#pragma unroll
            for (int r = 0; r < QL_ITEMS; ++r) {
                const uint32_t p = w * (CH_TILE / 4) + r * 64 + lane;
                it[r] = src[p < cnt ? p : 0u];
            }
#pragma unroll
            for (int r = 0; r < QL_ITEMS; ++r) it[r] |= ((uint32_t)(s_key[it[r]] >> shift) & (ndig - 1u)) << 12;
#pragma unroll
            for (int h = 0; h < QL_ITEMS; h += 4) {
                // peer masks in 32-bit halves: bit b of the digit set, the
                // lane keeps the ballot, else its complement (t = 0 / ~0)
                uint32_t plo[4], phi[4], nd[4];
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const uint32_t p = w * (CH_TILE / 4) + (h + rr) * 64 + lane;
                    const uint64_t v = __builtin_amdgcn_ballot_w64(p < cnt);
                    plo[rr] = (uint32_t)v;
                    phi[rr] = (uint32_t)(v >> 32);
                    nd[rr] = ~(it[h + rr] >> 12);
                }
                for (uint32_t b = 0; b < nb; ++b) {
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) {
                        const uint32_t t = (uint32_t)__builtin_amdgcn_sbfe((int32_t)nd[rr], (int32_t)b, 1);
                        const uint64_t bb = __builtin_amdgcn_ballot_w64(t == 0u);
                        plo[rr] &= (uint32_t)bb ^ t;
                        phi[rr] &= (uint32_t)(bb >> 32) ^ t;
                    }
                }
                uint32_t bf[4];
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int r = h + rr;
                    const uint32_t p = w * (CH_TILE / 4) + r * 64 + lane;
                    const bool valid = p < cnt;
                    const uint32_t prank = __builtin_amdgcn_mbcnt_hi(phi[rr], __builtin_amdgcn_mbcnt_lo(plo[rr], 0u));
                    const uint32_t lead = !valid ? lane : plo[rr] ? (uint32_t)__builtin_ctz(plo[rr])
                                                                   : 32u + (uint32_t)__builtin_ctz(phi[rr]);
                    rk[r] = prank | (lead << 6);
                    bf[rr] = 0;
                    if (valid && prank == 0)
                        bf[rr] = atomicAdd(&wcnt[it[r] >> 12], (uint32_t)(__popc(plo[rr]) + __popc(phi[rr])) << hsh);
                }
                // every lane: its group's count before this round (the
                // leader's half of the counter word) + its rank in the group
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int r = h + rr;
                    rk[r] = ((__shfl(bf[rr], (int)(rk[r] >> 6), 64) >> hsh) & 0xffffu) + (rk[r] & 63u);
                }
            }
            __syncthreads();
            {
                // digits [tid * per, tid * per + per): their first positions in
                // the new order, then every wavefront's (waves 0 and 1, 2 and 3
                // packed in 16-bit halves; positions are below 4096)
                const uint32_t per = ndig > CH_WG ? ndig / CH_WG : 1u;
                const uint32_t d0 = tid * per, d1 = d0 + per < ndig ? d0 + per : ndig;
                uint32_t sum = 0;
                for (uint32_t d = d0; d < d1; ++d) {
                    const uint32_t c01 = s_cnt[0][d], c23 = s_cnt[1][d];
                    sum += (c01 & 0xffffu) + (c01 >> 16) + (c23 & 0xffffu) + (c23 >> 16);
                }
                uint32_t tot;
                uint32_t st = block_excl_scan(sum, s_tmp, &tot);
                for (uint32_t d = d0; d < d1; ++d) {
                    const uint32_t c01 = s_cnt[0][d], c23 = s_cnt[1][d];
                    const uint32_t a0 = c01 & 0xffffu, a1 = c01 >> 16, a2 = c23 & 0xffffu;
                    s_cnt[0][d] = st | ((st + a0) << 16);
                    s_cnt[1][d] = (st + a0 + a1) | ((st + a0 + a1 + a2) << 16);
                    st += a0 + a1 + a2 + (c23 >> 16);
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < QL_ITEMS; ++r) {
                const uint32_t p = w * (CH_TILE / 4) + r * 64 + lane;
                if (p < cnt) dst[((wcnt[it[r] >> 12] >> hsh) & 0xffffu) + rk[r]] = (uint16_t)(it[r] & 0xfffu);
            }
            __syncthreads();
            uint16_t* t = src;
            src = dst;
            dst = t;
        }
        // Sorted positions [tid * 16, tid * 16 + 16) of this thread: links to
        // the next position of the same key (into s_link by point), run
        // starts counted for the items. A full tile needs no position checks.
        const uint32_t s0 = tid * QL_ITEMS;
        uint32_t first_bits = 0;
        auto link = [&](auto full_tag) {
            constexpr bool FULL = decltype(full_tag)::value;
            if (!FULL && s0 >= cnt) return;
            // positions >= cnt of src hold no sorted entry: never used as an index
            uint32_t ix[QL_ITEMS + 1];
            const uint4* v = reinterpret_cast<const uint4*>(src + s0);
            const uint4 v0 = v[0], v1 = v[1];
            const uint32_t wd[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                ix[2 * q] = wd[q] & 0xffffu;
                ix[2 * q + 1] = wd[q] >> 16;
            }
            ix[QL_ITEMS] = src[s0 + QL_ITEMS < cnt ? s0 + QL_ITEMS : 0u];
            if (!FULL) {
#pragma unroll
                for (int k = 0; k <= QL_ITEMS; ++k) ix[k] = s0 + k < cnt ? ix[k] : ix[0];   // a valid point
            }
            KeyT kp = s_key[src[s0 > 0 ? s0 - 1 : 0u]];
            KeyT kc = s_key[ix[0]];
#pragma unroll
            for (int k = 0; k < QL_ITEMS; ++k) {
                const uint32_t sp = s0 + k;
                const KeyT kn = s_key[ix[k + 1]];
                const bool has_next = sp + 1 < cnt && kn == kc;
                const bool first = (FULL || sp < cnt) && (sp == 0 || kp != kc);
                first_bits |= (first ? 1u : 0u) << k;
                const uint32_t dist = ix[k + 1] - ix[k];   // > 0 when has_next: the sort is stable
                const bool rel = has_next && (dist >> s_exp[ix[k]]) == 0u;
                if (FULL || sp < cnt) s_link[ix[k]] = (uint16_t)(rel ? dist : 0u);
                kp = kc;
                kc = kn;
            }
        };
        if (cnt == CH_TILE) link(std::true_type{});
        else link(std::false_type{});
        const uint32_t myruns = (uint32_t)__popc(first_bits);
        const uint32_t run0 = block_excl_scan(myruns, s_tmp, &nruns);
        // run r starts at dst[r] (the other buffer is free now)
        uint32_t rr = run0;
#pragma unroll
        for (int k = 0; k < QL_ITEMS; ++k)
            if ((first_bits >> k) & 1u) dst[rr++] = (uint16_t)(s0 + k);
        // items of this tile, placed after every earlier tile's (tile order)
        if (tid < 64) {
            const uint64_t e = lookback(status, tile, nruns, err);
            if (tid == 0) {
                s_excl = e;
                if (tile == ntiles - 1) *item_total = (uint32_t)(e + nruns);
            }
        }
        __syncthreads();
        {
            const uint64_t base = s_excl;
            for (uint32_t r = tid; r < nruns; r += CH_WG) {
                const uint32_t a = dst[r], b = (r + 1 < nruns ? dst[r + 1] : cnt) - 1u;
                const uint32_t en = src[a], ex = src[b];
                const KeyT ck = s_key[en];
                KeyT key = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) key |= (bmin[d] + ((ck >> csh[d]) & cmask[d])) << ksh[d];
                item_key[base + r] = key;
                item_pay[base + r] = ((uint64_t)(F.pt_off + start + ex) << 32) | (uint32_t)(F.pt_off + start + en);
            }
        }
        // the links in index order, written out coalesced (exits 0: they are
        // linked across tiles by k_link_items)
#pragma unroll
        for (int k = 0; k < QL_ITEMS; ++k) {
            const uint32_t j = k * CH_WG + tid;
            if (j < cnt) out[j] = s_link[j];
        }
    } else {
        // no L1 in this frame: every point is explicit, and the tile has no items
#pragma unroll
        for (int k = 0; k < QL_ITEMS; ++k) {
            const uint32_t j = k * CH_WG + tid;
            if (j < cnt) out[j] = 0;
        }
        if (tid < 64) {
            const uint64_t e = lookback(status, tile, 0, err);
            if (tid == 0 && tile == ntiles - 1) *item_total = (uint32_t)e;
        }
    }
}

// The cross-tile links: items sorted by key (stable), each carrying its last
// and first point (exit << 32 | entry), so item s + 1 of the same key is the
// next tile holding it and the last point of item s links to the first point
// of item s + 1 when both lie in one frame. Every read is coalesced. Grid
// sized for the largest item count; positions past the actual count exit.
template <typename KeyT, int D>
__global__ __launch_bounds__(CH_WG) void k_link_items(const KeyT* __restrict__ skey, const uint64_t* __restrict__ spay,
                                                      const uint32_t* __restrict__ item_total,
                                                      const FrameDev* __restrict__ fr,
                                                      const uint32_t* __restrict__ frame_pt_begin, uint32_t nframes,
                                                      uint32_t* __restrict__ idrel) {
    __shared__ uint32_t s_off[CH_MAXF_LDS + 1];
    const uint32_t* offs = frame_pt_begin;
    if (nframes <= CH_MAXF_LDS) {
        for (uint32_t k = threadIdx.x; k <= nframes; k += CH_WG) s_off[k] = frame_pt_begin[k];
        __syncthreads();
        offs = s_off;
    }
    const uint64_t M = *item_total;
    const uint64_t s0 = (uint64_t)blockIdx.x * CH_TILE;
    if (s0 >= M) return;
    uint64_t pa[CH_ITEMS], pb[CH_ITEMS];
    KeyT kv[CH_ITEMS], kn[CH_ITEMS];
#pragma unroll
    for (int q = 0; q < CH_ITEMS; ++q) {
        const uint64_t s = s0 + (uint64_t)q * CH_WG + threadIdx.x;
        // unconditional loads at valid positions (the tile's first one)
        const uint64_t sc = s < M ? s : s0, nc = s + 1 < M ? s + 1 : s0;
        pa[q] = spay[sc];
        kv[q] = skey[sc];
        pb[q] = spay[nc];
        kn[q] = skey[nc];
    }
#pragma unroll
    for (int q = 0; q < CH_ITEMS; ++q) {
        const uint64_t s = s0 + (uint64_t)q * CH_WG + threadIdx.x;
        if (s + 1 >= M || kn[q] != kv[q]) continue;
        const uint32_t ex = (uint32_t)(pa[q] >> 32), en = (uint32_t)pb[q];
        const uint32_t f = upper_find(offs, nframes, ex);
        const FrameDev& F = fr[f];
        if (en < offs[f + 1] && chain_relative<D>(F, (uint64_t)kv[q], (uint64_t)(en - ex))) idrel[ex] = en - ex;
    }
}

// keys wider than 64 bits: the sort runs word by word (least significant word
// first); before the passes over word w its values are gathered in the order
// the previous words left behind
__global__ void k_gather_word(const uint64_t* __restrict__ word, const uint32_t* __restrict__ order, uint64_t T,
                              uint64_t* __restrict__ out) {
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < T; s += (uint64_t)gridDim.x * blockDim.x)
        out[s] = word[order[s]];
}

// successor + decision for multi-word keys: equal tuples need every word equal
template <int D>
__global__ __launch_bounds__(CH_WG) void k_link_wide(const uint64_t* __restrict__ words,
                                                     const uint32_t* __restrict__ sval, uint64_t T, uint32_t nwords,
                                                     const FrameDev* __restrict__ fr,
                                                     const uint32_t* __restrict__ frame_pt_begin, uint32_t nframes,
                                                     uint32_t* __restrict__ idrel) {
    __shared__ uint32_t s_off[CH_MAXF_LDS + 1];
    const uint32_t* offs = frame_pt_begin;
    if (nframes <= CH_MAXF_LDS) {
        for (uint32_t k = threadIdx.x; k <= nframes; k += CH_WG) s_off[k] = frame_pt_begin[k];
        __syncthreads();
        offs = s_off;
    }
    for (uint64_t s = blockIdx.x * (uint64_t)CH_WG + threadIdx.x; s < T; s += (uint64_t)gridDim.x * CH_WG) {
        const uint32_t i = sval[s];
        uint32_t out = 0;
        if (s + 1 < T) {
            const uint32_t j = sval[s + 1];
            bool same = true;
            for (uint32_t w = 0; w < nwords && same; ++w) same = words[(uint64_t)w * T + i] == words[(uint64_t)w * T + j];
            const uint32_t f = upper_find(offs, nframes, i);
            const FrameDev& F = fr[f];
            if (same && (F.flags & FL_L1) && j < offs[f + 1]) {
                const uint64_t dist = (uint64_t)(j - i);
                uint64_t mb = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const uint32_t wd = F.key_width[d];
                    const uint64_t mask = wd >= 64 ? ~0ull : ((1ull << wd) - 1ull);
                    const uint64_t b = (words[(uint64_t)F.key_word[d] * T + i] >> F.key_shift[d]) & mask;
                    mb = b > mb ? b : mb;
                }
                const int64_t local_bits = bitlen64(mb);
                const int64_t tag_bits = bitlen64((uint64_t)local_bits);
                const int64_t a = local_bits * D + tag_bits;
                const int64_t c = (int64_t)D * (int64_t)F.max_bid_bits;
                const int64_t smaller = a < c ? a : c;
                const int64_t bl = bitlen64(dist);
                const int64_t lhs = bl + bitlen64((uint64_t)bl);
                if (lhs < smaller - F.threshold) out = (uint32_t)dist;
            }
        }
        idrel[i] = out;
    }
}

// ------------------------------------------------------------------------
// K5: ordered compaction of the explicit points' block ids (index order),
// per-frame point tiles, one decoupled look-back over groups of TPW
// consecutive tiles (one ticket and one look-back step per group, which
// halves the serial part per point against one tile per workgroup).

// tiles per workgroup: 4 with one-word 32-bit keys, 2 with 64-bit keys (registers)
template <typename KeyT> constexpr int expl_tpw() { return sizeof(KeyT) == 4 ? 4 : 2; }

// BT: element type of the block ids written (32-bit when every frame's block
// ids stay below 2^31)
template <typename KeyT, int D, int TPW, typename BT>
__global__ __launch_bounds__(CH_WG) void k_explicit(const uint32_t* __restrict__ idrel, const KeyT* __restrict__ keys,
                                                    const FrameDev* __restrict__ fr,
                                                    const uint32_t* __restrict__ tile_begin, uint32_t nframes,
                                                    uint32_t ntiles, uint64_t* status, uint32_t* ticket,
                                                    uint32_t* err, BT* __restrict__ blocks,
                                                    FrameOut* __restrict__ fo, uint64_t* __restrict__ rec_total,
                                                    uint32_t nwords, uint64_t T) {
    // striped: item k of thread t is point k*256 + t of its tile (coalesced);
    // record positions from ballots, scanned (tile, row, wave) counts and a
    // decoupled look-back over the groups
    constexpr uint32_t W = CH_WG / 64;
    constexpr uint32_t NSEG = TPW * CH_ITEMS * W;   // (tile, row, wave) segments
    constexpr uint32_t SPL = (NSEG + 63) / 64;       // segments scanned per lane
    __shared__ uint32_t s_grp, s_off[NSEG];
    __shared__ uint64_t s_excl;
    if (threadIdx.x == 0) s_grp = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t grp = s_grp;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    uint32_t tl[TPW], fl[TPW];
    uint64_t st[TPW], nn[TPW], po[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {   // a tile past the end has no live points
        tl[u] = grp * TPW + u;
        const uint32_t tc = tl[u] < ntiles ? tl[u] : ntiles - 1;
        fl[u] = upper_find(tile_begin, nframes, tc);
        const FrameDev& F = fr[fl[u]];
        st[u] = (uint64_t)(tc - F.tile0) * CH_TILE;
        nn[u] = tl[u] < ntiles ? F.n : 0;
        po[u] = F.pt_off;
    }
    uint32_t flags[TPW];
    // one-word keys of the tiles are loaded before the look-back waits
    KeyT kpre[TPW][CH_ITEMS];
    uint32_t idr[TPW][CH_ITEMS];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k) {
            const uint64_t i = st[u] + k * CH_WG + threadIdx.x;
            const bool live = i < nn[u];
            // unconditional loads at a valid index (the tile's first point)
            const uint64_t ic = po[u] + (live ? i : st[u]);
            const uint32_t v = idrel[ic];
            const KeyT kk = keys[ic];
            idr[u][k] = live ? v : 1u;
            kpre[u][k] = live && nwords == 1 ? kk : KeyT(0);
        }
    }
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        flags[u] = 0;
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k) {
            const bool ex = idr[u][k] == 0;
            const uint64_t m = __ballot(ex);
            flags[u] |= (ex ? 1u : 0u) << k;
            if (lane == 0) s_off[(u * CH_ITEMS + k) * W + wave] = __popcll(m);
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // lane l scans segments [SPL*l, SPL*l + SPL)
        uint32_t loc[SPL], sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < SPL; ++j) {
            const uint32_t ix = SPL * threadIdx.x + j;
            loc[j] = ix < NSEG ? s_off[ix] : 0u;
            sum += loc[j];
        }
        const uint32_t inc = wave_incl_scan(sum);
        uint32_t run = inc - sum;
        uint32_t mine[SPL];   // exclusive offset of each of the lane's segments
#pragma unroll
        for (uint32_t j = 0; j < SPL; ++j) {
            const uint32_t ix = SPL * threadIdx.x + j;
            mine[j] = run;
            if (ix < NSEG) s_off[ix] = run;
            run += loc[j];
        }
        const uint32_t tot = __shfl(inc, 63, 64);
        const uint64_t e = lookback(status, grp, tot, err);
        if (threadIdx.x == 0) s_excl = e;
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            // the tile's first segment is entry (u*CH_ITEMS*W) % SPL of lane (u*CH_ITEMS*W) / SPL
            const uint32_t seg = u * CH_ITEMS * W;
            uint32_t sel = 0;
#pragma unroll
            for (uint32_t j = 0; j < SPL; ++j) sel = j == seg % SPL ? mine[j] : sel;
            const uint32_t tile_pre = __shfl(sel, (int)(seg / SPL), 64);
            if (threadIdx.x == 0 && tl[u] < ntiles) {
                if (tl[u] == fr[fl[u]].tile0) fo[fl[u]].rec_base = e + tile_pre;
                if (tl[u] == ntiles - 1) *rec_total = e + tot;
            }
        }
    }
    __syncthreads();
    const uint64_t base = s_excl;
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        if (!flags[u]) continue;
        const FrameDev& F = fr[fl[u]];
        // the key layout in registers: the record stores then need no wait
        uint32_t sh[D], kw[D];
        uint64_t mk[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const uint32_t wd = F.key_width[d];
            mk[d] = wd >= 64 ? ~0ull : ((1ull << wd) - 1ull);
            sh[d] = F.key_shift[d];
            kw[d] = F.key_word[d];
        }
        if (nwords == 1) {
#pragma unroll
            for (int k = 0; k < CH_ITEMS; ++k) {
                if (!(flags[u] & (1u << k))) continue;
                // the active lanes are the row's explicit points: rank = active lanes below
                const uint32_t pre = __popcll(__ballot(1) & lt);
                const uint64_t rec = base + s_off[(u * CH_ITEMS + k) * W + wave] + pre;
#pragma unroll
                for (int d = 0; d < D; ++d) blocks[rec * D + d] = (BT)(((uint64_t)kpre[u][k] >> sh[d]) & mk[d]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < CH_ITEMS; ++k) {
                if (!(flags[u] & (1u << k))) continue;
                const uint64_t i = st[u] + k * CH_WG + threadIdx.x;
                const uint32_t pre = __popcll(__ballot(1) & lt);
                const uint64_t rec = base + s_off[(u * CH_ITEMS + k) * W + wave] + pre;
                uint64_t kd[D];
#pragma unroll
                for (int d = 0; d < D; ++d) kd[d] = (uint64_t)keys[(uint64_t)kw[d] * T + po[u] + i];
#pragma unroll
                for (int d = 0; d < D; ++d) blocks[rec * D + d] = (BT)((kd[d] >> sh[d]) & mk[d]);
            }
        }
    }
}

// record counts per frame from the bases (one thread per frame)
__global__ void k_frame_counts(FrameOut* fo, uint32_t nframes, const uint64_t* rec_total) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nframes) return;
    const uint64_t next = f + 1 < nframes ? fo[f + 1].rec_base : *rec_total;
    fo[f].nrec = next - fo[f].rec_base;
    fo[f].blk_min = ~0ull;
    fo[f].blk_max = 0;
    fo[f].blk_maxdiff = 0;
}

// ------------------------------------------------------------------------
// K6: per-frame L2 statistics over the explicit records: minimum, maximum and
// largest consecutive difference (cpp/chain/lcps_cpp.cpp:1338-1358).
// Each thread folds a run of records; runs of one frame reduce across the
// wavefront before one atomic per statistic.

__device__ inline uint32_t frame_of_rec(const uint64_t* s_base, uint32_t nframes, uint64_t r) {
    uint32_t lo = 0, hi = nframes;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (s_base[mid] <= r) lo = mid; else hi = mid;
    }
    return lo;
}

// Contiguous record ranges per workgroup, 256 consecutive records per step
// (coalesced D x u64 loads); per-thread min / max / max |difference| per
// frame, flushed when the frame changes and reduced per block at the end, so
// a frame costs a few atomics per block, not per record.
template <int D, typename BT>
__global__ __launch_bounds__(CH_WG) void k_l2_stats(const BT* __restrict__ blocks,
                                                    const uint64_t* __restrict__ rec_total, uint32_t nframes,
                                                    const uint64_t* __restrict__ frame_rec_base,
                                                    FrameOut* __restrict__ fo) {
    constexpr uint32_t W = CH_WG / 64;
    __shared__ uint64_t s_base[CH_MAXF_LDS];
    __shared__ uint32_t s_f[W];
    __shared__ uint64_t s_mn[W], s_mx[W], s_md[W];
    const uint64_t R = *rec_total;
    const bool lds = nframes <= CH_MAXF_LDS;
    if (lds)
        for (uint32_t k = threadIdx.x; k < nframes; k += CH_WG) s_base[k] = frame_rec_base[k];
    __syncthreads();
    const uint64_t* bases = lds ? s_base : frame_rec_base;
    uint32_t cur = 0xffffffffu;
    uint64_t mn = ~0ull, mx = 0, md = 0;
    auto flush = [&]() {
        if (mn != ~0ull) {
            atomicMin((unsigned long long*)&fo[cur].blk_min, mn);
            atomicMax((unsigned long long*)&fo[cur].blk_max, mx);
            atomicMax((unsigned long long*)&fo[cur].blk_maxdiff, md);
        }
        mn = ~0ull;
        mx = 0;
        md = 0;
    };
    // every workgroup walks one contiguous range of records, 256 at a time,
    // so a thread's frame changes only at frame boundaries
    const uint64_t per = ((R + gridDim.x - 1) / gridDim.x + CH_WG - 1) / CH_WG * CH_WG;
    const uint64_t b0 = (uint64_t)blockIdx.x * per, b1 = b0 + per < R ? b0 + per : R;
    // four records per thread and round: the records and their predecessors
    // are loaded first (unconditionally: past the end, the round's first
    // record; record 0 stands in for its own predecessor)
    for (uint64_t rr = b0 + threadIdx.x; rr < b1; rr += 4 * CH_WG) {
        uint64_t cv[4][D], pv[4][D];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t r = rr + (uint64_t)u * CH_WG;
            const uint64_t rc = r < b1 ? r : rr, rp = rc ? rc - 1 : rc;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                cv[u][d] = (uint64_t)blocks[rc * D + d];
                pv[u][d] = (uint64_t)blocks[rp * D + d];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t r = rr + (uint64_t)u * CH_WG;
            if (r >= b1) continue;
            const uint32_t f = frame_of_rec(bases, nframes, r);
            if (f != cur) {
                flush();
                cur = f;
            }
            const bool has_prev = r > bases[f];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const uint64_t v = cv[u][d];
                mn = v < mn ? v : mn;
                mx = v > mx ? v : mx;
                const uint64_t w = pv[u][d];
                const uint64_t ad = w > v ? w - v : v - w;
                md = has_prev && ad > md ? ad : md;
            }
        }
    }
    // waves whose lanes all hold one frame reduce by shuffles; others flush per lane
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t has = __ballot(mn != ~0ull);
    uint32_t c0 = 0xffffffffu;
    if (has) c0 = __shfl(cur, __ffsll((unsigned long long)has) - 1, 64);
    const bool uniform = has && __ballot(mn != ~0ull && cur != c0) == 0;
    if (uniform) {
        for (int k = 32; k >= 1; k >>= 1) {
            const uint64_t a = __shfl_xor(mn, k, 64), b = __shfl_xor(mx, k, 64), c = __shfl_xor(md, k, 64);
            mn = a < mn ? a : mn;
            mx = b > mx ? b : mx;
            md = c > md ? c : md;
        }
    } else {
        flush();
    }
    if (lane == 0) {
        s_f[wave] = uniform ? c0 : 0xffffffffu;
        s_mn[wave] = mn;
        s_mx[wave] = mx;
        s_md[wave] = md;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 0; w < W; ++w) {
            if (s_f[w] == 0xffffffffu) continue;
            uint64_t a = s_mn[w], b = s_mx[w], c = s_md[w];
            for (uint32_t w2 = w + 1; w2 < W; ++w2)
                if (s_f[w2] == s_f[w]) {
                    a = s_mn[w2] < a ? s_mn[w2] : a;
                    b = s_mx[w2] > b ? s_mx[w2] : b;
                    c = s_md[w2] > c ? s_md[w2] : c;
                    s_f[w2] = 0xffffffffu;
                }
            atomicMin((unsigned long long*)&fo[s_f[w]].blk_min, a);
            atomicMax((unsigned long long*)&fo[s_f[w]].blk_max, b);
            atomicMax((unsigned long long*)&fo[s_f[w]].blk_maxdiff, c);
        }
    }
}

// wrap parameters (cpp/chain/lcps_cpp.cpp:1221-1253) and fixed costs
template <int D>
__global__ void k_l2_wrap(FrameOut* fo, uint32_t nframes, uint32_t* err) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nframes) return;
    FrameOut& o = fo[f];
    const uint64_t mn = o.blk_min, mx = o.blk_max;
    if (mn > mx || mx - mn >= (uint64_t)INT64_MAX) { atomicOr(err, 2u); return; }
    o.wrap_min = (int64_t)mn;
    o.wrap_max = (int64_t)mx;
    o.wrap_range = (int64_t)(mx - mn + 1);
    o.wrap_max_corr = o.wrap_range / 2;
    o.wrap_min_corr = -o.wrap_max_corr;
    if ((o.wrap_range & 1) == 0) --o.wrap_max_corr;
    o.fixed_value_bits = (uint64_t)D * bitlen64(mx);
    o.fixed_diff_bits = (uint64_t)D * bitlen64(o.blk_maxdiff);
}

// ------------------------------------------------------------------------
// K7: L2 choice per record and ordered split into diff / keep streams
// (cpp/chain/lcps_cpp.cpp:1360-1454; keep-all when L2 is off, :1456-1478).

// Striped tiles of 4096 records (item k of thread t = record k*256 + t):
// choices by ballot, ordered positions from (row, wave) counts and a
// decoupled look-back over tiles; coalesced loads and stores.
#define L2S_GROUP 4
#define L2S_MAXF_LDS 1024u   // frame bases kept in LDS up to this many frames
#define L2S_WGS 2048u        // workgroups of a split launch at most
// DKT: element type of the diff and keep streams (32-bit when every frame's
// block ids, and so every wrapped difference, stay below 2^31)
// Occupancy: up to three dimensions the kernel is held to 168 VGPRs for three
// waves per SIMD (the 32-bit, three-dimension variant needs 178 unconstrained
// and spills 12 at 168: 0.091 vs 0.110 ms on a 50M-point frame); at four
// waves it would spill 66.
template <int D, typename DKT>
__global__ __launch_bounds__(CH_WG) __attribute__((amdgpu_waves_per_eu(D <= 3 ? 3 : 1))) void k_l2_split(const DKT* __restrict__ blocks,
                                                    const uint64_t* __restrict__ rec_total, uint32_t nframes,
                                                    const uint64_t* __restrict__ frame_rec_base,
                                                    const FrameDev* __restrict__ fr, FrameOut* __restrict__ fo,
                                                    uint64_t* status, uint32_t* ticket, uint32_t* err,
                                                    uint8_t* __restrict__ choice2, DKT* __restrict__ diff,
                                                    DKT* __restrict__ keep, uint64_t* __restrict__ diff_total) {
    constexpr uint32_t W = CH_WG / 64;
    __shared__ uint64_t s_base[L2S_MAXF_LDS];
    __shared__ uint32_t s_tile, s_off[CH_ITEMS * W];
    __shared__ uint64_t s_excl;
    const uint64_t R = *rec_total;
    const bool lds = nframes <= L2S_MAXF_LDS;
    if (lds)
        for (uint32_t k = threadIdx.x; k < nframes; k += CH_WG) s_base[k] = frame_rec_base[k];
    const uint64_t* bases = lds ? s_base : frame_rec_base;
    // At most L2S_WGS workgroups, each taking tiles by ticket until they run
    // past the record count (known only on the device): a frame with few
    // explicit records costs no workgroups for the tiles it does not have.
    for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if ((uint64_t)tile * CH_TILE >= R) break;
    const uint64_t t0 = (uint64_t)tile * CH_TILE;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    // A tile inside one frame (the usual case) reads that frame's parameters
    // once, into registers; a tile across frames looks them up per record.
    const uint64_t tl = (t0 + CH_TILE < R ? t0 + CH_TILE : R) - 1;
    const uint32_t fa = frame_of_rec(bases, nframes, t0);
    const bool one = fa == frame_of_rec(bases, nframes, tl);
    const uint64_t base_a = bases[fa];
    const bool l2_a = (fr[fa].flags & FL_L2) != 0;
    const uint64_t fvb_a = fo[fa].fixed_value_bits, fdb_a = fo[fa].fixed_diff_bits;
    const int64_t wminc_a = fo[fa].wrap_min_corr, wmaxc_a = fo[fa].wrap_max_corr, wr_a = fo[fa].wrap_range;
    if constexpr (sizeof(DKT) == 4) {
        // 32-bit streams (every block id below 2^31): the first pass decides
        // each record's choice and computes the value it writes (its block, or
        // the wrapped difference), kept in registers over the look-back, so the
        // second pass only stores. Blocks are 32-bit values here.
        uint32_t ov[CH_ITEMS][D];
        uint32_t chosen = 0;
#pragma unroll
        for (int kg = 0; kg < CH_ITEMS; kg += L2S_GROUP) {
            uint32_t cur[L2S_GROUP][D], prv[L2S_GROUP][D];
#pragma unroll
            for (int j = 0; j < L2S_GROUP; ++j) {
                const uint64_t r = t0 + (kg + j) * CH_WG + threadIdx.x;
                const uint64_t rc = r < R ? r : t0, rp = rc ? rc - 1 : rc;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    cur[j][d] = blocks[rc * D + d];
                    prv[j][d] = blocks[rp * D + d];
                }
            }
#pragma unroll
            for (int j = 0; j < L2S_GROUP; ++j) {
                const int k = kg + j;
                const uint64_t r = t0 + k * CH_WG + threadIdx.x;
                bool c = false;
                if (r < R) {
                    uint32_t f = fa;
                    uint64_t fbase = base_a, fvb = fvb_a, fdb = fdb_a;
                    bool l2 = l2_a;
                    int64_t wminc = wminc_a, wmaxc = wmaxc_a, wr = wr_a;
                    if (!one) {
                        f = frame_of_rec(bases, nframes, r);
                        fbase = bases[f];
                        l2 = (fr[f].flags & FL_L2) != 0;
                        fvb = fo[f].fixed_value_bits;
                        fdb = fo[f].fixed_diff_bits;
                        wminc = fo[f].wrap_min_corr;
                        wmaxc = fo[f].wrap_max_corr;
                        wr = fo[f].wrap_range;
                    }
                    if (r != fbase && l2) {
                        uint32_t mv = 0, mdf = 0;
#pragma unroll
                        for (int d = 0; d < D; ++d) {
                            const uint32_t v = cur[j][d], u = prv[j][d];
                            mv = v > mv ? v : mv;
                            const uint32_t ad = u > v ? u - v : v - u;
                            mdf = ad > mdf ? ad : mdf;
                        }
                        const uint64_t tvw = 32u - (uint32_t)__clz(mv), tdw = 32u - (uint32_t)__clz(mdf);
                        const uint64_t tv = (uint64_t)D * tvw + bitlen64(tvw);
                        const uint64_t td = (uint64_t)D * tdw + bitlen64(tdw);
                        const uint64_t value_bits = tv < fvb ? tv : fvb;
                        const uint64_t diff_bits = td > fdb ? td : fdb;
                        c = value_bits >= diff_bits;
                    }
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        if (c) {
                            // prediction = previous record clamped to [wrap_min, wrap_max] (an identity here)
                            int64_t corr = (int64_t)prv[j][d] - (int64_t)cur[j][d];
                            if (corr < wminc) corr += wr;
                            else if (corr > wmaxc) corr -= wr;
                            ov[k][d] = (uint32_t)(corr >= 0 ? ((uint64_t)corr << 1)
                                                            : (((uint64_t)(-(corr + 1)) << 1) | 1ull));
                        } else {
                            ov[k][d] = cur[j][d];
                        }
                    }
                }
                const uint64_t m = __ballot(c);
                chosen |= (c ? 1u : 0u) << k;
                if (lane == 0) s_off[k * W + wave] = __popcll(m);
            }
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            const uint32_t v = s_off[threadIdx.x];
            const uint32_t inc = wave_incl_scan(v);
            s_off[threadIdx.x] = inc - v;
            const uint32_t tot = __shfl(inc, 63, 64);
            const uint64_t e = lookback(status, tile, tot, err);
            if (threadIdx.x == 0) s_excl = e;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k) {
            const uint64_t r = t0 + k * CH_WG + threadIdx.x;
            const bool c = (chosen >> k) & 1u;
            const uint32_t pre = __popcll(__ballot(c) & lt);   // diffs of the row before this lane
            if (r >= R) continue;
            const uint32_t f = one ? fa : frame_of_rec(bases, nframes, r);
            const uint64_t dpos = s_excl + s_off[k * W + wave] + pre;   // diffs before record r
            if (r == (one ? base_a : bases[f])) {
                fo[f].diff_base = dpos;
                fo[f].keep_base = r - dpos;
            }
            choice2[r] = (uint8_t)c;
            if (!c) {
                const uint64_t kpos = r - dpos;
#pragma unroll
                for (int d = 0; d < D; ++d) keep[kpos * D + d] = (DKT)ov[k][d];
            } else {
#pragma unroll
                for (int d = 0; d < D; ++d) diff[dpos * D + d] = (DKT)ov[k][d];
            }
            if (r == R - 1) *diff_total = dpos + (c ? 1u : 0u);
        }
        __syncthreads();   // LDS reused by the next tile
        continue;
    } else {
    // Items are taken L2S_GROUP at a time: the blocks of the group's records
    // and of their predecessors are loaded unconditionally (clamped to the
    // tile's first record) before any is used.
    // per-item ballot prefix and frame live in LDS: the groups are a loop
    __shared__ uint16_t s_pre[CH_ITEMS][CH_WG];
    __shared__ uint32_t s_fk[CH_ITEMS][CH_WG];
    uint32_t chosen = 0;
#pragma unroll 1
    for (int kg = 0; kg < CH_ITEMS; kg += L2S_GROUP) {
        uint64_t cur[L2S_GROUP][D], prv[L2S_GROUP][D];
#pragma unroll
        for (int j = 0; j < L2S_GROUP; ++j) {
            const uint64_t r = t0 + (kg + j) * CH_WG + threadIdx.x;
            const uint64_t rc = r < R ? r : t0, rp = rc ? rc - 1 : rc;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                cur[j][d] = blocks[rc * D + d];
                prv[j][d] = blocks[rp * D + d];
            }
        }
#pragma unroll
        for (int j = 0; j < L2S_GROUP; ++j) {
            const int k = kg + j;
            const uint64_t r = t0 + k * CH_WG + threadIdx.x;
            bool c = false;
            uint32_t fkk = 0;
            if (r < R) {
                uint32_t f = fa;
                uint64_t fbase = base_a, fvb = fvb_a, fdb = fdb_a;
                bool l2 = l2_a;
                if (!one) {
                    f = frame_of_rec(bases, nframes, r);
                    fbase = bases[f];
                    l2 = (fr[f].flags & FL_L2) != 0;
                    fvb = fo[f].fixed_value_bits;
                    fdb = fo[f].fixed_diff_bits;
                }
                fkk = f;
                if (r != fbase && l2) {
                    uint64_t mv = 0, mdf = 0;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const uint64_t v = cur[j][d], u = prv[j][d];
                        mv = v > mv ? v : mv;
                        const uint64_t ad = u > v ? u - v : v - u;
                        mdf = ad > mdf ? ad : mdf;
                    }
                    const uint64_t tvw = bitlen64(mv), tdw = bitlen64(mdf);
                    const uint64_t tv = (uint64_t)D * tvw + bitlen64(tvw);
                    const uint64_t td = (uint64_t)D * tdw + bitlen64(tdw);
                    const uint64_t value_bits = tv < fvb ? tv : fvb;
                    const uint64_t diff_bits = td > fdb ? td : fdb;
                    c = value_bits >= diff_bits;
                }
            }
            const uint64_t m = __ballot(c);
            s_pre[k][threadIdx.x] = (uint16_t)__popcll(m & lt);
            s_fk[k][threadIdx.x] = fkk;
            chosen |= (c ? 1u : 0u) << k;
            if (lane == 0) s_off[k * W + wave] = __popcll(m);
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t v = s_off[threadIdx.x];
        const uint32_t inc = wave_incl_scan(v);
        s_off[threadIdx.x] = inc - v;
        const uint32_t tot = __shfl(inc, 63, 64);
        const uint64_t e = lookback(status, tile, tot, err);
        if (threadIdx.x == 0) s_excl = e;
    }
    __syncthreads();
#pragma unroll 1
    for (int kg = 0; kg < CH_ITEMS; kg += L2S_GROUP) {
        uint64_t cur[L2S_GROUP][D], prv[L2S_GROUP][D];
#pragma unroll
        for (int j = 0; j < L2S_GROUP; ++j) {
            const uint64_t r = t0 + (kg + j) * CH_WG + threadIdx.x;
            const uint64_t rc = r < R ? r : t0, rp = rc ? rc - 1 : rc;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                cur[j][d] = blocks[rc * D + d];
                prv[j][d] = blocks[rp * D + d];
            }
        }
#pragma unroll
        for (int j = 0; j < L2S_GROUP; ++j) {
            const int k = kg + j;
            const uint64_t r = t0 + k * CH_WG + threadIdx.x;
            if (r >= R) continue;
            const uint32_t f = one ? fa : s_fk[k][threadIdx.x];
            const uint64_t dpos = s_excl + s_off[k * W + wave] + s_pre[k][threadIdx.x];   // diffs before record r
            if (r == (one ? base_a : bases[f])) {
                fo[f].diff_base = dpos;
                fo[f].keep_base = r - dpos;
            }
            const bool c = (chosen >> k) & 1u;
            choice2[r] = (uint8_t)c;
            if (!c) {
                const uint64_t kpos = r - dpos;
#pragma unroll
                for (int d = 0; d < D; ++d) keep[kpos * D + d] = (DKT)cur[j][d];
            } else {
                int64_t wminc = wminc_a, wmaxc = wmaxc_a, wr = wr_a;
                if (!one) {
                    wminc = fo[f].wrap_min_corr;
                    wmaxc = fo[f].wrap_max_corr;
                    wr = fo[f].wrap_range;
                }
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    // prediction = previous record clamped to [wrap_min, wrap_max] (an identity here)
                    int64_t corr = (int64_t)prv[j][d] - (int64_t)cur[j][d];
                    if (corr < wminc) corr += wr;
                    else if (corr > wmaxc) corr -= wr;
                    diff[dpos * D + d] =
                        (DKT)(corr >= 0 ? ((uint64_t)corr << 1) : (((uint64_t)(-(corr + 1)) << 1) | 1ull));
                }
            }
            if (r == R - 1) *diff_total = dpos + (c ? 1u : 0u);
        }
    }
    __syncthreads();   // LDS reused by the next tile
    }
    }
}

__global__ void k_copy_rec_base(const FrameOut* fo, uint32_t nframes, uint64_t* rec_base) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nframes) rec_base[f] = fo[f].rec_base;
}

__global__ void k_frame_corr(FrameOut* fo, uint32_t nframes, const uint32_t* tile_begin,
                             const uint32_t* tile_corr_off, const uint64_t* corr_total) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nframes) return;
    const uint64_t b = tile_corr_off[tile_begin[f]];
    const uint64_t e = f + 1 < nframes ? tile_corr_off[tile_begin[f + 1]] : *corr_total;
    fo[f].corr_base = b;
    fo[f].ncorr = e - b;
}

__global__ void k_frame_split_counts(FrameOut* fo, uint32_t nframes, const uint64_t* diff_total,
                                     const uint64_t* rec_total) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nframes) return;
    const uint64_t dnext = f + 1 < nframes ? fo[f + 1].diff_base : *diff_total;
    const uint64_t knext = f + 1 < nframes ? fo[f + 1].keep_base : (*rec_total - *diff_total);
    fo[f].ndiff = dnext - fo[f].diff_base;
    fo[f].nkeep = knext - fo[f].keep_base;
}

}  // namespace chain
