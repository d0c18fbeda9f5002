// Entropy stage and LCH1 container assembly of the Chain compressor (gfx950).
//
// For each frame and each of the five components (idrel, diff, keep, choice2,
// rel) the engine sizes three candidate encodings, estimating the rANS-coded
// ones from their histograms (chain_int.hpp), encodes only the candidates
// whose estimate is within 1/256 of the best and keeps the smallest of those,
// ties in the order Tagged, Fixed, Raw-rANS (the reference's chooser order,
// cpp/chain/lcps_cpp.cpp:1140-1174, with Zstd replaced):
//   TAGR  per-record widths rANS-coded + MSB-first point-major values
//         (the Tagged payload of :1005-1061)
//   FIXP  per-dimension minimum and width + MSB-first payloads (:856-933)
//   RANS  point-major values rANS-coded when every value is < 4096 (:935-995)
// rANS sections (RVC1) use the RVN1 table of :560-684 and split the symbols
// into chunks of lane-interleaved states, one wavefront per chunk.
// Everything stays on the device: histograms, tables, encoding, sizes, the
// layout of the output and the final bytes.
#include "chain_common.hpp"

#pragma clang fp contract(off)

namespace chain {

enum : uint32_t { KIND_TAGR = 0, KIND_FIXP = 1, KIND_RANS = 2 };

struct CompInst {
    uint64_t base, n;          // first record in the component stream, records
    uint64_t dmin[CH_MAXD], dmax[CH_MAXD];
    uint64_t tag_bits;         // sum over records of dims * width
    uint32_t max_tag, overflow;
    uint32_t kind, dims;
    uint32_t width[CH_MAXD];
    uint64_t fix_bytes[CH_MAXD];
    uint64_t size;             // component bytes
    uint64_t off;              // absolute output offset of the component magic
    uint64_t payload_off;      // FIXP payload / TAGR packed values
    uint32_t cand, pad;        // contenders (chain_int.hpp lch1_contenders)
};

struct Section {               // RVC1 section of one component: s = 0 values, 1 widths
    uint64_t count;
    uint32_t eligible, pb, alphabet, lanes;
    uint32_t table_len, nchunks, chunk_base, chunk_symbols;
    uint64_t payload;          // bytes of all chunks (states + streams)
    uint64_t size;             // bytes of the whole section
    uint64_t off;              // absolute output offset of "RVC1"
    uint64_t est_bits;         // estimated stream bits, in 1/65536 (chain_int.hpp)
    uint32_t encode, pad;      // encoded only when its candidate contends
};

#define RVC1_HDR_FIXED 30u     // "RVC1" u8 u64 u64 u8 u32 u32
#define TABLE_BYTES_MAX 12304u // varuint(4096) + 4096 x 3 bytes
#define CHUNK_STATE_BYTES 256u // 64 lanes x u32

__device__ inline void put_u8(uint8_t* p, uint8_t v) { *p = v; }
__device__ inline void put_u32(uint8_t* p, uint32_t v) {
    for (int i = 0; i < 4; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
__device__ inline void put_u64(uint8_t* p, uint64_t v) {
    for (int i = 0; i < 8; ++i) p[i] = (uint8_t)(v >> (8 * i));
}

__device__ inline uint32_t comp_dims(int c, uint32_t D) { return (c == 0 || c == 3) ? 1u : D; }

// segmented decoupled look-back: tiles [seg_start, tile) form the segment
__device__ inline uint64_t lookback_seg(uint64_t* status, uint32_t tile, uint32_t seg_start, uint64_t aggregate,
                                        uint32_t* err) {
    const uint32_t lane = lane_id();
    if (tile == seg_start) {
        if (lane == 0) lb_store(&status[tile], LB_INC | aggregate);
        return 0;
    }
    if (lane == 0) lb_store(&status[tile], LB_AGG | aggregate);
    uint64_t excl = 0;
    int64_t t = (int64_t)tile - 1;
    uint32_t spins = 0;
    while (t >= (int64_t)seg_start) {
        const int64_t mine = t - (int64_t)lane;
        uint64_t s = mine >= (int64_t)seg_start ? lb_load(&status[mine]) : LB_INC;
        const uint64_t not_ready = __ballot((s >> 62) == 0);
        const uint64_t inc_mask = __ballot((s >> 62) == 2);
        const uint64_t first_inc = inc_mask ? (inc_mask & (~inc_mask + 1)) : 0;
        const uint64_t first_nr = not_ready ? (not_ready & (~not_ready + 1)) : 0;
        if (first_nr && (!first_inc || first_nr < first_inc)) {
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

// ------------------------------------------------------------------------
// component instances

// one workgroup per component instance: its statistics record and its two
// histograms start at zero
__global__ __launch_bounds__(CH_WG) void k_comp_init(const FrameDev* __restrict__ fr, const FrameOut* __restrict__ fo,
                                                     uint32_t F, uint32_t D, CompInst* __restrict__ ci,
                                                     uint32_t* __restrict__ vhist, uint32_t* __restrict__ thist) {
    const uint32_t i = blockIdx.x;
    for (uint32_t k = threadIdx.x; k < RANS_VALUE_LIMIT; k += CH_WG) vhist[(uint64_t)i * RANS_VALUE_LIMIT + k] = 0;
    if (threadIdx.x < RANS_TAG_LIMIT) thist[(uint64_t)i * RANS_TAG_LIMIT + threadIdx.x] = 0;
    if (threadIdx.x != 0) return;
    const int c = i / F;
    const uint32_t f = i % F;
    CompInst x;
    memset(&x, 0, sizeof x);
    switch (c) {
        case 0: x.base = fr[f].pt_off; x.n = fr[f].n; break;
        case 1: x.base = fo[f].diff_base; x.n = fo[f].ndiff; break;
        case 2: x.base = fo[f].keep_base; x.n = fo[f].nkeep; break;
        case 3: x.base = fo[f].rec_base; x.n = fo[f].nrec; break;
        default: x.base = fr[f].pt_off; x.n = fr[f].n; break;
    }
    x.dims = comp_dims(c, D);
    for (int d = 0; d < CH_MAXD; ++d) { x.dmin[d] = ~0ull; x.dmax[d] = 0; }
    ci[i] = x;
}

// Component statistics: range, value histogram (< 4096), width histogram and
// total Tagged bits.
// One workgroup walks a contiguous range of tiles and keeps its histograms
// and range statistics in LDS / registers until the frame changes.
// Small values and widths (< CS_PRIV_BINS, the common case: in-block offsets,
// choice flags, short links) are counted in per-thread LDS columns of 8-bit
// counters: bin b of thread t is byte (b & 3) of word [b >> 2][t], so the 64
// lanes of a wavefront add to 64 distinct words and never serialise on a
// shared bin. Larger values use a shared LDS histogram (256 bins for byte
// values). A fold moves the columns into 32-bit totals before a byte can
// overflow (at most CH_ITEMS * DIMS additions per tile).
// Publishing: the small bins and the range statistics of a (workgroup, frame)
// pair go to a slot of their own with plain stores (k_comp_reduce combines
// them), because thousands of workgroups adding into the same few words
// serialise on the device's atomic unit; larger bins are spread over many
// words and are added with global atomics directly.
#define CS_PRIV_BINS 32u
#define CS_PRIV_ROWS (CS_PRIV_BINS / 4)   // four 8-bit counters per word
#define CS_GRID 2048u
#define CS_REDUCE_SLOTS 64u

struct CompSlot {
    uint32_t vh[CS_PRIV_BINS];     // values 0 .. CS_PRIV_BINS-1
    uint32_t th[CS_PRIV_BINS];     // widths 1 .. CS_PRIV_BINS
    uint64_t dmin[CH_MAXD], dmax[CH_MAXD];
    uint64_t bits;
    uint32_t max_tag, overflow;
    uint32_t pad[12];
};
static_assert(sizeof(CompSlot) == 384, "slot layout");

// Tiles of a component: frame f of component c holds ceil(n / CH_TILE) tiles
// of its own records, so a component with fewer records than points (diffs,
// keeps, choices) spreads only its own tiles over the workgroups.
// ctb[c * (F + 1) + f] = first tile of frame f, [c * (F + 1) + F] = total.
__global__ __launch_bounds__(CH_WG) void k_comp_tiles(const CompInst* __restrict__ ci, uint32_t F,
                                                      uint32_t* __restrict__ ctb) {
    __shared__ uint32_t s_tmp[4];
    const uint32_t c = blockIdx.x;
    uint32_t* out = ctb + (uint64_t)c * (F + 1);
    uint32_t run = 0;
    for (uint32_t f0 = 0; f0 < F; f0 += CH_WG) {
        const uint32_t f = f0 + threadIdx.x;
        const uint32_t v = f < F ? (uint32_t)((ci[(uint64_t)c * F + f].n + CH_TILE - 1) / CH_TILE) : 0u;
        uint32_t tot;
        const uint32_t pre = block_excl_scan(v, s_tmp, &tot);
        if (f < F) out[f] = run + pre;
        run += tot;
    }
    if (threadIdx.x == 0) out[F] = run;
}

// slot of the (workgroup b, frame f) pair: b for the frame of b's first tile,
// G + f for a frame that starts inside b's range (at most one workgroup).
// tile_begin is the component's tile prefix (k_comp_tiles).
template <typename ValT, int DIMS>
__global__ __launch_bounds__(CH_WG) void k_comp_stats(const ValT* __restrict__ vals, int c, uint32_t F,
                                                      const uint32_t* __restrict__ tile_begin,
                                                      const CompInst* __restrict__ ci, CompSlot* __restrict__ slots,
                                                      uint32_t* __restrict__ vhist, uint32_t* __restrict__ thist) {
    // tiles between folds: a thread adds at most CH_ITEMS * DIMS to a byte per tile
    constexpr uint32_t FOLD_TILES = 255u / (CH_ITEMS * DIMS);
    static_assert(FOLD_TILES >= 2, "two tiles are counted between folds");
    // byte values need 256 value bins
    constexpr uint32_t VH = sizeof(ValT) == 1 ? 256u : RANS_VALUE_LIMIT;
    // values of at most 32 bits are compared and measured in 32-bit registers
    using AccT = typename std::conditional<(sizeof(ValT) <= 4), uint32_t, uint64_t>::type;
    __shared__ uint32_t s_vh[VH], s_th[RANS_TAG_LIMIT];
    __shared__ uint32_t s_sv[CS_PRIV_BINS], s_st[CS_PRIV_BINS];
    __shared__ uint32_t s_pv[CS_PRIV_ROWS][CH_WG], s_pt[CS_PRIV_ROWS][CH_WG];
    __shared__ uint64_t s_red[4][2 * CH_MAXD + 2];
    const uint32_t tid = threadIdx.x;
    const uint32_t ntiles = tile_begin[F];
    const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    const uint32_t tb0 = blockIdx.x * per, tb1 = tb0 + per < ntiles ? tb0 + per : ntiles;
    if (tb0 >= tb1) return;
    for (uint32_t k = tid; k < VH; k += CH_WG) s_vh[k] = 0;
    if (tid < RANS_TAG_LIMIT) s_th[tid] = 0;
    if (tid < CS_PRIV_BINS) {
        s_sv[tid] = 0;
        s_st[tid] = 0;
    }
    for (uint32_t k = 0; k < CS_PRIV_ROWS; ++k) {
        s_pv[k][tid] = 0;
        s_pt[k][tid] = 0;
    }
    __syncthreads();
    const uint32_t f_first = upper_find(tile_begin, F, tb0);
    AccT mn[DIMS], mx[DIMS];
    uint32_t ovf = 0;
    auto reset = [&]() {
#pragma unroll
        for (int d = 0; d < DIMS; ++d) {
            mn[d] = (AccT)~0ull;
            mx[d] = 0;
        }
        ovf = 0;
    };
    reset();
    const uint32_t w = tid >> 6;
    // private columns -> 32-bit totals (whole workgroup; ends synchronised).
    // Thread (row, g) sums 8 words of its row, four bytes each; the 32
    // lanes of the row reduce them.
    auto fold = [&]() {
        __syncthreads();
        const uint32_t row = tid >> 5, g = tid & 31u;
        uint32_t vs[4] = {0, 0, 0, 0}, ts[4] = {0, 0, 0, 0};
#pragma unroll
        for (uint32_t i = 0; i < CH_WG / 32; ++i) {
            const uint32_t col = g + 32u * i;
            const uint32_t a = s_pv[row][col], b = s_pt[row][col];
            s_pv[row][col] = 0;
            s_pt[row][col] = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                vs[q] += (a >> (8 * q)) & 0xffu;
                ts[q] += (b >> (8 * q)) & 0xffu;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            for (int k = 16; k >= 1; k >>= 1) {
                vs[q] += __shfl_xor(vs[q], k, 64);
                ts[q] += __shfl_xor(ts[q], k, 64);
            }
        if (g == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                s_sv[4 * row + q] += vs[q];
                s_st[4 * row + q] += ts[q];
            }
        }
        __syncthreads();
    };
    // publish the statistics of frame f (called by the whole workgroup)
    auto flush = [&](uint32_t f) {
        fold();
#pragma unroll
        for (int d = 0; d < DIMS; ++d) {
            for (int k = 32; k >= 1; k >>= 1) {
                const AccT a = __shfl_xor(mn[d], k, 64), b = __shfl_xor(mx[d], k, 64);
                mn[d] = a < mn[d] ? a : mn[d];
                mx[d] = b > mx[d] ? b : mx[d];
            }
        }
        for (int k = 32; k >= 1; k >>= 1) ovf |= __shfl_xor(ovf, k, 64);
        if (lane_id() == 0) {
            for (int d = 0; d < DIMS; ++d) {
                // a wavefront without records (minimum above maximum) keeps the 64-bit identity (~0, 0)
                s_red[w][2 * d] = mn[d] > mx[d] ? ~0ull : (uint64_t)mn[d];
                s_red[w][2 * d + 1] = mx[d];
            }
            s_red[w][2 * CH_MAXD + 1] = ovf;
        }
        // Tagged bits and the widest width from the width histogram (widths
        // 1..CS_PRIV_BINS in s_st, wider ones in s_th): wavefront 0, lane = width
        if (tid < 64) {
            const uint32_t t = tid;
            const uint32_t cnt = t >= 1 && t <= CS_PRIV_BINS ? s_st[t - 1] : (t > CS_PRIV_BINS ? s_th[t] : 0u);
            uint64_t bsum = (uint64_t)t * cnt * DIMS;
            uint32_t tmax = cnt ? t : 0u;
            for (int k = 32; k >= 1; k >>= 1) {
                bsum += __shfl_xor(bsum, k, 64);
                const uint32_t o = __shfl_xor(tmax, k, 64);
                tmax = o > tmax ? o : tmax;
            }
            if (tid == 0) s_red[0][2 * CH_MAXD] = ((uint64_t)tmax << 56) | bsum;
        }
        __syncthreads();
        CompSlot& S = slots[f == f_first ? blockIdx.x : gridDim.x + f];
        if (tid < CS_PRIV_BINS) {
            S.vh[tid] = s_sv[tid];
            S.th[tid] = s_st[tid];
            s_sv[tid] = 0;
            s_st[tid] = 0;
        }
        if (tid == CS_PRIV_BINS) {
            uint32_t of = 0;
            for (int k = 0; k < 4; ++k) of |= (uint32_t)s_red[k][2 * CH_MAXD + 1];
            for (int d = 0; d < CH_MAXD; ++d) {
                uint64_t a = ~0ull, b = 0;
                if (d < DIMS)
                    for (int k = 0; k < 4; ++k) {
                        a = s_red[k][2 * d] < a ? s_red[k][2 * d] : a;
                        b = s_red[k][2 * d + 1] > b ? s_red[k][2 * d + 1] : b;
                    }
                S.dmin[d] = a;
                S.dmax[d] = b;
            }
            const uint64_t bt = s_red[0][2 * CH_MAXD];
            S.bits = bt & ((1ull << 56) - 1ull);
            S.max_tag = (uint32_t)(bt >> 56);
            S.overflow = of;
        }
        const uint32_t ci_idx = c * F + f;
        for (uint32_t k = CS_PRIV_BINS + tid; k < VH; k += CH_WG)
            if (s_vh[k]) {
                atomicAdd(&vhist[(uint64_t)ci_idx * RANS_VALUE_LIMIT + k], s_vh[k]);
                s_vh[k] = 0;
            }
        if (tid > CS_PRIV_BINS && tid < RANS_TAG_LIMIT && s_th[tid]) {
            atomicAdd(&thist[(uint64_t)ci_idx * RANS_TAG_LIMIT + tid], s_th[tid]);
            s_th[tid] = 0;
        }
        __syncthreads();
        reset();
    };
    // one record's values; the value histogram is skipped once the
    // wavefront has seen a value >= RANS_VALUE_LIMIT in this frame (the
    // component then has no value-rANS candidate and the histogram is unused)
    const uint32_t pt_base = tid;
    auto record = [&](const AccT (&v)[DIMS], bool vals_on) {
        AccT any = 0;   // bit length of the OR = widest bit length
#pragma unroll
        for (int d = 0; d < DIMS; ++d) {
            mn[d] = v[d] < mn[d] ? v[d] : mn[d];
            mx[d] = v[d] > mx[d] ? v[d] : mx[d];
            any |= v[d];
            if (vals_on) {
                if (v[d] < CS_PRIV_BINS)
                    atomicAdd(&s_pv[(uint32_t)v[d] >> 2][pt_base], 1u << (((uint32_t)v[d] & 3u) << 3));
                else if (v[d] < VH)
                    atomicAdd(&s_vh[(uint32_t)v[d]], 1u);
                else
                    ovf = 1u;
            }
        }
        const uint32_t wb = sizeof(AccT) == 4 ? 32u - (uint32_t)__clz((uint32_t)any) : (uint32_t)bitlen64(any);
        const uint32_t tag = wb > 1u ? wb : 1u;
        if (tag <= CS_PRIV_BINS)
            atomicAdd(&s_pt[(tag - 1) >> 2][pt_base], 1u << (((tag - 1) & 3u) << 3));
        else
            atomicAdd(&s_th[tag], 1u);
    };
    // a full tile of values of at most 32 bits: thread t reads the CH_ITEMS
    // consecutive records [t*CH_ITEMS, t*CH_ITEMS + CH_ITEMS) of the tile as
    // whole words (the statistics do not depend on the order)
    constexpr uint32_t BPT = CH_ITEMS * DIMS * sizeof(ValT);
    constexpr uint32_t NW = BPT / 4;
    constexpr bool WORDS = sizeof(ValT) <= 4 && BPT % 4 == 0 && NW <= 48;
    // loads the words of the thread's chunk of the full tile starting at record r
    auto load_chunk = [&](const ValT* src, uint64_t r, uint32_t (&wv)[NW + 1]) {
        const uint8_t* p = (const uint8_t*)(src + (r + (uint64_t)tid * CH_ITEMS) * DIMS);
        const uint32_t al = (uint32_t)(uintptr_t)p & 3u;   // 0 unless values are narrower than 4 bytes
        const uint32_t* pa = (const uint32_t*)(p - al);
#pragma unroll
        for (uint32_t i = 0; i < NW; ++i) wv[i] = pa[i];
        // the word past the chunk holds chunk bytes only when the chunk is unaligned
        wv[NW] = (sizeof(ValT) < 4 && al) ? pa[NW] : 0u;
        return al;
    };
    auto count_chunk = [&](uint32_t (&wv)[NW + 1], uint32_t al, bool vals_on) {
        if (sizeof(ValT) < 4) {
#pragma unroll
            for (uint32_t i = 0; i < NW; ++i) wv[i] = __builtin_amdgcn_alignbyte(wv[i + 1], wv[i], al);
        }
#pragma unroll
        for (uint32_t q = 0; q < CH_ITEMS; ++q) {
            AccT v[DIMS];
#pragma unroll
            for (int d = 0; d < DIMS; ++d) {
                const uint32_t e = q * DIMS + d;
                if (sizeof(ValT) == 1) v[d] = (wv[e >> 2] >> ((e & 3u) * 8u)) & 0xffu;
                else if (sizeof(ValT) == 2) v[d] = (wv[e >> 1] >> ((e & 1u) * 16u)) & 0xffffu;
                else v[d] = wv[e];
            }
            record(v, vals_on);
        }
    };
    uint32_t cur = 0xffffffffu, acc = 0;   // acc: tiles counted since the last fold
    for (uint32_t tile = tb0; tile < tb1;) {
        const uint32_t f = upper_find(tile_begin, F, tile);
        const CompInst& C = ci[c * F + f];
        const uint64_t r0 = (uint64_t)(tile - tile_begin[f]) * CH_TILE;
        const uint64_t n = C.n;
        // two full tiles of one frame are loaded together (one-dimensional
        // components: wider ones would pass 128 registers, fewer waves per SIMD)
        const bool pair = WORDS && DIMS == 1 && tile + 1 < tb1 && r0 + 2ull * CH_TILE <= n;
        const uint32_t g = pair ? 2u : 1u;
        // every frame the range touches is published, even with no records
        if (f != cur) {
            if (cur != 0xffffffffu) flush(cur);
            cur = f;
            acc = 0;
        } else if (acc + g > FOLD_TILES) {
            fold();
            acc = 0;
        }
        acc += g;
        tile += g;
        if (r0 >= n) continue;
        const ValT* src = vals + C.base * DIMS;
        const bool vals_on = __ballot(ovf != 0) == 0;
        if (pair) {
            uint32_t w0[NW + 1], w1[NW + 1];
            const uint32_t a0 = load_chunk(src, r0, w0);
            const uint32_t a1 = load_chunk(src, r0 + CH_TILE, w1);
            count_chunk(w0, a0, vals_on);
            count_chunk(w1, a1, vals_on);
            continue;
        }
        if (WORDS && r0 + CH_TILE <= n) {
            uint32_t w0[NW + 1];
            const uint32_t a0 = load_chunk(src, r0, w0);
            count_chunk(w0, a0, vals_on);
            continue;
        }
        // striped: item j of thread t is record j*256 + t of the tile;
        // every load of the tile is issued before the first is used
        ValT x[CH_ITEMS][DIMS];
#pragma unroll
        for (int j = 0; j < CH_ITEMS; ++j) {
            const uint64_t r = r0 + (uint64_t)j * CH_WG + tid;
#pragma unroll
            for (int d = 0; d < DIMS; ++d) x[j][d] = src[(r < n ? r : r0) * DIMS + d];   // unconditional
        }
#pragma unroll
        for (int j = 0; j < CH_ITEMS; ++j) {
            const uint64_t r = r0 + (uint64_t)j * CH_WG + tid;
            if (r >= n) continue;
            AccT v[DIMS];
#pragma unroll
            for (int d = 0; d < DIMS; ++d) v[d] = (AccT)x[j][d];
            record(v, vals_on);
        }
    }
    if (cur != 0xffffffffu) flush(cur);
}

// Combines the slots of every frame of every component: block (f, y, c) takes
// up to CS_REDUCE_SLOTS slots of frame f of component c and adds them into
// that instance's statistics. Frame f's slots are G + f when it starts inside a workgroup's
// range, then the workgroups whose first tile lies in f.
__global__ __launch_bounds__(CH_WG) void k_comp_reduce(const CompSlot* __restrict__ all_slots, uint32_t G,
                                                       const uint32_t* __restrict__ ctb,
                                                       uint32_t F, CompInst* __restrict__ ci,
                                                       uint32_t* __restrict__ vhist, uint32_t* __restrict__ thist) {
    __shared__ uint32_t s_sum[4][2 * CS_PRIV_BINS];
    const int c = (int)blockIdx.z;   // component; its slots follow the previous components'
    const CompSlot* __restrict__ slots = all_slots + (uint64_t)c * (G + F);
    const uint32_t* __restrict__ tile_begin = ctb + (uint64_t)c * (F + 1);   // the component's tiles
    const uint32_t ntiles = tile_begin[F];
    const uint32_t f = blockIdx.x;
    const uint32_t t0 = tile_begin[f], t1 = tile_begin[f + 1];
    if (t0 == t1) return;
    const uint32_t per = (ntiles + G - 1) / G;
    const uint32_t b_lo = (t0 + per - 1) / per;
    uint32_t b_hi = (t1 + per - 1) / per;
    b_hi = b_hi < G ? b_hi : G;
    const uint32_t mid = t0 % per != 0 ? 1u : 0u;
    const uint32_t cnt = (b_hi > b_lo ? b_hi - b_lo : 0u) + mid;
    const uint32_t i0 = blockIdx.y * CS_REDUCE_SLOTS;
    if (i0 >= cnt) return;
    const uint32_t i1 = i0 + CS_REDUCE_SLOTS < cnt ? i0 + CS_REDUCE_SLOTS : cnt;
    auto slot_of = [&](uint32_t i) -> const CompSlot& {
        return slots[mid ? (i == 0 ? G + f : b_lo + i - 1) : b_lo + i];
    };
    const uint32_t tid = threadIdx.x, word = tid & 63u, grp = tid >> 6;
    // bins: thread (grp, word) sums word `word` of slots i0 + grp, + 4, ...
    uint32_t acc = 0;
#pragma unroll
    for (uint32_t k = 0; k < CS_REDUCE_SLOTS / 4; ++k) {
        const uint32_t i = i0 + grp + 4 * k;
        if (i < i1) {
            const CompSlot& S = slot_of(i);
            acc += word < CS_PRIV_BINS ? S.vh[word] : S.th[word - CS_PRIV_BINS];
        }
    }
    s_sum[grp][word] = acc;
    // range statistics: lane l of wavefront 0 takes slot i0 + l
    if (grp == 0) {
        const uint32_t i = i0 + word;
        const bool live = i < i1;
        uint64_t mn[CH_MAXD], mx[CH_MAXD], bits = 0;
        uint32_t mt = 0, of = 0;
        for (int d = 0; d < CH_MAXD; ++d) {
            mn[d] = ~0ull;
            mx[d] = 0;
        }
        if (live) {
            const CompSlot& S = slot_of(i);
            for (int d = 0; d < CH_MAXD; ++d) {
                mn[d] = S.dmin[d];
                mx[d] = S.dmax[d];
            }
            bits = S.bits;
            mt = S.max_tag;
            of = S.overflow;
        }
        for (int k = 32; k >= 1; k >>= 1) {
            for (int d = 0; d < CH_MAXD; ++d) {
                const uint64_t a = __shfl_xor(mn[d], k, 64), b = __shfl_xor(mx[d], k, 64);
                mn[d] = a < mn[d] ? a : mn[d];
                mx[d] = b > mx[d] ? b : mx[d];
            }
            bits += __shfl_xor(bits, k, 64);
            const uint32_t a = __shfl_xor(mt, k, 64);
            mt = a > mt ? a : mt;
            of |= __shfl_xor(of, k, 64);
        }
        if (word == 0) {
            CompInst& C = ci[c * F + f];
            for (uint32_t d = 0; d < C.dims; ++d) {
                if (mn[d] != ~0ull) atomicMin((unsigned long long*)&C.dmin[d], mn[d]);
                if (mx[d]) atomicMax((unsigned long long*)&C.dmax[d], mx[d]);
            }
            if (bits) atomicAdd((unsigned long long*)&C.tag_bits, bits);
            if (mt) atomicMax(&C.max_tag, mt);
            if (of) atomicOr(&C.overflow, 1u);
        }
    }
    __syncthreads();
    if (tid < 2 * CS_PRIV_BINS) {
        const uint32_t v = s_sum[0][tid] + s_sum[1][tid] + s_sum[2][tid] + s_sum[3][tid];
        const uint64_t ci_idx = (uint64_t)c * F + f;
        if (v) {
            if (tid < CS_PRIV_BINS) atomicAdd(&vhist[ci_idx * RANS_VALUE_LIMIT + tid], v);
            else atomicAdd(&thist[ci_idx * RANS_TAG_LIMIT + (tid - CS_PRIV_BINS) + 1], v);
        }
    }
}

// exclusive max over the lower threads of a 256-thread workgroup (0 for
// thread 0); s4 holds 4 entries
__device__ inline uint32_t block_excl_max(uint32_t v, uint32_t* s4) {
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t inc = wave_incl_dpp(v, 0u, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
    if (lane == 63) s4[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t k = 0; k < w; ++k) before = s4[k] > before ? s4[k] : before;
    const uint32_t ex = dpp_u32<0x138, 0xf>(0u, inc);   // wave_shr:1 (lane 0: 0)
    __syncthreads();
    return ex > before ? ex : before;
}

// exclusive min over the higher threads of a 256-thread workgroup (ident
// for the last thread); s4 holds 4 entries
__device__ inline uint32_t block_excl_suffix_min(uint32_t v, uint32_t* s4, uint32_t ident) {
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    uint32_t inc = v;
#pragma unroll
    for (int k = 1; k < 64; k <<= 1) {
        const uint32_t t = __shfl_down(inc, k, 64);
        if (lane + (uint32_t)k < 64) inc = t < inc ? t : inc;
    }
    if (lane == 0) s4[w] = inc;
    __syncthreads();
    uint32_t after = ident;
    for (uint32_t k = w + 1; k < CH_WG / 64; ++k) after = s4[k] < after ? s4[k] : after;
    uint32_t ex = __shfl_down(inc, 1, 64);
    if (lane == 63) ex = ident;
    __syncthreads();
    return ex < after ? ex : after;
}

// ------------------------------------------------------------------------
// RVN1 table of every section, built on the device (cpp/chain/lcps_cpp.cpp:560-729)

__global__ __launch_bounds__(CH_WG) void k_rans_table(const CompInst* __restrict__ ci, Section* __restrict__ sec,
                                                      const uint32_t* __restrict__ vhist,
                                                      const uint32_t* __restrict__ thist, uint32_t F,
                                                      uint64_t* __restrict__ pc_out,
                                                      uint8_t* __restrict__ tbytes, uint32_t* err) {
    const uint32_t si = blockIdx.x;
    const uint32_t cidx = si >> 1, s = si & 1;
    const CompInst& C = ci[cidx];
    const uint64_t count = s == 0 ? C.n * C.dims : C.n;
    const bool eligible = s == 1 || C.overflow == 0;
    __shared__ uint32_t s_prob[RANS_VALUE_LIMIT];
    __shared__ uint64_t s_key[RANS_VALUE_LIMIT];
    __shared__ uint32_t s_tmp[4];
    __shared__ uint64_t s_red[4];
    __shared__ uint32_t s_A, s_unique;
    Section& S = sec[si];
    if (!eligible || count == 0) {
        if (threadIdx.x == 0) {
            S.count = count; S.eligible = eligible; S.pb = 0; S.alphabet = 0; S.table_len = 0; S.est_bits = 0;
        }
        return;
    }
    const uint32_t HB = s == 0 ? RANS_VALUE_LIMIT : RANS_TAG_LIMIT;
    const uint32_t* hist = s == 0 ? vhist + (uint64_t)cidx * RANS_VALUE_LIMIT : thist + (uint64_t)cidx * RANS_TAG_LIMIT;
    if (threadIdx.x == 0) { s_A = 0; s_unique = 0; }
    __syncthreads();
    uint32_t amax = 0, uq = 0;
    for (uint32_t b = threadIdx.x; b < HB; b += CH_WG)
        if (hist[b]) { amax = b + 1 > amax ? b + 1 : amax; ++uq; }
    atomicMax(&s_A, amax);
    atomicAdd(&s_unique, uq);
    __syncthreads();
    const uint32_t A = s_A;
    const uint32_t pb = (uint32_t)rvn1_precision_bits(s_unique);
    const uint64_t M = 1ull << pb;
    uint64_t sum = 0;
    for (uint32_t b = threadIdx.x; b < A; b += CH_WG) {
        const uint64_t f = hist[b];
        uint32_t pr = rvn1_first_prob(f, pb, count);
        if (pr == 0 && f != 0) pr = 1;
        s_prob[b] = pr;
        sum += pr;
    }
    for (int k = 32; k >= 1; k >>= 1) sum += __shfl_xor(sum, k, 64);
    if (lane_id() == 0) s_red[threadIdx.x >> 6] = sum;
    __syncthreads();
    sum = s_red[0] + s_red[1] + s_red[2] + s_red[3];
    if (sum != M) {
        // std::stable_sort by probability == sort of (probability << 12 | index)
        uint32_t n2 = 1;
        while (n2 < A) n2 <<= 1;
        for (uint32_t b = threadIdx.x; b < n2; b += CH_WG) s_key[b] = b < A ? (((uint64_t)s_prob[b] << 12) | b) : ~0ull;
        __syncthreads();
        for (uint32_t k = 2; k <= n2; k <<= 1)
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < n2; i += CH_WG) {
                    const uint32_t l = i ^ j;
                    if (l > i) {
                        const uint64_t a = s_key[i], b = s_key[l];
                        const bool up = (i & k) == 0;
                        if ((a > b) == up) { s_key[i] = b; s_key[l] = a; }
                    }
                }
                __syncthreads();
            }
        if (sum < M) {
            if (threadIdx.x == 0) s_prob[s_key[A - 1] & 0xfffu] += (uint32_t)(M - sum);
        } else {
            // Passes over the symbols from the most probable down
            // (lcps_cpp.cpp:626-672), position q = A - 1 - j of the sorted
            // order. A pass stops at its first symbol of probability <= 1 and
            // once the excess is gone. Inside a pass a symbol's reduction
            // depends on its own probability and the sum at the pass's start
            // only, so the pass is a block-wide prefix sum: symbol q gives up
            // min(reduction_q, excess left after the symbols before it).
            // Sums stay below 2^32: M <= 2^20 and A <= 4096.
            __shared__ uint32_t s_qstop;
            uint32_t cur = (uint32_t)sum, error = (uint32_t)(sum - M);
            const uint32_t q0 = threadIdx.x * 16;
            while (error > 0) {
                if (threadIdx.x == 0) s_qstop = A - 1;   // positions 0 .. A-2; A-1 = no stop
                __syncthreads();
                uint32_t pq[16], qmin = A - 1;
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const uint32_t q = q0 + k;
                    pq[k] = q + 1 < A ? s_prob[s_key[A - 1 - q] & 0xfffu] : 0u;
                    if (q + 1 < A && pq[k] <= 1 && q < qmin) qmin = q;
                }
                if (qmin < A - 1) atomicMin(&s_qstop, qmin);
                __syncthreads();
                const uint32_t qstop = s_qstop;
                if (qstop == 0) {   // the most probable symbol cannot give up any
                    if (threadIdx.x == 0) atomicOr(err, 4u);
                    break;
                }
                uint32_t red[16], local = 0;
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const uint32_t q = q0 + k;
                    red[k] = 0;
                    if (q < qstop) {   // pq >= 2: the reduction lies in [1, pq - 1]
                        const uint64_t np = rvn1_reduced_prob((int)pb, cur, pq[k]);
                        int64_t r = (int64_t)pq[k] - (int64_t)np;
                        if (r == 0) r = 1;
                        if (r >= (int64_t)pq[k]) r = (int64_t)pq[k] - 1;
                        red[k] = (uint32_t)r;
                        local += red[k];
                    }
                }
                uint32_t tot;
                uint32_t before = block_excl_scan(local, s_tmp, &tot);
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    if (!red[k]) continue;
                    if (before < error) {
                        const uint32_t a = red[k] < error - before ? red[k] : error - before;
                        s_prob[s_key[A - 1 - (q0 + k)] & 0xfffu] -= a;
                    }
                    before += red[k];
                }
                const uint32_t applied = tot < error ? tot : error;
                error -= applied;
                cur -= applied;
                __syncthreads();   // this pass's probabilities before the next pass reads them
            }
        }
        __syncthreads();
    }
    // cumulative frequencies (exclusive scan, 16 entries per thread)
    {
        const uint32_t b0 = threadIdx.x * 16;
        uint32_t local = 0;
        for (uint32_t k = 0; k < 16; ++k) local += (b0 + k < A) ? s_prob[b0 + k] : 0u;
        uint32_t tot;
        uint32_t run = block_excl_scan(local, s_tmp, &tot);
        for (uint32_t k = 0; k < 16 && b0 + k < A; ++k) {
            pc_out[(uint64_t)si * RANS_VALUE_LIMIT + b0 + k] = ((uint64_t)run << 32) | s_prob[b0 + k];
            run += s_prob[b0 + k];
        }
        if (threadIdx.x == 0 && tot != M) atomicOr(err, 4u);
        // estimated stream bits: sum of count * (pb - log2 f) in 1/65536
        uint64_t eb = 0;
        for (uint32_t k = 0; k < 16 && b0 + k < A; ++k) {
            const uint32_t h = hist[b0 + k];
            if (h) eb += (uint64_t)h * rvc1_symbol_cost_q16(s_prob[b0 + k], pb);
        }
        for (int k = 32; k >= 1; k >>= 1) eb += __shfl_xor(eb, k, 64);
        __syncthreads();
        if (lane_id() == 0) s_red[threadIdx.x >> 6] = eb;
        __syncthreads();
        if (threadIdx.x == 0) S.est_bits = s_red[0] + s_red[1] + s_red[2] + s_red[3];
    }
    // compact table bytes, cpp/chain/lcps_cpp.cpp:686-729: varuint(A), then
    // per symbol (pr << 2) | e and e more bytes (e = 0, 1, 2 for pr < 2^6,
    // < 2^14, else), and a zero run as one byte ((n - 1) << 2) | 3 per 64
    // zeros. Every thread writes its own 16 symbols: a zero starts a byte
    // when it lies a multiple of 64 after its run's start (exclusive max-scan
    // of "last nonzero + 1"), whose count is bounded by the run's end
    // (exclusive suffix min-scan of nonzero positions); the byte offsets are
    // an exclusive sum over the symbols.
    {
        __shared__ uint32_t s_t2[4];
        uint8_t* o = tbytes + (uint64_t)si * TABLE_BYTES_MAX;
        uint32_t hl = 0;
        for (uint64_t v = A; v >= 0x80u; v >>= 7) ++hl;
        ++hl;   // varuint(A) bytes
        const uint32_t b0 = threadIdx.x * 16;
        uint32_t pr[16];
        uint32_t lmax = 0, lmin = A;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t i = b0 + k;
            pr[k] = i < A ? s_prob[i] : 0u;
            if (i < A && pr[k]) {
                lmax = i + 1;                  // entries ascend: the last wins
                lmin = lmin < i ? lmin : i;
            }
        }
        const uint32_t rs0 = block_excl_max(lmax, s_t2);        // last nonzero + 1 before b0 (or 0)
        const uint32_t re0 = block_excl_suffix_min(lmin, s_t2, A);   // first nonzero after b0 + 15 (or A)
        uint32_t re[16];
        {
            uint32_t rn = re0;
#pragma unroll
            for (int k = 15; k >= 0; --k) {
                re[k] = rn;
                const uint32_t i = b0 + k;
                if (i < A && pr[k]) rn = i;
            }
        }
        uint32_t nbytes[16], cnt = 0;
        {
            uint32_t rm = rs0;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const uint32_t i = b0 + k;
                uint32_t nb = 0;
                if (i < A) {
                    if (pr[k]) {
                        nb = 1u + (pr[k] >= (1u << 6) ? 1u : 0u) + (pr[k] >= (1u << 14) ? 1u : 0u);
                        rm = i + 1;
                    } else {
                        nb = ((i - rm) & 63u) == 0 ? 1u : 0u;
                    }
                }
                nbytes[k] = nb;
                cnt += nb;
            }
        }
        uint32_t total;
        uint32_t at = hl + block_excl_scan(cnt, s_tmp, &total);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (!nbytes[k]) continue;
            const uint32_t i = b0 + k;
            if (pr[k]) {
                const uint32_t eb = nbytes[k] - 1;
                o[at] = (uint8_t)((pr[k] << 2) | eb);
                for (uint32_t q = 0; q < eb; ++q) o[at + 1 + q] = (uint8_t)(pr[k] >> (8 * (q + 1) - 2));
            } else {
                const uint32_t run = re[k] - i;   // zeros from i to the run's end
                const uint32_t extra = run - 1 < 63u ? run - 1 : 63u;
                o[at] = (uint8_t)((extra << 2) | 3u);
            }
            at += nbytes[k];
        }
        if (threadIdx.x == 0) {
            uint32_t len = 0;
            uint64_t v = A;
            while (v >= 0x80u) { o[len++] = (uint8_t)((uint8_t)v | 0x80u); v >>= 7; }
            o[len++] = (uint8_t)v;
            S.count = count; S.eligible = 1; S.pb = pb; S.alphabet = A; S.table_len = hl + total;
        }
    }
}

// chunk plan of every section (one thread): lanes, chunk count, chunk bases
// Chunk layout of every section: one workgroup scans the sections' chunk
// counts in order (component c's sections start at 2cF).
__global__ __launch_bounds__(CH_WG) void k_rans_plan(Section* sec, uint32_t nsec, uint32_t F, uint32_t lanes_max,
                                                     uint32_t chunk_symbols,
                                                     uint32_t* comp_chunk /* [6]: begin of component c, [5] = total */) {
    __shared__ uint32_t s_tmp[4];
    uint32_t run = 0;
    for (uint32_t s0 = 0; s0 < nsec; s0 += CH_WG) {
        const uint32_t si = s0 + threadIdx.x;
        uint32_t nch = 0;
        if (si < nsec) {
            Section& S = sec[si];
            S.chunk_symbols = chunk_symbols;
            const bool used = S.eligible && S.count > 0 && S.encode;
            S.lanes = used ? rvc1_lanes_for(S.count, lanes_max) : 0;
            nch = used ? (uint32_t)((S.count + chunk_symbols - 1) / chunk_symbols) : 0;
            S.nchunks = nch;
        }
        uint32_t tot;
        const uint32_t pre = block_excl_scan(nch, s_tmp, &tot);
        if (si < nsec) {
            sec[si].chunk_base = run + pre;
            if (si % (2 * F) == 0) comp_chunk[si / (2 * F)] = run + pre;
        }
        run += tot;
    }
    if (threadIdx.x == 0) comp_chunk[5] = run;
}

__device__ inline uint32_t find_section(const Section* sec, uint32_t lo, uint32_t hi, uint32_t chunk) {
    // largest si in [lo, hi) with chunk_base <= chunk
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sec[mid].chunk_base <= chunk) lo = mid; else hi = mid;
    }
    return lo;
}

// rANS encoding, one wavefront per chunk; lanes beyond the section's lane
// count idle. Symbols are loaded one group of ENC_GROUP steps ahead of the
// group being coded. For alphabets <= 256 the table lives in LDS together
// with an exact reciprocal of every frequency, so the state update needs no
// integer division: with l = ceil(log2 f) and m = ceil(2^(31+l) / f) < 2^32
// (rcp31_* in chain_int.hpp), floor(x / f) = mulhi(x, m) >> (l - 1) for
// l >= 1. The update x' = (x / f) * M + x % f + cum is taken as
// x + (x / f) * (M - f) + cum, where x / f < 1024 and M - f < 2^20 make the
// product a 24-bit multiply. For f = 1 the entry holds m = 2^32 - 1 and no
// shift, so the quotient comes out as x - 1 (x >= 1 after renormalisation),
// and cum + M - 1 in place of cum: x + (x - 1) * (M - 1) + cum + M - 1 = x * M + cum.
#define ENC_GROUP 4
#define ENC_XMAX_IDLE (1u << 30)   // above every state: an idle lane emits nothing

struct EncLds {
    // per symbol: {1024 * f (bytes are emitted while x >= this), cum, m,
    //              (M - f) | (l - 1) << 24}, f = 1 as described above;
    // entry 256 is an idle step: no byte, and the update leaves x as it is
    uint4 tab[257];
};
#define ENC_IDLE_SYM 256u

// Raw input of steps t-1, ..., t-g of this lane (steps are coded
// backwards): the value, or for a tags section the DIMS values of the record.
// Loads are unconditional (a missing step reads the chunk's first symbol) and
// their results are not touched here, so they stay in flight while the group
// before is coded; bit j of *valid marks step t-1-j as present.
template <typename ValT, int RD>
__device__ inline void enc_group_load(const ValT* __restrict__ base, uint32_t Sc, uint32_t L, uint32_t lane,
                                      uint32_t t, uint32_t g, ValT (&raw)[ENC_GROUP][RD], uint32_t& valid) {
    valid = 0;
#pragma unroll
    for (int j = 0; j < ENC_GROUP; ++j) {
        const uint32_t k = (t - 1 - j) * L + lane;
        const bool v = (uint32_t)j < g && lane < L && k < Sc;
        valid |= v ? 1u << j : 0u;
        const uint32_t q = v ? k : 0u;
#pragma unroll
        for (int d = 0; d < RD; ++d) raw[j][d] = base[q * RD + d];
    }
}

// symbols of a loaded group; a missing step is ENC_IDLE_SYM (LDS table) or
// 0xffffffff (global table). A record's tag is
// max(1, bit length of its widest value) = max(1, bit length of the OR of its
// values).
template <typename ValT, int RD, bool TAGS, bool SMALL>
__device__ inline void enc_group_sym(const ValT (&raw)[ENC_GROUP][RD], uint32_t valid, uint32_t* sym) {
#pragma unroll
    for (int j = 0; j < ENC_GROUP; ++j) {
        uint32_t v;
        if (!TAGS) {
            v = (uint32_t)raw[j][0];
        } else {
            ValT o = raw[j][0];
#pragma unroll
            for (int d = 1; d < RD; ++d) o |= raw[j][d];
            const uint32_t b = sizeof(ValT) == 8 ? (uint32_t)bitlen64((uint64_t)o)
                                                 : 32u - (uint32_t)__clz((uint32_t)o);
            v = b > 1u ? b : 1u;
        }
        sym[j] = (valid >> j) & 1u ? v : (SMALL ? ENC_IDLE_SYM : 0xffffffffu);
    }
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// One wavefront codes one chunk, lane l holding the state of lane l. The
// renormalisation bytes of a step are placed plane by plane (every emitting
// lane's first byte in lane order, then the second bytes, then the third) at
// decreasing offsets. They are staged in an LDS ring of ENC_RING bytes and
// written out in aligned blocks of ENC_FLUSH bytes (16 per lane), so the
// global stores are wide and few; what is left at the end is copied byte by
// byte. Only the lanes with a byte for a plane write it.
#define ENC_RING 2048u
#define ENC_FLUSH 1024u
template <typename ValT, int DIMS, bool SMALL, bool TAGS>
__device__ inline void enc_chunk(const ValT* __restrict__ vals, uint32_t chunk, const CompInst& C,
                                 const Section& S, const uint64_t* __restrict__ gpc, EncLds& lds,
                                 uint8_t* __restrict__ ring, uint8_t* __restrict__ scratch, uint64_t stride,
                                 uint32_t* __restrict__ chunk_len, uint32_t* __restrict__ chunk_pos) {
    constexpr int RD = TAGS ? DIMS : 1;   // values per symbol
    const uint32_t lane = threadIdx.x, L = S.lanes, pb = S.pb;
    const uint32_t M = 1u << pb, lower = 4u * M;
    if (SMALL)
        for (uint32_t b = lane; b < S.alphabet; b += 64) {
            const uint64_t e = gpc[b];
            const uint32_t f = (uint32_t)e;
            const uint32_t l = f > 1 ? rcp31_shift(f) : 1u;
            const uint32_t m = f > 1 ? (uint32_t)rcp31_mult(f, l) : (f == 1 ? 0xffffffffu : 0u);
            const uint32_t cum = (uint32_t)(e >> 32) + (f == 1 ? M - 1 : 0u);
            lds.tab[b] = make_uint4(1024u * f, cum, m, (M - f) | ((l - 1) << 24));
        }
    if (SMALL && lane == 0) lds.tab[ENC_IDLE_SYM] = make_uint4(ENC_XMAX_IDLE, 0u, 0u, 0u);
    __syncthreads();
    const uint64_t first = (uint64_t)(chunk - S.chunk_base) * S.chunk_symbols;
    const uint32_t Sc = (uint32_t)(S.count - first < S.chunk_symbols ? S.count - first : S.chunk_symbols);
    const uint32_t T = (Sc + L - 1) / L;
    uint8_t* out = scratch + (uint64_t)chunk * stride;
    uint8_t* stream = out + CHUNK_STATE_BYTES;
    // stream offsets: [pos, end) written so far, [flushed, end) already global
    const uint32_t end = (uint32_t)(stride - CHUNK_STATE_BYTES);
    uint32_t pos = end, flushed = end;
    uint32_t x = lower;
    // the chunk's symbols: values from `first`, or records from `first` for a tags section
    const ValT* base = vals + C.base * DIMS + first * RD;
    uint32_t cur[ENC_GROUP];
    ValT raw[ENC_GROUP][RD];
    uint32_t valid;
    uint32_t t = T;
    uint32_t g = t < ENC_GROUP ? t : ENC_GROUP;
    enc_group_load<ValT, RD>(base, Sc, L, lane, t, g, raw, valid);
    enc_group_sym<ValT, RD, TAGS, SMALL>(raw, valid, cur);
    while (t > 0) {
        // table entries of this group first: a wait for them is then no
        // wait for the next group's loads, issued after them
        uint4 te[ENC_GROUP];       // alphabet <= 256: LDS entries
        uint64_t ent[ENC_GROUP];   // larger alphabets: (cum << 32) | f
#pragma unroll
        for (int j = 0; j < ENC_GROUP; ++j) {
            if (SMALL) {
                te[j] = lds.tab[cur[j]];   // an idle step reads the idle entry
            } else {
                ent[j] = gpc[cur[j] != 0xffffffffu ? cur[j] : 0u];
            }
        }
        const uint32_t tn = t - g;
        const uint32_t gn = tn < ENC_GROUP ? tn : ENC_GROUP;
        if (tn > 0) enc_group_load<ValT, RD>(base, Sc, L, lane, tn, gn, raw, valid);
#pragma unroll
        for (int j = 0; j < ENC_GROUP; ++j) {
            if ((uint32_t)j >= g) break;
            const bool active = SMALL || cur[j] != 0xffffffffu;
            const uint32_t f = SMALL ? 0u : (uint32_t)ent[j];
            // bytes to emit: x < 2^30 <= 256^3 * 1024 * f, so at most three
            const uint32_t xmax = SMALL ? te[j].x : (active ? 1024u * f : ENC_XMAX_IDLE);
            // byte i is emitted when x >> 8i >= xmax; the three compares are
            // also the lane masks of the three planes
            const bool e0 = x >= xmax, e1 = (x >> 8) >= xmax, e2 = (x >> 16) >= xmax;
            const uint32_t kb = (uint32_t)e0 + (uint32_t)e1 + (uint32_t)e2;
            const uint32_t xx = x >> (8 * kb);
            const uint64_t m0 = __ballot(e0), m1 = __ballot(e1), m2 = __ballot(e2);
            const uint32_t c0 = __popcll(m0), c1 = __popcll(m1), c2 = __popcll(m2);
            pos -= c0 + c1 + c2;
            // planes 1 and 2 are rare: their lanes (and whole blocks) are skipped
            if (e0) ring[(pos + mbcnt64(m0)) & (ENC_RING - 1)] = (uint8_t)(x >> (8 * (kb - 1)));
            if (e1) ring[(pos + c0 + mbcnt64(m1)) & (ENC_RING - 1)] = (uint8_t)(x >> (8 * (kb - 2)));
            if (e2) ring[(pos + c0 + c1 + mbcnt64(m2)) & (ENC_RING - 1)] = (uint8_t)x;
            uint32_t xn;
            if (SMALL) {
                const uint32_t w = te[j].w;
                const uint32_t q = __umulhi(xx, te[j].z) >> ((w >> 24) & 31u);
                xn = xx + __umul24(q, w & 0xffffffu) + te[j].y;
            } else {
                const uint32_t q = xx / (active ? f : 1u);
                xn = q * M + (xx - q * f) + (uint32_t)(ent[j] >> 32);
            }
            x = active ? xn : x;
        }
        // a group adds at most ENC_GROUP * 192 bytes: one flush keeps the
        // unflushed bytes below ENC_FLUSH + 768 < ENC_RING
        if (flushed - pos >= ENC_FLUSH) {
            __syncthreads();
            const uint32_t o = flushed - ENC_FLUSH + 16u * lane;
            const uint4 v = *reinterpret_cast<const uint4*>(ring + (o & (ENC_RING - 1)));
            *reinterpret_cast<uint4*>(stream + o) = v;
            flushed -= ENC_FLUSH;
            __syncthreads();
        }
        if (tn > 0) enc_group_sym<ValT, RD, TAGS, SMALL>(raw, valid, cur);
        t = tn;
        g = gn;
    }
    __syncthreads();
    for (uint32_t o = pos + lane; o < flushed; o += 64) stream[o] = ring[o & (ENC_RING - 1)];
    if (lane < L) put_u32(out + 4 * lane, x);
    if (lane == 0) {
        chunk_len[chunk] = end - pos;
        chunk_pos[chunk] = pos;
    }
}

template <typename ValT, int DIMS>
__device__ inline void enc_chunk_any(const ValT* __restrict__ vals, uint32_t chunk, uint32_t si, const CompInst& C,
                                     const Section& S, const uint64_t* __restrict__ gpc, EncLds& lds,
                                     uint8_t* __restrict__ ring, uint8_t* __restrict__ scratch, uint64_t stride,
                                     uint32_t* __restrict__ chunk_len, uint32_t* __restrict__ chunk_pos) {
    // section 2i holds the values of instance i, section 2i + 1 its record tags
    const bool small = S.alphabet <= 256;
    if (si & 1) {
        if (small)
            enc_chunk<ValT, DIMS, true, true>(vals, chunk, C, S, gpc, lds, ring, scratch, stride, chunk_len, chunk_pos);
        else
            enc_chunk<ValT, DIMS, false, true>(vals, chunk, C, S, gpc, lds, ring, scratch, stride, chunk_len, chunk_pos);
    } else {
        if (small)
            enc_chunk<ValT, DIMS, true, false>(vals, chunk, C, S, gpc, lds, ring, scratch, stride, chunk_len, chunk_pos);
        else
            enc_chunk<ValT, DIMS, false, false>(vals, chunk, C, S, gpc, lds, ring, scratch, stride, chunk_len, chunk_pos);
    }
}

// every component's chunks in one launch: block b codes chunk b of the
// batch (the chunks of component c are [comp_chunk[c], comp_chunk[c+1]))
#define ENC_WGS 8192u    // rANS coding workgroups (one wavefront each) at most
#define COPY_WGS 4096u   // chunk copy workgroups at most
// Occupancy: held to 96 VGPRs for five single-wavefront workgroups per SIMD
// (113 unconstrained; 88 and no spill under the bound).
template <int D, typename RelT, typename DKT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_rans_encode(const uint32_t* __restrict__ idrel,
                                                    const DKT* __restrict__ diff,
                                                    const DKT* __restrict__ keep,
                                                    const uint8_t* __restrict__ choice2,
                                                    const RelT* __restrict__ rel, uint32_t F,
                                                    const CompInst* __restrict__ ci, const Section* __restrict__ sec,
                                                    const uint32_t* __restrict__ comp_chunk,
                                                    const uint64_t* __restrict__ pc_tab, uint8_t* __restrict__ scratch,
                                                    uint64_t stride, uint32_t* __restrict__ chunk_len,
                                                    uint32_t* __restrict__ chunk_pos) {
    __shared__ EncLds lds;
    __shared__ __attribute__((aligned(16))) uint8_t ring[ENC_RING];
    // the grid is bounded by the most chunks the batch can have; workgroups
    // stride over the chunks the plan made (known on the device only)
    const uint32_t nchunks = comp_chunk[5];
    for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    int c = 0;
    while (c < 4 && chunk >= comp_chunk[c + 1]) ++c;
    const uint32_t si = find_section(sec, 2 * F * c, 2 * F * (c + 1), chunk);
    const Section& S = sec[si];
    const CompInst& C = ci[si >> 1];
    const uint64_t* gpc = pc_tab + (uint64_t)si * RANS_VALUE_LIMIT;
    switch (c) {
    case 0: enc_chunk_any<uint32_t, 1>(idrel, chunk, si, C, S, gpc, lds, ring, scratch, stride, chunk_len, chunk_pos); break;
    case 1: enc_chunk_any<DKT, D>(diff, chunk, si, C, S, gpc, lds, ring, scratch, stride, chunk_len, chunk_pos); break;
    case 2: enc_chunk_any<DKT, D>(keep, chunk, si, C, S, gpc, lds, ring, scratch, stride, chunk_len, chunk_pos); break;
    case 3: enc_chunk_any<uint8_t, 1>(choice2, chunk, si, C, S, gpc, lds, ring, scratch, stride, chunk_len, chunk_pos); break;
    default: enc_chunk_any<RelT, D>(rel, chunk, si, C, S, gpc, lds, ring, scratch, stride, chunk_len, chunk_pos); break;
    }
    __syncthreads();   // LDS tables and ring are reloaded for the next chunk
    }
}

// ------------------------------------------------------------------------
// candidate sizes and the choice, one thread per component instance

// bytes of every section and each chunk's offset inside it: one workgroup
// per section, block scan over its chunks
#define SEC_ITEMS 16
__global__ __launch_bounds__(CH_WG) void k_section_sizes(Section* __restrict__ sec,
                                                         const uint32_t* __restrict__ chunk_len,
                                                         uint32_t* __restrict__ chunk_off) {
    __shared__ uint32_t s_tmp[4];
    Section& S = sec[blockIdx.x];
    const uint32_t nch = S.nchunks, base = S.chunk_base, st = 4u * S.lanes;
    // SEC_ITEMS consecutive chunks per thread and round, all loaded first
    uint64_t run = 0;
    for (uint32_t k0 = 0; k0 < nch; k0 += CH_WG * SEC_ITEMS) {
        const uint32_t kb = k0 + threadIdx.x * SEC_ITEMS;
        uint32_t v[SEC_ITEMS], sum = 0;
#pragma unroll
        for (int i = 0; i < SEC_ITEMS; ++i) {
            const uint32_t k = kb + i;
            const uint32_t len = chunk_len[base + (k < nch ? k : 0u)];   // unconditional (chunk 0 exists)
            v[i] = k < nch ? st + len : 0u;
            sum += v[i];
        }
        uint32_t tot;
        const uint32_t pre = block_excl_scan(sum, s_tmp, &tot);
        uint64_t at = run + pre;
#pragma unroll
        for (int i = 0; i < SEC_ITEMS; ++i) {
            if (kb + i < nch) chunk_off[base + kb + i] = (uint32_t)at;
            at += v[i];
        }
        run += tot;
    }
    if (threadIdx.x == 0) {
        S.payload = run;
        S.size = RVC1_HDR_FIXED + S.table_len + 4ull * nch + run;
    }
}

// Estimates of every candidate and the contenders (chain_int.hpp): Fixed is
// sized exactly here, Tagged and Raw-rANS from their sections' estimates;
// only the contenders' sections are encoded. One thread per instance.
__global__ void k_prechoose(CompInst* __restrict__ ci, Section* __restrict__ sec, uint32_t ninst, uint32_t lanes_max,
                            uint32_t chunk_symbols) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    CompInst& C = ci[i];
    Section& sv = sec[2 * i];
    Section& st = sec[2 * i + 1];
    uint64_t fix_size = 4 + 4 + 8 + (uint64_t)C.dims * 17;
    for (uint32_t d = 0; d < C.dims; ++d) {
        uint32_t w = 1;
        if (C.n && C.dmin[d] != C.dmax[d]) w = (uint32_t)bitlen64(C.dmax[d] - C.dmin[d]);
        C.width[d] = w;
        C.fix_bytes[d] = (C.n * w + 7) / 8;
        fix_size += C.fix_bytes[d];
        if (C.n == 0) C.dmin[d] = 0;
    }
    const uint64_t tag_est = 4 + 4 + 8 + 1 + 8 +
                             rvc1_section_estimate(st.count, st.table_len, rvc1_lanes_for(st.count, lanes_max),
                                                   chunk_symbols, st.est_bits) +
                             8 + (C.tag_bits + 7) / 8;
    const uint64_t rans_est = sv.eligible ? 4 + 4 + 8 + 8 +
                                                rvc1_section_estimate(sv.count, sv.table_len,
                                                                      rvc1_lanes_for(sv.count, lanes_max),
                                                                      chunk_symbols, sv.est_bits)
                                          : ~0ull;
    const uint32_t mask = lch1_contenders(tag_est, fix_size, rans_est);
    C.cand = mask;
    st.encode = (mask & 1u) ? 1u : 0u;
    sv.encode = (mask & 4u) ? 1u : 0u;
}

// the exact smallest contender; ties Tagged > Fixed > Raw-rANS
__global__ void k_choose(CompInst* __restrict__ ci, Section* __restrict__ sec, uint32_t ninst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    CompInst& C = ci[i];
    const Section& sv = sec[2 * i];
    const Section& st = sec[2 * i + 1];
    uint32_t kind = ~0u;
    uint64_t best = 0;
    if (C.cand & 1u) {
        best = 4 + 4 + 8 + 1 + 8 + st.size + 8 + (C.tag_bits + 7) / 8;
        kind = KIND_TAGR;
    }
    if (C.cand & 2u) {
        uint64_t fix_size = 4 + 4 + 8 + (uint64_t)C.dims * 17;
        for (uint32_t d = 0; d < C.dims; ++d) fix_size += C.fix_bytes[d];
        if (kind == ~0u || fix_size < best) { best = fix_size; kind = KIND_FIXP; }
    }
    if (C.cand & 4u) {
        const uint64_t rans_size = 4 + 4 + 8 + 8 + sv.size;
        if (kind == ~0u || rans_size < best) { best = rans_size; kind = KIND_RANS; }
    }
    C.kind = kind;
    C.size = best;
}

// output layout: one thread walks the frames in order
#define TAG_TILE 4096   // records of a Tagged packing tile (16 per thread)
// 64-bit exclusive scan over the workgroup (the container offsets of a batch
// can pass 2^32)
__device__ inline uint64_t block_excl_scan_u64(uint64_t v, uint64_t* s_tmp64, uint64_t* total) {
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    uint64_t inc = v;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint64_t t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) s_tmp64[w] = inc;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (uint32_t k = 0; k < CH_WG / 64; ++k) {
        const uint64_t t = s_tmp64[k];
        before += k < w ? t : 0ull;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// Output layout: one workgroup scans the frames' container sizes in order
// (and, per component, the Tagged tiles of the frames); inside a frame the
// five components follow the 48 + 8D byte header, each after its 8-byte
// length, and the corrections close the frame (16 + 16 per correction).
__global__ __launch_bounds__(CH_WG) void k_layout(CompInst* __restrict__ ci, Section* __restrict__ sec,
                                                  const FrameOut* __restrict__ fo, uint32_t F, uint32_t D,
                                                  uint64_t* __restrict__ out_off, uint64_t* __restrict__ out_size,
                                                  uint64_t* __restrict__ total,
                                                  uint32_t* __restrict__ tag_tiles /* [5][F+1] */, uint64_t out_cap,
                                                  uint32_t* __restrict__ cap_flag) {
    __shared__ uint32_t s_tmp[4];
    __shared__ uint64_t s_tmp64[4];
    uint64_t run = 0;
    uint32_t trun[5] = {0, 0, 0, 0, 0};
    for (uint32_t f0 = 0; f0 < F; f0 += CH_WG) {
        const uint32_t f = f0 + threadIdx.x;
        const bool live = f < F;
        uint64_t size = 0;
        uint32_t tt[5] = {0, 0, 0, 0, 0};
        if (live) {
            size = 48 + 8ull * D + 16 + 16ull * fo[f].ncorr;
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const CompInst& C = ci[c * F + f];
                size += 8 + C.size;
                if (C.kind == KIND_TAGR) tt[c] = (uint32_t)((C.n + TAG_TILE - 1) / TAG_TILE);
            }
        }
        uint64_t tot;
        const uint64_t off = run + block_excl_scan_u64(size, s_tmp64, &tot);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            uint32_t ctot;
            const uint32_t pre = block_excl_scan(tt[c], s_tmp, &ctot);
            if (live) tag_tiles[c * (F + 1) + f] = trun[c] + pre;
            trun[c] += ctot;
        }
        if (live) {
            out_off[f] = off;
            out_size[f] = size;
            uint64_t p = off + 48 + 8ull * D;
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                CompInst& C = ci[c * F + f];
                p += 8;
                C.off = p;
                const uint32_t i = c * F + f;
                if (C.kind == KIND_TAGR) {
                    sec[2 * i + 1].off = p + 4 + 4 + 8 + 1 + 8;
                    C.payload_off = sec[2 * i + 1].off + sec[2 * i + 1].size + 8;
                } else if (C.kind == KIND_FIXP) {
                    C.payload_off = p + 16 + 17ull * C.dims;
                } else {
                    sec[2 * i].off = p + 16 + 8;
                }
                p += C.size;
            }
        }
        run += tot;
    }
    if (threadIdx.x == 0) {
        *total = run;
        if (run > out_cap) atomicOr(cap_flag, 2u);   // the writers skip; the call runs again with the worst-case bound
#pragma unroll
        for (int c = 0; c < 5; ++c) tag_tiles[c * (F + 1) + F] = trun[c];
    }
}

// ------------------------------------------------------------------------
// container bytes

__device__ inline void write_section_header(uint8_t* out, const Section& S, const uint8_t* tbytes, uint32_t si,
                                            const uint32_t* chunk_len, uint32_t tid, uint32_t nthreads) {
    uint8_t* p = out + S.off;
    const bool used = S.eligible && S.count > 0;
    if (tid == 0) {
        p[0] = 'R'; p[1] = 'V'; p[2] = 'C'; p[3] = '1';
        put_u8(p + 4, (uint8_t)(used ? S.pb : 0));
        put_u64(p + 5, S.count);
        put_u64(p + 13, S.table_len);
    }
    const uint8_t* tb = tbytes + (uint64_t)si * TABLE_BYTES_MAX;
    for (uint32_t k = tid; k < S.table_len; k += nthreads) p[21 + k] = tb[k];
    uint8_t* q = p + 21 + S.table_len;
    if (tid == 0) {
        put_u8(q, (uint8_t)S.lanes);
        put_u32(q + 1, used ? S.chunk_symbols : 0u);
        put_u32(q + 5, S.nchunks);
    }
    for (uint32_t k = tid; k < S.nchunks; k += nthreads)
        put_u32(q + 9 + 4 * k, 4u * S.lanes + chunk_len[S.chunk_base + k]);
}

__global__ __launch_bounds__(CH_WG) void k_write_components(const CompInst* __restrict__ ci,
                                                            const Section* __restrict__ sec, uint32_t F,
                                                            const uint8_t* __restrict__ tbytes,
                                                            const uint32_t* __restrict__ chunk_len,
                                                            const uint64_t* __restrict__ out_total, uint64_t out_cap,
                                                            uint8_t* __restrict__ out) {
    if (*out_total > out_cap) return;   // the layout overran the output bound: nothing is written
    const uint32_t i = blockIdx.x;
    const CompInst& C = ci[i];
    const uint32_t tid = threadIdx.x;
    uint8_t* p = out + C.off;
    if (tid == 0) {
        put_u64(p - 8, C.size);
        const char* mg = C.kind == KIND_TAGR ? "TAGR" : (C.kind == KIND_FIXP ? "FIXP" : "RANS");
        for (int k = 0; k < 4; ++k) p[k] = (uint8_t)mg[k];
        put_u32(p + 4, C.dims);
        put_u64(p + 8, C.n);
        if (C.kind == KIND_TAGR) {
            put_u8(p + 16, (uint8_t)C.max_tag);
            put_u64(p + 17, sec[2 * i + 1].size);
            put_u64(out + sec[2 * i + 1].off + sec[2 * i + 1].size, (C.tag_bits + 7) / 8);
        } else if (C.kind == KIND_FIXP) {
            for (uint32_t d = 0; d < C.dims; ++d) {
                put_u64(p + 16 + 17 * d, C.dmin[d]);
                put_u8(p + 16 + 17 * d + 8, (uint8_t)C.width[d]);
                put_u64(p + 16 + 17 * d + 9, C.fix_bytes[d]);
            }
        } else {
            put_u64(p + 16, sec[2 * i].size);
        }
    }
    if (C.kind == KIND_TAGR) write_section_header(out, sec[2 * i + 1], tbytes, 2 * i + 1, chunk_len, tid, CH_WG);
    if (C.kind == KIND_RANS) write_section_header(out, sec[2 * i], tbytes, 2 * i, chunk_len, tid, CH_WG);
}

// frame header and exact corrections (cpp/chain/lcps_cpp.cpp:2306-2329 layout, magic LCH1)
__global__ __launch_bounds__(CH_WG) void k_write_frames(const FrameDev* __restrict__ fr,
                                                        const FrameOut* __restrict__ fo, uint32_t D,
                                                        const uint64_t* __restrict__ out_off,
                                                        const uint64_t* __restrict__ out_size,
                                                        const uint64_t* __restrict__ corr_idx,
                                                        const double* __restrict__ corr_val,
                                                        const uint64_t* __restrict__ out_total, uint64_t out_cap,
                                                        uint8_t* __restrict__ out) {
    if (*out_total > out_cap) return;   // the layout overran the output bound: nothing is written
    const uint32_t f = blockIdx.x;
    const FrameDev& F = fr[f];
    const FrameOut& O = fo[f];
    uint8_t* p = out + out_off[f];
    if (threadIdx.x == 0) {
        p[0] = 'L'; p[1] = 'C'; p[2] = 'H'; p[3] = '1';
        put_u32(p + 4, D);
        put_u64(p + 8, F.n);
        put_u64(p + 16, F.p);
        put_u64(p + 24, (uint64_t)__double_as_longlong(F.eb));
        put_u64(p + 32, (uint64_t)O.wrap_min);
        put_u64(p + 40, (uint64_t)O.wrap_max);
        for (uint32_t d = 0; d < D; ++d) put_u64(p + 48 + 8 * d, (uint64_t)__double_as_longlong(F.mins[d]));
    }
    uint8_t* q = p + out_size[f] - 16 - 16 * O.ncorr;
    if (threadIdx.x == 0) {
        put_u64(q, O.ncorr);
        put_u64(q + 8, 16 * O.ncorr);
    }
    for (uint64_t k = threadIdx.x; k < O.ncorr; k += CH_WG) {
        put_u64(q + 16 + 16 * k, corr_idx[O.corr_base + k]);
        put_u64(q + 24 + 16 * k, (uint64_t)__double_as_longlong(corr_val[O.corr_base + k]));
    }
}

// copy every chunk of a chosen section into place
__global__ __launch_bounds__(CH_WG) void k_copy_chunks(const CompInst* __restrict__ ci, const Section* __restrict__ sec,
                                                       uint32_t nsec, const uint32_t* __restrict__ comp_chunk,
                                                       const uint8_t* __restrict__ scratch, uint64_t stride,
                                                       const uint32_t* __restrict__ chunk_len,
                                                       const uint32_t* __restrict__ chunk_pos,
                                                       const uint32_t* __restrict__ chunk_off,
                                                       const uint64_t* __restrict__ out_total, uint64_t out_cap,
                                                       uint8_t* __restrict__ out) {
    if (*out_total > out_cap) return;   // the layout overran the output bound: nothing is written
    const uint32_t nchunks = comp_chunk[5];
    for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint32_t si = find_section(sec, 0, nsec, chunk);
    const Section& S = sec[si];
    const CompInst& C = ci[si >> 1];
    const bool tags = (si & 1) != 0;
    const bool used = tags ? C.kind == KIND_TAGR : C.kind == KIND_RANS;
    if (!used) continue;
    uint8_t* dst = out + S.off + RVC1_HDR_FIXED + S.table_len + 4ull * S.nchunks + chunk_off[chunk];
    const uint8_t* src = scratch + (uint64_t)chunk * stride;
    const uint32_t nstate = 4u * S.lanes;
    for (uint32_t k = threadIdx.x; k < nstate; k += CH_WG) dst[k] = src[k];
    const uint8_t* st = src + CHUNK_STATE_BYTES + chunk_pos[chunk];
    const uint32_t len = chunk_len[chunk];
    // eight bytes per thread and round, all loaded before any is stored (a
    // byte past the end reads the round's first byte and is not stored)
    for (uint32_t k = threadIdx.x; k < len; k += 8 * CH_WG) {
        uint8_t b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t i = k + u * CH_WG;
            b[u] = st[i < len ? i : k];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t i = k + u * CH_WG;
            if (i < len) dst[nstate + i] = b[u];
        }
    }
    }
}

// Fixed payloads, MSB-first, byte-padded per dimension. Eight consecutive
// values of width w fill exactly w bytes, so thread g of a tile packs values
// 8g .. 8g+7 of each dimension into bytes [g w, g w + w) of the tile's part
// (a tile starts at a multiple of 4096 values, hence on a byte). Values past
// the end are zero, which is the padding of the last byte.
#define PACK_GROUPS (CH_TILE / 8)
template <typename ValT, int DIMS>
__global__ __launch_bounds__(CH_WG) void k_pack_fixed(const ValT* __restrict__ vals, int c, uint32_t F,
                                                      const uint32_t* __restrict__ tile_begin,
                                                      const CompInst* __restrict__ ci,
                                                      const uint64_t* __restrict__ out_total, uint64_t out_cap,
                                                      uint8_t* __restrict__ out) {
    if (*out_total > out_cap) return;   // the layout overran the output bound: nothing is written
    const uint32_t tile = blockIdx.x;
    const uint32_t f = upper_find(tile_begin, F, tile);
    const CompInst& C = ci[c * F + f];
    if (C.kind != KIND_FIXP) return;
    const uint64_t r0 = (uint64_t)(tile - tile_begin[f]) * CH_TILE;
    const uint64_t n = C.n;
    if (r0 >= n) return;
    const uint64_t r1 = r0 + CH_TILE < n ? r0 + CH_TILE : n;
    const ValT* src = vals + C.base * DIMS;
    uint64_t doff = C.payload_off;
    // byte values: the tile's bytes are staged in LDS with coalesced 4-byte
    // loads (the arena's arrays are 256-byte aligned and padded, so the words
    // covering the tile's range are readable), then cut from LDS; the
    // workgroup branches on the array's alignment as a whole
    constexpr uint32_t STAGE_WORDS = sizeof(ValT) == 1 ? CH_TILE * DIMS / 4 + 2 : 1;
    // byte values (width <= 8): every dimension's packed bytes of the tile
    // (an even word count keeps every dimension's run 8-byte aligned)
    constexpr uint32_t OUT_WORDS = sizeof(ValT) == 1 ? CH_TILE * 8 / 32 + 2 : 2;
    __shared__ uint32_t s_w[STAGE_WORDS];
    __shared__ __attribute__((aligned(16))) uint32_t s_o[sizeof(ValT) == 1 ? DIMS : 1][OUT_WORDS];
    const uint8_t* s_b = nullptr;
    if constexpr (sizeof(ValT) == 1) {
        if ((reinterpret_cast<uintptr_t>(vals) & 3u) == 0) {
            const uint64_t b0 = (C.base + r0) * DIMS, b1 = (C.base + r1) * DIMS;
            const uint64_t w0 = b0 >> 2, w1 = (b1 + 3) >> 2;   // w1 - w0 <= STAGE_WORDS
            const uint32_t* vw = reinterpret_cast<const uint32_t*>(vals);
            constexpr int NW = (STAGE_WORDS + CH_WG - 1) / CH_WG;
            uint32_t wv[NW];
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                const uint64_t wi = w0 + (uint64_t)q * CH_WG + threadIdx.x;
                wv[q] = vw[wi < w1 ? wi : w0];   // unconditional
            }
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                const uint64_t wi = w0 + (uint64_t)q * CH_WG + threadIdx.x;
                if (wi < w1) s_w[wi - w0] = wv[q];
            }
            __syncthreads();
            s_b = reinterpret_cast<const uint8_t*>(s_w) + (b0 - 4 * w0);
            // Group g (records 8g .. 8g + 7 of the tile) reads its 8 * DIMS
            // bytes as realigned LDS words; per dimension its eight values
            // fill exactly w bytes (w <= 8), big-endian bit order, which go to
            // that dimension's stream at byte g * w of the tile's run. Each
            // run then leaves as whole 4-byte output words (consecutive LDS
            // words joined at the run's byte offset in its first word), with
            // the bytes of the two edge words, which neighbouring tiles also
            // touch, written singly.
            const uint32_t sb0 = (uint32_t)(b0 - 4 * w0);
            constexpr uint32_t GW = (8 * DIMS + 3) / 4 + 1;   // words a group's bytes can span
            uint32_t wd[DIMS];
            uint64_t mnd[DIMS];
#pragma unroll
            for (int d = 0; d < DIMS; ++d) {
                wd[d] = C.width[d];
                mnd[d] = C.dmin[d];
            }
#pragma unroll
            for (int h = 0; h < PACK_GROUPS / CH_WG; ++h) {
                const uint32_t g = h * CH_WG + threadIdx.x;
                const uint32_t la = 8u * g;
                if (r0 + la >= r1) continue;
                const uint32_t nvals = r1 - r0 - la < 8 ? (uint32_t)(r1 - r0 - la) : 8u;
                const uint32_t q = sb0 + la * DIMS;
                const uint32_t qs = q & 3u;
                uint32_t rw[GW];
#pragma unroll
                for (uint32_t i = 0; i < GW; ++i) rw[i] = s_w[(q >> 2) + i];
#pragma unroll
                for (uint32_t i = 0; i + 1 < GW; ++i) rw[i] = __builtin_amdgcn_alignbyte(rw[i + 1], rw[i], qs);
#pragma unroll
                for (int d = 0; d < DIMS; ++d) {
                    const uint32_t w = wd[d], m32 = (uint32_t)mnd[d];
                    uint8_t* const so = reinterpret_cast<uint8_t*>(s_o[d]) + g * w;
                    if (w <= 4) {
                        uint32_t acc = 0;
#pragma unroll
                        for (uint32_t k = 0; k < 8; ++k) {
                            const uint32_t e = k * DIMS + d;
                            const uint32_t v = k < nvals ? ((rw[e >> 2] >> (8 * (e & 3u))) & 0xffu) - m32 : 0u;
                            acc = (acc << w) | v;
                        }
                        // 8 w bits, the first value in the top bits
                        if (w == 4) *reinterpret_cast<uint32_t*>(so) = __builtin_bswap32(acc);
                        else if (w == 2) *reinterpret_cast<uint16_t*>(so) = __builtin_bswap16((uint16_t)acc);
                        else if (w == 1) *so = (uint8_t)acc;
                        else
                            for (uint32_t b = 0; b < w; ++b) so[b] = (uint8_t)(acc >> (8 * (w - 1 - b)));
                    } else {
                        uint64_t acc = 0;
#pragma unroll
                        for (uint32_t k = 0; k < 8; ++k) {
                            const uint32_t e = k * DIMS + d;
                            const uint32_t v = k < nvals ? ((rw[e >> 2] >> (8 * (e & 3u))) & 0xffu) - m32 : 0u;
                            acc = (acc << w) | v;
                        }
                        if (w == 8) *reinterpret_cast<uint64_t*>(so) = __builtin_bswap64(acc);
                        else
                            for (uint32_t b = 0; b < w; ++b) so[b] = (uint8_t)(acc >> (8 * (w - 1 - b)));
                    }
                }
            }
            __syncthreads();
            uint64_t dof = doff;
#pragma unroll
            for (int d = 0; d < DIMS; ++d) {
                const uint32_t w = wd[d];
                const uint64_t o0 = dof + (r0 * w) / 8;   // r0 is a multiple of 8: a whole byte
                const uint32_t nb = (uint32_t)(((r1 - r0) * w + 7) / 8);
                const uint32_t a = (uint32_t)(o0 & 3u);
                const uint32_t* const L = s_o[d];
                const uint8_t* const Lb = reinterpret_cast<const uint8_t*>(L);
                uint8_t* const ob = out + (o0 - a);   // 4-byte aligned: output byte 4k + i = run byte 4k + i - a
                // whole words k in [k0, k1): run bytes [4k - a, 4k - a + 4) inside [0, nb)
                const uint32_t k0 = a ? 1u : 0u, k1 = (nb + a) / 4;
                if (k0 < k1) {
                    uint32_t* const ow = reinterpret_cast<uint32_t*>(ob);
                    for (uint32_t k = k0 + threadIdx.x; k < k1; k += CH_WG)
                        ow[k] = a ? __builtin_amdgcn_alignbyte(L[k], L[k - 1], 4 - a) : L[k];
                    if (threadIdx.x < 3 && threadIdx.x + a < 4 * k0) ob[a + threadIdx.x] = Lb[threadIdx.x];
                    if (threadIdx.x >= 4 && threadIdx.x < 8 && 4 * k1 + threadIdx.x - 4 < nb + a)
                        ob[4 * k1 + threadIdx.x - 4] = Lb[4 * k1 + threadIdx.x - 4 - a];
                } else {
                    for (uint32_t i = threadIdx.x; i < nb; i += CH_WG) ob[a + i] = Lb[i];
                }
                dof += C.fix_bytes[d];
            }
            return;
        }
    }
#pragma unroll
    for (int d = 0; d < DIMS; ++d) {
        const uint32_t w = C.width[d];
        const uint64_t mn = C.dmin[d];
        ValT x[PACK_GROUPS / CH_WG][8];
        if (s_b) {
#pragma unroll
            for (int h = 0; h < PACK_GROUPS / CH_WG; ++h) {
                const uint32_t la = 8u * (h * CH_WG + threadIdx.x);   // record ra - r0
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint32_t lr = r0 + la + k < r1 ? la + k : 0u;
                    x[h][k] = (ValT)s_b[lr * DIMS + d];
                }
            }
        } else {
#pragma unroll
            for (int h = 0; h < PACK_GROUPS / CH_WG; ++h) {
                const uint64_t ra = r0 + 8ull * (h * CH_WG + threadIdx.x);
#pragma unroll
                for (int k = 0; k < 8; ++k) x[h][k] = src[(ra + k < r1 ? ra + k : r0) * DIMS + d];   // unconditional
            }
        }
#pragma unroll
        for (int h = 0; h < PACK_GROUPS / CH_WG; ++h) {
            const uint64_t ra = r0 + 8ull * (h * CH_WG + threadIdx.x);
            if (ra >= r1) continue;
            const uint32_t nvals = r1 - ra < 8 ? (uint32_t)(r1 - ra) : 8u;
            const uint32_t nbytes = (nvals * w + 7) / 8;
            uint8_t* o = out + doff + (ra * w) / 8;
            uint64_t v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = (uint32_t)k < nvals ? (uint64_t)x[h][k] - mn : 0ull;
            if (w <= 8) {
                uint64_t acc = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) acc = (acc << w) | v[k];
                for (uint32_t b = 0; b < nbytes; ++b) o[b] = (uint8_t)(acc >> (8 * (w - 1 - b)));
            } else {
                // a 64-bit window, values pushed in pieces of at most 32 bits
                uint64_t acc = 0;
                uint32_t nb = 0, pos = 0;
                for (int k = 0; k < 8; ++k) {
                    uint32_t rem = w;
                    while (rem) {
                        const uint32_t take = rem < 32u ? rem : 32u;
                        acc = (acc << take) | ((v[k] >> (rem - take)) & ((1ull << take) - 1ull));
                        nb += take;
                        rem -= take;
                        while (nb >= 8) {
                            nb -= 8;
                            if (pos < nbytes) o[pos] = (uint8_t)(acc >> nb);
                            ++pos;
                        }
                    }
                }
            }
        }
        doff += C.fix_bytes[d];
    }
}

// Tagged payloads: TAG_TILE records per tile, bit offsets by a segmented look-back
// over each frame's tiles, bits assembled in LDS; the partial bytes at the
// two ends of a tile go to side arrays and are merged by k_tag_seams.
// tag_tiles[f] (F+1 entries) counts the tiles of the frames whose component
// is Tagged. Workgroups take tiles by ticket, in order, until they run out.
#define PACK_TAG_WGS 2048u
template <typename ValT, int DIMS>
__global__ __launch_bounds__(CH_WG) void k_pack_tagged(const ValT* __restrict__ vals, int c, uint32_t F,
                                                       const uint32_t* __restrict__ tag_tiles,
                                                       const CompInst* __restrict__ ci, uint64_t* status,
                                                       uint32_t* ticket, uint32_t* err,
                                                       uint8_t* __restrict__ head, uint8_t* __restrict__ tail,
                                                       uint64_t* __restrict__ head_pos,
                                                       const uint64_t* __restrict__ out_total, uint64_t out_cap,
                                                       uint8_t* __restrict__ out) {
    if (*out_total > out_cap) return;   // the layout overran the output bound: nothing is written
    __shared__ uint32_t s_tile, s_tmp[4];
    __shared__ uint64_t s_excl;
    // widest value: 63 bits (values below 2^63), else the type's width
    constexpr uint32_t MAXW = sizeof(ValT) == 8 ? 63u : 8u * (uint32_t)sizeof(ValT);
    constexpr uint32_t RPT = TAG_TILE / CH_WG;   // records per thread
    __shared__ uint32_t s_words[(TAG_TILE * DIMS * MAXW) / 32 + 4];
    const uint32_t* tile_begin = tag_tiles + c * (F + 1);
    // The grid is sized for the largest possible tile count, capped at
    // PACK_TAG_WGS workgroups; each workgroup takes tiles by ticket until
    // they run out, so an unused component costs one short launch.
    const uint32_t total = tile_begin[F];
    if (blockIdx.x >= total) return;
    for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= total) break;
    const uint32_t f = upper_find(tile_begin, F, tile);
    const CompInst& C = ci[c * F + f];
    if (threadIdx.x == 0) head_pos[tile] = ~0ull;
    const uint64_t r0 = (uint64_t)(tile - tile_begin[f]) * TAG_TILE;
    const ValT* src = vals + C.base * DIMS;
    const uint64_t rt = r0 + threadIdx.x * RPT;
    // the thread's RPT records, loaded once and unconditionally (the tile's
    // first record past the end)
    uint64_t xv[RPT][DIMS];
#pragma unroll
    for (int k = 0; k < (int)RPT; ++k)
#pragma unroll
        for (int d = 0; d < DIMS; ++d) xv[k][d] = (uint64_t)src[(rt + k < C.n ? rt + k : r0) * DIMS + d];
    uint32_t tg[RPT];
    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < (int)RPT; ++k) {
        tg[k] = 0;
        if (rt + k < C.n) {
            uint64_t any = xv[k][0];   // bit length of the OR = widest bit length
#pragma unroll
            for (int d = 1; d < DIMS; ++d) any |= xv[k][d];
            const uint32_t wb = (uint32_t)bitlen64(any);
            const uint32_t t = wb > 1u ? wb : 1u;
            tg[k] = t;
            bits += DIMS * t;
        }
    }
    uint32_t tot;
    const uint32_t pre = block_excl_scan(bits, s_tmp, &tot);
    if (threadIdx.x < 64) {
        const uint64_t e = lookback_seg(status, tile, tile_begin[f], tot, err);
        if (threadIdx.x == 0) s_excl = e;
    }
    __syncthreads();
    const uint64_t b0 = s_excl, b1 = b0 + tot;
    const uint32_t lb0 = (uint32_t)(b0 & 7);
    const uint64_t gb0 = b0 >> 3;
    // The tile's bits are staged in words aligned to the output: staged byte
    // a + i is output byte payload_off + gb0 + i, a = its offset in a 4-byte
    // word. Stream bit o (MSB first) is bit 7 - (o & 7) of staged byte o >> 3;
    // the words hold their bytes little-endian, so a value is placed as a
    // big-endian bit field, byte-swapped, with at most three word ORs.
    const uint64_t gbyte0 = C.payload_off + gb0;
    const uint32_t a = (uint32_t)(gbyte0 & 3u);
    const uint32_t nwords = (8 * a + lb0 + tot + 31) / 32 + 2;
    for (uint32_t k = threadIdx.x; k < nwords; k += CH_WG) s_words[k] = 0;
    __syncthreads();
    uint32_t o = 8 * a + lb0 + pre;
#pragma unroll
    for (int k = 0; k < (int)RPT; ++k) {
        if (!tg[k]) continue;
        const uint32_t w = tg[k];
#pragma unroll
        for (int d = 0; d < DIMS; ++d) {
            const uint64_t v = xv[k][d];   // below 2^w
            const uint32_t wi = o >> 5, sh = o & 31u;
            if (sh + w <= 64u) {
                const uint64_t x = v << (64u - sh - w);
                atomicOr(&s_words[wi], __builtin_bswap32((uint32_t)(x >> 32)));
                if (sh + w > 32u) atomicOr(&s_words[wi + 1], __builtin_bswap32((uint32_t)x));
            } else {
                // 64 - sh high bits fill the rest of words wi, wi + 1; the
                // low r bits start word wi + 2
                const uint32_t r = sh + w - 64u;
                const uint64_t x = v >> r;
                atomicOr(&s_words[wi], __builtin_bswap32((uint32_t)(x >> 32)));
                atomicOr(&s_words[wi + 1], __builtin_bswap32((uint32_t)x));
                atomicOr(&s_words[wi + 2], __builtin_bswap32((uint32_t)(v << (32u - r))));
            }
            o += w;
        }
    }
    __syncthreads();
    // staged bytes [a, a + nbytes): a partial first byte goes to head (merged
    // with the previous tile's tail by k_tag_seams), a partial last byte to
    // tail unless the component ends here; the bytes between are written as
    // aligned words where whole, byte by byte at the two ends
    const uint32_t nbytes = (uint32_t)(((b1 + 7) >> 3) - gb0);
    const bool last = r0 + TAG_TILE >= C.n;
    const bool has_head = lb0 != 0, has_tail = (b1 & 7) != 0 && !last;
    const uint8_t* sb = reinterpret_cast<const uint8_t*>(s_words);
    if (threadIdx.x == 0 && has_head) {
        head[tile] = sb[a];
        head_pos[tile] = gbyte0;
    }
    if (threadIdx.x == 1 && has_tail) tail[tile] = sb[a + nbytes - 1];
    const uint32_t i0 = a + (has_head ? 1u : 0u), i1 = a + nbytes - (has_tail ? 1u : 0u);
    uint8_t* const ob = out + (gbyte0 - a);   // 4-byte aligned: staged byte i -> ob[i]
    const uint32_t k0 = (i0 + 3) / 4, k1 = i1 / 4;
    if (k0 < k1) {
        uint32_t* const ow = reinterpret_cast<uint32_t*>(ob);
        for (uint32_t k = k0 + threadIdx.x; k < k1; k += CH_WG) ow[k] = s_words[k];
        // edge bytes: [i0, 4 k0) and [4 k1, i1)
        if (threadIdx.x < 3 && i0 + threadIdx.x < 4 * k0) ob[i0 + threadIdx.x] = sb[i0 + threadIdx.x];
        if (threadIdx.x >= 4 && threadIdx.x < 7 && 4 * k1 + threadIdx.x - 4 < i1)
            ob[4 * k1 + threadIdx.x - 4] = sb[4 * k1 + threadIdx.x - 4];
    } else {
        for (uint32_t i = i0 + threadIdx.x; i < i1; i += CH_WG) ob[i] = sb[i];
    }
    __syncthreads();   // LDS reused by the next tile
    }
}

// seams of every component (blockIdx.y): slice c of the seam arrays holds
// its tiles; tag_tiles[c * (F + 1) + F] is its tile count
__global__ void k_tag_seams(const uint8_t* __restrict__ head, const uint8_t* __restrict__ tail,
                            const uint64_t* __restrict__ head_pos, uint32_t stride,
                            const uint32_t* __restrict__ tag_tiles, uint32_t F,
                            const uint64_t* __restrict__ out_total, uint64_t out_cap,
                            uint8_t* __restrict__ out) {
    if (*out_total > out_cap) return;   // the layout overran the output bound: nothing is written
    const uint32_t c = blockIdx.y;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t o = (uint64_t)c * stride;
    if (t >= tag_tiles[c * (F + 1) + F] || head_pos[o + t] == ~0ull) return;
    out[head_pos[o + t]] = head[o + t] | tail[o + t - 1];
}

}  // namespace chain
