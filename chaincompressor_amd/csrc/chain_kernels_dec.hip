// Decoder kernels of the Chain engine (gfx950): LCH1 and LCP3 streams to
// coordinates. Follows the reference decoder's semantics
// (cpp/decompress/lcps_cpu_decompress_serial.cpp:1182-1338):
//   component decode -> L2 inverse (:1230-1275) -> backward chain pass
//   (:1277-1302) -> fma reconstruction and exact corrections (:1305-1326).
// Parallel form: rANS chunks per wavefront, a segmented modular scan with
// decoupled look-back for the L2 inverse, and an O(N) chain resolution that
// walks tiles from the end (every successor lies at a larger index).
#include "chain_common.hpp"

#pragma clang fp contract(off)

namespace chain {

enum : uint32_t {
    DERR_RANS = 1u << 4,        // rANS state / stream mismatch
    DERR_VALUE = 1u << 5,       // value out of range for its stream
    DERR_L2 = 1u << 6,          // invalid choice / correction / kept block
    DERR_CHAIN = 1u << 7,       // successor outside the frame / unresolved chain
    DERR_QUANT = 1u << 8,       // offset >= p or quantised overflow
    DERR_CORR = 1u << 9,        // invalid exact correction
    DERR_COUNT = 1u << 10,      // explicit points != records
    DERR_WAIT = 1u << 11,       // bounded wait expired
};

// one decoded component stream of the batch: values go to a typed array
struct DecFrame {
    uint64_t n;                 // points
    uint64_t pt_off;            // global point offset
    uint64_t rec_base, nrec;    // explicit records (choice2 count)
    uint64_t diff_base, ndiff, keep_base, nkeep;
    uint64_t p;
    double eb, denom;
    double mins[CH_MAXD];
    int64_t wrap_min, wrap_max;
    uint64_t range;             // wrap_max - wrap_min + 1
    int64_t min_corr, max_corr;
    uint64_t out_off;           // element offset of the frame in the output
    uint64_t corr_off, ncorr;   // byte offset of the raw correction records in the input pool
    uint32_t tile0, rtile0;     // first point tile, first record tile
    // LCG2 streams (the reference GPU compressor's container): float32
    // reconstruction fmaf(float(q) + 0.5f, 2 eb_f, mins_f[d]), q <= INT32_MAX
    uint32_t lcg2;
    float eb_f;
    float mins_f[CH_MAXD];
};

// destination of a decoded value: component c of frame f, value index k
// in-block offsets are stored as u8 when every frame of the batch has
// p <= 256 (rel8), else as u32; larger values are format errors either way.
// Diff and keep values are stored as u32 when every frame's wrap range lies
// in [0, 2^31) (dk32): a valid keep is at most wrap_max and a valid diff at
// most the range, so a value of 2^32 or more is the L2 error it would be
// found to be later.
__device__ inline void store_dk(void* dst, uint64_t gidx, uint64_t v, uint32_t dk32, uint32_t* err) {
    if (dk32) {
        if (v >> 32) atomicOr(err, DERR_L2);
        static_cast<uint32_t*>(dst)[gidx] = (uint32_t)v;
    } else {
        static_cast<uint64_t*>(dst)[gidx] = v;
    }
}
__device__ inline void store_value(int c, uint64_t gidx, uint64_t v, uint32_t* idrel, void* diff, void* keep,
                                   uint32_t dk32, uint8_t* choice, void* rel, uint32_t rel8, uint32_t* err) {
    switch (c) {
        case 0: if (v >> 32) atomicOr(err, DERR_VALUE); idrel[gidx] = (uint32_t)v; break;
        case 1: store_dk(diff, gidx, v, dk32, err); break;
        case 2: store_dk(keep, gidx, v, dk32, err); break;
        case 3: if (v > 1) atomicOr(err, DERR_VALUE); choice[gidx] = (uint8_t)v; break;
        default:
            if (rel8) {
                if (v > 0xffu) atomicOr(err, DERR_QUANT);
                static_cast<uint8_t*>(rel)[gidx] = (uint8_t)v;
            } else {
                if (v >> 32) atomicOr(err, DERR_VALUE);
                static_cast<uint32_t*>(rel)[gidx] = (uint32_t)v;
            }
            break;
    }
}

struct DecSection {
    uint32_t pb, lanes, alphabet, chunk_symbols;
    uint64_t count;
    uint64_t cum_off;           // alphabet + 1 cumulative frequencies in the u32 pool
    int32_t comp;               // 0..4 values of that component, -1 widths
    uint32_t pad;
    uint64_t dst;               // first global value index (values) or first width index (tags)
};

struct DecChunk {
    uint32_t sec;
    uint32_t bytes;             // chunk bytes (states + stream)
    uint64_t byte_off;          // chunk start (states) in the input pool
    uint64_t first;             // first symbol of the chunk within its section
};

// up to DEC_WAVES consecutive chunks of one section, decoded by one workgroup
#define DEC_WAVES 8
struct DecJob {
    uint32_t sec, chunk0, nchunks, pad;
};

// rANS decoding. A workgroup of DEC_WAVES wavefronts takes up to DEC_WAVES
// chunks of one section, one chunk per wavefront. The section's cumulative
// frequencies and a coarse index of 4096 buckets (the symbol at the start of
// every bucket of 2^(pb-12) slots) are built in LDS, so every precision up
// to 20 bits decodes without global table reads: the symbol of a slot is a
// bucket lookup, refined by binary search over the symbols inside the bucket
// when pb > 12. Each wavefront stages its chunk's byte stream through its own
// LDS ring of DEC_RING bytes, refilled in halves of 64 lanes x 16 bytes with
// the next half already in registers. Ring and tables fit four workgroups in
// a CU's LDS, and the register budget is held to four waves per SIMD.
#define DEC_RING 2048u
#define DEC_HALF 1024u
#define DEC_BUCKETS 4096u
__global__ __launch_bounds__(64 * DEC_WAVES, 4) void k_dec_rans(const uint8_t* __restrict__ in,
                                                             const DecSection* __restrict__ secs,
                                                             const DecChunk* __restrict__ chunks,
                                                             const DecJob* __restrict__ jobs,
                                                             const uint32_t* __restrict__ cum_pool,
                                                             uint8_t* __restrict__ tags, uint32_t* idrel,
                                                             void* diff, void* keep, uint32_t dk32, uint8_t* choice,
                                                             void* rel, uint32_t rel8, uint32_t* err) {
    __shared__ uint32_t s_cum[RANS_VALUE_LIMIT + 1];
    __shared__ uint16_t s_bucket[DEC_BUCKETS + 1];
    __shared__ uint8_t s_ring[DEC_WAVES][DEC_RING];
    const DecJob J = jobs[blockIdx.x];
    const DecSection S = secs[J.sec];
    const uint32_t A = S.alphabet, pb = S.pb, M = 1u << pb, lower = 4u * M;
    const uint32_t sh = pb > 12 ? pb - 12 : 0u;
    const uint32_t nb = M >> sh;   // buckets (<= 4096)
    for (uint32_t i = threadIdx.x; i <= A; i += blockDim.x) s_cum[i] = cum_pool[S.cum_off + i];
    __syncthreads();
    // bucket b: the largest symbol s < A with cum[s] <= b << sh
    for (uint32_t b = threadIdx.x; b <= nb; b += blockDim.x) {
        const uint32_t slot0 = b << sh;
        uint32_t lo = 0, hi = A - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (s_cum[mid] <= slot0) lo = mid; else hi = mid - 1;
        }
        s_bucket[b] = (uint16_t)lo;
    }
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (wave >= J.nchunks) return;
    const DecChunk ch = chunks[J.chunk0 + wave];
    uint8_t* ring = s_ring[wave];
    const uint32_t L = S.lanes;
    const uint64_t Sc = S.count - ch.first < S.chunk_symbols ? S.count - ch.first : S.chunk_symbols;
    const uint64_t T = (Sc + L - 1) / L;
    const uint8_t* p = in + ch.byte_off;
    uint32_t x = lower;
    if (lane < L) x = (uint32_t)p[4 * lane] | ((uint32_t)p[4 * lane + 1] << 8) | ((uint32_t)p[4 * lane + 2] << 16) |
                      ((uint32_t)p[4 * lane + 3] << 24);
    if (lane < L && (x < lower || x >= 256u * lower)) {
        atomicOr(err, DERR_RANS);
        x = lower;   // keep a corrupt stream's work bounded
    }
    const uint8_t* stream = p + 4 * L;
    const uint64_t stream_len = ch.bytes - 4ull * L;
    // ring staging: byte b of the stream lives at ring[b % DEC_RING]; this
    // lane moves bytes from + 64*j + lane, j < DEC_HALF / 64, of every half
    constexpr int PW = DEC_HALF / 256;   // staged words per lane
    uint32_t pre[PW];
    auto fetch = [&](uint64_t from) {
#pragma unroll
        for (int w = 0; w < PW; ++w) {
            uint32_t v = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t b = from + (uint64_t)(4 * w + k) * 64 + lane;
                // unconditional load (past the stream: the chunk's first state
                // byte) so that the half's loads are all in flight together
                const uint32_t by = *(b < stream_len ? stream + b : p);
                v |= (b < stream_len ? by : 0u) << (8 * k);
            }
            pre[w] = v;
        }
    };
    auto commit = [&](uint64_t from) {
#pragma unroll
        for (int w = 0; w < PW; ++w)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                ring[(uint32_t)(from + (uint64_t)(4 * w + k) * 64 + lane) & (DEC_RING - 1)] = (uint8_t)(pre[w] >> (8 * k));
        __builtin_amdgcn_wave_barrier();
    };
    uint64_t loaded = 0;
    fetch(0);
    commit(0);
    loaded = DEC_HALF;
    if (loaded < stream_len) fetch(loaded);
    uint64_t pos = 0;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint64_t dst0 = S.dst + ch.first;   // first output index of the chunk
    // The step loop, specialised for the section's output (widths, or one
    // component's values of its element type): symbol k of the chunk goes
    // to dst0 + k. Alphabets are below 4096, so only the one-bit choices
    // and one-byte offsets can be out of range.
    auto steps = [&](auto kind) {
        constexpr int K = decltype(kind)::value;
        for (uint32_t t = 0; t < (uint32_t)T; ++t) {
            // keep at least DEC_HALF staged bytes ahead of pos (a step takes <= 256)
            if (loaded - pos < DEC_HALF && loaded < stream_len) {
                commit(loaded);
                loaded += DEC_HALF;
                if (loaded < stream_len) fetch(loaded);
            }
            const uint32_t k = t * L + lane;
            const bool active = lane < L && k < Sc;
            if (active) {
                const uint32_t slot = x & (M - 1);
                const uint32_t b = slot >> sh;
                uint32_t s = s_bucket[b];
                if (sh) {
                    uint32_t hi = s_bucket[b + 1];
                    while (s < hi) {
                        const uint32_t mid = (s + hi + 1) >> 1;
                        if (s_cum[mid] <= slot) s = mid; else hi = mid - 1;
                    }
                }
                const uint32_t c0 = s_cum[s];
                x = (s_cum[s + 1] - c0) * (x >> pb) + slot - c0;
                const uint64_t g = dst0 + k;
                if constexpr (K < 0) {
                    tags[g] = (uint8_t)s;
                } else if constexpr (K == 0) {
                    idrel[g] = s;
                } else if constexpr (K == 1) {
                    static_cast<uint64_t*>(diff)[g] = s;
                } else if constexpr (K == 2) {
                    static_cast<uint64_t*>(keep)[g] = s;
                } else if constexpr (K == 6) {   // symbols < 4096: no range check
                    static_cast<uint32_t*>(diff)[g] = s;
                } else if constexpr (K == 7) {
                    static_cast<uint32_t*>(keep)[g] = s;
                } else if constexpr (K == 3) {
                    if (s > 1) atomicOr(err, DERR_VALUE);
                    choice[g] = (uint8_t)s;
                } else if constexpr (K == 4) {
                    if (s > 0xffu) atomicOr(err, DERR_QUANT);
                    static_cast<uint8_t*>(rel)[g] = (uint8_t)s;
                } else {
                    static_cast<uint32_t*>(rel)[g] = s;
                }
            }
            // Bytes this step reads. lower = 2^(pb+2), pb >= 12: x << 8 | byte
            // < lower exactly when x < lower >> 8, whatever the byte, and two
            // bytes fall short exactly when x < lower >> 16 (pb >= 14; below,
            // only x = 0 would, which no valid stream reaches: x >= 4 after a
            // step). So the count is known before any byte is read and the (at
            // most three) planes are read together; a state still below lower
            // afterwards is a corrupt stream.
            const uint32_t kb = active ? (uint32_t)(x < lower) + (uint32_t)(x < (lower >> 8)) +
                                             (uint32_t)(x < (lower >> 16))
                                       : 0u;
            const uint64_t m0 = __ballot(kb > 0), m1 = __ballot(kb > 1), m2 = __ballot(kb > 2);
            const uint32_t n0 = __popcll(m0), n1 = __popcll(m1);
            if (kb > 0) {
                auto byte_at = [&](uint64_t at) -> uint32_t {
                    return at < stream_len ? (uint32_t)ring[(uint32_t)at & (DEC_RING - 1)] : 0u;
                };
                uint32_t v = byte_at(pos + __popcll(m0 & lt));
                if (kb > 1) v = (v << 8) | byte_at(pos + n0 + __popcll(m1 & lt));
                if (kb > 2) v = (v << 8) | byte_at(pos + n0 + n1 + __popcll(m2 & lt));
                x = (x << (8 * kb)) | v;
            }
            pos += n0 + n1 + __popcll(m2);
            if (active && x < lower) {
                atomicOr(err, DERR_RANS);
                x = lower;
            }
        }
    };
    switch (S.comp) {
        case -1: steps(std::integral_constant<int, -1>{}); break;
        case 0: steps(std::integral_constant<int, 0>{}); break;
        case 1:
            if (dk32) steps(std::integral_constant<int, 6>{});
            else steps(std::integral_constant<int, 1>{});
            break;
        case 2:
            if (dk32) steps(std::integral_constant<int, 7>{});
            else steps(std::integral_constant<int, 2>{});
            break;
        case 3: steps(std::integral_constant<int, 3>{}); break;
        default:
            if (rel8) steps(std::integral_constant<int, 4>{});
            else steps(std::integral_constant<int, 5>{});
            break;
    }
    if (lane < L && x != lower) atomicOr(err, DERR_RANS);
    if (lane == 0 && pos != stream_len) atomicOr(err, DERR_RANS);
}

// Fixed payload job: one component of one frame (all its dimensions)
struct DecFixed {
    uint64_t byte_off[CH_MAXD]; // payload of each dimension in the input pool
    uint64_t minimum[CH_MAXD];
    uint64_t end[CH_MAXD];      // end of each dimension's payload in the input pool
    uint32_t width[CH_MAXD];
    uint64_t n;                 // records
    uint64_t dst;               // first value index (record r -> dst + r*dims + d)
    uint32_t dims;
    int32_t comp;
    uint32_t tile0;             // first tile of the job (prefix over jobs)
    uint32_t recs;              // records per tile: DEC_FIX_RECS, or DEC_FIX_BRECS on the byte path
};

// One workgroup per tile of DEC_FIX_RECS records of one job: every dimension
// of those records, dimension by dimension, so that width, payload and
// minimum are uniform (scalar) values. A value of width <= 33 bits is cut from
// two aligned 32-bit words (MSB-first: byte-swapped); every such pair of the
// tile is loaded before the first is used. Values whose pair would reach past
// the end of their payload, or wider ones, are assembled byte by byte
// afterwards.
#define DEC_FIX_LDS 1024u
#define DEC_FIX_ITEMS 4
#define DEC_FIX_RECS (CH_WG * DEC_FIX_ITEMS)
#define DEC_FIX_BRECS 4096u   // records per tile of a byte-path job

// J is a register copy of the job (uniform, indexed only by unrolled
// constants); Jg is the job in memory, read by the rare byte-by-byte path
template <typename OutT>
__device__ __forceinline__ void fixed_tile(const uint8_t* __restrict__ in, const DecFixed J,
                                          const DecFixed* __restrict__ Jg, uint32_t tile_in_job,
                                          OutT* __restrict__ out, uint64_t vmax, uint32_t verr, uint32_t* err) {
    const uint32_t dims = J.dims;
    const uint64_t rbase = (uint64_t)tile_in_job * DEC_FIX_RECS;
    // In-tile offsets are 32-bit lane values over per-tile scalar bases:
    // record rbase + lr, lr < DEC_FIX_RECS. rbase is a multiple of 1024, so
    // the tile's first value of every dimension starts on a byte.
    const uint32_t nr = J.n - rbase < DEC_FIX_RECS ? (uint32_t)(J.n - rbase) : DEC_FIX_RECS;
    uint32_t q0[CH_MAXD][DEC_FIX_ITEMS], q1[CH_MAXD][DEC_FIX_ITEMS];
    uint32_t slow = 0;   // bit d * DEC_FIX_ITEMS + j
#pragma unroll
    for (int d = 0; d < CH_MAXD; ++d) {
        if ((uint32_t)d >= dims) break;
        const uint32_t w = J.width[d];
        const bool narrow = w <= 33;
        const uint64_t tb = J.byte_off[d] + ((rbase * w) >> 3);
        const uint64_t ta = tb & ~3ull;
        const uint32_t ts = (uint32_t)(tb & 3ull) * 8;
        const uint64_t tail = (J.end[d] - 8) & ~3ull;
#pragma unroll
        for (int j = 0; j < DEC_FIX_ITEMS; ++j) {
            const uint32_t lr = (uint32_t)j * CH_WG + threadIdx.x;
            const uint64_t a4 = ta + 4ull * ((ts + lr * w) >> 5);
            const bool live = lr < nr;
            const bool fast = live && narrow && a4 + 8 <= J.end[d];
            // unconditional loads (a lane that does not use them reads the
            // payload's last aligned word pair, which a component header
            // always precedes) so that every load of the tile is in flight
            const uint32_t* q = reinterpret_cast<const uint32_t*>(in + (fast ? a4 : tail));
            q0[d][j] = q[0];
            q1[d][j] = q[1];
            if (live && !fast) slow |= 1u << (d * DEC_FIX_ITEMS + j);
        }
    }
    auto put = [&](uint64_t r, uint32_t d, uint64_t mn, uint64_t v) {
        if (v > (uint64_t)INT64_MAX - mn) atomicOr(err, DERR_VALUE);
        const uint64_t y = mn + v;
        if (y > vmax) atomicOr(err, verr);
        out[J.dst + r * dims + d] = (OutT)y;
    };
    OutT* const tout = out + J.dst + rbase * dims;
#pragma unroll
    for (int d = 0; d < CH_MAXD; ++d) {
        if ((uint32_t)d >= dims) break;
        const uint32_t w = J.width[d];
        const uint64_t mn = J.minimum[d];
        // every value of width w lands in range: the per-value checks can go
        const uint64_t vm = w >= 64 ? ~0ull : (1ull << w) - 1ull;
        const bool checked = !(mn <= vmax && vm <= vmax - mn && vm <= (uint64_t)INT64_MAX - mn);
        const uint32_t ts = (uint32_t)((J.byte_off[d] + ((rbase * w) >> 3)) & 3ull) * 8;
#pragma unroll
        for (int j = 0; j < DEC_FIX_ITEMS; ++j) {
            const uint32_t lr = (uint32_t)j * CH_WG + threadIdx.x;
            if (lr >= nr || ((slow >> (d * DEC_FIX_ITEMS + j)) & 1u)) continue;
            const uint32_t s = (ts + lr * w) & 31u;
            const uint64_t x = ((uint64_t)__builtin_bswap32(q0[d][j]) << 32) | __builtin_bswap32(q1[d][j]);
            const uint64_t v = (x << s) >> (64 - w);
            if (checked) put(rbase + lr, d, mn, v);
            else tout[lr * dims + d] = (OutT)(mn + v);
        }
    }
    while (slow) {
        const uint32_t b = (uint32_t)__builtin_ctz(slow);
        slow &= slow - 1;
        const uint32_t d = b / DEC_FIX_ITEMS, j = b % DEC_FIX_ITEMS;
        const uint64_t r = rbase + (uint64_t)j * CH_WG + threadIdx.x;
        const uint32_t w = Jg->width[d];
        const uint64_t bit = r * w;
        const uint8_t* p = in + Jg->byte_off[d] + (bit >> 3);
        const uint32_t sh = (uint32_t)(bit & 7);
        const uint32_t nb = (sh + w + 7) >> 3;
        uint64_t acc_hi = 0, acc = 0;
        for (uint32_t qq = 0; qq < nb; ++qq) {
            acc_hi = (acc_hi << 8) | (acc >> 56);
            acc = (acc << 8) | p[qq];
        }
        const uint32_t drop = 8 * nb - sh - w;
        uint64_t v = drop ? ((acc >> drop) | (acc_hi << (64 - drop))) : acc;
        v &= (w >= 64 ? ~0ull : ((1ull << w) - 1ull));
        put(r, d, Jg->minimum[d], v);
    }
}

// Byte outputs whose every value lands in range without a check (width <= 8,
// minimum + 2^w - 1 within the type): a group of 8 records of one dimension
// is exactly w payload bytes, read as three aligned words (bytes one by one
// for the groups at the payload's end), cut with 32-bit bit-field extracts
// for w <= 4, and written to LDS at their output positions; the tile's
// output run (records x dims bytes) then leaves as whole 4-byte words, with
// the bytes of its two edge words written singly.
__device__ __forceinline__ void fixed_tile_bytes(const uint8_t* __restrict__ in, const DecFixed* __restrict__ Jg,
                                                uint32_t dims, uint64_t n, uint64_t dst, uint32_t tile_in_job,
                                                uint8_t* __restrict__ out, uint8_t* __restrict__ s_out) {
    const uint64_t rbase = (uint64_t)tile_in_job * DEC_FIX_BRECS;
    const uint32_t nr = n - rbase < DEC_FIX_BRECS ? (uint32_t)(n - rbase) : DEC_FIX_BRECS;
    const uint64_t o0 = dst + rbase * dims;   // first output byte of the tile
    const uint32_t a2 = (uint32_t)(o0 & 3u);  // LDS byte a2 + i holds output byte o0 + i
    constexpr uint32_t NG = DEC_FIX_BRECS / 8;   // groups per dimension (a power of two)
    static_assert((NG & (NG - 1)) == 0 && NG % 64 == 0, "a wavefront's groups share a dimension");
    for (uint32_t pi = threadIdx.x; pi < dims * NG; pi += CH_WG) {
        const uint32_t d = __builtin_amdgcn_readfirstlane(pi / NG);
        const uint32_t g = pi & (NG - 1);
        if (8 * g >= nr) continue;
        const uint32_t w = Jg->width[d];
        const uint32_t mn = (uint32_t)Jg->minimum[d];
        const uint64_t src = Jg->byte_off[d] + ((rbase * w) >> 3) + (uint64_t)g * w;
        const uint64_t end = Jg->end[d];
        const uint32_t a = (uint32_t)(src & 3u);
        uint32_t hi, lo;   // the group's 8 bytes, big-endian bit order
        if ((src & ~3ull) + 12 <= end) {
            const uint32_t* q = reinterpret_cast<const uint32_t*>(in + (src & ~3ull));
            const uint32_t w0 = q[0], w1 = q[1], w2 = q[2];
            hi = __builtin_bswap32(__builtin_amdgcn_alignbyte(w1, w0, a));
            lo = __builtin_bswap32(__builtin_amdgcn_alignbyte(w2, w1, a));
        } else {
            uint64_t be = 0;
            for (uint32_t b = 0; b < 8; ++b) be = (be << 8) | (src + b < end ? (uint64_t)in[src + b] : 0ull);
            hi = (uint32_t)(be >> 32);
            lo = (uint32_t)be;
        }
        const uint32_t nk = nr - 8 * g < 8 ? nr - 8 * g : 8u;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            if (k >= nk) break;
            uint32_t v;
            if (w <= 4) {
                v = __builtin_amdgcn_ubfe(hi, 32 - w * (k + 1), w);
            } else {
                const uint64_t be = ((uint64_t)hi << 32) | lo;
                v = (uint32_t)(be >> (64 - w * (k + 1))) & ((1u << w) - 1u);
            }
            s_out[a2 + (8 * g + k) * dims + d] = (uint8_t)(mn + v);
        }
    }
    __syncthreads();
    const uint32_t nb = nr * dims;
    uint8_t* const ob = out + (o0 - a2);   // 4-byte aligned
    const uint32_t k0 = a2 ? 1u : 0u, k1 = (nb + a2) / 4;
    if (k0 < k1) {
        const uint32_t* L = reinterpret_cast<const uint32_t*>(s_out);
        uint32_t* const ow = reinterpret_cast<uint32_t*>(ob);
        for (uint32_t k = k0 + threadIdx.x; k < k1; k += CH_WG) ow[k] = L[k];
        if (threadIdx.x < 3 && a2 + threadIdx.x < 4 * k0) ob[a2 + threadIdx.x] = s_out[a2 + threadIdx.x];
        if (threadIdx.x >= 4 && threadIdx.x < 8 && 4 * k1 + threadIdx.x - 4 < nb + a2)
            ob[4 * k1 + threadIdx.x - 4] = s_out[4 * k1 + threadIdx.x - 4];
    } else {
        for (uint32_t i = threadIdx.x; i < nb; i += CH_WG) ob[a2 + i] = s_out[a2 + i];
    }
}

// One workgroup per tile (more tiles per workgroup serialise their load
// latencies: 0.196 against 0.153 ms for the 50M-point frame's offsets).
#define DEC_FIX_TPW 1u
__global__ __launch_bounds__(CH_WG) void k_dec_fixed(const uint8_t* __restrict__ in, const DecFixed* __restrict__ jobs,
                                                     uint32_t njobs, uint32_t ntiles, uint32_t* idrel, void* diff,
                                                     void* keep, uint32_t dk32, uint8_t* choice, void* rel,
                                                     uint32_t rel8, uint32_t* err) {
    __shared__ uint32_t s_tile0[DEC_FIX_LDS];
    __shared__ __attribute__((aligned(16))) uint8_t s_out[DEC_FIX_BRECS * CH_MAXD + 8];
    const bool lds = njobs <= DEC_FIX_LDS;
    if (lds) {
        for (uint32_t k = threadIdx.x; k < njobs; k += CH_WG) s_tile0[k] = jobs[k].tile0;
        __syncthreads();
    }
    for (uint32_t u = 0; u < DEC_FIX_TPW; ++u) {
        const uint32_t tile = blockIdx.x * DEC_FIX_TPW + u;
        if (tile >= ntiles) break;
        // the job holding the tile: every thread searches (uniform result)
        uint32_t lo = 0, hi = njobs;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((lds ? s_tile0[mid] : jobs[mid].tile0) <= tile) lo = mid; else hi = mid;
        }
        const DecFixed* Jg = jobs + __builtin_amdgcn_readfirstlane(lo);
        const DecFixed J = *Jg;
        const uint32_t t = tile - J.tile0;
        if (J.recs == DEC_FIX_BRECS) {
            // byte outputs whose values need no range check (decided by the host)
            fixed_tile_bytes(in, Jg, J.dims, J.n, J.dst, t, J.comp == 3 ? choice : static_cast<uint8_t*>(rel), s_out);
            __syncthreads();   // s_out is reused by the next tile
            continue;
        }
        switch (J.comp) {
            case 0: fixed_tile<uint32_t>(in, J, Jg, t, idrel, 0xffffffffull, DERR_VALUE, err); break;
            case 1:
            case 2: {
                void* dst = J.comp == 1 ? diff : keep;
                if (dk32) fixed_tile<uint32_t>(in, J, Jg, t, static_cast<uint32_t*>(dst), 0xffffffffull, DERR_L2, err);
                else fixed_tile<uint64_t>(in, J, Jg, t, static_cast<uint64_t*>(dst), ~0ull, 0, err);
                break;
            }
            case 3: fixed_tile<uint8_t>(in, J, Jg, t, choice, 1, DERR_VALUE, err); break;
            default:
                if (rel8) fixed_tile<uint8_t>(in, J, Jg, t, static_cast<uint8_t*>(rel), 0xffull, DERR_QUANT, err);
                else fixed_tile<uint32_t>(in, J, Jg, t, static_cast<uint32_t*>(rel), 0xffffffffull, DERR_VALUE, err);
                break;
        }
    }
}

// raw int64 components (LCP3's RAW1)
struct DecRaw {
    uint64_t byte_off, count, dst;
    int32_t comp;
    uint32_t pad;
    uint64_t first_thread;
};

__global__ __launch_bounds__(CH_WG) void k_dec_raw(const uint8_t* __restrict__ in, const DecRaw* __restrict__ jobs,
                                                   uint32_t njobs, uint64_t total, uint32_t* idrel, void* diff,
                                                   void* keep, uint32_t dk32, uint8_t* choice, void* rel, uint32_t rel8,
                                                   uint32_t* err) {
    for (uint64_t g = blockIdx.x * (uint64_t)CH_WG + threadIdx.x; g < total; g += (uint64_t)gridDim.x * CH_WG) {
        uint32_t lo = 0, hi = njobs;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (jobs[mid].first_thread <= g) lo = mid; else hi = mid;
        }
        const DecRaw J = jobs[lo];
        const uint64_t k = g - J.first_thread;
        const uint8_t* p = in + J.byte_off + 8 * k;
        uint64_t v = 0;
        for (int b = 0; b < 8; ++b) v |= (uint64_t)p[b] << (8 * b);
        if ((int64_t)v < 0) atomicOr(err, DERR_VALUE);
        store_value(J.comp, J.dst + k, v, idrel, diff, keep, dk32, choice, rel, rel8, err);
    }
}

// Tagged payload job (one component of one frame), decoded in tiles of DEC_TAG_TILE records
struct DecTagged {
    uint64_t byte_off;          // packed values in the input pool
    uint64_t bytes;             // packed bytes
    uint64_t n;                 // records
    uint64_t tag_off;           // first width in the width pool
    uint64_t dst;               // first value index
    uint32_t dims, tile0;       // first tile of this job
    int32_t comp;
    uint32_t maxw;              // the section's widest width (every tag is at most this)
    uint32_t recs;              // records per tile: DEC_TAG_TILE, or DEC_TAG_TILE2 for narrow jobs
    uint32_t pad;
};

// Tiles of DEC_TAG_TILE records, DEC_TAG_ITEMS consecutive records per
// thread (their bit offsets are one block scan). After the look-back the
// tile's payload bytes are staged in LDS with batched loads, and every value
// is cut from LDS.
#define DEC_TAG_TILE 1024u
#define DEC_TAG_ITEMS (DEC_TAG_TILE / CH_WG)
// Narrow jobs (maxw <= 32, maxw * dims <= 96 bits a record, dims <= 3) take tiles of
// DEC_TAG_TILE2 records: their payload (<= 24 KB) and their values staged as
// u32 (<= 24 KB) fit the same LDS (half the tiles, look-back steps and
// barriers per record).
#define DEC_TAG_TILE2 2048u
#define DEC_TAG_LDS (DEC_TAG_TILE * CH_MAXD * 63u / 8u + 16u)   // widest tile's bytes
// staging words: the widest tile's bytes plus alignment and the windows' two
// words past the end, or the tile's values as u64, whichever is more
#define DEC_TAG_WORDS ((DEC_TAG_TILE * CH_MAXD * 2u > DEC_TAG_LDS / 4u + 4u) ? DEC_TAG_TILE * CH_MAXD * 2u \
                                                                          : DEC_TAG_LDS / 4u + 4u)
// One tile of job J: ITEMS consecutive records per thread, values staged in
// LDS as VT before the coalesced stores.
template <uint32_t ITEMS, typename VT>
__device__ __forceinline__ void tagged_tile(const uint8_t* __restrict__ in, const DecTagged& J, uint32_t tile,
                                            const uint8_t* __restrict__ tags, uint64_t* status, uint32_t* s_tmp,
                                            uint64_t* s_excl, uint32_t* s_pay, uint32_t* idrel, void* diff,
                                            void* keep, uint32_t dk32, uint8_t* choice, void* rel, uint32_t rel8,
                                            uint32_t* err) {
    constexpr uint32_t RECS = ITEMS * CH_WG;
    const uint64_t rfirst = (uint64_t)(tile - J.tile0) * RECS;   // < J.n
    const uint64_t r0 = rfirst + (uint64_t)threadIdx.x * ITEMS;
    uint32_t tg[ITEMS];
#pragma unroll
    for (int k = 0; k < (int)ITEMS; ++k) tg[k] = tags[J.tag_off + (r0 + k < J.n ? r0 + k : rfirst)];
    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < (int)ITEMS; ++k) {
        if (r0 + k >= J.n) {
            tg[k] = 0;
            continue;
        }
        if (tg[k] < 1 || tg[k] > J.maxw) {
            atomicOr(err, DERR_VALUE);
            tg[k] = 1;   // keep the work bounded; the stream is rejected
        }
        bits += tg[k] * J.dims;
    }
    uint32_t tot;
    const uint32_t pre = block_excl_scan(bits, s_tmp, &tot);
    if (threadIdx.x < 64) {
        const uint64_t e = lookback_seg(status, tile, J.tile0, tot, err);
        if (threadIdx.x == 0) *s_excl = e;
    }
    __syncthreads();
    const uint64_t bit0 = *s_excl;
    const uint64_t B0 = bit0 >> 3;
    const uint64_t want = ((bit0 + tot + 7) >> 3) - B0;   // <= the staging bytes (widths <= maxw)
    const uint64_t have = J.bytes > B0 ? J.bytes - B0 : 0;
    const uint32_t nst = (uint32_t)(want < have ? want : have);
    if (want > have) {
        if (threadIdx.x == 0) atomicOr(err, DERR_VALUE);
    }
    // Stage payload bytes [B0, B0 + nst) as aligned words: staged byte
    // a + i is payload byte B0 + i. The last word may reach up to three bytes
    // past the payload, still inside its container (a payload is always
    // followed by the correction count); bits past the tile are never used.
    const uint64_t gb = J.byte_off + B0;
    const uint32_t a = (uint32_t)(gb & 3u);
    const uint32_t* gw = reinterpret_cast<const uint32_t*>(in + (gb - a));
    const uint32_t nw = (a + nst + 3) / 4;
    for (uint32_t q0 = 0; q0 < nw; q0 += CH_WG * 8) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t q = q0 + j * CH_WG + threadIdx.x;
            v[j] = gw[q < nw ? q : 0u];   // unconditional (the tile's first word)
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t q = q0 + j * CH_WG + threadIdx.x;
            if (q < nw) s_pay[q] = v[j];
        }
    }
    if (threadIdx.x < 2) s_pay[nw + threadIdx.x] = 0;   // read by the last windows
    __syncthreads();
    // Cut each value from a big-endian window of two staged words (a third
    // for a value that reaches past them): stream bit o is bit 7 - (o & 7)
    // of staged byte o >> 3, and the words hold their bytes little-endian.
    uint32_t pos = 8 * a + (uint32_t)(bit0 & 7) + pre;
    const uint32_t dims = J.dims;
    VT val[ITEMS][CH_MAXD];
#pragma unroll
    for (int k = 0; k < (int)ITEMS; ++k) {
        const uint32_t w = tg[k];
#pragma unroll
        for (int d = 0; d < CH_MAXD; ++d) {
            val[k][d] = 0;
            if (!w || (uint32_t)d >= dims) continue;
            const uint32_t wi = pos >> 5, sh = pos & 31u;
            const uint64_t x = ((uint64_t)__builtin_bswap32(s_pay[wi]) << 32) | __builtin_bswap32(s_pay[wi + 1]);
            uint64_t v = (x << sh) >> (64u - w);
            if (sh + w > 64u) v |= (uint64_t)(__builtin_bswap32(s_pay[wi + 2]) >> (96u - sh - w));
            val[k][d] = (VT)v;   // below 2^maxw: fits VT
            pos += w;
        }
    }
    // the tile's values in LDS (the staged payload is no longer needed), then
    // written in index order by consecutive threads
    __syncthreads();
    VT* const s_val = reinterpret_cast<VT*>(s_pay);
    const uint32_t nrec = J.n - rfirst < RECS ? (uint32_t)(J.n - rfirst) : RECS;
#pragma unroll
    for (int k = 0; k < (int)ITEMS; ++k) {
        const uint32_t lr = threadIdx.x * ITEMS + k;
        if (lr >= nrec) continue;
#pragma unroll
        for (int d = 0; d < CH_MAXD; ++d)
            if ((uint32_t)d < dims) s_val[lr * dims + d] = val[k][d];
    }
    __syncthreads();
    const uint64_t g0 = J.dst + rfirst * dims;
    for (uint32_t e = threadIdx.x; e < nrec * dims; e += CH_WG)
        store_value(J.comp, g0 + e, (uint64_t)s_val[e], idrel, diff, keep, dk32, choice, rel, rel8, err);
}

__global__ __launch_bounds__(CH_WG) void k_dec_tagged(const uint8_t* __restrict__ in,
                                                      const DecTagged* __restrict__ jobs, uint32_t njobs,
                                                      const uint8_t* __restrict__ tags, uint64_t* status,
                                                      uint32_t* ticket, uint32_t* idrel, void* diff, void* keep,
                                                      uint32_t dk32, uint8_t* choice, void* rel, uint32_t rel8,
                                                      uint32_t* err) {
    __shared__ uint32_t s_tile, s_tmp[4];
    __shared__ uint64_t s_excl;
    // staged payload words, then the tile's decoded values
    __shared__ uint32_t s_pay[DEC_TAG_WORDS];
    static_assert(DEC_TAG_TILE2 * 3 * 4 / 4 <= DEC_TAG_WORDS && DEC_TAG_TILE2 * 96 / 32 + 4 <= DEC_TAG_WORDS,
                  "a narrow tile's payload and values fit the staging words");
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].tile0 <= tile) lo = mid; else hi = mid;
    }
    const DecTagged J = jobs[lo];
    if (J.recs == DEC_TAG_TILE2)
        tagged_tile<DEC_TAG_TILE2 / CH_WG, uint32_t>(in, J, tile, tags, status, s_tmp, &s_excl, s_pay, idrel, diff,
                                                     keep, dk32, choice, rel, rel8, err);
    else
        tagged_tile<DEC_TAG_ITEMS, uint64_t>(in, J, tile, tags, status, s_tmp, &s_excl, s_pay, idrel, diff, keep,
                                             dk32, choice, rel, rel8, err);
}

// ------------------------------------------------------------------------
// per-tile counts: zeros of choice2 (record tiles) and explicit points (point tiles)

__global__ __launch_bounds__(CH_WG) void k_dec_tile_zero_counts(const uint8_t* __restrict__ v8,
                                                                const uint32_t* __restrict__ v32,
                                                                const uint32_t* __restrict__ tile_begin,
                                                                const uint64_t* __restrict__ seg_base,
                                                                const uint64_t* __restrict__ seg_n, uint32_t F,
                                                                uint32_t* __restrict__ counts) {
    const uint32_t tile = blockIdx.x;
    const uint32_t f = upper_find(tile_begin, F, tile);
    const uint64_t r0 = (uint64_t)(tile - tile_begin[f]) * CH_TILE;
    uint32_t c = 0;
    const uint64_t nf = seg_n[f], bf = seg_base[f];
    uint32_t xv[CH_ITEMS];
#pragma unroll
    for (int j = 0; j < CH_ITEMS; ++j) {   // unconditional loads (the tile's first entry past the end)
        const uint64_t r = r0 + (uint64_t)j * CH_WG + threadIdx.x;
        const uint64_t rc = bf + (r < nf ? r : r0);
        xv[j] = v8 ? (uint32_t)v8[rc] : v32[rc];
    }
#pragma unroll
    for (int j = 0; j < CH_ITEMS; ++j) {
        const uint64_t r = r0 + (uint64_t)j * CH_WG + threadIdx.x;
        c += r < nf && xv[j] == 0;
    }
    __shared__ uint32_t s_tmp[4];
    uint32_t tot;
    (void)block_excl_scan(c, s_tmp, &tot);
    if (threadIdx.x == 0) counts[tile] = tot;
}

// ------------------------------------------------------------------------
// L2 inverse: segmented modular scan over each frame's records.
// element = (reset, value): keep records reset to keep - wrap_min, diff
// records add (-correction) mod range. (r1,v1)+(r2,v2) = (r1|r2, r2 ? v2 : v1+v2 mod R)

// V is the value type of the scan: u32 when every range of the launch is at
// most 2^31 (sums of two values stay below 2^32), else u64 (ranges < 2^63)
template <typename V>
struct Agg {
    uint32_t reset;
    V v[CH_MAXD];
};

template <typename V>
__device__ inline V addmod(V a, V b, V R) {
    // a, b < R
    const V s = a + b;
    return s >= R ? s - R : s;
}

template <int D, typename V>
__device__ inline void agg_combine(Agg<V>& left, const Agg<V>& right, V R) {   // left = left (+) right
    if (right.reset) {
        left = right;
    } else {
#pragma unroll
        for (int d = 0; d < D; ++d) left.v[d] = addmod(left.v[d], right.v[d], R);
    }
}

// shuffles of a scan element (reset flag and D values)
template <int D, typename V>
__device__ inline void agg_shfl_up(uint32_t& r, V* v, uint32_t off) {
    r = __shfl_up(r, off, 64);
#pragma unroll
    for (int d = 0; d < D; ++d) v[d] = __shfl_up(v[d], off, 64);
}

template <int D, typename V>
__device__ inline void agg_shfl_down(uint32_t& r, V* v, uint32_t off) {
    r = __shfl_down(r, off, 64);
#pragma unroll
    for (int d = 0; d < D; ++d) v[d] = __shfl_down(v[d], off, 64);
}

// left = left (+) right
template <int D, typename V>
__device__ inline void agg_comb(uint32_t& lr, V* lv, uint32_t rr, const V* rv, V R) {
    if (rr) {
        lr = 1;
#pragma unroll
        for (int d = 0; d < D; ++d) lv[d] = rv[d];
    } else {
#pragma unroll
        for (int d = 0; d < D; ++d) lv[d] = addmod(lv[d], rv[d], R);
    }
}

// One DPP step of the wavefront scan of scan elements: every lane combines
// the element CTRL names (an earlier lane; the identity (0, 0) where there is
// none or its row is masked off) with its own, earlier on the left.
template <int CTRL, int ROW_MASK, typename V>
__device__ inline V dpp_val(V x) {
    if constexpr (sizeof(V) == 4) {
        return (V)dpp_u32<CTRL, ROW_MASK>(0u, (uint32_t)x);
    } else {
        const uint32_t lo = dpp_u32<CTRL, ROW_MASK>(0u, (uint32_t)x);
        const uint32_t hi = dpp_u32<CTRL, ROW_MASK>(0u, (uint32_t)((uint64_t)x >> 32));
        return (V)(((uint64_t)hi << 32) | lo);
    }
}
template <int CTRL, int ROW_MASK, int D, typename V>
__device__ inline void agg_dpp_step(uint32_t& r, V* v, V R) {
    uint32_t er = dpp_u32<CTRL, ROW_MASK>(0u, r);
    V ev[D];
#pragma unroll
    for (int d = 0; d < D; ++d) ev[d] = dpp_val<CTRL, ROW_MASK, V>(v[d]);
    agg_comb<D, V>(er, ev, r, v, R);   // earlier (+) mine
    r = er;
#pragma unroll
    for (int d = 0; d < D; ++d) v[d] = ev[d];
}

// Wave 0 of an L2 tile: publishes the tile aggregate, looks back over the
// earlier tiles of the frame (64 predecessors at a time: lane l reads tile
// t - l; the window up to the nearest inclusive or resetting aggregate is
// combined in order, farther tiles on the left), publishes the inclusive
// prefix and returns the exclusive one in every lane.
// flags[t]: 1 = aggregate, 2 = inclusive prefix, bit 8 = the value resets.
template <int D, typename V>
__device__ inline Agg<V> l2_lookback(uint32_t tile, uint32_t seg, const Agg<V>& tot, V R, uint32_t* flags,
                                     uint64_t* aggv, uint32_t* err) {
    const uint32_t lane = threadIdx.x & 63u;
    Agg<V> incoming;
    incoming.reset = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) incoming.v[d] = 0;
    if (tile != seg) {
        if (lane == 0) {
            for (int d = 0; d < D; ++d) aggv[((uint64_t)tile * 2 + 0) * CH_MAXD + d] = tot.v[d];
            __hip_atomic_store(&flags[tile], 1u | (tot.reset << 8), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        // look-back over 64 predecessors at a time: lane l reads tile
        // t - l; the window up to the nearest inclusive (or resetting)
        // aggregate is combined in order, farther tiles on the left
        uint32_t ar = 0;
        V av[D];
#pragma unroll
        for (int d = 0; d < D; ++d) av[d] = 0;
        int64_t t = (int64_t)tile - 1;
        uint32_t spins = 0;
        while (t >= (int64_t)seg) {
            const int64_t mine = t - (int64_t)lane;
            const bool beyond = mine < (int64_t)seg;
            const uint32_t fl =
                beyond ? 2u : __hip_atomic_load(&flags[mine], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t st = fl & 0xffu;
            const uint64_t not_ready = __ballot(!beyond && st == 0);
            const uint64_t stop = __ballot(st == 2 || (fl >> 8) != 0);
            const uint64_t first_stop = stop ? (stop & (~stop + 1)) : 0ull;
            const uint64_t need = first_stop ? (first_stop | (first_stop - 1)) : ~0ull;
            if (not_ready & need) {
                if (++spins > LB_SPIN_LIMIT) {
                    if (lane == 0) atomicOr(err, DERR_WAIT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            uint32_t yr = 0;
            V yv[D];
#pragma unroll
            for (int d = 0; d < D; ++d) yv[d] = 0;
            if (((need >> lane) & 1ull) && !beyond) {
                yr = fl >> 8;
                const uint64_t slot = st == 2 ? 1 : 0;
#pragma unroll
                for (int d = 0; d < D; ++d) yv[d] = (V)aggv[((uint64_t)mine * 2 + slot) * CH_MAXD + d];
            }
            for (uint32_t off = 1; off < 64; off <<= 1) {
                uint32_t orr = yr;
                V ov[D];
#pragma unroll
                for (int d = 0; d < D; ++d) ov[d] = yv[d];
                agg_shfl_down<D, V>(orr, ov, off);
                if (lane + off < 64 && (lane & (2 * off - 1)) == 0) {
                    agg_comb<D, V>(orr, ov, yr, yv, R);   // farther (+) nearer
                    yr = orr;
#pragma unroll
                    for (int d = 0; d < D; ++d) yv[d] = ov[d];
                }
            }
            uint32_t wr = __shfl(yr, 0, 64);
            V wv[D];
#pragma unroll
            for (int d = 0; d < D; ++d) wv[d] = __shfl(yv[d], 0, 64);
            agg_comb<D, V>(wr, wv, ar, av, R);   // window (+) what was nearer
            ar = wr;
#pragma unroll
            for (int d = 0; d < D; ++d) av[d] = wv[d];
            if (first_stop) break;
            t -= 64;
        }
        incoming.reset = ar;
#pragma unroll
        for (int d = 0; d < D; ++d) incoming.v[d] = av[d];
    }
    if (lane == 0) {
        Agg<V> inc = incoming;
        agg_combine<D, V>(inc, tot, R);
        for (int d = 0; d < D; ++d) aggv[((uint64_t)tile * 2 + 1) * CH_MAXD + d] = inc.v[d];
        __hip_atomic_store(&flags[tile], 2u | (inc.reset << 8), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    incoming.reset = __shfl(incoming.reset, 0, 64);
#pragma unroll
    for (int d = 0; d < D; ++d) incoming.v[d] = __shfl(incoming.v[d], 0, 64);
    return incoming;
}

// L2 inverse of one tile of CH_TILE records, blocked: every record's element
// is built from coalesced loads in striped order (record k*CH_WG + t by thread
// t) and stored in LDS; thread t then scans the CH_ITEMS consecutive records
// [CH_ITEMS*t, CH_ITEMS*t + CH_ITEMS) in registers, the threads' aggregates
// are scanned with shuffles (then over the wavefronts, then across tiles by
// look-back), and the tile's records leave LDS as one coalesced run of
// n*D words. LDS rows are padded by one word per CH_ITEMS records, so the
// blocked reads (stride CH_ITEMS+1 words) hit distinct banks. The reset flags
// are kept as one ballot word per 64 records, which holds the kernel's LDS at
// 53 KB for three dimensions of 32-bit values (three workgroups per CU).
#define L2_PAD(r) ((r) + ((r) / CH_ITEMS))
#define L2_LDS_WORDS (CH_TILE + CH_TILE / CH_ITEMS)
template <int D, typename V, typename DKT>
__global__ __launch_bounds__(CH_WG) void k_dec_l2(const DecFrame* __restrict__ fr, uint32_t F,
                                                  const uint32_t* __restrict__ rtile_begin,
                                                  const uint8_t* __restrict__ choice, const DKT* __restrict__ diff,
                                                  const DKT* __restrict__ keep,
                                                  const uint32_t* __restrict__ tile_zero_off,
                                                  uint32_t* flags, uint64_t* aggv, uint32_t* ticket,
                                                  DKT* __restrict__ compact, uint32_t* err) {
    constexpr uint32_t W = CH_WG / 64;
    constexpr uint32_t NS = CH_ITEMS * W;   // (row, wave) segments of the striped order
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_zoff[NS];         // keep records before each segment
    __shared__ V s_v[D][L2_LDS_WORDS];
    __shared__ uint64_t s_rsb[CH_TILE / 64];   // reset flags, bit r % 64 of word r / 64
    static_assert(64 % CH_ITEMS == 0 && CH_WG % 64 == 0, "a thread's records share one flag word");
    __shared__ uint32_t s_wr[W];
    __shared__ V s_wv[W][D];
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t f = upper_find(rtile_begin, F, tile);
    const DecFrame& Fr = fr[f];
    const V R = (V)Fr.range;
    const uint64_t t0 = (uint64_t)(tile - rtile_begin[f]) * CH_TILE;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    // keep records (choice 0): ballots per (row, wave) segment, in order
    uint32_t zbits = 0;
    uint32_t zpre[CH_ITEMS];
    {
        uint8_t chv[CH_ITEMS];
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k) {   // loads first (a tile holds at least one record)
            const uint64_t r = t0 + k * CH_WG + threadIdx.x;
            chv[k] = choice[Fr.rec_base + (r < Fr.nrec ? r : t0)];
        }
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k) {
            const uint64_t r = t0 + k * CH_WG + threadIdx.x;
            const bool z = r < Fr.nrec && chv[k] == 0;
            const uint64_t m = __ballot(z);
            zpre[k] = (uint32_t)__popcll(m & lt);
            zbits |= (z ? 1u : 0u) << k;
            if (lane == 0) s_zoff[k * W + wave] = __popcll(m);
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t v = s_zoff[threadIdx.x];
        s_zoff[threadIdx.x] = wave_incl_scan(v) - v;
    }
    __syncthreads();
    const uint64_t kbase = tile_zero_off[tile];
    // striped: elements into LDS. keep -> (reset, value - wrap_min),
    // diff -> (0, delta mod R); past the frame's end the identity (0, 0)
#define L2_GROUP 4
#pragma unroll
    for (int kg = 0; kg < CH_ITEMS; kg += L2_GROUP) {
        uint64_t raw[L2_GROUP][D];
#pragma unroll
        for (int j = 0; j < L2_GROUP; ++j) {   // the group's loads before any use
            const int k = kg + j;
            const uint64_t r = t0 + k * CH_WG + threadIdx.x;
            const uint64_t kix = kbase + s_zoff[k * W + wave] + zpre[k];   // global keep index (or keeps before)
            const bool z = (zbits >> k) & 1u;
            const uint64_t idx = z ? kix : Fr.rec_base + r - kix;
            const bool inb = r < Fr.nrec && (z ? kix >= Fr.keep_base && kix < Fr.keep_base + Fr.nkeep
                                               : idx >= Fr.diff_base && idx < Fr.diff_base + Fr.ndiff);
            const DKT* src = z ? keep : diff;
            const uint64_t ic = inb ? idx : 0ull;   // record 0 of the array always exists
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const uint64_t v = (uint64_t)src[ic * D + d];
                raw[j][d] = inb ? v : (z ? (uint64_t)Fr.wrap_min : 0ull);
            }
        }
#pragma unroll
        for (int j = 0; j < L2_GROUP; ++j) {
            const int k = kg + j;
            const uint32_t rl = k * CH_WG + threadIdx.x;
            const uint64_t r = t0 + rl;
            uint32_t rs = 0;
            V v[D];
#pragma unroll
            for (int d = 0; d < D; ++d) v[d] = 0;
            if (r < Fr.nrec) {
                const uint64_t kix = kbase + s_zoff[k * W + wave] + zpre[k];
                if ((zbits >> k) & 1u) {
                    if (kix < Fr.keep_base || kix >= Fr.keep_base + Fr.nkeep) atomicOr(err, DERR_L2);
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const uint64_t kv = raw[j][d];
                        if ((int64_t)kv < Fr.wrap_min || (int64_t)kv > Fr.wrap_max) atomicOr(err, DERR_L2);
                        v[d] = (V)(kv - (uint64_t)Fr.wrap_min);
                    }
                    rs = 1;
                } else {
                    if (r == 0) atomicOr(err, DERR_L2);
                    const uint64_t didx = Fr.rec_base + r - kix;   // global diff index
                    if (didx < Fr.diff_base || didx >= Fr.diff_base + Fr.ndiff) atomicOr(err, DERR_L2);
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const uint64_t sym = raw[j][d];
                        const int64_t corr = (sym & 1u) ? -(int64_t)(sym >> 1) - 1 : (int64_t)(sym >> 1);
                        if ((int64_t)sym < 0 || corr < Fr.min_corr || corr > Fr.max_corr) atomicOr(err, DERR_L2);
                        const uint64_t delta = corr > 0 ? (uint64_t)R - (uint64_t)corr : (uint64_t)(-corr);
                        v[d] = (V)(delta < (uint64_t)R ? delta : delta % (uint64_t)R);
                    }
                }
            }
            const uint32_t pos = L2_PAD(rl);
#pragma unroll
            for (int d = 0; d < D; ++d) s_v[d][pos] = v[d];
            const uint64_t rb = __ballot(rs != 0);   // records k*CH_WG + 64*wave + lane
            if (lane == 0) s_rsb[k * W + wave] = rb;
        }
    }
    __syncthreads();
    // blocked: inclusive scan of the thread's records in registers; rbits
    // marks the records at or after the thread's first reset
    const uint32_t pbase = L2_PAD(threadIdx.x * CH_ITEMS);
    // the thread's CH_ITEMS records share one flag word
    const uint32_t rsm = (uint32_t)(s_rsb[(threadIdx.x * CH_ITEMS) / 64] >> ((threadIdx.x * CH_ITEMS) % 64));
    V iv[CH_ITEMS][D];
    uint32_t ar = 0, rbits = 0;
    V av[D];
#pragma unroll
    for (int d = 0; d < D; ++d) av[d] = 0;
#pragma unroll
    for (int j = 0; j < CH_ITEMS; ++j) {
        const uint32_t rs = (rsm >> j) & 1u;
        V v[D];
#pragma unroll
        for (int d = 0; d < D; ++d) v[d] = s_v[d][pbase + j];
        agg_comb<D, V>(ar, av, rs, v, R);
        rbits |= ar << j;
#pragma unroll
        for (int d = 0; d < D; ++d) iv[j][d] = av[d];
    }
    // exclusive scan of the threads' aggregates within the wavefront
    uint32_t xr = ar;
    V xv[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xv[d] = av[d];
    // inclusive segmented scan: inside each row of 16 lanes, then rows 1 and
    // 3 take the end of the row before, rows 2 and 3 the end of row 1
    agg_dpp_step<DPP_ROW_SHR(1), 0xf, D, V>(xr, xv, R);
    agg_dpp_step<DPP_ROW_SHR(2), 0xf, D, V>(xr, xv, R);
    agg_dpp_step<DPP_ROW_SHR(4), 0xf, D, V>(xr, xv, R);
    agg_dpp_step<DPP_ROW_SHR(8), 0xf, D, V>(xr, xv, R);
    agg_dpp_step<DPP_ROW_BCAST15, 0xa, D, V>(xr, xv, R);
    agg_dpp_step<DPP_ROW_BCAST31, 0xc, D, V>(xr, xv, R);
    if (lane == 63) {
        s_wr[wave] = xr;
#pragma unroll
        for (int d = 0; d < D; ++d) s_wv[wave][d] = xv[d];
    }
    // inclusive of lane - 1 = exclusive of lane (wave_shr:1; lane 0 gets (0, 0))
    xr = dpp_u32<0x138, 0xf>(0u, xr);
#pragma unroll
    for (int d = 0; d < D; ++d) xv[d] = dpp_val<0x138, 0xf, V>(xv[d]);
    if (lane == 0) {
        xr = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) xv[d] = 0;
    }
    __syncthreads();
    // wave 0: the tile aggregate (the wavefronts in order), the look-back,
    // then every wavefront's exclusive prefix
    if (threadIdx.x < 64) {
        Agg<V> tot;
        tot.reset = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) tot.v[d] = 0;
        Agg<V> wagg[W];
#pragma unroll
        for (uint32_t w = 0; w < W; ++w) {
            wagg[w].reset = s_wr[w];
#pragma unroll
            for (int d = 0; d < D; ++d) wagg[w].v[d] = s_wv[w][d];
            agg_combine<D, V>(tot, wagg[w], R);
        }
        Agg<V> run = l2_lookback<D, V>(tile, rtile_begin[f], tot, R, flags, aggv, err);
        if (lane == 0) {
#pragma unroll
            for (uint32_t w = 0; w < W; ++w) {
                s_wr[w] = run.reset;
#pragma unroll
                for (int d = 0; d < D; ++d) s_wv[w][d] = run.v[d];
                agg_combine<D, V>(run, wagg[w], R);
            }
        }
    }
    __syncthreads();
    // the thread's prefix = wavefront prefix (+) exclusive of the lane; a
    // record at or after the thread's first reset keeps its own inclusive value
    uint32_t pr = s_wr[wave];
    V pv[D];
#pragma unroll
    for (int d = 0; d < D; ++d) pv[d] = s_wv[wave][d];
    agg_comb<D, V>(pr, pv, xr, xv, R);
#pragma unroll
    for (int j = 0; j < CH_ITEMS; ++j) {
#pragma unroll
        for (int d = 0; d < D; ++d)
            s_v[d][pbase + j] = ((rbits >> j) & 1u) ? iv[j][d] : addmod(pv[d], iv[j][d], R);
    }
    __syncthreads();
    // one coalesced run of the tile's records
    const uint64_t left = Fr.nrec - t0;
    const uint32_t nt = left < CH_TILE ? (uint32_t)left : CH_TILE;
    DKT* out = compact + (Fr.rec_base + t0) * D;   // 32-bit: every block id below 2^31
    for (uint32_t e = threadIdx.x; e < nt * D; e += CH_WG) {
        const uint32_t r = e / D, d = e - r * D;
        out[e] = (DKT)((uint64_t)Fr.wrap_min + (uint64_t)s_v[d][L2_PAD(r)]);
    }
}

// ------------------------------------------------------------------------
// Chain resolution: the explicit record of every point (the backward chain
// pass of serial.cpp:1277-1302, in parallel). Every successor lies at a larger
// index, so a point's record is its successor's, down to an explicit point.
//
// In a tile of 4096 points, wavefront w owns the 1024 consecutive points
// [1024 w, 1024 w + 1024) as 16 windows of 64 and walks them from the last to
// the first: a successor in a later window of its own range is already
// resolved in LDS (same wavefront, program order), one in the same window is
// resolved by pointer jumping between lanes with shuffles (no barrier), one
// in a later wavefront's range is left pending and resolved after one barrier
// (at most three hops, every hop into a later range). A chain that leaves the
// tile ends at its exit, the tile's last point on it: the exit keeps
// CH_EXT | its successor and joins the open list, the points that reach it
// keep CH_EXT | the exit. k_dec_chain_resolve then settles the exits, and
// dequantisation reads a follower's value at its exit.
#define CH_EXT 0x80000000u
// in-tile states: resolved ordinal (< 4096), exit at local e, pending on local t,
// in-window pointer to lane l
#define CS_EXIT 0x10000u
#define CS_PEND 0x20000u
#define CS_WIN 0x40000u
// Occupancy: held to 80 VGPRs for six waves per SIMD (81 unconstrained, five
// waves; no spill at 80): 0.219 vs 0.237 ms on a 50M-point frame.
template <int D>
__global__ __launch_bounds__(CH_WG) __attribute__((amdgpu_waves_per_eu(6))) void k_dec_chain_local(const DecFrame* __restrict__ fr, uint32_t F,
                                                           const uint32_t* __restrict__ tile_begin,
                                                           const uint32_t* __restrict__ idrel,
                                                           uint32_t* __restrict__ tile_expl_off, uint64_t* status,
                                                           uint32_t* ticket, uint64_t* __restrict__ expl_total,
                                                           uint32_t* __restrict__ res, uint32_t* __restrict__ open_list,
                                                           uint32_t* __restrict__ open_count, uint32_t* err) {
    constexpr uint32_t W = CH_WG / 64;
    constexpr uint32_t WIN = CH_TILE / CH_WG;   // windows per wavefront (16)
    __shared__ uint32_t s_val[CH_TILE];
    __shared__ uint32_t s_wtot[W], s_tile, s_rank0, s_base;
    // tiles are taken by ticket, so every tile a look-back waits on has started
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t f = upper_find(tile_begin, F, tile);
    const DecFrame& Fr = fr[f];
    const uint64_t p0 = (uint64_t)(tile - tile_begin[f]) * CH_TILE;
    const uint64_t p1 = p0 + CH_TILE < Fr.n ? p0 + CH_TILE : Fr.n;
    const uint32_t cnt = (uint32_t)(p1 - p0);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    const uint32_t* src = idrel + Fr.pt_off + p0;
    const uint32_t wbase = w * (CH_TILE / W);
    // every load first (unconditional: a tile holds at least one point)
    uint32_t dv[WIN];
#pragma unroll
    for (int k = 0; k < WIN; ++k) {
        const uint32_t j = wbase + k * 64 + lane;
        dv[k] = src[j < cnt ? j : 0u];
    }
    // explicit ordinals in index order: ballots inside the window, running
    // counts over the wavefront's windows, wavefront offsets
    uint32_t ord[WIN], run = 0, ebits = 0;
#pragma unroll
    for (int k = 0; k < WIN; ++k) {
        const uint32_t j = wbase + k * 64 + lane;
        const bool ex = j < cnt && dv[k] == 0;
        const uint64_t m = __ballot(ex);
        ord[k] = run + (uint32_t)__popcll(m & lt);
        run += (uint32_t)__popcll(m);
        ebits |= (ex ? 1u : 0u) << k;
    }
    if (lane == 0) s_wtot[w] = run;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (uint32_t q = 0; q < W; ++q) {
        woff += q < w ? s_wtot[q] : 0u;
        tot += s_wtot[q];
    }
    // this tile's explicit count is published now; the ranks of the earlier
    // tiles are waited for only when the results are written
    if (threadIdx.x == 0) lb_store(&status[tile], (tile == 0 ? LB_INC : LB_AGG) | tot);
    // windows from the last to the first
    const uint32_t wend = wbase + CH_TILE / W;
    for (int k = WIN - 1; k >= 0; --k) {
        const uint32_t j = wbase + (uint32_t)k * 64 + lane;
        const uint32_t lo = wbase + (uint32_t)k * 64, hi = lo + 64;
        uint32_t v = 0;
        if (j < cnt) {
            if ((ebits >> k) & 1u) {
                v = woff + ord[k];
            } else {
                const uint64_t t = (uint64_t)j + dv[k];
                if (p0 + t >= Fr.n) {
                    atomicOr(err, DERR_CHAIN);   // a successor outside the frame: the stream is rejected
                    v = 0;
                } else if (t >= cnt) {
                    v = CS_EXIT | j;             // leaves the tile: j is an exit
                } else if (t >= wend) {
                    v = CS_PEND | (uint32_t)t;   // a later wavefront's range
                } else if (t >= hi) {
                    v = s_val[t];                // a later window of this wavefront: resolved
                } else {
                    v = CS_WIN | (uint32_t)(t - lo);
                }
            }
        }
        // pointer jumping inside the window: a lane pointing at a later lane
        // takes that lane's state until none points inside the window
        while (__ballot((v & CS_WIN) != 0)) {
            const uint32_t src_lane = (v & CS_WIN) ? (v & 63u) : lane;
            const uint32_t u = __shfl(v, (int)src_lane, 64);
            if (v & CS_WIN) v = u;
        }
        if (j < cnt) s_val[j] = v;
    }
    __syncthreads();
    // pending entries: follow into later ranges (resolved there, or pending
    // further on); nothing is written until every wavefront has read
    uint32_t fin[WIN];
#pragma unroll
    for (int k = 0; k < WIN; ++k) {
        const uint32_t j = wbase + k * 64 + lane;
        uint32_t v = j < cnt ? s_val[j] : 0u;
        while (v & CS_PEND) v = s_val[v & 0xffffu];
        fin[k] = v;
    }
    // the tile's first rank: explicit points of all earlier tiles (every
    // frame's, in order)
    if (threadIdx.x < 64) {
        const uint64_t excl = lookback(status, tile, tot, err);
        if (threadIdx.x == 0) {
            s_rank0 = (uint32_t)excl;
            tile_expl_off[tile] = (uint32_t)excl;
            if (tile + 1 == gridDim.x) *expl_total = excl + tot;
        }
    }
    __syncthreads();
    const uint32_t rank0 = s_rank0;
    const uint32_t gbase = (uint32_t)(Fr.pt_off + p0);
    // results; exits go to the open list, one reservation per tile
    uint32_t* dst = res + Fr.pt_off + p0;
    uint32_t xbits = 0, nx = 0;
#pragma unroll
    for (int k = 0; k < WIN; ++k) {
        const uint32_t j = wbase + k * 64 + lane;
        if (j >= cnt) continue;
        const uint32_t v = fin[k];
        uint32_t out;
        if (v & CS_EXIT) {
            const uint32_t e = v & 0xffffu;
            if (e == j) {
                out = CH_EXT | (gbase + j + dv[k]);   // the exit: its successor outside the tile
                xbits |= 1u << k;
                ++nx;
            } else {
                out = CH_EXT | (gbase + e);           // a follower: its exit
            }
        } else {
            out = rank0 + v;
        }
        dst[j] = out;
    }
    {
        const uint32_t inc = wave_incl_scan(nx);
        if (lane == 63) s_wtot[w] = inc;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t all = s_wtot[0] + s_wtot[1] + s_wtot[2] + s_wtot[3];
            s_base = all ? atomicAdd(open_count, all) : 0u;
        }
        __syncthreads();
        if (xbits) {
            uint32_t at = s_base + inc - nx;
            for (uint32_t q = 0; q < w; ++q) at += s_wtot[q];
#pragma unroll
            for (int k = 0; k < WIN; ++k)
                if ((xbits >> k) & 1u) open_list[at++] = gbase + wbase + k * 64 + lane;
        }
    }
}

// one round of global pointer jumping over the open points: every point of
// list_in takes its target's value; those still pointing out of their tile
// go to list_out (one reservation per 4096 list entries). Rounds after
// convergence find an empty list.
__global__ __launch_bounds__(CH_WG) void k_dec_chain_jump(uint32_t* __restrict__ res,
                                                          const uint32_t* __restrict__ list_in,
                                                          const uint32_t* __restrict__ count_in,
                                                          uint32_t* __restrict__ list_out,
                                                          uint32_t* __restrict__ count_out) {
    constexpr uint32_t W = CH_WG / 64;
    __shared__ uint32_t s_off[CH_ITEMS * W];
    __shared__ uint32_t s_base;
    const uint32_t n = *count_in;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    for (uint32_t t0 = blockIdx.x * CH_TILE; t0 < n; t0 += gridDim.x * CH_TILE) {
        uint32_t idx[CH_ITEMS], pre[CH_ITEMS], nvs[CH_ITEMS];
        uint32_t open_bits = 0;
        // the three dependent loads of every item, each level for all items
        // at once (past the list's end: entry t0, which exists)
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k) {
            const uint32_t kk = t0 + k * CH_WG + threadIdx.x;
            idx[k] = list_in[kk < n ? kk : t0];
        }
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k) nvs[k] = res[idx[k]];
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k)
            nvs[k] = __hip_atomic_load(&res[nvs[k] & ~CH_EXT], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int k = 0; k < CH_ITEMS; ++k) {
            const uint32_t kk = t0 + k * CH_WG + threadIdx.x;
            bool open = false;
            if (kk < n) {
                res[idx[k]] = nvs[k];
                open = (nvs[k] & CH_EXT) != 0;
            }
            const uint64_t m = __ballot(open);
            pre[k] = __popcll(m & lt);
            open_bits |= (open ? 1u : 0u) << k;
            if (lane == 0) s_off[k * W + wave] = __popcll(m);
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            const uint32_t v = threadIdx.x < CH_ITEMS * W ? s_off[threadIdx.x] : 0u;
            const uint32_t inc = wave_incl_scan(v);
            if (threadIdx.x < CH_ITEMS * W) s_off[threadIdx.x] = inc - v;
            const uint32_t total = __shfl(inc, 63, 64);
            if (threadIdx.x == 0) s_base = total ? atomicAdd(count_out, total) : 0u;
        }
        __syncthreads();
        if (open_bits) {
            const uint32_t base = s_base;
#pragma unroll
            for (int k = 0; k < CH_ITEMS; ++k)
                if (open_bits & (1u << k)) list_out[base + s_off[k * W + wave] + pre[k]] = idx[k];
        }
        __syncthreads();   // s_off / s_base are reused by the next tile
    }
}

// The exits of every tile, settled in one launch: an exit follows pointers
// (a follower's points to its exit, an exit's to its successor) until it
// reaches an explicit record, writing what it found back as it goes so that
// the exits behind it take the longer jumps. The list (in tile order) is
// taken from its end, so the exits nearest their records settle first and
// a walk from an earlier exit mostly finds its target settled. Every pointer
// goes to a larger index, so reading a value before or after another
// thread's write-back is equally valid and every walk ends; the bound only
// guards corrupt input.
#define CH_WALK_LIMIT (1u << 22)
__global__ __launch_bounds__(CH_WG) void k_dec_chain_resolve(uint32_t* __restrict__ res,
                                                             const uint32_t* __restrict__ list,
                                                             const uint32_t* __restrict__ count, uint32_t* err) {
    const uint32_t n = *count;
    for (uint32_t k = blockIdx.x * CH_WG + threadIdx.x; k < n; k += gridDim.x * CH_WG) {
        const uint32_t idx = list[n - 1 - k];
        uint32_t v = __hip_atomic_load(&res[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t steps = 0;
        while (v & CH_EXT) {
            const uint32_t u = __hip_atomic_load(&res[v & ~CH_EXT], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((u & CH_EXT) && (u & ~CH_EXT) <= (v & ~CH_EXT)) {   // never on valid input
                atomicOr(err, DERR_CHAIN);
                break;
            }
            v = u;
            __hip_atomic_store(&res[idx], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (++steps > CH_WALK_LIMIT) {
                atomicOr(err, DERR_CHAIN);
                break;
            }
        }
    }
}

// every frame: explicit points == records
__global__ void k_dec_check_counts(const DecFrame* __restrict__ fr, uint32_t F, const uint32_t* __restrict__ tb,
                                   const uint32_t* __restrict__ tile_expl_off, const uint64_t* __restrict__ expl_total,
                                   uint32_t* err) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const uint64_t a = tile_expl_off[tb[f]];
    const uint64_t b = f + 1 < F ? tile_expl_off[tb[f + 1]] : *expl_total;
    if (b - a != fr[f].nrec || a != fr[f].rec_base) atomicOr(err, DERR_COUNT);
}

// q = block * p + offset; x = fma(q + 0.5, 2 eb, min) (serial.cpp:1294-1311)
// One workgroup per point tile of a frame (the frame's fields are uniform).
// Each pass over DQ_ITEMS points per thread loads every root first, then
// every block and offset those roots name, then computes and stores.
#define DQ_ITEMS 16
#define DQ_STAGE 4
// BT: element type of the L2-inverted block ids (32-bit when every frame's
// wrap range lies in [0, 2^31))
template <int D, typename OutT, bool SOA, typename RelT, typename BT>
__global__ __launch_bounds__(CH_WG) void k_dec_dequant(const DecFrame* __restrict__ fr, uint32_t F,
                                                       const uint32_t* __restrict__ tile_begin,
                                                       const uint32_t* __restrict__ root,
                                                       const BT* __restrict__ compact,
                                                       const RelT* __restrict__ rel, OutT* __restrict__ out,
                                                       uint32_t* err) {
    const uint32_t tile = blockIdx.x;
    const uint32_t f = upper_find(tile_begin, F, tile);
    const DecFrame& Fr = fr[f];
    const uint64_t n = Fr.n, pt_off = Fr.pt_off, rec_base = Fr.rec_base, nrec = Fr.nrec, p = Fr.p;
    const uint64_t out_off = Fr.out_off;
    const double step = 2.0 * Fr.eb;
    double mins[D];
#pragma unroll
    for (int d = 0; d < D; ++d) mins[d] = Fr.mins[d];
    const bool lcg2 = Fr.lcg2 != 0;
    const float step_f = 2.0f * Fr.eb_f;
    float mins_f[D];
#pragma unroll
    for (int d = 0; d < D; ++d) mins_f[d] = Fr.mins_f[d];
    const uint64_t i0 = (uint64_t)(tile - tile_begin[f]) * CH_TILE;
    __shared__ OutT s_out[SOA ? 1 : DQ_STAGE * CH_WG * D];
    uint32_t bad = 0;
    for (int h = 0; h < CH_ITEMS; h += DQ_ITEMS) {
        uint32_t rv[DQ_ITEMS];
#pragma unroll
        for (int j = 0; j < DQ_ITEMS; ++j) {
            const uint64_t i = i0 + (uint64_t)(h + j) * CH_WG + threadIdx.x;
            rv[j] = root[pt_off + (i < n ? i : i0)];   // unconditional (the tile's first point)
        }
        {
            // a follower holds CH_EXT | the exit of its chain segment, whose
            // value is final after the jump rounds: take it (the others load
            // their own entry again, a cache hit)
            uint32_t rx[DQ_ITEMS];
#pragma unroll
            for (int j = 0; j < DQ_ITEMS; ++j) {
                const uint64_t i = i0 + (uint64_t)(h + j) * CH_WG + threadIdx.x;
                rx[j] = root[(rv[j] & CH_EXT) ? (uint64_t)(rv[j] & ~CH_EXT) : pt_off + (i < n ? i : i0)];
            }
#pragma unroll
            for (int j = 0; j < DQ_ITEMS; ++j) rv[j] = (rv[j] & CH_EXT) ? rx[j] : rv[j];
        }
        uint64_t blk[DQ_ITEMS][D];
        RelT off[DQ_ITEMS][D];
        uint32_t ok = 0;
#pragma unroll
        for (int j = 0; j < DQ_ITEMS; ++j) {
            const uint64_t i = i0 + (uint64_t)(h + j) * CH_WG + threadIdx.x;
            const uint64_t rec = rv[j];
            const bool live = i < n;
            const bool good = live && !(rv[j] & CH_EXT) && rec - rec_base < nrec;
            if (live && !good) bad |= DERR_CHAIN;
            ok |= good ? 1u << j : 0u;
            // unconditional loads: the frame's first record (the record
            // array holds one spare) and the tile's first point otherwise
            const uint64_t rc = good ? rec : rec_base, ic = live ? i : i0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                blk[j][d] = (uint64_t)compact[rc * D + d];
                off[j][d] = rel[(pt_off + ic) * D + d];
            }
        }
        // interleaved output is staged in LDS, DQ_STAGE items (contiguous
        // points) at a time, and written back element by element, so every
        // store instruction covers one contiguous span
#pragma unroll
        for (int hs = 0; hs < DQ_ITEMS; hs += DQ_STAGE) {
#pragma unroll
            for (int jj = 0; jj < DQ_STAGE; ++jj) {
                const int j = hs + jj;
                const uint64_t i = i0 + (uint64_t)(h + j) * CH_WG + threadIdx.x;
                const bool good = (ok >> j) & 1u;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const uint64_t o = off[j][d], b = blk[j][d];
                    // b * p + o must fit in int64
                    if (good && (o >= p || __umul64hi(b, p) != 0 || b * p > (uint64_t)INT64_MAX - o))
                        bad |= DERR_QUANT;
                    const int64_t q = (int64_t)(b * p + o);
                    double x;
                    if (lcg2) {
                        // cu/decompress/lcps_gpu_decompress_ans.cu:2605-2625
                        const float v = __builtin_fmaf((float)(int32_t)q + 0.5f, step_f, mins_f[d]);
                        if (good && (q > (int64_t)INT32_MAX || !isfinite(v))) bad |= DERR_QUANT;
                        x = (double)v;
                    } else {
                        x = fma((double)q + 0.5, step, mins[d]);
                    }
                    if (SOA) {
                        if (good) out[out_off + (uint64_t)d * n + i] = (OutT)x;
                    } else {
                        s_out[(jj * CH_WG + threadIdx.x) * D + d] = (OutT)x;
                    }
                }
            }
            if (!SOA) {
                const uint64_t p0 = i0 + (uint64_t)(h + hs) * CH_WG;
                __syncthreads();
                if (p0 < n) {
                    const uint32_t np = n - p0 < DQ_STAGE * CH_WG ? (uint32_t)(n - p0) : DQ_STAGE * CH_WG;
                    OutT* dst = out + out_off + p0 * D;
                    for (uint32_t e = threadIdx.x; e < np * D; e += CH_WG) dst[e] = s_out[e];
                }
                __syncthreads();
            }
        }
    }
    if (bad) atomicOr(err, bad);
}

// exact corrections: (u64 index, f64 value) records, ascending, finite
template <typename OutT, bool SOA>
__global__ __launch_bounds__(CH_WG) void k_dec_corr(const uint8_t* __restrict__ in, const DecFrame* __restrict__ fr,
                                                    uint32_t F, uint32_t D, OutT* __restrict__ out, uint32_t* err) {
    const uint32_t f = blockIdx.y;
    const DecFrame& Fr = fr[f];
    for (uint64_t k = blockIdx.x * (uint64_t)CH_WG + threadIdx.x; k < Fr.ncorr; k += (uint64_t)gridDim.x * CH_WG) {
        const uint8_t* p = in + Fr.corr_off + 16 * k;
        uint64_t idx = 0, bits = 0, prev = 0;
        for (int b = 0; b < 8; ++b) idx |= (uint64_t)p[b] << (8 * b);
        for (int b = 0; b < 8; ++b) bits |= (uint64_t)p[8 + b] << (8 * b);
        if (k) for (int b = 0; b < 8; ++b) prev |= (uint64_t)p[b - 16] << (8 * b);
        const double v = __longlong_as_double((long long)bits);
        if (idx >= Fr.n * D || (k && idx <= prev) || !isfinite(v)) { atomicOr(err, DERR_CORR); continue; }
        const uint64_t i = idx / D, d = idx % D;
        if (SOA) out[Fr.out_off + d * Fr.n + i] = (OutT)v;
        else out[Fr.out_off + idx] = (OutT)v;
    }
}

}  // namespace chain
