/*
 * chain_oracle.c -- CPU oracle for the Chain compressor (TEST INFRASTRUCTURE).
 *
 * Used only by tests/, the smoke check and bench.py's CPU-baseline leg, as the
 * checker. The GPU engine never links or calls this file.
 *
 * Plain C restatement of the reference CPU path. Each function cites the
 * reference lines it follows (paths relative to the reference repository).
 * Build: gcc -O2 -std=c11 -ffp-contract=off (no fast-math; the quantiser and
 * error-bound test depend on IEEE round-to-nearest binary64 arithmetic, and
 * the rANS normalisation on x87 80-bit long double exactly like the
 * reference's g++ build on x86-64).
 */
#define _GNU_SOURCE
#include "chain_oracle.h"

#include <fenv.h>
#include <math.h>
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zstd.h>

#if defined(__FAST_MATH__)
#error "the oracle requires IEEE arithmetic (no -ffast-math)"
#endif

/* ------------------------------------------------------------------------ */
/* Arena + error handling: every allocation of one API call lives in an arena
 * that is released when the call returns; failures longjmp back to the entry
 * point with a message. */

typedef union ahdr { union ahdr* prev; long double align_; max_align_t a2_; } ahdr;
typedef struct { jmp_buf jb; ahdr* top; } octx;

static _Thread_local char g_err[512];

const char* oc_last_error(void) { return g_err; }
void oc_free(void* p) { free(p); }

static void arena_release(octx* c) {
    while (c->top) { ahdr* p = c->top->prev; free(c->top); c->top = p; }
}

static void fail(octx* c, const char* fmt, ...) {
    va_list ap; va_start(ap, fmt); vsnprintf(g_err, sizeof g_err, fmt, ap); va_end(ap);
    longjmp(c->jb, 1);
}
#define REQUIRE(c, cond, ...) do { if (!(cond)) fail((c), __VA_ARGS__); } while (0)

static void* ax(octx* c, size_t sz) {
    ahdr* h = (ahdr*)calloc(1, sizeof(ahdr) + (sz ? sz : 1));
    if (!h) fail(c, "out of host memory");
    h->prev = c->top; c->top = h;
    return h + 1;
}
static void* dup_out(octx* c, const void* src, size_t sz) {
    void* p = malloc(sz ? sz : 1);
    if (!p) fail(c, "out of host memory");
    if (sz) memcpy(p, src, sz);
    return p;
}

/* growable byte buffer inside the arena */
typedef struct { octx* c; uint8_t* p; size_t n, cap; } bb;
static bb bb_make(octx* c) { bb b = {c, NULL, 0, 0}; return b; }
static void bb_need(bb* b, size_t k) {
    if (b->n + k <= b->cap) return;
    size_t nc = b->cap ? b->cap * 2 : 256;
    while (nc < b->n + k) nc *= 2;
    uint8_t* q = (uint8_t*)ax(b->c, nc);
    if (b->n) memcpy(q, b->p, b->n);
    b->p = q; b->cap = nc;
}
static void bb_put(bb* b, const void* s, size_t k) { bb_need(b, k); if (k) memcpy(b->p + b->n, s, k); b->n += k; }
static void bb_u8(bb* b, uint8_t v) { bb_put(b, &v, 1); }
static void bb_u32(bb* b, uint32_t v) { uint8_t t[4]; for (int i = 0; i < 4; ++i) t[i] = (uint8_t)(v >> (8 * i)); bb_put(b, t, 4); }
static void bb_u64(bb* b, uint64_t v) { uint8_t t[8]; for (int i = 0; i < 8; ++i) t[i] = (uint8_t)(v >> (8 * i)); bb_put(b, t, 8); }
static void bb_i64(bb* b, int64_t v) { bb_u64(b, (uint64_t)v); }
static void bb_f64(bb* b, double v) { uint64_t u; memcpy(&u, &v, 8); bb_u64(b, u); }
static void bb_varuint(bb* b, uint64_t v) {           /* lcps_cpp.cpp:92-100 */
    while (v >= 0x80u) { bb_u8(b, (uint8_t)((uint8_t)v | 0x80u)); v >>= 7; }
    bb_u8(b, (uint8_t)v);
}
static void bb_blob(bb* b, const bb* x) { bb_u64(b, x->n); bb_put(b, x->p, x->n); }

/* little-endian reader, lcps_cpp.cpp:1543-1604 */
typedef struct { octx* c; const uint8_t* p; size_t n; } rd;
static const uint8_t* rd_bytes(rd* r, size_t k) {
    REQUIRE(r->c, k <= r->n, "truncated LCPS bitstream");
    const uint8_t* q = r->p; r->p += k; r->n -= k; return q;
}
static uint8_t rd_u8(rd* r) { return *rd_bytes(r, 1); }
static uint32_t rd_u32(rd* r) { const uint8_t* q = rd_bytes(r, 4); uint32_t v = 0; for (int i = 0; i < 4; ++i) v |= (uint32_t)q[i] << (8 * i); return v; }
static uint64_t rd_u64(rd* r) { const uint8_t* q = rd_bytes(r, 8); uint64_t v = 0; for (int i = 0; i < 8; ++i) v |= (uint64_t)q[i] << (8 * i); return v; }
static int64_t rd_i64(rd* r) { return (int64_t)rd_u64(r); }
static double rd_f64(rd* r) { uint64_t u = rd_u64(r); double d; memcpy(&d, &u, 8); return d; }
static uint64_t rd_varuint(rd* r) {
    uint64_t v = 0;
    for (int shift = 0; shift <= 63; shift += 7) {
        uint8_t byte = rd_u8(r);
        REQUIRE(r->c, shift != 63 || (byte & 0xfeu) == 0, "LCPS varuint overflow");
        v |= (uint64_t)(byte & 0x7fu) << shift;
        if ((byte & 0x80u) == 0) return v;
    }
    fail(r->c, "invalid LCPS varuint");
    return 0;
}
static rd rd_blob(rd* r) { uint64_t k = rd_u64(r); rd s = {r->c, rd_bytes(r, (size_t)k), (size_t)k}; return s; }
static void rd_end(rd* r) { REQUIRE(r->c, r->n == 0, "unexpected trailing LCPS bytes"); }
static int rd_magic(rd* r, const char* m) { const uint8_t* q = rd_bytes(r, 4); return memcmp(q, m, 4) == 0; }

static size_t mul_chk(octx* c, size_t a, size_t b) {
    REQUIRE(c, b == 0 || a <= SIZE_MAX / b, "LCPS size overflow");
    return a * b;
}

/* lcps_cpp.cpp:475-482 */
static int bitlen(uint64_t v) { int b = 0; while (v) { ++b; v >>= 1; } return b; }

/* ------------------------------------------------------------------------ */
/* Quantisation helpers, lcps_cpp.cpp:1494-1523 */

static double reconstruct(int64_t q, double mn, double eb) {
    return fma((double)q + 0.5, 2.0 * eb, mn);
}

static int within_eb(double restored, double original, double eb) {
    if (!isfinite(restored)) return 0;
    const double difference = restored - original;
    const double magnitude = fabs(difference);
    if (magnitude < eb) return 1;
    if (!(magnitude == eb)) return 0;
    /* TwoSum remainder of restored + (-original) at the boundary */
    const double virtual_original = difference - restored;
    const double remainder = (restored - (difference - virtual_original)) + (-original - virtual_original);
    return difference >= 0.0 ? remainder <= 0.0 : remainder >= 0.0;
}

/* lcps_cpp.cpp:1221-1253 */
typedef struct { int64_t min_value, max_value, max_difference, min_correction, max_correction; } wrap_t;
static wrap_t make_wrap(octx* c, int64_t mn, int64_t mx) {
    REQUIRE(c, mn <= mx, "invalid block-ID wrap range");
    uint64_t difference;
    if (mn < 0 && mx >= 0) difference = (uint64_t)mx + ((uint64_t)(-(mn + 1)) + 1);
    else difference = (uint64_t)(mx - mn);
    REQUIRE(c, difference < (uint64_t)INT64_MAX, "block-ID range is too wide for wrapped differences");
    wrap_t w;
    w.min_value = mn; w.max_value = mx;
    w.max_difference = (int64_t)(difference + 1);
    w.max_correction = w.max_difference / 2;
    w.min_correction = -w.max_correction;
    if ((w.max_difference & 1) == 0) --w.max_correction;
    return w;
}

/* lcps_cpp.cpp:1255-1262 */
static uint64_t zigzag(int64_t v) {
    if (v >= 0) return (uint64_t)v << 1;
    return ((uint64_t)(-(v + 1)) << 1) | 1u;
}

static uint64_t absdiff_nn(octx* c, int64_t a, int64_t b) {   /* :1286-1296 */
    REQUIRE(c, a >= 0 && b >= 0, "block IDs must be non-negative");
    uint64_t x = (uint64_t)a, y = (uint64_t)b;
    return x >= y ? x - y : y - x;
}

/* ------------------------------------------------------------------------ */
/* L1 successor table: hash of the block tuple, lcps_cpp.cpp:2114-2138.
 * The reference keeps the last index per block in an unordered_map; this uses
 * open addressing with the same meaning (slot -> last index + 1). */

static uint64_t tuple_hash(const int64_t* t, uint32_t dims) {
    uint64_t h = 1469598103934665603ull;
    for (uint32_t d = 0; d < dims; ++d) {
        uint64_t y = (uint64_t)t[d];
        for (int i = 0; i < 8; ++i) { h ^= (y >> (8 * i)) & 0xffu; h *= 1099511628211ull; }
    }
    return h ^ (h >> 29);
}

static void l1_successors(octx* c, const int64_t* bid, uint64_t n, uint32_t dims, uint64_t* next) {
    uint64_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    uint64_t* slot = (uint64_t*)ax(c, cap * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) {
        const int64_t* t = bid + i * dims;
        uint64_t h = tuple_hash(t, dims) & (cap - 1);
        for (;;) {
            if (slot[h] == 0) { slot[h] = i + 1; break; }
            uint64_t j = slot[h] - 1;
            if (memcmp(bid + j * dims, t, dims * sizeof(int64_t)) == 0) {
                next[j] = i; slot[h] = i + 1; break;
            }
            h = (h + 1) & (cap - 1);
        }
    }
}

/* ------------------------------------------------------------------------ */

int oc_compress_streams(const double* pts, uint64_t n, uint32_t dims,
                        double eb, int64_t p, int64_t thr,
                        int bl, int l1, int l2, oc_streams* out) {
    octx cx; cx.top = NULL; octx* c = &cx;
    memset(out, 0, sizeof *out);
    if (setjmp(cx.jb)) { arena_release(c); oc_streams_free(out); return -1; }

    /* lcps_cpp.cpp:2028-2065 */
    REQUIRE(c, fegetround() == FE_TONEAREST, "LCP2 encoding requires round-to-nearest floating-point mode");
    REQUIRE(c, eb > 0.0 && isfinite(2.0 * eb), "eb must be positive with finite 2*eb");
    REQUIRE(c, p > 0, "p must be positive");
    const int64_t pe = bl ? p : 1;
    REQUIRE(c, n > 0, "zero-size array to reduction operation minimum which has no identity");
    REQUIRE(c, dims > 0, "points must have at least one dimension");
    (void)mul_chk(c, (size_t)n, dims);
    const double denom = 2.0 * eb;

    /* minima, :2072-2079 (std::min keeps the earlier of two equal values) */
    double* mins = (double*)ax(c, dims * sizeof(double));
    for (uint32_t d = 0; d < dims; ++d) mins[d] = INFINITY;
    for (uint64_t i = 0; i < n; ++i)
        for (uint32_t d = 0; d < dims; ++d) {
            const double x = pts[i * dims + d];
            REQUIRE(c, isfinite(x), "LCPS coordinates must be finite");
            if (x < mins[d]) mins[d] = x;
        }

    /* quantise, split, corrections, :2091-2110 */
    int64_t* bid = (int64_t*)ax(c, n * dims * sizeof(int64_t));
    int64_t* rel = (int64_t*)ax(c, n * dims * sizeof(int64_t));
    bb corr_i = bb_make(c), corr_v = bb_make(c);
    uint64_t ncorr = 0;
    int64_t max_bid = 0;
    for (uint64_t i = 0; i < n; ++i)
        for (uint32_t d = 0; d < dims; ++d) {
            const double original = pts[i * dims + d];
            const double qv = floor((original - mins[d]) / denom);
            REQUIRE(c, qv >= 0.0 && qv < 0x1p63, "quantized coordinate does not fit int64");
            const int64_t q = (int64_t)qv;
            const double restored = reconstruct(q, mins[d], eb);
            if (!within_eb(restored, original, eb)) {
                uint64_t idx = i * dims + d;
                bb_put(&corr_i, &idx, 8); bb_put(&corr_v, &original, 8); ++ncorr;
            }
            const int64_t b = q / pe;
            bid[i * dims + d] = b;
            if (b > max_bid) max_bid = b;
            rel[i * dims + d] = q - b * pe;
        }
    const int64_t max_bid_bits = bitlen((uint64_t)max_bid);

    /* L1, :2114-2138 */
    uint64_t* next = (uint64_t*)ax(c, n * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) next[i] = n;
    if (l1) l1_successors(c, bid, n, dims, next);

    /* chain decision, :2155-2200 */
    int64_t* idrel = (int64_t*)ax(c, n * sizeof(int64_t));
    int64_t* blocks = (int64_t*)ax(c, n * dims * sizeof(int64_t));
    uint64_t B = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (next[i] != n) {
            const uint64_t rid = next[i] - i;
            int64_t mb = bid[i * dims];
            for (uint32_t d = 1; d < dims; ++d) if (bid[i * dims + d] > mb) mb = bid[i * dims + d];
            const int local_bits = bitlen((uint64_t)mb);
            const int tag_bits = bitlen((uint64_t)local_bits);
            int64_t a = (int64_t)local_bits * (int64_t)dims + tag_bits;
            int64_t bnd = (int64_t)dims * max_bid_bits;
            const int smaller = (int)(a < bnd ? a : bnd);
            const int64_t lhs = bitlen(rid) + bitlen((uint64_t)bitlen(rid));
            if (lhs < (int64_t)smaller - thr) { idrel[i] = (int64_t)rid; continue; }
        }
        idrel[i] = 0;
        memcpy(blocks + B * dims, bid + i * dims, dims * sizeof(int64_t));
        ++B;
    }

    /* L2, :1305-1454 (or keep-all, :1456-1478) */
    int64_t* choice2 = (int64_t*)ax(c, (B ? B : 1) * sizeof(int64_t));
    int64_t* diff = (int64_t*)ax(c, (B ? B : 1) * dims * sizeof(int64_t));
    int64_t* keep = (int64_t*)ax(c, (B ? B : 1) * dims * sizeof(int64_t));
    uint64_t ndiff = 0, nkeep = 0;
    wrap_t wrap;
    if (!l2) {
        int64_t mn = INT64_MAX, mx = 0;
        for (uint64_t r = 0; r < B * dims; ++r) { if (blocks[r] < mn) mn = blocks[r]; if (blocks[r] > mx) mx = blocks[r]; }
        wrap = make_wrap(c, B == 0 ? 0 : mn, mx);
        for (uint64_t r = 0; r < B; ++r) choice2[r] = 0;
        memcpy(keep, blocks, B * dims * sizeof(int64_t));
        nkeep = B;
    } else if (B == 0) {
        wrap = make_wrap(c, 0, 0);
    } else {
        int64_t mn = INT64_MAX, mx = INT64_MIN;
        uint64_t max_abs = 0;
        for (uint64_t r = 0; r < B; ++r)
            for (uint32_t d = 0; d < dims; ++d) {
                const int64_t v = blocks[r * dims + d];
                REQUIRE(c, v >= 0, "block IDs must be non-negative");
                if (v < mn) mn = v;
                if (v > mx) mx = v;
                if (r) { uint64_t ad = absdiff_nn(c, blocks[(r - 1) * dims + d], v); if (ad > max_abs) max_abs = ad; }
            }
        wrap = make_wrap(c, mn, mx);
        const uint64_t fixed_value_bits = (uint64_t)dims * (uint64_t)bitlen((uint64_t)mx);
        const uint64_t fixed_diff_bits = (uint64_t)dims * (uint64_t)bitlen(max_abs);
        for (uint64_t r = 0; r < B; ++r) {
            int use_diff = 0;
            if (r) {
                uint64_t mv = 0, md = 0;
                for (uint32_t d = 0; d < dims; ++d) {
                    const uint64_t v = (uint64_t)blocks[r * dims + d];
                    if (v > mv) mv = v;
                    const uint64_t ad = absdiff_nn(c, blocks[(r - 1) * dims + d], blocks[r * dims + d]);
                    if (ad > md) md = ad;
                }
                const int tvw = bitlen(mv), tdw = bitlen(md);
                const uint64_t tag_value_bits = (uint64_t)dims * tvw + bitlen((uint64_t)tvw);
                const uint64_t tag_diff_bits = (uint64_t)dims * tdw + bitlen((uint64_t)tdw);
                const uint64_t value_bits = tag_value_bits < fixed_value_bits ? tag_value_bits : fixed_value_bits;
                const uint64_t diff_bits = tag_diff_bits > fixed_diff_bits ? tag_diff_bits : fixed_diff_bits;
                use_diff = value_bits >= diff_bits;
            }
            choice2[r] = use_diff;
            if (!use_diff) {
                memcpy(keep + nkeep * dims, blocks + r * dims, dims * sizeof(int64_t));
                ++nkeep;
                continue;
            }
            for (uint32_t d = 0; d < dims; ++d) {
                int64_t prev = blocks[(r - 1) * dims + d];
                int64_t pred = prev < wrap.min_value ? wrap.min_value : prev;
                if (pred > wrap.max_value) pred = wrap.max_value;
                int64_t corr = pred - blocks[r * dims + d];
                if (corr < wrap.min_correction) corr += wrap.max_difference;
                else if (corr > wrap.max_correction) corr -= wrap.max_difference;
                const uint64_t sym = zigzag(corr);
                REQUIRE(c, sym <= (uint64_t)INT64_MAX, "ZigZag block difference exceeds int64");
                diff[ndiff * dims + d] = (int64_t)sym;
            }
            ++ndiff;
        }
    }

    out->dims = dims; out->n = n; out->p_eff = pe; out->eb = eb;
    out->mins = (double*)dup_out(c, mins, dims * sizeof(double));
    out->idrel = (int64_t*)dup_out(c, idrel, n * sizeof(int64_t));
    out->nblocks = B;
    out->blocks = (int64_t*)dup_out(c, blocks, B * dims * sizeof(int64_t));
    out->nchoice = B;
    out->choice2 = (int64_t*)dup_out(c, choice2, B * sizeof(int64_t));
    out->ndiff = ndiff; out->diff = (int64_t*)dup_out(c, diff, ndiff * dims * sizeof(int64_t));
    out->nkeep = nkeep; out->keep = (int64_t*)dup_out(c, keep, nkeep * dims * sizeof(int64_t));
    out->rel = (int64_t*)dup_out(c, rel, n * dims * sizeof(int64_t));
    out->wrap_min = wrap.min_value; out->wrap_max = wrap.max_value;
    out->ncorr = ncorr;
    out->corr_index = (uint64_t*)dup_out(c, corr_i.p, ncorr * 8);
    out->corr_value = (double*)dup_out(c, corr_v.p, ncorr * 8);
    out->max_bid = (uint64_t)max_bid;
    arena_release(c);
    return 0;
}

void oc_streams_free(oc_streams* s) {
    if (!s) return;
    free(s->mins); free(s->idrel); free(s->blocks); free(s->choice2); free(s->diff);
    free(s->keep); free(s->rel); free(s->corr_index); free(s->corr_value);
    memset(s, 0, sizeof *s);
}

/* ------------------------------------------------------------------------ */
/* Fixed and Tagged packing */

/* lcps_cpp.cpp:134-159 */
static int fixed_width(const int64_t* v, uint64_t n, uint32_t dims, uint32_t d, int64_t* mn_out) {
    if (n == 0) { *mn_out = 0; return 1; }
    int64_t mn = v[d], mx = v[d];
    for (uint64_t i = 1; i < n; ++i) { int64_t x = v[i * dims + d]; if (x < mn) mn = x; if (x > mx) mx = x; }
    *mn_out = mn;
    if (mn == mx) return 1;
    uint64_t range = (uint64_t)mx - (uint64_t)mn;
    int w = 1;
    while ((range >>= 1) != 0) ++w;
    return w;
}

/* MSB-first bit writer, lcps_cpp.cpp:161-222 and 484-537 */
typedef struct { bb* b; uint8_t cur; int filled; } bitw;
static void bw_put(bitw* w, uint64_t v, int width) {
    for (int bit = width - 1; bit >= 0; --bit) {
        w->cur = (uint8_t)((w->cur << 1) | ((v >> bit) & 1u));
        if (++w->filled == 8) { bb_u8(w->b, w->cur); w->cur = 0; w->filled = 0; }
    }
}
static void bw_flush(bitw* w) {
    if (w->filled > 0) { bb_u8(w->b, (uint8_t)(w->cur << (8 - w->filled))); w->cur = 0; w->filled = 0; }
}

static void pack_fixed_dim(octx* c, const int64_t* v, uint64_t n, uint32_t dims, uint32_t d,
                           int64_t mn, int width, bb* out) {
    REQUIRE(c, width >= 1 && width <= 64, "invalid fixed width");
    const uint64_t mask = width < 64 ? ((uint64_t)1 << width) - 1 : UINT64_MAX;
    bitw w = {out, 0, 0};
    for (uint64_t i = 0; i < n; ++i) {
        const int64_t raw = v[i * dims + d];
        REQUIRE(c, raw >= mn, "_pack_unsigned expects non-negative values");
        const uint64_t x = (uint64_t)raw - (uint64_t)mn;
        REQUIRE(c, x <= mask, "fixed value does not fit in its bit width");
        bw_put(&w, x, width);
    }
    bw_flush(&w);
}

/* per-record tags and the point-major packed values, lcps_cpp.cpp:1005-1061 */
static void tagged_parts(octx* c, const int64_t* v, uint64_t n, uint32_t dims,
                         uint64_t* tags, int* max_width, bb* packed) {
    bitw w = {packed, 0, 0};
    int mw = 0;
    for (uint64_t i = 0; i < n; ++i) {
        int tag = 1;
        for (uint32_t d = 0; d < dims; ++d) {
            REQUIRE(c, v[i * dims + d] >= 0, "Tagged encoding expects non-negative integer values");
            int b = bitlen((uint64_t)v[i * dims + d]);
            if (b > tag) tag = b;
        }
        tags[i] = (uint64_t)tag;
        if (tag > mw) mw = tag;
        for (uint32_t d = 0; d < dims; ++d) bw_put(&w, (uint64_t)v[i * dims + d], tag);
    }
    bw_flush(&w);
    *max_width = mw;
}

/* ------------------------------------------------------------------------ */
/* RVN1 normalised value-rANS, lcps_cpp.cpp:539-821 (encoder) and
 * lcps_cpu_decompress_serial.cpp:981-1057 (decoder). */

#define MAX_VALUE_ALPHABET (1u << 18)

static int rvn1_precision_bits(octx* c, size_t unique) {   /* :541-552 */
    REQUIRE(c, unique != 0, "cannot choose rANS precision for no symbols");
    int ub = bitlen((uint64_t)unique) + 2;
    if (ub > 18) ub = 18;
    int pb = (3 * ub) / 2;
    if (pb < 12) pb = 12;
    if (pb > 20) pb = 20;
    return pb;
}

typedef struct { int precision_bits; uint32_t alphabet; uint32_t* prob; uint32_t* cum; } rtable;

typedef struct { uint32_t prob; uint32_t idx; } pidx;
static int pidx_cmp(const void* a, const void* b) {
    const pidx* x = (const pidx*)a; const pidx* y = (const pidx*)b;
    if (x->prob != y->prob) return x->prob < y->prob ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static rtable rvn1_build_table(octx* c, const uint64_t* values, uint64_t n) {   /* :560-684 */
    REQUIRE(c, n > 0, "cannot build an rANS table for an empty stream");
    uint64_t maxv = 0;
    for (uint64_t i = 0; i < n; ++i) if (values[i] > maxv) maxv = values[i];
    REQUIRE(c, maxv < MAX_VALUE_ALPHABET, "Raw value-rANS alphabet exceeds the 18-bit limit");
    const uint32_t A = (uint32_t)maxv + 1;
    uint64_t* freq = (uint64_t*)ax(c, A * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) ++freq[values[i]];
    size_t unique = 0;
    for (uint32_t s = 0; s < A; ++s) unique += freq[s] != 0;
    rtable t;
    t.precision_bits = rvn1_precision_bits(c, unique);
    t.alphabet = A;
    const uint32_t precision = 1u << t.precision_bits;
    const long double total = (long double)n;
    t.prob = (uint32_t*)ax(c, A * sizeof(uint32_t));
    t.cum = (uint32_t*)ax(c, A * sizeof(uint32_t));
    int64_t sum = 0;
    for (uint32_t s = 0; s < A; ++s) {
        uint32_t pr = (uint32_t)floorl((long double)freq[s] * precision / total + 0.5L);
        if (pr == 0 && freq[s] != 0) pr = 1;
        t.prob[s] = pr;
        sum += pr;
    }
    if (sum != (int64_t)precision) {
        /* std::stable_sort by probability == sort by (probability, index) */
        pidx* sorted = (pidx*)ax(c, A * sizeof(pidx));
        for (uint32_t s = 0; s < A; ++s) { sorted[s].prob = t.prob[s]; sorted[s].idx = s; }
        qsort(sorted, A, sizeof(pidx), pidx_cmp);
        if (sum < (int64_t)precision) {
            t.prob[sorted[A - 1].idx] += (uint32_t)((int64_t)precision - sum);
            sum = precision;
        } else {
            int64_t error = sum - (int64_t)precision;
            while (error > 0) {
                const long double relative = (long double)precision / (long double)sum;
                int changed = 0;
                for (size_t j = A; j-- > 1 && error > 0;) {
                    uint32_t* pr = &t.prob[sorted[j].idx];
                    if (*pr <= 1) {
                        REQUIRE(c, j != (size_t)A - 1, "rANS normalization cannot retain every symbol");
                        break;
                    }
                    const uint32_t np = (uint32_t)floorl(relative * *pr);
                    int64_t reduction = (int64_t)*pr - np;
                    if (reduction == 0) reduction = 1;
                    if (reduction >= *pr) reduction = *pr - 1;
                    if (reduction > error) reduction = error;
                    *pr -= (uint32_t)reduction;
                    sum -= reduction;
                    error -= reduction;
                    changed = 1;
                }
                REQUIRE(c, changed, "rANS normalization did not converge");
            }
        }
    }
    uint32_t cu = 0;
    for (uint32_t s = 0; s < A; ++s) { t.cum[s] = cu; cu += t.prob[s]; }
    REQUIRE(c, cu == precision && sum == (int64_t)precision, "normalized rANS probability total mismatch");
    return t;
}

/* :686-729 */
static void rvn1_table_bytes(octx* c, const rtable* t, bb* out) {
    REQUIRE(c, t->alphabet > 0 && t->alphabet <= MAX_VALUE_ALPHABET, "invalid normalized rANS alphabet size");
    bb_varuint(out, t->alphabet);
    for (uint32_t i = 0; i < t->alphabet; ++i) {
        const uint32_t pr = t->prob[i];
        if (pr == 0) {
            uint32_t extra = 0;
            while (extra < 63 && i + extra + 1 < t->alphabet && t->prob[i + extra + 1] == 0) ++extra;
            bb_u8(out, (uint8_t)((extra << 2) | 3u));
            i += extra;
            continue;
        }
        int eb = 0;
        if (pr >= (1u << 6)) { ++eb; if (pr >= (1u << 14)) ++eb; }
        bb_u8(out, (uint8_t)((pr << 2) | (uint32_t)eb));
        for (int k = 0; k < eb; ++k) bb_u8(out, (uint8_t)(pr >> (8 * (k + 1) - 2)));
    }
}

/* parse the compact table, lcps_cpu_decompress_serial.cpp:1004-1039 */
static rtable rvn1_parse_table(octx* c, rd* tr, int precision_bits) {
    rtable t;
    t.precision_bits = precision_bits;
    const uint64_t A = rd_varuint(tr);
    REQUIRE(c, A > 0 && A <= MAX_VALUE_ALPHABET, "invalid rANS alphabet");
    t.alphabet = (uint32_t)A;
    t.prob = (uint32_t*)ax(c, A * sizeof(uint32_t));
    t.cum = (uint32_t*)ax(c, A * sizeof(uint32_t));
    for (uint64_t i = 0; i < A;) {
        const uint8_t first = rd_u8(tr);
        const int token = first & 3u;
        if (token == 3) {
            const uint64_t zeros = (first >> 2) + 1u;
            REQUIRE(c, zeros <= A - i, "invalid rANS zero run");
            i += zeros;
        } else {
            uint32_t pr = first >> 2;
            for (int b = 0; b < token; ++b) pr |= (uint32_t)rd_u8(tr) << (6 + 8 * b);
            REQUIRE(c, pr > 0, "invalid rANS probability");
            t.prob[i++] = pr;
        }
    }
    rd_end(tr);
    const uint32_t precision = 1u << precision_bits;
    uint32_t sum = 0;
    for (uint32_t s = 0; s < A; ++s) {
        t.cum[s] = sum;
        REQUIRE(c, t.prob[s] <= precision - sum, "invalid rANS probability sum");
        sum += t.prob[s];
    }
    REQUIRE(c, sum == precision, "incomplete rANS probability table");
    return t;
}

static uint32_t* slot_lookup(octx* c, const rtable* t) {
    const uint32_t precision = 1u << t->precision_bits;
    uint32_t* lk = (uint32_t*)ax(c, precision * sizeof(uint32_t));
    for (uint32_t s = 0; s < t->alphabet; ++s)
        for (uint32_t k = 0; k < t->prob[s]; ++k) lk[t->cum[s] + k] = s;
    return lk;
}

/* :739-821 */
static void rvn1_encode(octx* c, const uint64_t* values, uint64_t n, bb* out) {
    out->n = 0;
    bb_put(out, "RVN1", 4);
    if (n == 0) {
        bb_u8(out, 0); bb_u64(out, 0); bb_u32(out, 1); bb_u64(out, 0); bb_u64(out, 0);
        return;
    }
    rtable t = rvn1_build_table(c, values, n);
    bb table = bb_make(c);
    rvn1_table_bytes(c, &t, &table);
    const uint32_t precision = 1u << t.precision_bits;
    const uint32_t lower = precision * 4u;
    uint64_t state = lower;
    uint8_t* ren = (uint8_t*)ax(c, n * 4 + 16);
    size_t nren = 0;
    for (uint64_t k = n; k-- > 0;) {
        const uint64_t s = values[k];
        const uint64_t f = t.prob[s], start = t.cum[s];
        const uint64_t x_max = ((uint64_t)lower >> t.precision_bits) * 256u * f;
        while (state >= x_max) { ren[nren++] = (uint8_t)(state & 0xffu); state >>= 8; }
        state = (state / f) * precision + (state % f) + start;
        REQUIRE(c, state < (uint64_t)lower * 256u, "normalized value-rANS state overflow");
    }
    bb_u8(out, (uint8_t)t.precision_bits);
    bb_u64(out, n);
    bb_u32(out, (uint32_t)state);
    bb_blob(out, &table);
    bb_u64(out, nren);
    for (size_t k = nren; k-- > 0;) bb_u8(out, ren[k]);
}

static uint64_t* rvn1_decode(octx* c, rd* r, uint64_t expected) {
    REQUIRE(c, rd_magic(r, "RVN1"), "invalid RVN1 header");
    const int pb = rd_u8(r);
    REQUIRE(c, rd_u64(r) == expected, "rANS symbol count mismatch");
    uint64_t state = rd_u32(r);
    rd table = rd_blob(r);
    rd ren = rd_blob(r);
    rd_end(r);
    uint64_t* res = (uint64_t*)ax(c, (expected ? expected : 1) * sizeof(uint64_t));
    if (expected == 0) {
        REQUIRE(c, pb == 0 && state == 1 && table.n == 0 && ren.n == 0, "invalid empty rANS stream");
        return res;
    }
    REQUIRE(c, pb >= 12 && pb <= 20, "invalid rANS precision");
    const uint32_t precision = 1u << pb, lower = precision * 4u;
    REQUIRE(c, state >= lower && state < (uint64_t)lower * 256, "invalid rANS state");
    rtable t = rvn1_parse_table(c, &table, pb);
    uint32_t* lk = slot_lookup(c, &t);
    for (uint64_t k = 0; k < expected; ++k) {
        const uint32_t slot = (uint32_t)state & (precision - 1);
        const uint32_t s = lk[slot];
        res[k] = s;
        state = (uint64_t)t.prob[s] * (state >> pb) + slot - t.cum[s];
        while (state < lower) state = (state << 8) | rd_u8(&ren);
    }
    rd_end(&ren);
    REQUIRE(c, state == lower, "rANS final state mismatch");
    return res;
}

/* ------------------------------------------------------------------------ */
/* RVC1: this project's chunked, lane-interleaved form of the RVN1 coder.
 *
 * Same normalised table, same state interval [4M, 1024M) and byte-wise
 * renormalisation as RVN1. Symbol k of a chunk belongs to lane k % L at step
 * k / L. The decoder runs the steps forward; after a step, lanes whose state
 * fell below 4M read bytes in rounds, each round in lane order. The encoder
 * writes every step's bytes in exactly that order, so with L = 1 and one chunk
 * the stream equals RVN1's renormalisation bytes.
 *
 *   "RVC1" u8 precision_bits u64 count u64 table_bytes table
 *   u8 lanes u32 chunk_symbols u32 chunk_count
 *   chunk_count x u32 chunk_bytes
 *   chunks: lanes x u32 final_state, then the byte stream
 * An empty section has precision, lanes, chunk_symbols and chunk_count 0.
 */

uint32_t oc_rvc1_lanes_for(uint64_t n, uint32_t lanes_max) {
    uint32_t L = lanes_max ? lanes_max : 1;
    while (L > 1 && n < (uint64_t)L * 256u) L >>= 1;
    return L;
}

static void rvc1_encode(octx* c, const uint64_t* values, uint64_t n, uint32_t L, uint32_t S, bb* out) {
    bb_put(out, "RVC1", 4);
    if (n == 0) {
        bb_u8(out, 0); bb_u64(out, 0); bb_u64(out, 0); bb_u8(out, 0); bb_u32(out, 0); bb_u32(out, 0);
        return;
    }
    REQUIRE(c, L >= 1 && L <= 64 && (L & (L - 1)) == 0, "RVC1 lanes must be a power of two <= 64");
    REQUIRE(c, S >= 1, "RVC1 chunk size must be positive");
    rtable t = rvn1_build_table(c, values, n);
    bb table = bb_make(c);
    rvn1_table_bytes(c, &t, &table);
    const int pb = t.precision_bits;
    const uint32_t M = 1u << pb, lower = 4u * M;
    const uint64_t nchunks = (n + S - 1) / S;
    REQUIRE(c, nchunks <= UINT32_MAX, "too many RVC1 chunks");
    bb_u8(out, (uint8_t)pb); bb_u64(out, n); bb_blob(out, &table);
    bb_u8(out, (uint8_t)L); bb_u32(out, S); bb_u32(out, (uint32_t)nchunks);
    const size_t size_pos = out->n;
    for (uint64_t k = 0; k < nchunks; ++k) bb_u32(out, 0);
    uint8_t* tmp = (uint8_t*)ax(c, (size_t)S * 3 + 16);
    uint64_t x[64];
    int kj[64];
    for (uint64_t ch = 0; ch < nchunks; ++ch) {
        const uint64_t base = ch * S;
        const uint64_t Sc = (n - base) < S ? (n - base) : S;
        const uint64_t T = (Sc + L - 1) / L;
        for (uint32_t j = 0; j < L; ++j) x[j] = lower;
        size_t pos = (size_t)S * 3 + 16, end = pos;
        for (uint64_t t1 = T; t1-- > 0;) {
            int cnt[4] = {0, 0, 0, 0};
            for (uint32_t j = 0; j < L; ++j) {
                kj[j] = 0;
                const uint64_t k = t1 * L + j;
                if (k >= Sc) continue;
                const uint64_t f = t.prob[values[base + k]];
                const uint64_t x_max = (uint64_t)(lower >> pb) * 256u * f;
                uint64_t xx = x[j];
                while (xx >= x_max) { ++kj[j]; xx >>= 8; }
                for (int r = 0; r < kj[j]; ++r) ++cnt[r];
            }
            const size_t G = (size_t)cnt[0] + cnt[1] + cnt[2] + cnt[3];
            pos -= G;
            size_t round_off[4];
            round_off[0] = 0;
            for (int r = 1; r < 4; ++r) round_off[r] = round_off[r - 1] + cnt[r - 1];
            int rank[4] = {0, 0, 0, 0};
            for (uint32_t j = 0; j < L; ++j) {
                for (int r = 0; r < kj[j]; ++r) {
                    tmp[pos + round_off[r] + rank[r]] = (uint8_t)(x[j] >> (8 * (kj[j] - 1 - r)));
                    ++rank[r];
                }
            }
            for (uint32_t j = 0; j < L; ++j) {
                const uint64_t k = t1 * L + j;
                if (k >= Sc) continue;
                const uint64_t s = values[base + k];
                const uint64_t f = t.prob[s];
                uint64_t xx = x[j] >> (8 * kj[j]);
                xx = (xx / f) * M + (xx % f) + t.cum[s];
                REQUIRE(c, xx < (uint64_t)lower * 256u, "RVC1 state overflow");
                x[j] = xx;
            }
        }
        const size_t chunk_start = out->n;
        for (uint32_t j = 0; j < L; ++j) bb_u32(out, (uint32_t)x[j]);
        bb_put(out, tmp + pos, end - pos);
        const uint64_t cb = out->n - chunk_start;
        REQUIRE(c, cb <= UINT32_MAX, "RVC1 chunk too large");
        for (int i = 0; i < 4; ++i) out->p[size_pos + 4 * ch + i] = (uint8_t)(cb >> (8 * i));
    }
}

static uint64_t* rvc1_decode(octx* c, rd* r, uint64_t expected) {
    REQUIRE(c, rd_magic(r, "RVC1"), "invalid RVC1 header");
    const int pb = rd_u8(r);
    REQUIRE(c, rd_u64(r) == expected, "RVC1 symbol count mismatch");
    rd table = rd_blob(r);
    const uint32_t L = rd_u8(r);
    const uint32_t S = rd_u32(r);
    const uint32_t nchunks = rd_u32(r);
    uint64_t* res = (uint64_t*)ax(c, (expected ? expected : 1) * sizeof(uint64_t));
    if (expected == 0) {
        REQUIRE(c, pb == 0 && table.n == 0 && L == 0 && S == 0 && nchunks == 0, "invalid empty RVC1 section");
        rd_end(r);
        return res;
    }
    REQUIRE(c, pb >= 12 && pb <= 20, "invalid RVC1 precision");
    REQUIRE(c, L >= 1 && L <= 64 && (L & (L - 1)) == 0 && S >= 1, "invalid RVC1 lanes or chunk size");
    REQUIRE(c, nchunks == (expected + S - 1) / S, "RVC1 chunk count mismatch");
    rtable t = rvn1_parse_table(c, &table, pb);
    uint32_t* lk = slot_lookup(c, &t);
    const uint32_t M = 1u << pb, lower = 4u * M;
    const uint8_t* sizes = rd_bytes(r, (size_t)nchunks * 4);
    uint64_t x[64];
    for (uint32_t ch = 0; ch < nchunks; ++ch) {
        uint32_t cb = 0;
        for (int i = 0; i < 4; ++i) cb |= (uint32_t)sizes[4 * ch + i] << (8 * i);
        REQUIRE(c, cb >= 4u * L, "RVC1 chunk too small");
        rd chunk = {c, rd_bytes(r, cb), cb};
        for (uint32_t j = 0; j < L; ++j) {
            x[j] = rd_u32(&chunk);
            REQUIRE(c, x[j] >= lower && x[j] < (uint64_t)lower * 256u, "invalid RVC1 state");
        }
        const uint64_t base = (uint64_t)ch * S;
        const uint64_t Sc = (expected - base) < S ? (expected - base) : S;
        const uint64_t T = (Sc + L - 1) / L;
        for (uint64_t t1 = 0; t1 < T; ++t1) {
            for (uint32_t j = 0; j < L; ++j) {
                const uint64_t k = t1 * L + j;
                if (k >= Sc) continue;
                const uint32_t slot = (uint32_t)x[j] & (M - 1);
                const uint32_t s = lk[slot];
                res[base + k] = s;
                x[j] = (uint64_t)t.prob[s] * (x[j] >> pb) + slot - t.cum[s];
            }
            for (;;) {
                int any = 0;
                for (uint32_t j = 0; j < L; ++j)
                    if (x[j] < lower) { x[j] = (x[j] << 8) | rd_u8(&chunk); any = 1; }
                if (!any) break;
            }
        }
        rd_end(&chunk);
        for (uint32_t j = 0; j < L; ++j) REQUIRE(c, x[j] == lower, "RVC1 final state mismatch");
    }
    rd_end(r);
    return res;
}

/* ------------------------------------------------------------------------ */
/* Zstd (lcps_cpp.cpp:112-132, serial.cpp:969-979) */

static void zstd_compress(octx* c, const uint8_t* src, size_t n, bb* out) {
    const size_t bound = ZSTD_compressBound(n);
    out->n = 0;
    bb_need(out, bound);
    const size_t k = ZSTD_compress(out->p, bound, src, n, 22);
    REQUIRE(c, !ZSTD_isError(k), "ZSTD_compress failed: %s", ZSTD_getErrorName(k));
    out->n = k;
}

static uint8_t* zstd_decompress(octx* c, const uint8_t* src, size_t n, size_t size) {
    uint8_t* res = (uint8_t*)ax(c, size ? size : 1);
    const size_t k = ZSTD_decompress(res, size, src, n);
    REQUIRE(c, !ZSTD_isError(k), "ZSTD_decompress failed: %s", ZSTD_getErrorName(k));
    REQUIRE(c, k == size, "LCPS Zstd output size mismatch");
    return res;
}

/* ------------------------------------------------------------------------ */
/* Component encoders */

typedef struct { const int64_t* v; uint64_t n; uint32_t dims; } mat;

static uint64_t* to_u64(octx* c, const mat* m) {
    uint64_t* u = (uint64_t*)ax(c, ((m->n * m->dims) != 0 ? m->n * m->dims : 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < m->n * m->dims; ++i) u[i] = (uint64_t)m->v[i];
    return u;
}

static void raw1(const mat* m, bb* out) {                       /* :1184-1211 */
    bb_put(out, "RAW1", 4); bb_u32(out, m->dims); bb_u64(out, m->n);
    for (uint64_t i = 0; i < m->n * m->dims; ++i) bb_i64(out, m->v[i]);
}

static void ztg4(octx* c, const mat* m, bb* out) {              /* :1093-1138 */
    uint64_t* tags = (uint64_t*)ax(c, (m->n ? m->n : 1) * sizeof(uint64_t));
    bb packed = bb_make(c), tagp = bb_make(c), z = bb_make(c);
    int mw = 0;
    tagged_parts(c, m->v, m->n, m->dims, tags, &mw, &packed);
    rvn1_encode(c, tags, m->n, &tagp);
    zstd_compress(c, packed.p, packed.n, &z);
    bb_put(out, "ZTG4", 4); bb_u32(out, m->dims); bb_u64(out, m->n); bb_u8(out, (uint8_t)mw);
    bb_blob(out, &tagp);
    bb_u64(out, packed.n);
    bb_blob(out, &z);
}

static void zfmx(octx* c, const mat* m, bb* out) {              /* :856-933 */
    bb payload = bb_make(c), z = bb_make(c);
    int64_t* mn = (int64_t*)ax(c, m->dims * sizeof(int64_t));
    int* w = (int*)ax(c, m->dims * sizeof(int));
    uint64_t* sz = (uint64_t*)ax(c, m->dims * sizeof(uint64_t));
    for (uint32_t d = 0; d < m->dims; ++d) {
        w[d] = fixed_width(m->v, m->n, m->dims, d, &mn[d]);
        size_t before = payload.n;
        pack_fixed_dim(c, m->v, m->n, m->dims, d, mn[d], w[d], &payload);
        sz[d] = payload.n - before;
    }
    zstd_compress(c, payload.p, payload.n, &z);
    bb_put(out, "ZFMX", 4); bb_u32(out, m->dims); bb_u64(out, m->n);
    for (uint32_t d = 0; d < m->dims; ++d) { bb_i64(out, mn[d]); bb_u8(out, (uint8_t)w[d]); bb_u64(out, sz[d]); }
    bb_u64(out, payload.n);
    bb_blob(out, &z);
}

static int rvmx(octx* c, const mat* m, bb* out) {               /* :935-995 */
    for (uint64_t i = 0; i < m->n * m->dims; ++i) {
        REQUIRE(c, m->v[i] >= 0, "Raw rANS expects non-negative integer values");
        if ((uint64_t)m->v[i] >= MAX_VALUE_ALPHABET) return 0;
    }
    bb r = bb_make(c);
    rvn1_encode(c, to_u64(c, m), m->n * m->dims, &r);
    bb_put(out, "RVMX", 4); bb_u32(out, m->dims); bb_u64(out, m->n);
    bb_blob(out, &r);
    return 1;
}

/* :1140-1174: smallest wins, tie order Tagged+Zstd, Fixed+Zstd, Raw+rANS */
static void choose_lcp2(octx* c, const mat* m, bb* out) {
    bb best = bb_make(c), cand = bb_make(c);
    ztg4(c, m, &best);
    zfmx(c, m, &cand);
    if (cand.n < best.n) { bb t = best; best = cand; cand = t; }
    cand.n = 0;
    if (rvmx(c, m, &cand) && cand.n < best.n) best = cand;
    out->n = 0;
    bb_put(out, best.p, best.n);
}

/* LCH1 components (this project's GPU container). */
enum { K_TAGR = 0, K_FIXP = 1, K_RANS = 2 };
#define RANS_VALUE_LIMIT 4096u

static void lch1_tagr(octx* c, const mat* m, const oc_rans_params* rp, bb* out) {
    uint64_t* tags = (uint64_t*)ax(c, (m->n ? m->n : 1) * sizeof(uint64_t));
    bb packed = bb_make(c), sec = bb_make(c);
    int mw = 0;
    tagged_parts(c, m->v, m->n, m->dims, tags, &mw, &packed);
    rvc1_encode(c, tags, m->n, oc_rvc1_lanes_for(m->n, rp->lanes_max), rp->chunk_symbols, &sec);
    bb_put(out, "TAGR", 4); bb_u32(out, m->dims); bb_u64(out, m->n); bb_u8(out, (uint8_t)mw);
    bb_blob(out, &sec);
    bb_blob(out, &packed);
}

static void lch1_fixp(octx* c, const mat* m, bb* out) {
    bb payload = bb_make(c);
    bb_put(out, "FIXP", 4); bb_u32(out, m->dims); bb_u64(out, m->n);
    for (uint32_t d = 0; d < m->dims; ++d) {
        int64_t mn; int w = fixed_width(m->v, m->n, m->dims, d, &mn);
        size_t before = payload.n;
        pack_fixed_dim(c, m->v, m->n, m->dims, d, mn, w, &payload);
        bb_i64(out, mn); bb_u8(out, (uint8_t)w); bb_u64(out, payload.n - before);
    }
    bb_put(out, payload.p, payload.n);
}

static int lch1_rans(octx* c, const mat* m, const oc_rans_params* rp, bb* out) {
    const uint64_t k = m->n * m->dims;
    for (uint64_t i = 0; i < k; ++i) if (m->v[i] < 0 || (uint64_t)m->v[i] >= RANS_VALUE_LIMIT) return 0;
    bb sec = bb_make(c);
    rvc1_encode(c, to_u64(c, m), k, oc_rvc1_lanes_for(k, rp->lanes_max), rp->chunk_symbols, &sec);
    bb_put(out, "RANS", 4); bb_u32(out, m->dims); bb_u64(out, m->n);
    bb_blob(out, &sec);
    return 1;
}

/* The LCH1 chooser (this project's container): every candidate is first
 * estimated in integer arithmetic, a RVC1 section as sum(count * (pb - log2 f))
 * bits in 1/65536 (log2 from the exponent and a table of log2(1 + k/256) over
 * the 8 bits below the leading one) plus its exact header, table and chunk
 * words and the chunks' lane states; Fixed is exact. The candidates within
 * 1/256 (+ 64 bytes) of the smallest estimate are encoded and the smallest of
 * those wins, ties Tagged > Fixed > Raw-rANS. The GPU engine applies the same
 * rule (chain_int.hpp: log2_q16, rvc1_section_estimate, lch1_contenders). */
static const uint16_t LOG2_FRAC[256] = {
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

static uint32_t log2_q16(uint32_t f) {
    const int e = bitlen(f) - 1;
    const uint32_t k = e >= 8 ? (f >> (e - 8)) & 255u : (f << (8 - e)) & 255u;
    return ((uint32_t)e << 16) + LOG2_FRAC[k];
}

/* estimated bytes of the RVC1 section of values[0..n) with the given lanes */
static uint64_t rvc1_estimate(octx* c, const uint64_t* values, uint64_t n, uint32_t L, uint32_t S) {
    if (n == 0) return 30;
    rtable t = rvn1_build_table(c, values, n);
    bb table = bb_make(c);
    rvn1_table_bytes(c, &t, &table);
    uint64_t* hist = (uint64_t*)ax(c, (size_t)t.alphabet * sizeof(uint64_t));
    memset(hist, 0, (size_t)t.alphabet * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) ++hist[values[i]];
    uint64_t bits = 0;
    for (uint32_t s = 0; s < t.alphabet; ++s)
        if (hist[s]) bits += hist[s] * (uint64_t)(((uint32_t)t.precision_bits << 16) - log2_q16(t.prob[s]));
    const uint64_t nchunks = (n + S - 1) / S;
    return 30 + table.n + 4 * nchunks + 4ull * L * nchunks + ((bits + (8ull << 16) - 1) >> 19);
}

static uint32_t lch1_contenders(uint64_t tagr, uint64_t fixp, uint64_t rans) {
    uint64_t best = tagr < fixp ? tagr : fixp;
    best = rans < best ? rans : best;
    const uint64_t lim = best + (best >> 8) + 64;
    return (tagr <= lim ? 1u : 0u) | (fixp <= lim ? 2u : 0u) | (rans <= lim ? 4u : 0u);
}

static uint32_t choose_lch1(octx* c, const mat* m, const oc_rans_params* rp, bb* out) {
    const uint64_t k = m->n * m->dims;
    /* Tagged estimate: widths section + the packed bits */
    uint64_t* tags = (uint64_t*)ax(c, (m->n ? m->n : 1) * sizeof(uint64_t));
    uint64_t tag_bits = 0;
    for (uint64_t i = 0; i < m->n; ++i) {
        int tag = 1;
        for (uint32_t d = 0; d < m->dims; ++d) {
            const int b = m->v[i * m->dims + d] < 0 ? 64 : bitlen((uint64_t)m->v[i * m->dims + d]);
            if (b > tag) tag = b;
        }
        tags[i] = (uint64_t)tag;
        tag_bits += (uint64_t)tag * m->dims;
    }
    const uint64_t tagr = 4 + 4 + 8 + 1 + 8 + rvc1_estimate(c, tags, m->n, oc_rvc1_lanes_for(m->n, rp->lanes_max),
                                                            rp->chunk_symbols) + 8 + (tag_bits + 7) / 8;
    /* Fixed: exact */
    uint64_t fixp = 4 + 4 + 8 + 17ull * m->dims;
    for (uint32_t d = 0; d < m->dims; ++d) {
        int64_t mn;
        const int w = fixed_width(m->v, m->n, m->dims, d, &mn);
        fixp += (m->n * (uint64_t)w + 7) / 8;
    }
    /* Raw-rANS estimate, only when every value is < 4096 */
    int eligible = 1;
    for (uint64_t i = 0; i < k; ++i) if (m->v[i] < 0 || (uint64_t)m->v[i] >= RANS_VALUE_LIMIT) { eligible = 0; break; }
    uint64_t rans = UINT64_MAX;
    if (eligible)
        rans = 4 + 4 + 8 + 8 + rvc1_estimate(c, to_u64(c, m), k, oc_rvc1_lanes_for(k, rp->lanes_max), rp->chunk_symbols);
    const uint32_t mask = lch1_contenders(tagr, fixp, rans);
    /* the contenders, encoded exactly; ties Tagged > Fixed > Raw-rANS */
    bb best = bb_make(c), cand = bb_make(c);
    uint32_t kind = UINT32_MAX;
    if (mask & 1u) { lch1_tagr(c, m, rp, &best); kind = K_TAGR; }
    if (mask & 2u) {
        lch1_fixp(c, m, &cand);
        if (kind == UINT32_MAX || cand.n < best.n) { bb t = best; best = cand; cand = t; kind = K_FIXP; }
        cand.n = 0;
    }
    if (mask & 4u) {
        cand.n = 0;
        (void)lch1_rans(c, m, rp, &cand);
        if (kind == UINT32_MAX || cand.n < best.n) { best = cand; kind = K_RANS; }
    }
    out->n = 0;
    bb_put(out, best.p, best.n);
    return kind;
}

/* ------------------------------------------------------------------------ */
/* Container writers, lcps_cpp.cpp:2306-2329 */

static void header(bb* b, const char* magic, const oc_streams* s) {
    bb_put(b, magic, 4);
    bb_u32(b, s->dims); bb_u64(b, s->n); bb_i64(b, s->p_eff); bb_f64(b, s->eb);
    bb_i64(b, s->wrap_min); bb_i64(b, s->wrap_max);
    for (uint32_t d = 0; d < s->dims; ++d) bb_f64(b, s->mins[d]);
}

static void components(const oc_streams* s, mat m[5]) {
    m[0].v = s->idrel;   m[0].n = s->n;       m[0].dims = 1;
    m[1].v = s->diff;    m[1].n = s->ndiff;   m[1].dims = s->dims;
    m[2].v = s->keep;    m[2].n = s->nkeep;   m[2].dims = s->dims;
    m[3].v = s->choice2; m[3].n = s->nchoice; m[3].dims = 1;
    m[4].v = s->rel;     m[4].n = s->n;       m[4].dims = s->dims;
}

static void raw_corrections(const oc_streams* s, bb* b) {
    for (uint64_t k = 0; k < s->ncorr; ++k) { bb_u64(b, s->corr_index[k]); bb_f64(b, s->corr_value[k]); }
}

int oc_write_lcp3(const oc_streams* s, uint8_t** out, size_t* out_len) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    bb b = bb_make(c);
    header(&b, "LCP3", s);
    mat m[5]; components(s, m);
    for (int k = 0; k < 5; ++k) { bb comp = bb_make(c); raw1(&m[k], &comp); bb_blob(&b, &comp); }
    bb_u64(&b, s->ncorr);
    bb corr = bb_make(c); raw_corrections(s, &corr); bb_blob(&b, &corr);
    *out = (uint8_t*)dup_out(c, b.p, b.n); *out_len = b.n;
    arena_release(c);
    return 0;
}

int oc_write_lcp2(const oc_streams* s, uint8_t** out, size_t* out_len, uint64_t stats[5]) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    bb b = bb_make(c);
    header(&b, "LCP2", s);
    mat m[5]; components(s, m);
    uint64_t sizes[5];
    for (int k = 0; k < 5; ++k) { bb comp = bb_make(c); choose_lcp2(c, &m[k], &comp); sizes[k] = comp.n; bb_blob(&b, &comp); }
    bb_u64(&b, s->ncorr);
    bb corr = bb_make(c);
    if (s->ncorr) { bb raw = bb_make(c); raw_corrections(s, &raw); zstd_compress(c, raw.p, raw.n, &corr); }
    bb_blob(&b, &corr);
    if (stats) {     /* :2331-2337 */
        stats[0] = b.n;
        stats[1] = sizes[1] + sizes[2] + sizes[3];
        stats[2] = sizes[0];
        stats[3] = sizes[4];
        stats[4] = s->n - s->nblocks;
    }
    *out = (uint8_t*)dup_out(c, b.p, b.n); *out_len = b.n;
    arena_release(c);
    return 0;
}

int oc_write_lch1(const oc_streams* s, const oc_rans_params* rp, uint8_t** out, size_t* out_len, uint32_t kinds[5]) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    bb b = bb_make(c);
    header(&b, "LCH1", s);
    mat m[5]; components(s, m);
    for (int k = 0; k < 5; ++k) {
        bb comp = bb_make(c);
        uint32_t kind = choose_lch1(c, &m[k], rp, &comp);
        if (kinds) kinds[k] = kind;
        bb_blob(&b, &comp);
    }
    bb_u64(&b, s->ncorr);
    bb corr = bb_make(c); raw_corrections(s, &corr); bb_blob(&b, &corr);
    *out = (uint8_t*)dup_out(c, b.p, b.n); *out_len = b.n;
    arena_release(c);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Decoders */

static uint64_t read_bits(octx* c, const uint8_t* p, size_t n, size_t* pos, int width) {   /* serial.cpp:1059-1072 */
    REQUIRE(c, width >= 0 && width <= 64, "invalid packed bit width");
    uint64_t v = 0;
    for (int bit = 0; bit < width; ++bit) {
        REQUIRE(c, *pos / 8 < n, "truncated packed values");
        v = (v << 1) | ((p[*pos / 8] >> (7 - *pos % 8)) & 1u);
        ++*pos;
    }
    return v;
}

/* serial.cpp:1076-1174 (LCP2/LCP3 components) plus LCH1 components.
 * Returns a point-major matrix. */
static int64_t* decode_component(octx* c, rd r, uint32_t expected_dims, uint64_t max_records, uint64_t* n_out) {
    const uint8_t* magic = rd_bytes(&r, 4);
    const uint32_t dims = rd_u32(&r);
    const uint64_t n = rd_u64(&r);
    REQUIRE(c, dims == expected_dims && n <= max_records, "LCPS component shape mismatch");
    const size_t nv = mul_chk(c, (size_t)n, dims);
    int64_t* res = (int64_t*)ax(c, (nv ? nv : 1) * sizeof(int64_t));
    *n_out = n;
    if (!memcmp(magic, "RAW1", 4)) {
        REQUIRE(c, r.n == mul_chk(c, nv, 8), "RAW1 matrix size mismatch");
        for (size_t k = 0; k < nv; ++k) res[k] = rd_i64(&r);
        rd_end(&r);
        return res;
    }
    if (!memcmp(magic, "RVMX", 4) || !memcmp(magic, "RANS", 4)) {
        rd sec = rd_blob(&r);
        rd_end(&r);
        uint64_t* v = magic[1] == 'V' ? rvn1_decode(c, &sec, nv) : rvc1_decode(c, &sec, nv);
        for (size_t k = 0; k < nv; ++k) res[k] = (int64_t)v[k];
        return res;
    }
    if (!memcmp(magic, "ZFMX", 4) || !memcmp(magic, "FIXP", 4)) {
        const int z = magic[0] == 'Z';
        int64_t* mn = (int64_t*)ax(c, dims * sizeof(int64_t));
        int* w = (int*)ax(c, dims * sizeof(int));
        uint64_t* sz = (uint64_t*)ax(c, dims * sizeof(uint64_t));
        size_t total = 0;
        for (uint32_t d = 0; d < dims; ++d) {
            mn[d] = rd_i64(&r); w[d] = rd_u8(&r); sz[d] = rd_u64(&r);
            REQUIRE(c, w[d] >= 1 && w[d] <= 64, "invalid fixed width");
            const size_t bits = mul_chk(c, (size_t)n, (size_t)w[d]);
            REQUIRE(c, sz[d] == bits / 8 + (bits % 8 != 0), "invalid fixed payload size");
            REQUIRE(c, sz[d] <= SIZE_MAX - total, "fixed payload size overflow");
            total += sz[d];
        }
        const uint8_t* raw;
        if (z) {
            REQUIRE(c, rd_u64(&r) == total, "fixed raw size mismatch");
            rd zb = rd_blob(&r);
            raw = zstd_decompress(c, zb.p, zb.n, total);
        } else {
            raw = rd_bytes(&r, total);
        }
        rd_end(&r);
        size_t off = 0;
        for (uint32_t d = 0; d < dims; ++d) {
            size_t pos = 0;
            for (uint64_t i = 0; i < n; ++i) {
                const uint64_t o = read_bits(c, raw + off, sz[d], &pos, w[d]);
                REQUIRE(c, o <= (uint64_t)INT64_MAX - (uint64_t)mn[d], "fixed value overflow");
                res[i * dims + d] = (int64_t)((uint64_t)mn[d] + o);
            }
            off += sz[d];
        }
        return res;
    }
    const int ztg = !memcmp(magic, "ZTG4", 4), rtg = !memcmp(magic, "RTG4", 4), tagr = !memcmp(magic, "TAGR", 4);
    REQUIRE(c, ztg || rtg || tagr, "unknown LCPS encoding");
    const int maxw = rd_u8(&r);
    rd tsec = rd_blob(&r);
    uint64_t* tags = tagr ? rvc1_decode(c, &tsec, n) : rvn1_decode(c, &tsec, n);
    size_t total_bits = 0;
    int actual = 0;
    for (uint64_t i = 0; i < n; ++i) {
        REQUIRE(c, tags[i] >= 1 && tags[i] <= 63, "invalid Tagged width");
        if ((int)tags[i] > actual) actual = (int)tags[i];
        const size_t bits = mul_chk(c, dims, (size_t)tags[i]);
        REQUIRE(c, bits <= SIZE_MAX - total_bits, "Tagged size overflow");
        total_bits += bits;
    }
    REQUIRE(c, maxw == actual, "Tagged maximum width mismatch");
    const size_t raw_size = total_bits / 8 + (total_bits % 8 != 0);
    const uint8_t* packed;
    if (ztg) {
        REQUIRE(c, rd_u64(&r) == raw_size, "Tagged raw size mismatch");
        rd zb = rd_blob(&r);
        packed = zstd_decompress(c, zb.p, zb.n, raw_size);
    } else {
        rd pb = rd_blob(&r);
        REQUIRE(c, pb.n == raw_size, "Tagged payload size mismatch");
        packed = pb.p;
    }
    rd_end(&r);
    size_t pos = 0;
    for (uint64_t i = 0; i < n; ++i)
        for (uint32_t d = 0; d < dims; ++d)
            res[i * dims + d] = (int64_t)read_bits(c, packed, raw_size, &pos, (int)tags[i]);
    return res;
}

/* serial.cpp:1182-1338 and lcps_cpp.cpp:1838-1995 */
/* Integer decode shared by every container: L2 inverse (serial.cpp:1230-1275)
 * and the backward chain pass (serial.cpp:1277-1302); comp[4] (block offsets)
 * becomes the quantised coordinates q. */
static void rebuild_q(octx* c, int64_t* comp[5], const uint64_t cn[5], uint32_t dims, uint64_t n, int64_t p,
                      int64_t wmin, int64_t wmax, const wrap_t wrap) {
    const int64_t* idrel = comp[0];
    const int64_t* dif = comp[1];
    const int64_t* kept = comp[2];
    const int64_t* choices = comp[3];
    int64_t* qz = comp[4];
    const uint64_t B = cn[3];

    /* L2 inverse, serial.cpp:1230-1275 */
    int64_t* compact = (int64_t*)ax(c, (B ? B : 1) * dims * sizeof(int64_t));
    uint64_t di = 0, ki = 0;
    const uint64_t range = (uint64_t)wrap.max_difference;
    for (uint64_t i = 0; i < B; ++i) {
        const int64_t ch = choices[i];
        REQUIRE(c, ch == 0 || ch == 1, "invalid choice2 value");
        if (ch == 0) {
            REQUIRE(c, ki < cn[2], "missing kept block");
            for (uint32_t d = 0; d < dims; ++d) {
                const int64_t v = kept[ki * dims + d];
                REQUIRE(c, v >= wmin && v <= wmax, "kept block outside wrap range");
                compact[i * dims + d] = v;
            }
            ++ki;
            continue;
        }
        REQUIRE(c, i > 0 && di < cn[1], "missing block difference or initial block");
        for (uint32_t d = 0; d < dims; ++d) {
            const int64_t ss = dif[di * dims + d];
            REQUIRE(c, ss >= 0, "negative ZigZag symbol");
            const uint64_t sym = (uint64_t)ss;
            const int64_t corr = (sym & 1u) ? -(int64_t)(sym >> 1) - 1 : (int64_t)(sym >> 1);
            REQUIRE(c, corr >= wrap.min_correction && corr <= wrap.max_correction, "block correction outside wrap range");
            const uint64_t prev = (uint64_t)(compact[(i - 1) * dims + d] - wmin);
            uint64_t off;
            if (corr >= 0) {
                const uint64_t a = (uint64_t)corr;
                off = prev >= a ? prev - a : range - (a - prev);
            } else {
                const uint64_t a = (uint64_t)(-(corr + 1)) + 1;
                off = prev >= range - a ? prev - (range - a) : prev + a;
            }
            compact[i * dims + d] = wmin + (int64_t)off;
        }
        ++di;
    }
    REQUIRE(c, di == cn[1] && ki == cn[2], "unused block differences or kept blocks");

    /* backward chain pass, serial.cpp:1277-1302 */
    uint64_t ci = B;
    for (uint64_t i = n; i-- > 0;) {
        const int64_t jump = idrel[i];
        REQUIRE(c, jump >= 0, "negative successor distance");
        uint64_t succ = i;
        if (jump == 0) { REQUIRE(c, ci > 0, "missing explicit block"); --ci; }
        else { REQUIRE(c, (uint64_t)jump < n - i, "successor outside point array"); succ += (uint64_t)jump; }
        for (uint32_t d = 0; d < dims; ++d) {
            const int64_t off = qz[i * dims + d];
            REQUIRE(c, off >= 0 && off < p, "invalid block offset");
            const int64_t block = jump == 0 ? compact[ci * dims + d] : qz[succ * dims + d] / p;
            REQUIRE(c, block <= (INT64_MAX - off) / p, "quantized coordinate overflow");
            qz[i * dims + d] = block * p + off;
        }
    }
    REQUIRE(c, ci == 0, "unused explicit blocks");
}

int oc_decode(const uint8_t* bytes, size_t len, double** out, uint64_t* n_out, uint32_t* dims_out) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    REQUIRE(c, fegetround() == FE_TONEAREST, "LCP2 decoding requires round-to-nearest floating-point mode");
    rd r = {c, bytes, len};
    const uint8_t* magic = rd_bytes(&r, 4);
    const int lcp2 = !memcmp(magic, "LCP2", 4), lcp3 = !memcmp(magic, "LCP3", 4), lch1 = !memcmp(magic, "LCH1", 4);
    REQUIRE(c, lcp2 || lcp3 || lch1, "expected a self-contained LCP2, LCP3 or LCH1 bitstream");
    const uint32_t dims = rd_u32(&r);
    const uint64_t n = rd_u64(&r);
    REQUIRE(c, dims > 0 && n > 0, "invalid LCP2 shape");
    const size_t nv = mul_chk(c, (size_t)n, dims);
    const int64_t p = rd_i64(&r);
    const double eb = rd_f64(&r);
    REQUIRE(c, p > 0 && eb > 0.0 && isfinite(2.0 * eb), "invalid LCP2 quantization parameters");
    const int64_t wmin = rd_i64(&r), wmax = rd_i64(&r);
    REQUIRE(c, wmin >= 0, "invalid LCP2 block range");
    const wrap_t wrap = make_wrap(c, wmin, wmax);
    REQUIRE(c, dims <= r.n / 8, "truncated LCP2 minima");
    double* mins = (double*)ax(c, dims * sizeof(double));
    for (uint32_t d = 0; d < dims; ++d) { mins[d] = rd_f64(&r); REQUIRE(c, isfinite(mins[d]), "invalid LCP2 coordinate minimum"); }
    int64_t* comp[5];
    uint64_t cn[5];
    const uint32_t cd[5] = {1, dims, dims, 1, dims};
    for (int k = 0; k < 5; ++k) {
        rd blob = rd_blob(&r);
        if (lcp3) REQUIRE(c, blob.n >= 4 && !memcmp(blob.p, "RAW1", 4), "LCP3 requires raw component streams");
        comp[k] = decode_component(c, blob, cd[k], n, &cn[k]);
    }
    const uint64_t exact_count = rd_u64(&r);
    rd exact = rd_blob(&r);
    rd_end(&r);
    REQUIRE(c, cn[0] == n && cn[4] == n, "LCP2 point count mismatch");
    REQUIRE(c, exact_count <= nv, "invalid LCP2 correction count");
    rebuild_q(c, comp, cn, dims, n, p, wmin, wmax, wrap);
    const int64_t* qz = comp[4];

    double* vals = (double*)ax(c, nv * sizeof(double));
    for (uint64_t i = 0; i < n; ++i)
        for (uint32_t d = 0; d < dims; ++d)
            vals[i * dims + d] = reconstruct(qz[i * dims + d], mins[d], eb);
    if (exact_count == 0) {
        REQUIRE(c, exact.n == 0, "unexpected exact correction bytes");
    } else {
        const size_t raw_size = mul_chk(c, (size_t)exact_count, 16);
        const uint8_t* cb;
        if (lcp2) cb = zstd_decompress(c, exact.p, exact.n, raw_size);
        else { REQUIRE(c, exact.n == raw_size, "raw exact correction size mismatch"); cb = exact.p; }
        rd cr = {c, cb, raw_size};
        uint64_t prev = 0;
        for (uint64_t k = 0; k < exact_count; ++k) {
            const uint64_t idx = rd_u64(&cr);
            const double v = rd_f64(&cr);
            REQUIRE(c, idx < nv && (k == 0 || idx > prev) && isfinite(v), "invalid exact coordinate correction");
            vals[idx] = v;
            prev = idx;
        }
        rd_end(&cr);
    }
    *out = (double*)dup_out(c, vals, nv * sizeof(double));
    *n_out = n; *dims_out = dims;
    arena_release(c);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* LCG2, the reference GPU compressor's container (cu/chain/lcps_gpu_cpp.cu:5-35).
 * The writer is test infrastructure for the reader: it stores given integer
 * streams with the Fixed encoding only (one Zstd section per component, raw
 * chunks of 65536 bytes, lcps_gpu_decompress_ans.cu:2356-2358) or as an empty
 * component; the Tagged and Raw encodings need nvCOMP ANS, whose format is
 * nvCOMP's own. The decoder restates lcg2::decode
 * (cu/decompress/lcps_gpu_decompress_ans.cu:2345-2452, 2700-2900): the same
 * integer streams as LCP2, float32 reconstruction
 * fmaf(float(q) + 0.5f, 2.0f * eb, minimum) (:2605-2625), no corrections. Its
 * sections may also be stored (codec 0, the inflated form of this project). */

static void lcg2_fixed_component(octx* c, const mat* m, bb* out) {
    bb comp = bb_make(c);
    if (m->n == 0) {
        bb_u8(&comp, 0); bb_u32(&comp, m->dims); bb_u64(&comp, 0); bb_u8(&comp, 0);
        bb_blob(out, &comp);
        return;
    }
    bb payload = bb_make(c);
    bb_u8(&comp, 2); bb_u32(&comp, m->dims); bb_u64(&comp, m->n);
    for (uint32_t d = 0; d < m->dims; ++d) {
        int64_t mn; const int w = fixed_width(m->v, m->n, m->dims, d, &mn);
        REQUIRE(c, mn >= 0 && mn <= INT32_MAX && w <= 31, "LCG2 values are non-negative int32 of at most 31 bits");
        const size_t before = payload.n;
        pack_fixed_dim(c, m->v, m->n, m->dims, d, mn, w, &payload);
        bb_u32(&comp, (uint32_t)(int32_t)mn); bb_u8(&comp, (uint8_t)w); bb_u64(&comp, payload.n - before);
    }
    bb_u8(&comp, 1);                         /* one section */
    bb_u8(&comp, 2);                         /* Zstd */
    bb_u64(&comp, payload.n);
    const uint64_t limit = 65536, nchunks = (payload.n + limit - 1) / limit;
    bb_u32(&comp, (uint32_t)nchunks);
    bb* frames = (bb*)ax(c, nchunks * sizeof(bb));
    for (uint64_t k = 0; k < nchunks; ++k) {
        const uint64_t ub = payload.n - k * limit < limit ? payload.n - k * limit : limit;
        frames[k] = bb_make(c);
        zstd_compress(c, payload.p + k * limit, ub, &frames[k]);
        bb_u64(&comp, ub); bb_u64(&comp, frames[k].n);
    }
    for (uint64_t k = 0; k < nchunks; ++k) bb_put(&comp, frames[k].p, frames[k].n);
    bb_blob(out, &comp);
}

/* one Zstd frame of the given bytes (for hand-built test containers) */
int oc_zstd_frame(const uint8_t* src, size_t n, uint8_t** out, size_t* out_len) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    bb z = bb_make(c);
    zstd_compress(c, src, n, &z);
    *out = (uint8_t*)dup_out(c, z.p, z.n); *out_len = z.n;
    arena_release(c);
    return 0;
}

int oc_write_lcg2(const oc_streams* s, float eb, const float* mins, uint8_t** out, size_t* out_len) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    REQUIRE(c, s->n <= INT32_MAX && s->p_eff <= INT32_MAX && s->wrap_max <= INT32_MAX, "LCG2 holds int32 fields");
    bb b = bb_make(c);
    bb_put(&b, "LCG2", 4); bb_u32(&b, 1); bb_u32(&b, s->dims); bb_u64(&b, s->n);
    uint32_t u;
    bb_u32(&b, (uint32_t)(int32_t)s->p_eff);
    memcpy(&u, &eb, 4); bb_u32(&b, u);
    bb_u32(&b, (uint32_t)(int32_t)s->wrap_min); bb_u32(&b, (uint32_t)(int32_t)s->wrap_max);
    for (uint32_t d = 0; d < s->dims; ++d) { memcpy(&u, &mins[d], 4); bb_u32(&b, u); }
    bb_u32(&b, 5);
    mat m[5]; components(s, m);
    static const char* names[5] = {"IDR0", "DIF0", "KEP0", "CH20", "REL0"};
    for (int k = 0; k < 5; ++k) { bb_put(&b, names[k], 4); lcg2_fixed_component(c, &m[k], &b); }
    *out = (uint8_t*)dup_out(c, b.p, b.n); *out_len = b.n;
    arena_release(c);
    return 0;
}

/* a section (lcps_gpu_decompress_ans.cu:2345-2375), or the stored form */
static const uint8_t* lcg2_section(octx* c, rd* r, uint64_t expected_raw) {
    const uint8_t codec = rd_u8(r);
    REQUIRE(c, codec != 1, "LCG2 nvCOMP ANS sections are not supported");
    REQUIRE(c, codec == 2 || codec == 0, "unexpected LCG2 section codec");
    const uint64_t raw = rd_u64(r);
    REQUIRE(c, raw == expected_raw && raw > 0, "LCG2 section raw length mismatch");
    if (codec == 0) return rd_bytes(r, (size_t)raw);
    const uint32_t nchunks = rd_u32(r);
    const uint64_t limit = 65536;
    REQUIRE(c, nchunks == raw / limit + (raw % limit != 0), "LCG2 section chunk count mismatch");
    uint64_t* ub = (uint64_t*)ax(c, (nchunks ? nchunks : 1) * 8);
    uint64_t* cb = (uint64_t*)ax(c, (nchunks ? nchunks : 1) * 8);
    uint64_t remaining = raw;
    for (uint32_t k = 0; k < nchunks; ++k) {
        ub[k] = rd_u64(r); cb[k] = rd_u64(r);
        REQUIRE(c, ub[k] == (remaining < limit ? remaining : limit) && cb[k] > 0, "invalid LCG2 chunk lengths");
        remaining -= ub[k];
    }
    uint8_t* outb = (uint8_t*)ax(c, (size_t)raw);
    uint64_t at = 0;
    for (uint32_t k = 0; k < nchunks; ++k) {
        const uint8_t* chunk = rd_bytes(r, (size_t)cb[k]);
        const uint8_t* part = zstd_decompress(c, chunk, (size_t)cb[k], (size_t)ub[k]);
        memcpy(outb + at, part, (size_t)ub[k]);
        at += ub[k];
    }
    return outb;
}

/* a component (lcps_gpu_decompress_ans.cu:2377-2424, 2484-2495) */
static int64_t* lcg2_component(octx* c, rd* top, const char* name, uint32_t dims, uint64_t max_n, uint64_t* n_out) {
    REQUIRE(c, rd_magic(top, name), "unexpected LCG2 component name");
    rd r = rd_blob(top);
    const uint8_t kind = rd_u8(&r);
    REQUIRE(c, kind <= 3, "unknown LCG2 encoding");
    REQUIRE(c, kind != 1 && kind != 3, "LCG2 nvCOMP ANS sections are not supported");
    const uint32_t cd = rd_u32(&r);
    const uint64_t n = rd_u64(&r);
    REQUIRE(c, cd == dims && n <= max_n, "LCG2 component shape mismatch");
    *n_out = n;
    int64_t* res = (int64_t*)ax(c, ((n * dims) != 0 ? n * dims : 1) * sizeof(int64_t));
    if (kind == 0) {
        REQUIRE(c, n == 0 && rd_u8(&r) == 0, "invalid empty LCG2 component");
        rd_end(&r);
        return res;
    }
    REQUIRE(c, n > 0, "nonempty LCG2 encoding with no records");
    int32_t mn[64]; int w[64]; uint64_t sz[64];
    REQUIRE(c, dims <= 64, "too many LCG2 dimensions");
    uint64_t total = 0;
    for (uint32_t d = 0; d < dims; ++d) {
        mn[d] = (int32_t)rd_u32(&r); w[d] = rd_u8(&r); sz[d] = rd_u64(&r);
        REQUIRE(c, mn[d] >= 0 && w[d] >= 1 && w[d] <= 31, "invalid LCG2 fixed field");
        REQUIRE(c, sz[d] == (n * (uint64_t)w[d] + 7) / 8, "LCG2 fixed length mismatch");
        total += sz[d];
    }
    REQUIRE(c, rd_u8(&r) == 1, "LCG2 Fixed needs one Zstd section");
    const uint8_t* raw = lcg2_section(c, &r, total);
    rd_end(&r);
    uint64_t off = 0;
    for (uint32_t d = 0; d < dims; ++d) {
        size_t pos = 0;
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t v = read_bits(c, raw + off, (size_t)sz[d], &pos, w[d]) + (uint64_t)mn[d];
            REQUIRE(c, v <= INT32_MAX, "LCG2 value overflow");
            res[i * dims + d] = (int64_t)v;
        }
        off += sz[d];
    }
    return res;
}

int oc_decode_lcg2(const uint8_t* bytes, size_t len, float** out, uint64_t* n_out, uint32_t* dims_out) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    REQUIRE(c, fegetround() == FE_TONEAREST, "LCG2 decoding requires round-to-nearest floating-point mode");
    rd r = {c, bytes, len};
    REQUIRE(c, rd_magic(&r, "LCG2"), "not an LCG2 stream");
    REQUIRE(c, rd_u32(&r) == 1, "unsupported LCG2 version");
    const uint32_t dims = rd_u32(&r);
    const uint64_t n = rd_u64(&r);
    const int32_t p = (int32_t)rd_u32(&r);
    float eb; uint32_t u = rd_u32(&r); memcpy(&eb, &u, 4);
    const int32_t wmin = (int32_t)rd_u32(&r), wmax = (int32_t)rd_u32(&r);
    REQUIRE(c, dims > 0 && dims <= 64 && n > 0 && n <= INT32_MAX && p > 0, "invalid LCG2 shape or block size");
    REQUIRE(c, eb > 0 && isfinite(2.0f * eb), "invalid LCG2 error bound");
    REQUIRE(c, wmin >= 0 && wmax >= wmin, "invalid LCG2 wrap range");
    float* mins = (float*)ax(c, dims * sizeof(float));
    for (uint32_t d = 0; d < dims; ++d) {
        u = rd_u32(&r); memcpy(&mins[d], &u, 4);
        REQUIRE(c, isfinite(mins[d]), "nonfinite LCG2 minimum");
    }
    REQUIRE(c, rd_u32(&r) == 5, "expected five LCG2 components");
    static const char* names[5] = {"IDR0", "DIF0", "KEP0", "CH20", "REL0"};
    const uint32_t cd[5] = {1, dims, dims, 1, dims};
    int64_t* comp[5];
    uint64_t cn[5];
    for (int k = 0; k < 5; ++k) comp[k] = lcg2_component(c, &r, names[k], cd[k], n, &cn[k]);
    rd_end(&r);
    REQUIRE(c, cn[0] == n && cn[4] == n, "LCG2 coordinate stream count mismatch");
    REQUIRE(c, cn[3] > 0 && cn[2] > 0 && cn[1] + cn[2] == cn[3], "LCG2 diff/keep count mismatch");
    const wrap_t wrap = make_wrap(c, wmin, wmax);
    rebuild_q(c, comp, cn, dims, n, p, wmin, wmax, wrap);
    const int64_t* qz = comp[4];
    const size_t nv = mul_chk(c, (size_t)n, dims);
    float* vals = (float*)ax(c, nv * sizeof(float));
    const float step = 2.0f * eb;
    for (size_t k = 0; k < nv; ++k) {
        const int64_t q = qz[k];
        REQUIRE(c, q >= 0 && q <= INT32_MAX, "LCG2 quantized coordinate outside int32");
        const float mid = (float)(int32_t)q + 0.5f;
        vals[k] = fmaf(mid, step, mins[k % dims]);
        REQUIRE(c, isfinite(vals[k]), "nonfinite LCG2 coordinate");
    }
    *out = (float*)dup_out(c, vals, nv * sizeof(float));
    *n_out = n; *dims_out = dims;
    arena_release(c);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Exposed pieces for tests */

int oc_pack_fixed(const int64_t* v, uint64_t n, uint32_t dims, int64_t* mins_out, uint8_t* widths_out,
                  uint8_t** out, size_t* out_len) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    bb payload = bb_make(c);
    for (uint32_t d = 0; d < dims; ++d) {
        int64_t mn; int w = fixed_width(v, n, dims, d, &mn);
        pack_fixed_dim(c, v, n, dims, d, mn, w, &payload);
        mins_out[d] = mn; widths_out[d] = (uint8_t)w;
    }
    *out = (uint8_t*)dup_out(c, payload.p, payload.n); *out_len = payload.n;
    arena_release(c);
    return 0;
}

int oc_pack_tagged(const int64_t* v, uint64_t n, uint32_t dims, uint8_t** tags_out, uint8_t** out,
                   size_t* out_len, uint32_t* max_width) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    uint64_t* tags = (uint64_t*)ax(c, (n ? n : 1) * sizeof(uint64_t));
    bb packed = bb_make(c);
    int mw = 0;
    tagged_parts(c, v, n, dims, tags, &mw, &packed);
    uint8_t* t8 = (uint8_t*)ax(c, n ? n : 1);
    for (uint64_t i = 0; i < n; ++i) t8[i] = (uint8_t)tags[i];
    *tags_out = (uint8_t*)dup_out(c, t8, n);
    *out = (uint8_t*)dup_out(c, packed.p, packed.n); *out_len = packed.n;
    *max_width = (uint32_t)mw;
    arena_release(c);
    return 0;
}

int oc_rvn1_encode(const uint64_t* values, uint64_t n, uint8_t** out, size_t* out_len) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    bb b = bb_make(c);
    rvn1_encode(c, values, n, &b);
    *out = (uint8_t*)dup_out(c, b.p, b.n); *out_len = b.n;
    arena_release(c);
    return 0;
}

int oc_rvn1_decode(const uint8_t* bytes, size_t len, uint64_t n, uint64_t** out) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    rd r = {c, bytes, len};
    uint64_t* v = rvn1_decode(c, &r, n);
    *out = (uint64_t*)dup_out(c, v, n * sizeof(uint64_t));
    arena_release(c);
    return 0;
}

int oc_rvn1_table(const uint64_t* values, uint64_t n, uint32_t** probs, uint32_t* alphabet, int* precision_bits) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    rtable t = rvn1_build_table(c, values, n);
    *probs = (uint32_t*)dup_out(c, t.prob, t.alphabet * sizeof(uint32_t));
    *alphabet = t.alphabet;
    *precision_bits = t.precision_bits;
    arena_release(c);
    return 0;
}

int oc_rvc1_encode(const uint64_t* values, uint64_t n, uint32_t lanes, uint32_t chunk_symbols,
                   uint8_t** out, size_t* out_len) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    bb b = bb_make(c);
    rvc1_encode(c, values, n, lanes, chunk_symbols, &b);
    *out = (uint8_t*)dup_out(c, b.p, b.n); *out_len = b.n;
    arena_release(c);
    return 0;
}

int oc_rvc1_decode(const uint8_t* bytes, size_t len, uint64_t n, uint64_t** out) {
    octx cx; cx.top = NULL; octx* c = &cx;
    if (setjmp(cx.jb)) { arena_release(c); return -1; }
    rd r = {c, bytes, len};
    uint64_t* v = rvc1_decode(c, &r, n);
    *out = (uint64_t*)dup_out(c, v, (n ? n : 1) * sizeof(uint64_t));
    arena_release(c);
    return 0;
}
