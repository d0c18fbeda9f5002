// Request server of the Chain engine: continuous batching on one GPU.
//
// Callers submit compress and decompress requests at any time. W worker
// threads each own an engine (its own HIP stream and arenas), so the batches
// of different workers overlap on the device. A worker that becomes free
// takes the oldest waiting request and every later waiting request of the
// same kind and shape (dtype, layout, dims), in arrival order, up to the
// batch limits and its share of the HBM budget, and runs them through one
// launch per stage. Requests that arrive while batches are in flight wait
// only until a worker frees up; nothing waits for a batch to fill.
//
// HBM admission: the budget is divided evenly between the workers, and each
// worker's share is allocated once, at creation, as its two arenas: three
// quarters for compression, one quarter for decompression (device memory
// stays resident; nothing is allocated or freed while serving). A batch is
// formed only while the estimated device bytes of its requests fit in the
// arena of its kind; a request larger than that is refused with
// CHAIN_ENOMEM at submission.
//
// The integer and bitstream arithmetic is that of the engine: a request's
// result does not depend on the batch it ran in. If a batch fails, each of
// its requests is run again alone so that one bad request cannot fail others.
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

namespace chain {

struct ServeItem {
    uint64_t ticket;
    chain_request req;
    uint64_t est;          // estimated device bytes
    uint32_t dims;         // dims (decompress: from the stream header)
    double t_submit;
};

struct ServeDone {
    int status = CHAIN_OK;
    std::string error;
    chain_buf out{nullptr, 0};
    uint64_t n = 0;
    uint32_t dims = 0;
    double t_submit = 0, t_start = 0, t_end = 0;
    uint32_t batch = 0;
};

// Upper bound of the arena bytes of one frame's compression in a first run
// (compress_batch_try in chain_engine.hip, buffer by buffer): the host
// frame's copy; keys as 8-byte words (D words, or one word and the
// (tile, key) items), two sort buffers of keys and indices, rel (one byte
// when p <= 256, else four), idrel, block records, choice2, diff, keep and
// the radix counters; CORR_TILE_CAP corrections per tile twice; the rANS
// chunk scratch of the 3D + 7 sections' symbols; the output bound of
// out_bound_frame at the widest values a request allows (idrel 32 bits,
// diff and keep 64, rel bitlen(p - 1)); the per-frame tables (histograms,
// rANS tables, statistics slots) and the 256-byte rounding of every buffer.
// A rare second run with worst-case bounds (a tile with more corrections,
// an output over its bound) may need more; in a worker whose arena it
// exceeds it fails with CHAIN_ENOMEM.
inline uint64_t serve_estimate_compress(uint64_t n, uint32_t D, uint32_t es, int64_t p, uint32_t flags,
                                        const Tuning& rc) {
    const uint64_t pe = (flags & CHAIN_NO_BL) ? 1 : (uint64_t)p;
    const uint64_t rbytes = pe <= 256 ? 1 : 4;
    const uint64_t cs = rc.chunk_symbols;
    const uint64_t stride = CHUNK_STATE_BYTES + ((3 * cs + 255) / 256) * 256;
    const uint64_t tiles = (n + CH_TILE - 1) / CH_TILE;
    uint64_t b = n * D * es;
    b += n * (std::max<uint64_t>(24, 8ull * D) + 24 + D * rbytes + 4 + 24ull * D + 1) + n / 2 + n / 16;
    b += 4 * tiles * CORR_TILE_CAP * 8;
    const uint64_t chunks = ((3ull * D + 7) * n + cs - 1) / cs + 10;
    b += chunks * (stride + 12);
    const uint64_t relb = std::max<uint64_t>(1, bitlen64(pe - 1));
    uint64_t out = 2 * (8 + 16 + 17) + 3 * (8 + 16 + 17ull * D) + 4 * n + D * (8 * n + 1) + (n + 7) / 8 +
                   D * ((n * relb + 7) / 8);
    out += out / 32 + 5 * 4096;
    out += 48 + 8ull * D + 16 + 16 * std::min<uint64_t>(n * D, tiles * CORR_TILE_CAP);
    b += out;
    b += 5ull * RANS_VALUE_LIMIT * 4 + 5ull * RANS_TAG_LIMIT * 4 + 10ull * RANS_VALUE_LIMIT * 8 +
         10ull * TABLE_BYTES_MAX + 5ull * (CS_GRID + 1) * sizeof(CompSlot) + (1u << 20);
    return b;
}

// Decode carve per point, worst case: idrel, root, list_a, list_b (4 B each),
// rel (4 D), choice2 (1), diff + keep (8 D together: one record each per
// explicit block), compact blocks (8 D), Tagged widths (at most 2 per point:
// idrel and rel are per point, the three record components together hold at
// most 2 R <= 2 T widths) and the output (D es). The stream pool takes len;
// the parsed tables (a 24-byte chunk record per >= 4 stream bytes, the
// cumulative tables of at most 10 sections of 4097 entries) are bounded by
// 6 len + 10 * 16.5 KB; per-tile look-back and count words add < 1 B per point;
// every carved buffer is rounded up to 256 bytes.
inline uint64_t serve_estimate_decompress(uint64_t n, uint32_t D, uint32_t es_out, uint64_t len) {
    return n * (16ull + 1 + 2 + 1 + 20ull * D + (uint64_t)D * es_out) + 7 * len + 10 * 16512 + (1u << 20);
}

struct Server {
    int device = 0;
    uint64_t budget = 0, share = 0;
    uint64_t cap_c = 0, cap_d = 0;   // per-worker arena capacity per kind
    uint32_t max_batch = 64;
    uint64_t max_batch_points = 1ull << 27;
    std::vector<chain_engine*> engines;
    std::vector<std::thread> threads;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeItem> queue;
    std::unordered_map<uint64_t, ServeDone> done;
    uint64_t next_ticket = 1;
    uint64_t in_flight = 0;   // admitted, not yet done
    bool stop = false;
    std::string err;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    static bool same_class(const ServeItem& a, const ServeItem& b) {
        if (a.req.kind != b.req.kind || a.dims != b.dims) return false;
        if (a.req.kind == CHAIN_REQ_COMPRESS)
            return a.req.dtype == b.req.dtype && a.req.layout == b.req.layout && a.req.format == b.req.format;
        return a.req.out_dtype == b.req.out_dtype && a.req.out_layout == b.req.out_layout;
    }

    // called with mu held: the next batch for a free worker (empty if none)
    std::vector<ServeItem> take_batch() {
        std::vector<ServeItem> batch;
        if (queue.empty()) return batch;
        uint64_t bytes = 0, points = 0;
        for (auto it = queue.begin(); it != queue.end();) {
            const ServeItem& q = *it;
            const bool first = batch.empty();
            if (!first && !same_class(batch.front(), q)) {
                ++it;
                continue;
            }
            const uint64_t cap = q.req.kind == CHAIN_REQ_COMPRESS ? cap_c : cap_d;
            if (!first && (batch.size() >= max_batch || bytes + q.est > cap || points + q.req.n > max_batch_points))
                break;
            bytes += q.est;
            points += q.req.n;
            batch.push_back(q);
            it = queue.erase(it);
        }
        return batch;
    }

    void run_batch(chain_engine* e, std::vector<ServeItem>& batch, double t_start) {
        const uint32_t F = (uint32_t)batch.size();
        std::vector<chain_buf> outs(F, chain_buf{nullptr, 0});
        std::vector<uint64_t> ns(F, 0);
        uint32_t dims = batch[0].dims;
        int status = CHAIN_OK;
        std::string msg;
        try {
            (void)hipSetDevice(e->device);
            if (batch[0].req.kind == CHAIN_REQ_COMPRESS) {
                std::vector<chain_frame> fr(F);
                for (uint32_t k = 0; k < F; ++k) {
                    const chain_request& r = batch[k].req;
                    fr[k] = chain_frame{r.points, r.n, r.eb, r.p, r.threshold, r.flags};
                    ns[k] = r.n;
                }
                CompressResult res;
                res.host_out = outs.data();
                compress_batch_impl(e, fr.data(), F, batch[0].req.dtype, batch[0].req.layout, dims, false,
                                    batch[0].req.format, res);
                if (!res.containers.empty())
                    for (uint32_t k = 0; k < F; ++k) {
                        outs[k].data = (uint8_t*)std::malloc(res.containers[k].size() + 1);
                        if (!outs[k].data) throw std::bad_alloc();
                        std::memcpy(outs[k].data, res.containers[k].data(), res.containers[k].size());
                        outs[k].size = res.containers[k].size();
                    }
            } else {
                std::vector<const uint8_t*> bytes(F);
                std::vector<size_t> lens(F);
                for (uint32_t k = 0; k < F; ++k) {
                    bytes[k] = batch[k].req.bytes;
                    lens[k] = batch[k].req.len;
                }
                DecodeResult res;
                res.host_out = outs.data();
                decompress_batch_impl(e, bytes.data(), lens.data(), F, batch[0].req.out_dtype,
                                      batch[0].req.out_layout, res);
                for (uint32_t k = 0; k < F; ++k) ns[k] = res.ns[k];
                dims = res.dims;
            }
        } catch (const Error& x) {
            status = x.code;
            msg = x.what();
        } catch (const std::bad_alloc&) {
            status = CHAIN_ENOMEM;
            msg = "out of host memory";
        } catch (const std::exception& x) {
            status = CHAIN_ERUNTIME;
            msg = x.what();
        }
        if (status != CHAIN_OK && F > 1) {
            // isolate the failing request(s): every request alone
            for (uint32_t k = 0; k < F; ++k) chain_buf_free(&outs[k]);
            for (uint32_t k = 0; k < F; ++k) {
                std::vector<ServeItem> one{batch[k]};
                run_batch(e, one, t_start);
            }
            return;
        }
        const double t_end = now();
        std::lock_guard<std::mutex> lock(mu);
        for (uint32_t k = 0; k < F; ++k) {
            ServeDone d;
            d.status = status;
            d.error = msg;
            d.out = status == CHAIN_OK ? outs[k] : chain_buf{nullptr, 0};
            if (status != CHAIN_OK) chain_buf_free(&outs[k]);
            d.n = ns[k];
            d.dims = dims;
            d.t_submit = batch[k].t_submit;
            d.t_start = t_start;
            d.t_end = t_end;
            d.batch = F;
            done[batch[k].ticket] = d;
        }
        in_flight -= F;
        cv_done.notify_all();
    }

    void worker(chain_engine* e) {
        (void)hipSetDevice(device);
        for (;;) {
            std::vector<ServeItem> batch;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv_work.wait(lock, [&] { return stop || !queue.empty(); });
                if (stop && queue.empty()) return;
                batch = take_batch();
            }
            if (batch.empty()) continue;
            // the engine mutex keeps external users of this engine (none)
            // and this worker apart
            std::lock_guard<std::mutex> elock(e->mu);
            run_batch(e, batch, now());
        }
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lock(mu);
            stop = true;
        }
        cv_work.notify_all();
        for (auto& t : threads)
            if (t.joinable()) t.join();
        for (auto* e : engines) chain_engine_destroy(e);
        for (auto& kv : done) chain_buf_free(&kv.second.out);
    }
};

}  // namespace chain

struct chain_server : chain::Server {};

extern "C" {

int chain_server_create(int device, uint64_t hbm_budget_bytes, uint32_t workers, uint32_t max_batch,
                        chain_server** out) {
    if (!out) return CHAIN_EINVAL;
    *out = nullptr;
    if (workers == 0 || workers > 16 || hbm_budget_bytes == 0) {
        g_create_err = "the server needs 1 to 16 workers and a non-zero HBM budget";
        return CHAIN_EINVAL;
    }
    auto* s = new (std::nothrow) chain_server();
    if (!s) return CHAIN_ENOMEM;
    s->device = device;
    s->budget = hbm_budget_bytes;
    s->share = hbm_budget_bytes / workers;
    s->cap_d = s->share / 4;
    s->cap_c = s->share - s->cap_d;
    s->max_batch = max_batch ? max_batch : 64;
    for (uint32_t w = 0; w < workers; ++w) {
        chain_engine* e = nullptr;
        int rc = chain_engine_create(device, s->share, &e);
        if (rc == CHAIN_OK) {
            s->engines.push_back(e);
            e->shared_budget = true;
            e->timer.on = false;
            rc = guarded(e, [&] {
                e->arena.ensure(s->cap_c, s->cap_c);
                e->arena_d.ensure(s->cap_d, s->cap_d);
            });
            if (rc != CHAIN_OK) g_create_err = e->err;
        }
        if (rc != CHAIN_OK) {
            delete s;
            return rc;
        }
    }
    for (chain_engine* e : s->engines) s->threads.emplace_back([s, e] { s->worker(e); });
    *out = s;
    return CHAIN_OK;
}

void chain_server_destroy(chain_server* s) { delete s; }

const char* chain_server_last_error(const chain_server* s) { return s ? s->err.c_str() : g_create_err.c_str(); }

int chain_serve_submit(chain_server* s, const chain_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket) return CHAIN_EINVAL;
    std::lock_guard<std::mutex> lock(s->mu);
    auto fail = [&](int code, const char* m) {
        s->err = m;
        return code;
    };
    chain::ServeItem it;
    it.req = *r;
    it.t_submit = s->now();
    if (r->kind == CHAIN_REQ_COMPRESS) {
        if (!r->points || r->n == 0) return fail(CHAIN_EINVAL, "compress request without points");
        if (r->dims < 1 || r->dims > CH_MAXD) return fail(CHAIN_EINVAL, "the GPU engine supports 1 to 4 dimensions");
        if (r->dtype != CHAIN_F32 && r->dtype != CHAIN_F64) return fail(CHAIN_EINVAL, "dtype must be float32 or float64");
        if (r->layout != CHAIN_AOS && r->layout != CHAIN_SOA) return fail(CHAIN_EINVAL, "layout must be AoS or SoA");
        if (r->format != CHAIN_FMT_LCH1 && r->format != CHAIN_FMT_LCP3) return fail(CHAIN_EINVAL, "unknown format");
        if (!(r->eb > 0.0) || !std::isfinite(2.0 * r->eb)) return fail(CHAIN_EINVAL, "eb must be positive with finite 2*eb");
        if (r->p <= 0) return fail(CHAIN_EINVAL, "p must be positive");
        it.dims = r->dims;
        it.est = chain::serve_estimate_compress(r->n, r->dims, r->dtype == CHAIN_F32 ? 4 : 8, r->p, r->flags,
                                                s->engines[0]->tuning);
    } else if (r->kind == CHAIN_REQ_DECOMPRESS) {
        if (!r->bytes || r->len < 24) return fail(CHAIN_EINVAL, "decompress request without a stream");
        if (r->out_dtype != CHAIN_F32 && r->out_dtype != CHAIN_F64) return fail(CHAIN_EINVAL, "dtype must be float32 or float64");
        if (r->out_layout != CHAIN_AOS && r->out_layout != CHAIN_SOA) return fail(CHAIN_EINVAL, "layout must be AoS or SoA");
        uint32_t D = 0;
        uint64_t n = 0;
        for (int i = 0; i < 4; ++i) D |= (uint32_t)r->bytes[4 + i] << (8 * i);
        for (int i = 0; i < 8; ++i) n |= (uint64_t)r->bytes[8 + i] << (8 * i);
        if (D < 1 || D > CH_MAXD) return fail(CHAIN_EINVAL, "the GPU engine supports 1 to 4 dimensions");
        it.dims = D;
        it.req.n = n;
        it.est = chain::serve_estimate_decompress(n, D, r->out_dtype == CHAIN_F32 ? 4 : 8, r->len);
    } else {
        return fail(CHAIN_EINVAL, "unknown request kind");
    }
    if (it.est > (r->kind == CHAIN_REQ_COMPRESS ? s->cap_c : s->cap_d))
        return fail(CHAIN_ENOMEM, "request exceeds one worker's share of the HBM budget");
    it.ticket = s->next_ticket++;
    *ticket = it.ticket;
    s->queue.push_back(it);
    ++s->in_flight;
    s->cv_work.notify_one();
    return CHAIN_OK;
}

int chain_serve_poll(chain_server* s, uint64_t ticket, int wait_ms, chain_result* out) {
    if (!s || !out) return CHAIN_EINVAL;
    std::unique_lock<std::mutex> lock(s->mu);
    auto ready = [&] { return s->done.count(ticket) != 0; };
    if (!ready()) {
        if (wait_ms < 0) s->cv_done.wait(lock, ready);
        else if (wait_ms > 0) s->cv_done.wait_for(lock, std::chrono::milliseconds(wait_ms), ready);
        if (!ready()) return CHAIN_PENDING;
    }
    chain::ServeDone d = s->done[ticket];
    s->done.erase(ticket);
    out->status = d.status;
    out->out = d.out;
    out->n = d.n;
    out->dims = d.dims;
    out->t_submit = d.t_submit;
    out->t_start = d.t_start;
    out->t_end = d.t_end;
    out->batch = d.batch;
    std::strncpy(out->error, d.error.c_str(), sizeof out->error - 1);
    out->error[sizeof out->error - 1] = 0;
    return CHAIN_OK;
}

double chain_server_clock(chain_server* s) { return s ? s->now() : 0.0; }

uint64_t chain_serve_estimate_compress(uint64_t n, uint32_t dims, int dtype, int64_t p, uint32_t flags,
                                       uint32_t chunk_symbols) {
    if (n == 0 || dims < 1 || dims > CH_MAXD || p <= 0 || chunk_symbols == 0) return 0;
    chain::Tuning rc;
    rc.chunk_symbols = chunk_symbols;
    return chain::serve_estimate_compress(n, dims, dtype == CHAIN_F64 ? 8 : 4, p, flags, rc);
}

int chain_engine_memory_stats(chain_engine* e, uint64_t* compress_need, uint64_t* capacity_retries) {
    if (!e || !compress_need || !capacity_retries) return CHAIN_EINVAL;
    std::lock_guard<std::mutex> lock(e->mu);
    *compress_need = e->last_need_c;
    *capacity_retries = e->capacity_retries;
    return CHAIN_OK;
}

int chain_server_stats(chain_server* s, uint64_t* arena_allocations) {
    if (!s || !arena_allocations) return CHAIN_EINVAL;
    uint64_t a = 0;
    for (chain_engine* e : s->engines) {
        std::lock_guard<std::mutex> lock(e->mu);
        a += e->arena.allocs + e->arena_d.allocs;
    }
    *arena_allocations = a;
    return CHAIN_OK;
}

}  // extern "C"
