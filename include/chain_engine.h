/*
 * chain_engine.h -- C ABI of the MI355X (gfx950) Chain compressor engine.
 *
 * This is the drop-in boundary below the Python modules that replace the
 * reference's pybind11 extensions:
 *   lcps_gpu_cpp.lcps_compress2_gpu            cu/chain/lcps_gpu_cpp.cu:1095-1102
 *   lcps_gpu_decompress_ans.lcps_decompress_gpu cu/decompress/lcps_gpu_decompress_ans.cu:2898-2921
 * and, for the reference CPU format, lcps_cpp.lcps_compress2
 * (cpp/chain/lcps_cpp.cpp:2340-2355, ec=False -> LCP3).
 *
 * Conventions: plain pointers and sizes, integer status codes (0 = success),
 * no exceptions across the ABI, library-allocated outputs released with
 * chain_buf_free(). One engine owns one HIP stream (default priority) on one
 * device and a reusable device arena; calls on one engine are serialised by
 * an internal mutex.
 */
#ifndef CHAIN_ENGINE_H
#define CHAIN_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chain_engine chain_engine;

enum chain_status {
    CHAIN_OK = 0,
    CHAIN_EINVAL = 1,      /* invalid argument (the reference raises ValueError)      */
    CHAIN_ERUNTIME = 2,    /* data or format error (the reference raises RuntimeError) */
    CHAIN_EDEVICE = 3,     /* HIP runtime error or device error word                  */
    CHAIN_ENOMEM = 4,      /* device memory budget exceeded                           */
    CHAIN_PENDING = 5,     /* chain_serve_poll: the request has not finished yet      */
};

enum chain_dtype { CHAIN_F32 = 0, CHAIN_F64 = 1 };
enum chain_layout { CHAIN_AOS = 0, /* (N, D) */ CHAIN_SOA = 1 /* (D, N) */ };
enum chain_format {
    CHAIN_FMT_LCH1 = 0,    /* this engine's entropy-coded container (chunked interleaved rANS) */
    CHAIN_FMT_LCP3 = 1,    /* the reference's raw container, ec=False (lcps_cpp.cpp:2306-2329) */
};
/* frame flags: the reference's ablation switches, lcps_cpp.cpp:2023-2026 */
enum chain_flags { CHAIN_NO_BL = 1u, CHAIN_NO_L1 = 2u, CHAIN_NO_L2 = 4u };

typedef struct {
    uint8_t* data;   /* malloc'd by the library */
    size_t size;
} chain_buf;

/* one frame of a batch */
typedef struct {
    const void* points;    /* host pointer (or device pointer, see chain_compress_batch) */
    uint64_t n;            /* particles */
    double eb;             /* absolute error bound, > 0 */
    int64_t p;             /* block size in quantisation steps, > 0 */
    int64_t threshold;     /* chain threshold in bits */
    uint32_t flags;        /* chain_flags */
} chain_frame;

/* per-frame statistics, as lcps_compress2's return list (lcps_cpp.cpp:2331-2337) */
typedef struct {
    uint64_t bytes;                  /* container size */
    uint64_t block_replacement_size; /* diff + keep + choice2 component bytes */
    uint64_t idrel_size;
    uint64_t rel_size;
    uint64_t num_chained;            /* N - B */
    uint64_t num_corrections;
} chain_stats;

int chain_engine_create(int device, uint64_t hbm_budget_bytes, chain_engine** out);
void chain_engine_destroy(chain_engine* e);
/* message of the last failure on this engine (or of chain_engine_create when e is NULL) */
const char* chain_last_error(const chain_engine* e);
/* the engine's hipStream_t, for callers that time or order work around it */
void* chain_engine_stream(chain_engine* e);
void chain_buf_free(chain_buf* b);

/* Compress one frame from host memory. Replaces lcps_compress2_gpu
 * (cu/chain/lcps_gpu_cpp.cu:718-1091) and, with CHAIN_FMT_LCP3,
 * lcps_compress2(..., ec=False) (cpp/chain/lcps_cpp.cpp:2017-2338). */
int chain_compress(chain_engine* e, const void* points, int dtype, int layout, uint32_t dims, uint64_t n,
                   double eb, int64_t p, int64_t threshold, uint32_t flags, int format,
                   chain_buf* out, chain_stats* stats);

/* Compress a batch of frames (same dtype, layout and dims) through one launch
 * per stage. points_on_device != 0: every frame's points pointer is a device
 * pointer. outs/stats have nframes entries. */
int chain_compress_batch(chain_engine* e, const chain_frame* frames, uint32_t nframes, int dtype, int layout,
                         uint32_t dims, int points_on_device, int format, chain_buf* outs, chain_stats* stats);

/* Device-resident batch compression (LCH1): the frames' points are device
 * pointers (contiguous frames are read in place), the containers are written
 * into a device buffer owned by the engine and valid until its next call;
 * offsets/sizes give each frame's container inside *d_out. */
int chain_compress_batch_device(chain_engine* e, const chain_frame* frames, uint32_t nframes, int dtype, int layout,
                                uint32_t dims, void** d_out, uint64_t* offsets, uint64_t* sizes, chain_stats* stats);

/* (p, threshold) search on the device: compresses one frame (LCH1) under
 * every pair ps[i] x ts[j] and reports each configuration's statistics in
 * stats[i * nts + j] (bytes = container size; num_chained etc. as
 * lcps_compress2 returns them), as the reference harness's parameter sweeps
 * do with one CPU call per pair (reader.py:80-103, 227-248). The pairs run as
 * the frames of batches, so a sweep costs a few launches per stage. */
int chain_compress_search(chain_engine* e, const void* points, int dtype, int layout, uint32_t dims, uint64_t n,
                          double eb, const int64_t* ps, uint32_t nps, const int64_t* ts, uint32_t nts, uint32_t flags,
                          int points_on_device, chain_stats* stats);

/* RVC1 parameters of the containers this engine writes: the largest lane
 * count per chunk (power of two <= 64) and the symbols per chunk. Both are
 * recorded in every section, so any decoder reads any setting. */
int chain_engine_set_rans(chain_engine* e, uint32_t lanes_max, uint32_t chunk_symbols);

/* Decompress one LCH1 or LCP3 stream to float64 (or float32) coordinates.
 * Replaces lcps_decompress_gpu (cu/decompress/lcps_gpu_decompress_ans.cu:2701-2921)
 * and, for the reference's formats, lcps_decompress_serial
 * (cpp/decompress/lcps_cpu_decompress_serial.cpp:1377-1381). */
int chain_decompress(chain_engine* e, const uint8_t* bytes, size_t len, int out_dtype, int out_layout,
                     chain_buf* out, uint64_t* n, uint32_t* dims);

/* Batch decompression: one launch per stage over all streams. */
int chain_decompress_batch(chain_engine* e, const uint8_t* const* bytes, const size_t* lens, uint32_t nstreams,
                           int out_dtype, int out_layout, chain_buf* outs, uint64_t* ns, uint32_t* dims);

/* Device-resident batch decompression: bytes/lens are host copies of the
 * streams (their headers plan the launch) and d_streams holds the same
 * streams concatenated in device memory (NULL: upload the host copies). The
 * coordinates stay in an engine-owned device buffer, valid until the next
 * decompression call; out_offsets are element offsets of every frame. */
int chain_decompress_batch_device(chain_engine* e, const uint8_t* const* bytes, const size_t* lens,
                                  const void* d_streams, uint32_t nstreams, int out_dtype, int out_layout,
                                  void** d_out, uint64_t* out_offsets, uint64_t* ns, uint32_t* dims);

/* Decompression that allocates nothing. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same inputs as chain_decompress_batch_device (host copies of the streams
 * plan the launches; d_streams, required, holds the same streams concatenated
 * in device memory), but the coordinates go to the caller's device buffer
 * d_out of out_capacity bytes (frames back to back; out_offsets are element
 * offsets) and every intermediate buffer is carved from the library's fixed
 * device scratch of CHAIN_NOALLOC_SCRATCH_BYTES per GPU. A batch that needs
 * more scratch fails with CHAIN_ENOMEM, an output buffer that is too small with
 * CHAIN_EINVAL. The call makes no device or pinned-host allocation or release
 * (no hipMalloc, hipFree, hipHostMalloc, hipEventCreate). Every call on one
 * GPU uses that GPU's one scratch region, which keeps the last call's
 * contents: calls that overlap on one GPU (from several engines or threads)
 * race and must be serialised by the caller. */
#define CHAIN_NOALLOC_SCRATCH_BYTES (256ull << 20)
int chain_decompress_batch_noalloc(chain_engine* e, const uint8_t* const* bytes, const size_t* lens,
                                   const void* d_streams, uint32_t nstreams, int out_dtype, int out_layout,
                                   void* d_out, uint64_t out_capacity, uint64_t* out_offsets, uint64_t* ns,
                                   uint32_t* dims);

/* Device buffers for callers that keep frames in HBM without another
 * runtime (synchronous copies on the engine's stream; to_device: 1 host to
 * device, 0 device to host, 2 device to device). */
int chain_device_malloc(chain_engine* e, uint64_t bytes, void** out);
int chain_device_free(chain_engine* e, void* p);
int chain_copy(chain_engine* e, void* dst, const void* src, uint64_t bytes, int to_device);

/* Tunable settings: the RVC1 layout of the containers an engine writes
 * (largest lane count per chunk, symbols per chunk; both recorded in every
 * section, so any decoder reads any setting) and the L1 strategy (l1_tiles=1:
 * successors inside 4096-point tiles plus a sort of one item per (tile,
 * key); 0: one radix sort of every point). Only the RVC1 layout shows in the
 * output; the integer streams are the same under every setting. */
typedef struct {
    uint32_t lanes_max;
    uint32_t chunk_symbols;
    uint32_t l1_tiles;
    uint32_t reserved;
} chain_tuning;
int chain_engine_set_tuning(chain_engine* e, const chain_tuning* t);
int chain_engine_get_tuning(chain_engine* e, chain_tuning* t);

/* Autotuning: compress + decompress (a sample of up to 2^24 points of) the
 * given batch under candidate settings, keep the fastest whose output is
 * within 2 % of the smallest, apply it to the engine and remember it for the
 * batch's size bucket (log2 of the mean frame size and of the frame count).
 * The search is coordinate-wise: 8192, 16384, 32768 and 65536 symbols per
 * chunk (64 lanes, tile-local L1), then 32 lanes and then the global-sort L1
 * at the best setting so far. trials / trial_ms / trial_bytes (up to
 * max_trials entries each, may be NULL) report every trial, *ntrials their
 * count. With chain_engine_set_autotune(e, 1), every compression whose bucket
 * has no setting yet runs this first. Containers stay valid under any setting. */
int chain_engine_autotune(chain_engine* e, const chain_frame* frames, uint32_t nframes, int dtype, int layout,
                          uint32_t dims, int points_on_device, chain_tuning* best, chain_tuning* trials,
                          float* trial_ms, uint64_t* trial_bytes, uint32_t max_trials, uint32_t* ntrials);
int chain_engine_set_autotune(chain_engine* e, int on);

/* Compression through a HIP graph: every launch after the planning
 * synchronisation is captured once and the executable graph is replayed
 * while batches keep the same launch shape (frame sizes, dims, dtype, input
 * and arena addresses); a new shape is captured again and the executable
 * graph updated in place. Results are identical either way. Off by default. */
int chain_engine_set_graphs(chain_engine* e, int on);
int chain_engine_graph_stats(chain_engine* e, uint64_t* captures, uint64_t* replays);

/* Per-stage HIP event timing of every call (on by default; off saves one
 * event pair per stage, for latency-sensitive callers such as a server). */
int chain_engine_set_timing(chain_engine* e, int on);

/* Per-stage device times (milliseconds, from HIP events on the engine's
 * stream) of the last call. names is a ';'-separated list. Returns the count. */
int chain_last_stage_times(chain_engine* e, float* ms, int max, char* names, size_t names_len);

/* Algorithmic HBM bytes of the same timed regions (0 where none is stated):
 * every input a region must read and every output it must write, counted
 * once, so bytes / time is the region's achieved bandwidth against HBM peak.
 * Returns the count. */
int chain_last_stage_bytes(chain_engine* e, uint64_t* bytes, int max);

/* ------------------------------------------------------------------------
 * Request server: continuous batching on one GPU. No counterpart in the
 * reference, whose modules run one synchronous frame per call; this is the
 * serving form built on the same engine and bitstreams.
 *
 * W worker threads each own an engine (stream + arenas), so batches overlap
 * on the device. A free worker takes the oldest waiting request and every
 * later waiting request of the same kind and shape, up to max_batch
 * requests and its share (hbm_budget / W) of the device-memory budget, and
 * runs them through one launch per stage. Results are identical to running
 * each request alone. Input buffers are borrowed until the request's result
 * has been polled. */
typedef struct chain_server chain_server;

enum chain_request_kind { CHAIN_REQ_COMPRESS = 0, CHAIN_REQ_DECOMPRESS = 1 };

typedef struct {
    int kind;                  /* chain_request_kind */
    /* compress: host points, (N, D) AoS or (D, N) SoA */
    const void* points;
    int dtype, layout;
    uint32_t dims;
    uint64_t n;
    double eb;
    int64_t p, threshold;
    uint32_t flags;
    int format;                /* chain_format */
    /* decompress: one LCH1 / LCP3 stream in host memory */
    const uint8_t* bytes;
    size_t len;
    int out_dtype, out_layout;
} chain_request;

typedef struct {
    int status;                /* chain_status of the request */
    chain_buf out;             /* container or coordinates, release with chain_buf_free */
    uint64_t n;                /* particles */
    uint32_t dims;
    uint32_t batch;            /* requests in the batch it ran with */
    double t_submit, t_start, t_end;   /* seconds on chain_server_clock */
    char error[256];
} chain_result;

int chain_server_create(int device, uint64_t hbm_budget_bytes, uint32_t workers, uint32_t max_batch,
                        chain_server** out);
void chain_server_destroy(chain_server* s);   /* finishes queued requests first */
const char* chain_server_last_error(const chain_server* s);
int chain_serve_submit(chain_server* s, const chain_request* r, uint64_t* ticket);
/* wait_ms < 0 waits until done, 0 polls; CHAIN_PENDING if not finished.
 * On CHAIN_OK the result is handed over (and forgotten by the server). */
int chain_serve_poll(chain_server* s, uint64_t ticket, int wait_ms, chain_result* out);
double chain_server_clock(chain_server* s);
/* device allocations made for the workers' arenas since creation (two per
 * worker at creation; a larger count means a batch outgrew its arena) */
int chain_server_stats(chain_server* s, uint64_t* arena_allocations);

/* Memory diagnostics (no counterpart in the reference, which sizes nothing in
 * advance; for callers that budget device memory).
 * chain_serve_estimate_compress: the device bytes the server reserves for one
 * compress request of n points (dtype CHAIN_F32 / CHAIN_F64, block size p,
 * chain_flags) with the given rANS chunk length: an upper bound of what the
 * compression carves in its first run. 0 for invalid arguments.
 * chain_engine_memory_stats: arena bytes the last compression on e carved,
 * and how many compressions on e ran a second time with worst-case bounds (a
 * tile with more than 256 exact corrections, or an output over the bound of
 * its plan). */
uint64_t chain_serve_estimate_compress(uint64_t n, uint32_t dims, int dtype, int64_t p, uint32_t flags,
                                       uint32_t chunk_symbols);
int chain_engine_memory_stats(chain_engine* e, uint64_t* compress_need, uint64_t* capacity_retries);

#ifdef __cplusplus
}
#endif
#endif
