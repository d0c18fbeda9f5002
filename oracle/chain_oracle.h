/*
 * chain_oracle.h -- CPU oracle for the Chain compressor (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product. Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it. The GPU engine under
 * chaincompressor_amd/ never links or calls it.
 *
 * It restates, in plain C, the reference CPU path:
 *   cpp/chain/lcps_cpp.cpp:2017-2338           lcps_compress2_cpp (LCP2/LCP3 writer)
 *   cpp/decompress/lcps_cpu_decompress_serial.cpp:1182-1338  LCP2 decoder
 *   cpp/chain/lcps_cpp.cpp:1838-1995           LCP3 (raw) decoding
 * plus the CPU form of this project's own GPU container "LCH1" and its chunked,
 * lane-interleaved rANS sections "RVC1". A RVC1 section with one lane and one
 * chunk carries exactly the RVN1 renormalisation bytes and final state of
 * cpp/chain/lcps_cpp.cpp:739-821.
 *
 * Every entry point returns 0 on success and -1 on failure; oc_last_error()
 * returns the message of the last failure on the calling thread.
 */
#ifndef CHAIN_ORACLE_H
#define CHAIN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Integer streams of one frame, the exact content of an ec=False LCP3 stream
 * (cpp/chain/lcps_cpp.cpp:1184-1211, 2306-2329). Matrices are point-major
 * (record r, dimension d) at [r*dims + d]. */
typedef struct {
    uint32_t dims;
    uint64_t n;
    int64_t p_eff;
    double eb;
    double* mins;          /* [dims] */
    int64_t* idrel;        /* [n]   0 = explicit block, >0 = successor distance */
    uint64_t nblocks;      /* explicit block records B */
    int64_t* blocks;       /* [B*dims] explicit block ids in index order */
    uint64_t nchoice;      /* = B */
    int64_t* choice2;      /* [B] */
    uint64_t ndiff;
    int64_t* diff;         /* [ndiff*dims] ZigZag(wrapped correction) */
    uint64_t nkeep;
    int64_t* keep;         /* [nkeep*dims] */
    int64_t* rel;          /* [n*dims] offsets inside the block, in [0, p_eff) */
    int64_t wrap_min, wrap_max;
    uint64_t ncorr;
    uint64_t* corr_index;  /* [ncorr] i*dims+d, ascending */
    double* corr_value;    /* [ncorr] original coordinate */
    uint64_t max_bid;      /* largest block id over all points and dimensions */
} oc_streams;

/* Parameters of the LCH1 rANS sections (this project's GPU container). */
typedef struct {
    uint32_t lanes_max;      /* 1..64, power of two */
    uint32_t chunk_symbols;  /* symbols per chunk, >= 1 */
} oc_rans_params;

const char* oc_last_error(void);
void oc_free(void* p);

/* Quantisation, L1 successor links, chain decision and L2 split.
 * points: float64 AoS (n, dims). bl/l1/l2 are the ablation switches of
 * lcps_compress2 (cpp/chain/lcps_cpp.cpp:2023-2026). */
int oc_compress_streams(const double* points, uint64_t n, uint32_t dims,
                        double eb, int64_t p, int64_t threshold,
                        int bl, int l1, int l2, oc_streams* out);
void oc_streams_free(oc_streams* s);

/* Containers. *out is malloc'd; release with oc_free. */
int oc_write_lcp3(const oc_streams* s, uint8_t** out, size_t* out_len);
/* LCP2 with the Zstd-22 chooser; stats[5] = lcps_compress2's return list tail:
 * [block_replacement_size, idrel_bytes, rel_bytes, n - B] plus total length in stats[0]. */
int oc_write_lcp2(const oc_streams* s, uint8_t** out, size_t* out_len, uint64_t stats[5]);
int oc_write_lch1(const oc_streams* s, const oc_rans_params* rp,
                  uint8_t** out, size_t* out_len, uint32_t kinds[5]);

/* Decoders: LCP2, LCP3 and LCH1 -> float64 AoS (n, dims), malloc'd. */
int oc_decode(const uint8_t* bytes, size_t len, double** out, uint64_t* n, uint32_t* dims);

/* LCG2 (the reference GPU compressor's container, cu/chain/lcps_gpu_cpp.cu:5-35):
 * the writer stores given streams with Fixed (Zstd) components only, with the
 * given float32 error bound and minima; the decoder restates lcg2::decode of
 * cu/decompress/lcps_gpu_decompress_ans.cu (float32 AoS output, malloc'd) and
 * also reads stored (codec 0) sections. */
int oc_write_lcg2(const oc_streams* s, float eb, const float* mins, uint8_t** out, size_t* out_len);
int oc_decode_lcg2(const uint8_t* bytes, size_t len, float** out, uint64_t* n, uint32_t* dims);
/* one Zstd frame (level 22) of the given bytes, for hand-built containers */
int oc_zstd_frame(const uint8_t* src, size_t n, uint8_t** out, size_t* out_len);

/* Pre-entropy payloads of one component (matrix values[r*dims+d] >= 0). */
int oc_pack_fixed(const int64_t* values, uint64_t n, uint32_t dims,
                  int64_t* mins_out, uint8_t* widths_out,
                  uint8_t** out, size_t* out_len);
int oc_pack_tagged(const int64_t* values, uint64_t n, uint32_t dims,
                   uint8_t** tags_out, uint8_t** out, size_t* out_len, uint32_t* max_width);

/* RVN1 value-rANS (cpp/chain/lcps_cpp.cpp:787-821) and its decoder. */
int oc_rvn1_encode(const uint64_t* values, uint64_t n, uint8_t** out, size_t* out_len);
int oc_rvn1_decode(const uint8_t* bytes, size_t len, uint64_t n, uint64_t** out);
/* Normalised probabilities of cpp/chain/lcps_cpp.cpp:560-684. probs is malloc'd. */
int oc_rvn1_table(const uint64_t* values, uint64_t n, uint32_t** probs,
                  uint32_t* alphabet, int* precision_bits);

/* RVC1 chunked interleaved rANS section over values < 4096 (or tags). */
int oc_rvc1_encode(const uint64_t* values, uint64_t n, uint32_t lanes, uint32_t chunk_symbols,
                   uint8_t** out, size_t* out_len);
int oc_rvc1_decode(const uint8_t* bytes, size_t len, uint64_t n, uint64_t** out);
uint32_t oc_rvc1_lanes_for(uint64_t n, uint32_t lanes_max);

#ifdef __cplusplus
}
#endif
#endif
