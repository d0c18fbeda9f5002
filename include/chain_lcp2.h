/*
 * chain_lcp2.h -- host-side LCP2 export of Chain streams (libchain_lcp2.so).
 *
 * Converts the raw container LCP3, which the GPU engine writes byte for byte
 * like lcps_compress2(..., ec=False) (cpp/chain/lcps_cpp.cpp:1184-1211,
 * 2306-2329), into the entropy-coded LCP2 of lcps_compress2(..., ec=True)
 * (cpp/chain/lcps_cpp.cpp:2017-2338): the Zstd-22 / RVN1 chooser of
 * :1140-1174 over every component, Zstd-coded corrections. This is the
 * reference's CPU-side container, so lcps_cpu_decompress_serial
 * (cpp/decompress/lcps_cpu_decompress_serial.cpp:1377-1381) reads it; with the
 * reference's libzstd (1.4.9) the bytes equal lcps_compress2's.
 *
 * Host code only; it links libzstd and no GPU runtime.
 */
#ifndef CHAIN_LCP2_H
#define CHAIN_LCP2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum chain_lcp2_status {
    CHAIN_LCP2_OK = 0,
    CHAIN_LCP2_EINVAL = 1,    /* invalid argument */
    CHAIN_LCP2_EFORMAT = 2,   /* not a valid LCP3 stream */
    CHAIN_LCP2_EZSTD = 3,     /* libzstd error */
    CHAIN_LCP2_ENOMEM = 4,
    CHAIN_LCP2_EUNSUPPORTED = 5,   /* an LCG2 section coded with nvCOMP ANS */
};

/* lcp3/len: one LCP3 stream. zstd_level: 22 for the reference's bytes.
 * *out is malloc'd (release with chain_lcp2_free). stats (may be NULL) =
 * lcps_compress2's return list after the bytes: block_replacement_size,
 * idrel_size, rel_size, num_chained (lcps_cpp.cpp:2331-2337). err (may be
 * NULL) receives the message of a failure. */
int chain_lcp2_from_lcp3(const uint8_t* lcp3, size_t len, int zstd_level, uint8_t** out, size_t* out_len,
                         uint64_t stats[4], char* err, size_t err_len);
/* The reverse: an LCP2 stream (any of its component encodings RAW1, RVMX,
 * ZFMX, ZTG4; cpp/decompress/lcps_cpu_decompress_serial.cpp:981-1174) to the
 * LCP3 of the same integer streams, which the GPU engine decodes. */
int chain_lcp3_from_lcp2(const uint8_t* lcp2, size_t len, uint8_t** out, size_t* out_len, char* err,
                         size_t err_len);
/* An LCG2 stream of the reference GPU compressor (cu/chain/lcps_gpu_cpp.cu:5-35;
 * replaces the host parse of lcg2::decode, cu/decompress/
 * lcps_gpu_decompress_ans.cu:2345-2452, with every one of its checks) with
 * each Zstd section inflated to a stored section (codec 0: u8 0, u64 raw
 * bytes, the bytes), the form the GPU engine decodes (chain_decompress*).
 * Fixed and empty components are supported; a Tagged or Raw component,
 * whose section is coded with nvCOMP ANS, returns CHAIN_LCP2_EUNSUPPORTED.
 * A stream that is already inflated is returned unchanged. */
int chain_lcg2_inflate(const uint8_t* lcg2, size_t len, uint8_t** out, size_t* out_len, char* err,
                       size_t err_len);
void chain_lcp2_free(uint8_t* p);
/* version of the libzstd the library runs with */
const char* chain_lcp2_zstd_version(void);

#ifdef __cplusplus
}
#endif
#endif
