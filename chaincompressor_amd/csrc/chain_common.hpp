// Shared definitions of the Chain engine (host and device).
//
// Data layout in HBM (one batch = F frames of the same dims/dtype/layout):
//   points        caller's buffer, frame f at element offset pt_off[f]*D,
//                 AoS (N_f, D) or SoA (D, N_f)
//   key   [T]     concatenated block ids of each point (index order)
//   rel   [T*D]   in-block offsets, point-major (the REL0 matrix)
//   idrel [T]     0 = explicit block, >0 = successor distance (u32)
//   blocks[R*D]   explicit block ids in index order (u64), R = sum_f B_f
//   choice2[R]    u8; diff[R*D] / keep[R*D] u64 compacted per frame
// T = total points of the batch (< 2^32). Per-frame bases of the compacted
// streams live in FrameOut.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "chain_int.hpp"

#define CH_WG 256
#define CH_ITEMS 16
#define CH_TILE (CH_WG * CH_ITEMS)
#define CH_MAXD 4
#define CH_MAXF_LDS 4096

namespace chain {

// FL_Q32: every quantised coordinate of the frame is below 2^32;
// FL_P2: the block size p is a power of two
enum : uint32_t { FL_L1 = 1u, FL_L2 = 2u, FL_Q32 = 4u, FL_P2 = 8u };

struct FrameDev {
    uint64_t pt_off;        // first global point of the frame
    uint64_t n;             // points
    uint32_t tile0;         // first point tile
    uint32_t flags;         // FL_L1 | FL_L2 | FL_Q32 | FL_P2
    double eb, denom;       // denom = 2*eb
    double mins[CH_MAXD];
    uint64_t p;             // effective p (p, or 1 without blocking)
    uint64_t p_magic;       // ceil(2^(63+l)/p), l = ceil(log2 p)
    uint32_t p_l;           // l
    uint32_t max_bid_bits;  // bit length of the largest block id of the frame
    int64_t threshold;
    uint32_t key_shift[CH_MAXD];  // bit position of the dimension inside its key word
    uint32_t key_width[CH_MAXD];
    uint32_t key_word[CH_MAXD];   // key word of the dimension (keys wider than 64 bits)
    uint32_t pad1;
};

struct FrameOut {
    uint64_t rec_base, nrec;       // explicit records
    uint64_t diff_base, ndiff;
    uint64_t keep_base, nkeep;
    uint64_t blk_min, blk_max, blk_maxdiff;
    uint64_t corr_base, ncorr;
    int64_t wrap_min, wrap_max, wrap_range, wrap_min_corr, wrap_max_corr;
    uint64_t fixed_value_bits, fixed_diff_bits;
};

}  // namespace chain
