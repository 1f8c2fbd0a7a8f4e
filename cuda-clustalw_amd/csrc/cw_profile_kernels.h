// Shared between the gfx950 profile-alignment kernels and their host launcher.
#pragma once
#include <stdint.h>

#define CWP_PRF_W 35          // ints per profile row: codes 0..31, GAPCOL 32, LENCOL 33, pad
#define CWP_GAPCOL 32
#define CWP_LENCOL 33
#define CWP_WAVES 4           // waves per workgroup; one workgroup aligns one profile pair
#define CWP_THREADS (64 * CWP_WAVES)
#define CWP_RMAX 8            // rows per lane the sweep is instantiated for
#define CWP_NODE_INTS 8       // node record: A B M N flags type midj child
#define CWP_SCR_ROWS 8        // per-wave scratch rows: 4 ping-pong + 4 final
#define CWP_LOCAL_INTS 8192   // per-wave LDS for a subtree solved locally (records, profile rows, scores)

// Per profile pair; offsets are in int32 units into PrfArgs::pool.
struct PrfProb {
    int32_t L1, L2;           // prfLength1, prfLength2
    int32_t ldS;              // S row stride (L1 + 2)
    int32_t flags;            // bit0: go1, bit1: go2 of the top call
    int32_t node_cap;         // node records available
    int32_t rowcap;           // ints per scratch row (>= L2 + 2)
    int64_t prf1, prf2;       // raw profiles, (L + 2) x CWP_PRF_W
    int64_t row, col;         // int4 records {gap, len, mask, gap + len} per row / column
    int64_t pl1, pl2;         // prefix sums of LENCOL, (L + 2) ints
    int64_t S;                // score matrix S[J * ldS + I] = prfScore(I, J)
    int64_t nodes;            // node records
    int64_t scratch;          // CWP_WAVES x CWP_SCR_ROWS x rowcap
    int64_t lvl;              // final rows of a level's half sweeps: 4 x (L2 + 2 + node_cap)
};

struct PrfArgs {
    const PrfProb* probs;
    int32_t* pool;
    int32_t* out;             // per problem: score, node count, status, levels
    const int32_t* tile_base; // S kernel: first 64x64 tile of each problem
    int32_t nprob;
    int32_t rmax;
};
