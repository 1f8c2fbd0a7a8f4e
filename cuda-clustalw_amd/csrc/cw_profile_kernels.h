// Shared between the gfx950 profile-alignment kernels and their host launcher.
#pragma once
#include <stdint.h>

#define CWP_PRF_W 35          // ints per profile row: codes 0..31, GAPCOL 32, LENCOL 33, pad
#define CWP_GAPCOL 32
#define CWP_LENCOL 33
#define CWP_WAVES 4           // waves per cw_prf_mm workgroup; one workgroup aligns one profile pair
#define CWP_THREADS (64 * CWP_WAVES)
#define CWP_LOCAL_WAVES 4     // waves per cw_prf_local workgroup (CWP_LOCAL_INTS of LDS each)
#define CWP_RMAX 8            // rows per lane the sweep is instantiated for
#define CWP_NODE_INTS 8       // node record: A B M N flags type midj child
#define CWP_SCR_ROWS 8        // per-wave scratch rows: 4 ping-pong + 4 final
#define CWP_LOCAL_INTS 8192   // per-wave LDS for a subtree solved locally (records, profile rows, scores)
#define CWP_ROOT_WIDE 8       // PrfProb::flags: cw_prf_wide has swept the root's halves
#define CWP_WIDE_MAXG 8       // cw_prf_wide workgroups (one stripe each) per half
#define CWP_WIDE_MAXP 512     // profile pairs a batch may have for cw_prf_wide / cw_prf_level to run
#define CWP_WIDE_SETS 64      // most cw_prf_level tasks (scratch sets of CWP_SCR_ROWS rows) per pair

// Per profile pair; offsets are in int32 units into PrfArgs::pool.
struct PrfProb {
    int32_t L1, L2;           // prfLength1, prfLength2
    int32_t ldS;              // S row stride (L1 + 2)
    int32_t flags;            // bit0: go1, bit1: go2 of the top call; CWP_ROOT_WIDE
    int32_t node_cap;         // node records available
    int32_t rowcap;           // ints per scratch row (>= L2 + 2)
    int64_t prf1, prf2;       // raw profiles, (L + 2) x CWP_PRF_W
    int64_t row, col;         // int4 records {gap, len, mask, gap + len} per row / column
    int64_t pl1, pl2;         // prefix sums of LENCOL, (L + 2) ints
    int64_t S;                // score matrix S[J * ldS + I] = prfScore(I, J)
    int64_t nodes;            // node records
    int64_t scratch;          // CWP_WAVES (or 2 CWP_WIDE_MAXG) x CWP_SCR_ROWS x rowcap
    int64_t lvl;              // final rows of a level's half sweeps: 4 x (L2 + 2 + node_cap)
};

struct PrfArgs {
    const PrfProb* probs;
    int32_t* pool;
    int32_t* out;             // per problem: score, node count, status, levels
    const int32_t* tile_base; // S kernel: first 64x64 tile of each problem
    int32_t nprob;
    int32_t rmax;
    int32_t* local_list;      // {problem, node, first reserved slot} of every subtree cw_prf_mm
                              // leaves to cw_prf_local
    int32_t* local_ctr;       // [0] entries in local_list, [1] next entry to take (both 0 at launch)
    int32_t* state;           // per problem: lev_b, lev_e, nodes allocated, levels done
    int32_t resume;           // cw_prf_mm: continue at the level in state (else start at the root)
    int32_t ext;              // cw_prf_mm: this level's half sweeps are done (cw_prf_level); finish
                              // it and stop (else sweep and finish every remaining level)
};

// cw_prf_wide: one task (a one-wave workgroup) per stripe of a root's half.
struct WideArgs {
    const int32_t* tasks;     // 4 per task: problem, half, member m, members G of the half
    int32_t* prog;            // per task: chunks its stripe has flushed (0 at launch)
    int32_t* err;             // set to 1 when a stripe's wait timed out (0 at launch)
    int32_t wait_limit;       // sleeps a stripe waits for the one above before it gives up
};

// ---- profile construction on the device ------------------------------------
// One profile of a step (MyersMillerProfileAlign.cpp:202-240): the gap
// coefficients of ProfileBase::calcGapCoeff and the residue columns of
// ProfileWithSub::calcProfileWithSub (profile 1) or
// ProfileStandard::calcStandardProfile (profile 2), written where the DP
// kernels read them.  Offsets are in int32 units into PrfArgs::pool, except
// `rows` (bytes into BuildArgs::rows).
#define CWP_BUILD_COLS 64     // columns per workgroup of cw_prf_cols
#define CWP_BUILD_SLICES 4    // row slices per column (CWP_BUILD_COLS x CWP_BUILD_SLICES threads)
#define CWP_F_SUB 1           // profile with substitution matrix (else standard profile)
#define CWP_F_DNA 2           // DNA: no residue-specific, hydrophilic or variable penalties
#define CWP_F_VAR 4           // variable penalties (vW) of a two-row group with pcid > 60
#define CWP_F_HYD 8           // hydrophilic gap penalties
#define CWP_F_PREF 16         // residue-specific gap penalties
#define CWP_F_EGP 32          // UserParameters::endGapPenalties

struct PrfSide {
    int32_t n;                // rows of the group (numSeq)
    int32_t L;                // prfLength
    int32_t stride;           // bytes per row in the rows buffer (>= L)
    int32_t flags;            // CWP_F_*
    int32_t gdist;            // gap separation distance in effect (0: none)
    int32_t gapcoef, lencoef; // of the step
    int32_t sum2;             // sum of the rows' weights
    int32_t col0;             // first cw_prf_cols workgroup of this side
    int32_t pad;
    int64_t rows;             // row 0 of the group
    int64_t rinfo;            // n x {weight, first, end}: [first, end) are the columns whose gaps count
    int64_t mat;              // 32 x 32 substitution matrix (profile with substitution only)
    int64_t prf;              // (L + 2) x CWP_PRF_W profile
    int64_t rec;              // (L + 2) x {gap, len, mask, gap + len} for the DP
    int64_t pl;               // (L + 2) LENCOL prefix sums for the DP
    int64_t tmp;              // 4 x (L + 2): gap counts, hydrophilic, residue and variable weights
};

struct BuildArgs {
    const PrfSide* sides;
    const uint8_t* rows;
    int32_t* pool;
    int32_t nside;
    int32_t ncolblk;          // cw_prf_cols workgroups in all
    uint32_t hyd_mask;        // bit c: code c is a hydrophilic residue
    int32_t pen[32];          // residue-specific gap penalty of each code (180 - pprob)
};
