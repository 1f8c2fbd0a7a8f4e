// Shared between the gfx950 pairwise kernels and their host launcher.
#pragma once
#include <stdint.h>

#define CW_SEQ_PAD 64          // code-0 padding on each side of a per-wave LDS sequence
#define CW_SMALL_COLS 256      // Myers-Miller nodes with N <= this run lanes-own-columns, rows in LDS
#define CW_STACK_CAP 40        // Myers-Miller stack entries per wave (depth <= log2(L)+2, 6 of them queues)
#define CW_TINYQ 64            // queued tiny Myers-Miller subtrees per wave (one per lane)
#define CW_TAB_STRIDE 35       // ints per matrix row in LDS: an odd stride spreads a wave's
                               // (row, column) lookups over more banks than 32 would
#define CW_TAB_ROW (4 * CW_TAB_STRIDE)        // bytes per matrix row in LDS
#define CW_TAB_BYTES (32 * CW_TAB_ROW)        // int32 [32][CW_TAB_STRIDE] substitution matrix in LDS
#define CW_MAX_THREADS 256     // threads per workgroup (4 waves)
#define CW_RMAX 8              // rows per lane the kernels are instantiated for
#define CW_NCLASS 6            // tuning classes by the longer sequence: <=128 <=256 <=512 <=1024 <=2048 longer
#define CW_NBINS 136           // work-queue bins: 4 per octave of the DP area n*m (n, m < 2^16)
#define CW_QBINS (2 * CW_NBINS)            // queue bins: the short pairs' area bins, then the long pairs'
#define CW_BINS_INTS (2 * CW_QBINS + 2)    // counts, cursors, then the long pairs' count and the total

struct DevArgs {
    const uint8_t* codes;      // packed residue codes of all sequences
    const int64_t* offs;       // sequence s starts at codes[offs[s]]
    const int32_t* lens;       // gapped length of sequence s
    const int32_t* pairs;      // pair p = (pairs[2p], pairs[2p+1]), 0-based
    const int32_t* order;      // work order over pairs (may be null: identity)
    int64_t pair_base;         // plain kernel: first work item of this launch
    int64_t pair_end;          // number of work items
    const int32_t* go_by_len;  // gap open for min(n,m) == L
    const int32_t* matrix;     // [32][32]
    int32_t ge, gap1, gap2;
    int32_t rmax;              // rows-per-lane cap for stripes (1..CW_RMAX)
    int32_t lcap;              // LDS bytes per sequence buffer (with padding)
    int32_t wave_lds;          // LDS bytes per wave
    int32_t rowcap;            // ints per scratch row array (>= maxlen + 1)
    int32_t maxlen;            // longest sequence the LDS buffers, scratch rows and go_by_len cover
    int64_t scratch_stride;    // ints per wave slot (8 * rowcap)
    int32_t* scratch;          // per-wave DP boundary rows
    int32_t* out_count;        // per pair identity count (-1: empty sequence)
    int32_t* out_info;         // optional per pair [maxScore,se1,se2,sb1,sb2,status,nodes,mm cells]
    int32_t* status;           // OR of per-pair status bits (0 = consistent)
    unsigned int* queue;       // persistent kernel work counters (two: the long part, the short part)
    int32_t* bins;             // batched queue: CW_BINS_INTS ints (see CW_BINS_INTS)
    int32_t* order_out;        // batched queue: the binned work order the persistent kernel pops
    int32_t rmax_cls[CW_NCLASS];   // stripe rows per lane by tuning class
    int32_t colcap_cls[CW_NCLASS]; // Myers-Miller nodes with N <= this use the column path (<= CW_SMALL_COLS)
    int32_t nseq;              // sequences behind lens/offs: a pair naming another is refused
    int32_t lsplit;            // pairs whose longer sequence exceeds this queue first (the long part)
    int32_t qpart;             // persistent kernel: 0 the whole queue, 1 its long part, 2 its short part
    int32_t pad_;
};

// The short part's launch of a split batched run: its arguments (smaller
// sequence buffers) and shape.
struct CwSplitShape {
    DevArgs args;
    int blocks, threads;
    size_t lds;
};

// Status bits of DevArgs::status: 1 reverse pass found no start cell, 2
// Myers-Miller stack or tiny-node error, 4 a pair that names a sequence outside
// 0..nseq-1 or longer than maxlen (its count is set to -2 and it is not run).
#define CW_ST_BADPAIR 4

// Tuning class of a pair: by the longer sequence.
static inline __host__ __device__ int cw_class_of(int n, int m)
{
    const int l = n > m ? n : m;
    return l <= 128 ? 0 : l <= 256 ? 1 : l <= 512 ? 2 : l <= 1024 ? 3 : l <= 2048 ? 4 : 5;
}

// Queue bin of a pair: 4 bins per octave of n*m, larger area = larger bin.
static inline __host__ __device__ int cw_bin_of(int n, int m)
{
    const uint64_t a = (uint64_t)n * (uint64_t)m;
    if (a < 4) return (int)a;
    const int lg = 63 - __builtin_clzll(a);
    const int b = 4 * lg + (int)((a >> (lg - 2)) & 3);
    return b < CW_NBINS ? b : CW_NBINS - 1;
}

static inline int cw_wave_lds_bytes(int lcap)
{
    int b = 2 * lcap + 4 * (4 * (CW_SMALL_COLS + 1) + 6 * CW_STACK_CAP + 2 * CW_TINYQ);
    return (b + 15) & ~15;
}

static inline int cw_lcap(int maxlen)
{
    return ((maxlen + 2 * CW_SEQ_PAD + 1) + 15) & ~15;
}

// Waves per workgroup: `want`, or fewer where the waves' sequence buffers and
// the matrix would exceed a CU's 160 KiB of LDS (0: not even one wave fits).
static inline int cw_fit_waves(int want, int wave_lds)
{
    const int fit = (160 * 1024 - CW_TAB_BYTES) / wave_lds;
    return fit < want ? fit : want;
}
