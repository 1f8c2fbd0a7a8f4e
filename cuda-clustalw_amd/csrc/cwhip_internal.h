// cwhip_internal.h -- the context object and error helpers shared by the
// host files of libcwhip (pairwise: cwhip_api.cpp, profile: cwhip_prf.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <vector>

#include "../../include/cwhip.h"
#include "cw_nj_kernels.h"
#include "cw_pairwise_kernels.h"
#include "cw_profile_kernels.h"

// Buffers grow on demand: the first allocation is the size asked for, a later
// growth at least half again the current size, so a run of slowly growing
// batches (the steps of a progressive alignment) reallocates a few times
// rather than at every batch (a device free synchronises the device).
inline size_t grown(size_t have, size_t count) { return have ? std::max(count, have + have / 2) : count; }

// Device bytes a context holds, and an optional cap on them (a server sets
// the cap to the reservation a job was admitted with).  `total`/`peak` point
// at counters shared by several contexts (a server's slots), or are null.
struct MemAccount {
    int64_t bytes = 0;
    int64_t limit = 0;                         // 0: no cap
    std::atomic<int64_t>* total = nullptr;
    std::atomic<int64_t>* peak = nullptr;
    void add(int64_t d) {
        bytes += d;
        if (total) {
            const int64_t now = total->fetch_add(d) + d;
            if (peak) {
                int64_t p = peak->load();
                while (now > p && !peak->compare_exchange_weak(p, now)) {}
            }
        }
    }
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    MemAccount* acct = nullptr;
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        size_t want = grown(n, count);
        if (acct && acct->limit) {
            // within the cap: grow by exactly what is asked when the usual
            // half-again step would not fit, and fail when that does not fit
            const int64_t rest = acct->bytes - (int64_t)(n * sizeof(T));
            if (rest + (int64_t)(std::max<size_t>(want, 1) * sizeof(T)) > acct->limit) want = count;
            if (rest + (int64_t)(std::max<size_t>(want, 1) * sizeof(T)) > acct->limit) return hipErrorOutOfMemory;
        }
        release();
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) {
            n = want;
            if (acct) acct->add((int64_t)(std::max<size_t>(want, 1) * sizeof(T)));
        } else {
            p = nullptr;
        }
        return e;
    }
    void release() {
        if (p) {
            (void)hipFree(p);
            if (acct) acct->add(-(int64_t)(std::max<size_t>(n, 1) * sizeof(T)));
        }
        p = nullptr;
        n = 0;
    }
};

// Page-locked host staging (hipHostMalloc), so copies to and from the device
// are DMA transfers that do not block the calling thread.
template <typename T>
struct PinBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        const size_t want = grown(n, count);
        if (p) { (void)hipHostFree(p); p = nullptr; n = 0; }
        hipError_t e = hipHostMalloc((void**)&p, std::max<size_t>(want, 1) * sizeof(T), hipHostMallocDefault);
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() { if (p) (void)hipHostFree(p); p = nullptr; n = 0; }
};

struct cwhip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    std::string err;
    int num_cu = 0;
    // sequences
    int nseq = 0, maxlen = 0;
    std::vector<int32_t> h_lens, h_gapfree;
    std::vector<int64_t> h_offs;
    std::vector<uint8_t> h_codes;
    DevBuf<uint8_t> d_codes;
    DevBuf<int64_t> d_offs;
    DevBuf<int32_t> d_lens, d_gapfree;
    // staged work; counts_ready: d_count holds the finished run of the staged pairs
    int64_t npairs = 0;
    bool counts_ready = false;
    DevBuf<int32_t> d_pairs, d_order, d_count, d_info, d_matrix, d_go, d_status, d_scratch;
    DevBuf<unsigned int> d_queue;
    std::vector<int32_t> h_pairs;
    int gap1 = 30, gap2 = 31, ge = 1;
    // tuning
    int rmax = CW_RMAX, wpg = CW_MAX_THREADS / 64, groups_per_cu = 0;
    cwhip_tune_table tune;                        // batched kernel shapes by class
    // batched stages captured as a graph; gkey holds the launch it was captured
    // for, glast the last launch that ran directly
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    std::vector<uint8_t> gkey, glast;
    DevBuf<int32_t> d_bins;
    // profile alignment
    DevBuf<int32_t> p_pool, p_out, p_tiles;
    DevBuf<int32_t> p_wide;                     // cw_prf_wide: tasks, chunk counters, error flag
    DevBuf<PrfProb> p_probs;
    PinBuf<int32_t> h_prf_in, h_prf_nodes;      // staging of a profile batch's inputs and node tables
    DevBuf<int32_t> f_scratch, f_pairs;         // k-tuple distances: per-lane arrays, pair list
    DevBuf<double> f_dist;
    DevBuf<unsigned int> f_ctr;
    float fast_ms = 0.f;
    DevBuf<double> t_dist;                      // neighbour-joining: the (n+1)^2 DistMatrix,
    DevBuf<int32_t> t_desc, t_status;           // treeDesc and the kernel's status,
    DevBuf<double> t_branch;                    // leftBranch then rightBranch
    DevBuf<uint8_t> p_rows;                     // group rows of a profile batch
    DevBuf<PrfSide> p_sides;                    // its profiles to build
    PinBuf<uint8_t> h_rows;                     // staging of the rows
    std::vector<PrfSide> h_sides;               // host image of p_sides
    int prf_rmax = CWP_RMAX;
    int prf_wide_rows = 256;                    // rows per cw_prf_wide stripe (0: no cw_prf_wide)
    int prf_wide_levels = 5;                    // levels before cw_prf_mm's last launch (<= 6; 1: none)
    int prf_wide_pairs = CWP_WIDE_MAXP;         // batches of at most this many pairs use them
    int prf_wide_wait = 1 << 20;                // sleeps a cw_prf_wide stripe waits before it gives up
    float prf_score_ms = 0.f, prf_mm_ms = 0.f;
    double prf_cells = 0.0, prf_entries = 0.0;
    std::vector<std::vector<uint8_t>> msa_rows;   // rows of the last cwhip_msa
    double msa_stats[6] = {0, 0, 0, 0, 0, 0};
    MemAccount mem;                               // device bytes of the buffers above

    // every device buffer of the context, for the accounting and release_buffers
    template <typename F>
    void for_each_buf(F f)
    {
        f(d_codes); f(d_offs); f(d_lens); f(d_gapfree); f(d_pairs); f(d_order); f(d_count); f(d_info);
        f(d_matrix); f(d_go); f(d_status); f(d_scratch); f(d_queue); f(d_bins);
        f(p_pool); f(p_out); f(p_tiles); f(p_wide); f(p_probs); f(f_scratch); f(f_pairs); f(f_dist);
        f(f_ctr); f(t_dist); f(t_desc); f(t_status); f(t_branch); f(p_rows); f(p_sides);
    }
    void bind_account() { for_each_buf([this](auto& b) { b.acct = &mem; }); }
    // frees every device buffer (a later call allocates again); the loaded
    // sequences and a staged pair list are gone with them
    void release_buffers()
    {
        for_each_buf([](auto& b) { b.release(); });
        nseq = 0;
        npairs = 0;
        counts_ready = false;
        h_pairs.clear();
        if (gexec) {
            (void)hipGraphExecDestroy(gexec);
            gexec = nullptr;
        }
        gkey.clear();
        glast.clear();
    }
};

// The PHYLIP text of a neighbour-joining result (ClusterTreeOutput::
// printPhylipTree and twoWaySplit) from its treeDesc ((n+1)^2, consumed) and
// branch lengths (n+2 each); names[0 .. n-1].  Defined in cwhip_tree.cpp.
std::string cw_phylip_text(int n, std::vector<int32_t>& desc, const std::vector<double>& left,
                           const std::vector<double>& right, const char* const* names);

static inline int fail(cwhip_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

static inline int hipfail(cwhip_ctx* c, hipError_t e, const char* what)
{
    return fail(c, e == hipErrorOutOfMemory ? CWHIP_E_NOMEM : CWHIP_E_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(call)                                              \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return hipfail(ctx, e_, #call);    \
    } while (0)
