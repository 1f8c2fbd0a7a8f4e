// cwhip_internal.h -- the context object and error helpers shared by the
// host files of libcwhip (pairwise: cwhip_api.cpp, profile: cwhip_prf.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/cwhip.h"
#include "cw_pairwise_kernels.h"
#include "cw_profile_kernels.h"

// Buffers grow on demand: the first allocation is the size asked for, a later
// growth at least half again the current size, so a run of slowly growing
// batches (the steps of a progressive alignment) reallocates a few times
// rather than at every batch (a device free synchronises the device).
inline size_t grown(size_t have, size_t count) { return have ? std::max(count, have + have / 2) : count; }

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        const size_t want = grown(n, count);
        if (p) { (void)hipFree(p); p = nullptr; n = 0; }
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
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
    DevBuf<int32_t> d_lens;
    // staged work
    int64_t npairs = 0;
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
    DevBuf<PrfProb> p_probs;
    PinBuf<int32_t> h_prf_in, h_prf_nodes;      // staging of a profile batch's inputs and node tables
    DevBuf<int32_t> f_scratch, f_pairs;         // k-tuple distances: per-lane arrays, pair list
    DevBuf<double> f_dist;
    DevBuf<unsigned int> f_ctr;
    float fast_ms = 0.f;
    DevBuf<uint8_t> p_rows;                     // group rows of a profile batch
    DevBuf<PrfSide> p_sides;                    // its profiles to build
    PinBuf<uint8_t> h_rows;                     // staging of the rows
    std::vector<PrfSide> h_sides;               // host image of p_sides
    int prf_rmax = CWP_RMAX;
    float prf_score_ms = 0.f, prf_mm_ms = 0.f;
    double prf_cells = 0.0, prf_entries = 0.0;
    std::vector<std::vector<uint8_t>> msa_rows;   // rows of the last cwhip_msa
    double msa_stats[6] = {0, 0, 0, 0, 0, 0};
};

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
