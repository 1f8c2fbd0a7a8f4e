// cwhip_api.cpp -- host side of the C ABI declared in include/cwhip.h.
//
// Owns device memory and launches; all DP arithmetic runs in the gfx950
// kernels of cw_pairwise_kernels.hip.  There is no CPU fallback: every entry
// point that computes fails with CWHIP_E_HIP when the device is unusable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "cw_matrices.inc"
#include "cwhip_internal.h"

hipError_t cw_launch_plain(int blocks, int threads, size_t lds, hipStream_t st, const DevArgs& a);
hipError_t cw_launch_persistent(int blocks, int threads, size_t lds, hipStream_t st, const DevArgs& a);
hipError_t cw_launch_reset(hipStream_t st, unsigned int* q, int32_t* status);
hipError_t cw_occupancy(int threads, size_t lds, int* plain_blocks, int* pers_blocks);

namespace {

constexpr int kPlainChunk = 16384;   // pairs per plain-kernel launch (scratch slots)

} // namespace

extern "C" {

int cwhip_create(int device, cwhip_ctx** out)
{
    if (!out) return CWHIP_E_ARG;
    *out = nullptr;
    cwhip_ctx* ctx = new cwhip_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) { delete ctx; return CWHIP_E_HIP; }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) { delete ctx; return CWHIP_E_HIP; }
    ctx->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipEventCreate(&ctx->ev2) != hipSuccess) {
        delete ctx;
        return CWHIP_E_HIP;
    }
    *out = ctx;
    return CWHIP_OK;
}

void cwhip_destroy(cwhip_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    ctx->d_codes.release(); ctx->d_offs.release(); ctx->d_lens.release();
    ctx->d_pairs.release(); ctx->d_order.release(); ctx->d_count.release();
    ctx->d_info.release(); ctx->d_matrix.release(); ctx->d_go.release();
    ctx->d_status.release(); ctx->d_scratch.release(); ctx->d_queue.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* cwhip_last_error(const cwhip_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// SubMatrix::getMatrix (SubMatrix.cpp:689-829) for negFlag=true: scaled table,
// matrixAvgScore from the mean off-diagonal entry, zero gap rows/columns.
static int build_matrix(const short* tab, int nletters, int scale, int32_t* m, int* mat_avg)
{
    for (int i = 0; i < 32 * 32; ++i) m[i] = 0;
    for (int i = 0; i < 25; ++i)
        for (int j = 0; j < 25; ++j) m[i * 32 + j] = tab[i * 25 + j] * scale;
    const int max_res = nletters - 1;
    int off = 0;
    for (int i = 0; i <= 24; ++i)
        for (int j = 0; j < i; ++j) off += m[i * 32 + j];
    off = (int)(off / (((float)(max_res * max_res - max_res)) / 2));
    *mat_avg = -off;
    for (int i = 0; i < 30; ++i) {
        m[i * 32 + 30] = m[30 * 32 + i] = 0;
        m[i * 32 + 31] = m[31 * 32 + i] = 0;
    }
    m[30 * 32 + 30] = m[31 * 32 + 31] = m[31 * 32 + 30] = m[30 * 32 + 31] = 0;
    return max_res + 2;
}

int cwhip_pair_params_default(int dna, float pw_gap_open, float pw_gap_extend,
                              float transition_weight, int max_len, cwhip_pair_params* p,
                              int32_t* gap_open_buf)
{
    if (!p || !gap_open_buf || max_len < 0) return CWHIP_E_ARG;
    int mat_avg = 0;
    int int_scale = 100;              // SubMatrix.cpp:274-278
    const float go_scale = 1.0f, ge_scale = 1.0f;
    if (dna) {                        // "iub" (SubMatrix.cpp:290-318)
        build_matrix(cw_iub_dna, CW_IUB_DNA_NLETTERS, int_scale, p->matrix, &mat_avg);
        const int w = (int)(transition_weight * p->matrix[0]);
        p->matrix[0 * 32 + 4] = p->matrix[4 * 32 + 0] = w;
        p->matrix[2 * 32 + 11] = p->matrix[11 * 32 + 2] = w;
        p->matrix[2 * 32 + 12] = p->matrix[12 * 32 + 2] = w;
    } else {                          // "gonnet" (SubMatrix.cpp:348-354)
        int_scale /= 10;
        build_matrix(cw_gonnet250, CW_GONNET250_NLETTERS, int_scale, p->matrix, &mat_avg);
    }
    // FullPairwiseAlign.cpp:129-148
    for (int L = 0; L <= max_len; ++L) {
        int go;
        if (dna) {
            go = (int)(2 * pw_gap_open * int_scale * go_scale);
        } else if (L == 0) {
            go = 0;                   // never used: empty sequences do not reach the DP
        } else if (mat_avg <= 0) {
            go = 2 * (int)((pw_gap_open + log((double)L)) * int_scale);
        } else {
            go = (int)(2 * mat_avg * (pw_gap_open + log((double)L)) * go_scale);
        }
        gap_open_buf[L] = go;
    }
    p->gap_extend = dna ? (int)(pw_gap_extend * int_scale * ge_scale) : (int)(pw_gap_extend * int_scale);
    p->gap_pos1 = 30;
    p->gap_pos2 = 31;
    p->max_len = max_len;
    p->gap_open_by_len = gap_open_buf;
    return CWHIP_OK;
}

float cwhip_distance(int32_t count, int32_t len1, int32_t len2)
{
    // FullPairwiseAlign.cpp:172-181 (tracePath's score is 100.0*(float)count)
    float mm = (float)(100.0 * (float)count);
    if (len1 == 0 || len2 == 0) mm = 0;
    else mm /= (float)std::min(len1, len2);
    return ((float)100.0 - mm) / (float)100.0;
}

int cwhip_load_sequences(cwhip_ctx* ctx, const uint8_t* codes, const int64_t* offsets,
                         const int32_t* lens, int nseq)
{
    if (!ctx || nseq < 0 || (nseq > 0 && (!codes || !offsets || !lens))) return fail(ctx, CWHIP_E_ARG, "bad sequence arguments");
    HIPCHK(hipSetDevice(ctx->device));
    int64_t total = 0;
    int maxlen = 0;
    for (int s = 0; s < nseq; ++s) {
        if (lens[s] < 0 || offsets[s] < 0) return fail(ctx, CWHIP_E_ARG, "negative length/offset");
        if (lens[s] > CWHIP_MAX_LEN) return fail(ctx, CWHIP_E_TOOLONG, "sequence longer than CWHIP_MAX_LEN");
        maxlen = std::max(maxlen, (int)lens[s]);
        total = std::max<int64_t>(total, offsets[s] + lens[s]);
    }
    ctx->nseq = nseq;
    ctx->maxlen = maxlen;
    ctx->h_lens.assign(lens, lens + nseq);
    ctx->h_offs.assign(offsets, offsets + nseq);
    ctx->h_codes.assign(codes, codes + total);
    ctx->h_gapfree.resize(nseq);
    for (int s = 0; s < nseq; ++s) {
        int l = 0;
        for (int k = 0; k < lens[s]; ++k) {
            const int r = codes[offsets[s] + k];
            if (r != 30 && r != 31) ++l;
        }
        ctx->h_gapfree[s] = l;
    }
    HIPCHK(ctx->d_codes.ensure(total + 64));
    HIPCHK(ctx->d_offs.ensure(nseq + 1));
    HIPCHK(ctx->d_lens.ensure(nseq + 1));
    if (total) HIPCHK(hipMemcpyAsync(ctx->d_codes.p, codes, total, hipMemcpyHostToDevice, ctx->stream));
    if (nseq) {
        HIPCHK(hipMemcpyAsync(ctx->d_offs.p, offsets, nseq * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->d_lens.p, lens, nseq * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CWHIP_OK;
}

int64_t cwhip_range_npairs(int nseq, int i_start, int i_end, int j_start, int j_end)
{
    int64_t np = 0;
    for (int si = std::max(0, i_start); si < nseq && si < i_end; ++si) {
        const int lo = std::max(si + 1, j_start + 1);
        const int hi = std::min(nseq, j_end);
        if (hi > lo) np += hi - lo;
    }
    return np;
}

int cwhip_set_tuning(cwhip_ctx* ctx, int rmax, int waves_per_group, int groups_per_cu)
{
    if (!ctx) return CWHIP_E_ARG;
    if (rmax < 0 || rmax > CW_RMAX || waves_per_group < 0 || waves_per_group > CW_MAX_THREADS / 64 ||
        groups_per_cu < 0)
        return fail(ctx, CWHIP_E_ARG, "tuning value out of range");
    if (rmax) ctx->rmax = rmax;
    if (waves_per_group) ctx->wpg = waves_per_group;
    ctx->groups_per_cu = groups_per_cu;
    return CWHIP_OK;
}

int cwhip_stage_pairs(cwhip_ctx* ctx, const cwhip_pair_params* p, const int32_t* pairs, int64_t npairs)
{
    if (!ctx || !p || (npairs > 0 && !pairs) || npairs < 0) return fail(ctx, CWHIP_E_ARG, "bad pair arguments");
    if (ctx->nseq <= 0 && npairs > 0) return fail(ctx, CWHIP_E_NOSEQ, "no sequences loaded");
    if (p->max_len < 0 || !p->gap_open_by_len) return fail(ctx, CWHIP_E_ARG, "bad gap-open table");
    for (int64_t k = 0; k < 2 * npairs; ++k)
        if (pairs[k] < 0 || pairs[k] >= ctx->nseq) return fail(ctx, CWHIP_E_ARG, "pair index out of range");
    for (int64_t k = 0; k < npairs; ++k) {
        const int mn = std::min(ctx->h_lens[pairs[2 * k]], ctx->h_lens[pairs[2 * k + 1]]);
        if (mn > p->max_len) return fail(ctx, CWHIP_E_ARG, "gap-open table shorter than a pair's length");
    }
    HIPCHK(hipSetDevice(ctx->device));
    ctx->npairs = npairs;
    ctx->h_pairs.assign(pairs, pairs + 2 * npairs);
    ctx->gap1 = p->gap_pos1;
    ctx->gap2 = p->gap_pos2;
    ctx->ge = p->gap_extend;
    // work order of the persistent kernel: longest-processing-time first, by
    // the pair's DP area, so the queue drains evenly
    std::vector<int32_t> order(npairs);
    std::iota(order.begin(), order.end(), 0);
    std::vector<int64_t> cost(npairs);
    for (int64_t k = 0; k < npairs; ++k)
        cost[k] = (int64_t)ctx->h_lens[pairs[2 * k]] * ctx->h_lens[pairs[2 * k + 1]];
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
    HIPCHK(ctx->d_pairs.ensure(2 * npairs + 2));
    HIPCHK(ctx->d_order.ensure(npairs + 1));
    HIPCHK(ctx->d_count.ensure(npairs + 1));
    HIPCHK(ctx->d_info.ensure(8 * npairs + 8));
    HIPCHK(ctx->d_matrix.ensure(32 * 32));
    HIPCHK(ctx->d_go.ensure(p->max_len + 1));
    HIPCHK(ctx->d_status.ensure(4));
    HIPCHK(ctx->d_queue.ensure(4));
    if (npairs) {
        HIPCHK(hipMemcpyAsync(ctx->d_pairs.p, pairs, 2 * npairs * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->d_order.p, order.data(), npairs * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    }
    HIPCHK(hipMemcpyAsync(ctx->d_matrix.p, p->matrix, 32 * 32 * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_go.p, p->gap_open_by_len, (p->max_len + 1) * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CWHIP_OK;
}

static void fill_args(cwhip_ctx* ctx, DevArgs& a)
{
    std::memset(&a, 0, sizeof a);
    a.codes = ctx->d_codes.p;
    a.offs = ctx->d_offs.p;
    a.lens = ctx->d_lens.p;
    a.pairs = ctx->d_pairs.p;
    a.pair_end = ctx->npairs;
    a.go_by_len = ctx->d_go.p;
    a.matrix = ctx->d_matrix.p;
    a.ge = ctx->ge;
    a.gap1 = ctx->gap1;
    a.gap2 = ctx->gap2;
    a.rmax = ctx->rmax;
    a.lcap = cw_lcap(ctx->maxlen);
    a.wave_lds = cw_wave_lds_bytes(a.lcap);
    a.rowcap = ((ctx->maxlen + 2) + 63) & ~63;
    a.scratch_stride = 8 * (int64_t)a.rowcap;
    a.out_count = ctx->d_count.p;
    a.out_info = ctx->d_info.p;
    a.status = ctx->d_status.p;
    a.queue = ctx->d_queue.p;
}

int cwhip_run_staged(cwhip_ctx* ctx, int kernel, int reps, float* ms, float* kernel_ms)
{
    if (!ctx || reps < 1) return fail(ctx, CWHIP_E_ARG, "bad run arguments");
    if (kernel != CWHIP_KERNEL_PLAIN && kernel != CWHIP_KERNEL_BATCHED) return fail(ctx, CWHIP_E_ARG, "unknown kernel");
    HIPCHK(hipSetDevice(ctx->device));
    if (ctx->npairs == 0) { if (ms) *ms = 0; if (kernel_ms) *kernel_ms = 0; return CWHIP_OK; }
    DevArgs a;
    fill_args(ctx, a);
    const int threads = 64 * ctx->wpg;
    const size_t lds = CW_TAB_BYTES + (size_t)ctx->wpg * a.wave_lds;
    if (lds > 160 * 1024) return fail(ctx, CWHIP_E_TOOLONG, "per-workgroup LDS exceeds 160 KiB");
    int occ_plain = 0, occ_pers = 0;
    HIPCHK(cw_occupancy(threads, lds, &occ_plain, &occ_pers));
    if (occ_pers < 1 || occ_plain < 1) return fail(ctx, CWHIP_E_TOOLONG, "kernel does not fit on a CU");
    const int gpc = ctx->groups_per_cu ? std::min(ctx->groups_per_cu, occ_pers) : occ_pers;
    int64_t slots;
    int pers_blocks = 0;
    if (kernel == CWHIP_KERNEL_BATCHED) {
        // never more waves than pairs: a small job leaves the rest of the chip
        // (and its scratch) to whatever else runs beside it
        pers_blocks = (int)std::min<int64_t>((int64_t)ctx->num_cu * gpc,
                                             (ctx->npairs + ctx->wpg - 1) / ctx->wpg);
        slots = (int64_t)pers_blocks * ctx->wpg;
    } else {
        slots = std::min<int64_t>(kPlainChunk, ctx->npairs);
        slots = (slots + ctx->wpg - 1) / ctx->wpg * ctx->wpg;
    }
    HIPCHK(ctx->d_scratch.ensure((size_t)slots * a.scratch_stride));
    a.scratch = ctx->d_scratch.p;
    for (int rep = 0; rep < reps; ++rep) {
        HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(cw_launch_reset(ctx->stream, ctx->d_queue.p, ctx->d_status.p));
        HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        if (kernel == CWHIP_KERNEL_BATCHED) {
            a.order = ctx->d_order.p;
            a.pair_base = 0;
            HIPCHK(cw_launch_persistent(pers_blocks, threads, lds, ctx->stream, a));
        } else {
            a.order = nullptr;
            for (int64_t b0 = 0; b0 < ctx->npairs; b0 += slots) {
                a.pair_base = b0;
                a.pair_end = std::min<int64_t>(ctx->npairs, b0 + slots);
                const int64_t nb = (a.pair_end - b0 + ctx->wpg - 1) / ctx->wpg;
                HIPCHK(cw_launch_plain((int)nb, threads, lds, ctx->stream, a));
            }
        }
        HIPCHK(hipEventRecord(ctx->ev2, ctx->stream));
    }
    HIPCHK(hipEventSynchronize(ctx->ev2));
    float t_all = 0, t_k = 0;
    HIPCHK(hipEventElapsedTime(&t_all, ctx->ev0, ctx->ev2));
    HIPCHK(hipEventElapsedTime(&t_k, ctx->ev1, ctx->ev2));
    if (ms) *ms = t_all;
    if (kernel_ms) *kernel_ms = t_k;
    int32_t st = 0;
    HIPCHK(hipMemcpy(&st, ctx->d_status.p, sizeof st, hipMemcpyDeviceToHost));
    if (st) return fail(ctx, CWHIP_E_KERNEL, "kernel status " + std::to_string(st));
    return CWHIP_OK;
}

int cwhip_fetch_counts(cwhip_ctx* ctx, int32_t* out_count, int64_t npairs)
{
    if (!ctx || (!out_count && npairs) || npairs != ctx->npairs) return fail(ctx, CWHIP_E_ARG, "bad fetch arguments");
    HIPCHK(hipSetDevice(ctx->device));
    if (npairs) HIPCHK(hipMemcpy(out_count, ctx->d_count.p, npairs * sizeof(int32_t), hipMemcpyDeviceToHost));
    return CWHIP_OK;
}

// per-pair diagnostics of the last run: [maxScore,se1,se2,sb1,sb2,status,nodes,0]
int cwhip_fetch_info(cwhip_ctx* ctx, int32_t* out_info, int64_t npairs)
{
    if (!ctx || (!out_info && npairs) || npairs != ctx->npairs) return fail(ctx, CWHIP_E_ARG, "bad fetch arguments");
    HIPCHK(hipSetDevice(ctx->device));
    if (npairs) HIPCHK(hipMemcpy(out_info, ctx->d_info.p, 8 * npairs * sizeof(int32_t), hipMemcpyDeviceToHost));
    return CWHIP_OK;
}

int cwhip_pairwise_pairs(cwhip_ctx* ctx, const cwhip_pair_params* p, const int32_t* pairs,
                         int64_t npairs, int kernel, int32_t* out_count, float* out_dist)
{
    int rc = cwhip_stage_pairs(ctx, p, pairs, npairs);
    if (rc) return rc;
    if (npairs == 0) return CWHIP_OK;
    rc = cwhip_run_staged(ctx, kernel, 1, nullptr, nullptr);
    if (rc) return rc;
    std::vector<int32_t> cnt(npairs);
    rc = cwhip_fetch_counts(ctx, cnt.data(), npairs);
    if (rc) return rc;
    for (int64_t k = 0; k < npairs; ++k) {
        const int si = pairs[2 * k], sj = pairs[2 * k + 1];
        if (out_count) out_count[k] = cnt[k];
        if (out_dist)
            out_dist[k] = cnt[k] < 0 ? 1.0f : cwhip_distance(cnt[k], ctx->h_gapfree[si], ctx->h_gapfree[sj]);
    }
    return CWHIP_OK;
}

int cwhip_pairwise_range(cwhip_ctx* ctx, const cwhip_pair_params* p, int i_start, int i_end,
                         int j_start, int j_end, int kernel, int32_t* out_pairs,
                         int32_t* out_count, float* out_dist)
{
    if (!ctx) return CWHIP_E_ARG;
    // FullPairwiseAlign.cpp:53-58
    if (i_start < 0 || i_end < i_start || j_start < 0 || j_end < j_start)
        return fail(ctx, CWHIP_E_ARG, "The range for pairwise Alignment is incorrect.");
    const int n = ctx->nseq;
    std::vector<int32_t> pr;
    pr.reserve(2 * (size_t)cwhip_range_npairs(n, i_start, i_end, j_start, j_end));
    for (int si = std::max(0, i_start); si < n && si < i_end; ++si)
        for (int sj = std::max(si + 1, j_start + 1); sj < n && sj < j_end; ++sj) {
            pr.push_back(si);
            pr.push_back(sj);
        }
    const int64_t np = (int64_t)pr.size() / 2;
    if (out_pairs && np) std::memcpy(out_pairs, pr.data(), pr.size() * sizeof(int32_t));
    return cwhip_pairwise_pairs(ctx, p, pr.data(), np, kernel, out_count, out_dist);
}

} // extern "C"
