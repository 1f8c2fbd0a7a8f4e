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
hipError_t cw_launch_batched(int blocks, int threads, size_t lds, hipStream_t st, const DevArgs& a,
                             const CwSplitShape* s);
hipError_t cw_occupancy(int threads, size_t lds, int* plain_blocks, int* pers_blocks);

#include <map>
#include <mutex>

namespace {

constexpr int kPlainChunk = 16384;   // pairs per plain-kernel launch (scratch slots)

// Built-in batched shapes: every class at the plain kernel's stripe height and
// the full column path.  cwhip_autotune replaces them per device.
cwhip_tune_table default_table(int rmax)
{
    cwhip_tune_table t;
    for (int c = 0; c < CWHIP_TUNE_CLASSES; ++c) {
        t.rmax[c] = rmax;
        t.colcap[c] = CW_SMALL_COLS;
    }
    return t;
}

// tables found by cwhip_autotune, by device: contexts created later start from them
std::mutex g_tune_mu;
std::map<int, cwhip_tune_table> g_tuned;

bool table_ok(const cwhip_tune_table& t)
{
    for (int c = 0; c < CWHIP_TUNE_CLASSES; ++c)
        if (t.rmax[c] < 1 || t.rmax[c] > CW_RMAX || t.colcap[c] < 32 || t.colcap[c] > CW_SMALL_COLS)
            return false;
    return true;
}

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
    ctx->bind_account();
    ctx->tune = default_table(ctx->rmax);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(device);
        if (it != g_tuned.end()) ctx->tune = it->second;
    }
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
    ctx->release_buffers();
    ctx->h_prf_in.release(); ctx->h_prf_nodes.release(); ctx->h_rows.release();
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
    HIPCHK(ctx->d_gapfree.ensure(nseq + 1));
    if (total) HIPCHK(hipMemcpyAsync(ctx->d_codes.p, codes, total, hipMemcpyHostToDevice, ctx->stream));
    if (nseq) {
        HIPCHK(hipMemcpyAsync(ctx->d_offs.p, offsets, nseq * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->d_lens.p, lens, nseq * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->d_gapfree.p, ctx->h_gapfree.data(), nseq * sizeof(int32_t), hipMemcpyHostToDevice,
                              ctx->stream));
    }
    ctx->counts_ready = false;                    // distances of earlier sequences are stale
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
    if (rmax) {
        ctx->rmax = rmax;
        for (int c = 0; c < CWHIP_TUNE_CLASSES; ++c) ctx->tune.rmax[c] = rmax;
    }
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
    ctx->counts_ready = false;
    ctx->h_pairs.assign(pairs, pairs + 2 * npairs);
    ctx->gap1 = p->gap_pos1;
    ctx->gap2 = p->gap_pos2;
    ctx->ge = p->gap_extend;
    // the batched kernel's work order is built on the device at run time
    HIPCHK(ctx->d_pairs.ensure(2 * npairs + 2));
    HIPCHK(ctx->d_order.ensure(npairs + 1));
    HIPCHK(ctx->d_count.ensure(npairs + 1));
    HIPCHK(ctx->d_info.ensure(8 * npairs + 8));
    HIPCHK(ctx->d_matrix.ensure(32 * 32));
    HIPCHK(ctx->d_go.ensure(p->max_len + 1));
    HIPCHK(ctx->d_status.ensure(4));
    HIPCHK(ctx->d_queue.ensure(4));
    HIPCHK(ctx->d_bins.ensure(CW_BINS_INTS));
    if (npairs)
        HIPCHK(hipMemcpyAsync(ctx->d_pairs.p, pairs, 2 * npairs * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
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
    a.maxlen = ctx->maxlen;
    a.nseq = ctx->nseq;
    a.lsplit = 0x7fffffff;
    a.qpart = 0;
    a.scratch_stride = 8 * (int64_t)a.rowcap;
    a.out_count = ctx->d_count.p;
    a.out_info = ctx->d_info.p;
    a.status = ctx->d_status.p;
    a.queue = ctx->d_queue.p;
    a.bins = ctx->d_bins.p;
    a.order_out = ctx->d_order.p;
    for (int c = 0; c < CW_NCLASS; ++c) {
        a.rmax_cls[c] = ctx->rmax;
        a.colcap_cls[c] = CW_SMALL_COLS;
    }
}

// The batched stages as one graph launch.  A launch runs directly the first
// time it is seen and is captured when the same launch (buffers, sizes, shapes)
// comes again, so one-off jobs (a server's stream of different jobs) pay no
// capture and repeated runs of one staged job replay the graph.
static int launch_batched(cwhip_ctx* ctx, const DevArgs& a, int blocks, int threads, size_t lds,
                          const CwSplitShape* sp)
{
    std::vector<uint8_t> key(sizeof(DevArgs) + 3 * sizeof(int64_t) + (sp ? sizeof(CwSplitShape) : 0));
    std::memcpy(key.data(), &a, sizeof(DevArgs));
    const int64_t dims[3] = {blocks, threads, (int64_t)lds};
    std::memcpy(key.data() + sizeof(DevArgs), dims, sizeof dims);
    if (sp) std::memcpy(key.data() + sizeof(DevArgs) + sizeof dims, sp, sizeof(CwSplitShape));
    const bool replay = ctx->gexec && key == ctx->gkey;
    if (!ctx->use_graph || (!replay && key != ctx->glast)) {
        ctx->glast = key;
        HIPCHK(cw_launch_batched(blocks, threads, lds, ctx->stream, a, sp));
        return CWHIP_OK;
    }
    if (!replay) {
        if (ctx->gexec) {
            HIPCHK(hipGraphExecDestroy(ctx->gexec));
            ctx->gexec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t e1 = cw_launch_batched(blocks, threads, lds, ctx->stream, a, sp);
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hipfail(ctx, e1 != hipSuccess ? e1 : e2, "capturing the batched stages");
        }
        const hipError_t e3 = hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e3 != hipSuccess) {
            ctx->gexec = nullptr;
            return hipfail(ctx, e3, "hipGraphInstantiate");
        }
        ctx->gkey = key;
    }
    HIPCHK(hipGraphLaunch(ctx->gexec, ctx->stream));
    return CWHIP_OK;
}

// Persistent workgroups of a batched launch: as many as fit, never more waves
// than pairs (a small job leaves the rest of the chip to whatever runs beside it).
static int batched_shape(cwhip_ctx* ctx, const DevArgs& a, int64_t npairs, int* blocks, int* threads, size_t* lds)
{
    const int wpg = cw_fit_waves(ctx->wpg, a.wave_lds);
    if (wpg < 1) return fail(ctx, CWHIP_E_TOOLONG, "one wave's sequence buffers exceed the LDS");
    *threads = 64 * wpg;
    *lds = CW_TAB_BYTES + (size_t)wpg * a.wave_lds;
    int occ_plain = 0, occ_pers = 0;
    HIPCHK(cw_occupancy(*threads, *lds, &occ_plain, &occ_pers));
    if (occ_pers < 1 || occ_plain < 1) return fail(ctx, CWHIP_E_TOOLONG, "kernel does not fit on a CU");
    const int gpc = ctx->groups_per_cu ? std::min(ctx->groups_per_cu, occ_pers) : occ_pers;
    *blocks = (int)std::min<int64_t>((int64_t)ctx->num_cu * gpc, (npairs + wpg - 1) / wpg);
    return CWHIP_OK;
}

// Waves per CU of the batched kernel when every sequence is at most L long.
static int batched_waves_per_cu(cwhip_ctx* ctx, int L, int* wpg_out)
{
    const int wave_lds = cw_wave_lds_bytes(cw_lcap(L));
    const int wpg = cw_fit_waves(ctx->wpg, wave_lds);
    if (wpg < 1) return 0;
    int occ_plain = 0, occ_pers = 0;
    if (cw_occupancy(64 * wpg, CW_TAB_BYTES + (size_t)wpg * wave_lds, &occ_plain, &occ_pers) != hipSuccess)
        return 0;
    if (ctx->groups_per_cu) occ_pers = std::min(occ_pers, ctx->groups_per_cu);
    if (wpg_out) *wpg_out = wpg;
    return occ_pers * wpg;
}

// The longest sequence length at which the batched kernel still runs its most
// waves per CU, when the job's longest sequence would run fewer (0: no split).
// Pairs with a longer sequence then run first, in a launch of their own.
static int batched_split_len(cwhip_ctx* ctx)
{
    const int full = batched_waves_per_cu(ctx, ctx->maxlen, nullptr);
    const int most = batched_waves_per_cu(ctx, 64, nullptr);
    if (full >= most || ctx->maxlen <= 64) return 0;
    int lo = 64, hi = ctx->maxlen - 1;                 // lo keeps `most`, hi + 1 does not
    while (lo < hi) {
        const int mid = lo + (hi - lo + 1) / 2;
        if (batched_waves_per_cu(ctx, mid, nullptr) >= most) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

int cwhip_run_staged(cwhip_ctx* ctx, int kernel, int reps, float* ms, float* kernel_ms)
{
    if (!ctx || reps < 1) return fail(ctx, CWHIP_E_ARG, "bad run arguments");
    if (kernel != CWHIP_KERNEL_PLAIN && kernel != CWHIP_KERNEL_BATCHED) return fail(ctx, CWHIP_E_ARG, "unknown kernel");
    HIPCHK(hipSetDevice(ctx->device));
    if (ctx->npairs == 0) { if (ms) *ms = 0; if (kernel_ms) *kernel_ms = 0; return CWHIP_OK; }
    DevArgs a;
    fill_args(ctx, a);
    // fewer waves per workgroup where long sequences fill the LDS
    const int wpg = cw_fit_waves(ctx->wpg, a.wave_lds);
    if (wpg < 1) return fail(ctx, CWHIP_E_TOOLONG, "one wave's sequence buffers exceed the LDS");
    int threads = 64 * wpg, pers_blocks = 0;
    size_t lds = CW_TAB_BYTES + (size_t)wpg * a.wave_lds;
    int64_t slots;
    CwSplitShape split;
    std::memset(&split, 0, sizeof split);                // compared byte for byte as a graph key
    bool use_split = false;
    if (kernel == CWHIP_KERNEL_BATCHED) {
        for (int c = 0; c < CW_NCLASS; ++c) {
            a.rmax_cls[c] = ctx->tune.rmax[c];
            a.colcap_cls[c] = ctx->tune.colcap[c];
        }
        const int rc = batched_shape(ctx, a, ctx->npairs, &pers_blocks, &threads, &lds);
        if (rc) return rc;
        slots = (int64_t)pers_blocks * (threads / 64);
        // long sequences shrink the waves per CU: pairs with a sequence longer
        // than the split length run first with a's buffers, the rest after
        // them with buffers for the split length (more waves per CU)
        const int T = batched_split_len(ctx);
        if (T > 0) {
            a.lsplit = T;
            split.args = a;
            split.args.maxlen = T;
            split.args.lcap = cw_lcap(T);
            split.args.wave_lds = cw_wave_lds_bytes(split.args.lcap);
            const int rc2 = batched_shape(ctx, split.args, ctx->npairs, &split.blocks, &split.threads, &split.lds);
            if (rc2) return rc2;
            slots = std::max<int64_t>(slots, (int64_t)split.blocks * (split.threads / 64));
            use_split = true;
        }
    } else {
        int occ_plain = 0, occ_pers = 0;
        HIPCHK(cw_occupancy(threads, lds, &occ_plain, &occ_pers));
        if (occ_plain < 1) return fail(ctx, CWHIP_E_TOOLONG, "kernel does not fit on a CU");
        slots = std::min<int64_t>(kPlainChunk, ctx->npairs);
        slots = (slots + wpg - 1) / wpg * wpg;
    }
    HIPCHK(ctx->d_scratch.ensure((size_t)slots * a.scratch_stride));
    a.scratch = ctx->d_scratch.p;
    split.args.scratch = ctx->d_scratch.p;
    for (int rep = 0; rep < reps; ++rep) {
        HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
        if (kernel == CWHIP_KERNEL_BATCHED) {
            // queue reset, binning and the persistent kernel form one graph
            HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
            a.pair_base = 0;
            const int rc = launch_batched(ctx, a, pers_blocks, threads, lds, use_split ? &split : nullptr);
            if (rc) return rc;
        } else {
            HIPCHK(cw_launch_reset(ctx->stream, ctx->d_queue.p, ctx->d_status.p));
            HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
            a.order = nullptr;
            for (int64_t b0 = 0; b0 < ctx->npairs; b0 += slots) {
                a.pair_base = b0;
                a.pair_end = std::min<int64_t>(ctx->npairs, b0 + slots);
                const int64_t nb = (a.pair_end - b0 + wpg - 1) / wpg;
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
    HIPCHK(hipMemcpyAsync(&st, ctx->d_status.p, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (st) return fail(ctx, CWHIP_E_KERNEL, "kernel status " + std::to_string(st));
    ctx->counts_ready = true;
    return CWHIP_OK;
}

int cwhip_get_tune_table(cwhip_ctx* ctx, cwhip_tune_table* t)
{
    if (!ctx || !t) return fail(ctx, CWHIP_E_ARG, "null argument");
    *t = ctx->tune;
    return CWHIP_OK;
}

int cwhip_set_tune_table(cwhip_ctx* ctx, const cwhip_tune_table* t)
{
    if (!ctx || !t) return fail(ctx, CWHIP_E_ARG, "null argument");
    if (!table_ok(*t)) return fail(ctx, CWHIP_E_ARG, "tuning table value out of range");
    ctx->tune = *t;
    return CWHIP_OK;
}

int cwhip_set_graph(cwhip_ctx* ctx, int enable)
{
    if (!ctx) return CWHIP_E_ARG;
    ctx->use_graph = enable != 0;
    return CWHIP_OK;
}

// Candidate shapes of one class, without the ones that cannot differ for
// sequences no longer than lmax (a single stripe; every node on the column path).
static std::vector<std::pair<int, int>> tune_candidates(int lmax)
{
    static const int R[] = {2, 3, 4, 6, 8};
    static const int CC[] = {64, 128, 256};
    const int rcap = std::max(1, (lmax + 63) / 64);
    std::vector<std::pair<int, int>> out;
    for (int r : R) {
        const int re = std::min(r, rcap);
        for (int cc : CC) {
            const int ce = std::min(cc, std::max(32, lmax <= 32 ? 32 : lmax <= 64 ? 64 : lmax <= 128 ? 128 : 256));
            if (std::find(out.begin(), out.end(), std::make_pair(re, ce)) == out.end()) out.emplace_back(re, ce);
        }
    }
    return out;
}

int cwhip_autotune(cwhip_ctx* ctx, const cwhip_pair_params* p, const int32_t* pairs, int64_t npairs,
                   cwhip_tune_table* out, float* tune_ms)
{
    if (!ctx || !p || (npairs > 0 && !pairs) || npairs < 0) return fail(ctx, CWHIP_E_ARG, "bad autotune arguments");
    if (ctx->nseq <= 0) return fail(ctx, CWHIP_E_NOSEQ, "no sequences loaded");
    for (int64_t k = 0; k < 2 * npairs; ++k)
        if (pairs[k] < 0 || pairs[k] >= ctx->nseq) return fail(ctx, CWHIP_E_ARG, "pair index out of range");
    HIPCHK(hipSetDevice(ctx->device));
    const std::vector<int32_t> restage = ctx->h_pairs;
    const int64_t restage_n = ctx->npairs;
    // members of every class, and the longest sequence in each
    std::vector<std::vector<int64_t>> members(CW_NCLASS);
    std::vector<int> lmax(CW_NCLASS, 0);
    for (int64_t k = 0; k < npairs; ++k) {
        const int n = ctx->h_lens[pairs[2 * k]], m = ctx->h_lens[pairs[2 * k + 1]];
        if (n == 0 || m == 0) continue;
        const int c = cw_class_of(n, m);
        members[c].push_back(k);
        lmax[c] = std::max(lmax[c], std::max(n, m));
    }
    cwhip_tune_table best = ctx->tune;
    const cwhip_tune_table saved = ctx->tune;
    float spent = 0.f;
    DevArgs a0;
    fill_args(ctx, a0);
    int blocks = 0, threads = 0;
    size_t lds = 0;
    int rc = batched_shape(ctx, a0, INT64_MAX / 2, &blocks, &threads, &lds);
    if (rc) return rc;
    const int64_t waves = (int64_t)blocks * (threads / 64);
    for (int c = 0; c < CW_NCLASS; ++c) {
        const auto& mem = members[c];
        if (mem.empty()) continue;
        // a strided sample that fills every wave four times (fewer than that and
        // the queue's tail decides the time)
        const int64_t want = std::min<int64_t>((int64_t)mem.size(), 4 * waves);
        std::vector<int32_t> sp(2 * want);
        for (int64_t i = 0; i < want; ++i) {
            const int64_t k = mem[(size_t)(i * (int64_t)mem.size() / want)];
            sp[2 * i] = pairs[2 * k];
            sp[2 * i + 1] = pairs[2 * k + 1];
        }
        rc = cwhip_stage_pairs(ctx, p, sp.data(), want);
        if (rc) { ctx->tune = saved; return rc; }
        // the class's current shape first; a candidate replaces it only when it
        // is at least 3% faster (best of three runs each)
        std::vector<std::pair<int, int>> cands{{saved.rmax[c], saved.colcap[c]}};
        for (const auto& cand : tune_candidates(lmax[c]))
            if (cand != cands[0]) cands.push_back(cand);
        float best_ms = 1e30f;
        for (size_t ci = 0; ci < cands.size(); ++ci) {
            for (int k = 0; k < CW_NCLASS; ++k) {
                ctx->tune.rmax[k] = cands[ci].first;
                ctx->tune.colcap[k] = cands[ci].second;
            }
            float t = 0.f, tmin = 1e30f;
            for (int rep = 0; rep < 3; ++rep) {
                rc = cwhip_run_staged(ctx, CWHIP_KERNEL_BATCHED, 1, &t, nullptr);
                if (rc) { ctx->tune = saved; return rc; }
                spent += t;
                tmin = std::min(tmin, t);
            }
            if (ci == 0 || tmin < 0.97f * best_ms) {
                best_ms = tmin;
                best.rmax[c] = cands[ci].first;
                best.colcap[c] = cands[ci].second;
            }
        }
    }
    ctx->tune = best;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[ctx->device] = best;
    }
    if (out) *out = best;
    if (tune_ms) *tune_ms = spent;
    if (restage_n > 0) {
        rc = cwhip_stage_pairs(ctx, p, restage.data(), restage_n);
        if (rc) return rc;
    } else {
        ctx->npairs = 0;
        ctx->counts_ready = false;
        ctx->h_pairs.clear();
    }
    return CWHIP_OK;
}

int cwhip_fetch_counts(cwhip_ctx* ctx, int32_t* out_count, int64_t npairs)
{
    if (!ctx || (!out_count && npairs) || npairs != ctx->npairs) return fail(ctx, CWHIP_E_ARG, "bad fetch arguments");
    HIPCHK(hipSetDevice(ctx->device));
    if (npairs) {
        HIPCHK(hipMemcpyAsync(out_count, ctx->d_count.p, npairs * sizeof(int32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return CWHIP_OK;
}

// per-pair diagnostics of the last run: [maxScore,se1,se2,sb1,sb2,status,nodes,0]
int cwhip_fetch_info(cwhip_ctx* ctx, int32_t* out_info, int64_t npairs)
{
    if (!ctx || (!out_info && npairs) || npairs != ctx->npairs) return fail(ctx, CWHIP_E_ARG, "bad fetch arguments");
    HIPCHK(hipSetDevice(ctx->device));
    if (npairs) {
        HIPCHK(hipMemcpyAsync(out_info, ctx->d_info.p, 8 * npairs * sizeof(int32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
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

// ---------------------------------------------------------------------------
// Neighbour-joining on the device.

// the tree text into out (cwhip_nj_tree's sizing rules)
static int put_text(const std::string& s, char* out, int64_t cap, int64_t* out_len)
{
    *out_len = (int64_t)s.size();
    if (!out || cap < (int64_t)s.size() + 1) return CWHIP_E_ARG;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return CWHIP_OK;
}

// the DistMatrix of the last run, on the host, as FullPairwiseAlign leaves it
static int host_matrix_of_run(cwhip_ctx* ctx, std::vector<double>& D)
{
    const int n = ctx->nseq;
    std::vector<int32_t> cnt(ctx->npairs);
    const int rc = cwhip_fetch_counts(ctx, cnt.data(), ctx->npairs);
    if (rc) return rc;
    D.assign((size_t)(n + 1) * (n + 1), 0.0);
    for (int64_t k = 0; k < ctx->npairs; ++k) {
        const int si = ctx->h_pairs[2 * k], sj = ctx->h_pairs[2 * k + 1];
        const float d = cnt[k] < 0 ? 1.0f : cwhip_distance(cnt[k], ctx->h_gapfree[si], ctx->h_gapfree[sj]);
        D[(size_t)(si + 1) * (n + 1) + sj + 1] = D[(size_t)(sj + 1) * (n + 1) + si + 1] = (double)d;
    }
    return CWHIP_OK;
}

// cw_nj over D (built from the run's counts when `from_run`, else uploaded
// from `dist`); the text from the kernel's treeDesc and branch lengths.
// Returns 1 when the kernel left the tree to the host.
static int nj_device(cwhip_ctx* ctx, int n, bool from_run, const double* dist, const char* const* names,
                     std::string& text)
{
    const int64_t ld = n + 1;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(ctx->t_dist.ensure((size_t)ld * ld));
    HIPCHK(ctx->t_desc.ensure((size_t)ld * ld));
    HIPCHK(ctx->t_branch.ensure(2 * (size_t)(n + 2)));
    HIPCHK(ctx->t_status.ensure(1));
    NjArgs a{};
    if (from_run) {
        a.pairs = ctx->d_pairs.p;
        a.count = ctx->d_count.p;
        a.gapfree = ctx->d_gapfree.p;
        a.npairs = ctx->npairs;
    } else {
        HIPCHK(hipMemcpyAsync(ctx->t_dist.p, dist, (size_t)ld * ld * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
    }
    a.D = ctx->t_dist.p;
    a.desc = ctx->t_desc.p;
    a.left = ctx->t_branch.p;
    a.right = ctx->t_branch.p + (n + 2);
    a.status = ctx->t_status.p;
    a.n = n;
    HIPCHK(cw_launch_nj(a, ctx->stream));
    std::vector<int32_t> desc((size_t)ld * ld);
    std::vector<double> br(2 * (size_t)(n + 2));
    int32_t st = 0;
    HIPCHK(hipMemcpyAsync(desc.data(), ctx->t_desc.p, desc.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(br.data(), ctx->t_branch.p, br.size() * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(&st, ctx->t_status.p, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (st) return 1;
    std::vector<double> left(br.begin(), br.begin() + (n + 2)), right(br.begin() + (n + 2), br.end());
    text = cw_phylip_text(n, desc, left, right, names);
    return CWHIP_OK;
}

int cwhip_nj_tree_from_run(cwhip_ctx* ctx, const char* const* names, char* out, int64_t cap, int64_t* out_len)
{
    if (!ctx || !names || !out_len) return fail(ctx, CWHIP_E_ARG, "null argument");
    const int n = ctx->nseq;
    for (int i = 0; i < n; ++i)
        if (!names[i]) return fail(ctx, CWHIP_E_ARG, "null name");
    if (n < 2) return fail(ctx, CWHIP_E_ARG, "a tree needs two sequences");
    // the run must have covered every pair of the loaded sequences
    if (!ctx->counts_ready || ctx->npairs != (int64_t)n * (n - 1) / 2)
        return fail(ctx, CWHIP_E_ARG, "no finished run over all pairs of the loaded sequences");
    std::vector<bool> seen((size_t)n * n, false);
    for (int64_t k = 0; k < ctx->npairs; ++k) {
        int si = ctx->h_pairs[2 * k], sj = ctx->h_pairs[2 * k + 1];
        if (si > sj) std::swap(si, sj);
        if (si == sj || seen[(size_t)si * n + sj])
            return fail(ctx, CWHIP_E_ARG, "no finished run over all pairs of the loaded sequences");
        seen[(size_t)si * n + sj] = true;
    }
    if (n >= 4 && n <= CW_NJ_MAX_SEQS) {
        std::string s;
        const int rc = nj_device(ctx, n, true, nullptr, names, s);
        if (rc == CWHIP_OK) return put_text(s, out, cap, out_len);
        if (rc != 1) return rc;
    }
    std::vector<double> D;                          // two or three sequences, or past the kernel's size
    const int rc = host_matrix_of_run(ctx, D);
    if (rc) return rc;
    return cwhip_nj_tree(n, D.data(), names, out, cap, out_len);
}

int cwhip_nj_tree_on_device(cwhip_ctx* ctx, int nseq, const double* dist, const char* const* names, char* out,
                            int64_t cap, int64_t* out_len)
{
    if (!ctx || nseq < 2 || !dist || !names || !out_len) return fail(ctx, CWHIP_E_ARG, "bad tree arguments");
    for (int i = 0; i < nseq; ++i)
        if (!names[i]) return fail(ctx, CWHIP_E_ARG, "null name");
    // the kernel reads D(jj, ii) for D(ii, jj): a matrix that is not symmetric
    // goes to the host code, which reads it as the reference does
    bool sym = true;
    const int64_t ld = nseq + 1;
    for (int i = 1; i <= nseq && sym; ++i)
        for (int j = i + 1; j <= nseq; ++j)
            if (dist[i * ld + j] != dist[j * ld + i]) { sym = false; break; }
    if (sym && nseq >= 4 && nseq <= CW_NJ_MAX_SEQS) {
        std::string s;
        const int rc = nj_device(ctx, nseq, false, dist, names, s);
        if (rc == CWHIP_OK) return put_text(s, out, cap, out_len);
        if (rc != 1) return rc;
    }
    return cwhip_nj_tree(nseq, dist, names, out, cap, out_len);
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
