// cwhip_prf.cpp -- host side of the progressive (profile) alignment entry
// points of include/cwhip.h.
//
// Paths below are relative to CPU/clustalw-2.0.11/src of coldfunction/CUDA-clustalW.
//
// The Myers-Miller DP of every step runs in the gfx950 kernels of
// cw_profile_kernels.hip (cw_prf_score, cw_prf_mm); this file builds the two
// profiles of a step (ProfileBase / ProfileWithSub / ProfileStandard), turns
// the device's node table into the reference's edit script, applies it
// (progTracepath, addGGaps) and drives MSA::multiSeqAlign, batching the steps
// of disjoint subtrees into one device launch.  No step is computed on the
// host: without a device cwhip_msa fails like every other entry point.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "cw_matrices.inc"
#include "cwhip_internal.h"

hipError_t cwp_launch_score(int tiles, hipStream_t st, const PrfArgs& a);
hipError_t cwp_launch_mm(int nprob, hipStream_t st, const PrfArgs& a);
hipError_t cwp_launch_build(int ncolblk, int nside, hipStream_t st, const BuildArgs& a);
hipError_t cwp_launch_local(int groups, hipStream_t st, const PrfArgs& a);
hipError_t cwp_launch_wide(int ntask, hipStream_t st, const PrfArgs& a, const WideArgs& w);
hipError_t cwp_launch_level(int nprob, int T, hipStream_t st, const PrfArgs& a);

namespace {

constexpr int NUMRES = 32;
constexpr int MAXAA = 24;               // UserParameters.cpp:133
constexpr int GAP1 = 30, GAP2 = 31;     // UserParameters.cpp:38-39
constexpr int GAPCOL = CWP_GAPCOL, LENCOL = CWP_LENCOL, W = CWP_PRF_W;
constexpr int ENDALN = 127;             // general/clustalw.h:113
constexpr int INT_SCALE = 1000;         // INT_SCALE_FACTOR, general/clustalw.h:226 (OS_UNIX)

const char kCodes[] = "ABCDEFGHIKLMNOPQRSTUVWXYZ-";    // UserParameters.cpp:132
const char kAAOrder[] = "ABCDEFGHIKLMNPQRSTVWXYZ";     // SubMatrix.cpp:179
const char kDNAOrder[] = "ABCDGHKMNRSTUVWXY";          // SubMatrix.cpp:180

int code_of(char c)
{
    for (int i = 0; kCodes[i]; ++i)
        if (kCodes[i] == c) return i;
    return -1;
}

// ---------------------------------------------------------------------------
// Device batch of profile pairs.

struct DpJob {
    const int32_t* prf1;
    const int32_t* prf2;
    int L1, L2;
    int32_t* displ;        // 1-based scratch of L1 + L2 + 3 ints, filled by emit()
    int n_displ = 0;
    int score = 0;
    double cells = 0;
};

inline int64_t up4(int64_t x) { return (x + 3) & ~int64_t(3); }

// The edit-script writer of progDel / progAdd / progAlign (:1067-1117).
struct Emitter {
    int32_t* displ;
    int printPtr = 1, lastPrint = 0;
    void del(int k) {
        if (lastPrint < 0) lastPrint = displ[printPtr - 1] -= k;
        else lastPrint = displ[printPtr++] = -k;
    }
    void add(int k) {
        if (lastPrint < 0) {
            displ[printPtr - 1] = k;
            displ[printPtr++] = lastPrint;
        } else {
            lastPrint = displ[printPtr++] = k;
        }
    }
    void align() { displ[printPtr++] = lastPrint = 0; }
};

// Walks the node table in progDiff's depth-first order (:779-809) and emits
// the base cases of :540-600.  Returns false on a malformed table.
bool emit_script(const int32_t* nodes, int nnodes, Emitter& em, double* cells)
{
    std::vector<int> stack;   // >= 0: node to emit; -1: the del(2) of a type-2 split
    stack.push_back(0);
    double c = 0;
    while (!stack.empty()) {
        const int x = stack.back();
        stack.pop_back();
        if (x == -1) { em.del(2); continue; }
        if (x < 0 || x >= nnodes) return false;
        const int32_t* nd = nodes + (int64_t)x * CWP_NODE_INTS;
        const int M = nd[2], N = nd[3], type = nd[5], midj = nd[6], child = nd[7];
        if (type == 0) {
            if (N <= 0) {
                if (M > 0) em.del(M);
            } else if (M <= 0) {
                em.add(N);
            } else {                                   // M == 1
                if (midj == 0) {
                    em.add(N);
                    em.del(1);
                } else {
                    if (midj > 1) em.add(midj - 1);
                    em.align();
                    if (midj < N) em.add(N - midj);
                }
            }
            continue;
        }
        if (child < 0 || child + 1 >= nnodes) return false;
        c += (double)M * N;
        stack.push_back(child + 1);
        if (type == 2) stack.push_back(-1);
        stack.push_back(child);
    }
    *cells = c;
    return true;
}

// Host threads shared by every context of the process: up to 15 workers that
// sleep between jobs, so a parallel loop costs a wake-up, not thread start-ups.
// One loop runs on the pool at a time; a loop started from a pool worker (or
// while another loop holds the pool) runs on its caller alone.
class HostPool {
public:
    static HostPool& get() {
        static HostPool pool;
        return pool;
    }
    template <typename F>
    void run(size_t n, F& fn) {
        std::unique_lock<std::mutex> busy(run_mu_, std::try_to_lock);
        if (!busy.owns_lock() || in_worker_ || workers_.empty() || n <= 1) {
            for (size_t k = 0; k < n; ++k) fn(k);
            return;
        }
        std::atomic<size_t> next{0};
        std::function<void()> body = [&]() {
            for (size_t k; (k = next.fetch_add(1)) < n;) fn(k);
        };
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &body;
            pending_ = (int)workers_.size();
            ++gen_;
        }
        cv_.notify_all();
        body();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
    }
private:
    HostPool() {
        const int hw = (int)std::max<unsigned>(1, std::min<unsigned>(16, std::thread::hardware_concurrency()));
        for (int t = 1; t < hw; ++t)
            workers_.emplace_back([this]() {
                in_worker_ = true;
                uint64_t seen = 0;
                for (;;) {
                    std::function<void()>* job;
                    {
                        std::unique_lock<std::mutex> lk(mu_);
                        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                        if (stop_) return;
                        seen = gen_;
                        job = job_;
                    }
                    (*job)();
                    std::lock_guard<std::mutex> lk(mu_);
                    if (--pending_ == 0) done_cv_.notify_one();
                }
            });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    std::vector<std::thread> workers_;
    std::mutex mu_, run_mu_;
    std::condition_variable cv_, done_cv_;
    std::function<void()>* job_ = nullptr;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
    static thread_local bool in_worker_;
};
thread_local bool HostPool::in_worker_ = false;

// fn(0..n-1) on the host pool (the caller's thread included)
template <typename F>
void parallel_for(size_t n, F fn)
{
    HostPool::get().run(n, fn);
}

// Offsets of a batch of profile pairs in the pool: inputs (raw profiles,
// {gap, len, mask, gap + len} records per row and column, LENCOL prefix sums)
// from 0, then the work area: the node tables of all pairs first (one copy
// back), then per pair the score matrix, per-wave scratch rows and level buffers.
struct BatchLayout {
    std::vector<PrfProb> probs;
    std::vector<int32_t> tile_base;
    bool wide = false;               // few pairs: CWP_WIDE_SETS scratch sets each, for
                                     // cw_prf_wide and cw_prf_level
    int64_t in_sz = 0, nodes_sz = 0, total = 0;
    int64_t list = 0, ctr = 0;       // the subtrees left to cw_prf_local, and its two counters
    int64_t state = 0;               // cw_prf_mm's level state, 4 per pair
};

// Rows of the smaller half of a root that cw_prf_wide may sweep: each of its
// G >= 2 stripes keeps at least one row per lane.
constexpr int kWideMinHalf = 128;
constexpr int kWideMinCols = 64;

// cw_prf_level's tasks per pair at level k >= 1: both halves of up to 2^k nodes
int wide_tasks(int k) { return k < 1 ? 0 : 2 << k; }
// levels whose half sweeps run before cw_prf_mm's last launch (the root's
// through cw_prf_wide): every task of the deepest one has a scratch set
constexpr int kWideMaxLevels = 6;
static_assert(2 << (kWideMaxLevels - 1) <= CWP_WIDE_SETS, "scratch sets for cw_prf_level");
int wide_levels(const cwhip_ctx* ctx) { return std::max(1, std::min(ctx->prf_wide_levels, kWideMaxLevels)); }

// Batches of more pairs than this already give cw_prf_wide more stripes than
// the chip has SIMDs: their roots take as few stripes as fit (less pipeline
// lag for the same work; 0.55 ms less per cwhip_msa on 1000 x ~1002).
constexpr int kWideFewPairs = 64;

// cw_prf_wide's members per half for a root of L1 x L2 (0: leave it to
// cw_prf_mm): about `rows` rows per stripe, 2..CWP_WIDE_MAXG stripes of
// 64..512 rows in both halves; with `tallest`, the fewest stripes of at most
// 512 rows.
int wide_members(int L1, int L2, int rows, bool tallest = false)
{
    const int h = L1 / 2;                            // forward half; the reverse has h or h + 1
    if (rows <= 0 || h < kWideMinHalf || L2 < kWideMinCols) return 0;
    int G = tallest ? (L1 - h + 511) / 512 : std::min(CWP_WIDE_MAXG, std::max(2, h / rows));
    G = std::min(G, h / 64);
    if (G < (tallest ? 1 : 2) || (L1 - h + G - 1) / G > 512) return 0;
    return G;
}

void layout_batch(const std::vector<DpJob>& jobs, BatchLayout& lo, const cwhip_ctx* ctx)
{
    const int n = (int)jobs.size();
    lo.probs.assign(n, PrfProb{});
    lo.tile_base.assign(n + 1, 0);
    lo.wide = n <= ctx->prf_wide_pairs;
    // cw_prf_wide's 2 CWP_WIDE_MAXG stripes, cw_prf_level's 2^levels tasks per pair
    const int wide_sets = std::max({CWP_WAVES, 2 * CWP_WIDE_MAXG, wide_tasks(wide_levels(ctx) - 1)});
    int64_t in_sz = 0, work_sz = 0;
    for (int k = 0; k < n; ++k) {
        const int L1 = jobs[k].L1, L2 = jobs[k].L2;
        PrfProb& pp = lo.probs[k];
        pp.L1 = L1;
        pp.L2 = L2;
        pp.ldS = L1 + 2;
        pp.node_cap = 2 * L1 + 4;
        pp.rowcap = (int)up4(L2 + 2);
        pp.prf1 = in_sz; in_sz = up4(in_sz + (int64_t)(L1 + 2) * W);
        pp.prf2 = in_sz; in_sz = up4(in_sz + (int64_t)(L2 + 2) * W);
        pp.row = in_sz; in_sz += 4 * (int64_t)(L1 + 2);
        pp.col = in_sz; in_sz += 4 * (int64_t)(L2 + 2);
        pp.pl1 = in_sz; in_sz = up4(in_sz + L1 + 2);
        pp.pl2 = in_sz; in_sz = up4(in_sz + L2 + 2);
        lo.tile_base[k + 1] = lo.tile_base[k] + ((L1 + 63) / 64) * ((L2 + 63) / 64);
    }
    for (int k = 0; k < n; ++k) {
        lo.probs[k].nodes = work_sz;
        work_sz = up4(work_sz + (int64_t)lo.probs[k].node_cap * CWP_NODE_INTS);
    }
    lo.nodes_sz = work_sz;
    for (int k = 0; k < n; ++k) {
        PrfProb& pp = lo.probs[k];
        pp.S = work_sz; work_sz = up4(work_sz + (int64_t)(pp.L2 + 2) * pp.ldS);
        const int sets = lo.wide ? wide_sets : CWP_WAVES;
        pp.scratch = work_sz; work_sz += (int64_t)sets * CWP_SCR_ROWS * pp.rowcap;
        pp.lvl = work_sz; work_sz = up4(work_sz + 4 * ((int64_t)pp.L2 + 2 + pp.node_cap));
    }
    for (auto& pp : lo.probs) { pp.S += in_sz; pp.nodes += in_sz; pp.scratch += in_sz; pp.lvl += in_sz; }
    int64_t roots = 0;                 // a local subtree's root is a node of its problem
    for (const auto& pp : lo.probs) roots += pp.node_cap;
    lo.in_sz = in_sz;
    lo.list = in_sz + work_sz;
    lo.ctr = up4(lo.list + 3 * roots);
    lo.state = lo.ctr + 4;
    lo.total = lo.state + up4(4 * (int64_t)n);
}

// The score and Myers-Miller kernels over a batch whose inputs are in the pool
// (queued on the context's stream behind whatever wrote them; ev0 is recorded
// by the caller before its first kernel), then the node tables back in one
// copy and the edit scripts on host threads.
int dp_launch_collect(cwhip_ctx* ctx, BatchLayout& lo, std::vector<DpJob>& jobs)
{
    const int n = (int)jobs.size();
    // the roots whose half sweeps cw_prf_wide spreads over CUs
    std::vector<int32_t> wide;
    for (int k = 0; k < n; ++k) {
        PrfProb& pp = lo.probs[k];
        const int G = lo.wide ? wide_members(pp.L1, pp.L2, ctx->prf_wide_rows, n > kWideFewPairs) : 0;
        if (G == 0) continue;
        pp.flags |= CWP_ROOT_WIDE;
        for (int half = 0; half < 2; ++half)
            for (int m = 0; m < G; ++m) wide.insert(wide.end(), {k, half, m, G});
    }
    const int nwide = (int)wide.size() / 4;
    HIPCHK(ctx->p_probs.ensure(n));
    HIPCHK(ctx->p_out.ensure(4 * (size_t)n));
    HIPCHK(ctx->p_tiles.ensure(n + 1));
    HIPCHK(hipMemcpyAsync(ctx->p_probs.p, lo.probs.data(), n * sizeof(PrfProb), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->p_tiles.p, lo.tile_base.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                          ctx->stream));
    PrfArgs a;
    a.probs = ctx->p_probs.p;
    a.pool = ctx->p_pool.p;
    a.out = ctx->p_out.p;
    a.tile_base = ctx->p_tiles.p;
    a.nprob = n;
    a.rmax = ctx->prf_rmax;
    a.local_list = ctx->p_pool.p + lo.list;
    a.local_ctr = ctx->p_pool.p + lo.ctr;
    a.state = ctx->p_pool.p + lo.state;
    a.resume = 0;
    a.ext = 0;
    HIPCHK(hipMemsetAsync(a.local_ctr, 0, 2 * sizeof(int32_t), ctx->stream));
    WideArgs wa{};
    if (nwide > 0) {
        HIPCHK(ctx->p_wide.ensure(5 * (size_t)nwide + 1));
        HIPCHK(hipMemcpyAsync(ctx->p_wide.p, wide.data(), wide.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                              ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->p_wide.p + 4 * nwide, 0, ((size_t)nwide + 1) * sizeof(int32_t), ctx->stream));
        wa.tasks = ctx->p_wide.p;
        wa.prog = ctx->p_wide.p + 4 * nwide;
        wa.err = wa.prog + nwide;
        wa.wait_limit = ctx->prf_wide_wait;
    }
    if (lo.tile_base[n] > 0) HIPCHK(cwp_launch_score(lo.tile_base[n], ctx->stream, a));
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    if (nwide > 0) HIPCHK(cwp_launch_wide(nwide, ctx->stream, a, wa));
    if (lo.wide && wide_levels(ctx) > 1) {
        // levels 1 .. prf_wide_levels-1: their half sweeps one wave each over
        // the chip, then cw_prf_mm finishes the level; the rest in one launch
        a.ext = 1;
        HIPCHK(cwp_launch_mm(n, ctx->stream, a));
        a.resume = 1;
        for (int k = 1; k < wide_levels(ctx); ++k) {
            HIPCHK(cwp_launch_level(n, wide_tasks(k), ctx->stream, a));
            HIPCHK(cwp_launch_mm(n, ctx->stream, a));
        }
        a.ext = 0;
    }
    HIPCHK(cwp_launch_mm(n, ctx->stream, a));
    HIPCHK(cwp_launch_local(std::max(1, ctx->num_cu), ctx->stream, a));
    HIPCHK(hipEventRecord(ctx->ev2, ctx->stream));
    HIPCHK(ctx->h_prf_nodes.ensure(lo.nodes_sz + 4 * (size_t)n + 4));
    int32_t* h_nodes = ctx->h_prf_nodes.p;
    int32_t* out = h_nodes + lo.nodes_sz;
    HIPCHK(hipMemcpyAsync(out, ctx->p_out.p, 4 * (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    out[4 * n] = 0;
    if (nwide > 0)
        HIPCHK(hipMemcpyAsync(out + 4 * n, wa.err, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(h_nodes, ctx->p_pool.p + lo.in_sz, lo.nodes_sz * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms1 = 0, ms2 = 0;
    HIPCHK(hipEventElapsedTime(&ms1, ctx->ev0, ctx->ev1));
    HIPCHK(hipEventElapsedTime(&ms2, ctx->ev1, ctx->ev2));
    ctx->prf_score_ms += ms1;
    ctx->prf_mm_ms += ms2;
    if (out[4 * n] != 0) return fail(ctx, CWHIP_E_KERNEL, "profile DP stripe hand-off timed out");
    for (int k = 0; k < n; ++k) {
        const int nn = out[4 * k + 1];
        if (out[4 * k + 2] != 0 || nn < 1 || nn > lo.probs[k].node_cap)
            return fail(ctx, CWHIP_E_KERNEL, "profile DP node table overflow");
    }
    std::vector<char> bad(n, 0);
    std::vector<double> cells(n, 0.0);
    parallel_for(n, [&](size_t k) {
        Emitter em{jobs[k].displ};
        if (!emit_script(h_nodes + (lo.probs[k].nodes - lo.in_sz), out[4 * k + 1], em, &cells[k])) {
            bad[k] = 1;
            return;
        }
        jobs[k].n_displ = em.printPtr - 1;
        jobs[k].score = out[4 * k + 0];
        jobs[k].cells = cells[k];
    });
    for (int k = 0; k < n; ++k) {
        if (bad[k]) return fail(ctx, CWHIP_E_KERNEL, "profile DP node table is malformed");
        ctx->prf_cells += cells[k];
        ctx->prf_entries += (double)jobs[k].L1 * jobs[k].L2;
    }
    return CWHIP_OK;
}

// One device batch of profile pairs whose profiles the caller built: the
// inputs packed (on host threads) into page-locked staging and uploaded in one
// copy, then dp_launch_collect.
int run_dp_batch(cwhip_ctx* ctx, std::vector<DpJob>& jobs, int egp)
{
    const int n = (int)jobs.size();
    if (n == 0) return CWHIP_OK;
    HIPCHK(hipSetDevice(ctx->device));
    BatchLayout lo;
    layout_batch(jobs, lo, ctx);
    for (int k = 0; k < n; ++k)
        lo.probs[k].flags = (jobs[k].prf1[GAPCOL] != 0 ? 1 : 0) |
                            (jobs[k].prf1[(int64_t)jobs[k].L1 * W + GAPCOL] != 0 ? 2 : 0);
    // host images of the inputs: raw profiles, {gap, len, mask, gap+len}
    // records and LENCOL prefix sums (the mask is 0 at the profile ends when
    // end-gap penalties are off, the rule of openPenalty1/2 and extPenalty1/2)
    HIPCHK(ctx->h_prf_in.ensure(lo.in_sz + 4));
    int32_t* h_in = ctx->h_prf_in.p;
    parallel_for(n, [&](size_t k) {
        const PrfProb& pp = lo.probs[k];
        const DpJob& j = jobs[k];
        std::memcpy(&h_in[pp.prf1], j.prf1, sizeof(int32_t) * (size_t)(j.L1 + 2) * W);
        std::memcpy(&h_in[pp.prf2], j.prf2, sizeof(int32_t) * (size_t)(j.L2 + 2) * W);
        for (int side = 0; side < 2; ++side) {
            const int L = side ? j.L2 : j.L1;
            const int32_t* prf = side ? j.prf2 : j.prf1;
            int32_t* rec = &h_in[side ? pp.col : pp.row];
            int32_t* pl = &h_in[side ? pp.pl2 : pp.pl1];
            int32_t acc = 0;
            for (int x = 0; x <= L + 1; ++x) {
                const int g = prf[(int64_t)x * W + GAPCOL], l = prf[(int64_t)x * W + LENCOL];
                const bool end = !egp && (x == 0 || x == L);
                rec[4 * x + 0] = g;
                rec[4 * x + 1] = l;
                rec[4 * x + 2] = end ? 0 : -1;
                rec[4 * x + 3] = g + l;
                pl[x] = acc;
                acc += l;
            }
        }
    });
    HIPCHK(ctx->p_pool.ensure(lo.total + 4));
    HIPCHK(hipMemcpyAsync(ctx->p_pool.p, h_in, lo.in_sz * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    return dp_launch_collect(ctx, lo, jobs);
}

// ---------------------------------------------------------------------------
// Profiles of one step.

// SubMatrix::getMatrix (SubMatrix.cpp:689-829) on a code-indexed table whose
// alphabet is `order`, negFlag = neg.
int get_matrix(const short* tab, const char* order, bool neg, int scale, int mat[NUMRES][NUMRES],
               int* mat_avg)
{
    const int nlet = (int)std::strlen(order);
    const int max_res = nlet - 1;
    for (int i = 0; i < NUMRES; ++i)
        for (int j = 0; j < NUMRES; ++j) mat[i][j] = 0;
    for (int i = 0; i < 25; ++i)
        for (int j = 0; j < 25; ++j) mat[i][j] = tab[i * 25 + j] * scale;
    int av3 = 0;
    for (int i = 0; i <= MAXAA; ++i)
        for (int j = 0; j < i; ++j) av3 += mat[i][j];
    av3 = (int)(av3 / (((float)(max_res * max_res - max_res)) / 2));
    *mat_avg = -av3;
    int mn = mat[0][0];
    for (int i = 0; i <= MAXAA; ++i)
        for (int j = 1; j <= i; ++j) mn = std::min(mn, mat[i][j]);
    if (!neg && mn < 0)                                  // :777-798
        for (int a = 0; a < nlet; ++a)
            for (int b = 0; b < nlet; ++b) mat[code_of(order[a])][code_of(order[b])] -= mn;
    for (int i = 0; i < GAP1; ++i) {
        mat[i][GAP1] = mat[GAP1][i] = 0;
        mat[i][GAP2] = mat[GAP2][i] = 0;
    }
    mat[GAP1][GAP1] = mat[GAP2][GAP2] = mat[GAP1][GAP2] = mat[GAP2][GAP1] = 0;
    return max_res + 2;
}

// a caller's profile_matrix callback is not assumed to be reentrant (the
// reference's SubMatrix keeps state in members)
std::mutex g_matrix_cb_mu;

// SubMatrix::getProfileAlignMatrix (SubMatrix.cpp:403-677): the caller's
// callback when there is one, else the default matrix names ("gonnet" series,
// "iub").  Returns maxRes (<= 0: the step fails).
int profile_matrix(const cwhip_msa_params& p, double pcid, int minLen, int mat[NUMRES][NUMRES],
                   float* scale_out, int* mat_avg)
{
    if (p.profile_matrix) {
        std::lock_guard<std::mutex> lk(g_matrix_cb_mu);
        float sc = 1.0f;                 // MyersMillerProfileAlign.cpp:249-251
        int32_t avg = 0;
        for (int i = 0; i < NUMRES; ++i)
            for (int j = 0; j < NUMRES; ++j) mat[i][j] = 0;
        const int r = p.profile_matrix(p.profile_matrix_user, pcid, minLen, &mat[0][0], &sc, &avg);
        *scale_out = sc;
        *mat_avg = avg;
        return r;
    }
    float scale;
    float int_scale = 100;
    int r;
    if (p.dna) {
        scale = 1.0f;
        r = get_matrix(cw_iub_dna, kDNAOrder, p.neg_matrix != 0, (int)int_scale, mat, mat_avg);
        // the transitions through the default DNA cross reference: A-G, C-T, C-U
        const int tw = (int)(p.transition_weight * mat[0][0]);
        const int A = code_of('A'), G = code_of('G'), C = code_of('C'), T = code_of('T'), U = code_of('U');
        mat[A][G] = mat[G][A] = tw;
        mat[C][T] = mat[T][C] = tw;
        mat[C][U] = mat[U][C] = tw;
    } else {
        const short* tab;
        scale = 0.75f;
        scale /= 2.0;
        if (p.neg_matrix || !p.distance_tree) {
            tab = cw_gonnet250;
        } else if (pcid > 35.0) {
            tab = cw_gonnet80;
            scale /= 2.0;
        } else if (pcid > 25.0) {
            tab = minLen < 100 ? cw_gonnet250 : cw_gonnet120;
        } else {
            tab = minLen < 100 ? cw_gonnet350 : cw_gonnet160;
        }
        int_scale /= 10;
        r = get_matrix(tab, kAAOrder, p.neg_matrix != 0, (int)int_scale, mat, mat_avg);
    }
    *scale_out = scale;
    return r;
}

inline bool is_gap_code(int c) { return c < 0 || c > MAXAA; }

// The switches ProfileBase::calcGapCoeff takes for one profile (:64-95): a
// group of two rows more than 60% identical uses the variable penalties
// instead of the hydrophilic, residue-specific and gap-distance ones.
struct SideSpec {
    int first = 0, last = 0, L = 0;    // member rows [first, last) of the step, prfLength
    int flags = 0, gdist = 0, sum2 = 0;
};

// One profileAlign call (MyersMillerProfileAlign.cpp:36-505) split into the
// part before the DP (prepare: groups, pcid, the profile matrix, gap
// coefficients, the profile switches; the profiles are built on the device)
// and the part after it (apply).
struct Step {
    std::vector<int> members;          // group-1 sequences, then group-2 ones
    int n1 = 0, n2 = 0, L1 = 0, L2 = 0;
    std::vector<int> w;                // weights of the members
    std::vector<int32_t> mat;          // NUMRES x NUMRES profile matrix
    int gapcoef = 0, lencoef = 0;
    SideSpec s1, s2;
    std::vector<int32_t> displ;
    DpJob job;
};

using Seqs = std::vector<std::vector<uint8_t>>;

// row r of a group whose profile length is L: the sequence, then GAP1 up to L
inline int row_at(const std::vector<uint8_t>& r, int i) { return i < (int)r.size() ? r[i] : GAP1; }

void side_spec(const cwhip_msa_params& p, const Seqs& seqs, const Step& st, SideSpec& s, bool sub)
{
    const int numSeq = s.last - s.first;
    int pcid = 0;
    if (numSeq == 2) {                                   // percentId (ProfileBase.cpp:519-547)
        const std::vector<uint8_t>& s1 = seqs[st.members[s.first]];
        const std::vector<uint8_t>& s2 = seqs[st.members[s.first + 1]];
        int count = 0, total = 0;
        for (int i = 0; i < s.L; ++i) {
            const int c1 = row_at(s1, i);
            if (c1 < MAXAA) {
                ++total;
                if (c1 == row_at(s2, i)) ++count;
            }
        }
        pcid = total == 0 ? 0 : (int)(float)(100.0 * (float)count / (float)total);
    }
    bool nVarPen, nHydPen, nPrefPen;
    s.gdist = 0;
    if (!p.no_var_penalties && pcid > 60) {
        nHydPen = nPrefPen = true;
        nVarPen = false;
    } else {
        nVarPen = true;
        nHydPen = p.no_hyd_penalties != 0;
        nPrefPen = p.no_pref_penalties != 0;
        s.gdist = p.gap_dist;
    }
    s.flags = (sub ? CWP_F_SUB : 0) | (p.end_gap_penalties ? CWP_F_EGP : 0);
    if (p.dna) s.flags |= CWP_F_DNA;
    else
        s.flags |= (nVarPen ? 0 : CWP_F_VAR) | (nHydPen ? 0 : CWP_F_HYD) | (nPrefPen ? 0 : CWP_F_PREF);
    s.sum2 = 0;
    for (int k = s.first; k < s.last; ++k) s.sum2 += st.w[k];
}

// Returns 1 when the step runs, 0 when a group is empty (profileAlign returns
// 0, :96-99), -1 when its profile matrix fails (profileAlign returns -1).
// rescnt[s]: residues (codes other than GAP1/GAP2) of sequence s, which gaps
// inserted by earlier steps do not change.
int prepare_step(const cwhip_msa_params& p, const Seqs& seqs, const std::vector<int>& rescnt,
                  const double* sim, const std::vector<int>& weights, std::vector<int>& group,
                  const std::vector<int>& aligned, Step& st)
{
    const int nseq = (int)seqs.size();
    for (int i = 0; i < nseq; ++i)
        if (aligned[i + 1] == 0) group[i + 1] = 0;
    int n1 = 0, n2 = 0;
    for (int i = 0; i < nseq; ++i) {
        if (group[i + 1] == 1) ++n1;
        else if (group[i + 1] == 2) ++n2;
    }
    if (n1 == 0 || n2 == 0) return 0;
    if (n2 > n1) {                                       // :106-120
        for (int i = 0; i < nseq; ++i) {
            if (group[i + 1] == 1) group[i + 1] = 2;
            else if (group[i + 1] == 2) group[i + 1] = 1;
        }
        std::swap(n1, n2);
    }
    double pcid = 0.0;                                    // :124-137
    int count = 0;
    std::vector<int> in2;                                 // group 2, ascending: the same order of adds
    in2.reserve(n2);
    for (int j = 0; j < nseq; ++j)
        if (group[j + 1] == 2) in2.push_back(j + 1);
    for (int i = 0; i < nseq; ++i) {
        if (group[i + 1] != 1) continue;
        const double* row = sim + (size_t)(i + 1) * (nseq + 1);
        for (int j : in2) pcid += row[j];
        count += (int)in2.size();
    }
    pcid = pcid / (float)count;
    st.members.clear();
    st.w.clear();
    int L1 = 0, L2 = 0, len1 = 0, len2 = 0;               // :205-240 residues per row
    for (int g = 1; g <= 2; ++g)
        for (int i = 0; i < nseq; ++i)
            if (group[i + 1] == g) {
                st.members.push_back(i);
                st.w.push_back(weights[i]);
                (g == 1 ? L1 : L2) = std::max(g == 1 ? L1 : L2, (int)seqs[i].size());
                (g == 1 ? len1 : len2) += rescnt[i];
            }
    st.n1 = n1;
    st.n2 = n2;
    st.L1 = L1;
    st.L2 = L2;
    len1 = (int)(len1 / (float)n1);
    len2 = (int)(len2 / (float)n2);
    const int minLen = std::min(len1, len2);
    int mat[NUMRES][NUMRES];
    float scale = 1.0f;
    int matAvg = 0;
    if (profile_matrix(p, pcid, minLen, mat, &scale, &matAvg) <= 0) return -1;   // :258-263
    st.mat.assign(&mat[0][0], &mat[0][0] + NUMRES * NUMRES);
    int gapcoef, lencoef;                                 // :254-296
    if (p.dna) {
        gapcoef = (int)(100.0 * p.gap_open * scale);
        lencoef = (int)(100.0 * p.gap_extend * scale);
    } else {
        double logmin, logdiff;
        if (len1 == 0 || len2 == 0) {
            logmin = 1.0;
            logdiff = 1.0;
        } else {
            logmin = 1.0 / log10((double)minLen);
            if (len2 < len1) logdiff = 1.0 + 0.5 * log10((double)((float)len2 / (float)len1));
            else if (len1 < len2) logdiff = 1.0 + 0.5 * log10((double)((float)len1 / (float)len2));
            else logdiff = 1.0;
            if (logdiff < 0.9) logdiff = 0.9;
        }
        if (p.neg_matrix) gapcoef = (int)(100.0 * (float)(p.gap_open));
        else if (matAvg <= 0) gapcoef = (int)(100.0 * (float)(p.gap_open + logmin));
        else gapcoef = (int)(scale * matAvg * (float)(p.gap_open / (logdiff * logmin)));
        lencoef = (int)(100.0 * p.gap_extend);
    }
    st.gapcoef = gapcoef;
    st.lencoef = lencoef;
    st.s1.first = 0; st.s1.last = n1; st.s1.L = L1;
    st.s2.first = n1; st.s2.last = n1 + n2; st.s2.L = L2;
    side_spec(p, seqs, st, st.s1, true);
    side_spec(p, seqs, st, st.s2, false);
    st.displ.assign(L1 + L2 + 4, 0);
    st.job.prf1 = st.job.prf2 = nullptr;
    st.job.L1 = L1;
    st.job.L2 = L2;
    st.job.displ = st.displ.data();
    return 1;
}

// The residue-specific gap penalty of each code (calcResidueSpecificPen,
// ProfileBase.cpp:360-391) and the hydrophilic residue set.
void penalty_tables(const cwhip_msa_params& p, BuildArgs& ba)
{
    static const char pres[] = "ACDEFGHKILMNPQRSTVYW";
    static const int pprob[] = {87, 87, 104, 69, 80, 139, 100, 104, 68, 79,
                                71, 137, 126, 93, 128, 124, 111, 75, 100, 77};
    for (int c = 0; c < 32; ++c) ba.pen[c] = 0;
    for (int k = 0; k < 20; ++k) ba.pen[code_of(pres[k])] = 180 - pprob[k];
    ba.hyd_mask = 0;
    for (const char* h = p.hyd_residues ? p.hyd_residues : ""; *h; ++h) {
        const int c = code_of(*h);
        if (c >= 0 && c <= MAXAA) ba.hyd_mask |= 1u << c;
    }
}

// One device batch of steps: the member rows (padded with GAP1 to each
// profile's length), weights, gap ranges and distinct profile matrices packed
// on host threads into page-locked staging and uploaded in two copies; the
// profiles built by cw_prf_cols and cw_prf_gapcols straight into the DP's
// inputs; then (with `dp`) the score and Myers-Miller kernels.  `lo` receives
// the batch layout (the profiles of step k are at lo.probs[k].prf1 / prf2).
int run_built_batch(cwhip_ctx* ctx, const cwhip_msa_params& P, const Seqs& seqs,
                    const std::vector<Step*>& st, bool dp, BatchLayout& lo)
{
    const int n = (int)st.size();
    if (n == 0) return CWHIP_OK;
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<DpJob> jobs(n);
    for (int k = 0; k < n; ++k) jobs[k] = st[k]->job;
    layout_batch(jobs, lo, ctx);
    const bool egp = P.end_gap_penalties != 0;
    for (int k = 0; k < n; ++k)        // nonzero-ness of prf1 GAPCOL at rows 0 and L1 (the top call's go1/go2)
        lo.probs[k].flags = (egp && st[k]->gapcoef != 0 ? 1 : 0) |
                            (egp && (st[k]->L1 >= 1 || st[k]->gapcoef != 0) ? 2 : 0);
    // distinct profile matrices of the batch (a handful of series and scales)
    std::vector<int> mat_of(n);
    std::vector<int> uniq;
    for (int k = 0; k < n; ++k) {
        int u = 0;
        while (u < (int)uniq.size() &&
               std::memcmp(st[uniq[u]]->mat.data(), st[k]->mat.data(), sizeof(int32_t) * NUMRES * NUMRES) != 0)
            ++u;
        if (u == (int)uniq.size()) uniq.push_back(k);
        mat_of[k] = u;
    }
    // uploaded ints after the DP's layout: matrices, then per side the rows'
    // {weight, first, end}; then the builder's per-column sums (device only)
    const int64_t up0 = lo.total;
    int64_t off = up0 + (int64_t)uniq.size() * NUMRES * NUMRES;
    std::vector<PrfSide>& sides = ctx->h_sides;     // kept by the context: its upload is queued
    sides.assign(2 * n, PrfSide{});
    int64_t row_bytes = 0;
    int colblk = 0;
    for (int k = 0; k < n; ++k)
        for (int h = 0; h < 2; ++h) {
            const SideSpec& ss = h ? st[k]->s2 : st[k]->s1;
            PrfSide& sd = sides[2 * k + h];
            sd.n = ss.last - ss.first;
            sd.L = ss.L;
            sd.stride = (int)std::max<int64_t>(4, up4(ss.L));
            sd.flags = ss.flags;
            sd.gdist = ss.gdist;
            sd.gapcoef = st[k]->gapcoef;
            sd.lencoef = st[k]->lencoef;
            sd.sum2 = ss.sum2;
            sd.col0 = colblk;
            sd.pad = 0;
            colblk += (ss.L + CWP_BUILD_COLS - 1) / CWP_BUILD_COLS;
            sd.rows = row_bytes;
            row_bytes += (int64_t)sd.n * sd.stride;
            sd.rinfo = off;
            off += 3 * (int64_t)sd.n;
            sd.mat = up0 + (int64_t)mat_of[k] * NUMRES * NUMRES;
            sd.prf = h ? lo.probs[k].prf2 : lo.probs[k].prf1;
            sd.rec = h ? lo.probs[k].col : lo.probs[k].row;
            sd.pl = h ? lo.probs[k].pl2 : lo.probs[k].pl1;
        }
    const int64_t up_sz = off - up0;
    for (auto& sd : sides) {
        sd.tmp = off;
        off = up4(off + 4 * ((int64_t)sd.L + 2));
    }
    HIPCHK(ctx->h_prf_in.ensure(up_sz + 4));
    HIPCHK(ctx->h_rows.ensure(row_bytes + 4));
    int32_t* h_in = ctx->h_prf_in.p;
    uint8_t* h_rows = ctx->h_rows.p;
    for (size_t u = 0; u < uniq.size(); ++u)
        std::memcpy(h_in + u * NUMRES * NUMRES, st[uniq[u]]->mat.data(), sizeof(int32_t) * NUMRES * NUMRES);
    const bool trim = !P.use_end_gaps && !P.end_gap_penalties;
    parallel_for(n, [&](size_t k) {
        const Step& s = *st[k];
        for (int h = 0; h < 2; ++h) {
            const SideSpec& ss = h ? s.s2 : s.s1;
            const PrfSide& sd = sides[2 * k + h];
            int32_t* ri = h_in + (sd.rinfo - up0);
            for (int q = 0; q < sd.n; ++q) {
                const std::vector<uint8_t>& r = seqs[s.members[ss.first + q]];
                uint8_t* dst = h_rows + sd.rows + (int64_t)q * sd.stride;
                const int len = std::min((int)r.size(), sd.L);
                std::memcpy(dst, r.data(), len);
                std::memset(dst + len, GAP1, sd.stride - len);
                int is = 0, ie = sd.L;                   // calcGapCoeff :129-160
                if (trim) {
                    while (is < sd.L && is_gap_code(dst[is])) ++is;
                    if (is < sd.L)
                        while (ie > 0 && is_gap_code(dst[ie - 1])) --ie;
                    else
                        ie = 0;
                }
                ri[3 * q] = s.w[ss.first + q];
                ri[3 * q + 1] = is;
                ri[3 * q + 2] = ie;
            }
        }
    });
    HIPCHK(ctx->p_pool.ensure(off + 4));
    HIPCHK(ctx->p_rows.ensure(row_bytes + 4));
    HIPCHK(ctx->p_sides.ensure(sides.size()));
    HIPCHK(hipMemcpyAsync(ctx->p_rows.p, h_rows, row_bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->p_pool.p + up0, h_in, up_sz * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->p_sides.p, sides.data(), sides.size() * sizeof(PrfSide), hipMemcpyHostToDevice,
                          ctx->stream));
    BuildArgs ba;
    ba.sides = ctx->p_sides.p;
    ba.rows = ctx->p_rows.p;
    ba.pool = ctx->p_pool.p;
    ba.nside = (int)sides.size();
    ba.ncolblk = colblk;
    penalty_tables(P, ba);
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(cwp_launch_build(colblk, (int)sides.size(), ctx->stream, ba));
    if (!dp) return CWHIP_OK;
    const int rc = dp_launch_collect(ctx, lo, jobs);
    if (rc != CWHIP_OK) return rc;
    for (int k = 0; k < n; ++k) st[k]->job = jobs[k];
    return CWHIP_OK;
}

// progTracepath (:1023-1062), addGGaps (:818-934) and the write-back of :464-493.
// The members' rows are independent: with `par`, a large group's rows are
// rewritten on the host pool.
void apply_step(Step& st, Seqs& seqs, bool par = false)
{
    // the path of each group as runs: > 0 residues (path 2), < 0 gaps (path 1)
    std::vector<int> runs1, runs2;
    auto push = [](std::vector<int>& r, int v) {
        if (!r.empty() && (r.back() > 0) == (v > 0)) r.back() += v;
        else r.push_back(v);
    };
    for (int i = 1; i <= st.job.n_displ; ++i) {
        const int k = st.displ[i];
        if (k == 0) {                                 // one column of both
            push(runs1, 1);
            push(runs2, 1);
        } else {                                      // k > 0: k of group 2 against gaps
            push(runs1, -k);                          // k < 0: -k of group 1 against gaps
            push(runs2, k);
        }
    }
    int alen = 0, res1 = 0, res2 = 0;                 // columns; residue places of each group
    for (int v : runs1) {
        alen += v > 0 ? v : -v;
        res1 += v > 0 ? v : 0;
    }
    for (int v : runs2) res2 += v > 0 ? v : 0;
    // a row's residue ix is src[ix] below its length, GAP1 up to the profile
    // length, ENDALN past it.  The row grows in place from its end: an output
    // position is never below the source position it takes, so a run's source
    // bytes are still there when it is written.
    auto row = [&](size_t kk) {
        const int k = (int)kk;
        const bool g1 = k < st.n1;
        const std::vector<int>& runs = g1 ? runs1 : runs2;
        const int rowlen = g1 ? st.L1 : st.L2;
        std::vector<uint8_t>& v = seqs[st.members[k]];
        const int lim1 = std::min(rowlen, (int)v.size());
        if ((int)v.capacity() < alen) v.reserve((size_t)alen + alen / 4);   // room for later steps
        v.resize(alen);
        uint8_t* o = v.data();
        int op = alen, ix = g1 ? res1 : res2;
        for (size_t q = runs.size(); q-- > 0;) {
            const int r = runs[q];
            if (r < 0) {
                op += r;
                std::memset(o + op, GAP1, -r);
                continue;
            }
            op -= r;
            ix -= r;
            const int n1 = std::max(0, std::min(lim1 - ix, r));
            const int n2 = std::max(0, std::min(rowlen - (ix + n1), r - n1));
            std::memset(o + op + n1 + n2, ENDALN, r - n1 - n2);
            std::memset(o + op + n1, GAP1, n2);
            if (n1 > 0) std::memmove(o + op, o + ix, n1);
        }
    };
    const size_t nrow = (size_t)(st.n1 + st.n2);
    if (par && nrow * (size_t)alen > (1u << 16)) parallel_for(nrow, row);
    else
        for (size_t k = 0; k < nrow; ++k) row(k);
}

// residues (codes other than GAP1 / GAP2) of each row
std::vector<int> residue_counts(const Seqs& seqs)
{
    std::vector<int> r(seqs.size(), 0);
    for (size_t s = 0; s < seqs.size(); ++s)
        for (uint8_t c : seqs[s])
            if (c != GAP1 && c != GAP2) ++r[s];
    return r;
}

} // namespace

extern "C" {

int cwhip_profile_dp(cwhip_ctx* ctx, cwhip_prf_pair* pairs, int npairs, int end_gap_penalties)
{
    if (!ctx || npairs < 0 || (npairs > 0 && !pairs)) return fail(ctx, CWHIP_E_ARG, "bad profile pair arguments");
    std::vector<DpJob> jobs(npairs);
    std::vector<std::vector<int32_t>> displ(npairs);
    for (int k = 0; k < npairs; ++k) {
        const cwhip_prf_pair& q = pairs[k];
        if (!q.prf1 || !q.prf2 || !q.displ || q.len1 < 0 || q.len2 < 0 ||
            q.len1 > CWHIP_MAX_PROFILE_LEN || q.len2 > CWHIP_MAX_PROFILE_LEN)
            return fail(ctx, CWHIP_E_ARG, "bad profile pair");
        displ[k].assign(q.len1 + q.len2 + 4, 0);
        jobs[k].prf1 = q.prf1;
        jobs[k].prf2 = q.prf2;
        jobs[k].L1 = q.len1;
        jobs[k].L2 = q.len2;
        jobs[k].displ = displ[k].data();
    }
    const int rc = run_dp_batch(ctx, jobs, end_gap_penalties ? 1 : 0);
    if (rc != CWHIP_OK) return rc;
    for (int k = 0; k < npairs; ++k) {
        pairs[k].n_displ = jobs[k].n_displ;
        pairs[k].score = jobs[k].score;
        std::memcpy(pairs[k].displ, displ[k].data() + 1, sizeof(int32_t) * jobs[k].n_displ);
    }
    return CWHIP_OK;
}

int cwhip_profile_align(cwhip_ctx* ctx, const cwhip_msa_params* p, int nseq, const uint8_t* codes,
                        const int64_t* offsets, const int32_t* lens, const double* sim,
                        const int32_t* weights, int32_t* group, const int32_t* aligned,
                        uint8_t* rows_out, int64_t stride, int32_t* info)
{
    if (!ctx || !p || nseq < 1 || !codes || !offsets || !lens || !sim || !weights || !group ||
        !aligned || !info)
        return fail(ctx, CWHIP_E_ARG, "bad profile alignment arguments");
    Seqs seqs(nseq);
    for (int s = 0; s < nseq; ++s) {
        if (lens[s] < 0) return fail(ctx, CWHIP_E_ARG, "negative length");
        seqs[s].assign(codes + offsets[s], codes + offsets[s] + lens[s]);
    }
    std::vector<int> g(group, group + nseq + 1), al(aligned, aligned + nseq + 1);
    std::vector<int> w(weights, weights + nseq);
    Step st;
    for (int k = 0; k < 4; ++k) info[k] = 0;
    const int run = prepare_step(*p, seqs, residue_counts(seqs), sim, w, g, al, st);
    for (int i = 0; i <= nseq; ++i) group[i] = g[i];          // profileAlign mutates group
    if (run < 0) {                                             // the matrix failed (:258-263)
        info[0] = -1;
        return CWHIP_OK;
    }
    if (run == 0) return CWHIP_OK;                             // an empty group: score 0 (:96-99)
    if (rows_out && stride < st.L1 + st.L2 + 2) return fail(ctx, CWHIP_E_ARG, "row stride too small");
    BatchLayout lo;
    const int rc = run_built_batch(ctx, *p, seqs, {&st}, true, lo);
    if (rc != CWHIP_OK) return rc;
    apply_step(st, seqs);
    const int alen = (int)seqs[st.members[0]].size();
    if (rows_out)
        for (int m : st.members) std::memcpy(rows_out + (int64_t)m * stride, seqs[m].data(), alen);
    info[0] = st.job.score / 100;
    info[1] = alen;
    info[2] = st.L1;
    info[3] = st.L2;
    return CWHIP_OK;
}

int cwhip_step_profiles(cwhip_ctx* ctx, const cwhip_msa_params* p, int nseq, const uint8_t* codes,
                        const int64_t* offsets, const int32_t* lens, const double* sim,
                        const int32_t* weights, const int32_t* group, const int32_t* aligned,
                        int32_t* prf1, int64_t cap1, int32_t* prf2, int64_t cap2, int32_t* len12)
{
    if (!ctx || !p || nseq < 1 || !codes || !offsets || !lens || !sim || !weights || !group ||
        !aligned || !len12)
        return fail(ctx, CWHIP_E_ARG, "bad step profile arguments");
    Seqs seqs(nseq);
    for (int s = 0; s < nseq; ++s) {
        if (lens[s] < 0) return fail(ctx, CWHIP_E_ARG, "negative length");
        seqs[s].assign(codes + offsets[s], codes + offsets[s] + lens[s]);
    }
    std::vector<int> g(group, group + nseq + 1), al(aligned, aligned + nseq + 1);
    std::vector<int> w(weights, weights + nseq);
    Step st;
    len12[0] = len12[1] = -1;
    const int run = prepare_step(*p, seqs, residue_counts(seqs), sim, w, g, al, st);
    if (run < 0) return fail(ctx, CWHIP_E_MATRIX, "the profile matrix callback failed");
    if (run == 0) return CWHIP_OK;
    len12[0] = st.L1;
    len12[1] = st.L2;
    const int64_t n1 = (int64_t)(st.L1 + 2) * W, n2 = (int64_t)(st.L2 + 2) * W;
    if (!prf1 || !prf2 || cap1 < n1 || cap2 < n2) return fail(ctx, CWHIP_E_ARG, "profile buffers too small");
    BatchLayout lo;
    const int rc = run_built_batch(ctx, *p, seqs, {&st}, false, lo);
    if (rc != CWHIP_OK) return rc;
    HIPCHK(hipMemcpyAsync(prf1, ctx->p_pool.p + lo.probs[0].prf1, n1 * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(prf2, ctx->p_pool.p + lo.probs[0].prf2, n2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CWHIP_OK;
}

int cwhip_set_msa_tuning(cwhip_ctx* ctx, int wide_rows, int wide_levels, int wide_pairs, int wait_limit)
{
    if (!ctx) return CWHIP_E_ARG;
    if (wide_rows < -1 || wide_levels < -1 || wide_levels == 0 || wide_levels > kWideMaxLevels ||
        wide_pairs < -1 || wait_limit < -1)
        return fail(ctx, CWHIP_E_ARG, "msa tuning value out of range");
    if (wide_rows >= 0) ctx->prf_wide_rows = wide_rows;
    if (wide_levels >= 0) ctx->prf_wide_levels = wide_levels;
    if (wide_pairs >= 0) ctx->prf_wide_pairs = wide_pairs;
    if (wait_limit >= 0) ctx->prf_wide_wait = wait_limit > 0 ? wait_limit : 1 << 20;
    return CWHIP_OK;
}

int cwhip_msa_params_default(int dna, cwhip_msa_params* p)
{
    if (!p) return CWHIP_E_ARG;
    std::memset(p, 0, sizeof *p);
    p->dna = dna ? 1 : 0;
    p->gap_open = dna ? 15.0f : 10.0f;          // DNAGapOpen / AAGapOpen
    p->gap_extend = dna ? 6.66f : 0.2f;         // DNAGapExtend / AAGapExtend
    p->gap_dist = 4;
    p->no_var_penalties = 1;
    p->no_hyd_penalties = 0;
    p->no_pref_penalties = 0;
    p->use_end_gaps = 0;
    p->end_gap_penalties = 0;
    p->transition_weight = 0.5f;
    p->neg_matrix = 0;
    p->no_weights = 0;
    p->divergence_cutoff = 30;
    p->hyd_residues = "GPSNDQEKR";
    p->distance_tree = 1;
    p->profile_matrix = nullptr;
    p->profile_matrix_user = nullptr;
    return CWHIP_OK;
}

}  // extern "C"

extern "C" {

int cwhip_msa(cwhip_ctx* ctx, const cwhip_msa_params* p, int nseq, const uint8_t* codes,
              const int64_t* offsets, const int32_t* lens, const double* sim,
              const int32_t* seq_weight, const int32_t* steps, int nsteps, int32_t* out_order,
              int32_t* aln_len)
{
    if (!ctx || !p || nseq < 0 || nsteps < 0 || !out_order || !aln_len ||
        (nseq > 0 && (!codes || !offsets || !lens || !sim || !seq_weight)) || (nsteps > 0 && !steps))
        return fail(ctx, CWHIP_E_ARG, "bad multiple alignment arguments");
    const auto t0 = std::chrono::steady_clock::now();
    ctx->prf_score_ms = ctx->prf_mm_ms = 0.f;
    ctx->prf_cells = ctx->prf_entries = 0;
    Seqs& seqs = ctx->msa_rows;
    seqs.assign(nseq, {});
    for (int s = 0; s < nseq; ++s) {
        if (lens[s] < 0) return fail(ctx, CWHIP_E_ARG, "negative length");
        seqs[s].assign(codes + offsets[s], codes + offsets[s] + lens[s]);
    }
    const cwhip_msa_params& P = *p;
    auto SIM = [&](int a, int b) { return sim[(size_t)a * (nseq + 1) + b]; };
    std::vector<int> weights(seq_weight, seq_weight + nseq);
    const std::vector<int> rescnt = residue_counts(seqs);
    std::vector<int> aligned(nseq + 1, 0), maxid(nseq + 1, 0);
    parallel_for(nseq, [&](size_t k) {                    // MSA.cpp:67-78
        const int i = (int)k + 1;
        int m = -1;
        for (int j = 1; j <= nseq; ++j)
            if (j != i && m < SIM(i, j)) m = (int)SIM(i, j);
        maxid[i] = m;
    });
    std::vector<char> pass(nseq + 1, 0);                  // maxid above the cut-off
    for (int i = 1; i <= nseq; ++i) pass[i] = maxid[i] > P.divergence_cutoff;
    int ix = 0, calls = 0, batches = 0;
    // First pass (:80-157).  A step only touches its members whose maxid passes
    // the cut-off, and a member's rows change only in steps that contain it,
    // so a step can run as soon as the last earlier step sharing a member is
    // done: steps are grouped into dependency levels, each one device batch.
    std::vector<int> level(nsteps, -1), last(nseq + 1, -1);
    std::vector<std::vector<int>> members(nsteps);       // of each step, those past the cut-off
    int nlev = 0;
    for (int s = 0; s < nsteps; ++s) {
        const int32_t* row = steps + (size_t)s * (nseq + 1);
        std::vector<int>& mem = members[s];
        int lv = 0;
        for (int i = 1; i <= nseq; ++i) {
            if (row[i] == 0 || !pass[i]) continue;
            mem.push_back(i);
            if (aligned[i] == 0) {
                out_order[ix++] = i;                      // outputOrder ALIGNED
                aligned[i] = 1;
            }
            if (last[i] >= 0) lv = std::max(lv, level[last[i]] + 1);
        }
        if (mem.empty()) continue;
        level[s] = lv;
        for (int i : mem) last[i] = s;
        nlev = std::max(nlev, lv + 1);
    }
    std::vector<std::vector<int>> by_level(nlev);
    for (int s = 0; s < nsteps; ++s)
        if (level[s] >= 0) by_level[level[s]].push_back(s);
    std::vector<int> all_aligned(nseq + 1, 1);
    for (int lv = 0; lv < nlev; ++lv) {
        const size_t nst = by_level[lv].size();
        std::vector<Step> sts(nst);
        std::vector<signed char> ok(nst, 0);
        // the steps of a level touch disjoint sequences: build their profiles
        // (and later apply their edit scripts) on several host threads; a level
        // of few steps builds them in turn, each on all threads
        auto prep = [&](size_t k) {
            // the step's row of the group matrix, the members that fail the
            // cut-off left out (their aligned flag at this step is that test)
            const int s = by_level[lv][k];
            const int32_t* row = steps + (size_t)s * (nseq + 1);
            std::vector<int> grp(nseq + 1, 0);
            grp[0] = row[0];
            for (int i : members[s]) grp[i] = row[i];
            ok[k] = (signed char)prepare_step(P, seqs, rescnt, sim, weights, grp, all_aligned, sts[k]);
        };
        if (nst >= 4) parallel_for(nst, prep);
        else
            for (size_t k = 0; k < nst; ++k) prep(k);
        for (size_t k = 0; k < nst; ++k)              // MSA.cpp:147-152: a failed step ends the run
            if (ok[k] < 0) return fail(ctx, CWHIP_E_MATRIX, "the profile matrix callback failed");
        std::vector<Step*> live;
        for (size_t k = 0; k < nst; ++k)
            if (ok[k]) live.push_back(&sts[k]);
        BatchLayout lo;
        const int rc = run_built_batch(ctx, P, seqs, live, true, lo);
        if (rc != CWHIP_OK) return rc;
        ++batches;
        if (live.size() >= 4) parallel_for(live.size(), [&](size_t q) { apply_step(*live[q], seqs); });
        else
            for (Step* st : live) apply_step(*st, seqs, true);
        calls += (int)live.size();
    }
    // Second pass: the delayed (divergent) sequences one by one (:159-335).
    std::vector<int> treeWeight(seq_weight, seq_weight + nseq), group(nseq + 1, 0);
    if (ix == 0) {
        int mx = -1, iseq = 0;
        for (int i = 1; i <= nseq; ++i)
            for (int j = i + 1; j <= nseq; ++j)
                if (mx < SIM(i, j)) {
                    mx = (int)SIM(i, j);
                    iseq = i;
                }
        aligned[iseq] = 1;
        out_order[ix++] = iseq;
    }
    while (ix < nseq) {
        for (int i = 1; i <= nseq; ++i) {
            if (aligned[i] != 0) continue;
            maxid[i] = -1;
            for (int j = 1; j <= nseq; ++j)
                if (maxid[i] < SIM(i, j) && aligned[j] != 0) maxid[i] = (int)SIM(i, j);
        }
        int mx = -1, iseq = 0;
        for (int i = 1; i <= nseq; ++i)
            if (aligned[i] == 0 && maxid[i] > mx) {
                mx = maxid[i];
                iseq = i;
            }
        if (!P.no_weights) {
            for (int j = 0; j < nseq; ++j)
                if (aligned[j + 1] != 0) weights[j] = (int)(treeWeight[j] * SIM(j + 1, iseq));
            int sum = 0;
            for (int j = 0; j < nseq; ++j)
                if (aligned[j + 1] != 0) sum += weights[j];
            if (sum == 0) {
                for (int j = 0; j < nseq; ++j) weights[j] = 1;
                sum = nseq;
            }
            for (int j = 0; j < nseq; ++j)
                if (aligned[j + 1] != 0) {
                    weights[j] = (weights[j] * INT_SCALE) / sum;
                    if (weights[j] < 1) weights[j] = 1;
                }
        }
        for (int j = 1; j <= nseq; ++j) {
            if (aligned[j] != 0) group[j] = 1;
            else if (iseq == j) group[j] = 2;
        }
        aligned[iseq] = 1;
        Step st;
        const int run = prepare_step(P, seqs, rescnt, sim, weights, group, aligned, st);
        if (run < 0) return fail(ctx, CWHIP_E_MATRIX, "the profile matrix callback failed");
        if (run > 0) {
            BatchLayout lo;
            const int rc = run_built_batch(ctx, P, seqs, {&st}, true, lo);
            if (rc != CWHIP_OK) return rc;
            ++batches;
            apply_step(st, seqs, true);
            ++calls;
        }
        out_order[ix++] = iseq;
    }
    int alen = 0;
    for (int s = 0; s < nseq; ++s) alen = std::max(alen, (int)seqs[s].size());
    *aln_len = alen;
    const double host_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->msa_stats[0] = calls;
    ctx->msa_stats[1] = batches;
    ctx->msa_stats[2] = ctx->prf_cells;
    ctx->msa_stats[3] = host_ms;
    ctx->msa_stats[4] = ctx->prf_score_ms + ctx->prf_mm_ms;
    ctx->msa_stats[5] = ctx->prf_entries;
    return CWHIP_OK;
}

int cwhip_msa_rows(cwhip_ctx* ctx, uint8_t* out, int64_t stride)
{
    if (!ctx || !out) return fail(ctx, CWHIP_E_ARG, "bad row buffer");
    for (size_t s = 0; s < ctx->msa_rows.size(); ++s) {
        const std::vector<uint8_t>& r = ctx->msa_rows[s];
        if ((int64_t)r.size() > stride) return fail(ctx, CWHIP_E_ARG, "row stride too small");
        std::memcpy(out + (int64_t)s * stride, r.data(), r.size());
    }
    return CWHIP_OK;
}

int cwhip_msa_stats(cwhip_ctx* ctx, double* out6)
{
    if (!ctx || !out6) return CWHIP_E_ARG;
    for (int k = 0; k < 6; ++k) out6[k] = ctx->msa_stats[k];
    return CWHIP_OK;
}

}  // extern "C"
