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

// One device batch of profile pairs: inputs packed (on host threads) into
// page-locked staging and uploaded in one copy, the score and Myers-Miller
// kernels, the node tables (contiguous in the pool) fetched in one copy and
// turned into edit scripts on host threads.
int run_dp_batch(cwhip_ctx* ctx, std::vector<DpJob>& jobs, int egp)
{
    const int n = (int)jobs.size();
    if (n == 0) return CWHIP_OK;
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<PrfProb> probs(n);
    std::vector<int32_t> tile_base(n + 1, 0);
    int64_t in_sz = 0, work_sz = 0;
    for (int k = 0; k < n; ++k) {
        const int L1 = jobs[k].L1, L2 = jobs[k].L2;
        PrfProb& pp = probs[k];
        pp.L1 = L1;
        pp.L2 = L2;
        pp.ldS = L1 + 2;
        pp.flags = (jobs[k].prf1[GAPCOL] != 0 ? 1 : 0) |
                   (jobs[k].prf1[(int64_t)L1 * W + GAPCOL] != 0 ? 2 : 0);
        pp.node_cap = 2 * L1 + 4;
        pp.rowcap = (int)up4(L2 + 2);
        pp.prf1 = in_sz; in_sz = up4(in_sz + (int64_t)(L1 + 2) * W);
        pp.prf2 = in_sz; in_sz = up4(in_sz + (int64_t)(L2 + 2) * W);
        pp.row = in_sz; in_sz += 4 * (int64_t)(L1 + 2);
        pp.col = in_sz; in_sz += 4 * (int64_t)(L2 + 2);
        pp.pl1 = in_sz; in_sz = up4(in_sz + L1 + 2);
        pp.pl2 = in_sz; in_sz = up4(in_sz + L2 + 2);
        tile_base[k + 1] = tile_base[k] + ((L1 + 63) / 64) * ((L2 + 63) / 64);
    }
    // work area: node tables of all pairs first (one copy back), then per pair
    // the score matrix, per-wave scratch rows and level buffers
    for (int k = 0; k < n; ++k) {
        probs[k].nodes = work_sz;
        work_sz = up4(work_sz + (int64_t)probs[k].node_cap * CWP_NODE_INTS);
    }
    const int64_t nodes_sz = work_sz;
    for (int k = 0; k < n; ++k) {
        PrfProb& pp = probs[k];
        pp.S = work_sz; work_sz = up4(work_sz + (int64_t)(pp.L2 + 2) * pp.ldS);
        pp.scratch = work_sz; work_sz += (int64_t)CWP_WAVES * CWP_SCR_ROWS * pp.rowcap;
        pp.lvl = work_sz; work_sz = up4(work_sz + 4 * ((int64_t)pp.L2 + 2 + pp.node_cap));
    }
    for (auto& pp : probs) { pp.S += in_sz; pp.nodes += in_sz; pp.scratch += in_sz; pp.lvl += in_sz; }
    // host images of the inputs: raw profiles, {gap, len, mask, gap+len}
    // records and LENCOL prefix sums (the mask is 0 at the profile ends when
    // end-gap penalties are off, the rule of openPenalty1/2 and extPenalty1/2)
    HIPCHK(ctx->h_prf_in.ensure(in_sz + 4));
    int32_t* h_in = ctx->h_prf_in.p;
    parallel_for(n, [&](size_t k) {
        const PrfProb& pp = probs[k];
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
    HIPCHK(ctx->p_pool.ensure(in_sz + work_sz + 4));
    HIPCHK(ctx->p_probs.ensure(n));
    HIPCHK(ctx->p_out.ensure(4 * (size_t)n));
    HIPCHK(ctx->p_tiles.ensure(n + 1));
    HIPCHK(hipMemcpyAsync(ctx->p_pool.p, h_in, in_sz * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->p_probs.p, probs.data(), n * sizeof(PrfProb), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->p_tiles.p, tile_base.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    PrfArgs a;
    a.probs = ctx->p_probs.p;
    a.pool = ctx->p_pool.p;
    a.out = ctx->p_out.p;
    a.tile_base = ctx->p_tiles.p;
    a.nprob = n;
    a.rmax = ctx->prf_rmax;
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    if (tile_base[n] > 0) HIPCHK(cwp_launch_score(tile_base[n], ctx->stream, a));
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(cwp_launch_mm(n, ctx->stream, a));
    HIPCHK(hipEventRecord(ctx->ev2, ctx->stream));
    HIPCHK(ctx->h_prf_nodes.ensure(nodes_sz + 4 * (size_t)n + 4));
    int32_t* h_nodes = ctx->h_prf_nodes.p;
    int32_t* out = h_nodes + nodes_sz;
    HIPCHK(hipMemcpyAsync(out, ctx->p_out.p, 4 * (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(h_nodes, ctx->p_pool.p + in_sz, nodes_sz * sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms1 = 0, ms2 = 0;
    HIPCHK(hipEventElapsedTime(&ms1, ctx->ev0, ctx->ev1));
    HIPCHK(hipEventElapsedTime(&ms2, ctx->ev1, ctx->ev2));
    ctx->prf_score_ms += ms1;
    ctx->prf_mm_ms += ms2;
    for (int k = 0; k < n; ++k) {
        const int nn = out[4 * k + 1];
        if (out[4 * k + 2] != 0 || nn < 1 || nn > probs[k].node_cap)
            return fail(ctx, CWHIP_E_KERNEL, "profile DP node table overflow");
    }
    std::vector<char> bad(n, 0);
    std::vector<double> cells(n, 0.0);
    parallel_for(n, [&](size_t k) {
        Emitter em{jobs[k].displ};
        if (!emit_script(h_nodes + (probs[k].nodes - in_sz), out[4 * k + 1], em, &cells[k])) {
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

// SubMatrix::getProfileAlignMatrix (SubMatrix.cpp:403-677), default matrix
// names ("gonnet" series, "iub").
int profile_matrix(const cwhip_msa_params& p, double pcid, int minLen, int mat[NUMRES][NUMRES],
                   float* scale_out, int* mat_avg)
{
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
        if (p.neg_matrix) {
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

// One profile: rows[first..last) of the step, prfLength L.
struct ProfileSide {
    int first, last, L;
    std::vector<int32_t> prf;   // (L + 2) x W
    bool nVarPen = true, nHydPen = false, nPrefPen = false;
    int gdist = 0;
    int32_t& at(int i, int j) { return prf[(size_t)i * W + j]; }
};

// Per column of rows[first..last): the weighted residue counts and the plain
// counts of every code, gathered in one row-major pass (on host threads over
// column blocks when `par` and the group is large).  Integer sums, so the
// order of the additions does not matter.
struct ColHist {
    std::vector<int> wsum, cnt;   // [col * NUMRES + code]
};

void column_hist(const std::vector<const uint8_t*>& rows, const std::vector<int>& w, int first,
                 int last, int L, bool par, ColHist& h)
{
    h.wsum.assign((size_t)L * NUMRES, 0);
    h.cnt.assign((size_t)L * NUMRES, 0);
    constexpr int blk = 256;
    const size_t nb = (size_t)(L + blk - 1) / blk;
    auto body = [&](size_t b) {
        const int c0 = (int)b * blk, c1 = std::min(L, c0 + blk);
        for (int k = first; k < last; ++k) {
            const uint8_t* r = rows[k];
            const int wk = w[k];
            for (int col = c0; col < c1; ++col) {
                const int c = r[col];
                if (c < NUMRES) {
                    h.wsum[(size_t)col * NUMRES + c] += wk;
                    h.cnt[(size_t)col * NUMRES + c] += 1;
                }
            }
        }
    };
    if (par && (int64_t)(last - first) * L > (1 << 18)) parallel_for(nb, body);
    else
        for (size_t b = 0; b < nb; ++b) body(b);
}

// ProfileBase::calcGapCoeff (ProfileBase.cpp:52-298) and its helpers
// (:310-549).  `rows` are the step's padded rows; `gaps` is the per-column
// count the caller shares with calcProfileWithSub.
// sum over rows first..last-1 of per-row contributions to an L-long integer
// array: row chunks on the host pool (when `par`), each with its own array,
// added up at the end (integer sums: the order does not matter)
template <typename RowFn>
void rows_sum(int first, int last, int L, bool par, std::vector<int>& out, RowFn row_fn)
{
    const int n = last - first;
    const int nch = (par && (int64_t)n * L > (1 << 17)) ? std::min(16, std::max(1, n / 16)) : 1;
    if (nch == 1) {
        for (int k = first; k < last; ++k) row_fn(k, out.data());
        return;
    }
    std::vector<std::vector<int>> acc(nch, std::vector<int>(L, 0));
    parallel_for(nch, [&](size_t c) {
        const int a = first + (int)((int64_t)n * c / nch), b = first + (int)((int64_t)n * (c + 1) / nch);
        for (int k = a; k < b; ++k) row_fn(k, acc[c].data());
    });
    for (int c = 0; c < nch; ++c)
        for (int j = 0; j < L; ++j) out[j] += acc[c][j];
}

void calc_gap_coeff(const cwhip_msa_params& p, ProfileSide& s, const std::vector<const uint8_t*>& rows,
                    const ColHist& hist, std::vector<int>& gaps, int gapCoef, int lenCoef, bool par)
{
    const int L = s.L, first = s.first, last = s.last, numSeq = last - first;
    int pcid = 0;
    if (numSeq == 2) {                                   // percentId (:519-547)
        const uint8_t* s1 = rows[first];
        const uint8_t* s2 = rows[first + 1];
        int count = 0, total = 0;
        for (int i = 0; i < L; ++i) {
            if (s1[i] < MAXAA) {                          // (>= 0 always holds for codes)
                ++total;
                if (s1[i] == s2[i]) ++count;
            }
        }
        pcid = total == 0 ? 0 : (int)(float)(100.0 * (float)count / (float)total);
    }
    for (int j = 0; j < L; ++j) gaps[j] = 0;
    if (!p.no_var_penalties && pcid > 60) {
        s.nHydPen = s.nPrefPen = true;
        s.nVarPen = false;
    } else {
        s.nVarPen = true;
        s.nHydPen = p.no_hyd_penalties != 0;
        s.nPrefPen = p.no_pref_penalties != 0;
        s.gdist = p.gap_dist;
    }
    rows_sum(first, last, L, par, gaps, [&](int i, int* acc) {
        const uint8_t* r = rows[i];
        int is = 0, ie = L;
        if (!p.use_end_gaps && !p.end_gap_penalties) {
            while (is < L && is_gap_code(r[is])) ++is;
            if (is < L)
                while (ie > 0 && is_gap_code(r[ie - 1])) --ie;
            else
                ie = 0;
        }
        for (int j = is; j < ie; ++j)
            if (is_gap_code(r[j])) acc[j]++;
    });
    std::vector<int> vW, resW, hydW;
    if (!p.dna && !s.nVarPen) {                          // calcVPenalties (:310-353)
        vW.assign(L + 2, 0);
        const uint8_t* a = rows[first];
        const uint8_t* b = rows[first + 1];
        for (int i = 0; i < L; ++i) {
            int t = 0, w = 0;
            for (int j = i - 5; j < i + 5; ++j) {
                if (j < 0 || j >= L) continue;
                if (is_gap_code(a[j]) || is_gap_code(b[j])) continue;
                w += (a[j] == b[j]) ? 1 : 0;
                ++t;
            }
            w += t;
            w = t > 0 ? (w * 100) / (2 * t) : 100;
            vW[i] = w < 50 ? 50 : w;
        }
    }
    if (!p.dna && !s.nPrefPen) {                         // calcResidueSpecificPen (:360-391)
        static const char pres[] = "ACDEFGHKILMNPQRSTVYW";
        static const int pprob[] = {87, 87, 104, 69, 80, 139, 100, 104, 68, 79,
                                    71, 137, 126, 93, 128, 124, 111, 75, 100, 77};
        int pen[NUMRES];
        for (int c = 0; c < NUMRES; ++c) pen[c] = 0;
        for (int k = 0; k < 20; ++k) pen[code_of(pres[k])] = 180 - pprob[k];
        resW.assign(L + 2, 0);
        for (int i = 0; i < L; ++i) {
            int w = 0;
            for (int c = 0; c <= MAXAA; ++c) w += hist.cnt[(size_t)i * NUMRES + c] * pen[c];
            resW[i] = w / numSeq;
        }
    }
    if (!p.dna && !s.nHydPen) {                          // calcHydrophilicPen (:398-459)
        bool hyd[NUMRES];
        for (int c = 0; c < NUMRES; ++c) hyd[c] = false;
        for (const char* h = p.hyd_residues ? p.hyd_residues : ""; *h; ++h) {
            const int c = code_of(*h);
            if (c >= 0 && c <= MAXAA) hyd[c] = true;
        }
        hydW.assign(L + 2, 0);
        rows_sum(first, last, L, par, hydW, [&](int k, int* acc) {
            const uint8_t* r = rows[k];
            int i = 0;
            while (i < L) {
                if (is_gap_code(r[i]) || !hyd[r[i]]) { ++i; continue; }
                const int st = i;
                while (i < L && !is_gap_code(r[i]) && hyd[r[i]]) ++i;
                if (i - st > 3)
                    for (int j = st; j < i; ++j) acc[j] += 100;
            }
        });
        const float scale = (float)numSeq;
        for (int i = 0; i < L; ++i) hydW[i] = (int)(hydW[i] / scale);
    }
    std::vector<int> gapPos(L + 2, 0);
    if (p.dna || s.gdist <= 0) {
        for (int i = 0; i < L; ++i) gapPos[i] = gaps[i];
    } else {
        const int gdist = s.gdist;
        auto gapAt = [&](int x) { return x < (int)gaps.size() ? gaps[x] : 0; };
        int i = 0, j;
        while (i < L) {
            if (gaps[i] <= 0) {
                gapPos[i] = gaps[i];
                ++i;
            } else {
                for (j = -gdist + 1; j < 0; ++j)
                    if (i + j >= 0 && i + j < L && (gaps[i + j] == 0 || gaps[i + j] < j)) gapPos[i + j] = j;
                while (gapAt(i) > 0) {
                    if (i >= L) break;
                    gapPos[i] = gaps[i];
                    ++i;
                }
                for (j = 0; j < gdist; ++j) {
                    if (gapAt(i + j) > 0) break;
                    if (i + j >= 0 && i + j < L && (gaps[i + j] == 0 || gaps[i + j] < -j)) gapPos[i + j] = -j - 1;
                }
                i += j;
            }
        }
    }
    for (int j = 0; j < L; ++j) {
        int32_t& g = s.at(j + 1, GAPCOL);
        int32_t& l = s.at(j + 1, LENCOL);
        if (gapPos[j] <= 0) {
            if (!p.dna) {                                  // localPenalty (:468-512)
                bool h = false;
                float gw = 1.0f;
                if (!s.nVarPen) gw *= vW[j] / 100.0;
                if (!s.nHydPen && hydW[j] > 0) {
                    gw *= 0.5;
                    h = true;
                }
                if (!s.nPrefPen && !h) gw *= (resW[j] / 100.0);
                gw *= gapCoef;
                g = (int)gw;
            } else {
                g = gapCoef;
            }
            l = lenCoef;
            if (gapPos[j] < 0) g = (int)((g * (2.0 + 2.0 * (s.gdist + gapPos[j]) / s.gdist)));
        } else {
            const float scale = ((float)(numSeq - gaps[j]) / (float)numSeq) * 1.0f;
            g = (int)(scale * gapCoef);
            l = (int)(0.5 * lenCoef);
        }
        if (g <= 0) g = 1;
        if (l <= 0) l = 1;
    }
    if (p.end_gap_penalties) {
        s.at(0, GAPCOL) = gapCoef;
        s.at(0, LENCOL) = lenCoef;
    } else {
        s.at(0, GAPCOL) = 0;
        s.at(0, LENCOL) = 0;
        s.at(L, GAPCOL) = 0;
        s.at(L, LENCOL) = 0;
    }
}

// ProfileWithSub::calcProfileWithSub (ProfileWithSub.cpp:42-147)
void calc_profile_with_sub(ProfileSide& s, const ColHist& hist, const std::vector<int>& gaps,
                           const int mat[NUMRES][NUMRES], const std::vector<int>& w, bool par)
{
    const int numSeq = s.last - s.first;
    int sum2 = 0;
    for (int k = s.first; k < s.last; ++k) sum2 += w[k];
    auto column = [&](int pos) {
        if (gaps[pos] == numSeq) {
            for (int res = 0; res <= MAXAA; ++res) s.at(pos + 1, res) = mat[res][GAP1];
            s.at(pos + 1, GAP1) = mat[GAP1][GAP1];
            s.at(pos + 1, GAP2) = mat[GAP2][GAP1];
            return;
        }
        const int* wt = &hist.wsum[(size_t)pos * NUMRES];
        const float scale = (float)(numSeq - gaps[pos]) / (float)numSeq;
        // the weighted sum over the codes present in the column (integer, so
        // the order and the zero terms the reference also adds do not matter)
        int present[NUMRES], np = 0;
        for (int aa = 0; aa <= MAXAA; ++aa)
            if (wt[aa]) present[np++] = aa;
        if (wt[GAP1]) present[np++] = GAP1;
        if (wt[GAP2]) present[np++] = GAP2;
        auto one = [&](int res) {
            int f = 0;
            for (int k = 0; k < np; ++k) f += wt[present[k]] * mat[present[k]][res];
            return (int)(((float)f / (float)sum2) * scale);
        };
        for (int res = 0; res <= MAXAA; ++res) s.at(pos + 1, res) = one(res);
        s.at(pos + 1, GAP1) = one(GAP1);
        s.at(pos + 1, GAP2) = one(GAP2);
    };
    constexpr int blk = 128;
    const size_t nb = (size_t)(s.L + blk - 1) / blk;
    auto body = [&](size_t b) {
        for (int pos = (int)b * blk; pos < std::min(s.L, (int)(b + 1) * blk); ++pos) column(pos);
    };
    if (par && (int64_t)numSeq * s.L > (1 << 16)) parallel_for(nb, body);
    else
        for (size_t b = 0; b < nb; ++b) body(b);
}

// ProfileStandard::calcStandardProfile (ProfileStandard.cpp:41-117)
void calc_standard_profile(ProfileSide& s, const ColHist& hist, const std::vector<int>& w)
{
    int sum2 = 0;
    for (int k = s.first; k < s.last; ++k) sum2 += w[k];
    for (int r = 0; r < s.L; ++r) {
        if (sum2 == 0) {
            for (int d = 0; d <= MAXAA; ++d) s.at(r + 1, d) = 0;
            s.at(r + 1, GAP1) = 0;
            s.at(r + 1, GAP2) = 0;
            continue;
        }
        const int* wt = &hist.wsum[(size_t)r * NUMRES];
        for (int d = 0; d <= MAXAA; ++d) s.at(r + 1, d) = (int)(10 * (float)wt[d] / (float)sum2);
        s.at(r + 1, GAP1) = (int)(10 * (float)wt[GAP1] / (float)sum2);
        s.at(r + 1, GAP2) = (int)(10 * (float)wt[GAP2] / (float)sum2);
    }
}

// One profileAlign call (MyersMillerProfileAlign.cpp:36-505) split into the
// part before the DP (prepare) and the part after it (apply).
struct Step {
    std::vector<int> members;          // group-1 sequences, then group-2 ones
    int n1 = 0, n2 = 0, L1 = 0, L2 = 0;
    std::vector<uint8_t> padded;       // (n1 + n2) padded rows
    std::vector<const uint8_t*> rows;
    ProfileSide p1, p2;
    std::vector<int32_t> displ;
    DpJob job;
};

using Seqs = std::vector<std::vector<uint8_t>>;

// Returns false when a group is empty (profileAlign returns 0, :96-99).
bool prepare_step(const cwhip_msa_params& p, const Seqs& seqs, const double* sim,
                  const std::vector<int>& weights, std::vector<int>& group,
                  const std::vector<int>& aligned, Step& st, bool par)
{
    const int nseq = (int)seqs.size();
    for (int i = 0; i < nseq; ++i)
        if (aligned[i + 1] == 0) group[i + 1] = 0;
    int n1 = 0, n2 = 0;
    for (int i = 0; i < nseq; ++i) {
        if (group[i + 1] == 1) ++n1;
        else if (group[i + 1] == 2) ++n2;
    }
    if (n1 == 0 || n2 == 0) return false;
    if (n2 > n1) {                                       // :106-120
        for (int i = 0; i < nseq; ++i) {
            if (group[i + 1] == 1) group[i + 1] = 2;
            else if (group[i + 1] == 2) group[i + 1] = 1;
        }
        std::swap(n1, n2);
    }
    double pcid = 0.0;                                    // :124-137
    int count = 0;
    for (int i = 0; i < nseq; ++i) {
        if (group[i + 1] != 1) continue;
        const double* row = sim + (size_t)(i + 1) * (nseq + 1);
        for (int j = 0; j < nseq; ++j)
            if (group[j + 1] == 2) {
                ++count;
                pcid += row[j + 1];
            }
    }
    pcid = pcid / (float)count;
    st.members.clear();
    int L1 = 0, L2 = 0;
    for (int g = 1; g <= 2; ++g)
        for (int i = 0; i < nseq; ++i)
            if (group[i + 1] == g) {
                st.members.push_back(i);
                int& L = g == 1 ? L1 : L2;
                L = std::max(L, (int)seqs[i].size());
            }
    st.n1 = n1;
    st.n2 = n2;
    st.L1 = L1;
    st.L2 = L2;
    const int Lmax = std::max(L1, L2);
    const int stride = Lmax + 2;
    st.padded.assign((size_t)(n1 + n2) * stride, (uint8_t)GAP1);
    st.rows.resize(n1 + n2);
    std::vector<int> w(n1 + n2);
    int len1 = 0, len2 = 0;                               // :205-240 residues per row
    for (int k = 0; k < n1 + n2; ++k) {
        const std::vector<uint8_t>& s = seqs[st.members[k]];
        uint8_t* r = &st.padded[(size_t)k * stride];
        std::memcpy(r, s.data(), s.size());
        st.rows[k] = r;
        w[k] = weights[st.members[k]];
        int is = 0;
        for (uint8_t c : s)
            if (c != GAP1 && c != GAP2) ++is;
        (k < n1 ? len1 : len2) += is;
    }
    len1 = (int)(len1 / (float)n1);
    len2 = (int)(len2 / (float)n2);
    const int minLen = std::min(len1, len2);
    int mat[NUMRES][NUMRES];
    float scale = 1.0f;
    int matAvg = 0;
    profile_matrix(p, pcid, minLen, mat, &scale, &matAvg);
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
    std::vector<int> gaps(L1 + L2 + 3, 0);
    st.p1.first = 0; st.p1.last = n1; st.p1.L = L1;
    st.p1.prf.assign((size_t)(L1 + 2) * W, 0);
    st.p2.first = n1; st.p2.last = n1 + n2; st.p2.L = L2;
    st.p2.prf.assign((size_t)(L2 + 2) * W, 0);
    ColHist hist;
    column_hist(st.rows, w, 0, n1, L1, par, hist);
    calc_gap_coeff(p, st.p1, st.rows, hist, gaps, gapcoef, lencoef, par);
    calc_profile_with_sub(st.p1, hist, gaps, mat, w, par);
    column_hist(st.rows, w, n1, n1 + n2, L2, par, hist);
    calc_gap_coeff(p, st.p2, st.rows, hist, gaps, gapcoef, lencoef, par);
    calc_standard_profile(st.p2, hist, w);
    st.displ.assign(L1 + L2 + 4, 0);
    st.job.prf1 = st.p1.prf.data();
    st.job.prf2 = st.p2.prf.data();
    st.job.L1 = L1;
    st.job.L2 = L2;
    st.job.displ = st.displ.data();
    return true;
}

// progTracepath (:1023-1062), addGGaps (:818-934) and the write-back of :464-493.
void apply_step(Step& st, Seqs& seqs)
{
    std::vector<uint8_t> path1, path2;
    path1.reserve(st.L1 + st.L2 + 2);
    path2.reserve(st.L1 + st.L2 + 2);
    for (int i = 1; i <= st.job.n_displ; ++i) {
        const int k = st.displ[i];
        if (k == 0) {
            path1.push_back(2);
            path2.push_back(2);
        } else if (k > 0) {
            for (int j = 0; j < k; ++j) { path1.push_back(1); path2.push_back(2); }
        } else {
            for (int j = 0; j < -k; ++j) { path1.push_back(2); path2.push_back(1); }
        }
    }
    const int alen = (int)path1.size();
    const int stride = std::max(st.L1, st.L2) + 2;
    for (int k = 0; k < st.n1 + st.n2; ++k) {
        const bool g1 = k < st.n1;
        const std::vector<uint8_t>& path = g1 ? path1 : path2;
        const int rowlen = g1 ? st.L1 : st.L2;
        const uint8_t* src = &st.padded[(size_t)k * stride];
        std::vector<uint8_t> out(alen);
        int ix = 0;
        for (int i = 0; i < alen; ++i) {
            if (path[i] == 2) {
                out[i] = ix < rowlen ? src[ix] : (uint8_t)ENDALN;
                ++ix;
            } else {
                out[i] = GAP1;
            }
        }
        seqs[st.members[k]].swap(out);
    }
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
            q.len1 > CWHIP_MAX_LEN * 4 || q.len2 > CWHIP_MAX_LEN * 4)
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
    const bool run = prepare_step(*p, seqs, sim, w, g, al, st, true);
    for (int i = 0; i <= nseq; ++i) group[i] = g[i];          // profileAlign mutates group
    if (!run) return CWHIP_OK;                                 // an empty group: score 0 (:96-99)
    if (rows_out && stride < st.L1 + st.L2 + 2) return fail(ctx, CWHIP_E_ARG, "row stride too small");
    std::vector<DpJob> jobs{st.job};
    const int rc = run_dp_batch(ctx, jobs, p->end_gap_penalties ? 1 : 0);
    if (rc != CWHIP_OK) return rc;
    st.job = jobs[0];
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
    std::vector<int> aligned(nseq + 1, 0), maxid(nseq + 1, 0);
    for (int i = 1; i <= nseq; ++i) {                     // MSA.cpp:67-78
        maxid[i] = -1;
        for (int j = 1; j <= nseq; ++j)
            if (j != i && maxid[i] < SIM(i, j)) maxid[i] = (int)SIM(i, j);
    }
    int ix = 0, calls = 0, batches = 0;
    // First pass (:80-157).  A step only touches its members whose maxid passes
    // the cut-off, and a member's rows change only in steps that contain it,
    // so a step can run as soon as the last earlier step sharing a member is
    // done: steps are grouped into dependency levels, each one device batch.
    std::vector<int> level(nsteps, -1), last(nseq + 1, -1);
    std::vector<std::vector<int>> groups(nsteps);
    int nlev = 0;
    for (int s = 0; s < nsteps; ++s) {
        const int32_t* row = steps + (size_t)s * (nseq + 1);
        int entries = 0, lv = 0;
        for (int i = 1; i <= nseq; ++i) {
            if (row[i] != 0 && maxid[i] > P.divergence_cutoff) {
                ++entries;
                if (aligned[i] == 0) {
                    out_order[ix++] = i;                  // outputOrder ALIGNED
                    aligned[i] = 1;
                }
                if (last[i] >= 0) lv = std::max(lv, level[last[i]] + 1);
            }
        }
        if (entries == 0) continue;
        level[s] = lv;
        for (int i = 1; i <= nseq; ++i)
            if (row[i] != 0 && maxid[i] > P.divergence_cutoff) last[i] = s;
        nlev = std::max(nlev, lv + 1);
        // the member's aligned flag at this step is its own cut-off test
        groups[s].assign(row, row + nseq + 1);
        for (int i = 1; i <= nseq; ++i)
            if (!(maxid[i] > P.divergence_cutoff)) groups[s][i] = 0;
    }
    std::vector<std::vector<int>> by_level(nlev);
    for (int s = 0; s < nsteps; ++s)
        if (level[s] >= 0) by_level[level[s]].push_back(s);
    std::vector<int> all_aligned(nseq + 1, 1);
    for (int lv = 0; lv < nlev; ++lv) {
        const size_t nst = by_level[lv].size();
        std::vector<Step> sts(nst);
        std::vector<char> ok(nst, 0);
        // the steps of a level touch disjoint sequences: build their profiles
        // (and later apply their edit scripts) on several host threads; a level
        // of few steps builds them in turn, each on all threads
        auto prep = [&](size_t k) {
            std::vector<int> grp = groups[by_level[lv][k]];
            ok[k] = prepare_step(P, seqs, sim, weights, grp, all_aligned, sts[k], nst < 4) ? 1 : 0;
        };
        if (nst >= 4) parallel_for(nst, prep);
        else
            for (size_t k = 0; k < nst; ++k) prep(k);
        std::vector<DpJob> jobs;
        std::vector<int> live;
        for (size_t k = 0; k < nst; ++k)
            if (ok[k]) {
                live.push_back((int)k);
                jobs.push_back(sts[k].job);
            }
        const int rc = run_dp_batch(ctx, jobs, P.end_gap_penalties ? 1 : 0);
        if (rc != CWHIP_OK) return rc;
        ++batches;
        for (size_t q = 0; q < live.size(); ++q) sts[live[q]].job = jobs[q];
        parallel_for(live.size(), [&](size_t q) { apply_step(sts[live[q]], seqs); });
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
        if (prepare_step(P, seqs, sim, weights, group, aligned, st, true)) {
            std::vector<DpJob> jobs{st.job};
            const int rc = run_dp_batch(ctx, jobs, P.end_gap_penalties ? 1 : 0);
            if (rc != CWHIP_OK) return rc;
            ++batches;
            st.job = jobs[0];
            apply_step(st, seqs);
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
