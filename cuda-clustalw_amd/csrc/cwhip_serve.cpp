// cwhip_serve.cpp -- a request loop over the library for distance-matrix and
// full-alignment jobs of different sizes arriving over time.
//
// One device context per slot, each with its own stream and worker thread.
// Admission is by device memory, counted as it really is: an idle slot is
// charged the device bytes its context still holds from earlier jobs (buffers
// are kept for reuse), a running slot the reservation its job was admitted
// with, and the context of a running job cannot allocate beyond that
// reservation (its buffers refuse to grow past it).  The sum of the charges
// never exceeds the HBM budget, so neither do the bytes actually held.  Jobs
// start in arrival order: the oldest pending job starts as soon as a slot is
// free and its reservation fits (idle slots' buffers are released if that
// makes it fit); a later job never overtakes it, so a large job cannot starve.
// A job whose estimate turns out too small is run again with twice the
// reservation (up to the budget).  Jobs on different slots run on different
// streams, so their kernels overlap on the device.  A job's sequences are
// uploaded once to its slot and stay resident for all of its stages; an
// alignment job's guide tree is joined on the device from the distances its
// pairwise stage left in the slot's memory (only the tree text comes back
// before the progressive stage).  Every
// result is computed by the same entry points a caller uses directly
// (cwhip_pairwise_range, cwhip_nj_tree, cwhip_guide_from_tree, cwhip_msa,
// cwhip_clustal_text), so it is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cwhip_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int64_t id = 0;
    int kind = 0, dna = 0, nseq = 0;
    std::vector<uint8_t> codes;
    std::vector<int64_t> offs;
    std::vector<int32_t> lens;
    std::vector<std::string> names;
    int64_t footprint = 0;            // reservation (bytes); doubled when a run needs more
    int tries = 0;
    double cells = 0;
    // results
    std::vector<float> dist;
    std::string dnd, aln;
    int state = 0;                    // 0 pending, 1 running, 2 done, 3 failed
    int rc = CWHIP_OK;
    std::string err;
    double t_submit = 0, t_start = 0, t_end = 0;
};

// Device bytes a job holds while it runs: sequences, per-pair buffers, the
// batched kernel's per-wave scratch (8 rows of max length per wave, at most
// one wave per pair), and for an alignment the profile-DP pool of its widest
// level (score matrix, records and scratch of up to nseq/2 pairs of profiles
// of about 1.25 x the longest sequence) and the device guide tree's DistMatrix
// and treeDesc (12 bytes per entry, 16 counted).
int64_t footprint(const Job& j, int num_cu)
{
    int64_t total = 0, lmax = 0;
    for (int32_t L : j.lens) { total += L; lmax = std::max<int64_t>(lmax, L); }
    const int64_t n = j.nseq, pairs = n * (n - 1) / 2;
    const int64_t waves = std::min<int64_t>(pairs, (int64_t)num_cu * 16);
    int64_t b = total + 12 * n + 48 * pairs + waves * 8 * (lmax + 66) * 4 + (1 << 20);
    if (j.kind == CWHIP_JOB_ALIGNMENT) {
        const int64_t lp = lmax + lmax / 4 + 2;
        b += (n / 2 + 1) * (lp * lp * 4 + 112 * lp * 4 + 4 * 8 * lp * 4);
        b += (n + 1) * (n + 1) * 16;
    }
    return b;
}

}  // namespace

struct Slot {
    cwhip_ctx* ctx = nullptr;
    Job* job = nullptr;               // assigned job (running), or null (idle)
    int64_t charge = 0;               // idle: bytes its context holds; running: the reservation
};

struct cwhip_server {
    int device = 0;
    int64_t budget = 0, peak = 0;
    int num_cu = 0, running = 0, max_running = 0;
    std::vector<Slot> slots;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> pending;
    std::unordered_map<int64_t, std::unique_ptr<Job>> jobs;
    int64_t next_id = 1;
    bool stop = false;
    double done = 0, cells = 0;
    std::atomic<int64_t> real_total{0}, real_peak{0};   // device bytes the slots' contexts hold
    Clock::time_point t0;
    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    int64_t charged() const
    {
        int64_t c = 0;
        for (const Slot& sl : slots) c += sl.charge;
        return c;
    }
};

namespace {

int run_job(cwhip_ctx* ctx, Job& j)
{
    const int n = j.nseq;
    int rc = cwhip_load_sequences(ctx, j.codes.data(), j.offs.data(), j.lens.data(), n);
    if (rc) return rc;
    int lmax = 1;
    for (int32_t L : j.lens) lmax = std::max(lmax, (int)L);
    cwhip_pair_params p;
    std::vector<int32_t> go(lmax + 1);
    rc = cwhip_pair_params_default(j.dna, j.dna ? 15.0f : 10.0f, j.dna ? 6.66f : 0.1f, 0.5f, lmax,
                                   &p, go.data());
    if (rc) return rc;
    const int64_t np = cwhip_range_npairs(n, 0, n, 0, n);
    std::vector<int32_t> pairs(2 * std::max<int64_t>(np, 1)), cnt(std::max<int64_t>(np, 1));
    j.dist.assign(std::max<int64_t>(np, 1), 0.0f);
    rc = cwhip_pairwise_range(ctx, &p, 0, n, 0, n, CWHIP_KERNEL_BATCHED, pairs.data(), cnt.data(),
                              j.dist.data());
    if (rc) return rc;
    j.dist.resize(np);
    for (int64_t k = 0; k < np; ++k)
        j.cells += (double)j.lens[pairs[2 * k]] * j.lens[pairs[2 * k + 1]];
    if (j.kind != CWHIP_JOB_ALIGNMENT || n < 2) return CWHIP_OK;
    std::vector<const char*> nm(n);
    for (int i = 0; i < n; ++i) nm[i] = j.names[i].c_str();
    // the guide tree from the distances still in the slot's device memory;
    // one tree build: room for every name and two branch lengths per node, a
    // second call only if the text did not fit
    int64_t len = 0, cap = 80 * (int64_t)(n + 2) + 64;
    for (int i = 0; i < n; ++i) cap += (int64_t)j.names[i].size();
    j.dnd.assign(cap, '\0');
    rc = cwhip_nj_tree_from_run(ctx, nm.data(), &j.dnd[0], cap, &len);
    if (rc && len + 1 > cap) {
        j.dnd.assign(len + 1, '\0');
        rc = cwhip_nj_tree_from_run(ctx, nm.data(), &j.dnd[0], len + 1, &len);
    }
    if (rc) return rc;
    j.dnd.resize(len);
    std::vector<int32_t> w(n), steps((size_t)std::max(n - 1, 1) * (n + 1));
    std::vector<double> sim((size_t)(n + 1) * (n + 1));
    int32_t nsteps = 0, dtree = 1;
    rc = cwhip_guide_from_tree(n, j.dnd.c_str(), nm.data(), w.data(), sim.data(), steps.data(), &nsteps,
                               &dtree);
    if (rc) return rc;
    cwhip_msa_params mp;
    cwhip_msa_params_default(j.dna, &mp);
    mp.distance_tree = dtree;
    std::vector<int32_t> order(n);
    int32_t alen = 0;
    rc = cwhip_msa(ctx, &mp, n, j.codes.data(), j.offs.data(), j.lens.data(), sim.data(), w.data(),
                   steps.data(), nsteps, order.data(), &alen);
    if (rc) return rc;
    double st[6];
    cwhip_msa_stats(ctx, st);
    j.cells += st[2];
    const int64_t stride = std::max(alen, 1);
    std::vector<uint8_t> rows((size_t)n * stride);
    rc = cwhip_msa_rows(ctx, rows.data(), stride);
    if (rc) return rc;
    std::vector<int32_t> rl(n, alen);
    cwhip_clustal_text(n, nm.data(), rows.data(), stride, rl.data(), order.data(), j.dna, nullptr, 0, &len);
    j.aln.assign(len + 1, '\0');
    rc = cwhip_clustal_text(n, nm.data(), rows.data(), stride, rl.data(), order.data(), j.dna,
                            &j.aln[0], len + 1, &len);
    if (rc) return rc;
    j.aln.resize(len);
    return CWHIP_OK;
}

// Starts pending jobs in arrival order while the oldest one fits (called with
// the server's lock held).  The slot for a job is the idle one whose held
// buffers need the least growth; when the job does not fit the budget with
// what idle slots hold, their buffers are released, largest first.
void schedule(cwhip_server* srv)
{
    while (!srv->stop && !srv->pending.empty()) {
        Job* j = srv->pending.front();
        int best = -1;
        int64_t best_growth = 0;
        for (int s = 0; s < (int)srv->slots.size(); ++s) {
            const Slot& sl = srv->slots[s];
            if (sl.job) continue;
            const int64_t growth = std::max<int64_t>(0, j->footprint - sl.charge);
            if (best < 0 || growth < best_growth) {
                best = s;
                best_growth = growth;
            }
        }
        if (best < 0) return;                               // no idle slot
        if (srv->charged() + best_growth > srv->budget) {
            // release idle slots' buffers (the chosen slot's last) until it fits
            std::vector<int> idle;
            for (int s = 0; s < (int)srv->slots.size(); ++s)
                if (!srv->slots[s].job && s != best && srv->slots[s].charge > 0) idle.push_back(s);
            std::sort(idle.begin(), idle.end(),
                      [&](int a, int b) { return srv->slots[a].charge > srv->slots[b].charge; });
            idle.push_back(best);
            int64_t held_by_idle = 0;
            for (int s : idle) held_by_idle += srv->slots[s].charge;
            const int64_t running = srv->charged() - held_by_idle;
            if (running + j->footprint > srv->budget) return;   // wait for running jobs
            for (int s : idle) {
                if (srv->charged() + std::max<int64_t>(0, j->footprint - srv->slots[best].charge) <= srv->budget)
                    break;
                Slot& sl = srv->slots[s];
                (void)hipSetDevice(srv->device);
                sl.ctx->release_buffers();
                sl.charge = sl.ctx->mem.bytes;
            }
        }
        Slot& sl = srv->slots[best];
        srv->pending.pop_front();
        sl.job = j;
        sl.charge = std::max(sl.charge, j->footprint);
        sl.ctx->mem.limit = sl.charge;
        srv->peak = std::max(srv->peak, srv->charged());
        ++srv->running;
        srv->max_running = std::max(srv->max_running, srv->running);
        j->state = 1;
        if (j->tries == 0) j->t_start = srv->now();
    }
}

void worker(cwhip_server* srv, int slot)
{
    for (;;) {
        Job* j = nullptr;
        cwhip_ctx* ctx = nullptr;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_work.wait(lk, [&] { return srv->stop || srv->slots[slot].job != nullptr; });
            if (srv->stop) return;
            j = srv->slots[slot].job;
            ctx = srv->slots[slot].ctx;
        }
        const int rc = run_job(ctx, *j);
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            Slot& sl = srv->slots[slot];
            sl.job = nullptr;
            --srv->running;
            if (rc == CWHIP_E_NOMEM && j->footprint < srv->budget) {
                // the estimate was short: start again from empty buffers with
                // twice the reservation, ahead of every other pending job
                (void)hipSetDevice(srv->device);
                ctx->release_buffers();
                ctx->mem.limit = 0;
                sl.charge = ctx->mem.bytes;
                j->footprint = std::min(srv->budget, 2 * j->footprint);
                ++j->tries;
                j->cells = 0;
                j->state = 0;
                srv->pending.push_front(j);
            } else {
                ctx->mem.limit = 0;
                sl.charge = ctx->mem.bytes;                // what the context keeps
                j->rc = rc;
                if (rc) j->err = cwhip_last_error(ctx);
                j->state = rc ? 3 : 2;
                j->t_end = srv->now();
                srv->done += 1;
                srv->cells += j->cells;
            }
            schedule(srv);
        }
        srv->cv_done.notify_all();
        srv->cv_work.notify_all();
    }
}

Job* find(cwhip_server* srv, int64_t id)
{
    auto it = srv->jobs.find(id);
    return it == srv->jobs.end() ? nullptr : it->second.get();
}

}  // namespace

extern "C" {

int cwhip_server_create(int device, int slots, int64_t hbm_budget, cwhip_server** out)
{
    if (!out || slots < 1 || slots > 16 || hbm_budget <= 0) return CWHIP_E_ARG;
    *out = nullptr;
    auto* srv = new cwhip_server();
    srv->device = device;
    srv->budget = hbm_budget;
    srv->t0 = Clock::now();
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete srv;
        return CWHIP_E_HIP;
    }
    srv->num_cu = prop.multiProcessorCount;
    for (int s = 0; s < slots; ++s) {
        cwhip_ctx* c = nullptr;
        const int rc = cwhip_create(device, &c);
        if (rc) {
            for (Slot& x : srv->slots) cwhip_destroy(x.ctx);
            delete srv;
            return rc;
        }
        c->mem.total = &srv->real_total;
        c->mem.peak = &srv->real_peak;
        Slot sl;
        sl.ctx = c;
        srv->slots.push_back(sl);
    }
    for (int s = 0; s < slots; ++s) srv->workers.emplace_back(worker, srv, s);
    *out = srv;
    return CWHIP_OK;
}

void cwhip_server_destroy(cwhip_server* srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto& t : srv->workers) t.join();
    for (Slot& sl : srv->slots) cwhip_destroy(sl.ctx);
    delete srv;
}

int cwhip_server_submit(cwhip_server* srv, int kind, int dna, int nseq, const uint8_t* codes,
                        const int64_t* offsets, const int32_t* lens, const char* const* names,
                        int64_t* job)
{
    if (!srv || !job || nseq < 1 || !codes || !offsets || !lens ||
        (kind != CWHIP_JOB_DISTANCES && kind != CWHIP_JOB_ALIGNMENT) ||
        (kind == CWHIP_JOB_ALIGNMENT && !names))
        return CWHIP_E_ARG;
    auto j = std::make_unique<Job>();
    j->kind = kind;
    j->dna = dna ? 1 : 0;
    j->nseq = nseq;
    int64_t total = 0;
    j->offs.resize(nseq);
    j->lens.assign(lens, lens + nseq);
    for (int s = 0; s < nseq; ++s) {
        if (lens[s] < 0 || lens[s] > CWHIP_MAX_LEN) return CWHIP_E_ARG;
        j->offs[s] = total;
        total += lens[s];
    }
    j->codes.resize(std::max<int64_t>(total, 1));
    for (int s = 0; s < nseq; ++s) std::memcpy(&j->codes[j->offs[s]], codes + offsets[s], lens[s]);
    if (names)
        for (int s = 0; s < nseq; ++s) j->names.emplace_back(names[s] ? names[s] : "");
    j->footprint = footprint(*j, srv->num_cu);
    if (j->footprint > srv->budget) return CWHIP_E_NOMEM;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        j->id = srv->next_id++;
        j->t_submit = srv->now();
        *job = j->id;
        srv->pending.push_back(j.get());
        srv->jobs[j->id] = std::move(j);
        schedule(srv);
    }
    srv->cv_work.notify_all();
    return CWHIP_OK;
}

int cwhip_server_wait(cwhip_server* srv, int64_t id, double timeout_s)
{
    if (!srv) return CWHIP_E_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    Job* j = find(srv, id);
    if (!j) return CWHIP_E_ARG;
    const auto until = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                          std::chrono::duration<double>(timeout_s < 0 ? 1e9 : timeout_s));
    if (!srv->cv_done.wait_until(lk, until, [&] { return j->state >= 2; })) return CWHIP_E_TIMEOUT;
    return j->state == 2 ? CWHIP_OK : j->rc;
}

int cwhip_server_distances(cwhip_server* srv, int64_t id, float* dist)
{
    if (!srv || !dist) return CWHIP_E_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    Job* j = find(srv, id);
    if (!j || j->state != 2) return CWHIP_E_ARG;
    std::memcpy(dist, j->dist.data(), j->dist.size() * sizeof(float));
    return CWHIP_OK;
}

int cwhip_server_text(cwhip_server* srv, int64_t id, int which, char* out, int64_t cap, int64_t* len)
{
    if (!srv || !len) return CWHIP_E_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    Job* j = find(srv, id);
    if (!j || which < 0 || which > 2) return CWHIP_E_ARG;
    if (which == 2 ? j->state != 3 : (j->state != 2 || j->kind != CWHIP_JOB_ALIGNMENT))
        return CWHIP_E_ARG;
    const std::string& s = which == 2 ? j->err : which ? j->aln : j->dnd;
    *len = (int64_t)s.size();
    if (!out || cap < *len + 1) return CWHIP_E_ARG;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return CWHIP_OK;
}

int cwhip_server_times(cwhip_server* srv, int64_t id, double* t4)
{
    if (!srv || !t4) return CWHIP_E_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    Job* j = find(srv, id);
    if (!j) return CWHIP_E_ARG;
    t4[0] = j->t_submit;
    t4[1] = j->t_start;
    t4[2] = j->t_end;
    t4[3] = j->cells;
    return CWHIP_OK;
}

int cwhip_server_release(cwhip_server* srv, int64_t id)
{
    if (!srv) return CWHIP_E_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    Job* j = find(srv, id);
    if (!j || j->state < 2) return CWHIP_E_ARG;
    srv->jobs.erase(id);
    return CWHIP_OK;
}

int cwhip_server_stats(cwhip_server* srv, double* out6)
{
    if (!srv || !out6) return CWHIP_E_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    out6[0] = srv->done;
    out6[1] = srv->cells;
    out6[2] = (double)std::max<int64_t>(srv->real_peak.load(), 0);
    out6[3] = srv->max_running;
    out6[4] = (double)srv->budget;
    out6[5] = (double)srv->slots.size();
    return CWHIP_OK;
}

}  // extern "C"
