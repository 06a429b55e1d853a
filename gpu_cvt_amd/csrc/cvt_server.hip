// cvt_server.hip -- continuous-batching request loop over the public C ABI.
//
// The reference processes one mesh per process invocation (main.cu:371) and
// has no server.  Here requests (mesh + sites + parameters) arrive at any
// time; an ingest thread turns them into device-resident jobs while their
// footprint fits the HBM budget, and a scheduler thread advances every active
// job together, E iterations per epoch, as one batch graph (cvt_batch).  At
// each epoch boundary the scheduler checks convergence -- the one-sided
// Hausdorff distance of the remeshed surface (cvt_job_hausdorff), the check
// the reference runs every DUMP_STRIDE = 10 iterations (main.cu:395,1034) --
// retires finished jobs (freeing their memory) and admits jobs that became
// ready.  A group's batch has a fixed launch shape over job slots
// (cvt::SlotBatch): joining and retiring jobs edit its device slot table in
// place, and its graph is captured again only when the tables outgrow their
// capacity or shrink to a quarter of it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "cvt.h"
#include "cvt_internal.h"

using cvt::fail;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    unsigned long long ticket = 0;
    std::vector<double> V, S;
    std::vector<int> F;
    cvt_params p{};
    double tol_rel = 1e-3;
    int max_iters = 100;
    // lifecycle
    cvt_result res{};
    cvt_job* job = nullptr;
    unsigned long long bytes = 0;
    double diag = 0;
    int next_check = 0;
    double checked_h = -1;  // this epoch's Hausdorff distance (-1: not computed)
    int checked_in = 0;     // this epoch's check found every probe within tolerance
    bool in_batch = false;  // stepped by its group's slot batch (else on its own)
    std::vector<double> S_out;
};

double secs(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

}  // namespace

struct cvt_server {
    cvt_server_config cfg{};
    // device memory of every job, LBVH, batch and check temporary of this
    // server (declared before its users, so it is released after them)
    cvt::DevArena arena;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request*> pending;  // submitted, not yet on the device
    std::vector<Request*> ready;   // device job created, waiting for the next epoch
    std::vector<Request*> active;  // in the batch
    std::map<unsigned long long, Request*> all;
    unsigned long long next_ticket = 1;
    unsigned long long used = 0, peak = 0;
    bool stop = false;
    std::vector<std::thread> ingest_th;
    std::thread sched_th;
    std::vector<cvt::TempPool> check_pools;  // device temporaries, one set per check worker
    unsigned long long pool_bytes = 0;       // their size, counted in used
    // counters
    double site_iters = 0, busy_s = 0, rebuild_s = 0, step_s = 0, check_s = 0;
    long epochs = 0, rebuilds = 0, checks = 0;
    cvt::SlotBatchCounters batch_n{};  // summed over the groups' batches
    int failures = 0;

    void ingest_loop();
    void sched_loop();
    void finish(Request* r, int status);
};

// Conservative estimate of a job's device footprint before creating it (the
// exact figure replaces it once the job exists).
static unsigned long long estimate_bytes(int nV, int nF, int nS, int K) {
    const double v = 16.0 * 2 * nV + 4.0 * (nV + 1) + 32.0 * nV;   // V, N, CSR offsets, vertex grid
    const double f = 48.0 * nF + 12.0 * nF + 12.0 * nF + 16.0 * (3.0 * nF + nV) + 4.0 * nV;  // faces, F, CSR, rings
    const double s = nS * (16.0 * 2 + 4.0 * K + 20 + 3 + 4 + 4 + 16 + 4 * 2 + 4 + 4 * 32 + 20 + 32 + 4);
    return (unsigned long long)((v + f + s) * 1.25) + (1ull << 20);
}

void cvt_server::finish(Request* r, int status) {
    // called by the scheduler without the lock held
    if (r->job) {
        r->S_out.resize(3 * (size_t)r->res.nS);
        if (cvt_job_get_sites(r->job, r->S_out.data()) != CVT_OK) status = 4;
        cvt_job_destroy(r->job);
        r->job = nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    used -= r->bytes;
    r->res.status = status;
    r->res.t_done = secs(t0);
    cv.notify_all();
}

void cvt_server::ingest_loop() {
    cvt::current_arena() = &arena;
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] {
                if (stop) return true;
                if (pending.empty()) return false;
                Request* f = pending.front();
                const unsigned long long est = estimate_bytes((int)f->V.size() / 3, (int)f->F.size() / 3,
                                                              (int)f->S.size() / 3, f->p.K);
                return used + est <= cfg.hbm_budget_bytes || (active.empty() && ready.empty() && used == pool_bytes);
            });
            if (stop) return;
            r = pending.front();
            pending.pop_front();
            r->bytes = estimate_bytes((int)r->V.size() / 3, (int)r->F.size() / 3, (int)r->S.size() / 3, r->p.K);
            used += r->bytes;
        }
        cvt_job* job = nullptr;
        const int rc = cvt_job_create(r->V.data(), (int)r->V.size() / 3, r->F.data(), (int)r->F.size() / 3,
                                      r->S.data(), (int)r->S.size() / 3, &r->p, &job);
        std::lock_guard<std::mutex> lk(mu);
        if (rc != CVT_OK) {
            used -= r->bytes;
            r->res.status = 4;
            r->res.error = rc;
            r->res.t_done = secs(t0);
            ++failures;
        } else {
            cvt_job_info info;
            cvt_job_info_get(job, &info);
            // the job's own buffers plus the LBVH its first convergence check
            // builds over the input mesh (about 96 bytes per face)
            const unsigned long long b = info.device_bytes + 96ull * (unsigned long long)info.nF;
            used = used - r->bytes + b;
            peak = std::max(peak, used);
            r->bytes = b;
            r->diag = info.bbox_diag;
            r->res.bbox_diag = info.bbox_diag;
            r->job = job;
            r->res.t_admit = secs(t0);
            r->res.status = 1;
            r->next_check = cfg.epoch_iters;
            // the request's host copies are no longer needed
            std::vector<double>().swap(r->V);
            std::vector<int>().swap(r->F);
            std::vector<double>().swap(r->S);
            ready.push_back(r);
        }
        cv.notify_all();
    }
}

void cvt_server::sched_loop() {
    cvt::current_arena() = &arena;
    // One slot batch per (K, kproj) group, edited in place as jobs join and retire.
    struct Group {
        std::vector<Request*> members;  // in joining order
        cvt::SlotBatch* batch = nullptr;
    };
    std::map<std::pair<int, int>, Group> groups;
    // (a group's batch lives until the server stops)
    auto counters_now = [&] {
        cvt::SlotBatchCounters c{};
        for (auto& kv : groups)
            if (kv.second.batch) {
                const cvt::SlotBatchCounters g = cvt::slot_batch_counters(kv.second.batch);
                c.captures += g.captures;
                c.first += g.first;
                c.grow += g.grow;
                c.compactions += g.compactions;
                c.extent_updates += g.extent_updates;
            }
        return c;
    };
    for (;;) {
        std::vector<Request*> joined;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stop || !ready.empty() || !active.empty(); });
            if (stop) break;
            joined.swap(ready);
            active.insert(active.end(), joined.begin(), joined.end());
        }
        const auto e0 = Clock::now();
        double t_rebuild = 0, t_step = 0, t_check = 0;
        for (Request* r : joined) {
            Group& g = groups[{r->p.K, r->p.kproj}];
            g.members.push_back(r);
            const auto r0 = Clock::now();
            if (!g.batch && cvt::slot_batch_create(r->p.K, r->p.kproj, cfg.batch_jobs0, cfg.batch_site_blocks0, &g.batch) != CVT_OK) g.batch = nullptr;
            r->in_batch = g.batch && cvt::slot_batch_add(g.batch, r->job) == CVT_OK;
            t_rebuild += secs(r0);
        }
        double si = 0;
        long ncheck = 0;
        std::vector<Request*> done;
        std::vector<int> status;
        for (auto& kv : groups) {
            Group& g = kv.second;
            if (g.members.empty()) continue;
            const auto r1 = Clock::now();
            if (g.batch) cvt::slot_batch_step(g.batch, cfg.epoch_iters);
            for (Request* r : g.members)
                if (!r->in_batch) cvt_job_step(r->job, cfg.epoch_iters, nullptr);
            t_step += secs(r1);
            // convergence at the epoch boundary: gate each job, then run the
            // Hausdorff checks of the gated jobs in parallel host threads (each
            // on its job's own stream, with its own device temporaries)
            std::vector<Request*> check;
            std::vector<char> is_last(g.members.size());
            for (size_t k = 0; k < g.members.size(); ++k) {
                Request* r = g.members[k];
                {
                    std::lock_guard<std::mutex> lk(mu);  // res is read by cvt_server_poll
                    r->res.iterations += cfg.epoch_iters;
                }
                si += (double)r->res.nS * cfg.epoch_iters;
                cvt_iter_stats st;
                std::memset(&st, 0, sizeof st);
                cvt_job_stats(r->job, 1, &st);
                const bool gate = r->res.iterations >= cfg.min_check_iters &&
                                  ((double)st.max_disp <= cfg.check_disp_rel * r->diag ||
                                   st.frozen_total >= cfg.check_frozen_frac * r->res.nS);
                is_last[k] = r->res.iterations >= r->max_iters;
                r->checked_h = -1;
                r->checked_in = 0;
                if (gate || is_last[k]) check.push_back(r);
            }
            const int nth = (int)std::min<size_t>(check_pools.size(), check.size());
            std::atomic<size_t> next{0};
            std::atomic<long> nok{0};
            auto worker = [&](int w) {
                for (size_t c = next++; c < check.size(); c = next++) {
                    Request* r = check[c];
                    // bounded test first; the exact distance only for a job that
                    // converged or reached its last iteration
                    const bool last = r->res.iterations >= r->max_iters;
                    double h = -1;
                    int in = 0;
                    if (cvt::job_hausdorff(r->job, 1e-6, check_pools[w], r->tol_rel * r->diag, &in, last, &h,
                                           nullptr) == CVT_OK) {
                        r->checked_h = h;
                        r->checked_in = in;
                        ++nok;
                    }
                }
            };
            const auto c0 = Clock::now();
            std::vector<std::thread> ths;
            for (int w = 1; w < nth; ++w) ths.emplace_back(worker, w);
            if (!check.empty()) worker(0);
            for (std::thread& t : ths) t.join();
            t_check += secs(c0);
            ncheck += nok;
            std::vector<Request*> keep;
            for (size_t k = 0; k < g.members.size(); ++k) {
                Request* r = g.members[k];
                int st_done = 0;
                if (r->checked_h >= 0) {
                    std::lock_guard<std::mutex> lk(mu);
                    r->res.hausdorff = r->checked_h;
                }
                if (r->checked_h >= 0 && r->checked_in) st_done = 2;
                else if (is_last[k]) st_done = 3;
                if (st_done) {
                    done.push_back(r);
                    status.push_back(st_done);
                    // leaves the slot table before finish() destroys the job
                    const auto r2 = Clock::now();
                    if (r->in_batch) cvt::slot_batch_remove(g.batch, r->job);
                    r->in_batch = false;
                    t_rebuild += secs(r2);
                } else {
                    keep.push_back(r);
                }
            }
            g.members.swap(keep);
        }
        unsigned long long pb = 0;
        for (const cvt::TempPool& tp : check_pools) pb += tp.bytes();
        const cvt::SlotBatchCounters bn = counters_now();
        const double dt = secs(e0);
        {
            std::lock_guard<std::mutex> lk(mu);
            site_iters += si;
            busy_s += dt;
            ++epochs;
            rebuilds = bn.captures;
            batch_n = bn;
            checks += ncheck;
            rebuild_s += t_rebuild;
            step_s += t_step;
            check_s += t_check;
            // the check workers' temporaries count against the budget as well
            used = used - pool_bytes + pb;
            pool_bytes = pb;
            peak = std::max(peak, used);
            for (Request* r : done) active.erase(std::find(active.begin(), active.end(), r));
        }
        for (size_t k = 0; k < done.size(); ++k) finish(done[k], status[k]);
    }
    for (auto& kv : groups)
        if (kv.second.batch) cvt::slot_batch_destroy(kv.second.batch);
}

extern "C" {

void cvt_server_config_default(cvt_server_config* c) {
    std::memset(c, 0, sizeof *c);
    c->hbm_budget_bytes = 200ull << 30;
    c->epoch_iters = 10;
    c->min_check_iters = 10;
    c->check_disp_rel = 1e-3;
    c->check_frozen_frac = 0.5;
    c->ingest_threads = 4;
    c->check_threads = 8;
}

int cvt_server_create(const cvt_server_config* cfg, cvt_server** out) {
    if (!out) return fail(CVT_EINVAL, "null argument");
    cvt_server* s = new cvt_server();
    if (cfg) s->cfg = *cfg;
    else cvt_server_config_default(&s->cfg);
    if (s->cfg.epoch_iters < 1) s->cfg.epoch_iters = 1;
    s->t0 = Clock::now();
    s->check_pools.resize(std::max(1, std::min(s->cfg.check_threads, 64)));
    for (cvt::TempPool& tp : s->check_pools) tp.arena = &s->arena;
    const int nth = std::max(1, std::min(s->cfg.ingest_threads, 64));
    for (int k = 0; k < nth; ++k) s->ingest_th.emplace_back([s] { s->ingest_loop(); });
    s->sched_th = std::thread([s] { s->sched_loop(); });
    *out = s;
    return CVT_OK;
}

int cvt_server_submit(cvt_server* s, const double* V, int nV, const int* F, int nF, const double* S0, int nS,
                      const cvt_params* p, double tol_hausdorff_rel, int max_iters, unsigned long long* ticket) {
    if (!s || !V || !F || !S0 || !ticket || nV < 3 || nF < 1 || nS < 1 || max_iters < 1 || !(tol_hausdorff_rel >= 0))
        return fail(CVT_EINVAL, "bad argument");
    Request* r = new Request();
    r->V.assign(V, V + 3L * nV);
    r->F.assign(F, F + 3L * nF);
    r->S.assign(S0, S0 + 3L * nS);
    if (p) r->p = *p;
    else cvt_params_default(&r->p);
    r->tol_rel = tol_hausdorff_rel;
    r->max_iters = max_iters;
    r->res.nS = nS;
    r->res.hausdorff = -1;
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    r->res.t_arrive = secs(s->t0);
    s->all[r->ticket] = r;
    s->pending.push_back(r);
    *ticket = r->ticket;
    s->cv.notify_all();
    return CVT_OK;
}

int cvt_server_poll(cvt_server* s, unsigned long long ticket, cvt_result* out) {
    if (!s || !out) return fail(CVT_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(CVT_EINVAL, "unknown ticket %llu", ticket);
    *out = it->second->res;
    return CVT_OK;
}

int cvt_server_get_sites(cvt_server* s, unsigned long long ticket, double* S_out) {
    if (!s || !S_out) return fail(CVT_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(CVT_EINVAL, "unknown ticket %llu", ticket);
    Request* r = it->second;
    if (r->res.status < 2 || r->S_out.empty()) return fail(CVT_ESTATE, "request %llu has not finished", ticket);
    std::memcpy(S_out, r->S_out.data(), sizeof(double) * r->S_out.size());
    return CVT_OK;
}

int cvt_server_wait_idle(cvt_server* s, double timeout_s) {
    if (!s) return fail(CVT_EINVAL, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
        return s->pending.empty() && s->ready.empty() && s->active.empty() && s->used == s->pool_bytes;
    });
    return ok ? CVT_OK : fail(CVT_ESTATE, "timed out waiting for the server to drain");
}

int cvt_server_info_get(cvt_server* s, cvt_server_info* out) {
    if (!s || !out) return fail(CVT_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof *out);
    out->site_iters = s->site_iters;
    out->busy_seconds = s->busy_s;
    out->peak_bytes = s->peak;
    out->used_bytes = s->used;
    out->epochs = s->epochs;
    out->rebuilds = s->rebuilds;
    out->rebuilds_grow = (int)s->batch_n.grow;
    out->compactions = (int)s->batch_n.compactions;
    out->extent_updates = (int)s->batch_n.extent_updates;
    out->hausdorff_checks = s->checks;
    out->pending = (int)s->pending.size();
    out->active = (int)(s->active.size() + s->ready.size());
    out->failures = s->failures;
    out->step_seconds = s->step_s;
    out->check_seconds = s->check_s;
    out->rebuild_seconds = s->rebuild_s;
    return CVT_OK;
}

void cvt_server_destroy(cvt_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    for (std::thread& t : s->ingest_th) t.join();
    s->sched_th.join();
    for (auto& kv : s->all) {
        if (kv.second->job) cvt_job_destroy(kv.second->job);
        delete kv.second;
    }
    delete s;
}

}  // extern "C"
