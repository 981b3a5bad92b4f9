// server.hip -- served mode: continuous batching of partition jobs.
//
// Requests queue until a slot and their HBM footprint are free.  Every
// scheduling step advances all admitted jobs together, whatever stage each
// one is in: newly admitted jobs get their degrees copied out and their hub
// search (a host sort) runs on worker threads while the device keeps
// working; jobs whose hubs are ready run their seed search; every job in
// expansion advances a few rounds in one job set; every job in phase 2
// advances a few LP rounds in another.  Finished jobs release their memory
// and slot at once, so a request that arrives while others are on the GPU
// joins them at the next step.  Each job touches only its own buffers, so
// its result is the same alone and in any mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "phase1.hpp"
#include "phase2.hpp"

namespace dmolp {
void gen_device(dmolp_ctx *ctx, const dmolp_gen_spec *spec, dmolp_graph *g);
}  // namespace dmolp

namespace {

using namespace dmolp;

enum JobState { JS_QUEUED, JS_HUBS, JS_EXPAND, JS_P2INIT, JS_P2, JS_DONE };

struct SrvJob {
    int64_t id = 0;
    dmolp_request req{};
    dmolp_graph *g = nullptr;
    bool own_graph = false;
    JobState st = JS_QUEUED;
    int64_t footprint = 0;
    P1Stage p1;
    dmolp_phase1_metrics m1{};
    P2Item p2;
    std::atomic<int> hubs_ready{0};
    dmolp_served out{};
    std::vector<int32_t> labels;
};

// Vertices and CSR entries (upper bound) of a generator spec.
void spec_size(const dmolp_gen_spec &s, int64_t &n, int64_t &entries) {
    if (s.kind == DMOLP_GEN_RMAT) {
        n = s.scale >= 0 && s.scale < 31 ? (int64_t)1 << s.scale : 0;
        entries = 2 * (int64_t)s.edge_factor * n;
    } else if (s.kind == DMOLP_GEN_GRID) {
        n = (int64_t)s.rows * s.cols;
        entries = 4 * n;
    } else {
        n = s.n;
        entries = 2 * s.m;
    }
}

// Device bytes a job holds while admitted: CSR, labels and the phase-1 and
// phase-2 work arrays (about 276 bytes per vertex, 64 of them the kept
// light-row histograms), plus the generator's key buffers for generated
// graphs.
int64_t job_footprint(const dmolp_request &r) {
    const int64_t per_vertex = 276;
    if (r.graph) return per_vertex * r.graph->n;
    int64_t n = 0, e = 0;
    spec_size(r.spec, n, e);
    return 8 * (n + 1) + 4 * e + per_vertex * n + 16 * e;
}

}  // namespace

struct dmolp_server {
    dmolp_ctx *ctx = nullptr;
    int64_t budget = 0;
    int64_t used = 0;
    int max_active = 0;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;  // queue, finished, next_id
    int64_t next_id = 1;
    std::deque<SrvJob *> queue;
    std::vector<SrvJob *> active;
    std::unordered_map<int64_t, SrvJob *> finished;
    // Hub-search workers.
    std::mutex hmu;
    std::condition_variable hcv;
    std::deque<SrvJob *> hubq;
    bool stop = false;
    std::vector<std::thread> workers;
    // Device records of the per-step job sets.
    P1SetDev seeds_set, expand_set;
    BatchState init_set, round_set;
    // Rounds a step advances the jobs in expansion and in phase 2 (even for
    // expansion).  On the 96-job R-MAT mix 32 and 4 served 250 jobs/s against
    // 230 for 16 and 8: shorter phase-2 stretches let jobs that finish their
    // expansion join the phase-2 set sooner.
    int p1_rounds_per_step = 32;
    int p2_rounds_per_step = 4;

    double clock() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void worker() {
        for (;;) {
            SrvJob *j = nullptr;
            {
                std::unique_lock<std::mutex> lk(hmu);
                hcv.wait(lk, [&] { return stop || !hubq.empty(); });
                if (stop && hubq.empty()) return;
                j = hubq.front();
                hubq.pop_front();
            }
            int ready = 1;
            try {
                p1_hub_search(j->p1);
            } catch (const std::bad_alloc &) {
                ready = DMOLP_E_NOMEM;  // negative: the job finishes with this status
            } catch (...) {
                ready = DMOLP_E_INTERNAL;
            }
            j->hubs_ready.store(ready, std::memory_order_release);
        }
    }
    ~dmolp_server() {
        {
            std::lock_guard<std::mutex> lk(hmu);
            stop = true;
        }
        hcv.notify_all();
        for (auto &t : workers) t.join();
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        auto drop = [&](SrvJob *j) {
            if (j->own_graph) delete j->g;
            delete j;
        };
        for (SrvJob *j : queue) drop(j);
        for (SrvJob *j : active) drop(j);
        for (auto &kv : finished) drop(kv.second);
    }
};

namespace dmolp {

dmolp_server *server_create(dmolp_ctx *ctx, int64_t budget, int32_t max_active) {
    DMOLP_REQUIRE(budget > 0 && max_active > 0, DMOLP_E_ARG, "budget and max_active must be positive");
    std::unique_ptr<dmolp_server> s(new dmolp_server());
    s->ctx = ctx;
    s->budget = budget;
    s->max_active = max_active;
    s->t0 = std::chrono::steady_clock::now();
    unsigned T = std::thread::hardware_concurrency();
    if (T > 8) T = 8;
    if (T < 1) T = 1;
    for (unsigned i = 0; i < T; ++i) s->workers.emplace_back([p = s.get()] { p->worker(); });
    return s.release();
}

double server_clock(const dmolp_server *s) { return s->clock(); }

int64_t server_submit(dmolp_server *s, const dmolp_request &r) {
    DMOLP_REQUIRE(r.k >= 1 && r.k <= DMOLP_MAX_K, DMOLP_E_ARG, "k outside [1, 1024]");
    DMOLP_REQUIRE(r.theta > 0, DMOLP_E_ARG, "theta must be positive");
    DMOLP_REQUIRE(r.max_rounds >= 0, DMOLP_E_ARG, "max_rounds must be >= 0");
    if (!r.graph) {
        int64_t n = 0, e = 0;
        spec_size(r.spec, n, e);
        DMOLP_REQUIRE(n > 0, DMOLP_E_ARG, "request without a graph");
    }
    const int64_t fp = job_footprint(r);
    DMOLP_REQUIRE(fp <= s->budget, DMOLP_E_UNSUPPORTED, "request larger than the server's memory budget");
    SrvJob *j = new SrvJob();
    j->req = r;
    j->footprint = fp;
    j->out.arrival_s = r.arrival_s >= 0 ? r.arrival_s : s->clock();
    std::lock_guard<std::mutex> lk(s->mu);
    j->id = s->next_id++;
    s->queue.push_back(j);
    return j->id;
}

namespace {

void finish_job(dmolp_server *s, SrvJob *j, int status) {
    j->out.result.status = status;
    j->out.done_s = s->clock();
    if (j->g) {
        j->out.n = j->g->n;
        j->out.nnz = j->g->nnz;
    }
    if (j->own_graph) {
        (void)hipStreamSynchronize(s->ctx->stream);
        delete j->g;
        j->g = nullptr;
    }
    s->used -= j->footprint;
    j->st = JS_DONE;
    std::lock_guard<std::mutex> lk(s->mu);
    s->finished[j->id] = j;
}

}  // namespace

void server_step(dmolp_server *s, int32_t *active_out, int32_t *queued_out) {
    dmolp_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    PoolScope scope(&ctx->pool, ctx->stream);
    DMOLP_HIP_CHECK(hipSetDevice(ctx->device));
    const double now = s->clock();
    // 1. Admission: first come, first served, while a slot and the memory
    //    budget allow.
    std::vector<SrvJob *> admitted;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        while (!s->queue.empty()) {
            SrvJob *j = s->queue.front();
            if (j->out.arrival_s > now) break;
            if ((int)(s->active.size() + admitted.size()) >= s->max_active) break;
            if (s->used + j->footprint > s->budget) break;
            s->queue.pop_front();
            s->used += j->footprint;
            admitted.push_back(j);
        }
    }
    std::vector<P1Stage *> fresh;
    std::vector<SrvJob *> fresh_jobs;
    for (SrvJob *j : admitted) {
        j->out.admit_s = s->clock();
        try {
            if (j->req.graph) {
                j->g = j->req.graph;
                note_use(ctx, j->g);
            } else {
                j->g = new dmolp_graph();
                j->g->pool = ctx->pool.get();
                j->g->device = ctx->device;
                j->own_graph = true;
                gen_device(ctx, &j->req.spec, j->g);
                graph_max_degree(ctx, j->g);
            }
            P1Item it;
            it.g = j->g;
            it.k = j->req.k;
            it.theta = j->req.theta;
            p1_stage_init(ctx, it, j->p1);
            j->st = JS_HUBS;
            s->active.push_back(j);
            fresh.push_back(&j->p1);
            fresh_jobs.push_back(j);
        } catch (const Error &e) {
            finish_job(s, j, e.status);
        }
    }
    if (!fresh.empty()) {
        p1_set_build(ctx, fresh, s->seeds_set);
        p1_degrees(ctx, fresh, s->seeds_set);
        // Long graphs: partition steps on the device, now (a few ms); short
        // ones: host sorts on the workers while the device goes on.
        int dev_status = 1;
        try {
            p1_hub_search_device(ctx, fresh);
        } catch (const Error &e) {
            dev_status = e.status < 0 ? e.status : DMOLP_E_INTERNAL;
        } catch (const std::bad_alloc &) {
            dev_status = DMOLP_E_NOMEM;
        }
        {
            std::lock_guard<std::mutex> lk(s->hmu);
            for (SrvJob *j : fresh_jobs) {
                if (j->p1.dev_hubs) j->hubs_ready.store(dev_status, std::memory_order_release);
                else s->hubq.push_back(j);
            }
        }
        s->hcv.notify_all();
    }
    // 2. Seeds for jobs whose hub search is done; they enter expansion.
    std::vector<P1Stage *> seeded;
    std::vector<SrvJob *> seeded_jobs;
    std::vector<SrvJob *> failed;
    for (SrvJob *j : s->active) {
        if (j->st != JS_HUBS) continue;
        const int ready = j->hubs_ready.load(std::memory_order_acquire);
        if (ready > 0) {
            seeded.push_back(&j->p1);
            seeded_jobs.push_back(j);
        } else if (ready < 0) {
            failed.push_back(j);
        }
    }
    if (!failed.empty()) {
        for (SrvJob *j : failed) finish_job(s, j, j->hubs_ready.load(std::memory_order_acquire));
        s->active.erase(std::remove_if(s->active.begin(), s->active.end(),
                                       [](SrvJob *j) { return j->st == JS_DONE; }),
                        s->active.end());
    }
    if (!seeded.empty()) {
        p1_seeds(ctx, seeded, s->seeds_set);
        p1_expand_begin(ctx, seeded, s->seeds_set);
        for (SrvJob *j : seeded_jobs) j->st = JS_EXPAND;
    }
    // 3. Expansion rounds for every job in expansion.
    std::vector<P1Stage *> exp;
    std::vector<SrvJob *> exp_jobs;
    for (SrvJob *j : s->active)
        if (j->st == JS_EXPAND) { exp.push_back(&j->p1); exp_jobs.push_back(j); }
    if (!exp.empty()) {
        p1_set_build(ctx, exp, s->expand_set);
        // Rounds that need no selection, one workgroup per job (a bounded
        // number per step, so every step stays short); then, if a job
        // stopped at a round that needs one, rounds as separate launches.
        int ran = 0, stopped = 0;
        p1_rounds_wg(ctx, s->expand_set, 256, ran, stopped);
        if (stopped > 0) p1_expand_rounds(ctx, s->expand_set, s->p1_rounds_per_step);
        std::vector<int32_t> done;
        p1_done_flags(ctx, s->expand_set, done);
        for (size_t i = 0; i < exp_jobs.size(); ++i) {
            if (!done[i]) continue;
            SrvJob *j = exp_jobs[i];
            p1_finish(ctx, j->p1, j->m1, 0.0);
            j->p2.g = j->g;
            j->p2.k = j->req.k;
            j->p2.max_rounds = j->req.max_rounds;
            j->p2.epsilon = j->req.epsilon;
            j->st = JS_P2INIT;
        }
    }
    // 4. Phase-2 start (statistics sweep) for jobs leaving expansion.
    std::vector<P2Item *> init;
    for (SrvJob *j : s->active)
        if (j->st == JS_P2INIT) init.push_back(&j->p2);
    if (!init.empty()) {
        const P2Set set = p2_set_build(ctx, init, s->init_set, nullptr);
        phase2_init_enqueue(ctx, set);
        for (SrvJob *j : s->active)
            if (j->st == JS_P2INIT) j->st = JS_P2;
    }
    // 5. LP rounds for every job in phase 2; finished jobs leave.
    std::vector<P2Item *> p2;
    std::vector<SrvJob *> p2_jobs;
    for (SrvJob *j : s->active)
        if (j->st == JS_P2) { p2.push_back(&j->p2); p2_jobs.push_back(j); }
    if (!p2.empty()) {
        const P2Set set = p2_set_build(ctx, p2, s->round_set, nullptr);
        for (int r = 0; r < s->p2_rounds_per_step; ++r) phase2_round_enqueue(ctx, set);
        std::vector<int32_t> done;
        phase2_done_flags(ctx, set, s->round_set.done.as<int32_t>(), done);
        std::vector<P2Item *> fin;
        std::vector<SrvJob *> fin_jobs;
        for (size_t i = 0; i < p2_jobs.size(); ++i)
            if (done[i]) { fin.push_back(p2[i]); fin_jobs.push_back(p2_jobs[i]); }
        if (!fin.empty()) {
            p2_collect(ctx, fin);
            for (SrvJob *j : fin_jobs) {
                if (j->req.want_labels && j->p2.err == 0) {
                    j->labels.resize((size_t)j->g->n);
                    DMOLP_HIP_CHECK(hipMemcpyAsync(j->labels.data(), j->g->labels.p,
                                                   sizeof(int32_t) * (size_t)j->g->n,
                                                   hipMemcpyDeviceToHost, st));
                }
            }
            DMOLP_HIP_CHECK(hipStreamSynchronize(st));
            for (SrvJob *j : fin_jobs) {
                dmolp_result &r = j->out.result;
                r.rounds = j->p2.m.rounds;
                r.exit_reason = j->p2.m.exit_reason;
                r.n_seeds_fallback = j->m1.n_seeds_fallback;
                r.phase1_cut = j->p2.m.initial_cut;
                r.edge_cut = j->p2.m.edge_cut;
                r.lp_edges = j->p2.m.lp_edges;
                r.vertex_balance = j->p2.m.vertex_balance;
                r.edge_balance = j->p2.m.edge_balance;
                finish_job(s, j, j->p2.err);
            }
            s->active.erase(std::remove_if(s->active.begin(), s->active.end(),
                                           [](SrvJob *j) { return j->st == JS_DONE; }),
                            s->active.end());
        }
    }
    if (active_out) *active_out = (int32_t)s->active.size();
    if (queued_out) {
        std::lock_guard<std::mutex> lk(s->mu);
        *queued_out = (int32_t)s->queue.size();
    }
}

int server_take(dmolp_server *s, int64_t ticket, dmolp_served *out, int32_t *labels_out) {
    SrvJob *j = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->finished.find(ticket);
        if (it == s->finished.end()) return 0;
        j = it->second;
        s->finished.erase(it);
    }
    if (out) *out = j->out;
    if (labels_out && !j->labels.empty()) std::copy(j->labels.begin(), j->labels.end(), labels_out);
    delete j;
    return 1;
}

void server_destroy(dmolp_server *s) { delete s; }

}  // namespace dmolp
