// batch.hip -- many graphs partitioned together.
//
// Every job's phase 1 and phase 2 advance in lock step: one launch per
// kernel step serves all jobs (each job owns a share of the launch's blocks,
// in proportion to its size), the phase-2 round is captured once as a HIP
// graph and replayed, and the host checks a single counter of unfinished
// jobs.  A job's kernels touch only that job's buffers, so each result
// equals the job's plain run.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <vector>

#include "batch.hpp"
#include "opbatch.hpp"
#include "phase1.hpp"
#include "phase2.hpp"

namespace dmolp {
namespace {

uint64_t mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h;
}

// PartitioningMetrics balances (phase2.cpp:424-441) from final counts.
void final_balance(int k, const std::vector<int64_t> &cv, const std::vector<int64_t> &ci,
                   double &max_v, double &max_e) {
    int64_t tv = 0, te = 0;
    for (int i = 0; i < k; ++i) { tv += cv[(size_t)i]; te += ci[(size_t)i]; }
    const double exp_v = static_cast<double>(tv) / k;
    const double exp_e = (te > 0) ? static_cast<double>(te) / k : 1.0;
    max_v = 0.0;
    max_e = 0.0;
    for (int i = 0; i < k; ++i) {
        const double rv = (exp_v > 0) ? static_cast<double>(cv[(size_t)i]) / exp_v : 1.0;
        const double r_e = (exp_e > 0) ? static_cast<double>(ci[(size_t)i]) / exp_e : 1.0;
        max_v = std::max(max_v, rv);
        max_e = std::max(max_e, r_e);
    }
}

}  // namespace

void batch_check_job(const dmolp_job &j) {
    DMOLP_REQUIRE(j.graph != nullptr, DMOLP_E_ARG, "job without a graph");
    DMOLP_REQUIRE(j.k >= 1 && j.k <= DMOLP_MAX_K, DMOLP_E_ARG, "k outside [1, 1024]");
    DMOLP_REQUIRE(j.theta > 0, DMOLP_E_ARG, "theta must be positive");
    DMOLP_REQUIRE(j.max_rounds >= 0, DMOLP_E_ARG, "max_rounds must be >= 0");
}

P2Set p2_set_build(dmolp_ctx *ctx, std::vector<P2Item *> &items, BatchState &bs, uint64_t *key_out) {
    const int G = (int)items.size();
    hipStream_t s = ctx->stream;
    std::vector<P2Job> jobs((size_t)G);
    std::vector<SelArgs> sels((size_t)G);
    std::vector<std::array<int32_t, P_NPOL>> blocks((size_t)G);
    int passes = 0;
    bool any_heavy = false;
    bs.remaining.grow(sizeof(int32_t) * 2);
    for (int j = 0; j < G; ++j) {
        P2Item &it = *items[(size_t)j];
        Phase2Work &w = phase2_work(ctx, it.g, it.k);
        if (!it.have_job) {
            phase2_round_table(ctx, w, it.max_rounds);
            it.job = phase2_job(ctx, it.g, w, it.max_rounds, it.epsilon, nullptr);
            it.job.trace = nullptr;
            it.job.trace_cap = 0;
            it.have_job = true;
        }
        it.job.remaining = bs.remaining.as<int32_t>();
        DMOLP_REQUIRE(it.job.lh_on == (ctx->tune.light_hist ? 1 : 0), DMOLP_E_ARG,
                      "light_hist tuning changed while a job was in flight");
        jobs[(size_t)j] = it.job;
        sels[(size_t)j] = phase2_sel_args(ctx, it.g, w);
        job_blocks(ctx, it.g->n, it.g->nnz, it.k, it.g->n_heavy, G == 1, blocks[(size_t)j].data());
        passes = std::max(passes, w.sel.passes);
        any_heavy = any_heavy || it.g->n_heavy > 0;
    }
    launch_set_build(ctx, bs.ls, blocks);
    bs.jobs.grow(sizeof(P2Job) * (size_t)G);
    bs.sel.grow(sizeof(SelArgs) * (size_t)G);
    bs.done.grow(sizeof(int32_t) * (size_t)G);
    DMOLP_HIP_CHECK(hipMemcpyAsync(bs.jobs.p, jobs.data(), sizeof(P2Job) * (size_t)G,
                                   hipMemcpyHostToDevice, s));
    DMOLP_HIP_CHECK(hipMemcpyAsync(bs.sel.p, sels.data(), sizeof(SelArgs) * (size_t)G,
                                   hipMemcpyHostToDevice, s));
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));  // the host vectors go out of scope
    P2Set set;
    set.jobs = bs.jobs.as<P2Job>();
    set.sel = bs.sel.as<SelArgs>();
    set.ls = &bs.ls;
    set.passes = passes;
    set.any_heavy = any_heavy;
    set.lh_on = ctx->tune.light_hist;
    set.remaining = bs.remaining.as<int32_t>();
    if (key_out) {
        uint64_t key = 0x51ed270b27c1f1f3ull;
        for (const void *p : {bs.jobs.p, bs.sel.p, bs.ls.bstart.p, bs.remaining.p})
            key = mix(key, (uint64_t)(uintptr_t)p);
        key = mix(key, (uint64_t)G);
        key = mix(key, (uint64_t)passes);
        key = mix(key, any_heavy ? 1 : 0);
        key = mix(key, set.lh_on ? 1 : 0);
        for (int p = 0; p < P_NPOL; ++p) key = mix(key, (uint64_t)bs.ls.total[p]);
        *key_out = key;
    }
    return set;
}

void p2_collect(dmolp_ctx *ctx, std::vector<P2Item *> &items) {
    hipStream_t s = ctx->stream;
    const int G = (int)items.size();
    std::vector<Ctl> cs((size_t)G);
    OpBatch ob;  // the jobs' results in one copy
    std::vector<std::vector<int64_t>> cv((size_t)G), ci((size_t)G);
    for (int j = 0; j < G; ++j) {
        P2Item &it = *items[(size_t)j];
        Phase2Work &w = *it.g->p2;
        const int k = it.k;
        cv[(size_t)j].assign((size_t)k, 0);
        ci[(size_t)j].assign((size_t)k, 0);
        ob.download(&cs[(size_t)j], w.ctl.p, sizeof(Ctl));
        ob.download(cv[(size_t)j].data(), w.vc.p, sizeof(int64_t) * (size_t)k);
        ob.download(ci[(size_t)j].data(), w.ic.p, sizeof(int64_t) * (size_t)k);
    }
    ob.run(ctx);
    for (int j = 0; j < G; ++j) {
        P2Item &it = *items[(size_t)j];
        const Ctl &c = cs[(size_t)j];
        it.err = c.err;
        dmolp_phase2_metrics &m = it.m;
        std::memset(&m, 0, sizeof(m));
        final_balance(it.k, cv[(size_t)j], ci[(size_t)j], m.vertex_balance, m.edge_balance);
        int64_t tv = 0, te = 0;
        for (int i = 0; i < it.k; ++i) { tv += cv[(size_t)j][(size_t)i]; te += ci[(size_t)j][(size_t)i]; }
        m.edge_cut = (int64_t)c.v[C_PREV_CUT];
        m.total_vertices = tv;
        m.total_edges = te;
        m.rounds = c.rounds;
        m.exit_reason = it.max_rounds == 0 ? DMOLP_EXIT_MAX_ROUNDS : c.reason;
        m.lp_edges = (int64_t)c.v[C_LP_EDGES];
        m.initial_cut = (int64_t)c.v[C_INIT_CUT];
    }
}

void phase2_run_set(dmolp_ctx *ctx, std::vector<P2Item> &items, BatchState &bs) {
    const int G = (int)items.size();
    std::vector<P2Item *> ip((size_t)G);
    int Rmax = 0;
    for (int j = 0; j < G; ++j) {
        ip[(size_t)j] = &items[(size_t)j];
        items[(size_t)j].have_job = false;
        Rmax = std::max(Rmax, items[(size_t)j].max_rounds);
    }
    uint64_t key = 0;
    const P2Set set = p2_set_build(ctx, ip, bs, &key);
    phase2_init_enqueue(ctx, set);
    for (int t = 1; t <= Rmax; ++t) {
        phase2_round_launch(ctx, set, bs.round, key);
        if (t < Rmax && round_poll(ctx, set.remaining, t, true)) break;
    }
    p2_collect(ctx, ip);
}

void batch_run(dmolp_ctx *ctx, int32_t n_jobs, const dmolp_job *jobs, dmolp_result *results,
               double *device_ms) {
    for (int32_t i = 0; i < n_jobs; ++i) batch_check_job(jobs[i]);
    for (int32_t i = 0; i < n_jobs; ++i)
        for (int32_t j = 0; j < i; ++j)
            DMOLP_REQUIRE(jobs[i].graph != jobs[j].graph, DMOLP_E_ARG,
                          "a graph appears twice in one batch");
    std::memset(results, 0, sizeof(dmolp_result) * (size_t)n_jobs);
    if (n_jobs == 0) {
        if (device_ms) *device_ms = 0.0;
        return;
    }
    hipStream_t s = ctx->stream;
    hipEvent_t e0, e1;
    DMOLP_HIP_CHECK(hipEventCreate(&e0));
    DMOLP_HIP_CHECK(hipEventCreate(&e1));
    DMOLP_HIP_CHECK(hipEventRecord(e0, s));
    std::vector<P1Item> p1((size_t)n_jobs);
    for (int32_t i = 0; i < n_jobs; ++i) {
        p1[(size_t)i].g = jobs[i].graph;
        p1[(size_t)i].k = jobs[i].k;
        p1[(size_t)i].theta = jobs[i].theta;
    }
    // The phase-2 work arrays and degree bins depend on the graphs only: they
    // are set up while the host searches the hubs.
    phase1_run_set(ctx, p1, [&] {
        for (int32_t i = 0; i < n_jobs; ++i) {
            Phase2Work &w = phase2_work(ctx, jobs[i].graph, jobs[i].k);
            phase2_round_table(ctx, w, jobs[i].max_rounds);
        }
    });
    std::vector<P2Item> p2((size_t)n_jobs);
    for (int32_t i = 0; i < n_jobs; ++i) {
        p2[(size_t)i].g = jobs[i].graph;
        p2[(size_t)i].k = jobs[i].k;
        p2[(size_t)i].max_rounds = jobs[i].max_rounds;
        p2[(size_t)i].epsilon = jobs[i].epsilon;
    }
    phase2_run_set(ctx, p2, ctx->batch);
    DMOLP_HIP_CHECK(hipEventRecord(e1, s));
    for (int32_t i = 0; i < n_jobs; ++i)
        if (jobs[i].labels_out && p2[(size_t)i].err == 0)
            DMOLP_HIP_CHECK(hipMemcpyAsync(jobs[i].labels_out, jobs[i].graph->labels.p,
                                           sizeof(int32_t) * (size_t)jobs[i].graph->n,
                                           hipMemcpyDeviceToHost, s));
    DMOLP_HIP_CHECK(hipEventSynchronize(e1));
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    DMOLP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int32_t i = 0; i < n_jobs; ++i) {
        dmolp_result &r = results[i];
        const P2Item &b = p2[(size_t)i];
        r.status = b.err;
        r.rounds = b.m.rounds;
        r.exit_reason = b.m.exit_reason;
        r.n_seeds_fallback = p1[(size_t)i].m.n_seeds_fallback;
        r.phase1_cut = b.m.initial_cut;
        r.edge_cut = b.m.edge_cut;
        r.lp_edges = b.m.lp_edges;
        r.vertex_balance = b.m.vertex_balance;
        r.edge_balance = b.m.edge_balance;
    }
    if (device_ms) *device_ms = ms;
}

void batch_autotune(dmolp_ctx *ctx, int32_t n_jobs, const dmolp_job *jobs, int32_t max_trials,
                    dmolp_tuning *best_out) {
    DMOLP_REQUIRE(n_jobs > 0, DMOLP_E_ARG, "autotune needs at least one job");
    // Candidates: degree-bin threshold x chunk length x blocks per unit of
    // work.  Every candidate must give the same results; the fastest
    // (device time of the whole batch) is kept in the context.
    const int heavy[] = {256, 1024, 4096};
    const int chunk[] = {2048, 4096};
    const int64_t per[] = {8192, 32768, 131072};
    const Tuning orig = ctx->tune;
    std::vector<dmolp_result> ref((size_t)n_jobs), cur((size_t)n_jobs);
    std::vector<dmolp_job> js(jobs, jobs + n_jobs);
    for (auto &j : js) j.labels_out = nullptr;
    double best_ms = -1.0;
    Tuning best = orig;
    int trials = 0;
    bool have_ref = false;
    try {
        for (int h : heavy)
            for (int c : chunk)
                for (int64_t p : per) {
                    if (max_trials > 0 && trials >= max_trials) break;
                    ctx->tune.heavy_degree = h;
                    ctx->tune.chunk_edges = c;
                    ctx->tune.batch_work_per_block = p;
                    double ms = 0.0;
                    batch_run(ctx, n_jobs, js.data(), cur.data(), &ms);  // warm (allocations, capture)
                    batch_run(ctx, n_jobs, js.data(), cur.data(), &ms);
                    ++trials;
                    if (!have_ref) {
                        ref = cur;
                        have_ref = true;
                    } else {
                        for (int32_t i = 0; i < n_jobs; ++i)
                            DMOLP_REQUIRE(cur[(size_t)i].edge_cut == ref[(size_t)i].edge_cut &&
                                              cur[(size_t)i].rounds == ref[(size_t)i].rounds &&
                                              cur[(size_t)i].lp_edges == ref[(size_t)i].lp_edges,
                                          DMOLP_E_HIP, "a tuning candidate changed a result");
                    }
                    if (best_ms < 0 || ms < best_ms) {
                        best_ms = ms;
                        best = ctx->tune;
                    }
                }
    } catch (...) {
        ctx->tune = orig;
        throw;
    }
    ctx->tune = best;
    if (best_out) {
        std::memset(best_out, 0, sizeof(*best_out));
        best_out->heavy_degree = best.heavy_degree;
        best_out->chunk_edges = best.chunk_edges;
        best_out->block_threads = best.block_threads;
        best_out->grid_blocks = ctx->grid_blocks();
        best_out->pull_per_mille = (int32_t)(best.pull_fraction * 1000.0 + 0.5);
        best_out->batch_work_per_block = (int32_t)best.batch_work_per_block;
    }
}

}  // namespace dmolp
