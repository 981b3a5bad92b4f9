// batch.hip -- many graphs partitioned in one call.
#include <cstring>
#include <vector>

#include "internal.hpp"

extern "C" int dmolp_batch_run(dmolp_ctx *ctx, int32_t n_jobs, const dmolp_job *jobs,
                               dmolp_result *results, double *device_ms) {
    if (!ctx || n_jobs < 0 || (n_jobs > 0 && (!jobs || !results))) return DMOLP_E_ARG;
    double total_ms = 0.0;
    int first_err = DMOLP_OK;
    for (int32_t i = 0; i < n_jobs; ++i) {
        const dmolp_job &j = jobs[i];
        dmolp_result &r = results[i];
        std::memset(&r, 0, sizeof(r));
        dmolp_phase1_metrics m1{};
        int rc = dmolp_phase1(ctx, j.graph, j.k, j.theta, nullptr, nullptr, nullptr, nullptr, &m1);
        dmolp_phase2_metrics m2{};
        if (rc == DMOLP_OK)
            rc = dmolp_phase2(ctx, j.graph, j.k, j.max_rounds, j.epsilon, nullptr, nullptr, &m2,
                              nullptr);
        if (rc == DMOLP_OK && j.labels_out) rc = dmolp_graph_get_labels(ctx, j.graph, j.labels_out);
        r.status = rc;
        if (rc != DMOLP_OK) {
            if (first_err == DMOLP_OK) first_err = rc;
            continue;
        }
        r.rounds = m2.rounds;
        r.exit_reason = m2.exit_reason;
        r.n_seeds_fallback = m1.n_seeds_fallback;
        r.phase1_cut = m1.initial_edge_cut;
        r.edge_cut = m2.edge_cut;
        r.lp_edges = m2.lp_edges;
        r.vertex_balance = m2.vertex_balance;
        r.edge_balance = m2.edge_balance;
        total_ms += m1.device_ms + m2.device_ms;
    }
    if (device_ms) *device_ms = total_ms;
    return first_err;
}
