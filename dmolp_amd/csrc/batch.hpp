// batch.hpp -- job sets: many graphs advancing through phase 1 and phase 2
// together, one launch per kernel step.
#pragma once
#include <vector>

#include "internal.hpp"
#include "phase2.hpp"

namespace dmolp {

// One job of a phase-2 set: inputs and results.
struct P2Item {
    dmolp_graph *g = nullptr;
    int k = 0;
    int max_rounds = 0;
    double epsilon = 0.0;
    int err = 0;  // DMOLP_E_* raised on the device, 0 when fine
    dmolp_phase2_metrics m{};
    P2Job job{};  // device record, made when the item enters phase 2
    bool have_job = false;
};

// Upload a set of phase-2 items (records made on first use); key_out
// (nullable) receives the launch-shape key for round capture.
P2Set p2_set_build(dmolp_ctx *ctx, std::vector<P2Item *> &items, BatchState &bs, uint64_t *key_out);
// Final metrics of finished items (phase2.cpp:418-441).
void p2_collect(dmolp_ctx *ctx, std::vector<P2Item *> &items);
// Phase 2 for every item on its graph's resident labels (state in bs).
void phase2_run_set(dmolp_ctx *ctx, std::vector<P2Item> &items, BatchState &bs);
void batch_check_job(const dmolp_job &j);
// dmolp_batch_run behind the ABI guard.
void batch_run(dmolp_ctx *ctx, int32_t n_jobs, const dmolp_job *jobs, dmolp_result *results,
               double *device_ms);
// Try launch parameters on the batch, keep the fastest in ctx.
void batch_autotune(dmolp_ctx *ctx, int32_t n_jobs, const dmolp_job *jobs, int32_t max_trials,
                    dmolp_tuning *best_out);

// Served mode (server.hip).
dmolp_server *server_create(dmolp_ctx *ctx, int64_t budget, int32_t max_active);
double server_clock(const dmolp_server *s);
int64_t server_submit(dmolp_server *s, const dmolp_request &r);
void server_step(dmolp_server *s, int32_t *active_out, int32_t *queued_out);
int server_take(dmolp_server *s, int64_t ticket, dmolp_served *out, int32_t *labels_out);
void server_destroy(dmolp_server *s);

}  // namespace dmolp
