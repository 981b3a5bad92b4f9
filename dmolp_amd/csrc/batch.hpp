// batch.hpp -- job sets: many graphs advancing through phase 1 and phase 2
// together, one launch per kernel step.
#pragma once
#include <vector>

#include "internal.hpp"

namespace dmolp {

// One job of a phase-2 set: inputs and results.
struct P2Item {
    dmolp_graph *g = nullptr;
    int k = 0;
    int max_rounds = 0;
    double epsilon = 0.0;
    int err = 0;  // DMOLP_E_* raised on the device, 0 when fine
    dmolp_phase2_metrics m{};
};

// Phase 2 for every item on its graph's resident labels (state in bs).
void phase2_run_set(dmolp_ctx *ctx, std::vector<P2Item> &items, BatchState &bs);
void batch_check_job(const dmolp_job &j);
// dmolp_batch_run behind the ABI guard.
void batch_run(dmolp_ctx *ctx, int32_t n_jobs, const dmolp_job *jobs, dmolp_result *results,
               double *device_ms);

}  // namespace dmolp
