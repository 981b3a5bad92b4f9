// noalloc.hpp -- the no-allocation statistics entry point (noalloc.hip).
#pragma once
#include "internal.hpp"

namespace dmolp {

// Takes the address of the scratch array on `device` (current), once.
void noalloc_init_device(int device);
// computePartitionStats + computeEdgeCut of host labels through the scratch
// array: no device allocation.
void stats_noalloc(dmolp_ctx *ctx, const dmolp_graph *g, int32_t k, const int32_t *labels,
                   int64_t *vc, int64_t *ic, int64_t *cut);

}  // namespace dmolp
