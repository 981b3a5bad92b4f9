// lp_rule.hpp -- the LP decision rule of one vertex (gpu_lp_boundary.cu:
// 388-421), shared by the phase-2 kernels and the rank-share kernels.
#pragma once
#include "dev_common.hpp"

namespace dmolp {

// Scores h*(1+pen) over labels below E; among equal best scores the
// smallest 5-bit bit-reversed label wins (the order the 32-lane
// shuffle-down tree yields); the winner must beat the raw own-label count.
template <class Count>
__device__ __forceinline__ int32_t lp_decide_by(Count count, int E, const double *pen, int32_t own) {
    const double base = (own >= 0 && own < E) ? (double)count(own) : -1e300;
    int best_l = 0;
    double best_s = 0.0;
    for (int l = 0; l < E; ++l) {
        double s = (double)count(l);
        if (s > 0.0) s *= (1.0 + pen[l]);
        if (l == 0 || s > best_s || (s == best_s && bitrev5(l) < bitrev5(best_l))) {
            best_l = l;
            best_s = s;
        }
    }
    return (best_s > base) ? best_l : own;
}

}  // namespace dmolp
