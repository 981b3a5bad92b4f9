// phase1.hpp -- device state of the hub-seed expansion phase.
#pragma once
#include <utility>
#include <vector>

#include "internal.hpp"
#include "select.hpp"

namespace dmolp {

struct P1Ctl {
    int32_t done;
    int32_t seed_stop;    // the hub search found no further seed
    int32_t seeds_found;
    int32_t fill_part;    // partition receiving the fill this round, -1 if none
    int64_t labelled;
    int64_t rounds;
    int32_t fill_round;
    int32_t pad;
};

struct Phase1Work {
    int k = 0;
    int64_t cap_heavy = -1, cap_chunks = -1;
    DevMem claims, touched, touched_cnt, covered, deg, ctl, best, sizes, seeds;
    Recs F[2];
    List sel_list, lvl1;
    SelectWork sel;
};

Phase1Work &phase1_work(dmolp_ctx *ctx, dmolp_graph *g, int k);
std::vector<int32_t> find_hubs_host(const std::vector<int32_t> &deg);
std::pair<int32_t, int32_t> first_seed_host(const std::vector<int32_t> &deg);
// Seeds + expansion; labels become resident in g.  Metrics fields about
// seeds, rounds and timing are filled (cut and counts come from the stats
// sweep).
void phase1_random(dmolp_ctx *ctx, dmolp_graph *g, int k);
void phase1_run(dmolp_ctx *ctx, dmolp_graph *g, int k, int theta, std::vector<int32_t> &seeds_out,
                dmolp_phase1_metrics *m);

}  // namespace dmolp
