// phase1.hpp -- device state of the hub-seed expansion phase.
#pragma once
#include <utility>
#include <functional>
#include <vector>

#include "hubs.hpp"
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
    int32_t seed_cursor;  // hubs before it are used or covered
};

struct Phase1Work {
    int k = 0;
    int64_t cap_heavy = -1, cap_chunks = -1;
    DevMem claims, touched, touched_cnt, covered, deg, ctl, best, sizes, seeds;
    DevMem lab;  // bit v: vertex v is labelled (the expansion's test, L2-resident)
    DevMem hubbits;  // bit h: h is a listed hub (the only vertices the balls mark)
    DevMem hubs, used;  // hub list of the current run and its used flags
    Recs F[2];
    List sel_list, lvl1;
    SelectWork sel;
};

// Everything one graph's phase-1 kernels read and write (device pointers).
struct P1Job {
    GraphDev g;
    ListDev lvl1, sel_list;
    RecDev F[2];
    int32_t *L, *claims, *touched;
    int64_t *touched_cnt;
    uint32_t *covered;
    uint32_t *lab;  // labelled-vertex bitmap, kept in step with L
    uint32_t *hubbits;  // listed hubs
    int32_t *deg;
    unsigned long long *best, *sizes;
    int32_t *seeds;
    const int32_t *hubs;
    uint8_t *used;
    int64_t nh;
    P1Ctl *ctl;
    int32_t *remaining;  // unfinished jobs of the set (shared by the set)
    int32_t k;
    int32_t theta;       // frontier members a part expands per round
    int32_t seed_steps;  // steps of the hub search (k, or 1 when there are no hubs)
    int32_t heavy_degree, chunk_edges;
};

// One job of a phase-1 set: inputs and results.
struct P1Item {
    dmolp_graph *g = nullptr;
    int k = 0;
    int theta = 0;
    std::vector<int32_t> seeds;
    dmolp_phase1_metrics m{};
};

// Host state of one job while it is in phase 1.
struct P1Stage {
    dmolp_graph *g = nullptr;
    Phase1Work *w = nullptr;
    int k = 0, theta = 0;
    // deg: the degrees on the host (host hub searches, the seed fallback);
    // hubs: the first of the n_hubs hubs, hub_deg their degrees.
    std::vector<int32_t> deg, hubs, hub_deg, seeds;
    int64_t n_hubs = 0;
    int64_t hub_prefix = 0;  // hubs ordered by the first search (0: all)
    bool dev_hubs = false;   // hub search on the device (p1_hub_search_device)
    std::vector<int32_t> top;  // first 2k vertices by (degree desc, id asc)
    std::vector<uint8_t> used;
    std::pair<int32_t, int32_t> fs{-1, -1};
    double host_ms = 0.0;
    int32_t found = 0, fallback = 0;
    P1Job job{};
    bool have_job = false;
};

// Stages of phase 1 for a set of jobs (phase1_run_set runs them in order;
// the serving loop interleaves them with other jobs' work).
void p1_stage_init(dmolp_ctx *ctx, P1Item &it, P1Stage &h);
void p1_set_build(dmolp_ctx *ctx, std::vector<P1Stage *> &H, P1SetDev &sd);
void p1_degrees(dmolp_ctx *ctx, std::vector<P1Stage *> &H, P1SetDev &sd);  // degrees to the host
void p1_hub_search(P1Stage &h, int threads = 1);       // host: hubs, first seed (dev_hubs: none)
void p1_hub_search_device(dmolp_ctx *ctx, std::vector<P1Stage *> &H);  // the jobs with dev_hubs
void p1_seeds(dmolp_ctx *ctx, std::vector<P1Stage *> &H, P1SetDev &sd);    // hub-ball seeds + fallback
void p1_expand_begin(dmolp_ctx *ctx, std::vector<P1Stage *> &H, P1SetDev &sd);
void p1_expand_rounds(dmolp_ctx *ctx, P1SetDev &sd, int rounds);          // rounds: even
// Rounds that need no selection, one workgroup per job (k_p1_rounds_wg).
void p1_rounds_wg(dmolp_ctx *ctx, P1SetDev &sd, int max_rounds, int &ran, int &stopped);
int32_t p1_remaining(dmolp_ctx *ctx, P1SetDev &sd);
void p1_done_flags(dmolp_ctx *ctx, P1SetDev &sd, std::vector<int32_t> &out);
// ctl: the job's control block when the caller has already downloaded it.
void p1_finish(dmolp_ctx *ctx, P1Stage &h, dmolp_phase1_metrics &m, double device_ms,
               const P1Ctl *ctl = nullptr);

Phase1Work &phase1_work(dmolp_ctx *ctx, dmolp_graph *g, int k);
// Seeds + expansion; labels become resident in g.  Metrics fields about
// seeds, rounds and timing are filled (cut and counts come from the stats
// sweep).
void phase1_random(dmolp_ctx *ctx, dmolp_graph *g, int k);
void phase1_run(dmolp_ctx *ctx, dmolp_graph *g, int k, int theta, std::vector<int32_t> &seeds_out,
                dmolp_phase1_metrics *m);
// The same for a set of graphs advancing together: one launch per kernel
// step serves every job.  Each job's result equals its plain run.
// during_hubs (may be empty) runs on the calling thread while the host hub
// searches run on worker threads: device work queued there overlaps them.
void phase1_run_set(dmolp_ctx *ctx, std::vector<P1Item> &items,
                    const std::function<void()> &during_hubs);

}  // namespace dmolp
