// phase2.hpp -- device state of the boundary label-propagation phase.
#pragma once
#include "internal.hpp"
#include "select.hpp"

namespace dmolp {

// 64-bit control counters (device).
enum CtlIndex {
    C_CUT = 0,       // current cut (undirected)
    C_PREV_CUT,      // cut of the previous completed round
    C_INIT_CUT,      // cut before round 1
    C_CAND,          // first-touched candidates this round
    C_LP_EDGES,      // sum over rounds of sum deg(B_t)
    C_INVALID,       // entries with a -1 label (initial sweep)
    C_CUT2,          // differing entries (initial sweep)
    C_PREV_N,        // |B_{t-1}|
    C_PREV_DEG,      // sum deg over B_{t-1}
    C_KERNEL_ERR,    // a kernel found an impossible state
    C_DEFERRED,      // heavy rows left to the pull kernel's second pass
    C_PENDING,       // rows the pull pass left to k_cand_pull_rest
    C_BADLAB,        // vertices whose label lies outside [0, k) (packing pass)
    C_NCTL = 16
};

// How a round finds its candidates: every vertex (round 1, B_0 = V),
// pulled from the boundary side, or pushed from B_{t-1}'s rows.
enum CandMode : int32_t { CAND_ALL = 0, CAND_PULL = 1, CAND_PUSH = 2 };
// How the heavy-row histograms follow a round's changes: scan the rows
// against the changed-vertex bitmap, recount them, or push the changes.
enum HistMode : int32_t { HIST_SCAN = 0, HIST_RECOUNT = 1, HIST_PUSH = 2 };

constexpr int kHhCopies = 8;

struct Ctl {
    unsigned long long v[C_NCTL];
    int32_t done;
    int32_t reason;
    int32_t rounds;
    int32_t cc;
    int32_t err;  // DMOLP_E_* raised on the device
    int32_t cand_mode;  // CAND_ALL (round 1), CAND_PULL or CAND_PUSH
    int32_t t;           // current round, 1-based
    int32_t epoch;       // stamp value of the current round (epoch_base + t)
    int32_t epoch_base;
    int32_t max_rounds;
    int32_t hist_mode;   // HistMode of this round
    int32_t diff_ok;     // diff[v] holds v's count of differing neighbours
    double epsilon;
};

// Trace columns written by the round-end kernel.
enum TraceCol {
    T_PREV_N = 0, T_PREV_DEG, T_CAND, T_BD, T_B, T_BDEG, T_CHG, T_CHGDEG, T_CUT, T_BL, T_BLDEG,
    T_DL, T_DLDEG, T_NCOL
};

// Everything one graph's phase-2 kernels read and write (device pointers).
struct P2Job {
    GraphDev g;
    ListDev all, B, D;
    int32_t *L;
    uint8_t *L8;
    uint32_t *L5;   // packed labels (lbits bits, 32 / lbits a word) for the initial sweep
    int32_t *chg, *diff, *stamp, *hub_hist;
    int64_t *deferred;  // long rows the pull pass left undecided: (flags << 32) | v
    int32_t *rdeg;    // dense selection records: degree of each record vertex, else -1
    uint32_t *bbits;  // two bitmaps of n bits: B_t packed into half (t & 1)
    int64_t bwords;   // words per bitmap
    // Heavy rows keep their label histograms across rounds: hh[hid(v)][l]
    // counts v's neighbours with label l < min(k, 32), updated after each
    // round from the vertices that changed (bit set in cbits; old and new
    // byte labels in ON as old | new << 8).
    int32_t *hh;
    // Push-mode moves go to kHhCopies copies of hh (copy blockIdx % copies):
    // the highest-degree rows take thousands of moves a round, and atomics
    // on one cache line serialise.  k_hist_fold adds the copies into hh.
    int32_t *hhc;
    const int32_t *hid;  // vertex -> heavy row index of the graph, -1 for light rows
    uint32_t *cbits;
    uint16_t *ON;
    uint16_t *LN;        // label byte | this round's new label byte << 8 (255: unchanged)
    // Light rows keep their label counts too (when lh_on): row v of lh holds
    // kLpCap / 2 words, the count of label l < min(k, 32) in the 16-bit half
    // l % 2 of word l / 2.  Built by the initial sweep; after each round's
    // decisions every changed vertex moves one count of each light
    // neighbour's row from its old label to its new one.
    uint32_t *lh;
    const uint32_t *hbits;  // heavy-row bitmap of the graph
    const int32_t *first;   // each row's first (smallest) neighbour, -1 if empty
    const int32_t *deg;     // each row's degree
    int32_t *wit;           // per run: a neighbour that last put the row in the candidates
    RecDev bd;
    Ctl *ctl;
    unsigned long long *vc, *ic;
    double *pen, *re;
    const double *rtab;
    int64_t *trace;
    int32_t *remaining;  // unfinished jobs of the set (shared by the set)
    int32_t trace_cap, k, heavy_degree, chunk_edges;
    int32_t epoch_base, max_rounds;
    int32_t lh_on, lbits;  // lbits: bits of a packed label, 0 without the packed mirror
    double epsilon;
    double pull_frac;
    double recount_share, push_share;  // heavy-histogram modes (see k_hist_mode)
};

// A set of jobs that advance together, one launch per kernel step.
struct P2Set {
    const P2Job *jobs = nullptr;    // device
    const SelArgs *sel = nullptr;   // device, one per job
    const LaunchSet *ls = nullptr;
    int passes = 0;                 // largest selection digit-pass count
    bool any_heavy = false;
    bool lh_on = false;             // light rows decide from kept histograms
    int32_t *remaining = nullptr;   // device counter of unfinished jobs
};

struct Phase2Work {
    int k = 0;
    int64_t cap_heavy = -1, cap_chunks = -1;  // heavy capacities the lists were sized for
    int32_t epoch = 0;  // stamp value of the current candidate pass
    DevMem chg, diff, stamp, rdeg, bbits, deferred, wit;
    DevMem L8;  // byte mirror of the labels for the LP gathers
    DevMem L5;  // packed labels for the initial sweep (k <= 32)
    DevMem hh, hhc, cbits, ON, LN;
    DevMem lh;  // uint32 [n][kLpCap / 2]: kept light-row label counts
    List B, D;          // boundary list (B_{t-1} read, then B_t written), Delta
    DevMem rtab;        // per-round constants: [0..R] 1 + log(t), [R+1..2R+1] exp(-t/5)
    int rtab_rounds = -1;
    RoundGraph round;  // one captured round of this graph as a job set of one
    ~Phase2Work() { round.release(); }  // before any buffer its kernels read is freed
    // This graph as a job set of one (the plain path).
    DevMem job_dev, sel_dev, remaining;
    LaunchSet ls;
    P2Job job_host{};
    SelArgs sel_host{};
    Recs bd;  // record counter of the round (the records themselves are dense, in rdeg)
    SelectWork sel;
    DevMem hub_hist, hub_done;
    DevMem ctl;
    DevMem vc, ic;  // unsigned long long [k]
    DevMem pen, re; // double [k]
    DevMem trace;   // int64 [trace_cap * T_NCOL]
    int trace_cap = 0;
};

Phase2Work &phase2_work(dmolp_ctx *ctx, dmolp_graph *g, int k);
// Per-round constant table for rounds 1..R (host libm, as the reference).
void phase2_round_table(dmolp_ctx *ctx, Phase2Work &w, int R);

// The job record of g for a run of at most max_rounds rounds (advances the
// stamp epoch).  `remaining` is the set's counter of unfinished jobs.
P2Job phase2_job(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w, int max_rounds, double epsilon,
                 int32_t *remaining);
SelArgs phase2_sel_args(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w);
// g as a job set of one, uploaded; returns the set view.
P2Set phase2_single_set(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w, int max_rounds,
                        double epsilon);

// Enqueue: reset state, full sweep of the labels (vc, ic, cut, per-vertex
// differing-neighbour counts).  Labels are each graph's resident labels.
void phase2_init_enqueue(dmolp_ctx *ctx, const P2Set &set);
// Enqueue one LP round of every job; the round index lives in each job's
// control block, so the same enqueue (or its captured graph) serves every
// round.
void phase2_round_enqueue(dmolp_ctx *ctx, const P2Set &set);
// Launch one round through a captured HIP graph (captured on first use,
// cached in *exec under `key`).
void phase2_round_launch(dmolp_ctx *ctx, const P2Set &set, RoundGraph &rg, uint64_t key);
// Every job's done flag (d_out: device scratch of one int per job).
void phase2_done_flags(dmolp_ctx *ctx, const P2Set &set, int32_t *d_out, std::vector<int32_t> &out);
// Shape key of the single-graph set for phase2_round_launch.
uint64_t phase2_single_key(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w);

}  // namespace dmolp
