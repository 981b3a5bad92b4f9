/*
 * ref_cpu.h -- CPU oracle for the DMOLP two-phase partitioner, single rank.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links,
 * loads or calls it.
 *
 * It restates the reference's single-rank semantics (one MPI rank,
 * OMP_NUM_THREADS=1) with four documented decisions:
 *   - strict IEEE double arithmetic in source order, no contraction
 *     (the reference's Release build uses -ffast-math, CMakeLists.txt:33);
 *   - the unchunked GPU path (gpu_lp_boundary.cu:558-560), never the chunked
 *     one (gpu_lp_boundary.cu:571-628), which drops cross-chunk edges;
 *   - int64 offsets and counters where the reference's int would overflow
 *     (graph_types.h:127-130, phase2.cpp:112-113, phase1.cpp:269);
 *   - max_rounds and epsilon as parameters instead of the constants at
 *     phase2.cpp:263-264, and a documented seed fallback where the reference
 *     has undefined behaviour (seed.cpp:214-217 then partition.cpp:347).
 *
 * Graph input everywhere: symmetric CSR, n vertices, off[n+1] (int64),
 * col[off[n]] (int32), rows sorted ascending, no duplicates, no self-loops.
 */
#ifndef DMOLP_REF_CPU_H
#define DMOLP_REF_CPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared with the product ABI's meaning. */
#define REF_OK 0
#define REF_E_ARG (-1)
#define REF_E_SEEDS (-2)
#define REF_E_INVALID_LABEL (-3)
#define REF_E_ODD_CUT (-4)

typedef struct {
    int64_t initial_edge_cut;
    double initial_vertex_balance;
    double initial_edge_balance;
    int64_t total_vertices;
    int64_t total_edges;      /* nnz / 2 */
    int32_t first_seed;
    int32_t first_seed_degree;
    int64_t n_hubs;
    int32_t n_seeds_found;    /* seeds found by the hub search (<= k) */
    int32_t n_seeds_fallback; /* seeds added by the documented extension */
    int64_t rounds;           /* expansion rounds run */
    int32_t fill_round;       /* round index where the fill step fired, -1 if never */
    int32_t reserved;
} ref_p1_metrics;

typedef struct {
    int64_t edge_cut;
    double vertex_balance;
    double edge_balance;
    int64_t total_vertices;
    int64_t total_edges;
    int32_t rounds;       /* 1-based index of the last round started */
    int32_t exit_reason;  /* 0 max rounds, 1 empty boundary, 2 no change, 3 converged */
    int64_t lp_edges;     /* sum over rounds of sum_{b in B_t} deg b */
    int64_t initial_cut;
} ref_p2_metrics;

/* Per-round trace, each array of length max_rounds (entries past `rounds`
 * untouched).  Any pointer may be NULL. */
typedef struct {
    int64_t *prev_count;   /* |B_{t-1}| */
    int64_t *prev_deg;     /* sum deg over B_{t-1} */
    int64_t *cand_count;   /* |Cand_t| */
    int64_t *scan_len;     /* sum over Cand_t of p_x (neighbours scanned up to and incl. first differing label, or deg if none) */
    int64_t *bd_count;     /* |Bd_t| boundary vertices before selection */
    int64_t *b_count;      /* |B_t| */
    int64_t *b_deg;        /* sum deg over B_t */
    int64_t *chg_count;    /* |Delta_t| */
    int64_t *chg_deg;      /* sum deg over Delta_t */
    int64_t *cut;          /* cut after round t (or -1 when the round stopped before stats) */
} ref_p2_trace;

/* OpenMP threads for the parallel sweeps (results do not depend on it). */
void ref_set_threads(int32_t t);
int32_t ref_get_threads(void);

/* ---- Phase 1 pieces (init.cpp, seed.cpp, partition.cpp, phase1.cpp) ---- */

/* Iteration order of the degree map the reference builds at one rank
 * (init.cpp:43-44,68,90-93 then init.cpp:102-136).  Writes n ids. */
int ref_probe_degree_map_order(int32_t n, int32_t *order_out);

/* A.1 first seed, init.cpp:179-187. */
int ref_first_seed(int64_t n, const int64_t *off, int32_t *seed_out, int32_t *deg_out);

/* A.2 hubs, init.cpp:139-177.  hubs_out has capacity n. */
int ref_hubs(int64_t n, const int64_t *off, int32_t *hubs_out, int64_t *n_hubs_out);

/* A.3 seeds, seed.cpp:138-242, plus the documented fallback when fewer than
 * k seeds exist.  seeds_out has capacity k. */
int ref_seeds(int64_t n, const int64_t *off, const int32_t *col, int32_t k,
              int32_t *seeds_out, int32_t *n_found_out, int32_t *n_fallback_out);

/* Depth-<=2 ball bitmap of one vertex (seed.cpp:20-101); writes n bytes 0/1. */
int ref_ball2(int64_t n, const int64_t *off, const int32_t *col, int32_t s, uint8_t *ball_out);

/* A.4 expansion, partition.cpp:323-604.  labels_out has n entries. */
int ref_expand(int64_t n, const int64_t *off, const int32_t *col, int32_t k, int32_t theta,
               const int32_t *seeds, int64_t *rounds_out, int32_t *fill_round_out,
               int32_t *labels_out);

/* A.5 metrics, phase1.cpp:72-153.  vc, ec have k entries. */
int ref_phase1_metrics(int64_t n, const int64_t *off, const int32_t *col, int32_t k,
                       const int32_t *labels, int64_t *vc, int64_t *ec, ref_p1_metrics *m);

/* Whole phase 1 (phase1.cpp:18-161).  labels_out: n; vc, ec: k (nullable). */
int ref_phase1(int64_t n, const int64_t *off, const int32_t *col, int32_t k, int32_t theta,
               int32_t *labels_out, int32_t *seeds_out, int64_t *vc, int64_t *ec,
               ref_p1_metrics *m);

/* ---- Phase 2 pieces (phase2.cpp, gpu_lp_boundary.cu) ---- */

/* A.6 stats + cut, phase2.cpp:55-117 and 214-253.  vc, ic: k entries. */
int ref_stats(int64_t n, const int64_t *off, const int32_t *col, int32_t k,
              const int32_t *labels, int64_t *vc, int64_t *ic, int64_t *cut_out);

/* A.7 penalties, phase2.cpp:158-209.  pen, re: k entries; t >= 1. */
void ref_penalties(int32_t k, const int64_t *vc, const int64_t *ic, int32_t t,
                   double *pen, double *re);

/* A.8 boundary selection, gpu_lp_boundary.cu:185-305.  out: capacity n,
 * written in ascending id order. */
int ref_select_boundary(int64_t n, const int64_t *off, const int32_t *col,
                        const int32_t *labels, const int32_t *prev, int64_t n_prev,
                        const double *re, int32_t k, int32_t t,
                        int32_t *out, int64_t *n_out);

/* A.9 decision for each b in B (gpu_lp_boundary.cu:327-427), Jacobi:
 * labels_out must hold a copy of labels on entry; only B entries change. */
int ref_lp_round(int64_t n, const int64_t *off, const int32_t *col, const int32_t *labels,
                 const int32_t *b, int64_t nb, const double *pen, int32_t k,
                 int32_t *labels_out);

/* Whole phase 2 (phase2.cpp:256-449).  labels: in/out.  vc, ic: k (nullable). */
int ref_phase2(int64_t n, const int64_t *off, const int32_t *col, int32_t k,
               int32_t max_rounds, double epsilon, int32_t *labels,
               int64_t *vc, int64_t *ic, ref_p2_metrics *m, ref_p2_trace *trace);

/* Phase 2 run as P ranks on their shares of the graph (owner(v) = v mod P),
 * ghost labels refreshed from the allgathered deltas each round.  scope 0:
 * each rank selects over its own boundary vertices (the reference);
 * scope 1: selection over all ranks (equals ref_phase2 for every P).
 * labels: global, in/out.  ghost_updates (nullable): per round, the ghost
 * labels the refresh rewrote, summed over the ranks.  The trace's
 * cand_count and scan_len are -1. */
int ref_phase2_ranks(int64_t n, const int64_t *off, const int32_t *col, int32_t k, int32_t P,
                     int32_t scope, int32_t max_rounds, double epsilon, int32_t *labels,
                     int64_t *vc, int64_t *ic, ref_p2_metrics *m, ref_p2_trace *trace,
                     int64_t *ghost_updates);

#ifdef __cplusplus
}
#endif
#endif
