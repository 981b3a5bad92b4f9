/*
 * dmolp_hip.h -- C ABI of the MI355X (gfx950) implementation of DMOLP's
 * two-phase partitioner path.
 *
 * The reference exposes its path as C++ functions; each entry point below
 * names the reference interface it replaces.  The C++ wrappers with the
 * reference's own signatures (run_phase1 / run_phase2 over graph_types.h)
 * live in include/phase1/phase1.h and include/phase2/phase2.h and call this
 * ABI.
 *
 * Conventions
 *  - Every function returns DMOLP_OK (0) or a negative dmolp_status; no C++
 *    exception crosses the ABI.  dmolp_last_error(ctx) gives the message of
 *    the last failure on that context.
 *  - Host buffers are owned by the caller.  Device memory is owned by the
 *    context and graph handles.
 *  - One context per host thread; a context is not re-entrant.  Several
 *    contexts may share a device.
 *  - Graph layout: symmetric CSR, rows sorted ascending with no duplicates
 *    and no self-loops; offsets int32 or int64 (offset_bytes 4 or 8),
 *    columns int32, labels int32.  This is graph_types.h's Graph
 *    (row_ptr / col_indices / vertex_labels, graph_types.h:126-132) with
 *    offsets widened to int64 where int32 would overflow.
 */
#ifndef DMOLP_HIP_H
#define DMOLP_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMOLP_ABI_VERSION 1

typedef enum {
    DMOLP_OK = 0,
    DMOLP_E_ARG = -1,           /* invalid argument */
    DMOLP_E_SEEDS = -2,         /* k larger than the vertex count */
    DMOLP_E_INVALID_LABEL = -3, /* a -1 label reached the cut (phase2.cpp:239-244) */
    DMOLP_E_ODD_CUT = -4,       /* odd cut: asymmetric CSR (phase2.cpp:246-250) */
    DMOLP_E_HIP = -5,           /* HIP runtime failure */
    DMOLP_E_NOMEM = -6,         /* device or host allocation failed */
    DMOLP_E_UNSUPPORTED = -7,   /* valid request outside the library's limits */
    DMOLP_E_STATE = -8,         /* call out of order (e.g. no labels resident) */
    DMOLP_E_GRAPH = -9,         /* CSR fails validation (self-loop, unsorted row, bad offset) */
    DMOLP_E_INTERNAL = -10      /* a device-side consistency check failed (a library fault) */
} dmolp_status;

/* Phase-2 exit reasons (phase2.cpp:309-312, 387-390, 409-415). */
#define DMOLP_EXIT_MAX_ROUNDS 0
#define DMOLP_EXIT_EMPTY_BOUNDARY 1
#define DMOLP_EXIT_NO_CHANGE 2
#define DMOLP_EXIT_CONVERGED 3

/* Labels at or above this value are invisible to the LP decision
 * (gpu_lp_boundary.cu:360). */
#define DMOLP_LP_LABEL_CAP 32
/* Largest k the library accepts. */
#define DMOLP_MAX_K 1024

typedef struct dmolp_ctx dmolp_ctx;
typedef struct dmolp_graph dmolp_graph;

/* ---------------- context ---------------- */

int dmolp_abi_version(void);
const char *dmolp_status_string(int status);
/* Message of the last failure on ctx (never NULL). */
const char *dmolp_last_error(const dmolp_ctx *ctx);

/* Owns the device binding, one HIP stream, kernel scratch and the
 * autotuned launch parameters.  Replaces allocateGPU (main.cpp:25-121) for
 * one GPU and the per-call stream of gpu_lp_boundary.cu:448-449. */
int dmolp_ctx_create(int device, dmolp_ctx **out);
int dmolp_ctx_destroy(dmolp_ctx *ctx);
/* Stream the context launches on (a hipStream_t), for callers that time or
 * order work around it. */
void *dmolp_ctx_stream(dmolp_ctx *ctx);

/* Per-round device times of kernel groups, recorded with HIP events on the
 * context stream during dmolp_phase2 while profiling is on (no counterpart
 * in the reference, which times whole GPU calls with std::chrono,
 * phase2.cpp:324-339).  group: 0 light-row LP decision, 1 heavy-row LP
 * decision, 2 round start + candidates, 3 selection, 4 delta statistics +
 * apply, 5 whole round, 6 heavy-row histogram update, 7 light-row delta
 * kernel, 8 initial light-row sweep (once per run). */
#define DMOLP_PROF_LP_LIGHT 0
#define DMOLP_PROF_LP_HEAVY 1
#define DMOLP_PROF_CANDIDATES 2
#define DMOLP_PROF_SELECT 3
#define DMOLP_PROF_DELTA 4
#define DMOLP_PROF_ROUND 5
#define DMOLP_PROF_HIST 6   /* heavy-row histogram update after the apply */
#define DMOLP_PROF_DELTA_LIGHT 7  /* the light-row Delta kernel alone */
#define DMOLP_PROF_INIT_LIGHT 8   /* the initial light-row sweep (one entry per run) */
int dmolp_ctx_set_profiling(dmolp_ctx *ctx, int on);
int dmolp_ctx_round_times(dmolp_ctx *ctx, int group, double *ms_out, int cap, int *n_out);

/* Measurement hook, no counterpart in the reference: the device's rate of
 * scattered 32-bit atomic adds, the ceiling of the Delta pass (which moves
 * counts in the 64-byte kept rows of random neighbours).  Groups of `share`
 * lanes (a divisor of 64) pick a random row of `rows` (1 to 2^32) 64-byte rows and each
 * lane adds 1 to a word of it, `reps` times per lane (reps <= 4096);
 * ms_out: the fastest of two timed launches; adds_out (nullable): adds per
 * launch.  Allocates and frees rows * 64 bytes. */
int dmolp_probe_atomics(dmolp_ctx *ctx, int64_t rows, int32_t share, int32_t reps, double *ms_out,
                        int64_t *adds_out);

/* Launch parameters the autotuner chooses; 0 means "keep". */
typedef struct {
    int32_t heavy_degree;  /* rows longer than this are split into chunks */
    int32_t chunk_edges;   /* edges per chunk of a heavy row (64 .. 2^21) */
    int32_t block_threads; /* threads per workgroup for the row kernels */
    int32_t grid_blocks;   /* workgroups of the persistent row kernels */
    int32_t pull_per_mille; /* find candidates from the boundary side when B_{t-1}
                               covers at least this many per mille of the CSR entries */
    int32_t batch_work_per_block; /* batched launches: (vertices + entries/4) per workgroup */
    int32_t hist_recount_per_mille; /* heavy-row histograms: recount after a round whose
                                       changed vertices hold at least this share of entries */
    int32_t hist_push_per_mille;    /* ... push the changes when they hold at most this
                                       share (in 1/1000); between the two, scan a bitmap */
    int32_t light_hist;  /* light rows' label counts: 1 kept in HBM across rounds and
                            moved by the changed vertices, 2 gathered afresh each round */
} dmolp_tuning;
int dmolp_ctx_get_tuning(const dmolp_ctx *ctx, dmolp_tuning *out);
int dmolp_ctx_set_tuning(dmolp_ctx *ctx, const dmolp_tuning *t);

/* ---------------- graphs ---------------- */

/* Generator description (no counterpart in the reference, whose inputs are
 * WebGraph text files; post-processing matches tool/convert.sh:76-79,
 * 106-116: symmetrise, drop self-loops, dedupe, sort rows). */
#define DMOLP_GEN_RMAT 1     /* R-MAT, no noise: scale, edge_factor, a, b, c */
#define DMOLP_GEN_CHUNG_LU 2 /* n, m directed samples, gamma, dmax */
#define DMOLP_GEN_GRID 3     /* 2-D 5-point grid: rows x cols */
typedef struct {
    int32_t kind;
    int32_t scale;
    int32_t edge_factor;
    int32_t rows;
    int32_t cols;
    int32_t reserved;
    double a, b, c;
    double gamma;
    double dmax;
    int64_t n;
    int64_t m;
    uint64_t seed;
} dmolp_gen_spec;

/* Host CSR owned by the library (generator and file-reader output). */
typedef struct {
    int64_t n;
    int64_t nnz;
    int64_t *off; /* n + 1 */
    int32_t *col; /* nnz */
} dmolp_csr;

/* Generate on the host (threads); free with dmolp_csr_free. */
int dmolp_gen_host(const dmolp_gen_spec *spec, dmolp_csr *out);
void dmolp_csr_free(dmolp_csr *csr);

/* Reference text adjacency format (init.cpp:25-100; tool/convert.sh:132-176):
 * first line V, then one line per vertex listing its neighbours.  Rows are
 * sorted and deduplicated on load, self-loops dropped, as
 * build_local_graph_and_ghosts does (partition.cpp:231-232). */
int dmolp_read_adjacency_text(const char *path, dmolp_csr *out);
int dmolp_write_adjacency_text(const char *path, int64_t n, const int64_t *off, const int32_t *col);

/* Binary CSR cache (no counterpart in the reference, whose loader parses the
 * text file on every run, init.cpp:25-100).  Layout, native byte order: the
 * 8 bytes "DMOLPCSR", uint32 version (1), uint32 offset width (8), int64 n,
 * int64 nnz, then off[n + 1] (int64) and col[nnz] (int32).  The reader
 * checks the header, the sizes against the file length and the offsets'
 * bounds; it returns DMOLP_E_GRAPH for a file that fails them. */
int dmolp_write_csr_binary(const char *path, int64_t n, const int64_t *off, const int32_t *col);
int dmolp_read_csr_binary(const char *path, dmolp_csr *out);
/* The text file through its cache: reads <path>.csr when that file exists,
 * is at least as new as the text file and passes the reader's checks;
 * otherwise parses the text and writes <path>.csr for the next run (a cache
 * that cannot be written is skipped).  *from_cache (nullable) tells which. */
int dmolp_load_adjacency_cached(const char *path, dmolp_csr *out, int32_t *from_cache);

/* Make a CSR device-resident (replaces the per-round H2D copies of
 * gpu_lp_boundary.cu:456-471).  Validates offsets, sorted rows and the
 * absence of self-loops on the device. */
int dmolp_graph_upload(dmolp_ctx *ctx, int64_t n, int64_t nnz, const void *offsets,
                       int offset_bytes, const int32_t *cols, dmolp_graph **out);
/* Generate directly on the device; same graph as dmolp_gen_host. */
int dmolp_graph_generate(dmolp_ctx *ctx, const dmolp_gen_spec *spec, dmolp_graph **out);
int dmolp_graph_info(const dmolp_graph *g, int64_t *n, int64_t *nnz);
/* Degree binning of g under ctx's tuning (no counterpart in the reference):
 * rows longer than heavy_degree, and the CSR entries they hold. */
int dmolp_graph_bins(dmolp_ctx *ctx, dmolp_graph *g, int64_t *heavy_rows, int64_t *heavy_entries);
/* Copy the CSR back to the host (off: n + 1 int64, col: nnz). */
int dmolp_graph_download(dmolp_ctx *ctx, const dmolp_graph *g, int64_t *off, int32_t *col);
int dmolp_graph_free(dmolp_ctx *ctx, dmolp_graph *g);

/* Resident labels (n int32) of a graph: phase 1 writes them, phase 2 reads
 * and updates them in place on the device. */
int dmolp_graph_set_labels(dmolp_ctx *ctx, dmolp_graph *g, const int32_t *labels);
int dmolp_graph_get_labels(dmolp_ctx *ctx, const dmolp_graph *g, int32_t *labels);
/* Device-side checkpoint of the resident labels: op 0 saves, op 1 restores.
 * Lets a caller rerun phase 2 from the same start without host copies. */
int dmolp_graph_labels_checkpoint(dmolp_ctx *ctx, dmolp_graph *g, int op);

/* ---------------- phase 1 ---------------- */

/* find_hub_nodes (init.cpp:139-177) as the library runs it on the host
 * for dmolp_phase1: the hubs of a degree sequence deg[0, n) in the
 * reference's order (its std::sort's order of equal degrees included).
 * hubs_out has capacity n (nullable); no device work. */
int dmolp_hubs_host(const int32_t *deg, int64_t n, int32_t *hubs_out, int64_t *n_hubs);
/* find_first_seed (init.cpp:179-187) as the library runs it on the host for
 * dmolp_phase1: the first seed of a degree sequence deg[0, n) and its
 * degree (-1, -1 when n is 0); no device work. */
int dmolp_first_seed_host(const int32_t *deg, int64_t n, int32_t *seed, int32_t *degree);
/* How many hubs dmolp_phase1's host search orders before the seed search
 * (default 65536; 0: all of them).  A seed search that runs out of the
 * ordered hubs is repeated with all of them, so the result does not depend
 * on this; it only bounds the host sort (on near-regular graphs every
 * vertex is a hub).  Test hook; no counterpart in the reference. */
int dmolp_ctx_set_hub_prefix(dmolp_ctx *ctx, int64_t prefix);
/* find_hub_nodes and find_first_seed (init.cpp:139-187) as dmolp_phase1
 * runs them for graphs of at least dev_min vertices: the introsort's
 * partition steps on the device for ranges of at least host_max entries,
 * shorter ranges finished on the host.  The same results as
 * dmolp_hubs_host and dmolp_first_seed_host; limit >= 0 lists only the
 * first min(limit, n_hubs) hubs (hubs_out capacity n, nullable). */
int dmolp_hubs_device(dmolp_ctx *ctx, const int32_t *deg, int64_t n, int64_t limit, int64_t host_max,
                      int32_t *hubs_out, int64_t *n_hubs, int32_t *seed, int32_t *degree);
/* Graphs of at least dev_min vertices get dmolp_phase1's hub search on the
 * device, finishing ranges shorter than host_max on the host (defaults
 * 2^18 and 2^15).  The results do not depend on either.  Test hook; no
 * counterpart in the reference. */
int dmolp_ctx_set_hub_device(dmolp_ctx *ctx, int64_t dev_min, int64_t host_max);
/* The seed fallback's candidates as dmolp_phase1 selects them on the device
 * for graphs it searches there: the first min(m, n) vertices of deg[0, n)
 * by (degree descending, id ascending), in that order, into out (capacity
 * min(m, n)).  Ranges shorter than host_max are sorted on the host.  The
 * fallback extends seed.cpp:214-217 (fewer than k seeds); test hook. */
int dmolp_top_degree_device(dmolp_ctx *ctx, const int32_t *deg, int64_t n, int64_t m, int64_t host_max,
                            int32_t *out);

/* Phase1Metrics (graph_types.h:276-309) with int64 counters. */
typedef struct {
    int64_t initial_edge_cut;
    double initial_vertex_balance;
    double initial_edge_balance;
    int64_t total_vertices;
    int64_t total_edges;
    int32_t first_seed;
    int32_t first_seed_degree;
    int64_t n_hubs;
    int32_t n_seeds_found;    /* seeds from the hub search */
    int32_t n_seeds_fallback; /* seeds added by the documented fallback */
    int64_t rounds;           /* expansion rounds */
    int32_t fill_round;       /* round of the fill step, -1 if none */
    int32_t reserved;
    double device_ms;         /* device time of phase 1 (hipEvents) */
    double host_ms;           /* host time of the hub sort */
} dmolp_phase1_metrics;

/* Replaces run_phase1's hub-seed path (phase1.cpp:18-161 with mode=bfs;
 * find_hub_nodes init.cpp:139-177, find_first_seed init.cpp:179-187,
 * find_seeds seed.cpp:192-242, partition_expansion partition.cpp:323-621).
 * Labels become resident in g; labels_out (n, nullable) receives a copy;
 * seeds_out (k, nullable); vc, ec (k, nullable). */
int dmolp_phase1(dmolp_ctx *ctx, dmolp_graph *g, int32_t k, int32_t theta, int32_t *labels_out,
                 int32_t *seeds_out, int64_t *vc, int64_t *ec, dmolp_phase1_metrics *m);

/* Replaces random_partition (random.cpp:24-127, mode = random): label =
 * vertex id mod k.  Metrics as dmolp_phase1 (seed fields zero). */
int dmolp_phase1_random(dmolp_ctx *ctx, dmolp_graph *g, int32_t k, int32_t *labels_out,
                        int64_t *vc, int64_t *ec, dmolp_phase1_metrics *m);

/* ---------------- phase 2 ---------------- */

/* PartitioningMetrics (graph_types.h:312-346) with int64 counters. */
typedef struct {
    int64_t edge_cut;
    double vertex_balance;
    double edge_balance;
    int64_t total_vertices;
    int64_t total_edges;
    int32_t rounds;      /* 1-based index of the last round started */
    int32_t exit_reason; /* DMOLP_EXIT_* */
    int64_t lp_edges;    /* sum over rounds of sum_{b in B_t} deg b */
    int64_t initial_cut;
    double device_ms;    /* device time from phase-2 start to final labels */
} dmolp_phase2_metrics;

/* Per-round counters, arrays of max_rounds entries (any pointer may be
 * NULL).  Same meaning as the oracle's trace. */
typedef struct {
    int64_t *prev_count; /* |B_{t-1}| */
    int64_t *prev_deg;   /* sum deg over B_{t-1} */
    int64_t *cand_count; /* |B_{t-1} u N(B_{t-1})| */
    int64_t *bd_count;   /* boundary vertices before selection */
    int64_t *b_count;    /* |B_t| */
    int64_t *b_deg;      /* sum deg over B_t */
    int64_t *chg_count;  /* |Delta_t| */
    int64_t *chg_deg;    /* sum deg over Delta_t */
    int64_t *cut;        /* cut after round t, -1 if the round stopped before stats */
    int64_t *b_light_count; /* entries of B_t processed by the light-row kernel */
    int64_t *b_light_deg;   /* their degree sum */
    int64_t *chg_light_count; /* entries of Delta_t processed by the light-row kernels */
    int64_t *chg_light_deg;   /* their degree sum */
} dmolp_trace;

/* Replaces run_phase2 (phase2.cpp:256-449) with max_iter and epsilon as
 * parameters (phase2.cpp:263-264).  Runs on g's resident labels; the whole
 * round loop is enqueued without host synchronisation.  vc, ic (k,
 * nullable) receive the final counts. */
int dmolp_phase2(dmolp_ctx *ctx, dmolp_graph *g, int32_t k, int32_t max_rounds, double epsilon,
                 int64_t *vc, int64_t *ic, dmolp_phase2_metrics *m, dmolp_trace *trace);

/* ---------------- pieces (test hooks and the reference's GPU seam) ------- */

/* computePartitionStats + computeEdgeCut (phase2.cpp:55-117, 214-253) on
 * host labels. */
int dmolp_stats(dmolp_ctx *ctx, dmolp_graph *g, int32_t k, const int32_t *labels, int64_t *vc,
                int64_t *ic, int64_t *cut);
/* No counterpart in the reference: dmolp_stats for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and
 * statuses as dmolp_stats; its device memory beyond g is one static array
 * per GPU (the labels and the counters), reused by every call, so calls
 * that overlap on one GPU race.  n above DMOLP_NOALLOC_MAX_N fails with
 * DMOLP_E_UNSUPPORTED. */
#define DMOLP_NOALLOC_MAX_N (1 << 24)
int dmolp_stats_noalloc(dmolp_ctx *ctx, const dmolp_graph *g, int32_t k, const int32_t *labels,
                        int64_t *vc, int64_t *ic, int64_t *cut);
/* calculatePenalties (phase2.cpp:158-209) evaluated on the device. */
int dmolp_penalties(dmolp_ctx *ctx, int32_t k, const int64_t *vc, const int64_t *ic, int32_t t,
                    double *pen, double *re);
/* expandBoundaryNodes (gpu_lp_boundary.cu:185-305): B_t from B_{t-1}.
 * out (capacity n) is written in ascending id order. */
int dmolp_select_boundary(dmolp_ctx *ctx, dmolp_graph *g, const int32_t *labels,
                          const int32_t *prev, int64_t n_prev, const double *re, int32_t k,
                          int32_t t, int32_t *out, int64_t *n_out);
/* runBoundaryLPOnGPU_Streaming's decision (gpu_lp_boundary.cu:525-565 with
 * boundaryLPKernel_unified :327-427): labels_out = labels with every b in B
 * replaced by its decision (Jacobi). */
int dmolp_lp_round(dmolp_ctx *ctx, dmolp_graph *g, const int32_t *labels, const int32_t *b,
                   int64_t nb, const double *pen, int32_t k, int32_t *labels_out);

/* ---------------- rank shares with an on-device halo ---------------- */

/* P ranks' shares of one graph on one GPU (P in [1, 64]).  Rank r owns
 * the vertices v with v mod P == r (partition.cpp:55-57); its share holds
 * their rows with neighbours as local ids (v / P) or ghost ids nl + slot,
 * slots in order of first appearance over the rows (partition.cpp:182-262),
 * and its ghosts' labels fetched from their owners (partition.cpp:264-318),
 * all in device memory.  Built from g's resident labels. */
typedef struct dmolp_shares dmolp_shares;
int dmolp_shares_create(dmolp_ctx *ctx, dmolp_graph *g, int32_t P, dmolp_shares **out);
int dmolp_shares_destroy(dmolp_ctx *ctx, dmolp_shares *s);
/* Owned rows, ghosts and entries of share r. */
int dmolp_shares_info(const dmolp_shares *s, int32_t r, int64_t *n_local, int64_t *n_ghost,
                      int64_t *nnz);
/* Share r's layout and halo (nullable outputs): off (n_local + 1), col
 * (nnz, unified ids), ghost_gid (n_ghost, slot order), labels (n_local +
 * n_ghost: owned, then ghosts). */
int dmolp_shares_download(dmolp_ctx *ctx, const dmolp_shares *s, int32_t r, int64_t *off,
                          int32_t *col, int32_t *ghost_gid, int32_t *labels);
/* The halo refresh of share r (phase2.cpp:368-376): every Delta{gid, label}
 * (host arrays of count entries) whose vertex is one of r's ghosts rewrites
 * that ghost's label on the device; *updated (nullable) receives how many. */
int dmolp_halo_refresh(dmolp_ctx *ctx, dmolp_shares *s, int32_t r, const int32_t *gid,
                       const int32_t *label, int64_t count, int64_t *updated);
/* Phase 2 (phase2.cpp:256-449) run rank by rank on the shares: partial
 * counts summed (the Allreduce of phase2.cpp:100-117, 229-236), each
 * share's selection and decisions on its owned rows over its unified
 * labels, and every round the shares' deltas concatenated in rank order
 * (the Allgatherv of phase2.cpp:122-154) and applied to every share's
 * ghost labels.  scope 0: each rank selects over its own boundary vertices
 * (gpu_lp_boundary.cu:274-302 on the rank's vectors); scope 1: selection
 * over all ranks' boundary vertices (the run then equals dmolp_phase2).
 * labels_out (n, nullable); the final labels also become g's resident
 * labels.  ghost_updates (max_rounds, nullable): ghost labels rewritten by
 * each round's refresh, summed over the shares.  The trace's candidate and
 * light-part columns are -1. */
int dmolp_phase2_shares(dmolp_ctx *ctx, dmolp_shares *s, int32_t k, int32_t max_rounds,
                        double epsilon, int32_t scope, int32_t *labels_out, int64_t *vc,
                        int64_t *ic, dmolp_phase2_metrics *m, dmolp_trace *trace,
                        int64_t *ghost_updates);

/* ---------------- batched and served modes ---------------- */

/* One job of a batch: a graph, its parameters, and where results go. */
typedef struct {
    dmolp_graph *graph;
    int32_t k;
    int32_t theta;
    int32_t max_rounds;
    int32_t reserved;
    double epsilon;
    int32_t *labels_out;  /* n, nullable */
} dmolp_job;

typedef struct {
    int32_t status;
    int32_t rounds;
    int32_t exit_reason;
    int32_t n_seeds_fallback;
    int64_t phase1_cut;
    int64_t edge_cut;
    int64_t lp_edges;
    double vertex_balance;
    double edge_balance;
} dmolp_result;

/* Partition many graphs: every job's phase 1 + phase 2 rounds are advanced
 * together, one multi-graph launch per kernel step, captured once as a HIP
 * graph per step shape and replayed.  Each result equals the job run alone. */
int dmolp_batch_run(dmolp_ctx *ctx, int32_t n_jobs, const dmolp_job *jobs, dmolp_result *results,
                    double *device_ms);

/* Autotuner (no counterpart in the reference): runs the batch under
 * candidate launch parameters (degree-bin threshold, heavy-row chunk
 * length, blocks per unit of work; at most max_trials, 0 = all), checks that
 * every candidate gives the same results, and keeps the fastest in ctx.
 * best (nullable) receives it. */
int dmolp_batch_autotune(dmolp_ctx *ctx, int32_t n_jobs, const dmolp_job *jobs, int32_t max_trials,
                         dmolp_tuning *best);

/* Served mode (no counterpart in the reference, whose driver partitions one
 * file per process, main.cpp:123-193): a request loop with continuous
 * batching.  Requests wait until a slot and their device-memory footprint
 * are free; every dmolp_server_step advances all admitted jobs together,
 * whatever stage each is in, and finished jobs free their slot and memory at
 * once.  A job's result is the same as its plain run, alone or in any mix. */
typedef struct dmolp_server dmolp_server;

typedef struct {
    dmolp_graph *graph;    /* resident graph (kept by the caller until the result is
                              taken), or NULL: the graph is generated from spec on
                              admission and freed when the job finishes */
    dmolp_gen_spec spec;
    int32_t k;
    int32_t theta;
    int32_t max_rounds;
    int32_t want_labels;   /* keep the final labels for dmolp_server_take */
    double epsilon;
    double arrival_s;      /* arrival time on the server clock; < 0: now */
} dmolp_request;

typedef struct {
    dmolp_result result;
    int64_t n, nnz;
    double arrival_s, admit_s, done_s; /* server clock, seconds */
} dmolp_served;

/* budget: device bytes the admitted jobs may hold together; max_active:
 * jobs admitted at once.  The server drives ctx's stream: call
 * dmolp_server_step from ctx's thread; submit and take may come from any
 * thread. */
int dmolp_server_create(dmolp_ctx *ctx, int64_t budget, int32_t max_active, dmolp_server **out);
int dmolp_server_destroy(dmolp_server *srv);
/* Seconds since the server was created. */
double dmolp_server_clock(const dmolp_server *srv);
int dmolp_server_submit(dmolp_server *srv, const dmolp_request *req, int64_t *ticket);
/* One scheduling step; active/queued (nullable) receive the job counts after it. */
int dmolp_server_step(dmolp_server *srv, int32_t *active, int32_t *queued);
/* 1 when the job has finished (result copied out and released), 0 while it
 * is pending; labels_out (n, nullable) receives the labels if requested. */
int dmolp_server_take(dmolp_server *srv, int64_t ticket, dmolp_served *out, int32_t *labels_out);

#ifdef __cplusplus
}
#endif
#endif
