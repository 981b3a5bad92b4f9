// select.hpp -- per-group deterministic top-m selection by (degree, id).
//
// Used for the boundary selection of gpu_lp_boundary.cu:277-302 (per label,
// the first ceil(ratio*|group|) by degree descending when RE > 1.0 else
// ascending, ties by id) and for the frontier throttle of
// partition.cpp:400-431 (per partition, the theta smallest (degree, id)).
//
// No sort: a level-1 histogram over min(degree, B1-1) finds the degree bin
// where the cut falls; if that bin must be split, 12-bit MSD digit passes
// over the key K = (dk << IB) | id (dk = degree, or maxdeg - degree for the
// descending order) find the exact threshold key.  Every pass is a
// histogram, so the selected set does not depend on record order.
#pragma once
#include <array>
#include <vector>

#include "internal.hpp"
#include "rows.hpp"

namespace dmolp {

struct RecDev {
    int32_t *x;  // vertex id
    int32_t *g;  // group (label / partition)
    int32_t *d;  // degree
    unsigned long long *cnt;
};

__device__ __forceinline__ RecDev gload(const RecDev &v) {
    RecDev r = v;
    r.x = gload(v.x);
    r.g = gload(v.g);
    r.d = gload(v.d);
    r.cnt = gload(v.cnt);
    return r;
}

struct Recs {
    DevMem x, g, d, cnt;
    RecDev dev{};
    void alloc(int64_t cap) {
        x.alloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
        g.alloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
        d.alloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
        cnt.alloc(sizeof(unsigned long long) * 2);
        dev = RecDev{x.as<int32_t>(), g.as<int32_t>(), d.as<int32_t>(), cnt.as<unsigned long long>()};
    }
};

enum SelMode : int32_t {
    SEL_NONE = 0,     // take nothing
    SEL_ALL = 1,      // take the whole group
    SEL_THROUGH = 2,  // take every record up to and including level-1 bin b1
    SEL_REFINE = 3,   // threshold inside bin b1 still being refined
    SEL_KEY = 4       // take records before bin b1, and in b1 those with K <= prefix|(2^r-1)
};

struct SelState {
    int32_t mode;
    int32_t b1;
    int32_t r;
    int32_t desc;
    int64_t need;
    uint64_t prefix;
    int64_t take;
    int64_t count;
};

// How the per-group take count is derived from the group size c.
struct TakeRule {
    int32_t kind;          // 0: ceil(ratio * c) capped at c, desc = re[g] > 1.0; 1: min(theta, c), asc
    int32_t theta;
    const double *ratio_tab;  // device: ratio of round t at ratio_tab[*t_ptr] (kind 0)
    const int32_t *t_ptr;
    const double *re;      // device, groups entries (kind 0)
};

struct SelectWork {
    int groups = 0;
    int b1_bins = 0;
    int id_bits = 0;
    int deg_bits = 0;
    int passes = 0;
    DevMem h1, h2, st, refining;
    Recs ref;  // records still being refined after the level-1 pick
    int64_t ref_cap = 0;
    void prepare(dmolp_ctx *ctx, int groups, int64_t n, int32_t max_degree);
};

// Everything one job's selection reads and writes (device pointers).
struct SelArgs {
    RecDev in;         // records to select from
    ListDev sel;       // selected vertices (reset before writing)
    RecDev carry;      // unselected records are appended here when has_carry
    RecDev ref;
    int32_t *h1, *h2;
    SelState *st;
    int32_t *refining;
    const int32_t *done;      // the job's stop flag (nullable)
    int32_t *mark;            // selected vertices get mark[x] = *mark_val (nullable)
    const int32_t *mark_val;
    TakeRule rule;
    int32_t groups, b1_bins, id_bits, deg_bits;
    int32_t has_carry, heavy_degree, chunk_edges, dense;
    // Dense records (dense != 0): record v exists for 0 <= v < rn iff
    // rdeg[v] >= 0, with id v, group rlab[v] and degree rdeg[v]; `in` is
    // unused.  The selected vertices then go to half (*rule.t_ptr & 1) of
    // the double bitmap `bits` (bwords words per half), every word of it
    // written; sel gets the counts only (LC_LIGHT, LC_DEG, LC_SLOTS as the
    // heavy member count, LC_HDEG), no lists.
    const int32_t *rdeg;
    const int32_t *rlab;
    int64_t rn;
    uint32_t *bits;
    int64_t bwords;
};

SelArgs select_args(dmolp_ctx *ctx, SelectWork &w, const RecDev &in, const TakeRule &rule,
                    const int32_t *done, const ListDev &sel, const RecDev *carry,
                    int32_t *mark = nullptr, const int32_t *mark_val = nullptr);
// Switch a job's selection to dense records (see SelArgs::dense).
void select_args_dense(SelArgs &a, const int32_t *rdeg, const int32_t *rlab, int64_t rn,
                       uint32_t *bits, int64_t bwords);

// Enqueue the whole selection for a set of jobs (args: device array, one
// SelArgs per job): level-1 histogram, pick, digit passes, compaction.
// `passes` is the largest digit-pass count over the jobs.
void select_enqueue(dmolp_ctx *ctx, const SelArgs *args, const LaunchSet &ls, int passes);

// Block counts of one job for the launch policies (jobs with n vertices
// and nnz entries, `groups` selection groups, `heavy` heavy rows).
void job_blocks(const dmolp_ctx *ctx, int64_t n, int64_t nnz, int groups, int64_t heavy,
                bool single, int32_t out[P_NPOL]);
// Upload the block-start arrays of a job set.
void launch_set_build(dmolp_ctx *ctx, LaunchSet &ls, const std::vector<std::array<int32_t, P_NPOL>> &blocks);

}  // namespace dmolp
