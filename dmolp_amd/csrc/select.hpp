// select.hpp -- per-group deterministic top-m selection by (degree, id).
//
// Used for the boundary selection of gpu_lp_boundary.cu:277-302 (per label,
// the first ceil(ratio*|group|) by degree descending when RE > 1.0 else
// ascending, ties by id) and for the frontier throttle of
// partition.cpp:400-431 (per partition, the theta smallest (degree, id)).
//
// No sort: a level-1 histogram over min(degree, B1-1) finds the degree bin
// where the cut falls; if that bin must be split, MSD digit passes (8 bits
// wide, narrower when there are more than 64 groups)
// over the key K = (dk << IB) | id (dk = degree, or maxdeg - degree for the
// descending order) find the exact threshold key.  Every pass is a
// histogram, so the selected set does not depend on record order.
#pragma once
#include <array>
#include <type_traits>
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
    SEL_KEY = 4,      // take records before bin b1, and in b1 those with K <= prefix|(2^r-1)
    SEL_RANK = 5      // dense, b1 an exact degree: records before bin b1, and of b1 the
                      // first `need` by id (id chunks before chunk `prefix`, and in that
                      // chunk those ranked below its remaining `need`)
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
    int digit_bits = 0;
    int passes = 0;
    DevMem h1, h2, st, refining;
    Recs ref;  // records still being refined after the level-1 pick
    int64_t ref_cap = 0;
    // Dense rank cuts: per group (the first kSelLdsGroups) and 1024-id chunk,
    // the count of records in the group's cut bin; per chunk, the groups
    // whose cut falls in it.
    DevMem rank_cnt, rank_mark;
    int64_t rank_chunks = 0;
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
    unsigned long long *fault;  // counted when a pass finds an impossible state (nullable)
    int32_t *mark;            // selected vertices get mark[x] = *mark_val (nullable)
    const int32_t *mark_val;
    TakeRule rule;
    int32_t groups, b1_bins, id_bits, deg_bits;
    int32_t digit_bits;  // width of a digit pass: groups << digit_bits bins fit the LDS histogram
    int32_t has_carry, heavy_degree, chunk_edges, dense;
    // Dense records (dense != 0): record v exists for 0 <= v < rn iff
    // rdeg[v] >= 0, with id v, group rlab[v] (rlab8[v] when rlab8 is set:
    // every group id below 255) and degree rdeg[v]; `in` is
    // unused.  The selected vertices then go to half (*rule.t_ptr & 1) of
    // the double bitmap `bits` (bwords words per half), every word of it
    // written; sel gets the counts only (LC_LIGHT, LC_DEG, LC_SLOTS as the
    // heavy member count, LC_HDEG), no lists.
    const int32_t *rdeg;
    const int32_t *rlab;
    const uint8_t *rlab8;
    int64_t rn;
    uint32_t *bits;
    int64_t bwords;
    int32_t *rank_cnt;             // [rank groups][rank_stride(rank_chunks)]
    unsigned long long *rank_mark; // [rank_chunks]: bit g when group g's cut chunk is this one
    int64_t rank_chunks;
};

// Device helpers shared by the selection kernels and phase 2's fused
// take-and-decide pass.
__device__ __forceinline__ uint64_t sel_key(int32_t x, int32_t d, int desc, int id_bits,
                                            int deg_bits) {
    const uint64_t dmax = (deg_bits >= 63) ? ~0ull : ((1ull << deg_bits) - 1);
    const uint64_t dk = desc ? dmax - (uint64_t)d : (uint64_t)d;
    return (dk << id_bits) | (uint64_t)(uint32_t)x;
}

__device__ __forceinline__ int bin1_of(int32_t d, int b1_bins) {
    return d < b1_bins - 1 ? d : b1_bins - 1;
}

// What the per-record passes need of a group's state, cached in LDS for the
// first kSelLdsGroups groups (a global gather per record otherwise).
constexpr int kSelLdsGroups = 64;
struct SelLite {
    int32_t mode, b1, desc, pad;
    uint64_t kmax;  // SEL_KEY: largest key taken inside bin b1
};
__device__ __forceinline__ SelLite sel_lite(const SelState &s) {
    SelLite l;
    l.mode = s.mode;
    l.b1 = s.b1;
    l.desc = s.desc;
    l.pad = 0;
    // SEL_RANK: the cut chunk in the high word, the need inside it below.
    l.kmax = s.mode == SEL_RANK ? ((s.prefix << 32) | (uint64_t)(uint32_t)s.need)
                                : s.prefix | ((s.r >= 64) ? ~0ull : ((1ull << s.r) - 1));
    return l;
}
__device__ __forceinline__ void sel_lite_load(SelLite *lds, const SelState *st, int groups) {
    const int ng = groups < kSelLdsGroups ? groups : kSelLdsGroups;
    for (int i = threadIdx.x; i < ng; i += blockDim.x) lds[i] = sel_lite(st[i]);
    __syncthreads();
}
__device__ __forceinline__ SelLite sel_get(const SelLite *lds, const SelState *st, int g) {
    return g < kSelLdsGroups ? lds[g] : sel_lite(st[g]);
}

__device__ __forceinline__ SelArgs gload(const SelArgs &v) {
    SelArgs r = v;
    r.in = gload(v.in);
    r.sel = gload(v.sel);
    r.carry = gload(v.carry);
    r.ref = gload(v.ref);
    r.h1 = gload(v.h1);
    r.h2 = gload(v.h2);
    r.st = gload(v.st);
    r.refining = gload(v.refining);
    r.done = gload(v.done);
    r.mark = gload(v.mark);
    r.mark_val = gload(v.mark_val);
    r.rule.ratio_tab = gload(v.rule.ratio_tab);
    r.rule.t_ptr = gload(v.rule.t_ptr);
    r.rule.re = gload(v.rule.re);
    r.rdeg = gload(v.rdeg);
    r.rlab = gload(v.rlab);
    r.rlab8 = gload(v.rlab8);
    r.bits = gload(v.bits);
    r.rank_cnt = gload(v.rank_cnt);
    r.rank_mark = gload(v.rank_mark);
    return r;
}

// Group of dense record x.
__device__ __forceinline__ int32_t dense_group(const SelArgs &a, int64_t x) {
    return a.rlab8 ? (int32_t)a.rlab8[x] : a.rlab[x];
}

constexpr int kDenseItems = 4;  // dense passes: 64-id groups per wave per iteration
// Dense passes over ids run block-uniform iterations of 1024 ids (4 waves of
// kDenseItems 64-id groups): the rank chunks of SEL_RANK.
constexpr int kRankChunkLog = 10;
// Row stride (in counts) of a group's chunk counts: a multiple of 4, so rows
// start 16-byte aligned.
__host__ __device__ __forceinline__ int64_t rank_stride(int64_t chunks) { return (chunks + 3) & ~3ll; }

// Degree and group of dense records i0 + j * stride (j < N), d = -1 past rn
// or for a non-record.  The indices are clamped into [0, rn), so all 2N
// loads are issued together and waited for once (a load under a lane
// condition becomes a branch and a wait of its own).  The group of a
// non-record is read but meaningless.  Byte: groups from rlab8.
template <bool Byte, int N>
__device__ __forceinline__ void dense_load(const SelArgs &a, int64_t i0, int64_t stride, int32_t (&d)[N],
                                           int32_t (&g)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int64_t i = i0 + (int64_t)j * stride;
        const int64_t ic = i < a.rn ? i : a.rn - 1;
        const int32_t dv = a.rdeg[ic];
        g[j] = Byte ? (int32_t)a.rlab8[ic] : a.rlab[ic];
        d[j] = i < a.rn ? dv : -1;
    }
}

// Whether dense record v (degree d, its group's state s) is taken.
__device__ __forceinline__ bool dense_take(const SelLite &s, int32_t v, int32_t d, int b1_bins,
                                           int id_bits, int deg_bits) {
    const int b = bin1_of(d, b1_bins);
    const bool earlier = s.desc ? b > s.b1 : b < s.b1;
    switch (s.mode) {
        case SEL_ALL: return true;
        case SEL_THROUGH: return earlier || b == s.b1;
        case SEL_KEY: return earlier || (b == s.b1 && sel_key(v, d, s.desc, id_bits, deg_bits) <= s.kmax);
        // In the cut chunk itself the records are ranked by rank_cut.
        case SEL_RANK: return earlier || (b == s.b1 && (uint32_t)(v >> kRankChunkLog) < (uint32_t)(s.kmax >> 32));
        default: return false;
    }
}

// Take decisions of one block iteration over the 1024 ids of chunk c
// (thread layout of the dense passes: wave w, item j, lane: id
// (c << 10) + (w * kDenseItems + j) * 64 + lane) for the groups whose rank
// cut falls in this chunk (bits of `mark`): a record of such a group in its
// cut bin is taken when fewer than the chunk's need of the group's cut-bin
// records precede it in id order.  Block-uniform call (4 waves).
__device__ __forceinline__ void rank_cut(const SelLite *s_st, unsigned long long mark,
                                         const int32_t (&d)[kDenseItems], const int32_t (&gg)[kDenseItems],
                                         bool (&take)[kDenseItems], int32_t *s_wc) {
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint64_t lt = (1ull << lane) - 1ull;
    while (mark) {
        const int g = __ffsll((long long)mark) - 1;
        mark &= mark - 1;
        const SelLite s = s_st[g];
        if (s.mode != SEL_RANK) continue;  // uniform: a mark left from another round
        const int32_t need = (int32_t)(uint32_t)s.kmax;
        uint64_t m[kDenseItems];
        int32_t tot = 0;
#pragma unroll
        for (int j = 0; j < kDenseItems; ++j) {
            m[j] = __ballot(d[j] >= 0 && gg[j] == g && d[j] == s.b1);
            tot += __popcll(m[j]);
        }
        if (lane == 0) s_wc[w] = tot;
        __syncthreads();
        int32_t run = 0;
        for (int x = 0; x < w; ++x) run += s_wc[x];
#pragma unroll
        for (int j = 0; j < kDenseItems; ++j) {
            if ((m[j] >> lane) & 1ull) take[j] = run + __popcll(m[j] & lt) < need;
            run += __popcll(m[j]);
        }
        __syncthreads();
    }
}

SelArgs select_args(dmolp_ctx *ctx, SelectWork &w, const RecDev &in, const TakeRule &rule,
                    const int32_t *done, const ListDev &sel, const RecDev *carry,
                    int32_t *mark = nullptr, const int32_t *mark_val = nullptr);
// Switch a job's selection to dense records (see SelArgs::dense).
void select_args_dense(SelArgs &a, const int32_t *rdeg, const int32_t *rlab, const uint8_t *rlab8,
                       int64_t rn, uint32_t *bits, int64_t bwords);

// Enqueue the whole selection for a set of jobs (args: device array, one
// SelArgs per job): level-1 histogram, pick, digit passes, compaction.
// `passes` is the largest digit-pass count over the jobs.
// `dense`: every job's records are dense (select_args_dense), so only the
// dense kernels are launched.  `compact` false leaves the dense jobs' final take pass (bitmap and
// counts) to the caller, which must then run dense_take and rank_cut over
// every record itself.
void select_enqueue(dmolp_ctx *ctx, const SelArgs *args, const LaunchSet &ls, int passes,
                    bool compact = true, bool dense = false);

// Block counts of one job for the launch policies (jobs with n vertices
// and nnz entries, `groups` selection groups, `heavy` heavy rows).
void job_blocks(const dmolp_ctx *ctx, int64_t n, int64_t nnz, int groups, int64_t heavy,
                bool single, int32_t out[P_NPOL]);
// Upload the block-start arrays of a job set.
void launch_set_build(dmolp_ctx *ctx, LaunchSet &ls, const std::vector<std::array<int32_t, P_NPOL>> &blocks);

}  // namespace dmolp
