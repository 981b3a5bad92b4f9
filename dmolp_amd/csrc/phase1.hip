// phase1.hip -- hub seeds and throttled multi-source expansion
// (phase1.cpp:18-161 with mode = bfs), device-resident after the hub sort.
//
//   hubs      find_hub_nodes (init.cpp:139-177) on the host: the tie order
//             of libstdc++ std::sort is part of the result, so the same call
//             runs on the same (id, degree) sequence.
//   seeds     find_seeds (seed.cpp:192-242): depth-2 ball bitmaps of the
//             chosen seeds (atomicOr), next seed = unused, uncovered hub of
//             largest degree, first in hub order (one 64-bit atomicMax).
//   expand    partition_expansion (partition.cpp:323-604): per round the
//             theta smallest (degree, id) of each frontier are expanded; an
//             unlabelled neighbour goes to the lowest claiming partition
//             (atomicMin); frontiers become carry + newly labelled.
#include <algorithm>
#include <mutex>
#include <exception>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <array>
#include <functional>
#include <memory>
#include <thread>
#include <utility>
#include <vector>

#include "phase1.hpp"
#include "phase2.hpp"
#include "rows.hpp"
#include "select.hpp"
#include "gen_host.hpp"
#include "opbatch.hpp"

namespace dmolp {
namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr int32_t kNoClaim = 0x7f7f7f7f;  // claims are reset with a 0x7f byte fill
constexpr int kExpUnroll = 4;  // row entries per lane per step, loads issued together

__device__ __forceinline__ bool p1_done(const P1Ctl *c) { return c->done != 0; }

// Every kernel takes the job set and runs on its job's share of the launch
// (see VGrid): one graph, or many graphs in one launch.
using ::dmolp::gload;
__device__ __forceinline__ P1Job gload(const P1Job &v) {
    P1Job r = v;
    r.g = gload(v.g);
    r.lvl1 = gload(v.lvl1);
    r.sel_list = gload(v.sel_list);
    r.F[0] = gload(v.F[0]);
    r.F[1] = gload(v.F[1]);
    r.L = gload(v.L);
    r.claims = gload(v.claims);
    r.touched = gload(v.touched);
    r.touched_cnt = gload(v.touched_cnt);
    r.covered = gload(v.covered);
    r.lab = gload(v.lab);
    r.hubbits = gload(v.hubbits);
    r.deg = gload(v.deg);
    r.best = gload(v.best);
    r.sizes = gload(v.sizes);
    r.seeds = gload(v.seeds);
    r.hubs = gload(v.hubs);
    r.used = gload(v.used);
    r.ctl = gload(v.ctl);
    r.remaining = gload(v.remaining);
    return r;
}

#define P1_JOB()                                   \
    const int job_ = job_of_block(bstart, jobs);   \
    const VGrid vg = vgrid_of(bstart, job_);       \
    const P1Job J = gload(JS[job_]);               \
    (void)vg


__device__ __forceinline__ void mark_bit(uint32_t *bits, int32_t v) {
    const uint32_t m = 1u << (v & 31);
    if (!(bits[v >> 5] & m)) atomicOr(&bits[v >> 5], m);
}

// The labelled bitmap: the expansion tests bits (n / 8 bytes, which stay in
// L2) instead of gathering 4-byte labels.  Every phase-1 kernel that writes
// labels sets their bits in the same kernel (k_p1_rounds_wg rebuilds the
// bitmap from the labels at its end).
__device__ __forceinline__ void set_labelled(uint32_t *lab, int32_t v) {
    atomicOr(&lab[v >> 5], 1u << (v & 31));
}
__device__ __forceinline__ bool is_labelled(const uint32_t *lab, int32_t v) {
    return (lab[v >> 5] >> (v & 31)) & 1u;
}

// ---------------------------------------------------------------- seeds

// The seed search asks only whether a listed hub lies in a ball (seed.cpp:
// 154-169), so the balls mark listed hubs only: their other members cost a
// read of the hub bitmap (n / 8 bytes, L2-resident) instead of an atomic,
// and the atomics left fall on the hubs' few bitmap words.
__device__ __forceinline__ bool is_hub(const uint32_t *hubbits, int32_t v) {
    return (hubbits[v >> 5] >> (v & 31)) & 1u;
}
__global__ void k_hub_bits(const P1Job *JS, const int32_t *bstart, int jobs) {
    P1_JOB();
    for (int64_t i = vthread(vg); i < J.nh; i += vthreads(vg)) {
        const int32_t h = J.hubs[i];
        atomicOr(&J.hubbits[h >> 5], 1u << (h & 31));
    }
}

// Level 0 and 1 of the ball of seeds[j] (seed.cpp:26-29, 39-95).
__global__ void k_ball1(const P1Job *JS, const int32_t *bstart, int jobs, int step) {
    P1_JOB();
    const GraphDev g = J.g;
    const int32_t *seeds = J.seeds;
    uint32_t *covered = J.covered;
    const ListDev lvl1 = J.lvl1;
    const int heavy_degree = J.heavy_degree;
    const int chunk_edges = J.chunk_edges;
    const P1Ctl *ctl = J.ctl;
    if (step >= J.seed_steps) return;
    const int j = step - 1;
    if (ctl->seed_stop) return;
    const int32_t s = seeds[j];
    if (s < 0) return;
    if (vg.block == 0 && threadIdx.x == 0) mark_bit(covered, s);
    const int64_t e0 = g.off[s], e1 = g.off[s + 1];
    for (int64_t b = e0 + (vwave(vg) * kWave); b < e1; b += vwaves(vg) * kWave) {
        const int64_t e = b + lane_id();
        const bool on = e < e1;
        int32_t u = 0;
        int64_t d = 0;
        if (on) {
            u = g.col[e];
            if (is_hub(J.hubbits, u)) mark_bit(covered, u);
            d = g.off[u + 1] - g.off[u];
        }
        list_append(lvl1, on, u, d, heavy_degree, chunk_edges);
    }
}

// Level 2: every neighbour of a level-1 vertex.
__global__ __launch_bounds__(kThreads) void k_ball2_light(const P1Job *JS, const int32_t *bstart, int jobs, int step) {
    P1_JOB();
    const GraphDev g = J.g;
    const ListDev lvl1 = J.lvl1;
    uint32_t *covered = J.covered;
    const uint32_t *hubbits = J.hubbits;
    const P1Ctl *ctl = J.ctl;
    if (step >= J.seed_steps) return;
    if (ctl->seed_stop) return;
    __shared__ BatchLds s_b[kWavesPerBlock];
    __shared__ int32_t s_mark[kWavesPerBlock][kWave];
    const int w = threadIdx.x / kWave, lane = lane_id();
    const int64_t count = (int64_t)lvl1.cnt[LC_LIGHT];
    for (int64_t b0 = vwave(vg) * kWave; b0 < count; b0 += vwaves(vg) * kWave) {
        int32_t v, deg;
        bool valid;
        const int32_t total = batch_load(g, lvl1.light, count, b0, s_b[w], v, deg, valid);
        s_mark[w][lane] = -1;
        const int32_t rend = s_b[w].incl[lane], rstart = rend - deg;  // this lane's row
        wave_sync();
        // kExpUnroll entries a lane per step: columns, then the bitmap words,
        // each stage's loads issued together; the marks go out without
        // waiting for their results.
        for (int32_t q0 = 0; q0 < total; q0 += kWave * kExpUnroll) {
            int32_t u[kExpUnroll];
            bool on[kExpUnroll];
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) {
                const int32_t q = q0 + j * kWave + lane;
                on[j] = q < total;
                const int ws = batch_slot_window(s_mark[w], q0 + j * kWave, rstart, rend);
                u[j] = g.col[on[j] ? s_b[w].base[on[j] ? ws : 0] + q : 0];
            }
            bool hub[kExpUnroll];
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) hub[j] = on[j] && is_hub(hubbits, u[j]);
            uint32_t word[kExpUnroll];
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) word[j] = covered[hub[j] ? u[j] >> 5 : 0];
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) {
                const uint32_t m = 1u << (u[j] & 31);
                if (hub[j] && !(word[j] & m)) atomicOr(&covered[u[j] >> 5], m);
            }
        }
        wave_sync();
    }
}

__global__ void k_ball2_heavy(const P1Job *JS, const int32_t *bstart, int jobs, int step) {
    P1_JOB();
    const GraphDev g = J.g;
    const ListDev lvl1 = J.lvl1;
    const int chunk_edges = J.chunk_edges;
    uint32_t *covered = J.covered;
    const uint32_t *hubbits = J.hubbits;
    const P1Ctl *ctl = J.ctl;
    if (step >= J.seed_steps) return;
    if (ctl->seed_stop) return;
    const int64_t nch = (int64_t)lvl1.cnt[LC_CHUNKS];
    for (int64_t j = vwave(vg); j < nch; j += vwaves(vg)) {
        int32_t slot, v;
        int64_t e0, e1, deg;
        chunk_bounds(g, lvl1, lvl1.hc[j], chunk_edges, slot, v, e0, e1, deg);
        for (int64_t eb = e0; eb < e1; eb += (int64_t)kWave * kExpUnroll) {
            int32_t u[kExpUnroll];
            bool on[kExpUnroll];
#pragma unroll
            for (int q = 0; q < kExpUnroll; ++q) {
                const int64_t e = eb + q * kWave + lane_id();
                on[q] = e < e1;
                u[q] = g.col[on[q] ? e : e0];
            }
            bool hub[kExpUnroll];
#pragma unroll
            for (int q = 0; q < kExpUnroll; ++q) hub[q] = on[q] && is_hub(hubbits, u[q]);
            uint32_t word[kExpUnroll];
#pragma unroll
            for (int q = 0; q < kExpUnroll; ++q) word[q] = covered[hub[q] ? u[q] >> 5 : 0];
#pragma unroll
            for (int q = 0; q < kExpUnroll; ++q) {
                const uint32_t m = 1u << (u[q] & 31);
                if (hub[q] && !(word[q] & m)) atomicOr(&covered[u[q] >> 5], m);
            }
        }
    }
}

// The next seed (seed.cpp:154-169): the unused hub outside every ball with
// the largest degree, ties to the earliest in hub order (key = degree << 32
// | ~index).  The hub list
// holds the top of the degree order read downward, so its degrees never
// rise: that hub is the first unused, uncovered one.  Used and covered
// hubs stay so, so the search resumes where the last one stopped.  One
// block per job; the key (degree, position) goes to k_take_seed.
__global__ __launch_bounds__(kThreads) void k_next_seed(const P1Job *JS, const int32_t *bstart, int jobs,
                                                        int step) {
    P1_JOB();
    const GraphDev g = J.g;
    const int32_t *hubs = J.hubs;
    const int64_t nh = J.nh;
    const uint8_t *used = J.used;
    const uint32_t *covered = J.covered;
    unsigned long long *best = J.best;
    P1Ctl *ctl = J.ctl;
    if (step >= J.seed_steps) return;
    if (ctl->seed_stop) return;
    if (vg.block != 0) return;
    __shared__ int64_t s_first;
    for (int64_t i0 = ctl->seed_cursor; i0 < nh; i0 += blockDim.x) {
        if (threadIdx.x == 0) s_first = nh;
        __syncthreads();
        const int64_t i = i0 + threadIdx.x;
        if (i < nh && !used[i]) {
            const int32_t h = hubs[i];
            if (!(covered[h >> 5] & (1u << (h & 31)))) atomicMin((unsigned long long *)&s_first, (unsigned long long)i);
        }
        __syncthreads();
        const int64_t f = s_first;
        if (f < nh) {
            if (threadIdx.x == 0) {
                const int32_t h = hubs[f];
                const unsigned long long d = (unsigned long long)(g.off[h + 1] - g.off[h]);
                *best = (d << 32) | (0xffffffffull - (unsigned long long)f);
                ctl->seed_cursor = (int32_t)f;
            }
            return;
        }
        __syncthreads();  // s_first is reset by thread 0 next
    }
}

__global__ void k_take_seed(const P1Job *JS, const int32_t *bstart, int jobs, int step) {
    P1_JOB();
    unsigned long long *best = J.best;
    const int32_t *hubs = J.hubs;
    uint8_t *used = J.used;
    int32_t *seeds = J.seeds;
    P1Ctl *ctl = J.ctl;
    const ListDev lvl1 = J.lvl1;
    if (step >= J.seed_steps) return;
    const int j = step;
    if (ctl->seed_stop) return;
    for (int i = 0; i < LC_N; ++i) lvl1.cnt[i] = 0;
    const unsigned long long b = *best;
    *best = 0;
    if (b == 0) {  // seed.cpp:214-217
        ctl->seed_stop = 1;
        ctl->seeds_found = j;
        return;
    }
    const int64_t i = (int64_t)(0xffffffffull - (b & 0xffffffffull));
    seeds[j] = hubs[i];
    used[i] = 1;
    ctl->seeds_found = j + 1;
}

// ---------------------------------------------------------------- expansion

__global__ void k_expand_init(const P1Job *JS, const int32_t *bstart, int jobs) {
    P1_JOB();
    const GraphDev g = J.g;
    const int32_t *seeds = J.seeds;
    const int k = J.k;
    int32_t *L = J.L;
    const RecDev f = J.F[0];
    unsigned long long *sizes = J.sizes;
    P1Ctl *ctl = J.ctl;
    for (int p = threadIdx.x; p < k; p += blockDim.x) {
        const int32_t s = seeds[p];
        L[s] = p;  // partition.cpp:346-352
        set_labelled(J.lab, s);
        f.x[p] = s;
        f.g[p] = p;
        f.d[p] = (int32_t)(g.off[s + 1] - g.off[s]);
        sizes[p] = 1;
    }
    if (threadIdx.x == 0) {
        *f.cnt = (unsigned long long)k;
        ctl->labelled = k;
        ctl->rounds = 0;
        ctl->fill_round = -1;
        ctl->done = (int64_t)k >= g.n ? 1 : 0;
        if (ctl->done) atomicSub(J.remaining, 1);
    }
}

// Round start: stop when everything is labelled (partition.cpp:371); when
// every frontier is empty pick the fill partition (partition.cpp:376-387).
__global__ void k_p1_round_begin(const P1Job *JS, const int32_t *bstart, int jobs, int parity) {
    P1_JOB();
    P1Ctl *ctl = J.ctl;
    const RecDev fcur = parity ? J.F[1] : J.F[0];
    const RecDev fnext = parity ? J.F[0] : J.F[1];
    const ListDev sel = J.sel_list;
    const unsigned long long *sizes = J.sizes;
    const int k = J.k;
    const int64_t n = J.g.n;
    int64_t *touched_cnt = J.touched_cnt;
    if (p1_done(ctl)) return;
    if (ctl->labelled >= n) {
        ctl->done = 1;
        atomicSub(J.remaining, 1);
        return;
    }
    *fnext.cnt = 0;
    for (int i = 0; i < LC_N; ++i) sel.cnt[i] = 0;
    *touched_cnt = 0;
    ctl->fill_part = -1;
    if (*fcur.cnt == 0) {
        int b = 0;
        for (int p = 0; p < k; ++p)
            if (sizes[p] < sizes[b] || (sizes[p] == sizes[b] && p < b)) b = p;
        ctl->fill_part = b;
        if (ctl->fill_round < 0) ctl->fill_round = ctl->rounds;
    }
}

// partition.cpp:379-392: every unlabelled vertex joins the fill partition
// and its frontier.
__global__ void k_p1_fill(const P1Job *JS, const int32_t *bstart, int jobs, int parity) {
    P1_JOB();
    const GraphDev g = J.g;
    int32_t *L = J.L;
    const RecDev fcur = parity ? J.F[1] : J.F[0];
    unsigned long long *sizes = J.sizes;
    P1Ctl *ctl = J.ctl;
    if (p1_done(ctl)) return;
    const int b = ctl->fill_part;
    if (b < 0) return;  // uniform per job, so per block
    // kFillItems vertices a thread per step, the block's new members
    // reserved with one atomic (one per wave serialises on the counter).
    constexpr int kFillItems = 8;
    __shared__ BlockAppendLds s_app;
    const int64_t span = (int64_t)blockDim.x * kFillItems;
    int64_t added = 0;
    for (int64_t b0 = vg.block * span; b0 < g.n; b0 += vg.blocks * span) {
        bool un[kFillItems];
        int32_t nu = 0;
#pragma unroll
        for (int j = 0; j < kFillItems; ++j) {
            const int64_t v = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            const int32_t lv = L[v < g.n ? v : 0];
            un[j] = v < g.n && lv == -1;
            nu += un[j] ? 1 : 0;
        }
        int64_t pos = block_reserve(nu, fcur.cnt, s_app);
#pragma unroll
        for (int j = 0; j < kFillItems; ++j) {
            if (!un[j]) continue;
            const int64_t v = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            L[v] = b;
            set_labelled(J.lab, (int32_t)v);
            fcur.x[pos] = (int32_t)v;
            fcur.g[pos] = b;
            fcur.d[pos] = (int32_t)(g.off[v + 1] - g.off[v]);
            ++pos;
            ++added;
        }
    }
    added = wave_sum(added);
    if (lane_id() == 0 && added) {
        atomicAdd(&sizes[b], (unsigned long long)added);
        atomicAdd((unsigned long long *)&ctl->labelled, (unsigned long long)added);
    }
}

// Claims of one round (partition.cpp:451-476): an entry (u selected by part
// p, neighbour v) with v unlabelled lowers claims[v] to p; the thread whose
// atomicMin replaced "no claim" records v as touched.  Touched vertices are
// staged per wave in LDS and reserved in the global list once per
// kTouchStage - kWave entries (one counter hit by every wave each step
// serialises at the L2).
constexpr int kTouchStage = 1024;

struct TouchLds {
    int32_t q[kTouchStage];
    int32_t n;
};

// Wave-level: every lane calls it.
__device__ __forceinline__ void touch_flush(TouchLds &t, int32_t *touched, int64_t *touched_cnt) {
    wave_sync();
    const int32_t c = t.n;
    unsigned long long base = 0;
    if (lane_id() == 0 && c) base = atomicAdd((unsigned long long *)touched_cnt, (unsigned long long)c);
    base = __shfl(base, 0, kWave);
    for (int i = lane_id(); i < c; i += kWave) touched[base + (unsigned long long)i] = t.q[i];
    wave_sync();
    if (lane_id() == 0) t.n = 0;
    wave_sync();
}

// Stage the lanes with `first` set (every lane calls it); flushes when the
// next step could overflow the stage.
__device__ __forceinline__ void touch_stage(bool first, int32_t v, TouchLds &t, int32_t *touched,
                                            int64_t *touched_cnt) {
    const uint64_t m = __ballot(first);
    if (m == 0) return;  // wave-uniform
    const int32_t pos = t.n;
    if (first) t.q[pos + lanes_below(m)] = v;
    wave_sync();
    if (lane_id() == 0) t.n = pos + __popcll(m);
    wave_sync();
    if (t.n > kTouchStage - kWave) touch_flush(t, touched, touched_cnt);
}

// partition.cpp:451-476: claims by the selected frontier members.
__global__ __launch_bounds__(kThreads) void k_expand_light(const P1Job *JS, const int32_t *bstart, int jobs) {
    P1_JOB();
    const GraphDev g = J.g;
    const ListDev sel = J.sel_list;
    const int32_t *L = J.L;
    const uint32_t *lab = J.lab;
    int32_t *claims = J.claims;
    int32_t *touched = J.touched;
    int64_t *touched_cnt = J.touched_cnt;
    const P1Ctl *ctl = J.ctl;
    if (p1_done(ctl)) return;
    __shared__ BatchLds s_b[kWavesPerBlock];
    __shared__ TouchLds s_t[kWavesPerBlock];
    __shared__ int32_t s_mark[kWavesPerBlock][kWave];
    const int w = threadIdx.x / kWave, lane = lane_id();
    if (lane == 0) s_t[w].n = 0;
    wave_sync();
    const int64_t count = (int64_t)sel.cnt[LC_LIGHT];
    for (int64_t b0 = vwave(vg) * kWave; b0 < count; b0 += vwaves(vg) * kWave) {
        int32_t v, deg;
        bool valid;
        const int32_t total = batch_load(g, sel.light, count, b0, s_b[w], v, deg, valid);
        s_b[w].aux[lane] = valid ? L[v] : 0;
        s_mark[w][lane] = -1;
        const int32_t rend = s_b[w].incl[lane], rstart = rend - deg;  // this lane's row
        wave_sync();
        for (int32_t q0 = 0; q0 < total; q0 += kWave * kExpUnroll) {
            // Columns, then the neighbours' labels and claims, each stage's
            // loads from in-bounds indices so they issue together.
            int32_t u[kExpUnroll], p[kExpUnroll], cl[kExpUnroll];
            bool on[kExpUnroll], lb[kExpUnroll];
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) {
                const int32_t q = q0 + j * kWave + lane;
                on[j] = q < total;
                const int ws = batch_slot_window(s_mark[w], q0 + j * kWave, rstart, rend);
                const int slot = on[j] ? ws : 0;
                const int64_t e = s_b[w].base[slot] + q;
                u[j] = g.col[on[j] ? e : 0];
                p[j] = s_b[w].aux[slot];
            }
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) lb[j] = is_labelled(lab, u[j]);
            // Claims only of unlabelled neighbours: a labelled one's lane
            // reads claims[0] (one line shared by the wave), so most of a
            // late round's entries cost one bitmap read, not a gather.
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) cl[j] = claims[lb[j] ? 0 : u[j]];
            // The claims' atomics issued together, then their results.
            bool want[kExpUnroll];
            int32_t prev[kExpUnroll];
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) {
                want[j] = on[j] && !lb[j] && cl[j] > p[j];
                prev[j] = want[j] ? atomicMin(&claims[u[j]], p[j]) : 0;
            }
            // (Keeps each result's compare out of its atomic's branch, where
            // it would wait for that atomic before the next one issues.)
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) asm volatile("" : "+v"(prev[j]));
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j)
                touch_stage(want[j] && prev[j] == kNoClaim, u[j], s_t[w], touched, touched_cnt);
        }
        wave_sync();
    }
    touch_flush(s_t[w], touched, touched_cnt);
}

__global__ __launch_bounds__(kThreads) void k_expand_heavy(const P1Job *JS, const int32_t *bstart, int jobs) {
    P1_JOB();
    const GraphDev g = J.g;
    const ListDev sel = J.sel_list;
    const int chunk_edges = J.chunk_edges;
    const int32_t *L = J.L;
    const uint32_t *lab = J.lab;
    int32_t *claims = J.claims;
    int32_t *touched = J.touched;
    int64_t *touched_cnt = J.touched_cnt;
    const P1Ctl *ctl = J.ctl;
    if (p1_done(ctl)) return;
    __shared__ TouchLds s_t[kWavesPerBlock];
    const int w = threadIdx.x / kWave, lane = lane_id();
    if (lane == 0) s_t[w].n = 0;
    wave_sync();
    const int64_t nch = (int64_t)sel.cnt[LC_CHUNKS];
    for (int64_t j0 = vwave(vg); j0 < nch; j0 += vwaves(vg)) {
        int32_t slot, v;
        int64_t e0, e1, deg;
        chunk_bounds(g, sel, sel.hc[j0], chunk_edges, slot, v, e0, e1, deg);
        const int32_t p = L[v];
        for (int64_t eb = e0; eb < e1; eb += (int64_t)kWave * kExpUnroll) {
            int32_t u[kExpUnroll], cl[kExpUnroll];
            bool on[kExpUnroll], lb[kExpUnroll];
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) {
                const int64_t e = eb + j * kWave + lane;
                on[j] = e < e1;
                u[j] = g.col[on[j] ? e : e0];
            }
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) lb[j] = is_labelled(lab, u[j]);
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) cl[j] = claims[lb[j] ? 0 : u[j]];
            bool want[kExpUnroll];
            int32_t prev[kExpUnroll];
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) {
                want[j] = on[j] && !lb[j] && cl[j] > p;
                prev[j] = want[j] ? atomicMin(&claims[u[j]], p) : 0;
            }
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j) asm volatile("" : "+v"(prev[j]));
#pragma unroll
            for (int j = 0; j < kExpUnroll; ++j)
                touch_stage(want[j] && prev[j] == kNoClaim, u[j], s_t[w], touched, touched_cnt);
        }
    }
    touch_flush(s_t[w], touched, touched_cnt);
}

// partition.cpp:539-597: label the claimed vertices, extend the frontiers,
// recount sizes.
__global__ __launch_bounds__(kThreads) void k_p1_apply(const P1Job *JS, const int32_t *bstart, int jobs, int parity) {
    P1_JOB();
    const GraphDev g = J.g;
    int32_t *L = J.L;
    int32_t *claims = J.claims;
    const int32_t *touched = J.touched;
    const int64_t *touched_cnt = J.touched_cnt;
    const RecDev fnext = parity ? J.F[0] : J.F[1];
    unsigned long long *sizes = J.sizes;
    const int k = J.k;
    P1Ctl *ctl = J.ctl;
    if (p1_done(ctl)) return;
    __shared__ unsigned long long s_sz[DMOLP_MAX_K];
    for (int i = threadIdx.x; i < k; i += blockDim.x) s_sz[i] = 0;
    __syncthreads();
    const int64_t count = *touched_cnt;
    for (int64_t b0 = vwave(vg) * kWave; b0 < count; b0 += vwaves(vg) * kWave) {
        const int64_t i = b0 + lane_id();
        const bool on = i < count;
        int32_t v = 0, p = 0;
        if (on) {
            v = touched[i];
            p = claims[v];
            claims[v] = kNoClaim;
            L[v] = p;
            set_labelled(J.lab, v);
            atomicAdd(&s_sz[p], 1ull);
        }
        const int64_t pos = wave_append(on, fnext.cnt);
        if (on) {
            fnext.x[pos] = v;
            fnext.g[pos] = p;
            fnext.d[pos] = (int32_t)(g.off[v + 1] - g.off[v]);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += blockDim.x)
        if (s_sz[i]) atomicAdd(&sizes[i], s_sz[i]);
    if (vg.block == 0 && threadIdx.x == 0)
        atomicAdd((unsigned long long *)&ctl->labelled, (unsigned long long)count);
}

__global__ void k_p1_round_end(const P1Job *JS, const int32_t *bstart, int jobs) {
    P1_JOB();
    P1Ctl *ctl = J.ctl;
    if (p1_done(ctl)) return;
    ctl->rounds += 1;
}

// ------------------------------------------------------- rounds in one workgroup

// Rounds in which a job's every part keeps its whole frontier (at most theta
// members per part, so the per-part selection takes them all,
// partition.cpp:419-425), every member's row is light and the frontier has
// at most kWgMaxMembers members, run for that job by one workgroup: its
// begin step, fill, claims and their application, separated by workgroup
// barriers only (a grid-wide barrier costs more than such a round).  Jobs
// are independent, one block each.  A job stops before the first round that
// needs a real selection, a heavy row or more row entries than one
// workgroup gets through faster than the separate launches, and leaves it, with
// its frontier moved back to F[0], to the launches of p1_expand_rounds,
// which start every batch there.  The state left after every round is what
// those launches leave.  Claims and labels, which other waves change within
// the launch, are read and written with device-scope atomic operations
// (they bypass the vector L1).
constexpr int kWgThreads = 1024;
constexpr int kWgProbeRounds = 32;  // first one-workgroup attempt of a batch (even)
constexpr int kWgWaves = kWgThreads / kWave;
constexpr int64_t kWgMaxMembers = 16384;  // frontier members of a round run here
constexpr int64_t kWgMaxEntries = 32768;  // and their row entries

__device__ __forceinline__ int32_t ld_rel(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(int32_t *p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// out[0]: largest number of rounds a job ran here; out[1]: jobs that stopped
// at a round these blocks cannot run.
__global__ __launch_bounds__(kWgThreads) void k_p1_rounds_wg(const P1Job *JS, int max_rounds, int32_t *out) {
    __shared__ P1Job s_job;
    __shared__ int32_t s_cnt[2][DMOLP_MAX_K];   // per-part frontier counts by parity
    __shared__ unsigned long long s_sz[DMOLP_MAX_K];
    __shared__ BatchLds s_b[kWgWaves];
    __shared__ int32_t s_maxdeg[2];
    __shared__ unsigned long long s_entries[2];  // row entries of the frontier by parity
    __shared__ int32_t s_state;
    __shared__ unsigned long long s_touched, s_next, s_added;
    const int tid = threadIdx.x, w = tid / kWave, lane = lane_id();
    {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(JS + blockIdx.x);
        uint32_t *dst = reinterpret_cast<uint32_t *>(&s_job);
        for (int i = tid; i < (int)(sizeof(P1Job) / 4); i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const P1Job &J = s_job;
    P1Ctl *ctl = J.ctl;
    const int k = J.k;
    const int64_t n = J.g.n;
    // out[2 + job]: 0 the job ran max_rounds rounds here, 1 it stopped at a
    // round the separate launches must run, 2 it is finished.
    if (ctl->done) {  // uniform: nothing writes it before this point
        if (tid == 0) out[2 + blockIdx.x] = 2;
        return;
    }
    if ((int64_t)*J.F[0].cnt > kWgMaxMembers) {  // too large to count here
        if (tid == 0) {
            atomicAdd(&out[1], 1);
            out[2 + blockIdx.x] = 1;
        }
        return;
    }
    // The entry frontier (F[0]): per-part counts and largest degree.
    for (int p = tid; p < k; p += blockDim.x) {
        s_cnt[0][p] = 0;
        s_cnt[1][p] = 0;
        s_sz[p] = J.sizes[p];
    }
    if (tid < 2) {
        s_maxdeg[tid] = 0;
        s_entries[tid] = 0;
    }
    __syncthreads();
    {
        const RecDev f = J.F[0];
        const int64_t c = (int64_t)*f.cnt;
        for (int64_t i = tid; i < c; i += blockDim.x) {
            atomicAdd(&s_cnt[0][f.g[i]], 1);
            atomicMax(&s_maxdeg[0], f.d[i]);
            atomicAdd(&s_entries[0], (unsigned long long)f.d[i]);
        }
    }
    __syncthreads();
    int t = 0, st = 0;
    for (; t < max_rounds; ++t) {
        const int par = t & 1;
        const RecDev cur = J.F[par], nxt = J.F[par ^ 1];
        // Begin (k_p1_round_begin) and whether the round runs here:
        // 0 runs, 1 stops (the round is left to the separate launches),
        // 2 the job is finished, 3 fill round.
        if (tid == 0) {
            int s = 0;
            int b = -1;
            if (ctl->labelled >= n) {
                ctl->done = 1;
                atomicSub(J.remaining, 1);
                s = 2;
            } else {
                const int64_t c = (int64_t)*cur.cnt;
                if (c == 0) {
                    b = 0;
                    for (int p = 0; p < k; ++p)
                        if (s_sz[p] < s_sz[b] || (s_sz[p] == s_sz[b] && p < b)) b = p;
                    s = 3;
                } else if (c > kWgMaxMembers || s_entries[par] > (unsigned long long)kWgMaxEntries ||
                           s_maxdeg[par] > J.heavy_degree) {
                    s = 1;
                } else {
                    for (int p = 0; p < k; ++p)
                        if (s_cnt[par][p] > J.theta) s = 1;
                }
                if (s != 1) {
                    *nxt.cnt = 0;
                    for (int i = 0; i < LC_N; ++i) J.sel_list.cnt[i] = 0;
                    *J.touched_cnt = 0;
                    ctl->fill_part = b;
                    if (s == 3 && ctl->fill_round < 0) ctl->fill_round = (int32_t)ctl->rounds;
                }
            }
            s_state = s;
            s_touched = 0;
            s_next = 0;
            s_added = 0;
            s_maxdeg[par ^ 1] = 0;
            s_entries[par ^ 1] = 0;
        }
        for (int p = tid; p < k; p += blockDim.x) s_cnt[par ^ 1][p] = 0;
        __syncthreads();
        st = s_state;
        if (st == 1 || st == 2) break;
        if (st == 3) {
            // Fill (k_p1_fill): every unlabelled vertex joins part b; nothing
            // is left to claim this round.
            const int b = ctl->fill_part;
            for (int64_t b0 = (int64_t)w * kWave; b0 < n; b0 += (int64_t)kWgWaves * kWave) {
                const int64_t v = b0 + lane;
                const bool un = v < n && ld_rel(J.L + v) == -1;
                const int64_t pos = wave_append(un, cur.cnt);
                if (un) {
                    st_rel(J.L + v, b);
                    cur.x[pos] = (int32_t)v;
                    cur.g[pos] = b;
                    cur.d[pos] = (int32_t)(J.g.off[v + 1] - J.g.off[v]);
                }
                const uint64_t m = __ballot(un);
                if (lane == 0 && m) atomicAdd(&s_added, (unsigned long long)__popcll(m));
            }
            __syncthreads();
            if (tid == 0) {
                s_sz[b] += s_added;
                atomicAdd((unsigned long long *)&ctl->labelled, s_added);
            }
        } else {
            // Claims (k_expand_light) of every member.
            const int64_t c = (int64_t)*cur.cnt;
            for (int64_t b0 = (int64_t)w * kWave; b0 < c; b0 += (int64_t)kWgWaves * kWave) {
                int32_t v, deg;
                bool valid;
                const int32_t total = batch_load(J.g, cur.x, c, b0, s_b[w], v, deg, valid);
                s_b[w].aux[lane] = valid ? ld_rel(J.L + v) : 0;
                wave_sync();
                // kExpUnroll entries a lane per step, each stage's loads
                // issued together (columns, then labels, then the claims of
                // unlabelled neighbours), so a round waits for about three
                // memory latencies rather than three per entry.
                for (int32_t q0 = 0; q0 < total; q0 += kWave * kExpUnroll) {
                    int32_t u[kExpUnroll], p[kExpUnroll], lv[kExpUnroll], cl[kExpUnroll];
                    bool on[kExpUnroll];
#pragma unroll
                    for (int j = 0; j < kExpUnroll; ++j) {
                        const int32_t q = q0 + j * kWave + lane;
                        on[j] = q < total;
                        const int slot = on[j] ? batch_slot(s_b[w], q) : 0;
                        u[j] = J.g.col[on[j] ? s_b[w].base[slot] + q : 0];
                        p[j] = s_b[w].aux[slot];
                    }
#pragma unroll
                    for (int j = 0; j < kExpUnroll; ++j) lv[j] = ld_rel(J.L + u[j]);
#pragma unroll
                    for (int j = 0; j < kExpUnroll; ++j) cl[j] = ld_rel(J.claims + (lv[j] == -1 ? u[j] : 0));
#pragma unroll
                    for (int j = 0; j < kExpUnroll; ++j)
                        if (on[j] && lv[j] == -1 && cl[j] > p[j] && atomicMin(J.claims + u[j], p[j]) == kNoClaim)
                            J.touched[atomicAdd(&s_touched, 1ull)] = u[j];
                }
                wave_sync();
            }
            __syncthreads();
            // Application (k_p1_apply).
            const int64_t nt = (int64_t)s_touched;
            for (int64_t i = tid; i < nt; i += blockDim.x) {
                const int32_t v = J.touched[i];
                const int32_t p = ld_rel(J.claims + v);
                st_rel(J.claims + v, kNoClaim);
                st_rel(J.L + v, p);
                const int32_t d = (int32_t)(J.g.off[v + 1] - J.g.off[v]);
                const unsigned long long pos = atomicAdd(&s_next, 1ull);
                nxt.x[pos] = v;
                nxt.g[pos] = p;
                nxt.d[pos] = d;
                atomicAdd(&s_cnt[par ^ 1][p], 1);
                atomicAdd(&s_sz[p], 1ull);
                atomicMax(&s_maxdeg[par ^ 1], d);
                atomicAdd(&s_entries[par ^ 1], (unsigned long long)d);
            }
            __syncthreads();
            if (tid == 0) {
                *nxt.cnt = s_next;
                *J.touched_cnt = nt;
                atomicAdd((unsigned long long *)&ctl->labelled, (unsigned long long)nt);
            }
        }
        if (tid == 0) ctl->rounds += 1;  // k_p1_round_end
    }
    __syncthreads();
    // The labelled bitmap the separate launches test, rebuilt from the labels
    // once the rounds here have changed them (one atomic a labelled vertex
    // inside the rounds cost more than this pass).
    if (t > 0 && st != 2) {
        const int64_t nw = (n + 31) / 32;
        for (int64_t i = tid; i < nw; i += blockDim.x) {
            uint32_t word = 0;
            for (int b = 0; b < 32; ++b) {
                const int64_t v = i * 32 + b;
                if (v < n && ld_rel(J.L + v) != -1) word |= 1u << b;
            }
            J.lab[i] = word;
        }
    }
    // Leave the frontier in F[0] for the separate launches.
    if ((t & 1) && st != 2) {
        const RecDev src = J.F[1], dst = J.F[0];
        const int64_t c = (int64_t)*src.cnt;
        for (int64_t i = tid; i < c; i += blockDim.x) {
            dst.x[i] = src.x[i];
            dst.g[i] = src.g[i];
            dst.d[i] = src.d[i];
        }
        if (tid == 0) *dst.cnt = (unsigned long long)c;
    }
    for (int p = tid; p < k; p += blockDim.x) J.sizes[p] = s_sz[p];
    if (tid == 0) {
        atomicMax(&out[0], t);
        if (st == 1) atomicAdd(&out[1], 1);
        out[2 + blockIdx.x] = st == 1 ? 1 : (st == 2 ? 2 : 0);
    }
}

// random.cpp:71-74: label = global id mod k.
__global__ void k_mod_labels(int64_t n, int k, int32_t *L) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        L[v] = (int32_t)(v % k);
}

__global__ void k_degrees(const P1Job *JS, const int32_t *bstart, int jobs) {
    P1_JOB();
    const GraphDev g = J.g;
    int32_t *deg = J.deg;
    for (int64_t v = vthread(vg); v < g.n;
         v += vthreads(vg))
        deg[v] = (int32_t)(g.off[v + 1] - g.off[v]);
}

}  // namespace

void phase1_random(dmolp_ctx *ctx, dmolp_graph *g, int k) {
    DMOLP_REQUIRE(k >= 1 && k <= DMOLP_MAX_K, DMOLP_E_ARG, "k outside [1, 1024]");
    if (!g->labels.p) g->labels.alloc(sizeof(int32_t) * (size_t)g->n);
    hipLaunchKernelGGL(k_mod_labels, dim3(ctx->cu_count * 4), dim3(256), 0, ctx->stream, g->n, k,
                       g->labels.as<int32_t>());
    DMOLP_HIP_CHECK(hipGetLastError());
    g->labels_valid = true;
}

// ---------------------------------------------------------------- host side

Phase1Work &phase1_work(dmolp_ctx *ctx, dmolp_graph *g, int k) {
    graph_prepare_bins(ctx, g);
    if (!g->p1) g->p1.reset(new Phase1Work());
    Phase1Work &w = *g->p1;
    const int64_t n = g->n;
    if (!w.claims.p) {
        w.claims.alloc(sizeof(int32_t) * (size_t)n);
        DMOLP_HIP_CHECK(hipMemsetAsync(w.claims.p, 0x7f, w.claims.bytes, ctx->stream));  // kNoClaim
        w.touched.alloc(sizeof(int32_t) * (size_t)n);
        w.touched_cnt.alloc(sizeof(int64_t) * 2);
        w.F[0].alloc(n);
        w.F[1].alloc(n);
        w.covered.alloc(sizeof(uint32_t) * (size_t)((n + 31) / 32));
        w.lab.alloc(sizeof(uint32_t) * (size_t)((n + 31) / 32));
        w.hubbits.alloc(sizeof(uint32_t) * (size_t)((n + 31) / 32));
        w.deg.alloc(sizeof(int32_t) * (size_t)n);
        w.ctl.alloc(sizeof(P1Ctl));
        w.best.alloc(sizeof(unsigned long long));
    }
    if (w.cap_heavy != g->n_heavy || w.cap_chunks != g->n_heavy_chunks) {
        w.sel_list.alloc(n, g->n_heavy, g->n_heavy_chunks);
        w.lvl1.alloc(n, g->n_heavy, g->n_heavy_chunks);
        w.cap_heavy = g->n_heavy;
        w.cap_chunks = g->n_heavy_chunks;
    }
    if (w.k != k) {
        w.k = k;
        w.sizes.alloc(sizeof(unsigned long long) * (size_t)k);
        w.seeds.alloc(sizeof(int32_t) * (size_t)k);
    }
    w.sel.prepare(ctx, k, n, g->max_degree);
    return w;
}

namespace {

__global__ void k_p1_set_remaining(int32_t *remaining, int32_t v) { *remaining = v; }

// done flag of every job of a set, gathered for one host copy.
__global__ void k_p1_gather_done(const P1Job *JS, int jobs, int32_t *out) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < jobs; j += gridDim.x * blockDim.x)
        out[j] = gload(JS[j].ctl)->done;
}

}  // namespace

P1Job p1_job(dmolp_ctx *ctx, dmolp_graph *g, Phase1Work &w, int32_t *remaining) {
    P1Job j{};
    j.g = g->dev();
    j.lvl1 = w.lvl1.dev;
    j.sel_list = w.sel_list.dev;
    j.F[0] = w.F[0].dev;
    j.F[1] = w.F[1].dev;
    j.L = g->labels.as<int32_t>();
    j.claims = w.claims.as<int32_t>();
    j.touched = w.touched.as<int32_t>();
    j.touched_cnt = w.touched_cnt.as<int64_t>();
    j.covered = w.covered.as<uint32_t>();
    j.lab = w.lab.as<uint32_t>();
    j.hubbits = w.hubbits.as<uint32_t>();
    j.deg = w.deg.as<int32_t>();
    j.best = w.best.as<unsigned long long>();
    j.sizes = w.sizes.as<unsigned long long>();
    j.seeds = w.seeds.as<int32_t>();
    j.hubs = w.hubs.as<int32_t>();
    j.used = w.used.as<uint8_t>();
    j.nh = 0;
    j.ctl = w.ctl.as<P1Ctl>();
    j.remaining = remaining;
    j.k = w.k;
    j.seed_steps = 1;
    j.heavy_degree = ctx->tune.heavy_degree;
    j.chunk_edges = ctx->tune.chunk_edges;
    return j;
}

void p1_hub_search(P1Stage &h, int threads) {
    if (h.dev_hubs) return;  // p1_hub_search_device
    const auto t0 = std::chrono::steady_clock::now();
    h.fs = first_seed_host(h.deg, threads);
    // Only the first hub_prefix hubs are ordered: a seed search walks the
    // hub list from the top and stops at its k-th seed, which on graphs with
    // many hubs (near-regular ones, where every vertex is a hub) lies far
    // above that; a search that runs out of listed hubs is repeated with
    // the whole list (p1_seeds).
    h.hubs = find_hubs_host(h.deg, threads, h.hub_prefix > 0 ? h.hub_prefix : -1, &h.n_hubs);
    h.hub_deg.resize(h.hubs.size());
    for (size_t i = 0; i < h.hubs.size(); ++i) h.hub_deg[i] = h.deg[(size_t)h.hubs[i]];
    const auto t1 = std::chrono::steady_clock::now();
    h.host_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
}

void p1_hub_search_device(dmolp_ctx *ctx, std::vector<P1Stage *> &H) {
    std::vector<HubSearch> hs;
    std::vector<P1Stage *> on;
    for (P1Stage *h : H)
        if (h->dev_hubs) on.push_back(h);
    if (on.empty()) return;
    const auto t0 = std::chrono::steady_clock::now();
    hs.resize(on.size());
    std::vector<HubSearch *> ptr;
    for (size_t i = 0; i < on.size(); ++i) {
        hs[i].n = on[i]->g->n;
        hs[i].deg = on[i]->w->deg.as<int32_t>();
        hs[i].limit = on[i]->hub_prefix > 0 ? on[i]->hub_prefix : -1;  // as p1_hub_search
        ptr.push_back(&hs[i]);
    }
    hub_search_device(ctx, ptr, ctx->hub_host_max, std::min(16, host_threads()));
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (size_t i = 0; i < on.size(); ++i) {
        P1Stage &h = *on[i];
        h.fs = hs[i].fs;
        h.hubs = std::move(hs[i].hubs);
        h.hub_deg = std::move(hs[i].hub_deg);
        h.n_hubs = hs[i].n_hubs;
        h.host_ms = ms;
    }
}

// Candidates of the seed fallback (p1_seeds, only for a search that found
// fewer than k seeds): the first 2k vertices by (degree descending, id
// ascending).  At most k of them can already be seeds, so the first
// k - found unchosen vertices of the full order are here.
void p1_fallback_candidates(P1Stage &h, int threads) {
    const int64_t n = (int64_t)h.deg.size();
    const size_t m = (size_t)std::min<int64_t>(n, 2 * (int64_t)h.k);
    const std::vector<int32_t> &deg = h.deg;
    auto better = [&](int32_t a, int32_t b) {
        if (deg[(size_t)a] != deg[(size_t)b]) return deg[(size_t)a] > deg[(size_t)b];
        return a < b;
    };
    // Bounded heaps with the worst kept candidate on top, one pass over a
    // slice of the ids each; the order is total (ids break ties), so the
    // first m of the slices' candidates together are the first m overall.
    const int T = n >= (int64_t(1) << 20) ? std::max(1, threads) : 1;
    std::vector<std::vector<int32_t>> part((size_t)T);
    run_threads(T, [&](int t) {
        std::vector<int32_t> &heap = part[(size_t)t];
        heap.reserve(m);
        for (int64_t v = n * t / T; v < n * (t + 1) / T; ++v) {
            if (heap.size() < m) {
                heap.push_back((int32_t)v);
                std::push_heap(heap.begin(), heap.end(), better);
            } else if (m > 0 && better((int32_t)v, heap.front())) {
                std::pop_heap(heap.begin(), heap.end(), better);
                heap.back() = (int32_t)v;
                std::push_heap(heap.begin(), heap.end(), better);
            }
        }
    });
    std::vector<int32_t> all;
    for (const auto &p : part) all.insert(all.end(), p.begin(), p.end());
    std::sort(all.begin(), all.end(), better);
    if (all.size() > m) all.resize(m);
    h.top = std::move(all);
}

namespace {

// Hubs and first seed of every job searched on the host; jobs in parallel
// threads.  `during` (may be empty) runs on the calling thread while the
// searches run on worker threads.  Jobs are taken largest first from a shared cursor, so
// the longest search starts at once and the threads finish close together.
void host_hub_search(std::vector<P1Stage> &H, const std::function<void()> &during) {
    const size_t G = H.size();
    size_t T = (size_t)std::min(16, host_threads());
    if (T > G) T = G;
    // The first exception of a worker or of `during` is rethrown once every
    // thread has been joined.
    std::exception_ptr err;
    std::mutex err_mu;
    const auto keep = [&](std::exception_ptr e) {
        std::lock_guard<std::mutex> lk(err_mu);
        if (!err) err = e;
    };
    std::vector<std::thread> pool;
    std::vector<size_t> order(G);  // outlives the workers: joined below
    std::atomic<size_t> next{0};
    if (T <= 1) {  // one job: its passes and long partitions use the threads
        const auto all = [&] {
            try {
                for (size_t j = 0; j < G; ++j) p1_hub_search(H[j], std::min(16, host_threads()));
            } catch (...) {
                keep(std::current_exception());
            }
        };
        if (during) pool.emplace_back(all);
        else all();
    } else {
        for (size_t j = 0; j < G; ++j) order[j] = j;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return H[a].deg.size() > H[b].deg.size(); });
        // A job of at least 2^22 vertices splits its own long partitions over 4
        // threads: the largest search is the critical path.
        for (size_t t = 0; t < T; ++t)
            pool.emplace_back([&]() {
                try {
                    for (size_t i = next.fetch_add(1); i < G; i = next.fetch_add(1)) {
                        P1Stage &h = H[order[i]];
                        p1_hub_search(h, h.deg.size() >= (size_t(1) << 22) ? 4 : 1);
                    }
                } catch (...) {
                    keep(std::current_exception());
                }
            });
    }
    if (during) {
        try {
            during();
        } catch (...) {
            keep(std::current_exception());
        }
    }
    for (auto &th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

}  // namespace

void p1_stage_init(dmolp_ctx *ctx, P1Item &it, P1Stage &h) {
    dmolp_graph *g = it.g;
    DMOLP_REQUIRE(it.k >= 1 && it.k <= DMOLP_MAX_K, DMOLP_E_ARG, "k outside [1, 1024]");
    DMOLP_REQUIRE(it.theta > 0, DMOLP_E_ARG, "theta must be positive (theta <= 0 never terminates)");
    DMOLP_REQUIRE((int64_t)it.k <= g->n, DMOLP_E_SEEDS, "k larger than the vertex count");
    h.w = &phase1_work(ctx, g, it.k);
    if (!g->labels.p) g->labels.alloc(sizeof(int32_t) * (size_t)g->n);
    h.g = g;
    h.k = it.k;
    h.theta = it.theta;
    h.hub_prefix = ctx->hub_prefix;
}

void p1_set_build(dmolp_ctx *ctx, std::vector<P1Stage *> &H, P1SetDev &sd) {
    const int G = (int)H.size();
    hipStream_t s = ctx->stream;
    sd.G = G;
    sd.kmax = 0;
    sd.any_heavy = false;
    sd.passes = 0;
    sd.max_n = 0;
    sd.remaining.grow(sizeof(int32_t) * 2);
    std::vector<P1Job> jobs((size_t)G);
    std::vector<SelArgs> sel((size_t)2 * G);
    std::vector<std::array<int32_t, P_NPOL>> blocks((size_t)G);
    for (int j = 0; j < G; ++j) {
        P1Stage &h = *H[(size_t)j];
        Phase1Work &w = *h.w;
        if (!h.have_job) {
            h.job = p1_job(ctx, h.g, w, nullptr);
            h.have_job = true;
        }
        h.job.remaining = sd.remaining.as<int32_t>();
        h.job.theta = h.theta;
        jobs[(size_t)j] = h.job;
        P1Ctl *ctl = w.ctl.as<P1Ctl>();
        TakeRule rule{};
        rule.kind = 1;
        rule.theta = h.theta;
        sel[(size_t)j] = select_args(ctx, w.sel, w.F[0].dev, rule, &ctl->done, w.sel_list.dev, &w.F[1].dev);
        sel[(size_t)G + j] = select_args(ctx, w.sel, w.F[1].dev, rule, &ctl->done, w.sel_list.dev, &w.F[0].dev);
        job_blocks(ctx, h.g->n, h.g->nnz, h.k, h.g->n_heavy, G == 1, blocks[(size_t)j].data());
        sd.any_heavy = sd.any_heavy || h.g->n_heavy > 0;
        sd.passes = std::max(sd.passes, w.sel.passes);
        sd.max_n = std::max(sd.max_n, h.g->n);
        sd.kmax = std::max(sd.kmax, h.k);
    }
    launch_set_build(ctx, sd.ls, blocks);
    sd.jobs.grow(sizeof(P1Job) * (size_t)G);
    sd.sel.grow(sizeof(SelArgs) * 2 * (size_t)G);
    sd.done.grow(sizeof(int32_t) * (size_t)G);
    DMOLP_HIP_CHECK(hipMemcpyAsync(sd.jobs.p, jobs.data(), sizeof(P1Job) * (size_t)G,
                                   hipMemcpyHostToDevice, s));
    DMOLP_HIP_CHECK(hipMemcpyAsync(sd.sel.p, sel.data(), sizeof(SelArgs) * 2 * (size_t)G,
                                   hipMemcpyHostToDevice, s));
    // The host vectors are released below: finish the copies first.
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
}

// The degrees go to the host through the context's pinned staging buffer
// (a copy into pageable memory is staged by the runtime in small pieces, at
// a fraction of the link's rate), then into each job's vector on threads.
void p1_degrees(dmolp_ctx *ctx, std::vector<P1Stage *> &H, P1SetDev &sd) {
    hipStream_t s = ctx->stream;
    const LaunchSet &ls = sd.ls;
    hipLaunchKernelGGL(k_degrees, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, sd.jobs.as<P1Job>(),
                       ls.bs(P_GRID), sd.G);
    DMOLP_HIP_CHECK(hipGetLastError());
    // Jobs searched on the device keep their degrees there.
    std::vector<P1Stage *> Hh;
    for (P1Stage *h : H) {
        h->dev_hubs = h->g->n >= ctx->hub_dev_min;
        h->deg.clear();
        if (!h->dev_hubs) Hh.push_back(h);
    }
    if (Hh.empty()) return;
    size_t total = 0;
    std::vector<size_t> at(Hh.size());
    for (size_t j = 0; j < Hh.size(); ++j) {
        at[j] = total;
        total += (size_t)Hh[j]->g->n;
    }
    int32_t *pin = ctx->pinned_i32(total);
    for (size_t j = 0; j < Hh.size(); ++j)
        DMOLP_HIP_CHECK(hipMemcpyAsync(pin + at[j], Hh[j]->w->deg.p, sizeof(int32_t) * (size_t)Hh[j]->g->n,
                                       hipMemcpyDeviceToHost, s));
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    for (P1Stage *h : Hh) h->deg.resize((size_t)h->g->n);
    // Copy out in slices of about equal size across the jobs.
    const int T = total >= (size_t(1) << 22) ? std::min(16, host_threads()) : 1;
    run_threads(T, [&](int t) {
        const size_t a = total * (size_t)t / (size_t)T, b = total * (size_t)(t + 1) / (size_t)T;
        for (size_t j = 0; j < Hh.size(); ++j) {
            const size_t lo = std::max(a, at[j]), hi = std::min(b, at[j] + Hh[j]->deg.size());
            if (lo < hi)
                std::memcpy(Hh[j]->deg.data() + (lo - at[j]), pin + lo, sizeof(int32_t) * (hi - lo));
        }
    });
}

void p1_seeds(dmolp_ctx *ctx, std::vector<P1Stage *> &H, P1SetDev &sd) {
    hipStream_t s = ctx->stream;
    const int G = (int)H.size();
    int steps = 1;
    OpBatch ob;  // the jobs' clears and uploads together
    for (P1Stage *hp : H) {
        P1Stage &h = *hp;
        Phase1Work &w = *h.w;
        const int k = h.k;
        const int64_t nh = (int64_t)h.hubs.size();
        ob.fill(w.ctl.p, 0, sizeof(P1Ctl));
        ob.fill(w.covered.p, 0, w.covered.bytes);
        ob.fill(w.hubbits.p, 0, w.hubbits.bytes);
        ob.fill(w.best.p, 0, w.best.bytes);
        ob.fill(w.seeds.p, 0xff, w.seeds.bytes);
        ob.upload(w.seeds.p, &h.fs.first, sizeof(int32_t));
        for (int64_t i = 1; i < nh; ++i)  // k_next_seed relies on the order
            DMOLP_REQUIRE(h.hub_deg[(size_t)i] <= h.hub_deg[(size_t)i - 1], DMOLP_E_INTERNAL,
                          "hub degrees out of order");
        h.used.assign((size_t)(nh > 0 ? nh : 1), 0);
        for (int64_t i = 0; i < nh; ++i)
            if (h.hubs[(size_t)i] == h.fs.first) { h.used[(size_t)i] = 1; break; }
        w.hubs.grow(sizeof(int32_t) * (size_t)(nh > 0 ? nh : 1));
        w.used.grow((size_t)(nh > 0 ? nh : 1));
        if (nh > 0) {
            ob.upload(w.hubs.p, h.hubs.data(), sizeof(int32_t) * (size_t)nh);
            ob.upload(w.used.p, h.used.data(), (size_t)nh);
            ob.fill(w.lvl1.cnt.p, 0, w.lvl1.cnt.bytes);
        }
        h.job.hubs = w.hubs.as<int32_t>();
        h.job.used = w.used.as<uint8_t>();
        h.job.nh = nh;
        h.job.seed_steps = nh > 0 ? k : 1;
        steps = std::max(steps, h.job.seed_steps);
    }
    ob.run(ctx);
    p1_set_build(ctx, H, sd);  // uploads the hub fields
    const LaunchSet &ls = sd.ls;
    P1Job *d_jobs = sd.jobs.as<P1Job>();
    hipLaunchKernelGGL(k_hub_bits, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, d_jobs, ls.bs(P_GRID), G);
    for (int step = 1; step < steps; ++step) {
        // Ball of seed step-1, then the next seed.
        hipLaunchKernelGGL(k_ball1, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, d_jobs, ls.bs(P_GRID), G, step);
        hipLaunchKernelGGL(k_ball2_light, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, d_jobs, ls.bs(P_GRID), G, step);
        if (sd.any_heavy)
            hipLaunchKernelGGL(k_ball2_heavy, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, d_jobs, ls.bs(P_GRID), G, step);
        hipLaunchKernelGGL(k_next_seed, dim3(ls.total[P_ONE]), dim3(kThreads), 0, s, d_jobs, ls.bs(P_ONE), G, step);
        hipLaunchKernelGGL(k_take_seed, dim3(ls.total[P_ONE]), dim3(1), 0, s, d_jobs, ls.bs(P_ONE), G, step);
    }
    DMOLP_HIP_CHECK(hipGetLastError());
    std::vector<P1Ctl> hc((size_t)G);
    for (int j = 0; j < G; ++j) {
        P1Stage &h = *H[(size_t)j];
        h.seeds.assign((size_t)h.k, -1);
        ob.download(h.seeds.data(), h.w->seeds.p, sizeof(int32_t) * h.seeds.size());
        ob.download(&hc[(size_t)j], h.w->ctl.p, sizeof(P1Ctl));
    }
    ob.run(ctx);
    // A search that ran out of hubs on a shortened hub list would have gone
    // on into the rest of the list: such jobs are searched again, from the
    // start, with the whole list (a search that found all its seeds stopped
    // inside the prefix, where both lists agree).
    std::vector<P1Stage *> redo;
    std::vector<uint8_t> redone((size_t)G, 0);
    for (int j = 0; j < G; ++j) {
        P1Stage &h = *H[(size_t)j];
        if (hc[(size_t)j].seed_stop && (int64_t)h.hubs.size() < h.n_hubs) {
            redo.push_back(&h);
            redone[(size_t)j] = 1;
        }
    }
    if (!redo.empty()) {
        for (P1Stage *h : redo) h->hub_prefix = 0;
        p1_hub_search_device(ctx, redo);
        for (P1Stage *h : redo) p1_hub_search(*h, std::min(16, host_threads()));
        p1_seeds(ctx, redo, sd);
        p1_set_build(ctx, H, sd);  // the caller's set again
    }
    // Seed fallback candidates of the jobs that found fewer than k seeds, the
    // jobs on host threads (a lone job splits its own pass over them).
    std::vector<P1Stage *> short_of;
    for (int j = 0; j < G; ++j) {
        if (redone[(size_t)j]) continue;
        P1Stage &h = *H[(size_t)j];
        const int k = h.k;
        const int64_t nh = (int64_t)h.hubs.size();
        int32_t found = (nh > 0) ? (hc[(size_t)j].seed_stop ? hc[(size_t)j].seeds_found : k) : 1;
        if (k == 1) found = 1;
        h.found = found;
        h.fallback = 0;
        if (found < k && h.top.empty()) short_of.push_back(&h);
    }
    if (!short_of.empty()) {
        // Graphs searched on the device select their candidates there.
        std::vector<TopSearch> ts;
        std::vector<P1Stage *> on_dev, on_host;
        for (P1Stage *h : short_of) (h->dev_hubs ? on_dev : on_host).push_back(h);
        if (!on_dev.empty()) {
            ts.resize(on_dev.size());
            std::vector<TopSearch *> tp;
            for (size_t i = 0; i < on_dev.size(); ++i) {
                ts[i].n = on_dev[i]->g->n;
                ts[i].deg = on_dev[i]->w->deg.as<int32_t>();
                ts[i].m = std::min<int64_t>(ts[i].n, 2 * (int64_t)on_dev[i]->k);
                tp.push_back(&ts[i]);
            }
            top_by_degree_device(ctx, tp, ctx->hub_host_max, std::min(16, host_threads()));
            for (size_t i = 0; i < on_dev.size(); ++i) on_dev[i]->top = std::move(ts[i].top);
        }
        const int T = std::min(16, host_threads());
        if (on_host.size() == 1) {
            p1_fallback_candidates(*on_host[0], T);
        } else if (!on_host.empty()) {
            std::atomic<size_t> next{0};
            run_threads((int)std::min<size_t>((size_t)T, on_host.size()), [&](int) {
                for (size_t i = next.fetch_add(1); i < on_host.size(); i = next.fetch_add(1))
                    p1_fallback_candidates(*on_host[i], 1);
            });
        }
    }
    for (int j = 0; j < G; ++j) {
        if (redone[(size_t)j]) continue;
        P1Stage &h = *H[(size_t)j];
        const int k = h.k;
        const int32_t found = h.found;
        if (found < k) {
            // Documented extension for seed.cpp:214-217 followed by the
            // out-of-range seeds[p] of partition.cpp:347: the missing seeds are
            // the remaining vertices by (degree descending, id ascending).
            std::vector<int32_t> chosen(h.seeds.begin(), h.seeds.begin() + found);
            std::sort(chosen.begin(), chosen.end());
            int32_t at = found;
            for (int32_t v : h.top) {
                if (at == k) break;
                if (!std::binary_search(chosen.begin(), chosen.end(), v)) h.seeds[(size_t)at++] = v;
            }
            DMOLP_REQUIRE(at == k, DMOLP_E_HIP, "seed fallback ran out of candidates");
            h.fallback = k - found;
            ob.upload(h.w->seeds.p, h.seeds.data(), sizeof(int32_t) * (size_t)k);
        }
    }
    ob.run(ctx);
}

void p1_expand_begin(dmolp_ctx *ctx, std::vector<P1Stage *> &H, P1SetDev &sd) {
    hipStream_t s = ctx->stream;
    OpBatch ob;
    for (P1Stage *h : H) {
        ob.fill(h->g->labels.p, 0xff, sizeof(int32_t) * (size_t)h->g->n);
        ob.fill(h->w->lab.p, 0, h->w->lab.bytes);
        ob.fill(h->w->sizes.p, 0, h->w->sizes.bytes);
        ob.fill(h->w->ctl.p, 0, sizeof(P1Ctl));
    }
    ob.run(ctx);
    hipLaunchKernelGGL(k_p1_set_remaining, dim3(1), dim3(1), 0, s, sd.remaining.as<int32_t>(), (int32_t)sd.G);
    hipLaunchKernelGGL(k_expand_init, dim3(sd.ls.total[P_ONE]), dim3(256), 0, s, sd.jobs.as<P1Job>(),
                       sd.ls.bs(P_ONE), sd.G);
    DMOLP_HIP_CHECK(hipGetLastError());
}

namespace {

void p1_expand_enqueue(dmolp_ctx *ctx, P1SetDev &sd, int rounds) {
    hipStream_t s = ctx->stream;
    const LaunchSet &ls = sd.ls;
    const int G = sd.G;
    P1Job *d_jobs = sd.jobs.as<P1Job>();
    const SelArgs *d_sel0 = sd.sel.as<SelArgs>();
    const SelArgs *d_sel1 = d_sel0 + G;
    for (int r = 0; r < rounds; ++r) {
        const int par = r & 1;
        const SelArgs *sa = par ? d_sel1 : d_sel0;
        hipLaunchKernelGGL(k_p1_round_begin, dim3(ls.total[P_ONE]), dim3(1), 0, s, d_jobs, ls.bs(P_ONE), G, par);
        hipLaunchKernelGGL(k_p1_fill, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, d_jobs, ls.bs(P_GRID), G, par);
        select_enqueue(ctx, sa, ls, sd.passes);
        hipLaunchKernelGGL(k_expand_light, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, d_jobs, ls.bs(P_GRID), G);
        if (sd.any_heavy)
            hipLaunchKernelGGL(k_expand_heavy, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, d_jobs, ls.bs(P_GRID), G);
        hipLaunchKernelGGL(k_p1_apply, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, d_jobs, ls.bs(P_GRID), G, par);
        hipLaunchKernelGGL(k_p1_round_end, dim3(ls.total[P_ONE]), dim3(1), 0, s, d_jobs, ls.bs(P_ONE), G);
    }
    DMOLP_HIP_CHECK(hipGetLastError());
}

// Everything the captured launches bake in: pointers, grid sizes, pass
// count, job count and the number of rounds.
uint64_t p1_rounds_key(const P1SetDev &sd, int rounds) {
    uint64_t h = 0x6a09e667f3bcc909ull;
    const auto mix = [&h](uint64_t v) { h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); };
    mix((uint64_t)(uintptr_t)sd.jobs.p);
    mix((uint64_t)(uintptr_t)sd.sel.p);
    mix((uint64_t)(uintptr_t)sd.ls.bstart.p);
    mix((uint64_t)sd.G);
    mix((uint64_t)sd.ls.jobs);
    mix((uint64_t)sd.passes);
    mix(sd.any_heavy ? 1u : 0u);
    mix((uint64_t)rounds);
    for (int p = 0; p < P_NPOL; ++p) mix((uint64_t)sd.ls.total[p]);
    return h;
}

}  // namespace

// `rounds` expansion rounds of the set.  Each round is about twenty short
// launches, so a batch launched kernel by kernel spends as long between
// kernels as in them; a launch shape seen before runs as one captured graph
// (kernels only: no memset or copy is recorded).
void p1_expand_rounds(dmolp_ctx *ctx, P1SetDev &sd, int rounds) {
    const uint64_t key = p1_rounds_key(sd, rounds);
    RoundGraph &rg = sd.rounds;
    if (ctx->profiling || !graphs_enabled() || (!(rg.exec && rg.key == key) && sd.seen_key != key)) {
        sd.seen_key = key;
        p1_expand_enqueue(ctx, sd, rounds);
        return;
    }
    if (!rg.exec || rg.key != key) {
        if (rg.exec || rg.graph) {
            // Launches of the old executable may still be queued.
            DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            rg.release();
        }
        hipGraph_t graph = nullptr;
        DMOLP_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            p1_expand_enqueue(ctx, sd, rounds);
        } catch (...) {
            (void)hipStreamEndCapture(ctx->stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        DMOLP_HIP_CHECK(hipStreamEndCapture(ctx->stream, &graph));
        rg.graph = graph;
        DMOLP_HIP_CHECK(hipGraphInstantiate(&rg.exec, rg.graph, nullptr, nullptr, 0));
        rg.key = key;
    }
    DMOLP_HIP_CHECK(hipGraphLaunch(rg.exec, ctx->stream));
}

// Up to max_rounds (even) rounds of every job of the set that k_p1_rounds_wg
// can run, one block per job.  ran: the most rounds a job ran; stopped: jobs
// that stopped at a round the separate launches must run.
namespace {
void p1_rounds_wg_launch(P1SetDev &sd, int max_rounds, hipStream_t s) {
    sd.run_out.grow(sizeof(int32_t) * (2 + (size_t)sd.G));
    DMOLP_HIP_CHECK(hipMemsetAsync(sd.run_out.p, 0, sizeof(int32_t) * (2 + (size_t)sd.G), s));
    hipLaunchKernelGGL(k_p1_rounds_wg, dim3(sd.G), dim3(kWgThreads), 0, s, sd.jobs.as<P1Job>(),
                       max_rounds, sd.run_out.as<int32_t>());
    DMOLP_HIP_CHECK(hipGetLastError());
}
// ran, stopped and (states) every job's state after the launch on s.
void p1_rounds_wg_results(P1SetDev &sd, hipStream_t s, int &ran, int &stopped,
                          std::vector<int32_t> *states) {
    std::vector<int32_t> out(2 + (size_t)sd.G, 0);
    DMOLP_HIP_CHECK(hipMemcpyAsync(out.data(), sd.run_out.p, sizeof(int32_t) * out.size(),
                                   hipMemcpyDeviceToHost, s));
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    ran = out[0];
    stopped = out[1];
    if (states) states->assign(out.begin() + 2, out.end());
}
}  // namespace

void p1_rounds_wg(dmolp_ctx *ctx, P1SetDev &sd, int max_rounds, int &ran, int &stopped) {
    ran = 0;
    stopped = 0;
    if (sd.G < 1) return;
    p1_rounds_wg_launch(sd, max_rounds, ctx->stream);
    p1_rounds_wg_results(sd, ctx->stream, ran, stopped, nullptr);
}

int32_t p1_remaining(dmolp_ctx *ctx, P1SetDev &sd) {
    int32_t left = 0;
    DMOLP_HIP_CHECK(hipMemcpyAsync(&left, sd.remaining.p, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return left;
}

void p1_done_flags(dmolp_ctx *ctx, P1SetDev &sd, std::vector<int32_t> &out) {
    out.assign((size_t)sd.G, 0);
    hipLaunchKernelGGL(k_p1_gather_done, dim3((sd.G + 255) / 256), dim3(256), 0, ctx->stream,
                       sd.jobs.as<P1Job>(), sd.G, sd.done.as<int32_t>());
    DMOLP_HIP_CHECK(hipMemcpyAsync(out.data(), sd.done.p, sizeof(int32_t) * (size_t)sd.G,
                                   hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void p1_finish(dmolp_ctx *ctx, P1Stage &h, dmolp_phase1_metrics &m, double device_ms, const P1Ctl *ctl) {
    P1Ctl hc{};
    if (ctl) {
        hc = *ctl;
    } else {
        DMOLP_HIP_CHECK(hipMemcpyAsync(&hc, h.w->ctl.p, sizeof(P1Ctl), hipMemcpyDeviceToHost, ctx->stream));
        DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    std::memset(&m, 0, sizeof(m));
    m.first_seed = h.fs.first;
    m.first_seed_degree = h.fs.second;
    m.n_hubs = h.n_hubs;
    m.n_seeds_found = h.found;
    m.n_seeds_fallback = h.fallback;
    m.rounds = hc.rounds;
    m.fill_round = hc.fill_round;
    m.device_ms = device_ms;
    m.host_ms = h.host_ms;
    h.g->labels_valid = true;
}

void phase1_run(dmolp_ctx *ctx, dmolp_graph *g, int k, int theta, std::vector<int32_t> &seeds_out,
                dmolp_phase1_metrics *m) {
    std::vector<P1Item> items(1);
    items[0].g = g;
    items[0].k = k;
    items[0].theta = theta;
    phase1_run_set(ctx, items, {});
    seeds_out = items[0].seeds;
    if (m) *m = items[0].m;
}

void phase1_run_set(dmolp_ctx *ctx, std::vector<P1Item> &items, const std::function<void()> &during_hubs) {
    const int G = (int)items.size();
    DMOLP_REQUIRE(G >= 1, DMOLP_E_ARG, "empty job set");
    hipStream_t s = ctx->stream;
    std::vector<P1Stage> H((size_t)G);
    std::vector<P1Stage *> HP((size_t)G);
    for (int j = 0; j < G; ++j) {
        p1_stage_init(ctx, items[(size_t)j], H[(size_t)j]);
        HP[(size_t)j] = &H[(size_t)j];
    }
    P1SetDev &sd = ctx->p1_set;
    p1_set_build(ctx, HP, sd);
    DMOLP_HIP_CHECK(hipEventRecord(ctx->ev0, s));
    p1_degrees(ctx, HP, sd);
    // The device searches run on this thread while host threads search the
    // short graphs; then the caller's work.
    host_hub_search(H, [&] {
        p1_hub_search_device(ctx, HP);
        if (during_hubs) during_hubs();
    });
    p1_seeds(ctx, HP, sd);
    p1_expand_begin(ctx, HP, sd);
    int64_t enq = 0;
    const int batch = 16;        // even: every batch starts on the same frontier parity
    const int run_rounds = 4096;  // rounds of one k_p1_rounds_wg launch (even, for the same reason)
    std::vector<int32_t> done;
    // The rounds of the jobs of `act` (the jobs of sd) on the main stream,
    // until every one has finished.
    const auto run_main = [&](std::vector<P1Stage *> act) {
        int skip = 0, backoff = 1;
        while (!act.empty()) {
            // Rounds that need no selection run in one workgroup per job; a
            // round that does (and those after it, for a batch) runs as
            // separate launches.  After an attempt in which no job ran a
            // round, the next ones wait for 1, 2, 4, up to 16 batches.
            int ran = 0, stopped = 1;
            if (skip > 0) {
                --skip;
            } else {
                p1_rounds_wg(ctx, sd, run_rounds, ran, stopped);
                if (ran == 0) {
                    skip = backoff;
                    backoff = std::min(2 * backoff, 16);
                } else {
                    backoff = 1;
                }
            }
            enq += ran;
            int32_t left = p1_remaining(ctx, sd);
            if (left <= 0) break;
            if (stopped > 0) {
                p1_expand_rounds(ctx, sd, batch);
                enq += batch;
                left = p1_remaining(ctx, sd);
                if (left <= 0) break;
            }
            DMOLP_REQUIRE(enq < 64ll * (sd.max_n + 16), DMOLP_E_HIP, "phase-1 expansion did not finish");
            // A finished job keeps its blocks in every launch (they return at
            // once) and its selection pass count: once half of the set has
            // finished, the set is rebuilt from the jobs still expanding.
            if (act.size() > 1 && (size_t)left * 2 <= act.size()) {
                p1_done_flags(ctx, sd, done);
                std::vector<P1Stage *> keep;
                for (size_t j = 0; j < act.size(); ++j)
                    if (!done[j]) keep.push_back(act[j]);
                DMOLP_REQUIRE((int32_t)keep.size() == left, DMOLP_E_INTERNAL,
                              "phase-1 job count disagrees with the done flags");
                act.swap(keep);
                p1_set_build(ctx, act, sd);
                hipLaunchKernelGGL(k_p1_set_remaining, dim3(1), dim3(1), 0, s, sd.remaining.as<int32_t>(),
                                   (int32_t)act.size());
                DMOLP_HIP_CHECK(hipGetLastError());
            }
        }
    };
    // A short first attempt in one workgroup per job sorts the jobs: those
    // still running rounds that need no selection after it (small frontiers:
    // grids, long paths) go on in one long one-workgroup launch on a second
    // stream, while the others advance through the separate launches on
    // the main stream; the two touch disjoint jobs.  Whatever the long
    // launch leaves unfinished joins the main stream afterwards.
    std::vector<P1Stage *> act = HP, later;
    if (G > 1) {
        int ran = 0, stopped = 0;
        std::vector<int32_t> st;
        p1_rounds_wg_launch(sd, kWgProbeRounds, s);
        p1_rounds_wg_results(sd, s, ran, stopped, &st);
        enq += ran;
        std::vector<P1Stage *> mainj, sidej;
        for (int j = 0; j < G; ++j) {
            if (st[(size_t)j] == 0) sidej.push_back(HP[(size_t)j]);
            else if (st[(size_t)j] == 1) mainj.push_back(HP[(size_t)j]);
        }
        if (!sidej.empty()) {
            ctx->ensure_side();
            P1SetDev &sw = ctx->p1_side;
            p1_set_build(ctx, sidej, sw);
            hipLaunchKernelGGL(k_p1_set_remaining, dim3(1), dim3(1), 0, s, sw.remaining.as<int32_t>(),
                               (int32_t)sidej.size());
            if (!mainj.empty()) {
                p1_set_build(ctx, mainj, sd);
                hipLaunchKernelGGL(k_p1_set_remaining, dim3(1), dim3(1), 0, s, sd.remaining.as<int32_t>(),
                                   (int32_t)mainj.size());
            }
            DMOLP_HIP_CHECK(hipGetLastError());
            // The side stream starts after everything queued so far (the
            // set uploads, and the work of whatever memory it reuses).
            DMOLP_HIP_CHECK(hipEventRecord(ctx->ev_fork, s));
            DMOLP_HIP_CHECK(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
            p1_rounds_wg_launch(sw, run_rounds, ctx->side);
            act = mainj;
            later = sidej;
        }
    }
    try {
        run_main(act);
    } catch (...) {
        if (!later.empty()) (void)hipStreamSynchronize(ctx->side);  // no work left on its jobs
        throw;
    }
    if (!later.empty()) {
        int ran = 0, stopped = 0;
        std::vector<int32_t> st;
        p1_rounds_wg_results(ctx->p1_side, ctx->side, ran, stopped, &st);  // waits for the side stream
        enq += ran;
        std::vector<P1Stage *> rest;
        for (size_t j = 0; j < later.size(); ++j)
            if (st[j] != 2) rest.push_back(later[j]);
        if (!rest.empty()) {
            p1_set_build(ctx, rest, sd);
            hipLaunchKernelGGL(k_p1_set_remaining, dim3(1), dim3(1), 0, s, sd.remaining.as<int32_t>(),
                               (int32_t)rest.size());
            DMOLP_HIP_CHECK(hipGetLastError());
            run_main(rest);
        }
    }
    DMOLP_HIP_CHECK(hipEventRecord(ctx->ev1, s));
    DMOLP_HIP_CHECK(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    DMOLP_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    std::vector<P1Ctl> hc((size_t)G);
    OpBatch ob;
    for (int j = 0; j < G; ++j) ob.download(&hc[(size_t)j], H[(size_t)j].w->ctl.p, sizeof(P1Ctl));
    ob.run(ctx);
    for (int j = 0; j < G; ++j) {
        p1_finish(ctx, H[(size_t)j], items[(size_t)j].m, ms, &hc[(size_t)j]);
        items[(size_t)j].seeds = H[(size_t)j].seeds;
    }
}

}  // namespace dmolp
