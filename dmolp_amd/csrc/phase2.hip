// phase2.hip -- boundary label propagation (phase2.cpp:256-449 with
// expandBoundaryNodes gpu_lp_boundary.cu:185-305 and the decision of
// boundaryLPKernel_unified gpu_lp_boundary.cu:327-427), device-resident.
//
// Per round t the host enqueues, without waiting:
//   round_begin   penalties (phase2.cpp:158-209) from the device counts
//   candidates    B_{t-1} u N(B_{t-1}); a candidate is a record iff it has a
//                 neighbour with another label (round 1: the initial sweep's
//                 differing-neighbour counts; later rounds: a row scan)
//   select        per label top ceil(ratio*c) by (degree, id)  -> B_t
//   lp            decision for every b in B_t (Jacobi)        -> Delta
//   delta         incremental vc, ic and cut
//   apply         commit Delta to the labels
//   round_end     exit tests (phase2.cpp:309-312, 384-415), trace
// Every kernel returns at once when the control block says "done", so a
// finished run costs empty launches only.
#include <cmath>
#include <array>
#include <cstring>
#include <vector>

#include "lp_rule.hpp"
#include "phase2.hpp"
#include "rows.hpp"

namespace dmolp {
namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr int kHistStride = kLpCap + 1;  // odd stride: no LDS bank conflicts across slots
// Per-block label counters (vc, ic) live in LDS for labels below kLdsK;
// larger labels (k above kLdsK) add to the global counters directly.  Every
// k the byte mirror covers fits: a global counter takes the adds of rows
// that share its label one after another (consecutive rows mostly carry
// one label after phase 1), at the memory-side atomic rate.
constexpr int kLdsK = 256;
constexpr int kLpUnroll = 4;     // light rows: gathers in flight per lane
constexpr int kHeavyUnroll = 8;  // heavy rows: gathers in flight per lane
constexpr int kLightUnroll = 4;  // light rows in the initial sweep: gathers in flight per lane
constexpr int32_t kLabel8None = 255;  // byte-mirror value of labels the LP cannot see
constexpr int kScanItems = 4;         // vertices per thread per block iteration in scans
// Up to this k every valid label fits the byte mirror exactly (255 stands
// for -1 and out-of-range labels), so the sweeps gather bytes, not words.
constexpr int kExact8 = 254;
// Light rows count in 16-bit halves of 32-bit words (a light row has at most
// 65535 entries): label l in word l / 2, half l % 2.  LDS rows of a wave are
// kPackedStride words apart; kept rows in HBM are kPackedWords (64 bytes).
constexpr int kPackedWords = kLpCap / 2;
constexpr int kPackedStride = kPackedWords + 1;  // odd: rows of a wave fall in different banks


__device__ __forceinline__ bool is_done(const Ctl *c) { return c->done != 0; }
// T is the LDS counter type: 64-bit, or 32-bit where a block's count cannot
// reach 2^32 (one add per vertex the block handles).
template <typename T>
__device__ __forceinline__ void label_add(T *lds, unsigned long long *glob, int32_t l, T x) {
    if (l < kLdsK) atomicAdd(&lds[l], x);
    else atomicAdd(&glob[l], (unsigned long long)x);
}
template <typename T>
__device__ __forceinline__ void label_flush(const T *lds, unsigned long long *glob, int k) {
    const int kk = k < kLdsK ? k : kLdsK;
    for (int i = threadIdx.x; i < kk; i += blockDim.x)
        if (lds[i]) atomicAdd(&glob[i], (unsigned long long)lds[i]);
}
__device__ __forceinline__ bool bit_of(const uint32_t *bits, int32_t v) {
    return (bits[v >> 5] >> (v & 31)) & 1u;
}

// Neighbour label for the sweeps: from the byte mirror when k <= kExact8
// (255 maps back to -1), else the int32 label.
__device__ __forceinline__ int32_t nb_label(bool exact8, const uint8_t *L8, const int32_t *L,
                                            int32_t u) {
    if (exact8) {
        const int32_t b = L8[u];
        return b == kLabel8None ? -1 : b;
    }
    return L[u];
}

// Add each thread's x to *dst with one atomic per block (block-uniform
// call, after the kernel's loops).  Atomics on one cache line serialise
// (about 12 ns each on MI355X): one per block, not one per wave or per loop
// iteration.
__device__ __forceinline__ void block_count_add(unsigned long long *dst, int64_t x) {
    __shared__ unsigned long long s_acc;
    if (threadIdx.x == 0) s_acc = 0;
    __syncthreads();
    x = wave_sum(x);
    if (lane_id() == 0 && x) atomicAdd(&s_acc, (unsigned long long)x);
    __syncthreads();
    if (threadIdx.x == 0 && s_acc) atomicAdd(dst, s_acc);
}

// Word q of lane `lane`'s packed LDS counter row.
__device__ __forceinline__ uint32_t r_word(const uint32_t *h, int lane, int q) {
    return h[lane * kPackedStride + q];
}

// LN entry of a vertex deciding to move from own to nl (byte mirrors).
__device__ __forceinline__ uint16_t ln_pair(int32_t own, int32_t nl) {
    const uint32_t o = (own >= 0 && own < kLabel8None) ? (uint32_t)own : (uint32_t)kLabel8None;
    const uint32_t nw = (nl >= 0 && nl < kLabel8None) ? (uint32_t)nl : (uint32_t)kLabel8None;
    return (uint16_t)(o | (nw << 8));
}

// Every round kernel takes the job set and runs on its job's share of the
// launch (see VGrid): one graph, or many graphs in one launch.
using ::dmolp::gload;
__device__ __forceinline__ P2Job gload(const P2Job &v) {
    P2Job r = v;
    r.g = gload(v.g);
    r.all = gload(v.all);
    r.B = gload(v.B);
    r.D = gload(v.D);
    r.L = gload(v.L);
    r.L8 = gload(v.L8);
    r.L5 = gload(v.L5);
    r.chg = gload(v.chg);
    r.diff = gload(v.diff);
    r.stamp = gload(v.stamp);
    r.deferred = gload(v.deferred);
    r.hub_hist = gload(v.hub_hist);
    r.rdeg = gload(v.rdeg);
    r.bbits = gload(v.bbits);
    r.hh = gload(v.hh);
    r.hhc = gload(v.hhc);
    r.hid = gload(v.hid);
    r.cbits = gload(v.cbits);
    r.ON = gload(v.ON);
    r.LN = gload(v.LN);
    r.lh = gload(v.lh);
    r.hbits = gload(v.hbits);
    r.first = gload(v.first);
    r.deg = gload(v.deg);
    r.wit = gload(v.wit);
    r.bd = gload(v.bd);
    r.ctl = gload(v.ctl);
    r.vc = gload(v.vc);
    r.ic = gload(v.ic);
    r.pen = gload(v.pen);
    r.re = gload(v.re);
    r.rtab = gload(v.rtab);
    r.trace = gload(v.trace);
    r.remaining = gload(v.remaining);
    return r;
}

#define P2_JOB()                                   \
    const int job_ = job_of_block(bstart, jobs);   \
    const VGrid vg = vgrid_of(bstart, job_);       \
    const P2Job J = gload(JS[job_]);               \
    (void)vg


// ---------------------------------------------------------------- init sweep

// Label of vertex u for the initial sweep's gathers: the byte mirror (B = 8)
// or the packed mirror, B bits a label, 32 / B labels a word.  A smaller
// array keeps more of the gathered labels in L2: the sweep's gathers are
// bound by the lines their L2 misses move.
template <int B>
__device__ __forceinline__ int32_t sweep_label(const uint8_t *L8, const uint32_t *L5, uint32_t u) {
    if constexpr (B == 8) {
        return L8[u];
    } else {
        constexpr uint32_t per = 32 / B;
        return (int32_t)((L5[u / per] >> ((u % per) * B)) & ((1u << B) - 1u));
    }
}

// Packed label mirror (lbits bits a label) and the count of vertices whose
// label lies outside [0, k).  The sweep reads the packed mirror only when
// that count is zero: then no entry has an invalid label and the packed
// value is the label itself.
__global__ void k_pack_labels(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    if (J.lbits == 0) return;  // uniform per job
    const int64_t n = J.g.n;
    const int B = J.lbits, per = 32 / B;
    const int64_t nw = (n + per - 1) / per;
    int64_t bad = 0;
    for (int64_t i = vthread(vg); i < nw; i += vthreads(vg)) {
        uint32_t word = 0;
        for (int j = 0; j < per; ++j) {
            const int64_t v = i * per + j;
            if (v >= n) break;
            const int32_t l = J.L[v];
            bad += (l < 0 || l >= J.k) ? 1 : 0;
            word |= ((uint32_t)l & ((1u << B) - 1u)) << (j * B);
        }
        J.L5[i] = word;
    }
    block_count_add(&J.ctl->v[C_BADLAB], bad);
}

// phase2.cpp:55-117 and 214-253 over all vertices, plus the per-vertex count
// of differing neighbours used as the boundary test, and (lh_on) each light
// row's label counts for the kept histograms.
__global__ __launch_bounds__(kThreads, 6) void k_init_light(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const ListDev all = J.all;
    const int32_t *L = J.L;
    const uint8_t *L8 = J.L8;
    const int k = J.k;
    const bool exact8 = k <= kExact8;
    int32_t *diff = J.diff;
    Ctl *ctl = J.ctl;
    unsigned long long *vc = J.vc;
    unsigned long long *ic = J.ic;
    const bool lh_on = J.lh_on != 0;
    uint32_t *lh = J.lh;
    const int E = k < kLpCap ? k : kLpCap;
    const int nvec = (E + 7) / 8;  // 16-byte vectors of a row that hold labels below E
    // Whenever labels fit the byte mirror (k <= kExact8) a row's differing,
    // intra and invalid counts follow from per-row counts (count_rows): the
    // kept light row takes the entries whose label is below E, and s_cnt
    // counts, in its low half, the entries whose label is -1 and, in its
    // high half, those carrying the row's own label when that label has no
    // place in the kept row (k > 32, or no kept rows).
    const bool from_hist = exact8;
    // Packed gathers when every label is valid (see k_pack_labels).
    const int packed = (J.lbits != 0 && ctl->v[C_BADLAB] == 0) ? J.lbits : 0;
    __shared__ BatchLds s_b[kWavesPerBlock];
    __shared__ int32_t s_cnt[kWavesPerBlock][kWave];
    __shared__ uint32_t s_h[kWavesPerBlock][kWave * kPackedStride];
    // One vertex count a row: 32-bit counters keep the block within the LDS
    // budget of six blocks a CU.
    __shared__ uint32_t s_vc[kLdsK];
    __shared__ unsigned long long s_ic[kLdsK];
    for (int i = threadIdx.x; i < kLdsK; i += blockDim.x) { s_vc[i] = 0; s_ic[i] = 0; }
    __syncthreads();
    const int w = threadIdx.x / kWave, lane = lane_id();
    uint32_t *h = s_h[w];
    int64_t invalid = 0, cut2 = 0;
    const int64_t count = (int64_t)all.cnt[LC_LIGHT];
    for (int64_t b0 = vwave(vg) * kWave; b0 < count; b0 += vwaves(vg) * kWave) {
        int32_t v, deg;
        bool valid;
        const int32_t total = batch_load(g, all.light, count, b0, s_b[w], v, deg, valid);
        s_cnt[w][lane] = 0;
        s_b[w].aux[lane] = valid ? (int32_t)L8[v] : -2;  // the row's own label byte
        if (lh_on)
            for (int l = 0; l < 4 * nvec; ++l) h[lane * kPackedStride + l] = 0u;
        wave_sync();
        // Counting only: kLightUnroll entries per lane per step, their slot
        // searches, column loads and label gathers issued together.
        const auto count_rows = [&](auto bits) {
            constexpr int B = decltype(bits)::value;
            for (int32_t q0 = 0; q0 < total; q0 += kWave * kLightUnroll) {
                int slot[kLightUnroll];
                int32_t u[kLightUnroll];
#pragma unroll
                for (int j = 0; j < kLightUnroll; ++j) {
                    const int32_t q = q0 + j * kWave + lane;
                    slot[j] = q < total ? batch_slot(s_b[w], q) : -1;
                }
                // Loads from in-bounds indices, issued unconditionally (a
                // lane-conditional load is a branch and a wait of its own).
#pragma unroll
                for (int j = 0; j < kLightUnroll; ++j) {
                    const int32_t q = q0 + j * kWave + lane;
                    const int64_t e = s_b[w].base[slot[j] >= 0 ? slot[j] : 0] + q;
                    const int32_t cu = g.col[slot[j] >= 0 ? e : 0];
                    u[j] = slot[j] >= 0 ? cu : -1;
                }
                int32_t b[kLightUnroll];
#pragma unroll
                for (int j = 0; j < kLightUnroll; ++j) {
                    const int32_t lb = sweep_label<B>(L8, J.L5, (uint32_t)(u[j] >= 0 ? u[j] : 0));
                    b[j] = u[j] >= 0 ? lb : -1;
                }
#pragma unroll
                for (int j = 0; j < kLightUnroll; ++j) {
                    if (b[j] < 0) continue;
                    if (lh_on && b[j] < E)
                        atomicAdd(&h[slot[j] * kPackedStride + (b[j] >> 1)], 1u << ((b[j] & 1) * 16));
                    else if (b[j] == kLabel8None)
                        atomicAdd(&s_cnt[w][slot[j]], 1);
                    else if (b[j] == s_b[w].aux[slot[j]])
                        atomicAdd(&s_cnt[w][slot[j]], 0x10000);
                }
            }
        };
        if (from_hist) {
            if (packed == 5) count_rows(std::integral_constant<int, 5>{});
            else if (packed == 4) count_rows(std::integral_constant<int, 4>{});
            else count_rows(std::integral_constant<int, 8>{});
        }
        for (int32_t q0 = 0; q0 < (from_hist ? 0 : total); q0 += kWave) {
            const int32_t q = q0 + lane;
            if (q < total) {
                const int slot = batch_slot(s_b[w], q);
                const int64_t e = s_b[w].base[slot] + q;
                const int32_t vv = s_b[w].v[slot];
                const int32_t u = g.col[e];
                const int32_t b = L8[u];
                const int32_t lv = L[vv], lu = exact8 ? (b == kLabel8None ? -1 : b) : L[u];
                if (lh_on && b < E) atomicAdd(&h[slot * kPackedStride + (b >> 1)], 1u << ((b & 1) * 16));
                if (lv == -1 || lu == -1) {
                    ++invalid;
                } else if (lv != lu) {
                    ++cut2;
                    atomicAdd(&s_cnt[w][slot], 1);
                } else if (lv >= 0 && lv < k) {
                    label_add(s_ic, ic, lv, 1ull);
                }
            }
        }
        wave_sync();
        if (valid) {
            const int32_t lv = L[v];
            if (lv >= 0 && lv < k) label_add(s_vc, vc, lv, 1u);
            else if (lv != -1) atomicAdd(&ctl->v[C_KERNEL_ERR], 1ull);
            if (lh_on) {
                const uint32_t *r = &h[lane * kPackedStride];
                uint4 *dst = reinterpret_cast<uint4 *>(lh + (int64_t)v * kPackedWords);
                for (int q = 0; q < nvec; ++q)
                    dst[q] = make_uint4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
            }
            if (from_hist) {
                // The row's counts: `same` entries share v's label (from the
                // kept row when the label has a place there, else from
                // s_cnt's high half), `inv` have label -1 (byte 255), the
                // rest differ.  Light rows are shorter than 2^16 entries, so
                // the 16-bit halves do not overflow.
                const uint32_t oc = (uint32_t)s_cnt[w][lane];
                const int32_t inv = (int32_t)(oc & 0xffffu);
                if (lv == -1) {
                    invalid += deg;
                    diff[v] = 0;
                } else {
                    const int32_t same =
                        (lh_on && lv >= 0 && lv < E)
                            ? (int32_t)((r_word(h, lane, lv >> 1) >> ((lv & 1) * 16)) & 0xffffu)
                            : (int32_t)(oc >> 16);
                    const int32_t dif = deg - inv - same;
                    invalid += inv;
                    cut2 += dif;
                    diff[v] = dif;
                    if (same && lv >= 0 && lv < k) label_add(s_ic, ic, lv, (unsigned long long)same);
                }
            } else {
                diff[v] = s_cnt[w][lane];
            }
        }
        wave_sync();
    }
    block_count_add(&ctl->v[C_INVALID], invalid);
    block_count_add(&ctl->v[C_CUT2], cut2);
    __syncthreads();
    label_flush(s_vc, vc, k);
    label_flush(s_ic, ic, k);
}

__global__ __launch_bounds__(kThreads, 6) void k_init_heavy(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const ListDev all = J.all;
    const int32_t *L = J.L;
    const int k = J.k;
    const int chunk_edges = J.chunk_edges;
    int32_t *diff = J.diff;
    Ctl *ctl = J.ctl;
    unsigned long long *vc = J.vc;
    unsigned long long *ic = J.ic;
    const uint8_t *L8 = J.L8;
    int32_t *hh = J.hh;
    const bool exact8 = k <= kExact8;
    const int E = k < kLpCap ? k : kLpCap;
    // Per-lane histogram rows of the row's neighbour labels (below E), the
    // heavy row's starting histogram for the LP rounds: 16-bit counts, two
    // labels a word (a lane counts at most chunk_edges / 64 < 2^16 entries
    // of a chunk), which halves the LDS and so raises the occupancy.
    __shared__ uint32_t s_h[kWavesPerBlock][kWave * kPackedStride];
    const int w = threadIdx.x / kWave, lane = lane_id();
    uint32_t *row = &s_h[w][lane * kPackedStride];
    const int nw = (E + 1) / 2;  // words of a row in use
    // Packed gathers when every label is valid (see k_pack_labels).
    const int packed = (J.lbits != 0 && ctl->v[C_BADLAB] == 0) ? J.lbits : 0;
    // Column lane of the chunk's label-lane count (lane < E).
    const auto col_count = [&](int l) {
        int32_t c = 0;
        for (int r = 0; r < kWave; ++r) c += (int32_t)((s_h[w][r * kPackedStride + (l >> 1)] >> ((l & 1) * 16)) & 0xffffu);
        return c;
    };
    int64_t invalid = 0, cut2 = 0;
    const int64_t nch = (int64_t)all.cnt[LC_CHUNKS];
    for (int64_t j = vwave(vg); j < nch; j += vwaves(vg)) {
        int32_t slot, v;
        int64_t e0, e1, deg;
        const int64_t item = all.hc[j];
        chunk_bounds(g, all, item, chunk_edges, slot, v, e0, e1, deg);
        const int32_t lv = L[v];
        int64_t dcount = 0, intra = 0;
        for (int l = 0; l < nw; ++l) row[l] = 0u;
        if (k <= kLpCap) {
            // Every label is counted (E = k): the chunk's same, invalid and
            // differing counts follow from its histogram, so the entry loop
            // only counts, with kHeavyUnroll gathers in flight per lane.
            // Labels in [k, 255) count as differing, as the loop below
            // would; their vertices raise the kernel error anyway.
            int32_t inv = 0;
            const auto count_chunk = [&](auto bits) {
            constexpr int B = decltype(bits)::value;
            for (int64_t eb = e0; eb < e1; eb += kWave * kHeavyUnroll) {
                // In-bounds indices: the loads issue unconditionally.
                int32_t u[kHeavyUnroll];
#pragma unroll
                for (int q = 0; q < kHeavyUnroll; ++q) {
                    const int64_t e = eb + q * kWave + lane;
                    const int32_t cu = __builtin_nontemporal_load(&g.col[e < e1 ? e : e1 - 1]);
                    u[q] = e < e1 ? cu : -1;
                }
                int32_t lab[kHeavyUnroll];
#pragma unroll
                for (int q = 0; q < kHeavyUnroll; ++q) {
                    const int32_t lb = sweep_label<B>(L8, J.L5, (uint32_t)(u[q] >= 0 ? u[q] : 0));
                    lab[q] = u[q] >= 0 ? lb : -1;
                }
#pragma unroll
                for (int q = 0; q < kHeavyUnroll; ++q) {
                    if (lab[q] >= 0 && lab[q] < E) row[lab[q] >> 1] += 1u << ((lab[q] & 1) * 16);
                    inv += lab[q] == kLabel8None ? 1 : 0;
                }
            }
            };
            if (packed == 5) count_chunk(std::integral_constant<int, 5>{});
            else if (packed == 4) count_chunk(std::integral_constant<int, 4>{});
            else count_chunk(std::integral_constant<int, 8>{});
            inv = (int32_t)wave_sum((int64_t)inv);
            wave_sync();
            int32_t c = 0;
            if (lane < E) {
                c = col_count(lane);
                if (c) atomicAdd(&hh[(int64_t)slot * kLpCap + lane], c);
            }
            const int32_t same = __shfl(c, lv >= 0 && lv < E ? lv : 0, kWave);
            wave_sync();
            if (lane == 0) {
                const int64_t len = e1 - e0;
                if (lv == -1) {
                    invalid += len;
                } else {
                    const int32_t s = lv >= 0 && lv < E ? same : 0;
                    const int64_t d = len - inv - s;
                    invalid += inv;
                    cut2 += d;
                    if (d) atomicAdd(&diff[v], (int32_t)d);
                    if (s) atomicAdd(&ic[lv], (unsigned long long)s);
                }
                if ((item & 0xffffffffll) == 0) {
                    if (lv >= 0 && lv < k) atomicAdd(&vc[lv], 1ull);
                    else if (lv != -1) atomicAdd(&ctl->v[C_KERNEL_ERR], 1ull);
                }
            }
            continue;
        }
        for (int64_t e = e0 + lane; e < e1; e += kWave) {
            const int32_t u = g.col[e];
            const int32_t b = L8[u];
            if (b < E) row[b >> 1] += 1u << ((b & 1) * 16);
            const int32_t lu = exact8 ? (b == kLabel8None ? -1 : b) : L[u];
            if (lv == -1 || lu == -1) ++invalid;
            else if (lv != lu) { ++cut2; ++dcount; }
            else if (lv >= 0 && lv < k) ++intra;
        }
        dcount = wave_sum(dcount);
        intra = wave_sum(intra);
        wave_sync();
        if (lane < E) {
            const int32_t c = col_count(lane);
            if (c) atomicAdd(&hh[(int64_t)slot * kLpCap + lane], c);
        }
        wave_sync();
        if (lane == 0) {
            if (dcount) atomicAdd(&diff[v], (int32_t)dcount);
            if (intra) atomicAdd(&ic[lv], (unsigned long long)intra);
            if ((item & 0xffffffffll) == 0) {
                if (lv >= 0 && lv < k) atomicAdd(&vc[lv], 1ull);
                else if (lv != -1) atomicAdd(&ctl->v[C_KERNEL_ERR], 1ull);
            }
        }
    }
    block_count_add(&ctl->v[C_INVALID], invalid);
    block_count_add(&ctl->v[C_CUT2], cut2);
}

__global__ void k_init_finish(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    Ctl *ctl = J.ctl;
    unsigned long long *ic = J.ic;
    const int k = J.k;
    const int32_t epoch_base = J.epoch_base;
    const int32_t max_rounds = J.max_rounds;
    const double epsilon = J.epsilon;
    for (int i = threadIdx.x; i < k; i += blockDim.x) ic[i] = ic[i] / 2;  // phase2.cpp:110
    if (threadIdx.x == 0) {
        ctl->t = 1;
        ctl->epoch_base = epoch_base;
        ctl->epoch = epoch_base;
        ctl->max_rounds = max_rounds;
        ctl->epsilon = epsilon;
        const unsigned long long cut2 = ctl->v[C_CUT2];
        if (ctl->v[C_KERNEL_ERR] != 0) { ctl->err = DMOLP_E_ARG; ctl->done = 1; }
        else if (ctl->v[C_INVALID] != 0) { ctl->err = DMOLP_E_INVALID_LABEL; ctl->done = 1; }
        else if (cut2 & 1ull) { ctl->err = DMOLP_E_ODD_CUT; ctl->done = 1; }
        ctl->v[C_CUT] = cut2 >> 1;
        ctl->v[C_PREV_CUT] = cut2 >> 1;
        ctl->v[C_INIT_CUT] = cut2 >> 1;
        // Every label is valid from here on, so the sweep's differing counts
        // are exact; Delta keeps them exact after each round.
        ctl->diff_ok = 1;
        if (ctl->done) atomicSub(J.remaining, 1);
    }
}

// ---------------------------------------------------------------- round begin

// phase2.cpp:158-209 in source order; lg = 1 + log(t) from the host's libm.
// Called by every lane of a one-wave block.  Element-wise steps run one
// label per lane; the means and variances are summed by lane 0 in label
// order, so every double matches the sequential formula bit for bit.
__device__ void penalties(const unsigned long long *vc, const unsigned long long *ic, int k,
                          double lg, double *pen, double *re) {
    __shared__ double s_rv[DMOLP_MAX_K], s_re[DMOLP_MAX_K];
    __shared__ double s_imb[2];
    const int lane = lane_id();
    int64_t tv = 0, te = 0;  // integer sums: any order
    for (int i = lane; i < k; i += kWave) { tv += (int64_t)vc[i]; te += (int64_t)ic[i]; }
    tv = wave_sum(tv);
    te = wave_sum(te);
    const double exp_v = static_cast<double>(tv) / k;
    const double exp_e = (te > 0) ? static_cast<double>(te) / k : 1.0;
    for (int i = lane; i < k; i += kWave) {
        s_rv[i] = (exp_v > 0) ? static_cast<double>((int64_t)vc[i]) / exp_v : 1.0;  // RV
        s_re[i] = (exp_e > 0) ? static_cast<double>((int64_t)ic[i]) / exp_e : 1.0;
    }
    wave_sync();
    if (lane == 0) {
        double rv_mean = 0.0, re_mean = 0.0;
        for (int i = 0; i < k; ++i) { rv_mean += s_rv[i]; re_mean += s_re[i]; }
        rv_mean /= k;
        re_mean /= k;
        double rv_var = 0.0, re_var = 0.0;
        for (int i = 0; i < k; ++i) {
            rv_var += (s_rv[i] - rv_mean) * (s_rv[i] - rv_mean);
            re_var += (s_re[i] - re_mean) * (s_re[i] - re_mean);
        }
        rv_var /= k;
        re_var /= k;
        const double total_var = rv_var + re_var;
        s_imb[0] = (total_var > 0) ? rv_var / total_var : 0.0;
        s_imb[1] = (total_var > 0) ? re_var / total_var : 0.0;
    }
    wave_sync();
    const double imb_rv = s_imb[0], imb_re = s_imb[1];
    for (int i = lane; i < k; i += kWave) {
        const double g_rv = (1.0 - s_rv[i]) * lg;
        const double g_re = (1.0 - s_re[i]) * lg;
        pen[i] = imb_rv * g_rv + imb_re * g_re;
        re[i] = s_re[i];
    }
}

__global__ void k_penalties(const unsigned long long *vc, const unsigned long long *ic, int k,
                            double lg, double *pen, double *re) {
    if (threadIdx.x < kWave) penalties(vc, ic, k, lg, pen, re);
}

// Round start: penalties, counter resets, and the candidate direction.
// Pushing from B_{t-1} reads every edge of B_{t-1} at random; pulling reads
// every vertex in order and scans its row only until it has found both a
// B_{t-1} member and a differing label.  Pull when B_{t-1} covers at least
// pull_frac of the CSR entries.
// B holds B_{t-1} here; its counts are kept for the trace before the
// selection overwrites the list with B_t.
__global__ void k_round_begin(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    Ctl *ctl = J.ctl;
    const unsigned long long *vc = J.vc;
    const unsigned long long *ic = J.ic;
    const int k = J.k;
    const double *rtab = J.rtab;
    double *pen = J.pen;
    double *re = J.re;
    const ListDev B = J.B;
    const ListDev dl = J.D;
    const RecDev bd = J.bd;
    const int64_t n = J.g.n;
    const int64_t nnz = J.g.nnz;
    const double pull_frac = J.pull_frac;
    if (is_done(ctl)) return;
    if (threadIdx.x < LC_N) dl.cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        const int t = ctl->t;
        const unsigned long long prev_deg = t == 1 ? (unsigned long long)nnz : B.cnt[LC_DEG];
        ctl->v[C_PREV_N] = t == 1 ? (unsigned long long)n : B.cnt[LC_LIGHT] + B.cnt[LC_SLOTS];
        ctl->v[C_PREV_DEG] = prev_deg;
        ctl->epoch = ctl->epoch_base + t;
        *bd.cnt = 0;
        B.cnt[LC_FILL] = 0;
        ctl->v[C_CAND] = 0;
        ctl->v[C_DEFERRED] = 0;
        ctl->v[C_PENDING] = 0;
        ctl->cand_mode = t == 1 ? CAND_ALL : ((double)prev_deg >= pull_frac * (double)nnz ? CAND_PULL : CAND_PUSH);
    }
    if (threadIdx.x < kWave) penalties(vc, ic, k, rtab[ctl->t], pen, re);
}

// ---------------------------------------------------------------- candidates


// Round 1: B_0 is every vertex, so the candidates are every vertex.
__device__ __forceinline__ void cand_all(const P2Job &J, const VGrid &vg) {
    const GraphDev g = J.g;
    const int32_t *diff = J.diff;
    int32_t *rdeg = J.rdeg;
    int64_t nrec = 0;
    const int64_t span = (int64_t)blockDim.x * kScanItems;
    for (int64_t b0 = vg.block * span; b0 < g.n; b0 += vg.blocks * span) {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const int64_t v = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            const int64_t vi = v < g.n ? v : g.n - 1;  // in bounds: the loads issue together
            const int32_t dv = diff[vi];
            const int32_t d = J.deg[vi];
            if (v >= g.n) continue;
            const bool rec = dv > 0;
            rdeg[v] = rec ? d : -1;
            nrec += rec ? 1 : 0;
        }
    }
    block_count_add(J.bd.cnt, nrec);
}

// First touch of u in this pass; a touched boundary vertex becomes a record.
// Boundary test (gpu_lp_boundary.cu:252-263): does v (label lv) have a
// neighbour with another label among entries [s, e)?  Stops at the first.
__device__ __forceinline__ bool differs_in(const GraphDev &g, int32_t lv, int64_t s, int64_t e,
                                           bool exact8, const uint8_t *L8, const int32_t *L) {
    for (int64_t x = s; x < e; ++x)
        if (nb_label(exact8, L8, L, g.col[x]) != lv) return true;
    return false;
}

// Boundary test of v (label lv, row [s, e)) in rounds after the first, once
// the initial sweep has run (ctl->diff_ok): a light row's differing count is
// kept exact by Delta; a heavy row counts its own label in its kept
// histogram when every label is counted (k <= 32); otherwise the row is
// scanned.  Heavy rows take no differing-count updates: thousands of
// changed neighbours would all add to the same word.
__device__ __forceinline__ bool on_boundary(const P2Job &J, bool diff_ok, int32_t v, int32_t lv,
                                            int64_t s, int64_t e, bool exact8) {
    if (diff_ok) {
        if (!bit_of(J.hbits, v)) return J.diff[v] > 0;
        if (J.k <= kLpCap) return (int64_t)J.hh[(int64_t)J.hid[v] * kLpCap + lv] < e - s;
    }
    return differs_in(J.g, lv, s, e, exact8, J.L8, J.L);
}

// First touch of u in this pass; a touched boundary vertex becomes a record.
__device__ __forceinline__ void touch(const P2Job &J, bool diff_ok, bool pred, int32_t u,
                                      int32_t *stamp, int32_t epoch, const int32_t *L,
                                      bool exact8, int64_t &cand, int64_t &nrec) {
    bool first = false;
    if (pred && stamp[u] != epoch) first = atomicExch(&stamp[u], epoch) != epoch;
    if (first) ++cand;
    int32_t lu = 0;
    int64_t s = 0, e = 0;
    if (first) {
        lu = L[u];
        s = J.g.off[u];
        e = J.g.off[u + 1];
    }
    const bool rec = first && on_boundary(J, diff_ok, u, lu, s, e, exact8);
    if (rec) {
        J.rdeg[u] = (int32_t)(e - s);
        ++nrec;
    }
}

// Pull: a vertex is a candidate iff it or one of its neighbours was selected
// in round t-1 (bit set in the B_{t-1} bitmap), and a record iff it is on the
// boundary.  This pass decides what the vertex itself and its first
// neighbour answer; rows still undecided are staged in LDS and appended to a
// pending list that k_cand_pull_rest scans.  Scanning them here would hold
// nearly every wave: a few undecided rows are spread over all of them.
constexpr int kPendStage = 2048;
__device__ __forceinline__ void pend_flush(int32_t *out, unsigned long long *cnt, const int32_t *s_q,
                                           int32_t &s_n, unsigned long long &s_base) {
    const int32_t c = s_n;
    if (threadIdx.x == 0) s_base = atomicAdd(cnt, (unsigned long long)c);
    __syncthreads();
    for (int i = threadIdx.x; i < c; i += blockDim.x) out[s_base + (unsigned long long)i] = s_q[i];
    __syncthreads();
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_cand_pull(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const int32_t *L = J.L;
    const RecDev bd = J.bd;
    int64_t *deferred = J.deferred;
    const int heavy_degree = J.heavy_degree;
    Ctl *ctl = J.ctl;
    if (is_done(ctl)) return;
    if (ctl->cand_mode == CAND_ALL) {  // round 1, in the same launch
        cand_all(J, vg);
        return;
    }
    if (ctl->cand_mode != CAND_PULL) return;
    const uint32_t *inb = J.bbits + (int64_t)((ctl->t - 1) & 1) * J.bwords;  // B_{t-1}
    // Once the differing counts are kept (see on_boundary), interior vertices
    // are skipped outright and boundary rows are scanned for B_{t-1} only.
    const bool use_diff = ctl->diff_ok != 0;
    const int32_t *diff = J.diff;
    int32_t *rdeg = J.rdeg;
    // Pending rows go to the Delta list, unused until the LP fills it.
    int32_t *pend = J.D.light;
    __shared__ int32_t s_q[kPendStage];
    __shared__ int64_t s_dq[kThreads * kScanItems];  // one block iteration's long rows
    __shared__ int32_t s_n, s_dn;
    __shared__ unsigned long long s_base;
    if (threadIdx.x == 0) { s_n = 0; s_dn = 0; }
    __syncthreads();
    int64_t nrec_total = 0;
    const int64_t span = (int64_t)blockDim.x * kScanItems;
    for (int64_t b0 = vg.block * span; b0 < g.n; b0 += vg.blocks * span) {
        uint32_t rec = 0, pen = 0;
        int32_t deg[kScanItems], fnb[kScanItems];
        bool heavy_v[kScanItems], cand[kScanItems], known[kScanItems], hit[kScanItems];
        // Two stages of loads, each stage's loads for all items issued
        // together (a branch on one item's load before the next item's loads
        // would serialise them): first every per-vertex word, the row bounds
        // and witness included even for interior vertices, then the
        // witness's bit.
        // Loads from a clamped index (in bounds for every lane), so they are
        // issued unconditionally, back to back, and waited for once.
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const int64_t v = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            const bool in = v < g.n;
            const int64_t vi = in ? v : g.n - 1;
            const int32_t dv = diff[vi];
            const uint32_t hw = J.hbits[vi >> 5], bw = inb[vi >> 5];
            deg[j] = J.deg[vi];
            fnb[j] = J.wit[vi];
            heavy_v[j] = in && use_diff && ((hw >> (vi & 31)) & 1u);
            hit[j] = in && ((bw >> (vi & 31)) & 1u);
            cand[j] = in && (heavy_v[j] || !use_diff || dv != 0);  // light interior: never a record
            known[j] = use_diff && !heavy_v[j];
        }
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const int64_t v = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            if (cand[j] && heavy_v[j] && J.k <= kLpCap) {  // every label counted in the kept histogram
                if (J.hh[(int64_t)J.hid[v] * kLpCap + L[v]] == deg[j]) cand[j] = false;
                known[j] = true;
            }
        }
        // v itself, then its witness: the neighbour that last made v a
        // candidate (at first the row's smallest id, in these graphs usually
        // a hub).  The witnesses' bitmap words are gathered together.
        uint32_t wword[kScanItems];
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) wword[j] = inb[(fnb[j] >= 0 ? fnb[j] : 0) >> 5];
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const int64_t v = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            if (!cand[j]) continue;
            // The row's differing label is left to the scan unless the kept
            // counts already tell.
            if (!hit[j] && fnb[j] >= 0) hit[j] = (wword[j] >> (fnb[j] & 31)) & 1u;
            if (hit[j] && known[j]) {
                rec |= 1u << j;
            } else if (deg[j] > heavy_degree) {
                // Long rows are scanned one per wave by k_cand_pull_heavy.
                s_dq[atomicAdd(&s_dn, 1)] = ((int64_t)((known[j] ? 1 : 0) | (hit[j] ? 2 : 0)) << 32) | v;
            } else {
                pen |= 1u << j;
            }
        }
        nrec_total += __popc(rec);
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            const int64_t v = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            if (v < g.n) rdeg[v] = (rec & (1u << j)) ? deg[j] : -1;
            if (pen & (1u << j)) s_q[atomicAdd(&s_n, 1)] = (int32_t)v;
        }
        __syncthreads();
        if (s_n > kPendStage - (int)span) pend_flush(pend, &ctl->v[C_PENDING], s_q, s_n, s_base);
        if (s_dn > 0) {  // block-uniform: s_dn was read after the barrier
            const int32_t c = s_dn;
            if (threadIdx.x == 0) s_base = atomicAdd(&ctl->v[C_DEFERRED], (unsigned long long)c);
            __syncthreads();
            for (int i = threadIdx.x; i < c; i += blockDim.x) deferred[s_base + (unsigned long long)i] = s_dq[i];
            __syncthreads();
            if (threadIdx.x == 0) s_dn = 0;
            __syncthreads();
        }
    }
    __syncthreads();
    if (s_n > 0) pend_flush(pend, &ctl->v[C_PENDING], s_q, s_n, s_base);
    block_count_add(bd.cnt, nrec_total);
}

// The pull pass's pending rows (at most heavy_degree entries), one lane per
// row: kRestUnroll entries of the row per step, their column loads and
// tests issued together (the outcome is an OR over entries, so testing a
// few past the deciding one changes nothing).  Rows still undecided after
// kRestLane entries are scanned by the whole wave, 64 entries per step.
constexpr int kRestUnroll = 4;
constexpr int kRestLane = 16;
__global__ __launch_bounds__(kThreads) void k_cand_pull_rest(const P2Job *JS, const int32_t *bstart,
                                                             int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const int32_t *L = J.L;
    const uint8_t *L8 = J.L8;
    const RecDev bd = J.bd;
    const bool exact8 = J.k <= kExact8;
    Ctl *ctl = J.ctl;
    if (is_done(ctl) || ctl->cand_mode != CAND_PULL) return;
    const uint32_t *inb = J.bbits + (int64_t)((ctl->t - 1) & 1) * J.bwords;  // B_{t-1}
    const bool use_diff = ctl->diff_ok != 0;
    const int32_t *pend = J.D.light;
    const int64_t np = (int64_t)ctl->v[C_PENDING];
    int64_t nrec = 0;
    for (int64_t b0 = vwave(vg) * kWave; b0 < np; b0 += vwaves(vg) * kWave) {
        const int64_t i = b0 + lane_id();
        const bool live = i < np;
        // Loads from in-bounds indices, issued unconditionally (a
        // lane-conditional load is a branch and a wait of its own).
        const int32_t v = pend[live ? i : np - 1];
        const int64_t s0 = g.off[v];
        const int64_t e0 = g.off[v + 1];
        const int32_t lab = L[v];
        const uint32_t hw = J.hbits[v >> 5], bw = inb[v >> 5];
        const int64_t lim = live ? e0 : s0;
        int64_t xs = s0;
        const bool heavy_v = live && use_diff && ((hw >> (v & 31)) & 1u);
        // As in k_cand_pull: a pending row passed its boundary test, which
        // the kept counts answer for light rows and (k <= 32) heavy ones.
        bool dif = !live || (use_diff && (!heavy_v || J.k <= kLpCap));
        bool hit = !live || ((bw >> (v & 31)) & 1u);
        int32_t found = -1;  // the neighbour that answers `hit`, kept as v's witness
        const int64_t stop = xs + kRestLane < lim ? xs + kRestLane : lim;
        while (xs < stop && !(hit && dif)) {
            int32_t u[kRestUnroll];
#pragma unroll
            for (int q = 0; q < kRestUnroll; ++q) {
                const int32_t cu = g.col[xs + q < stop ? xs + q : xs];  // xs < stop: in the row
                u[q] = xs + q < stop ? cu : -1;
            }
            uint32_t uw[kRestUnroll];
#pragma unroll
            for (int q = 0; q < kRestUnroll; ++q) uw[q] = inb[(u[q] >= 0 ? u[q] : 0) >> 5];
#pragma unroll
            for (int q = 0; q < kRestUnroll; ++q) {
                if (u[q] < 0) continue;
                if (!dif) dif = nb_label(exact8, L8, L, u[q]) != lab;
                if (!hit && ((uw[q] >> (u[q] & 31)) & 1u)) {
                    hit = true;
                    found = u[q];
                }
            }
            xs = xs + kRestUnroll < stop ? xs + kRestUnroll : stop;
        }
        uint64_t todo = __ballot(!(hit && dif) && xs < lim);
        while (todo) {
            const int src = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            const int64_t x0 = __shfl(xs, src, kWave), x1 = __shfl(lim, src, kWave);
            const int32_t lv = __shfl(lab, src, kWave);
            bool h = __shfl(hit ? 1 : 0, src, kWave) != 0;
            bool d = __shfl(dif ? 1 : 0, src, kWave) != 0;
            int32_t fw = -1;
            int64_t b = x0;
            for (; b < x1 && !(h && d); b += kWave) {
                const int64_t x = b + lane_id();
                const int32_t u = x < x1 ? g.col[x] : -1;
                const bool dd = u >= 0 && !d && nb_label(exact8, L8, L, u) != lv;
                const uint64_t hm = __ballot(u >= 0 && !h && bit_of(inb, u));
                d = d || __ballot(dd) != 0;
                if (hm) {
                    fw = __shfl(u, __ffsll((long long)hm) - 1, kWave);
                    h = true;
                }
            }
            if (lane_id() == src) {
                hit = h;
                dif = d;
                xs = b < x1 ? b : x1;
                if (fw >= 0) found = fw;
            }
        }
        if (found >= 0) J.wit[v] = found;
        if (!live || !(hit && dif)) continue;
        J.rdeg[v] = (int32_t)(lim - s0);
        ++nrec;
    }
    block_count_add(bd.cnt, nrec);
}

// Rows longer than heavy_degree that the pull pass left undecided, one per
// wave, 64 entries per step.
__global__ void k_cand_pull_heavy(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const int32_t *L = J.L;
    const uint8_t *L8 = J.L8;
    const RecDev bd = J.bd;
    const int64_t *deferred = J.deferred;
    const bool exact8 = J.k <= kExact8;
    Ctl *ctl = J.ctl;
    if (is_done(ctl) || ctl->cand_mode != CAND_PULL) return;
    const uint32_t *inb = J.bbits + (int64_t)((ctl->t - 1) & 1) * J.bwords;  // B_{t-1}
    const int64_t nd = (int64_t)ctl->v[C_DEFERRED];
    for (int64_t j = vwave(vg); j < nd; j += vwaves(vg)) {
        const int64_t item = deferred[j];
        const int32_t v = (int32_t)(item & 0xffffffffll);
        bool dif = (item >> 32) & 1, hit = (item >> 33) & 1;
        const int32_t lv = L[v];
        const int64_t s = g.off[v], e = g.off[v + 1];
        for (int64_t b = s; b < e && !(dif && hit); b += kWave) {
            const int64_t x = b + lane_id();
            const int32_t u = x < e ? g.col[x] : -1;
            const bool d = u >= 0 && !dif && nb_label(exact8, L8, L, u) != lv;
            const uint64_t hm = __ballot(u >= 0 && !hit && bit_of(inb, u));
            dif = dif || __ballot(d) != 0;
            if (hm) {  // wave-uniform; the first such neighbour becomes v's witness
                const int32_t fw = __shfl(u, __ffsll((long long)hm) - 1, kWave);
                if (lane_id() == 0) J.wit[v] = fw;
                hit = true;
            }
        }
        const bool rec = dif && hit && lane_id() == 0;
        if (rec) {
            J.rdeg[v] = (int32_t)(e - s);
            atomicAdd(bd.cnt, 1ull);
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_cand_light(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const ListDev prev = J.B;
    const int32_t *L = J.L;
    const bool exact8 = J.k <= kExact8;
    int32_t *stamp = J.stamp;
    const RecDev bd = J.bd;
    Ctl *ctl = J.ctl;
    if (is_done(ctl) || ctl->cand_mode != CAND_PUSH) return;
    const bool diff_ok = ctl->diff_ok != 0;
    const int32_t epoch = ctl->epoch;
    __shared__ BatchLds s_b[kWavesPerBlock];
    const int w = threadIdx.x / kWave, lane = lane_id();
    int64_t cand = 0, nrec = 0;
    const int64_t count = (int64_t)prev.cnt[LC_LIGHT];
    for (int64_t b0 = vwave(vg) * kWave; b0 < count; b0 += vwaves(vg) * kWave) {
        int32_t v, deg;
        bool valid;
        const int32_t total = batch_load(g, prev.light, count, b0, s_b[w], v, deg, valid);
        touch(J, diff_ok, valid, v, stamp, epoch, L, exact8, cand, nrec);
        for (int32_t q0 = 0; q0 < total; q0 += kWave) {
            const int32_t q = q0 + lane;
            const bool on = q < total;
            int32_t u = 0;
            if (on) u = g.col[s_b[w].base[batch_slot(s_b[w], q)] + q];
            touch(J, diff_ok, on, u, stamp, epoch, L, exact8, cand, nrec);
        }
        wave_sync();
    }
    block_count_add(&ctl->v[C_CAND], cand);
    block_count_add(bd.cnt, nrec);
}

__global__ __launch_bounds__(kThreads) void k_cand_heavy(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const int32_t *L = J.L;
    const bool exact8 = J.k <= kExact8;
    int32_t *stamp = J.stamp;
    const int chunk_edges = J.chunk_edges;
    const RecDev bd = J.bd;
    Ctl *ctl = J.ctl;
    if (is_done(ctl) || ctl->cand_mode != CAND_PUSH) return;
    const bool diff_ok = ctl->diff_ok != 0;
    const int32_t epoch = ctl->epoch;
    const int lane = lane_id();
    int64_t cand = 0, nrec = 0;
    // Every chunk of the graph's heavy rows; rows in B_{t-1} (bitmap) push.
    const uint32_t *prevb = J.bbits + (int64_t)((ctl->t - 1) & 1) * J.bwords;
    const ListDev all = J.all;
    const int64_t nch = (int64_t)all.cnt[LC_CHUNKS];
    for (int64_t j = vwave(vg); j < nch; j += vwaves(vg)) {
        int32_t slot, v;
        int64_t e0, e1, deg;
        const int64_t item = all.hc[j];
        chunk_bounds(g, all, item, chunk_edges, slot, v, e0, e1, deg);
        if (!bit_of(prevb, v)) continue;  // wave-uniform
        touch(J, diff_ok, lane == 0 && (item & 0xffffffffll) == 0, v, stamp, epoch, L, exact8, cand,
              nrec);
        for (int64_t eb = e0; eb < e1; eb += kWave) {
            const int64_t e = eb + lane;
            const bool on = e < e1;
            touch(J, diff_ok, on, on ? g.col[e] : 0, stamp, epoch, L, exact8, cand, nrec);
        }
    }
    block_count_add(&ctl->v[C_CAND], cand);
    block_count_add(bd.cnt, nrec);
}

// ---------------------------------------------------------------- LP decision
// (lp_decide_by, the rule of gpu_lp_boundary.cu:388-421: lp_rule.hpp)

__device__ __forceinline__ int32_t lp_decide(const int32_t *h, int stride, int E, const double *pen,
                                             int32_t own) {
    return lp_decide_by([&](int l) { return h[l * stride]; }, E, pen, own);
}


__global__ __launch_bounds__(kThreads) void k_lp_light(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const int32_t *L = J.L;
    const ListDev B = J.B;
    const uint8_t *L8 = J.L8;
    const double *pen = J.pen;
    const int k = J.k;
    int32_t *chg = J.chg;
    uint16_t *LN = J.LN;
    const ListDev D = J.D;
    const int heavy_degree = J.heavy_degree;
    const int chunk_edges = J.chunk_edges;
    const Ctl *ctl = J.ctl;
    if (is_done(ctl)) return;
    __shared__ BatchLds s_b[kWavesPerBlock];
    __shared__ uint32_t s_h[kWavesPerBlock][kWave * kPackedStride];
    __shared__ double s_pen[kLpCap];
    const int E = k < kLpCap ? k : kLpCap;
    const int EW = (E + 1) / 2;
    if ((int)threadIdx.x < E) s_pen[threadIdx.x] = pen[threadIdx.x];
    __syncthreads();
    const int w = threadIdx.x / kWave, lane = lane_id();
    uint32_t *h = s_h[w];
    __shared__ BlockAppendLds s_app;
    const int64_t count = (int64_t)B.cnt[LC_LIGHT];
    // Block-uniform loop over groups of kWavesPerBlock batches (one per
    // wave), so the Delta append can be aggregated per block.
    for (int64_t g0 = vg.block * kWavesPerBlock * kWave; g0 < count;
         g0 += vg.blocks * kWavesPerBlock * kWave) {
        const int64_t b0 = g0 + (int64_t)w * kWave;
        int32_t v, deg;
        bool valid;
        for (int l = 0; l < EW; ++l) h[lane * kPackedStride + l] = 0u;
        const int32_t total = batch_load(g, B.light, count, b0, s_b[w], v, deg, valid);
        // kLpUnroll independent column loads, then kLpUnroll label gathers,
        // per lane per step: several misses in flight per lane.
        for (int32_t q0 = 0; q0 < total; q0 += kWave * kLpUnroll) {
            int slot[kLpUnroll];
            int32_t u[kLpUnroll];
#pragma unroll
            for (int j = 0; j < kLpUnroll; ++j) {
                const int32_t q = q0 + j * kWave + lane;
                slot[j] = q < total ? batch_slot(s_b[w], q) : -1;
                u[j] = slot[j] >= 0 ? __builtin_nontemporal_load(&g.col[s_b[w].base[slot[j]] + q])
                                    : -1;
            }
            int32_t lab[kLpUnroll];
#pragma unroll
            for (int j = 0; j < kLpUnroll; ++j) lab[j] = u[j] >= 0 ? (int32_t)L8[u[j]] : kLabel8None;
#pragma unroll
            for (int j = 0; j < kLpUnroll; ++j)
                if (lab[j] < E)
                    atomicAdd(&h[slot[j] * kPackedStride + (lab[j] >> 1)], 1u << ((lab[j] & 1) * 16));
        }
        wave_sync();
        bool changed = false;
        int32_t nl = 0;
        if (valid) {
            const int32_t own = L[v];
            const uint32_t *row = &h[lane * kPackedStride];
            nl = lp_decide_by([&](int l) { return (int32_t)((row[l >> 1] >> ((l & 1) * 16)) & 0xffffu); },
                              E, s_pen, own);
            changed = nl != own;
            if (changed) {
                chg[v] = nl;
                LN[v] = ln_pair(own, nl);
            }
        }
        block_list_append(D, changed, v, deg, heavy_degree, chunk_edges, s_app);
    }
}

// Word i (0..15) of a row held in registers: a select tree, so a
// lane-dependent index needs no scratch memory.
__device__ __forceinline__ uint32_t pick16(const uint32_t (&c)[kPackedWords], int i) {
    const bool b0 = i & 1, b1 = i & 2, b2 = i & 4, b3 = i & 8;
    const uint32_t a0 = b0 ? c[1] : c[0], a1 = b0 ? c[3] : c[2], a2 = b0 ? c[5] : c[4],
                   a3 = b0 ? c[7] : c[6], a4 = b0 ? c[9] : c[8], a5 = b0 ? c[11] : c[10],
                   a6 = b0 ? c[13] : c[12], a7 = b0 ? c[15] : c[14];
    const uint32_t d0 = b1 ? a1 : a0, d1 = b1 ? a3 : a2, d2 = b1 ? a5 : a4, d3 = b1 ? a7 : a6;
    const uint32_t e0 = b2 ? d1 : d0, e1 = b2 ? d3 : d2;
    return b3 ? e1 : e0;
}

// lp_decide_by for a row whose own label is counted (0 <= own < E), with
// fac[l] = 1 + pen[l] (the same double the loop multiplies by) and fmax the
// largest fac.  The loop's winner is the lexicographic maximum of
// (score, -bitrev5(l)), and it replaces own only when its score exceeds
// base = count(own); so only labels scoring above base matter, and the
// winner is the lexicographic maximum among them.  count * fac <= count *
// fmax, and rounding is monotone, so a label whose count is at most c_skip
// (the largest count c with c * fmax rounding to at most base) cannot score
// above base: only the few larger counts are scored.  row(q) returns word q
// of the row (16-bit halves, label l in word l / 2).
template <class Word>
__device__ __forceinline__ int32_t lp_decide_kept(Word row, int E, const double *fac, double fmax,
                                                  int32_t own) {
    const int32_t own_cnt = (int32_t)((row(own >> 1) >> ((own & 1) * 16)) & 0xffffu);
    const double base = (double)own_cnt;
    if (!(fmax > 0.0)) return own;  // every score is at most 0 <= base
    const double q0 = base / fmax;  // a first guess, made exact below
    int32_t cs = q0 >= 65535.0 ? 65535 : (int32_t)q0;
    while (cs < 65535 && (double)(cs + 1) * fmax <= base) ++cs;
    while (cs >= 0 && (double)cs * fmax > base) --cs;
    uint32_t surv = 0;
#pragma unroll
    for (int q = 0; q < kPackedWords; ++q) {
        const uint32_t wq = row(q);
        surv |= ((int32_t)(wq & 0xffffu) > cs ? 1u : 0u) << (2 * q);
        surv |= ((int32_t)(wq >> 16) > cs ? 1u : 0u) << (2 * q + 1);
    }
    if (E < 32) surv &= (1u << E) - 1u;
    int32_t best_l = -1;
    double best_s = 0.0;
    while (surv) {
        const int l = __builtin_ctz(surv);
        surv &= surv - 1u;
        const int32_t cnt = (int32_t)((row(l >> 1) >> ((l & 1) * 16)) & 0xffffu);
        const double sc = (double)cnt * fac[l];
        if (sc > base && (best_l < 0 || sc > best_s || (sc == best_s && bitrev5(l) < bitrev5(best_l)))) {
            best_l = l;
            best_s = sc;
        }
    }
    return best_l >= 0 ? best_l : own;
}

// Delta entries staged in LDS by a block: one reservation, a coalesced copy.
// Block-uniform call.
constexpr int kDStage = 2048;
__device__ __forceinline__ void d_flush(const ListDev &D, const int32_t *s_dq, int32_t &s_dn,
                                        unsigned long long &s_base) {
    const int32_t cnt = s_dn;
    if (threadIdx.x == 0) s_base = atomicAdd(&D.cnt[LC_LIGHT], (unsigned long long)cnt);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) D.light[s_base + (unsigned long long)i] = s_dq[i];
    __syncthreads();
    if (threadIdx.x == 0) s_dn = 0;
    __syncthreads();
}

// The selection's last pass and the light LP decisions in one pass over the
// vertices in id order.  Each wave takes 64 consecutive vertices per item:
// their records (degree, label byte) give the take decisions (B_t, written
// as two bitmap words per item with the member counts, as the selection's
// own dense pass would), and each light member decides at once from its
// kept histogram: one 64-byte row read (32 bytes or less for k <= 16) in
// place of a column read and a label gather per entry.  The decision needs
// only the vertex's own row and the round's penalties, so it can follow its
// take decision directly.  The rows are read as ascending addresses, which
// keeps the DRAM pages open: the rows table is far larger than the caches.
__global__ __launch_bounds__(kThreads, 6) void k_lp_light_take(const P2Job *JS, const SelArgs *SA,
                                                            const int32_t *bstart, int jobs) {
    P2_JOB();
    const SelArgs a = gload(SA[job_]);
    const uint32_t *lh = J.lh;
    const double *pen = J.pen;
    const int k = J.k;
    int32_t *chg = J.chg;
    uint16_t *LN = J.LN;
    const ListDev D = J.D;
    const ListDev sel = a.sel;
    const Ctl *ctl = J.ctl;
    if (is_done(ctl)) return;
    const SelState *st = a.st;
    const int b1_bins = a.b1_bins, id_bits = a.id_bits, deg_bits = a.deg_bits;
    const int heavy_degree = a.heavy_degree;
    const int64_t rn = a.rn, bwords = a.bwords;
    uint32_t *half = a.bits + (int64_t)(ctl->t & 1) * bwords;  // B_t
    __shared__ double s_pen[kLpCap], s_fac[kLpCap];
    __shared__ double s_fmax;
    __shared__ SelLite s_st[kSelLdsGroups];
    __shared__ unsigned long long s_sum[4];
    __shared__ int32_t s_dq[kDStage];
    __shared__ int32_t s_dn;
    __shared__ unsigned long long s_base;
    __shared__ int32_t s_av[kWavesPerBlock][kWave * kDenseItems];  // light members: id,
    __shared__ int32_t s_ao[kWavesPerBlock][kWave * kDenseItems];  // own label,
    __shared__ int32_t s_ad[kWavesPerBlock][kWave * kDenseItems];  // degree
    __shared__ int32_t s_wc[kWavesPerBlock];
    const int E = k < kLpCap ? k : kLpCap;
    const int nvec = (E + 7) / 8;
    if ((int)threadIdx.x < E) {
        s_pen[threadIdx.x] = pen[threadIdx.x];
        s_fac[threadIdx.x] = 1.0 + pen[threadIdx.x];
    }
    if (threadIdx.x < 4) s_sum[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_dn = 0;
    sel_lite_load(s_st, st, a.groups);  // ends with a barrier
    if (threadIdx.x == 0) {
        double f = s_fac[0];
        for (int l = 1; l < E; ++l) f = s_fac[l] > f ? s_fac[l] : f;
        s_fmax = f;
    }
    __syncthreads();
    const double fmax = s_fmax;
    const int w = threadIdx.x / kWave, lane = lane_id();
    const int64_t ng = (rn + kWave - 1) / kWave;
    const int64_t per_block = (int64_t)(blockDim.x / kWave) * kDenseItems;  // 64-id groups
    int64_t n_light = 0, deg_sum = 0, n_heavy = 0, heavy_deg = 0, dsum = 0;
    // Block-uniform loop, so the Delta append can be aggregated per block.
    for (int64_t base = vg.block * per_block; base < ng; base += vg.blocks * per_block) {
        const int64_t g0 = base + (int64_t)w * kDenseItems;
        int32_t d[kDenseItems], gg[kDenseItems];
        if (a.rlab8) dense_load<true>(a, g0 * kWave + lane, kWave, d, gg);
        else dense_load<false>(a, g0 * kWave + lane, kWave, d, gg);
        bool tk[kDenseItems];
#pragma unroll
        for (int j = 0; j < kDenseItems; ++j) {
            const int64_t v = (g0 + j) * kWave + lane;
            tk[j] = d[j] >= 0 && dense_take(sel_get(s_st, st, gg[j]), (int32_t)v, d[j], b1_bins, id_bits, deg_bits);
        }
        const unsigned long long mark = a.rank_mark[base / per_block];  // this chunk's rank cuts
        if (mark) rank_cut(s_st, mark, d, gg, tk, s_wc);
        uint32_t act = 0;
#pragma unroll
        for (int j = 0; j < kDenseItems; ++j) {
            const int64_t gi = g0 + j;
            if (gi >= ng) break;  // wave-uniform
            const bool take = tk[j];
            const bool light = take && d[j] <= heavy_degree;
            const uint64_t m = __ballot(take);  // the bitmap holds heavy members too
            if (lane == 0) half[2 * gi] = (uint32_t)m;
            if (lane == 1 && 2 * gi + 1 < bwords) half[2 * gi + 1] = (uint32_t)(m >> 32);
            if (take) {
                deg_sum += d[j];
                if (light) {
                    ++n_light;
                    act |= 1u << j;
                } else {
                    ++n_heavy;
                    heavy_deg += d[j];
                }
            }
        }
        // The wave's light members are packed into LDS, then decided one per
        // lane: the rows of a full batch are loaded together instead of one
        // item's (sparse) lanes at a time.
        int32_t na = 0;
#pragma unroll
        for (int j = 0; j < kDenseItems; ++j) {
            const bool on = (act >> j) & 1u;
            const uint64_t m = __ballot(on);
            if (on) {
                const int32_t pos = na + __popcll(m & ((1ull << lane) - 1ull));
                s_av[w][pos] = (int32_t)((g0 + j) * kWave + lane);
                s_ao[w][pos] = gg[j];
                s_ad[w][pos] = d[j];
            }
            na += __popcll(m);
        }
        wave_sync();
        for (int32_t b = 0; b < na; b += kWave) {  // wave-uniform
            const int32_t i = b + lane;
            if (i >= na) continue;
            const int32_t v = s_av[w][i];
            const int32_t own = s_ao[w][i];  // records carry valid labels
            const uint4 *row = reinterpret_cast<const uint4 *>(lh + (int64_t)v * kPackedWords);
            uint32_t c[kPackedWords];
#pragma unroll
            for (int q = 0; q < kPackedWords / 4; ++q) {
                const uint4 x = q < nvec ? row[q] : make_uint4(0u, 0u, 0u, 0u);
                c[4 * q] = x.x;
                c[4 * q + 1] = x.y;
                c[4 * q + 2] = x.z;
                c[4 * q + 3] = x.w;
            }
            const auto word = [&](int q) { return pick16(c, q); };
            const int32_t nl =
                (own >= 0 && own < E)
                    ? lp_decide_kept(word, E, s_fac, fmax, own)
                    : lp_decide_by([&](int l) { return (int32_t)((word(l >> 1) >> ((l & 1) * 16)) & 0xffffu); },
                                   E, s_pen, own);
            if (nl != own) {
                chg[v] = nl;
                LN[v] = ln_pair(own, nl);
                dsum += s_ad[w][i];
                // Changed vertices are staged in LDS and flushed to Delta
                // with one global reservation per kDStage / 2 or more entries.
                s_dq[atomicAdd(&s_dn, 1)] = v;
            }
        }
        __syncthreads();
        if (s_dn > kDStage - (int)(per_block * kWave)) d_flush(D, s_dq, s_dn, s_base);
    }
    __syncthreads();
    if (s_dn > 0) d_flush(D, s_dq, s_dn, s_base);
    block_count_add(&D.cnt[LC_DEG], dsum);
    n_light = wave_sum(n_light);
    deg_sum = wave_sum(deg_sum);
    n_heavy = wave_sum(n_heavy);
    heavy_deg = wave_sum(heavy_deg);
    if (lane == 0) {
        if (n_light) atomicAdd(&s_sum[0], (unsigned long long)n_light);
        if (deg_sum) atomicAdd(&s_sum[1], (unsigned long long)deg_sum);
        if (n_heavy) atomicAdd(&s_sum[2], (unsigned long long)n_heavy);
        if (heavy_deg) atomicAdd(&s_sum[3], (unsigned long long)heavy_deg);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_sum[0]) atomicAdd(&sel.cnt[LC_LIGHT], s_sum[0]);
        if (s_sum[1]) atomicAdd(&sel.cnt[LC_DEG], s_sum[1]);
        if (s_sum[2]) atomicAdd(&sel.cnt[LC_SLOTS], s_sum[2]);
        if (s_sum[3]) atomicAdd(&sel.cnt[LC_HDEG], s_sum[3]);
    }
}

// Heavy rows, pass 1: each wave counts one chunk with a peel loop over the
// distinct labels of each 64-edge step and adds the counts to the row's
// histogram in global memory.  Pass 2 (k_lp_heavy_decide) runs after the
// kernel boundary, so no fence or completion counter is needed.
__global__ __launch_bounds__(kThreads) void k_lp_heavy(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const uint8_t *L8 = J.L8;
    const ListDev B = J.B;
    const int k = J.k;
    const int chunk_edges = J.chunk_edges;
    int32_t *hub_hist = J.hub_hist;
    const Ctl *ctl = J.ctl;
    if (is_done(ctl)) return;
    // One private histogram row per lane (odd stride: rows fall in different
    // banks), summed across the wave once per chunk.
    __shared__ int32_t s_h[kWavesPerBlock][kWave * kHistStride];
    const int E = k < kLpCap ? k : kLpCap;
    const int w = threadIdx.x / kWave, lane = lane_id();
    int32_t *row = &s_h[w][lane * kHistStride];
    const int64_t nch = (int64_t)B.cnt[LC_CHUNKS];
    for (int64_t j = vwave(vg); j < nch; j += vwaves(vg)) {
        int32_t slot, v;
        int64_t e0, e1, deg;
        chunk_bounds(g, B, B.hc[j], chunk_edges, slot, v, e0, e1, deg);
        for (int l = 0; l < E; ++l) row[l] = 0;
        for (int64_t eb = e0; eb < e1; eb += kWave * kHeavyUnroll) {
            int32_t u[kHeavyUnroll];
#pragma unroll
            for (int j = 0; j < kHeavyUnroll; ++j) {
                const int64_t e = eb + j * kWave + lane;
                u[j] = e < e1 ? __builtin_nontemporal_load(&g.col[e]) : -1;
            }
            int32_t lab[kHeavyUnroll];
#pragma unroll
            for (int j = 0; j < kHeavyUnroll; ++j) lab[j] = u[j] >= 0 ? (int32_t)L8[u[j]] : kLabel8None;
#pragma unroll
            for (int j = 0; j < kHeavyUnroll; ++j)
                if (lab[j] < E) row[lab[j]] += 1;
        }
        wave_sync();
        if (lane < E) {
            int32_t c = 0;
            for (int r = 0; r < kWave; ++r) c += s_h[w][r * kHistStride + lane];
            if (c) atomicAdd(&hub_hist[(int64_t)slot * kLpCap + lane], c);
        }
        wave_sync();
    }
}

// Heavy rows, pass 2: one thread per heavy row decides from its histogram
// and clears it for the next round.
__global__ void k_lp_heavy_decide(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const int32_t *L = J.L;
    const ListDev B = J.B;
    const double *pen = J.pen;
    const int k = J.k;
    int32_t *chg = J.chg;
    uint16_t *LN = J.LN;
    const ListDev D = J.D;
    const int heavy_degree = J.heavy_degree;
    const int chunk_edges = J.chunk_edges;
    int32_t *hub_hist = J.hub_hist;
    const int64_t *off = J.g.off;
    const Ctl *ctl = J.ctl;
    if (is_done(ctl)) return;
    __shared__ double s_pen[kLpCap];
    const int E = k < kLpCap ? k : kLpCap;
    if ((int)threadIdx.x < E) s_pen[threadIdx.x] = pen[threadIdx.x];
    __syncthreads();
    const int64_t ns = (int64_t)B.cnt[LC_SLOTS];
    for (int64_t b0 = vwave(vg) * kWave; b0 < ns; b0 += vwaves(vg) * kWave) {
        const int64_t slot = b0 + lane_id();
        bool changed = false;
        int32_t v = 0;
        int64_t deg = 0;
        if (slot < ns) {
            v = B.hv[slot];
            deg = off[v + 1] - off[v];
            int32_t *h = hub_hist + slot * kLpCap;
            const int32_t own = L[v];
            const int32_t nl = lp_decide(h, 1, E, s_pen, own);
            for (int l = 0; l < E; ++l) h[l] = 0;
            changed = nl != own;
            if (changed) {
                chg[v] = nl;
                LN[v] = ln_pair(own, nl);
            }
        }
        list_append(D, changed, v, deg, heavy_degree, chunk_edges);
    }
}

// ---------------------------------------------------------------- delta stats


// One entry (v, u) of a changed vertex v (label a -> c).  Returns the
// entry's intra-partition effect: bit 0 the old pair was intra (a == lu),
// bit 16 the new one is (c == lu2); the cut change through `dd_out`.  An
// entry whose ends both changed is counted from the smaller id only.
// Callers fold the counts once per vertex, since a and c are the vertex's
// own old and new labels.  `same_out` tells whether u's new label is c (every
// entry: v's new differing count is its degree minus these).  When u keeps
// its label, its differing count moves here (keep_diff).
// u's old label lu, its label after this round lu2, and whether it changed
// (uchg) come from the caller: pair_effect below, or a caller that gathered
// the LN words of several entries together.
__device__ __forceinline__ int32_t pair_effect_of(int32_t v, int32_t a, int32_t c, int32_t u, int32_t lu,
                                                  int32_t lu2, bool uchg, bool keep_diff, int32_t *diff,
                                                  int32_t &dd_out, bool &same_out) {
    dd_out = 0;
    if (uchg) {
        same_out = lu2 == c;
        if (u < v) return 0;  // counted from u's side
    }
    if (!uchg) {
        same_out = lu == c;
        const int32_t du = (c != lu ? 1 : 0) - (a != lu ? 1 : 0);
        if (keep_diff && du) atomicAdd(&diff[u], du);
    }
    const int32_t dd = (c != lu2 ? 1 : 0) - (a != lu ? 1 : 0);
    dd_out = dd;
    return (a == lu ? 1 : 0) | (c == lu2 ? 1 << 16 : 0);
}
// LN word of a neighbour: old label byte | this round's new one << 8 (255:
// unchanged).
__device__ __forceinline__ void ln_labels(uint32_t ln, int32_t &lu, int32_t &lu2, bool &uchg) {
    const int32_t nu = (int32_t)(ln >> 8);
    uchg = nu != kLabel8None;
    lu = (int32_t)(ln & 0xffu);
    lu2 = uchg ? nu : lu;
}
__device__ __forceinline__ int32_t pair_effect(int32_t v, int32_t a, int32_t c, int32_t u,
                                               const int32_t *L, const int32_t *chg,
                                               const uint16_t *LN, bool exact8,
                                               bool keep_diff, int32_t *diff, int32_t &dd_out,
                                               bool &same_out) {
    int32_t lu, lu2;
    bool uchg;
    if (exact8) {  // one 16-bit gather
        ln_labels(LN[u], lu, lu2, uchg);
    } else {
        const int32_t cu = chg[u];
        uchg = cu >= 0;
        lu = L[u];
        lu2 = uchg ? cu : lu;
    }
    return pair_effect_of(v, a, c, u, lu, lu2, uchg, keep_diff, diff, dd_out, same_out);
}

// Row upkeep of the kept histograms, done while Delta's rows are read (the
// round's decisions are final, and the histograms are next read by the next
// round).  A changed vertex v (label a -> c) appears once in the row of each
// neighbour u (the CSR is symmetric), so it moves one count of u's row from
// a to c: u's kept light row (lh_on) or, in push mode, u's heavy histogram.
// Labels outside [0, E) are not counted.  The 16-bit halves of a light row
// stay within [0, 65535], so one 32-bit add moves both halves of a word.
__device__ __forceinline__ void lh_move(uint32_t *lh, int32_t u, int32_t a, int32_t c, int E) {
    uint32_t *row = lh + (int64_t)u * kPackedWords;
    const bool da = a >= 0 && a < E, dc = c >= 0 && c < E;
    const uint32_t ma = da ? 1u << ((a & 1) * 16) : 0u, mc = dc ? 1u << ((c & 1) * 16) : 0u;
    if (da && dc && (a >> 1) == (c >> 1)) {
        atomicAdd(&row[a >> 1], mc - ma);
        return;
    }
    if (da) atomicAdd(&row[a >> 1], 0u - ma);
    if (dc) atomicAdd(&row[c >> 1], mc);
}
// The same moves for one entry per lane of a full wave, two lanes an entry:
// in each of two instructions lane 2m moves entry m's count out of the old
// label's word and lane 2m+1 into the new label's word (entries of lanes
// 0-31, then 32-63), so each instruction touches a row's line once.  The
// rate of scattered atomics follows the number of lines an instruction
// touches, so this halves the cost of moves whose labels sit in different
// words.  u: the lane's neighbour, or -1; ac: old | new << 8, each 0xff
// when not counted.  Called by every lane of the wave.
__device__ __forceinline__ void lh_move_paired(uint32_t *lh, int32_t u, uint32_t ac, int lane) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int src = r * (kWave / 2) + (lane >> 1);
        const int32_t su = __shfl(u, src, kWave);
        const uint32_t sac = (uint32_t)__shfl((int32_t)ac, src, kWave);
        const uint32_t a = sac & 0xffu, c = (sac >> 8) & 0xffu;
        const bool da = a != 0xffu, dc = c != 0xffu;
        const bool odd = (lane & 1) != 0;
        uint32_t word, val;
        bool act;
        if (da && dc && (a >> 1) == (c >> 1)) {  // one add moves both halves
            act = !odd;
            word = a >> 1;
            val = (1u << ((c & 1) * 16)) - (1u << ((a & 1) * 16));
        } else if (!odd) {
            act = da;
            word = a >> 1;
            val = 0u - (1u << ((a & 1) * 16));
        } else {
            act = dc;
            word = c >> 1;
            val = 1u << ((c & 1) * 16);
        }
        if (su >= 0 && act) atomicAdd(lh + (int64_t)su * kPackedWords + word, val);
    }
}
__device__ __forceinline__ void hh_move(const int32_t *hid, int32_t *hhc, int64_t nheavy, int32_t u,
                                        int32_t a, int32_t c, int E) {
    int32_t *row = hhc + ((int64_t)(blockIdx.x % kHhCopies) * nheavy + hid[u]) * kLpCap;
    if (a >= 0 && a < E) atomicAdd(&row[a], -1);
    if (c >= 0 && c < E) atomicAdd(&row[c], 1);
}
__device__ __forceinline__ void row_upkeep(const P2Job &J, bool lh_on, bool hh_push, bool u_heavy,
                                           int32_t u, int32_t a, int32_t c, int E) {
    if (u_heavy) {
        if (hh_push) hh_move(J.hid, J.hhc, (int64_t)J.all.cnt[LC_SLOTS], u, a, c, E);
    } else if (lh_on) {
        lh_move(J.lh, u, a, c, E);
    }
}

__global__ __launch_bounds__(kThreads, 4) void k_delta_light(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const ListDev D = J.D;
    const int32_t *L = J.L;
    const int32_t *chg = J.chg;
    const uint16_t *LN = J.LN;
    const int k = J.k;
    const bool exact8 = k <= kExact8;
    Ctl *ctl = J.ctl;
    unsigned long long *vc = J.vc;
    unsigned long long *ic = J.ic;
    if (is_done(ctl)) return;
    const bool lh_on = J.lh_on != 0, hh_push = ctl->hist_mode == HIST_PUSH;
    const bool keep_diff = ctl->diff_ok != 0;
    const int64_t nheavy = (int64_t)J.all.cnt[LC_SLOTS];
    int32_t *diff = J.diff;
    const int E = k < kLpCap ? k : kLpCap;
    __shared__ BatchLds s_b[kWavesPerBlock];
    __shared__ int32_t s_io[kWavesPerBlock][kWave];
    __shared__ int32_t s_new[kWavesPerBlock][kWave];
    __shared__ int32_t s_same[kWavesPerBlock][kWave];
    __shared__ unsigned long long s_vc[kLdsK], s_ic[kLdsK];
    for (int i = threadIdx.x; i < kLdsK; i += blockDim.x) { s_vc[i] = 0; s_ic[i] = 0; }
    __syncthreads();
    const int w = threadIdx.x / kWave, lane = lane_id();
    int64_t cut_delta = 0;
    const int64_t count = (int64_t)D.cnt[LC_LIGHT];
    for (int64_t b0 = vwave(vg) * kWave; b0 < count; b0 += vwaves(vg) * kWave) {
        int32_t v, deg;
        bool valid;
        const int32_t total = batch_load(g, D.light, count, b0, s_b[w], v, deg, valid);
        s_io[w][lane] = 0;
        s_same[w][lane] = 0;
        s_b[w].aux[lane] = valid ? L[v] : 0;  // old label of the slot's vertex
        s_new[w][lane] = valid ? chg[v] : 0;  // its new label
        wave_sync();
        // Two versions of the entry loop, with and without the heavy-slot
        // gather, so that each issues its loads unconditionally.
        auto entries = [&](auto push) {
            constexpr bool Push = decltype(push)::value;
            for (int32_t q0 = 0; q0 < total; q0 += kWave * kLpUnroll) {
                // Three stages of loads, each issued for all kLpUnroll entries
                // together from in-bounds indices (a lane-conditional load would
                // become a branch and a wait of its own): the columns, then each
                // neighbour's heavy bit word and LN word (and heavy slot when
                // heavy rows take pushed counts), then the updates.
                int slot[kLpUnroll];
                int32_t u[kLpUnroll];
    #pragma unroll
                for (int j = 0; j < kLpUnroll; ++j) {
                    const int32_t q = q0 + j * kWave + lane;
                    slot[j] = q < total ? batch_slot(s_b[w], q) : -1;
                    const int64_t e = s_b[w].base[slot[j] >= 0 ? slot[j] : 0] + q;
                    const int32_t cu = g.col[slot[j] >= 0 ? e : 0];
                    u[j] = slot[j] >= 0 ? cu : -1;
                }
                uint32_t hw[kLpUnroll], ln[kLpUnroll], hs[kLpUnroll];  // hs: unsigned, no sign extension to wait for
    #pragma unroll
                for (int j = 0; j < kLpUnroll; ++j) {
                    const int32_t uc = u[j] >= 0 ? u[j] : 0;
                    hw[j] = J.hbits[uc >> 5];
                    ln[j] = exact8 ? (uint32_t)LN[uc] : 0u;
                    hs[j] = Push ? (uint32_t)J.hid[uc] : 0u;  // unconditional within the version
                }
                // Each entry's counts for its row (old label hits, new
                // label hits, same-label neighbours), summed over the lanes
                // of the row after the loop: one LDS add per row instead of
                // one per entry (the lanes of a row hit one LDS word).
                uint32_t sx[kLpUnroll];
    #pragma unroll
                for (int j = 0; j < kLpUnroll; ++j) {
                    sx[j] = 0xff000000u;
                    int32_t mu = -1;        // kept-row move of the entry: neighbour or -1
                    uint32_t mac = 0xffffu; // and its old | new << 8 (0xff: not counted)
                    if (slot[j] >= 0) {
                    const int32_t vv = s_b[w].v[slot[j]];
                    const int32_t a = s_b[w].aux[slot[j]], c = s_new[w][slot[j]];
                    int32_t dd, lu, lu2;
                    bool same, uchg;
                    const bool uh = (hw[j] >> (u[j] & 31)) & 1u;
                    if (exact8) {
                        ln_labels(ln[j], lu, lu2, uchg);
                    } else {
                        const int32_t cu = chg[u[j]];
                        uchg = cu >= 0;
                        lu = L[u[j]];
                        lu2 = uchg ? cu : lu;
                    }
                    const int32_t io = pair_effect_of(vv, a, c, u[j], lu, lu2, uchg, keep_diff && !uh, diff, dd, same);
                    cut_delta += dd;
                    sx[j] = (uint32_t)slot[j] << 24 | (uint32_t)(io & 1) | (uint32_t)(io >> 16) << 7 |
                            (same ? 1u << 14 : 0u);
                    if (uh) {
                        if (Push) {
                            int32_t *row = J.hhc + ((int64_t)(blockIdx.x % kHhCopies) * nheavy + (int64_t)hs[j]) * kLpCap;
                            if (a >= 0 && a < E) atomicAdd(&row[a], -1);
                            if (c >= 0 && c < E) atomicAdd(&row[c], 1);
                        }
                    } else if (lh_on) {
                        mu = u[j];
                        mac = (a >= 0 && a < E ? (uint32_t)a : 0xffu) |
                              (c >= 0 && c < E ? (uint32_t)c : 0xffu) << 8;
                    }
                    }
                    if (lh_on) lh_move_paired(J.lh, mu, mac, lane);  // every lane, convergent
                }
                // Fields of 7 bits: at most 64 lanes add a 0 or 1 to each.
    #pragma unroll
                for (int j = 0; j < kLpUnroll; ++j) {
                    const uint32_t t = wave_seg_incl(sx[j]);
                    const int32_t q = q0 + j * kWave + lane;
                    if (slot[j] >= 0 && (lane == kWave - 1 || q + 1 == s_b[w].incl[slot[j]])) {
                        const uint32_t io = (t & 0x7fu) | ((t >> 7) & 0x7fu) << 16, same = (t >> 14) & 0x7fu;
                        if (io) atomicAdd(&s_io[w][slot[j]], (int32_t)io);
                        if (same) atomicAdd(&s_same[w][slot[j]], (int32_t)same);
                    }
                }
            }
        };
        if (hh_push) entries(std::true_type{});
        else entries(std::false_type{});
        wave_sync();
        if (valid) {
            const int32_t a = L[v], c = chg[v];
            const int32_t io = s_io[w][lane];
            if (keep_diff) diff[v] = deg - s_same[w][lane];
            if (a >= 0 && a < k) {
                label_add(s_vc, vc, a, ~0ull);
                if (io & 0xffff) label_add(s_ic, ic, a, (unsigned long long)(-(int64_t)(io & 0xffff)));
            }
            if (c >= 0 && c < k) {
                label_add(s_vc, vc, c, 1ull);
                if (io >> 16) label_add(s_ic, ic, c, (unsigned long long)(io >> 16));
            }
        }
        wave_sync();
    }
    block_count_add(&ctl->v[C_CUT], cut_delta);
    __syncthreads();
    label_flush(s_vc, vc, k);
    label_flush(s_ic, ic, k);
}

__global__ __launch_bounds__(kThreads) void k_delta_heavy(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const ListDev D = J.D;
    const int32_t *L = J.L;
    const int32_t *chg = J.chg;
    const uint16_t *LN = J.LN;
    const int k = J.k;
    const bool exact8 = k <= kExact8;
    const int chunk_edges = J.chunk_edges;
    Ctl *ctl = J.ctl;
    unsigned long long *vc = J.vc;
    unsigned long long *ic = J.ic;
    if (is_done(ctl)) return;
    const bool lh_on = J.lh_on != 0, hh_push = ctl->hist_mode == HIST_PUSH;
    const bool upkeep = lh_on || hh_push;
    const bool keep_diff = ctl->diff_ok != 0;
    int32_t *diff = J.diff;
    const int E = k < kLpCap ? k : kLpCap;
    __shared__ unsigned long long s_ic[kLdsK];
    for (int i = threadIdx.x; i < kLdsK; i += blockDim.x) s_ic[i] = 0;
    __syncthreads();
    const int lane = lane_id();
    int64_t cut_delta = 0;
    const int64_t nch = (int64_t)D.cnt[LC_CHUNKS];
    for (int64_t j = vwave(vg); j < nch; j += vwaves(vg)) {
        int32_t slot, v;
        int64_t e0, e1, deg;
        const int64_t item = D.hc[j];
        chunk_bounds(g, D, item, chunk_edges, slot, v, e0, e1, deg);
        const int32_t a = L[v], c = chg[v];
        int64_t dv = 0, oi = 0, ni = 0;
        for (int64_t e = e0 + lane; e < e1; e += kWave) {
            int32_t dd;
            bool same;
            const int32_t u = g.col[e];
            const bool uh = bit_of(J.hbits, u);
            const int32_t io = pair_effect(v, a, c, u, L, chg, LN, exact8, keep_diff && !uh, diff,
                                           dd, same);
            if (upkeep) row_upkeep(J, lh_on, hh_push, uh, u, a, c, E);
            dv += dd;
            oi += io & 0xffff;
            ni += io >> 16;
        }
        dv = wave_sum(dv);
        oi = wave_sum(oi);
        ni = wave_sum(ni);
        cut_delta += lane == 0 ? dv : 0;
        if (lane == 0) {
            if (a >= 0 && a < k && oi) label_add(s_ic, ic, a, (unsigned long long)(-oi));
            if (c >= 0 && c < k && ni) label_add(s_ic, ic, c, (unsigned long long)ni);
            if ((item & 0xffffffffll) == 0) {
                if (a >= 0 && a < k) atomicAdd(&vc[a], ~0ull);
                if (c >= 0 && c < k) atomicAdd(&vc[c], 1ull);
            }
        }
    }
    block_count_add(&ctl->v[C_CUT], cut_delta);
    __syncthreads();
    label_flush(s_ic, ic, k);
}

// Byte mirror of the labels: the label if it fits a byte, else kLabel8None.
// The LP ignores entries at or above its label cap (gpu_lp_boundary.cu:360).
__global__ void k_mirror8(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const int64_t n = J.g.n;
    const int32_t *L = J.L;
    uint8_t *L8 = J.L8;
    uint16_t *LN = J.LN;
    for (int64_t v = vthread(vg); v < n;
         v += vthreads(vg)) {
        const int32_t l = L[v];
        const uint32_t b = (l >= 0 && l < kLabel8None) ? (uint32_t)l : (uint32_t)kLabel8None;
        L8[v] = (uint8_t)b;
        LN[v] = (uint16_t)(b | (kLabel8None << 8));
        J.wit[v] = J.first[v];
    }
}

// `round` (the LP round, graph with heavy rows when heavy != 0): the
// changed-vertex state the round's later kernels read is written only when
// they read it: the old|new pairs and changed bits only for a scanning
// heavy-histogram update, the new-label reset only where the word labels
// (k above the byte range) tell changed vertices apart.  The test hook
// passes 0 and gets every write.
__global__ void k_apply(const P2Job *JS, const int32_t *bstart, int jobs, int round, int heavy) {
    P2_JOB();
    // Recount rounds: the heavy histograms are cleared here, in the same
    // launch (nothing reads them between the Delta pass and the recount).
    if (round && heavy && !is_done(J.ctl) && J.ctl->hist_mode == HIST_RECOUNT) {
        const int64_t cnt = (int64_t)J.all.cnt[LC_SLOTS] * kLpCap;
        for (int64_t i = vthread(vg); i < cnt; i += vthreads(vg)) J.hh[i] = 0;
    }
    const ListDev D = J.D;
    int32_t *L = J.L;
    uint8_t *L8 = J.L8;
    uint16_t *LN = J.LN;
    uint16_t *ON = J.ON;
    uint32_t *cbits = J.cbits;
    int32_t *chg = J.chg;
    const Ctl *ctl = J.ctl;
    if (is_done(ctl)) return;
    const int64_t nl = (int64_t)D.cnt[LC_LIGHT], ns = (int64_t)D.cnt[LC_SLOTS];
    const int64_t total = nl + ns;
    const bool scan = !round || (heavy && ctl->hist_mode == HIST_SCAN);
    const bool reset = !round || J.k > kExact8;
    constexpr int kApplyItems = 4;  // changed vertices per thread per step, loads issued together
    const int64_t step = vthreads(vg);
    for (int64_t i0 = vthread(vg); i0 < total; i0 += kApplyItems * step) {
        int32_t v[kApplyItems], nv[kApplyItems];
        uint32_t ob[kApplyItems];
#pragma unroll
        for (int j = 0; j < kApplyItems; ++j) {
            const int64_t i = i0 + j * step;
            const int64_t ic = i < total ? i : total - 1;  // in bounds: the load issues unconditionally
            const int32_t a = D.light[ic < nl ? ic : 0], b = D.hv[ic < nl ? 0 : ic - nl];
            v[j] = ic < nl ? a : b;
        }
#pragma unroll
        for (int j = 0; j < kApplyItems; ++j) {
            nv[j] = chg[v[j]];
            ob[j] = L8[v[j]];
        }
#pragma unroll
        for (int j = 0; j < kApplyItems; ++j) {
            if (i0 + j * step >= total) continue;
            const int32_t u = v[j];
            const int32_t nl_v = nv[j];
            L[u] = nl_v;
            const uint32_t nb = (nl_v >= 0 && nl_v < kLabel8None) ? (uint32_t)nl_v : (uint32_t)kLabel8None;
            L8[u] = (uint8_t)nb;
            LN[u] = (uint16_t)(nb | (kLabel8None << 8));
            if (reset) chg[u] = -1;
            if (scan) {
                ON[u] = (uint16_t)(ob[j] | (nb << 8));
                atomicOr(&cbits[u >> 5], 1u << (u & 31));
            }
        }
    }
}

// ---------------------------------------------------------------- heavy-row histograms

// Heavy rows of B_t decide from their kept histograms (one thread per row).
__global__ void k_lp_heavy_hist(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const int32_t *L = J.L;
    const int k = J.k;
    int32_t *chg = J.chg;
    uint16_t *LN = J.LN;
    const ListDev D = J.D;
    const int64_t *off = J.g.off;
    if (is_done(J.ctl)) return;
    __shared__ double s_pen[kLpCap];
    const int E = k < kLpCap ? k : kLpCap;
    if ((int)threadIdx.x < E) s_pen[threadIdx.x] = J.pen[threadIdx.x];
    __syncthreads();
    // Every heavy row of the graph; those in B_t (bitmap) decide.
    const uint32_t *bt = J.bbits + (int64_t)(J.ctl->t & 1) * J.bwords;
    const ListDev all = J.all;
    const int64_t ns = (int64_t)all.cnt[LC_SLOTS];
    for (int64_t b0 = vwave(vg) * kWave; b0 < ns; b0 += vwaves(vg) * kWave) {
        const int64_t slot = b0 + lane_id();
        bool changed = false;
        int32_t v = 0;
        int64_t deg = 0;
        if (slot < ns && bit_of(bt, all.hv[slot])) {
            v = all.hv[slot];
            deg = off[v + 1] - off[v];
            const int32_t *h = J.hh + (int64_t)J.hid[v] * kLpCap;
            const int32_t own = L[v];
            const int32_t nl = lp_decide(h, 1, E, s_pen, own);
            changed = nl != own;
            if (changed) {
                chg[v] = nl;
                LN[v] = ln_pair(own, nl);
            }
        }
        list_append(D, changed, v, deg, J.heavy_degree, J.chunk_edges);
    }
}

// After the apply: how the heavy histograms follow this round's changes.
// Many changes: recount them.  Few: push each changed vertex's move to its
// heavy neighbours.  Otherwise: scan the heavy rows against the
// changed-vertex bitmap.

__global__ void k_hist_mode(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    Ctl *ctl = J.ctl;
    if (is_done(ctl)) return;
    const double chg_deg = (double)J.D.cnt[LC_DEG], nnz = (double)J.g.nnz;
    ctl->hist_mode = chg_deg >= J.recount_share * nnz ? HIST_RECOUNT
                     : (chg_deg <= J.push_share * nnz ? HIST_PUSH : HIST_SCAN);
}


// Recount mode (after rounds where many vertices changed): every heavy
// row's label histogram from scratch, one chunk per wave, per-lane LDS rows.
__global__ __launch_bounds__(kThreads) void k_hist_recount(const P2Job *JS, const int32_t *bstart,
                                                           int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const ListDev all = J.all;
    const uint8_t *L8 = J.L8;
    int32_t *hh = J.hh;
    if (is_done(J.ctl) || J.ctl->hist_mode != HIST_RECOUNT) return;
    const int E = J.k < kLpCap ? J.k : kLpCap;
    const int chunk_edges = J.chunk_edges;
    __shared__ int32_t s_h[kWavesPerBlock][kWave * kHistStride];
    const int w = threadIdx.x / kWave, lane = lane_id();
    int32_t *row = &s_h[w][lane * kHistStride];
    const int64_t nch = (int64_t)all.cnt[LC_CHUNKS];
    for (int64_t j = vwave(vg); j < nch; j += vwaves(vg)) {
        int32_t slot, v;
        int64_t e0, e1, deg;
        chunk_bounds(g, all, all.hc[j], chunk_edges, slot, v, e0, e1, deg);
        for (int l = 0; l < E; ++l) row[l] = 0;
        for (int64_t eb = e0; eb < e1; eb += kWave * kHeavyUnroll) {
            // In-bounds indices: the loads issue unconditionally.
            int32_t u[kHeavyUnroll];
#pragma unroll
            for (int q = 0; q < kHeavyUnroll; ++q) {
                const int64_t e = eb + q * kWave + lane;
                const int32_t cu = __builtin_nontemporal_load(&g.col[e < e1 ? e : e1 - 1]);
                u[q] = e < e1 ? cu : -1;
            }
            int32_t lab[kHeavyUnroll];
#pragma unroll
            for (int q = 0; q < kHeavyUnroll; ++q) {
                const int32_t lb = L8[u[q] >= 0 ? u[q] : 0];
                lab[q] = u[q] >= 0 ? lb : kLabel8None;
            }
#pragma unroll
            for (int q = 0; q < kHeavyUnroll; ++q)
                if (lab[q] < E) row[lab[q]] += 1;
        }
        wave_sync();
        if (lane < E) {
            int32_t c = 0;
            for (int r = 0; r < kWave; ++r) c += s_h[w][r * kHistStride + lane];
            if (c) atomicAdd(&hh[(int64_t)slot * kLpCap + lane], c);
        }
        wave_sync();
    }
}

// Push mode: add the copies of this round's moves into hh and clear them.
__device__ __forceinline__ void hist_fold(const P2Job &J, const VGrid &vg) {
    const int64_t nhh = (int64_t)J.all.cnt[LC_SLOTS] * kLpCap;
    for (int64_t i = vthread(vg); i < nhh; i += vthreads(vg)) {
        int32_t sum = 0;
#pragma unroll
        for (int c = 0; c < kHhCopies; ++c) {
            int32_t *p = &J.hhc[(int64_t)c * nhh + i];
            sum += *p;
            *p = 0;
        }
        if (sum) J.hh[i] += sum;
    }
}

// Update mode: one chunk of a heavy row per wave.  Each neighbour's bit in
// the changed-vertex bitmap (n bits: a small, well-cached array) is tested;
// a changed neighbour moves one count from its old to its new label (one
// 16-bit gather: old | new << 8) in a per-wave LDS row.  Few neighbours
// change, so the LDS atomics rarely collide, and the kernel needs almost no
// LDS, which keeps occupancy high for the column stream.
__global__ __launch_bounds__(kThreads) void k_hist_update(const P2Job *JS, const int32_t *bstart,
                                                          int jobs) {
    P2_JOB();
    const GraphDev g = J.g;
    const ListDev all = J.all;
    const uint16_t *ON = J.ON;
    const uint32_t *cbits = J.cbits;
    int32_t *hh = J.hh;
    if (is_done(J.ctl)) return;
    if (J.ctl->hist_mode == HIST_PUSH) {  // push rounds fold in the same launch
        hist_fold(J, vg);
        return;
    }
    if (J.ctl->hist_mode != HIST_SCAN) return;
    const int E = J.k < kLpCap ? J.k : kLpCap;
    const int chunk_edges = J.chunk_edges;
    __shared__ int32_t s_w[kWavesPerBlock][kLpCap];
    const int w = threadIdx.x / kWave, lane = lane_id();
    int32_t *wrow = s_w[w];
    if (lane < kLpCap) wrow[lane] = 0;
    wave_sync();
    const int64_t nch = (int64_t)all.cnt[LC_CHUNKS];
    for (int64_t j = vwave(vg); j < nch; j += vwaves(vg)) {
        int32_t slot, v;
        int64_t e0, e1, deg;
        chunk_bounds(g, all, all.hc[j], chunk_edges, slot, v, e0, e1, deg);
        bool touched = false;
        for (int64_t eb = e0; eb < e1; eb += kWave * kHeavyUnroll) {
            // In-bounds indices: the loads issue unconditionally.
            int32_t u[kHeavyUnroll];
#pragma unroll
            for (int q = 0; q < kHeavyUnroll; ++q) {
                const int64_t e = eb + q * kWave + lane;
                const int32_t cu = __builtin_nontemporal_load(&g.col[e < e1 ? e : e1 - 1]);
                u[q] = e < e1 ? cu : -1;
            }
            uint32_t word[kHeavyUnroll];
#pragma unroll
            for (int q = 0; q < kHeavyUnroll; ++q) {
                const uint32_t cw = cbits[(u[q] >= 0 ? u[q] : 0) >> 5];
                word[q] = u[q] >= 0 ? cw : 0u;
            }
            // Unchanged entries read ON[0] (one shared line), so these loads
            // too issue unconditionally.
            uint32_t onw[kHeavyUnroll];
#pragma unroll
            for (int q = 0; q < kHeavyUnroll; ++q) {
                const bool ch = u[q] >= 0 && ((word[q] >> (u[q] & 31)) & 1u);
                onw[q] = ON[ch ? u[q] : 0];
            }
#pragma unroll
            for (int q = 0; q < kHeavyUnroll; ++q) {
                if (u[q] < 0 || !((word[q] >> (u[q] & 31)) & 1u)) continue;
                const uint32_t on = onw[q];
                const int32_t o = (int32_t)(on & 0xffu), nw = (int32_t)(on >> 8);
                if (o < E) atomicAdd(&wrow[o], -1);
                if (nw < E) atomicAdd(&wrow[nw], 1);
                touched = true;
            }
        }
        if (__ballot(touched) == 0) continue;  // wave-uniform
        wave_sync();
        if (lane < E) {
            const int32_t c = wrow[lane];
            if (c) atomicAdd(&hh[(int64_t)slot * kLpCap + lane], c);
            wrow[lane] = 0;
        }
        wave_sync();
    }
}

// Clear the changed-vertex bits of this round (every writer stores zero).
// `round`: the LP round, which sets the bits only in scanning rounds.
__global__ void k_cbits_clear(const P2Job *JS, const int32_t *bstart, int jobs, int round, int heavy) {
    P2_JOB();
    const ListDev D = J.D;
    if (is_done(J.ctl)) return;
    if (round && !(heavy && J.ctl->hist_mode == HIST_SCAN)) return;  // uniform per job
    const int64_t nl = (int64_t)D.cnt[LC_LIGHT], ns = (int64_t)D.cnt[LC_SLOTS];
    for (int64_t i = vthread(vg); i < nl + ns; i += vthreads(vg)) {
        const int32_t v = i < nl ? D.light[i] : D.hv[i - nl];
        J.cbits[v >> 5] = 0u;
    }
}

// ---------------------------------------------------------------- round end

__global__ void k_round_end(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    Ctl *ctl = J.ctl;
    const ListDev cur = J.B;
    const ListDev D = J.D;
    const RecDev bd = J.bd;
    int64_t *trace = J.trace;
    const int trace_cap = J.trace_cap;
    const int64_t n = J.g.n;
    if (is_done(ctl)) return;
    const int t = ctl->t;
    const int max_rounds = ctl->max_rounds;
    const double epsilon = ctl->epsilon;
    const int64_t nprev = (int64_t)ctl->v[C_PREV_N];
    const int64_t dprev = (int64_t)ctl->v[C_PREV_DEG];
    const int64_t ncand = ctl->cand_mode == CAND_ALL ? n : (ctl->cand_mode == CAND_PULL ? -1 : (int64_t)ctl->v[C_CAND]);
    const int64_t nb = (int64_t)(cur.cnt[LC_LIGHT] + cur.cnt[LC_SLOTS]);
    const int64_t bdeg = (int64_t)cur.cnt[LC_DEG];
    const int64_t nchg = (int64_t)(D.cnt[LC_LIGHT] + D.cnt[LC_SLOTS]);
    const int64_t cdeg = (int64_t)D.cnt[LC_DEG];
    ctl->rounds = t;
    ctl->v[C_LP_EDGES] += (unsigned long long)bdeg;
    int64_t cut_rec = -1;
    if (ctl->v[C_KERNEL_ERR] != 0) {  // a kernel of this round found an impossible state
        ctl->done = 1;
        ctl->err = DMOLP_E_INTERNAL;
    } else if (nb == 0) {
        ctl->done = 1;
        ctl->reason = DMOLP_EXIT_EMPTY_BOUNDARY;
    } else if (nchg == 0) {
        ctl->done = 1;
        ctl->reason = DMOLP_EXIT_NO_CHANGE;
    } else {
        const int64_t curc = (int64_t)ctl->v[C_CUT];
        const int64_t prevc = (int64_t)ctl->v[C_PREV_CUT];
        const double delta =
            (prevc > 0) ? fabs((double)(curc - prevc) / (double)prevc) : 1.0;  // phase2.cpp:397
        if (delta < epsilon) ctl->cc += 1; else ctl->cc = 0;
        ctl->v[C_PREV_CUT] = (unsigned long long)curc;
        cut_rec = curc;
        if (ctl->cc >= 10) { ctl->done = 1; ctl->reason = DMOLP_EXIT_CONVERGED; }
    }
    if (!ctl->done && t >= max_rounds) { ctl->done = 1; ctl->reason = DMOLP_EXIT_MAX_ROUNDS; }
    if (trace && t <= trace_cap) {
        int64_t *r = trace + (int64_t)(t - 1) * T_NCOL;
        r[T_PREV_N] = nprev;
        r[T_PREV_DEG] = dprev;
        r[T_CAND] = ncand;
        r[T_BD] = (int64_t)*bd.cnt;
        r[T_B] = nb;
        r[T_BDEG] = bdeg;
        r[T_CHG] = nb == 0 ? 0 : nchg;
        r[T_CHGDEG] = nb == 0 ? 0 : cdeg;
        r[T_CUT] = cut_rec;
        r[T_BL] = (int64_t)cur.cnt[LC_LIGHT];
        r[T_BLDEG] = (int64_t)(cur.cnt[LC_DEG] - cur.cnt[LC_HDEG]);
        r[T_DL] = nb == 0 ? 0 : (int64_t)D.cnt[LC_LIGHT];
        r[T_DLDEG] = nb == 0 ? 0 : (int64_t)(D.cnt[LC_DEG] - D.cnt[LC_HDEG]);
    }
    if (!ctl->done) ctl->t = t + 1;
    else atomicSub(J.remaining, 1);
}

// Zero a job's per-run state before the initial sweep.
__global__ void k_init_zero(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    const int64_t n = J.g.n;
    for (int64_t v = vthread(vg); v < n; v += vthreads(vg)) J.diff[v] = 0;
    // Word labels (k above the byte range) tell changed neighbours apart by
    // chg >= 0; byte-label rounds leave chg unreset (k_apply), so start clean.
    if (J.k > kExact8)
        for (int64_t v = vthread(vg); v < n; v += vthreads(vg)) J.chg[v] = -1;
    for (int64_t i = vthread(vg); i < 2 * J.bwords; i += vthreads(vg)) J.bbits[i] = 0u;
    for (int64_t i = vthread(vg); i < J.bwords; i += vthreads(vg)) J.cbits[i] = 0u;
    const int64_t nhh = (int64_t)J.all.cnt[LC_SLOTS] * kLpCap;
    for (int64_t i = vthread(vg); i < nhh; i += vthreads(vg)) J.hh[i] = 0;
    for (int64_t i = vthread(vg); i < kHhCopies * nhh; i += vthreads(vg)) J.hhc[i] = 0;
    if (vg.block == 0) {
        uint32_t *c = reinterpret_cast<uint32_t *>(J.ctl);
        for (int i = threadIdx.x; i < (int)(sizeof(Ctl) / 4); i += blockDim.x) c[i] = 0;
        for (int i = threadIdx.x; i < J.k; i += blockDim.x) { J.vc[i] = 0; J.ic[i] = 0; }
    }
}

__global__ void k_set_remaining(int32_t *remaining, int32_t v) { *remaining = v; }

// After the selection: pack B_t (vertices stamped with this round's epoch)
// into the bitmap of half (t & 1), which the next round's pull candidates
// test.  One ballot per 64 vertices: coalesced stamp reads, two word stores.

// done flag of every job of a set, gathered for one host copy.
// Push rounds write record entries only for the vertices they touch: the
// rest of the dense record array is cleared first.
__global__ void k_rdeg_clear(const P2Job *JS, const int32_t *bstart, int jobs) {
    P2_JOB();
    if (is_done(J.ctl) || J.ctl->cand_mode != CAND_PUSH) return;
    for (int64_t v = vthread(vg); v < J.g.n; v += vthreads(vg)) J.rdeg[v] = -1;
}

// The light-vertex list of a boundary set from its bitmap (the selection
// writes bitmaps), for the kernels that walk the set as a list: the
// gathering LP (after the selection, shift 0: B_t) and push candidates
// (at round start, shift 1: B_{t-1}; only_push: skip unless this round
// pushes).  Heavy rows keep their slots from the selection.
__global__ void k_b_list(const P2Job *JS, const int32_t *bstart, int jobs, int shift, int only_push,
                         int heavy) {
    P2_JOB();
    const Ctl *ctl = J.ctl;
    if (is_done(ctl) || (only_push && ctl->cand_mode != CAND_PUSH)) return;
    const ListDev B = J.B;
    const int64_t bwords = J.bwords, n = J.g.n;
    // Push rounds start from clear record entries (the candidate kernels
    // write the records they find); one launch does both.
    if (only_push)
        for (int64_t v = vthread(vg); v < n; v += vthreads(vg)) J.rdeg[v] = -1;
    const uint32_t *half = J.bbits + (int64_t)((ctl->t - shift) & 1) * bwords;
    const int lane = lane_id();
    for (int64_t base = vwave(vg) * kWave; base < n; base += vwaves(vg) * kWave) {
        const int64_t wi = base >> 5;
        uint64_t m = half[wi] & ~J.hbits[wi], mh = half[wi] & J.hbits[wi];
        if (wi + 1 < bwords) {
            m |= (uint64_t)(half[wi + 1] & ~J.hbits[wi + 1]) << 32;
            mh |= (uint64_t)(half[wi + 1] & J.hbits[wi + 1]) << 32;
        }
        if (m) {  // wave-uniform
            const bool on = (m >> lane) & 1ull;
            const int64_t pos = wave_append(on, &B.cnt[LC_FILL]);
            if (on) B.light[pos] = (int32_t)(base + lane);
        }
        if (heavy && mh) {  // heavy members too (slots only, for a caller reading the list)
            const bool on = (mh >> lane) & 1ull;
            const int64_t pos = wave_append(on, &B.cnt[LC_HFILL]);
            if (on) B.hv[pos] = (int32_t)(base + lane);
        }
    }
}

// A boundary list given as a list (light ids and heavy slots) into half
// ((t - shift) & 1) of the bitmap, for the test hooks.
__global__ void k_list_bits(const P2Job *JS, const int32_t *bstart, int jobs, int shift) {
    P2_JOB();
    uint32_t *half = J.bbits + (int64_t)((J.ctl->t - shift) & 1) * J.bwords;
    for (int64_t i = vthread(vg); i < J.bwords; i += vthreads(vg)) half[i] = 0u;
}
__global__ void k_list_bits_set(const P2Job *JS, const int32_t *bstart, int jobs, int shift) {
    P2_JOB();
    const ListDev B = J.B;
    uint32_t *half = J.bbits + (int64_t)((J.ctl->t - shift) & 1) * J.bwords;
    const int64_t nl = (int64_t)B.cnt[LC_LIGHT], ns = (int64_t)B.cnt[LC_SLOTS];
    for (int64_t i = vthread(vg); i < nl + ns; i += vthreads(vg)) {
        const int32_t v = i < nl ? B.light[i] : B.hv[i - nl];
        atomicOr(&half[v >> 5], 1u << (v & 31));
    }
}

__global__ void k_p2_gather_done(const P2Job *JS, int jobs, int32_t *out) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < jobs; j += gridDim.x * blockDim.x)
        out[j] = gload(JS[j].ctl)->done;
}

}  // namespace

// ---------------------------------------------------------------- host side

Phase2Work &phase2_work(dmolp_ctx *ctx, dmolp_graph *g, int k) {
    DMOLP_REQUIRE(k >= 1 && k <= DMOLP_MAX_K, DMOLP_E_ARG, "k outside [1, 1024]");
    graph_prepare_bins(ctx, g);
    graph_prepare_first(ctx, g);
    if (!g->p2) g->p2.reset(new Phase2Work());
    Phase2Work &w = *g->p2;
    const int64_t n = g->n;
    if (!w.chg.p) {
        w.chg.alloc(sizeof(int32_t) * (size_t)n);
        w.diff.alloc(sizeof(int32_t) * (size_t)n);
        w.stamp.alloc(sizeof(int32_t) * (size_t)n);
        w.rdeg.alloc(sizeof(int32_t) * (size_t)n);
        w.wit.alloc(sizeof(int32_t) * (size_t)n);
        w.bbits.alloc(sizeof(uint32_t) * 2 * (size_t)((n + 31) / 32));
        w.L8.alloc((size_t)n);
        w.L5.alloc(sizeof(uint32_t) * (size_t)((n + 5) / 6));  // 6 labels a word at 5 bits, 8 at 4
        w.ON.alloc(sizeof(uint16_t) * (size_t)n);
        w.LN.alloc(sizeof(uint16_t) * (size_t)n);
        w.cbits.alloc(sizeof(uint32_t) * (size_t)((n + 31) / 32));
        DMOLP_HIP_CHECK(hipMemsetAsync(w.LN.p, 0xff, w.LN.bytes, ctx->stream));
        DMOLP_HIP_CHECK(hipMemsetAsync(w.cbits.p, 0, w.cbits.bytes, ctx->stream));
        w.bd.alloc(1);  // records are dense (rdeg): only bd's counter is used
        DMOLP_HIP_CHECK(hipMemsetAsync(w.stamp.p, 0, w.stamp.bytes, ctx->stream));
        DMOLP_HIP_CHECK(hipMemsetAsync(w.bbits.p, 0, w.bbits.bytes, ctx->stream));
        DMOLP_HIP_CHECK(hipMemsetAsync(w.chg.p, 0xff, w.chg.bytes, ctx->stream));
        w.ctl.alloc(sizeof(Ctl));
    }
    if (w.cap_heavy != g->n_heavy || w.cap_chunks != g->n_heavy_chunks) {
        w.B.alloc(n, g->n_heavy, g->n_heavy_chunks);
        w.D.alloc(n, g->n_heavy, g->n_heavy_chunks);
        w.hub_hist.alloc(sizeof(int32_t) * (size_t)(g->n_heavy > 0 ? g->n_heavy : 1) * kLpCap);
        w.hh.alloc(sizeof(int32_t) * (size_t)(g->n_heavy > 0 ? g->n_heavy : 1) * kLpCap);
        w.hhc.alloc(sizeof(int32_t) * (size_t)(g->n_heavy > 0 ? g->n_heavy : 1) * kLpCap * kHhCopies);
        w.deferred.alloc(sizeof(int64_t) * (size_t)(g->n_heavy > 0 ? g->n_heavy : 1));
        DMOLP_HIP_CHECK(hipMemsetAsync(w.hub_hist.p, 0, w.hub_hist.bytes, ctx->stream));
        DMOLP_HIP_CHECK(hipMemsetAsync(w.B.cnt.p, 0, w.B.cnt.bytes, ctx->stream));
        DMOLP_HIP_CHECK(hipMemsetAsync(w.D.cnt.p, 0, w.D.cnt.bytes, ctx->stream));
        w.cap_heavy = g->n_heavy;
        w.cap_chunks = g->n_heavy_chunks;
    }
    if (w.k != k) {
        w.k = k;
        w.vc.alloc(sizeof(unsigned long long) * (size_t)k);
        w.ic.alloc(sizeof(unsigned long long) * (size_t)k);
        w.pen.alloc(sizeof(double) * (size_t)k);
        w.re.alloc(sizeof(double) * (size_t)k);
    }
    if (ctx->tune.light_hist) w.lh.grow(sizeof(uint32_t) * kPackedWords * (size_t)(n > 0 ? n : 1));
    w.sel.prepare(ctx, k, n, g->max_degree);
    return w;
}

void phase2_round_table(dmolp_ctx *ctx, Phase2Work &w, int R) {
    if (w.rtab_rounds >= R) return;
    std::vector<double> tab(2 * ((size_t)R + 1), 0.0);
    for (int t = 1; t <= R; ++t) {
        tab[(size_t)t] = 1.0 + log((double)t);                 // phase2.cpp:199-200
        tab[(size_t)R + 1 + (size_t)t] = exp((-t) / (5.0));    // gpu_lp_boundary.cu:244
    }
    w.rtab.alloc(sizeof(double) * tab.size());
    DMOLP_HIP_CHECK(hipMemcpyAsync(w.rtab.p, tab.data(), w.rtab.bytes, hipMemcpyHostToDevice,
                                   ctx->stream));
    DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    w.rtab_rounds = R;
    w.round.key = 0;  // the captured round reads rtab: capture again
}

P2Job phase2_job(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w, int max_rounds, double epsilon,
                 int32_t *remaining) {
    P2Job j{};
    j.g = g->dev();
    j.all = g->all.dev;
    j.B = w.B.dev;
    j.D = w.D.dev;
    j.L = g->labels.as<int32_t>();
    j.L8 = w.L8.as<uint8_t>();
    j.L5 = w.L5.as<uint32_t>();
    j.lbits = w.k <= 16 ? 4 : (w.k <= kLpCap ? 5 : 0);
    j.chg = w.chg.as<int32_t>();
    j.diff = w.diff.as<int32_t>();
    j.stamp = w.stamp.as<int32_t>();
    j.rdeg = w.rdeg.as<int32_t>();
    j.bbits = w.bbits.as<uint32_t>();
    j.bwords = (g->n + 31) / 32;
    j.hh = w.hh.as<int32_t>();
    j.hhc = w.hhc.as<int32_t>();
    j.hid = g->hid.as<int32_t>();
    j.cbits = w.cbits.as<uint32_t>();
    j.ON = w.ON.as<uint16_t>();
    j.LN = w.LN.as<uint16_t>();
    j.lh_on = ctx->tune.light_hist ? 1 : 0;
    j.lh = j.lh_on ? w.lh.as<uint32_t>() : nullptr;
    j.hbits = g->hbits.as<uint32_t>();
    j.first = g->first.as<int32_t>();
    j.deg = g->deg.as<int32_t>();
    j.wit = w.wit.as<int32_t>();
    j.deferred = w.deferred.as<int64_t>();
    j.hub_hist = w.hub_hist.as<int32_t>();
    j.bd = w.bd.dev;
    j.ctl = w.ctl.as<Ctl>();
    j.vc = w.vc.as<unsigned long long>();
    j.ic = w.ic.as<unsigned long long>();
    j.pen = w.pen.as<double>();
    j.re = w.re.as<double>();
    j.rtab = w.rtab.as<double>();
    j.trace = w.trace.as<int64_t>();
    j.trace_cap = w.trace_cap;
    j.remaining = remaining;
    j.k = w.k;
    j.heavy_degree = ctx->tune.heavy_degree;
    j.chunk_edges = ctx->tune.chunk_edges;
    // Stamp values only grow, so stamps left by earlier runs never match.
    j.epoch_base = w.epoch;
    w.epoch += max_rounds + 2;
    j.max_rounds = max_rounds;
    j.epsilon = epsilon;
    j.pull_frac = ctx->tune.pull_fraction;
    j.recount_share = ctx->tune.hist_recount_share;
    j.push_share = ctx->tune.hist_push_share;
    return j;
}

SelArgs phase2_sel_args(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w) {
    Ctl *ctl = w.ctl.as<Ctl>();
    TakeRule rule{};
    rule.kind = 0;
    rule.ratio_tab = w.rtab.p ? w.rtab.as<double>() + w.rtab_rounds + 1 : nullptr;
    rule.t_ptr = &ctl->t;
    rule.re = w.re.as<double>();
    // Records are dense (one entry per vertex, labels read in place: the
    // byte mirror when it holds every label) and the selected set goes
    // straight to the B_t bitmap.  Records have valid labels: a run with
    // any label outside [0, k) stops after the initial sweep.
    SelArgs a = select_args(ctx, w.sel, w.bd.dev, rule, &ctl->done, w.B.dev, nullptr);
    a.fault = &ctl->v[C_KERNEL_ERR];
    select_args_dense(a, w.rdeg.as<int32_t>(), g->labels.as<int32_t>(),
                      w.k <= kExact8 ? w.L8.as<uint8_t>() : nullptr, g->n, w.bbits.as<uint32_t>(),
                      (g->n + 31) / 32);
    return a;
}

P2Set phase2_single_set(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w, int max_rounds,
                        double epsilon) {
    hipStream_t s = ctx->stream;
    if (!w.remaining.p) {
        w.remaining.alloc(sizeof(int32_t) * 2);
        w.job_dev.alloc(sizeof(P2Job));
        w.sel_dev.alloc(sizeof(SelArgs));
    }
    w.job_host = phase2_job(ctx, g, w, max_rounds, epsilon, w.remaining.as<int32_t>());
    w.sel_host = phase2_sel_args(ctx, g, w);
    DMOLP_HIP_CHECK(hipMemcpyAsync(w.job_dev.p, &w.job_host, sizeof(P2Job), hipMemcpyHostToDevice, s));
    DMOLP_HIP_CHECK(hipMemcpyAsync(w.sel_dev.p, &w.sel_host, sizeof(SelArgs), hipMemcpyHostToDevice, s));
    std::array<int32_t, P_NPOL> blk{};
    job_blocks(ctx, g->n, g->nnz, w.k, g->n_heavy, true, blk.data());
    bool same = w.ls.jobs == 1;
    for (int p = 0; same && p < P_NPOL; ++p) same = w.ls.total[p] == blk[(size_t)p];
    if (!same) launch_set_build(ctx, w.ls, {blk});
    P2Set set;
    set.jobs = w.job_dev.as<P2Job>();
    set.sel = w.sel_dev.as<SelArgs>();
    set.ls = &w.ls;
    set.passes = w.sel.passes;
    set.any_heavy = g->n_heavy > 0;
    set.lh_on = ctx->tune.light_hist;
    set.remaining = w.remaining.as<int32_t>();
    return set;
}

void phase2_done_flags(dmolp_ctx *ctx, const P2Set &set, int32_t *d_out, std::vector<int32_t> &out) {
    const int G = set.ls->jobs;
    out.assign((size_t)G, 0);
    hipLaunchKernelGGL(k_p2_gather_done, dim3((G + 255) / 256), dim3(256), 0, ctx->stream, set.jobs, G,
                       d_out);
    DMOLP_HIP_CHECK(hipMemcpyAsync(out.data(), d_out, sizeof(int32_t) * (size_t)G,
                                   hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void phase2_init_enqueue(dmolp_ctx *ctx, const P2Set &set) {
    hipStream_t s = ctx->stream;
    const LaunchSet &ls = *set.ls;
    const int J = ls.jobs;
    hipLaunchKernelGGL(k_set_remaining, dim3(1), dim3(1), 0, s, set.remaining, (int32_t)J);
    hipLaunchKernelGGL(k_init_zero, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, set.jobs,
                       ls.bs(P_GRID), J);
    hipLaunchKernelGGL(k_mirror8, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, set.jobs,
                       ls.bs(P_GRID), J);
    hipLaunchKernelGGL(k_pack_labels, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, set.jobs,
                       ls.bs(P_GRID), J);
    // The heavy rows' sweep runs on the side stream beside the light rows'
    // (disjoint rows; the counters both add to take atomic adds), and the
    // finish waits for both.
    if (set.any_heavy) {
        ctx->ensure_side();
        DMOLP_HIP_CHECK(hipEventRecord(ctx->ev_fork, s));
        DMOLP_HIP_CHECK(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
        hipLaunchKernelGGL(k_init_heavy, dim3(ls.total[P_GRID]), dim3(kThreads), 0, ctx->side,
                           set.jobs, ls.bs(P_GRID), J);
        DMOLP_HIP_CHECK(hipEventRecord(ctx->ev_join, ctx->side));
    }
    size_t i0 = 0;
    if (ctx->profiling) {
        DMOLP_HIP_CHECK(hipEventRecord(ctx->prof_event(), s));
        i0 = ctx->ev_used - 1;
    }
    hipLaunchKernelGGL(k_init_light, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, set.jobs,
                       ls.bs(P_GRID), J);
    if (ctx->profiling) {  // group 8: the initial light-row sweep, once per run
        DMOLP_HIP_CHECK(hipEventRecord(ctx->prof_event(), s));
        if (ctx->prof_marks.size() < 9) ctx->prof_marks.resize(9);
        ctx->prof_marks[8].push_back({i0, ctx->ev_used - 1});
    }
    if (set.any_heavy) DMOLP_HIP_CHECK(hipStreamWaitEvent(s, ctx->ev_join, 0));
    hipLaunchKernelGGL(k_init_finish, dim3(ls.total[P_ONE]), dim3(256), 0, s, set.jobs, ls.bs(P_ONE),
                       J);
    DMOLP_HIP_CHECK(hipGetLastError());
}

void phase2_round_enqueue(dmolp_ctx *ctx, const P2Set &set) {
    hipStream_t s = ctx->stream;
    const LaunchSet &ls = *set.ls;
    const int J = ls.jobs;
    const P2Job *jobs = set.jobs;
    const unsigned grid = (unsigned)ls.total[P_GRID];
    const int32_t *bg = ls.bs(P_GRID);

    auto mark = [&]() -> size_t {
        if (!ctx->profiling) return 0;
        hipEvent_t e = ctx->prof_event();
        DMOLP_HIP_CHECK(hipEventRecord(e, s));
        return ctx->ev_used - 1;
    };
    const size_t m0 = mark();
    hipLaunchKernelGGL(k_round_begin, dim3(ls.total[P_ONE]), dim3(64), 0, s, jobs, ls.bs(P_ONE), J);
    // Round 1 takes every vertex; later rounds pull or push (ctl->cand_mode).
    // Each writes one dense record entry per vertex (push: after a clear).
    hipLaunchKernelGGL(k_cand_pull, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J);  // and round 1's
    hipLaunchKernelGGL(k_cand_pull_rest, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J);
    if (set.any_heavy)
        hipLaunchKernelGGL(k_cand_pull_heavy, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J);
    hipLaunchKernelGGL(k_b_list, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J, 1, 1, 0);  // B_{t-1} list
    hipLaunchKernelGGL(k_cand_light, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J);
    if (set.any_heavy)
        hipLaunchKernelGGL(k_cand_heavy, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J);
    const size_t m1 = mark();
    // B_t: the bitmap and counts.  With kept histograms the selection's last
    // pass is fused with the light LP below.
    select_enqueue(ctx, set.sel, ls, set.passes, !set.lh_on, true);
    if (!set.lh_on)  // the gathering LP walks B_t as a list
        hipLaunchKernelGGL(k_b_list, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J, 0, 0, 0);
    const size_t m2 = mark();
    if (set.lh_on)
        hipLaunchKernelGGL(k_lp_light_take, dim3(grid), dim3(kThreads), 0, s, jobs, set.sel, bg, J);
    else
        hipLaunchKernelGGL(k_lp_light, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J);
    const size_t m3 = mark();
    if (set.any_heavy) {  // heavy rows decide from their kept histograms
        hipLaunchKernelGGL(k_lp_heavy_hist, dim3(ls.total[P_HEAVY]), dim3(kThreads), 0, s, jobs,
                           ls.bs(P_HEAVY), J);
        // How the heavy histograms follow this round's changes (push mode
        // moves their counts in the Delta pass).
        hipLaunchKernelGGL(k_hist_mode, dim3(ls.total[P_ONE]), dim3(1), 0, s, jobs, ls.bs(P_ONE), J);
    }
    const size_t m4 = mark();
    hipLaunchKernelGGL(k_delta_light, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J);
    const size_t m4b = mark();
    if (set.any_heavy)
        hipLaunchKernelGGL(k_delta_heavy, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J);
    hipLaunchKernelGGL(k_apply, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J, 1, set.any_heavy ? 1 : 0);
    const size_t m5 = mark();
    if (set.any_heavy) {  // recount, or fold the pushed moves or scan (one launch)
        hipLaunchKernelGGL(k_hist_recount, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J);
        hipLaunchKernelGGL(k_hist_update, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J);
    }
    hipLaunchKernelGGL(k_cbits_clear, dim3(grid), dim3(kThreads), 0, s, jobs, bg, J, 1, set.any_heavy ? 1 : 0);
    const size_t m6 = mark();
    hipLaunchKernelGGL(k_round_end, dim3(ls.total[P_ONE]), dim3(1), 0, s, jobs, ls.bs(P_ONE), J);
    DMOLP_HIP_CHECK(hipGetLastError());
    if (ctx->profiling) {
        const size_t m7 = mark();
        // Groups: 0 lp_light, 1 lp_heavy, 2 candidates, 3 select, 4 delta+apply,
        // 5 round, 6 heavy-row histogram update, 7 k_delta_light alone.
        const std::pair<size_t, size_t> spans[8] = {{m2, m3}, {m3, m4}, {m0, m1}, {m1, m2},
                                                    {m4, m5}, {m0, m7}, {m5, m6}, {m4, m4b}};
        if (ctx->prof_marks.size() < 9) ctx->prof_marks.resize(9);
        for (int i = 0; i < 8; ++i) ctx->prof_marks[(size_t)i].push_back(spans[i]);
    }
}

// Device penalties for the test hook.
void phase2_penalties_device(dmolp_ctx *ctx, int k, const unsigned long long *d_vc,
                             const unsigned long long *d_ic, int t, double *d_pen, double *d_re) {
    const double lg = 1.0 + log((double)t);
    hipLaunchKernelGGL(k_penalties, dim3(1), dim3(64), 0, ctx->stream, d_vc, d_ic, k, lg, d_pen,
                       d_re);
    DMOLP_HIP_CHECK(hipGetLastError());
}

namespace {

uint64_t mix_key(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h;
}

__global__ void k_hook_state(Ctl *ctl, int32_t t, int32_t epoch, int32_t mode) {
    ctl->done = 0;
    ctl->diff_ok = 0;
    ctl->t = t;
    ctl->epoch = epoch;
    ctl->cand_mode = mode;
}

}  // namespace

uint64_t phase2_single_key(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w) {
    // Everything the captured launches bake in.
    uint64_t key = 0x2545f4914f6cdd1dull;
    const void *ptrs[] = {w.job_dev.p, w.sel_dev.p, w.ls.bstart.p, w.remaining.p, g->labels.p};
    for (const void *p : ptrs) key = mix_key(key, (uint64_t)(uintptr_t)p);
    int64_t ints[4 + P_NPOL] = {w.sel.passes, g->n_heavy > 0 ? 1 : 0, w.ls.jobs,
                                ctx->tune.light_hist ? 1 : 0};
    for (int p = 0; p < P_NPOL; ++p) ints[4 + p] = w.ls.total[p];
    for (int64_t v : ints) key = mix_key(key, (uint64_t)v);
    return key;
}

void phase2_round_launch(dmolp_ctx *ctx, const P2Set &set, RoundGraph &rg, uint64_t key) {
    if (ctx->profiling || !graphs_enabled()) {  // events between kernels, or no graphs: launch directly
        phase2_round_enqueue(ctx, set);
        return;
    }
    if (!rg.exec || rg.key != key) {
        if (rg.exec || rg.graph) {
            // Launches of the old executable may still be queued: let the
            // stream drain before it is destroyed.
            DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            rg.release();
        }
        hipGraph_t graph = nullptr;
        DMOLP_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            phase2_round_enqueue(ctx, set);
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

// Test hook: the candidate + selection steps of round t from an explicit
// previous boundary already uploaded into w.B, with re on the device (the
// candidates are pushed from B's rows).  B_t is left in w.B.
void phase2_select_enqueue(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w, int t) {
    hipStream_t s = ctx->stream;
    Ctl *ctl = w.ctl.as<Ctl>();
    phase2_round_table(ctx, w, t > w.rtab_rounds ? t : w.rtab_rounds);
    const P2Set set = phase2_single_set(ctx, g, w, t, 0.0);
    const LaunchSet &ls = *set.ls;
    hipLaunchKernelGGL(k_hook_state, dim3(1), dim3(1), 0, s, ctl, t, w.epoch, (int32_t)CAND_PUSH);
    DMOLP_HIP_CHECK(hipMemsetAsync(w.bd.cnt.p, 0, w.bd.cnt.bytes, s));
    // B_{t-1} arrives as a list: its bitmap for the heavy-row push.
    hipLaunchKernelGGL(k_list_bits, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, set.jobs,
                       ls.bs(P_GRID), ls.jobs, 1);
    hipLaunchKernelGGL(k_list_bits_set, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, set.jobs,
                       ls.bs(P_GRID), ls.jobs, 1);
    hipLaunchKernelGGL(k_rdeg_clear, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, set.jobs,
                       ls.bs(P_GRID), ls.jobs);
    hipLaunchKernelGGL(k_cand_light, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, set.jobs,
                       ls.bs(P_GRID), ls.jobs);
    if (set.any_heavy)
        hipLaunchKernelGGL(k_cand_heavy, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, set.jobs,
                           ls.bs(P_GRID), ls.jobs);
    select_enqueue(ctx, set.sel, ls, set.passes, true, true);
    hipLaunchKernelGGL(k_b_list, dim3(ls.total[P_GRID]), dim3(kThreads), 0, s, set.jobs,
                       ls.bs(P_GRID), ls.jobs, 0, 0, 1);
    DMOLP_HIP_CHECK(hipGetLastError());
}

// Test hook: decisions for the list in w.B; Delta applied to g's labels.
void phase2_lp_enqueue(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w) {
    hipStream_t s = ctx->stream;
    const P2Set set = phase2_single_set(ctx, g, w, 1, 0.0);
    const LaunchSet &ls = *set.ls;
    const unsigned grid = (unsigned)ls.total[P_GRID];
    const int32_t *bg = ls.bs(P_GRID);
    DMOLP_HIP_CHECK(hipMemsetAsync(w.D.cnt.p, 0, w.D.cnt.bytes, s));
    hipLaunchKernelGGL(k_mirror8, dim3(grid), dim3(kThreads), 0, s, set.jobs, bg, ls.jobs);
    hipLaunchKernelGGL(k_lp_light, dim3(grid), dim3(kThreads), 0, s, set.jobs, bg, ls.jobs);
    if (set.any_heavy) {
        hipLaunchKernelGGL(k_lp_heavy, dim3(grid), dim3(kThreads), 0, s, set.jobs, bg, ls.jobs);
        hipLaunchKernelGGL(k_lp_heavy_decide, dim3(ls.total[P_HEAVY]), dim3(kThreads), 0, s,
                           set.jobs, ls.bs(P_HEAVY), ls.jobs);
    }
    hipLaunchKernelGGL(k_apply, dim3(grid), dim3(kThreads), 0, s, set.jobs, bg, ls.jobs, 0, 0);
    hipLaunchKernelGGL(k_cbits_clear, dim3(grid), dim3(kThreads), 0, s, set.jobs, bg, ls.jobs, 0, 0);
    DMOLP_HIP_CHECK(hipGetLastError());
}

}  // namespace dmolp
