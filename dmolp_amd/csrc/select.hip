// select.hip -- per-group top-m selection by (degree, id), see select.hpp.
#include <cmath>

#include "rows.hpp"
#include "select.hpp"

namespace dmolp {
namespace {

constexpr int kSelThreads = 256;
constexpr int kHist1LdsInts = 8192;  // 32 KB of LDS for the level-1 histogram
constexpr int kCompactItems = 8;     // records per thread per block iteration
constexpr int kHist2LdsInts = 16384; // 64 KB of LDS for the digit histograms

// Running count over bins in scan order (reverse: from the last bin down);
// finds the first bin where it reaches target (1 <= target <= total).
// Returns through LDS variables; all threads must call.
__device__ void block_crossing(const int32_t *bins, int nb, bool reverse, int64_t target,
                               int *s_bin, int64_t *s_before, int64_t *s_tmp) {
    const int t = threadIdx.x, T = blockDim.x;
    const int per = (nb + T - 1) / T;
    const int j0 = t * per, j1 = min(nb, j0 + per);
    int64_t sum = 0;
    for (int j = j0; j < j1; ++j) sum += bins[reverse ? nb - 1 - j : j];
    s_tmp[t] = sum;
    __syncthreads();
    for (int o = 1; o < T; o <<= 1) {
        int64_t add = t >= o ? s_tmp[t - o] : 0;
        __syncthreads();
        s_tmp[t] += add;
        __syncthreads();
    }
    const int64_t incl = s_tmp[t], excl = incl - sum;
    if (excl < target && target <= incl) {
        int64_t run = excl;
        for (int j = j0; j < j1; ++j) {
            const int b = reverse ? nb - 1 - j : j;
            const int64_t c = bins[b];
            if (run + c >= target) { *s_bin = b; *s_before = run; break; }
            run += c;
        }
    }
    __syncthreads();
}

// Kernels take the job set (one SelArgs per job) and run on the job's
// share of the launch.
#define SEL_JOB()                                      \
    const int job_ = job_of_block(bstart, jobs);       \
    const VGrid vg = vgrid_of(bstart, job_);           \
    const SelArgs a = gload(A[job_]);                  \
    (void)vg

__global__ __launch_bounds__(kSelThreads) void k_sel_hist1(const SelArgs *A, const int32_t *bstart,
                                                           int jobs) {
    SEL_JOB();
    if (a.done && *a.done) return;
    __shared__ int32_t s_h[kHist1LdsInts];
    const RecDev in = a.in;
    int32_t *h1 = a.h1;
    const int b1_bins = a.b1_bins;
    const int nb = a.groups * b1_bins;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s_h[i] = 0;
    if (vg.block == 0 && threadIdx.x == 0) a.refining[1] = 0;  // rank-cut groups of this selection
    __syncthreads();
    const bool dense = a.dense != 0;
    const int64_t count = dense ? a.rn : (int64_t)*in.cnt;
    constexpr int kH1Items = 8;  // independent record loads in flight per thread
    const int64_t step = vthreads(vg);
    // M: 0 listed records, 1 dense with int32 groups, 2 dense with group bytes.
    auto run = [&](auto mode) {
        constexpr int M = decltype(mode)::value;
        for (int64_t i0 = vthread(vg); i0 < count; i0 += kH1Items * step) {
            int32_t gg[kH1Items], dd[kH1Items];
            if (M == 0) {
#pragma unroll
                for (int j = 0; j < kH1Items; ++j) {
                    const int64_t i = i0 + j * step;
                    const int64_t ic = i < count ? i : count - 1;  // loads issued together
                    const int32_t g = in.g[ic];
                    dd[j] = in.d[ic];
                    gg[j] = i < count ? g : -1;
                }
            } else {
                dense_load<M == 2>(a, i0, step, dd, gg);
#pragma unroll
                for (int j = 0; j < kH1Items; ++j)
                    if (dd[j] < 0) gg[j] = -1;
            }
#pragma unroll
            for (int j = 0; j < kH1Items; ++j)
                if (gg[j] >= 0) atomicAdd(&s_h[gg[j] * b1_bins + bin1_of(dd[j], b1_bins)], 1);
        }
    };
    if (!dense) run(std::integral_constant<int, 0>{});
    else if (a.rlab8) run(std::integral_constant<int, 2>{});
    else run(std::integral_constant<int, 1>{});
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if (s_h[i]) atomicAdd(&h1[i], s_h[i]);
}

__global__ __launch_bounds__(kSelThreads) void k_sel_pick1(const SelArgs *A, const int32_t *bstart,
                                                           int jobs) {
    SEL_JOB();
    if (a.done && *a.done) return;
    const ListDev sel = a.sel;
    if (vg.block == 0 && threadIdx.x < LC_N) sel.cnt[threadIdx.x] = 0;
    if (vg.block == 0 && threadIdx.x == 0) *a.ref.cnt = 0;
    __shared__ int64_t s_tmp[kSelThreads];
    __shared__ int s_bin;
    __shared__ int64_t s_before;
    const int g = (int)vg.block;
    const int b1_bins = a.b1_bins, id_bits = a.id_bits, deg_bits = a.deg_bits;
    const TakeRule rule = a.rule;
    int32_t *bins = a.h1 + (int64_t)g * b1_bins;
    // Group size.
    int64_t part = 0;
    for (int j = threadIdx.x; j < b1_bins; j += blockDim.x) part += bins[j];
    s_tmp[threadIdx.x] = part;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s_tmp[threadIdx.x] += s_tmp[threadIdx.x + o];
        __syncthreads();
    }
    const int64_t c = s_tmp[0];
    __syncthreads();
    int64_t m;
    int desc;
    if (rule.kind == 0) {
        // gpu_lp_boundary.cu:282,296-297.
        desc = rule.re[g] > 1.0 ? 1 : 0;
        const double ratio = rule.ratio_tab[*rule.t_ptr];
        m = c > 0 ? (int64_t)(int32_t)ceil(ratio * (double)c) : 0;
        if (m > c) m = c;
    } else {
        // partition.cpp:426-430.
        desc = 0;
        m = c < (int64_t)rule.theta ? c : (int64_t)rule.theta;
    }
    SelState s{};
    s.desc = desc;
    s.take = m;
    s.count = c;
    if (m <= 0) {
        s.mode = SEL_NONE;
    } else if (m == c) {
        s.mode = SEL_ALL;
    } else {
        block_crossing(bins, b1_bins, desc != 0, m, &s_bin, &s_before, s_tmp);
        const int b = s_bin;
        const int64_t before = s_before;
        s.b1 = b;
        if (before + bins[b] == m) {
            s.mode = SEL_THROUGH;
        } else if (a.dense && b < b1_bins - 1 && g < kSelLdsGroups) {
            // Dense records, cut inside one exact degree: the first `need`
            // of that degree by id, found by counts per id chunk.
            s.mode = SEL_RANK;
            s.need = m - before;
            if (threadIdx.x == 0) atomicAdd(&a.refining[1], 1);
        } else {
            s.mode = SEL_REFINE;
            s.need = m - before;
            if (b < b1_bins - 1) {
                const uint64_t dmax = (1ull << deg_bits) - 1;
                const uint64_t dk = desc ? dmax - (uint64_t)b : (uint64_t)b;
                s.prefix = dk << id_bits;
                s.r = id_bits;
            } else {
                s.prefix = 0;
                s.r = id_bits + deg_bits;
            }
            if (threadIdx.x == 0) atomicAdd(a.refining, 1);
        }
    }
    if (threadIdx.x == 0) a.st[g] = s;
    __syncthreads();
    for (int j = threadIdx.x; j < b1_bins; j += blockDim.x) bins[j] = 0;
}

// The records the digit passes still have to look at: those in the chosen
// level-1 bin of a group whose cut falls inside that bin.  Usually a small
// share of all records, so the passes below scan this list only.
__global__ __launch_bounds__(kSelThreads) void k_sel_collect(const SelArgs *A, const int32_t *bstart,
                                                             int jobs) {
    SEL_JOB();
    if ((a.done && *a.done) || a.dense) return;
    if (*a.refining == 0) return;
    __shared__ BlockAppendLds s_app;
    __shared__ SelLite s_st[kSelLdsGroups];
    const RecDev in = a.in, ref = a.ref;
    const SelState *st = a.st;
    sel_lite_load(s_st, st, a.groups);
    const int b1_bins = a.b1_bins;
    const int64_t count = (int64_t)*in.cnt;
    const int64_t span = (int64_t)blockDim.x * kCompactItems;
    for (int64_t b0 = vg.block * span; b0 < count; b0 += vg.blocks * span) {
        int32_t x[kCompactItems], gg[kCompactItems], d[kCompactItems];
        uint32_t keep_mask = 0;
        int32_t nk = 0;
#pragma unroll
        for (int j = 0; j < kCompactItems; ++j) {
            const int64_t i = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            x[j] = 0; gg[j] = 0; d[j] = 0;
            if (i >= count) continue;
            gg[j] = in.g[i];
            d[j] = in.d[i];
            const SelLite s = sel_get(s_st, st, gg[j]);
            if (s.mode == SEL_REFINE && bin1_of(d[j], b1_bins) == s.b1) {
                x[j] = in.x[i];
                keep_mask |= 1u << j;
                ++nk;
            }
        }
        int64_t pos = block_reserve(nk, ref.cnt, s_app);
#pragma unroll
        for (int j = 0; j < kCompactItems; ++j) {
            if (!(keep_mask & (1u << j))) continue;
            ref.x[pos] = x[j];
            ref.g[pos] = gg[j];
            ref.d[pos] = d[j];
            ++pos;
        }
    }
}

// Dense records: the refine list is staged in LDS (ids only) and flushed
// with one reservation per kRefStage - span or more entries; the flush
// re-reads the staged records' group and degree (recently read lines).
constexpr int kRefStage = 8192;
__device__ __forceinline__ void ref_flush(const SelArgs &a, const int32_t *s_x, int32_t &s_n,
                                          unsigned long long &s_base) {
    const int32_t cnt = s_n;
    if (threadIdx.x == 0) s_base = atomicAdd(a.ref.cnt, (unsigned long long)cnt);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const int32_t x = s_x[i];
        const unsigned long long p = s_base + (unsigned long long)i;
        a.ref.x[p] = x;
        a.ref.g[p] = dense_group(a, x);
        a.ref.d[p] = a.rdeg[x];
    }
    __syncthreads();
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
}
__global__ __launch_bounds__(kSelThreads) void k_sel_collect_dense(const SelArgs *A,
                                                                   const int32_t *bstart, int jobs) {
    SEL_JOB();
    if ((a.done && *a.done) || !a.dense) return;
    if (*a.refining == 0) return;
    __shared__ SelLite s_st[kSelLdsGroups];
    __shared__ int32_t s_x[kRefStage];
    __shared__ int32_t s_n;
    __shared__ unsigned long long s_base;
    if (threadIdx.x == 0) s_n = 0;
    const SelState *st = a.st;
    sel_lite_load(s_st, st, a.groups);
    const int b1_bins = a.b1_bins;
    const int64_t count = a.rn;
    const int64_t span = (int64_t)blockDim.x * kCompactItems;
    auto run = [&](auto byte) {
        for (int64_t b0 = vg.block * span; b0 < count; b0 += vg.blocks * span) {
            int32_t d[kCompactItems], gg[kCompactItems];
            dense_load<decltype(byte)::value>(a, b0 + threadIdx.x, blockDim.x, d, gg);
#pragma unroll
            for (int j = 0; j < kCompactItems; ++j) {
                if (d[j] < 0) continue;
                const SelLite s = sel_get(s_st, st, gg[j]);
                if (s.mode == SEL_REFINE && bin1_of(d[j], b1_bins) == s.b1)
                    s_x[atomicAdd(&s_n, 1)] = (int32_t)(b0 + (int64_t)j * blockDim.x + threadIdx.x);
            }
            __syncthreads();
            if (s_n > kRefStage - (int)span) ref_flush(a, s_x, s_n, s_base);
        }
    };
    if (a.rlab8) run(std::true_type{});
    else run(std::false_type{});
    __syncthreads();
    if (s_n > 0) ref_flush(a, s_x, s_n, s_base);
}

// Rank cuts, pass 1: per 1024-id chunk, the count of each rank group's
// records in its cut bin (b1, an exact degree), and the chunk's cut marks
// cleared.  One block iteration per chunk, in the dense passes' layout.
__global__ __launch_bounds__(kSelThreads) void k_sel_rank_count(const SelArgs *A, const int32_t *bstart,
                                                                int jobs) {
    SEL_JOB();
    if ((a.done && *a.done) || !a.dense) return;
    if (a.refining[1] == 0) return;
    __shared__ SelLite s_st[kSelLdsGroups];
    __shared__ int32_t s_c[kSelLdsGroups];
    sel_lite_load(s_st, a.st, a.groups);  // ends with a barrier
    const int G = a.groups < kSelLdsGroups ? a.groups : kSelLdsGroups;
    const int w = threadIdx.x / kWave, lane = lane_id();
    const int64_t nc = a.rank_chunks;
    for (int64_t c = vg.block; c < nc; c += vg.blocks) {
        if ((int)threadIdx.x < G) s_c[threadIdx.x] = 0;
        __syncthreads();
        int32_t d[kDenseItems], gg[kDenseItems];
        const int64_t i0 = (c << kRankChunkLog) + (int64_t)w * kDenseItems * kWave + lane;
        if (a.rlab8) dense_load<true>(a, i0, kWave, d, gg);
        else dense_load<false>(a, i0, kWave, d, gg);
#pragma unroll
        for (int j = 0; j < kDenseItems; ++j) {
            if (d[j] < 0 || gg[j] >= G) continue;
            const SelLite s = s_st[gg[j]];
            if (s.mode == SEL_RANK && d[j] == s.b1) atomicAdd(&s_c[gg[j]], 1);
        }
        __syncthreads();
        if ((int)threadIdx.x < G) a.rank_cnt[(int64_t)threadIdx.x * rank_stride(nc) + c] = s_c[threadIdx.x];
        if (threadIdx.x == 0) a.rank_mark[c] = 0ull;
        __syncthreads();
    }
}

// Rank cuts, pass 2, one block per group: the chunk where the running count
// of the group's cut-bin records reaches its need, and the need left inside
// that chunk; the chunk gets the group's mark.
__global__ __launch_bounds__(kSelThreads) void k_sel_rank_scan(const SelArgs *A, const int32_t *bstart,
                                                               int jobs) {
    SEL_JOB();
    if ((a.done && *a.done) || !a.dense) return;
    const int g = (int)vg.block;
    if (g >= kSelLdsGroups || a.st[g].mode != SEL_RANK) return;  // uniform per block
    __shared__ int64_t s_sum[kSelThreads];
    const int64_t nc = a.rank_chunks;
    const int32_t *cnt = a.rank_cnt + (int64_t)g * rank_stride(nc);  // 16-byte aligned rows
    // Each thread's contiguous range of chunks, a multiple of 4 long, read
    // as 16-byte vectors, four in flight.
    const int64_t per = ((nc + blockDim.x - 1) / blockDim.x + 3) & ~3ll;
    const int64_t c0 = (int64_t)threadIdx.x * per;
    const int64_t c1 = c0 + per < nc ? c0 + per : nc;
    int64_t sum = 0;
    int64_t c = c0;
    for (; c + 16 <= c1; c += 16) {
        int4 x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = *reinterpret_cast<const int4 *>(cnt + c + 4 * q);
#pragma unroll
        for (int q = 0; q < 4; ++q) sum += (int64_t)x[q].x + x[q].y + x[q].z + x[q].w;
    }
    for (; c < c1; ++c) sum += cnt[c];
    s_sum[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {  // exclusive prefix of the threads' ranges, in chunk order
        int64_t run = 0;
        for (int t = 0; t < (int)blockDim.x; ++t) {
            const int64_t x = s_sum[t];
            s_sum[t] = run;
            run += x;
        }
    }
    __syncthreads();
    const int64_t need = a.st[g].need;
    __shared__ int s_owner;
    __shared__ int64_t s_run;
    if (threadIdx.x == 0) s_owner = -1;
    __syncthreads();
    const int64_t mine = s_sum[threadIdx.x];
    if (mine < need && need <= mine + sum) {  // exactly one thread: its range holds the cut
        s_owner = (int)threadIdx.x;
        s_run = mine;
    }
    __syncthreads();
    // The counts sum to the cut bin's size, which holds the need; a group
    // without an owner (no such chunk) is a counting fault: report it (the
    // run stops with DMOLP_E_INTERNAL at the round's end) and read nothing.
    if (s_owner < 0) {  // uniform
        if (threadIdx.x == 0 && a.fault) atomicAdd(a.fault, 1ull);
        return;
    }
    // The owner's range, walked by the whole block kSelThreads chunks a step
    // (an inclusive scan per step) instead of one chunk a load.
    const int64_t o0 = (int64_t)s_owner * per;
    const int64_t o1 = o0 + per < nc ? o0 + per : nc;
    int64_t run = s_run;
    for (int64_t b = o0; b < o1; b += blockDim.x) {  // uniform trip count
        const int64_t c = b + threadIdx.x;
        const int64_t x = c < o1 ? cnt[c] : 0;
        s_sum[threadIdx.x] = x;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            const int64_t y = (int)threadIdx.x >= o ? s_sum[threadIdx.x - o] : 0;
            __syncthreads();
            s_sum[threadIdx.x] += y;
            __syncthreads();
        }
        const int64_t incl = run + s_sum[threadIdx.x];
        if (c < o1 && incl - x < need && need <= incl) {  // the first chunk reaching the need
            a.st[g].prefix = (uint64_t)c;
            a.st[g].need = need - (incl - x);
            atomicOr(&a.rank_mark[c], 1ull << g);
        }
        run += s_sum[blockDim.x - 1];
        __syncthreads();
        if (run >= need) break;  // uniform
    }
}

// Digit histogram of the records still being refined.  The block
// accumulates in LDS and adds its counts once: the digit width is chosen so
// that every group's bins fit (adding every record to the global histogram
// instead serialises on the hot bins: the frontier records of one part often
// share their top id digits).
__global__ __launch_bounds__(kSelThreads) void k_sel_hist2(const SelArgs *A, const int32_t *bstart,
                                                           int jobs) {
    SEL_JOB();
    if (a.done && *a.done) return;
    if (*a.refining == 0) return;
    const RecDev in = a.ref;
    int32_t *h2 = a.h2;
    const SelState *st = a.st;
    const int b1_bins = a.b1_bins, id_bits = a.id_bits, deg_bits = a.deg_bits;
    const int64_t count = (int64_t)*in.cnt;
    const int db = a.digit_bits;
    const int nb = a.groups << db;
    const bool lds = nb <= kHist2LdsInts;
    constexpr int kH2Items = 4;  // records per thread per step, loads issued together
    // Only as many blocks as the refine list needs (at least kH2MinRecs
    // records a block) clear and flush an LDS histogram; the launch is sized
    // for the job's whole frontier, usually far more than the refine list.
    constexpr int64_t kH2MinRecs = 8 * kSelThreads * kH2Items;
    int64_t eff = (count + kH2MinRecs - 1) / kH2MinRecs;
    eff = eff < 1 ? 1 : (eff > vg.blocks ? vg.blocks : eff);
    if (vg.block >= eff) return;  // uniform per block
    __shared__ int32_t s_h[kHist2LdsInts];
    __shared__ SelState s_st[kSelLdsGroups];  // the first groups' states
    for (int i = threadIdx.x; i < a.groups && i < kSelLdsGroups; i += blockDim.x) s_st[i] = st[i];
    if (lds)
        for (int i = threadIdx.x; i < nb; i += blockDim.x) s_h[i] = 0;
    __syncthreads();
    const int64_t step = eff * (int64_t)blockDim.x;
    for (int64_t i0 = vg.block * (int64_t)blockDim.x + threadIdx.x; i0 < count; i0 += kH2Items * step) {
        int32_t gg[kH2Items], dd[kH2Items], xx[kH2Items];
#pragma unroll
        for (int j = 0; j < kH2Items; ++j) {
            const int64_t i = i0 + j * step;
            const int64_t ic = i < count ? i : count - 1;  // in bounds: loads issue unconditionally
            gg[j] = in.g[ic];
            dd[j] = in.d[ic];
            xx[j] = in.x[ic];
        }
#pragma unroll
        for (int j = 0; j < kH2Items; ++j) {
            if (i0 + j * step >= count) continue;
            const SelState s = gg[j] < kSelLdsGroups ? s_st[gg[j]] : st[gg[j]];
            if (s.mode != SEL_REFINE) continue;
            const int32_t d = dd[j];
            if (bin1_of(d, b1_bins) != s.b1) continue;
            const uint64_t K = sel_key(xx[j], d, s.desc, id_bits, deg_bits);
            if ((K >> s.r) != (s.prefix >> s.r)) continue;
            const int w = s.r < db ? s.r : db;
            const int digit = (int)((K >> (s.r - w)) & ((1ull << w) - 1));
            if (lds) atomicAdd(&s_h[(gg[j] << db) + digit], 1);
            else atomicAdd(&h2[((int64_t)gg[j] << db) + digit], 1);
        }
    }
    if (lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < nb; i += blockDim.x)
            if (s_h[i]) atomicAdd(&h2[i], s_h[i]);
    }
}

__global__ __launch_bounds__(kSelThreads) void k_sel_pick2(const SelArgs *A, const int32_t *bstart,
                                                           int jobs) {
    SEL_JOB();
    if (a.done && *a.done) return;
    __shared__ int64_t s_tmp[kSelThreads];
    __shared__ int s_bin;
    __shared__ int64_t s_before;
    const int g = (int)vg.block;
    SelState s = a.st[g];
    if (s.mode != SEL_REFINE) return;  // uniform per block
    const int w = s.r < a.digit_bits ? s.r : a.digit_bits;
    const int nb = 1 << w;
    int32_t *bins = a.h2 + ((int64_t)g << a.digit_bits);
    block_crossing(bins, nb, false, s.need, &s_bin, &s_before, s_tmp);
    const int digit = s_bin;
    const int64_t cnt = bins[digit];
    s.need -= s_before;
    s.prefix |= (uint64_t)digit << (s.r - w);
    s.r -= w;
    if (s.need == cnt) {
        s.mode = SEL_KEY;
        if (threadIdx.x == 0) atomicSub(a.refining, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) a.st[g] = s;
    for (int j = threadIdx.x; j < nb; j += blockDim.x) bins[j] = 0;
}

__global__ __launch_bounds__(kSelThreads) void k_sel_compact(const SelArgs *A, const int32_t *bstart,
                                                             int jobs) {
    SEL_JOB();
    if ((a.done && *a.done) || a.dense) return;
    int32_t *mark = a.mark;
    const int32_t mark_val = mark ? *a.mark_val : 0;
    const RecDev in = a.in, carry = a.carry;
    const ListDev sel = a.sel;
    const SelState *st = a.st;
    const int b1_bins = a.b1_bins, id_bits = a.id_bits, deg_bits = a.deg_bits;
    const int heavy_degree = a.heavy_degree, chunk_edges = a.chunk_edges;
    const int has_carry = a.has_carry;
    __shared__ BlockAppendLds s_app;
    __shared__ SelLite s_st[kSelLdsGroups];
    sel_lite_load(s_st, st, a.groups);
    const int64_t count = (int64_t)*in.cnt;
    const int64_t span = (int64_t)blockDim.x * kCompactItems;
    // Block-uniform loop; each thread takes kCompactItems records (strided
    // by the block size, so loads stay coalesced) and reserves its output
    // slots with one block-level atomic per iteration.
    for (int64_t b0 = vg.block * span; b0 < count; b0 += vg.blocks * span) {
        int32_t x[kCompactItems], gg[kCompactItems], d[kCompactItems];
        uint32_t take_mask = 0, keep_mask = 0;
        int32_t n_light = 0, n_keep = 0;
        int64_t deg_sum = 0;
#pragma unroll
        for (int j = 0; j < kCompactItems; ++j) {
            const int64_t i = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            x[j] = 0; gg[j] = 0; d[j] = 0;
            if (i >= count) continue;
            x[j] = in.x[i];
            gg[j] = in.g[i];
            d[j] = in.d[i];
            const SelLite s = sel_get(s_st, st, gg[j]);
            const int b = bin1_of(d[j], b1_bins);
            const bool earlier = s.desc ? b > s.b1 : b < s.b1;
            bool take = false;
            switch (s.mode) {
                case SEL_ALL: take = true; break;
                case SEL_THROUGH: take = earlier || b == s.b1; break;
                case SEL_KEY: {
                    if (earlier) { take = true; break; }
                    if (b != s.b1) break;
                    take = sel_key(x[j], d[j], s.desc, id_bits, deg_bits) <= s.kmax;
                    break;
                }
                default: break;
            }
            if (take) {
                take_mask |= 1u << j;
                deg_sum += d[j];
                if (d[j] <= heavy_degree) ++n_light;
                if (mark) mark[x[j]] = mark_val;
            } else {
                keep_mask |= 1u << j;
                ++n_keep;
            }
        }
        int64_t pos = block_reserve(n_light, &sel.cnt[LC_LIGHT], s_app, deg_sum, &sel.cnt[LC_DEG]);
#pragma unroll
        for (int j = 0; j < kCompactItems; ++j) {
            if (!(take_mask & (1u << j))) continue;
            if (d[j] <= heavy_degree) {
                sel.light[pos++] = x[j];
            } else {  // heavy rows are rare: one slot each
                const int64_t slot = (int64_t)atomicAdd(&sel.cnt[LC_SLOTS], 1ull);
                const int64_t nch = ((int64_t)d[j] + chunk_edges - 1) / chunk_edges;
                const int64_t base = (int64_t)atomicAdd(&sel.cnt[LC_CHUNKS], (unsigned long long)nch);
                sel.hv[slot] = x[j];
                sel.hn[slot] = (int32_t)nch;
                atomicAdd(&sel.cnt[LC_HDEG], (unsigned long long)d[j]);
                for (int64_t c = 0; c < nch; ++c) sel.hc[base + c] = (slot << 32) | c;
            }
        }
        if (has_carry) {
            int64_t cp = block_reserve(n_keep, carry.cnt, s_app);
#pragma unroll
            for (int j = 0; j < kCompactItems; ++j) {
                if (!(keep_mask & (1u << j))) continue;
                carry.x[cp] = x[j];
                carry.g[cp] = gg[j];
                carry.d[cp] = d[j];
                ++cp;
            }
        }
    }
}

// Dense records: one wave per 64 consecutive ids and item; the wave's take
// decisions (light and heavy) become two words of the B bitmap (plain stores:
// every word is written by exactly one wave), and the counts (light and
// heavy members, their degree sums) are summed per block and added once per
// block.  Heavy members get no list slots: their kernels walk the graph's
// heavy rows and test the bitmap.
__global__ __launch_bounds__(kSelThreads) void k_sel_compact_dense(const SelArgs *A,
                                                                   const int32_t *bstart, int jobs) {
    SEL_JOB();
    if ((a.done && *a.done) || !a.dense) return;
    const ListDev sel = a.sel;
    const SelState *st = a.st;
    const int b1_bins = a.b1_bins, id_bits = a.id_bits, deg_bits = a.deg_bits;
    const int heavy_degree = a.heavy_degree;
    const int64_t rn = a.rn, bwords = a.bwords;
    uint32_t *half = a.bits + (int64_t)(*a.rule.t_ptr & 1) * bwords;
    __shared__ SelLite s_st[kSelLdsGroups];
    __shared__ unsigned long long s_sum[4];
    if (threadIdx.x < 4) s_sum[threadIdx.x] = 0;
    __shared__ int32_t s_wc[kSelThreads / kWave];
    sel_lite_load(s_st, st, a.groups);
    const int w = threadIdx.x / kWave, lane = lane_id();
    const int64_t ng = (rn + kWave - 1) / kWave;
    const int64_t per_block = (int64_t)(blockDim.x / kWave) * kDenseItems;  // one rank chunk
    int64_t n_light = 0, deg_sum = 0, n_heavy = 0, heavy_deg = 0;
    // Block-uniform loop over 1024-id chunks (rank cuts rank across the block).
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
        const unsigned long long mark = a.rank_mark[base / per_block];
        if (mark) rank_cut(s_st, mark, d, gg, tk, s_wc);
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
                } else {
                    ++n_heavy;
                    heavy_deg += d[j];
                }
            }
        }
    }
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

#undef SEL_JOB

}  // namespace

void SelectWork::prepare(dmolp_ctx *ctx, int ngroups, int64_t n, int32_t max_degree) {
    const int ib = (int)(bits_for(n > 1 ? n - 1 : 1));
    const int db = (int)(bits_for(max_degree > 1 ? max_degree : 1));
    if (ref_cap < n) {
        ref.alloc(n);
        ref_cap = n;
    }
    const int64_t chunks = (n + (1 << kRankChunkLog) - 1) >> kRankChunkLog;
    const int rg = ngroups < kSelLdsGroups ? ngroups : kSelLdsGroups;
    if (rank_chunks != chunks || (int64_t)rank_cnt.bytes < (int64_t)sizeof(int32_t) * rg * rank_stride(chunks)) {
        rank_cnt.alloc(sizeof(int32_t) * (size_t)rg * (size_t)rank_stride(chunks > 0 ? chunks : 1));
        rank_mark.alloc(sizeof(unsigned long long) * (size_t)(chunks > 0 ? chunks : 1));
        DMOLP_HIP_CHECK(hipMemsetAsync(rank_mark.p, 0, rank_mark.bytes, ctx->stream));
        rank_chunks = chunks;
    }
    if (groups == ngroups && id_bits == ib && deg_bits == db) return;
    groups = ngroups;
    id_bits = ib;
    deg_bits = db;
    int b1 = kHist1LdsInts / ngroups;
    if (b1 > kSelBins) b1 = kSelBins;
    DMOLP_REQUIRE(b1 >= 2, DMOLP_E_UNSUPPORTED, "too many groups for the selection histogram");
    b1_bins = b1;
    digit_bits = kDigitBits;
    while (digit_bits > 1 && ((int64_t)groups << digit_bits) > kHist2LdsInts) --digit_bits;
    DMOLP_REQUIRE(((int64_t)groups << digit_bits) <= kHist2LdsInts, DMOLP_E_UNSUPPORTED,
                  "too many groups for the digit histogram");
    passes = (id_bits + deg_bits + digit_bits - 1) / digit_bits;
    h1.alloc(sizeof(int32_t) * (size_t)groups * b1_bins);
    h2.alloc(sizeof(int32_t) * ((size_t)groups << digit_bits));
    st.alloc(sizeof(SelState) * (size_t)groups);
    refining.alloc(sizeof(int32_t) * 4);
    DMOLP_HIP_CHECK(hipMemsetAsync(h1.p, 0, h1.bytes, ctx->stream));
    DMOLP_HIP_CHECK(hipMemsetAsync(h2.p, 0, h2.bytes, ctx->stream));
    DMOLP_HIP_CHECK(hipMemsetAsync(refining.p, 0, refining.bytes, ctx->stream));
}

SelArgs select_args(dmolp_ctx *ctx, SelectWork &w, const RecDev &in, const TakeRule &rule,
                    const int32_t *done, const ListDev &sel, const RecDev *carry, int32_t *mark,
                    const int32_t *mark_val) {
    SelArgs a{};
    a.in = in;
    a.sel = sel;
    a.carry = carry ? *carry : RecDev{nullptr, nullptr, nullptr, nullptr};
    a.has_carry = carry ? 1 : 0;
    a.ref = w.ref.dev;
    a.h1 = w.h1.as<int32_t>();
    a.h2 = w.h2.as<int32_t>();
    a.st = w.st.as<SelState>();
    a.refining = w.refining.as<int32_t>();
    a.rank_cnt = w.rank_cnt.as<int32_t>();
    a.rank_mark = w.rank_mark.as<unsigned long long>();
    a.rank_chunks = w.rank_chunks;
    a.done = done;
    a.mark = mark;
    a.mark_val = mark_val;
    a.rule = rule;
    a.groups = w.groups;
    a.b1_bins = w.b1_bins;
    a.id_bits = w.id_bits;
    a.deg_bits = w.deg_bits;
    a.digit_bits = w.digit_bits;
    a.heavy_degree = ctx->tune.heavy_degree;
    a.chunk_edges = ctx->tune.chunk_edges;
    a.dense = 0;
    return a;
}

void select_args_dense(SelArgs &a, const int32_t *rdeg, const int32_t *rlab, const uint8_t *rlab8,
                       int64_t rn, uint32_t *bits, int64_t bwords) {
    a.dense = 1;
    a.rdeg = rdeg;
    a.rlab = rlab;
    a.rlab8 = rlab8;
    a.rn = rn;
    a.bits = bits;
    a.bwords = bwords;
}

void select_enqueue(dmolp_ctx *ctx, const SelArgs *args, const LaunchSet &ls, int passes,
                    bool compact, bool dense) {
    hipStream_t s = ctx->stream;
    const int J = ls.jobs;
    hipLaunchKernelGGL(k_sel_hist1, dim3(ls.total[P_HIST1]), dim3(kSelThreads), 0, s, args,
                       ls.bs(P_HIST1), J);
    hipLaunchKernelGGL(k_sel_pick1, dim3(ls.total[P_GROUPS]), dim3(kSelThreads), 0, s, args,
                       ls.bs(P_GROUPS), J);
    if (dense) {
        // Dense jobs only: rank cuts, then digit passes for cuts in the last bin.
        hipLaunchKernelGGL(k_sel_rank_count, dim3(ls.total[P_GRID]), dim3(kSelThreads), 0, s, args,
                           ls.bs(P_GRID), J);
        hipLaunchKernelGGL(k_sel_rank_scan, dim3(ls.total[P_GROUPS]), dim3(kSelThreads), 0, s, args,
                           ls.bs(P_GROUPS), J);
        hipLaunchKernelGGL(k_sel_collect_dense, dim3(ls.total[P_GRID]), dim3(kSelThreads), 0, s, args,
                           ls.bs(P_GRID), J);
        // The refine list holds only cuts in the last bin here (few records):
        // the level-1 histogram's smaller grid suffices and launches faster.
        for (int p = 0; p < passes; ++p) {
            hipLaunchKernelGGL(k_sel_hist2, dim3(ls.total[P_HIST1]), dim3(kSelThreads), 0, s, args,
                               ls.bs(P_HIST1), J);
            hipLaunchKernelGGL(k_sel_pick2, dim3(ls.total[P_GROUPS]), dim3(kSelThreads), 0, s, args,
                               ls.bs(P_GROUPS), J);
        }
        if (compact)
            hipLaunchKernelGGL(k_sel_compact_dense, dim3(ls.total[P_GRID]), dim3(kSelThreads), 0, s, args,
                               ls.bs(P_GRID), J);
        DMOLP_HIP_CHECK(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_sel_collect, dim3(ls.total[P_GRID]), dim3(kSelThreads), 0, s, args,
                       ls.bs(P_GRID), J);
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL(k_sel_hist2, dim3(ls.total[P_GRID]), dim3(kSelThreads), 0, s, args,
                           ls.bs(P_GRID), J);
        hipLaunchKernelGGL(k_sel_pick2, dim3(ls.total[P_GROUPS]), dim3(kSelThreads), 0, s, args,
                           ls.bs(P_GROUPS), J);
    }
    hipLaunchKernelGGL(k_sel_compact, dim3(ls.total[P_GRID]), dim3(kSelThreads), 0, s, args,
                       ls.bs(P_GRID), J);
    DMOLP_HIP_CHECK(hipGetLastError());
}

void job_blocks(const dmolp_ctx *ctx, int64_t n, int64_t nnz, int groups, int64_t heavy, bool single,
                int32_t out[P_NPOL]) {
    const int64_t full = ctx->grid_blocks();
    if (single) {
        out[P_GRID] = (int32_t)full;
        out[P_HIST1] = ctx->cu_count * 4;
    } else {
        // Blocks in proportion to the job's work, so small graphs of a batch
        // do not hold idle blocks.
        const int64_t per = ctx->tune.batch_work_per_block > 0 ? ctx->tune.batch_work_per_block : 32768;
        int64_t b = (n + nnz / 4 + per - 1) / per;
        out[P_GRID] = (int32_t)(b < 1 ? 1 : (b > full ? full : b));
        int64_t h = (n + 4 * per - 1) / (4 * per);
        out[P_HIST1] = (int32_t)(h < 1 ? 1 : (h > ctx->cu_count * 2 ? ctx->cu_count * 2 : h));
    }
    out[P_GROUPS] = groups > 0 ? groups : 1;
    out[P_ONE] = 1;
    int64_t hb = (heavy + 255) / 256;
    out[P_HEAVY] = (int32_t)(hb < 1 ? 1 : (hb > 1024 ? 1024 : hb));
}

void launch_set_build(dmolp_ctx *ctx, LaunchSet &ls, const std::vector<std::array<int32_t, P_NPOL>> &blocks_in) {
    const int J = (int)blocks_in.size();
    // Grid-stride shares of a set whose jobs together take less than the
    // full grid (a few jobs of a served mix) are scaled up in proportion,
    // so the set's launches still fill the CUs; a job's results do not
    // depend on its share.
    std::vector<std::array<int32_t, P_NPOL>> blocks = blocks_in;
    const int64_t full = ctx->grid_blocks();
    int64_t sum = 0;
    for (const auto &b : blocks) sum += b[P_GRID];
    if (J > 1 && sum > 0 && sum < full)
        for (auto &b : blocks) {
            const int64_t x = (int64_t)b[P_GRID] * full / sum;
            b[P_GRID] = (int32_t)(x < 1 ? 1 : (x > full ? full : x));
        }
    std::vector<int32_t> h((size_t)P_NPOL * (J + 1));
    for (int p = 0; p < P_NPOL; ++p) {
        int64_t run = 0;
        for (int j = 0; j < J; ++j) {
            h[(size_t)p * (J + 1) + j] = (int32_t)run;
            run += blocks[(size_t)j][(size_t)p];
        }
        h[(size_t)p * (J + 1) + J] = (int32_t)run;
        DMOLP_REQUIRE(run < (1ll << 31), DMOLP_E_UNSUPPORTED, "launch too large");
        ls.total[p] = (int)run;
    }
    if (ls.jobs != J || !ls.bstart.p) ls.bstart.alloc(sizeof(int32_t) * h.size());
    ls.jobs = J;
    DMOLP_HIP_CHECK(hipMemcpyAsync(ls.bstart.p, h.data(), sizeof(int32_t) * h.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
    DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace dmolp
