// rows.hpp -- device-side traversal of CSR rows named by a work list.
//
// Light rows (degree <= heavy_degree) are processed 64 list entries per
// wave: each lane loads one entry's row bounds, a wave scan concatenates the
// rows, and the wave then walks the concatenated edge range 64 edges at a
// time, so short rows do not leave lanes idle.  Heavy rows are split into
// chunks of chunk_edges edges, one chunk per wave.
#pragma once
#include "dev_common.hpp"
#include "internal.hpp"

namespace dmolp {

__device__ __forceinline__ GraphDev gload(const GraphDev &v) {
    GraphDev r = v;
    r.off = gload(v.off);
    r.col = gload(v.col);
    return r;
}
__device__ __forceinline__ ListDev gload(const ListDev &v) {
    ListDev r = v;
    r.light = gload(v.light);
    r.hv = gload(v.hv);
    r.hn = gload(v.hn);
    r.hc = gload(v.hc);
    r.cnt = gload(v.cnt);
    return r;
}

// LDS block of one wave's batch: inclusive degree prefix and the value
// e - q for each slot, so edge index = s_base[slot] + q.
struct BatchLds {
    int32_t incl[kWave];
    int32_t v[kWave];
    int32_t aux[kWave];  // free per-slot word for the caller
    int64_t base[kWave];
};

// Load a batch of up to 64 list entries starting at b0.  `ids == nullptr`
// means the identity list 0..count-1.  Returns the batch's total edge count.
__device__ __forceinline__ int32_t batch_load(const GraphDev &g, const int32_t *ids, int64_t count,
                                              int64_t b0, BatchLds &lds, int32_t &v, int32_t &deg,
                                              bool &valid) {
    const int lane = lane_id();
    int64_t i = b0 + lane;
    valid = i < count;
    v = valid ? (ids ? ids[i] : (int32_t)i) : 0;
    int64_t s = valid ? g.off[v] : 0;
    deg = valid ? (int32_t)(g.off[v + 1] - s) : 0;
    int64_t incl = wave_incl_scan((int64_t)deg);
    lds.incl[lane] = (int32_t)incl;
    lds.v[lane] = v;
    lds.base[lane] = s - (incl - deg);
    int32_t total = (int32_t)wave_bcast(incl, kWave - 1);
    wave_sync();
    return total;
}

// Slot owning concatenated position q (q < total): the first slot whose
// inclusive prefix exceeds q.
__device__ __forceinline__ int batch_slot(const BatchLds &lds, int32_t q) {
    int lo = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1)
        if (lds.incl[lo + step - 1] <= q) lo += step;
    return lo;
}

// Slots owning the 64 positions w0 + lane of a batch, all lanes at once
// (every lane must call; positions at or past the batch's total get some
// slot below 64).  Lane r holds its row's concatenated range [start, end);
// the row covering w0 and every row starting inside the window mark their
// first position in `mark` (kWave ints of LDS, set to -1 at each batch
// start), and a max-scan over the window carries each mark forward.  Marks
// an earlier window of the same batch left name earlier rows, so they never
// win the max.  One LDS write and read and six DPP steps instead of a
// six-step binary search of dependent LDS reads per position.
__device__ __forceinline__ int batch_slot_window(int32_t *mark, int32_t w0, int32_t start, int32_t end) {
    const int lane = lane_id();
    const int32_t p = start > w0 ? start : w0;
    if (p < w0 + kWave && end > p) mark[p - w0] = lane;
    wave_sync();
    const int32_t x = wave_max_incl(mark[lane]);
    return x < 0 ? 0 : x;
}

// Append v to a list (all lanes of the wave must call; pred selects lanes).
__device__ __forceinline__ void list_append(const ListDev &L, bool pred, int32_t v, int64_t deg,
                                            int heavy_degree, int chunk_edges) {
    const bool light = pred && deg <= heavy_degree;
    int64_t pos = wave_append(light, &L.cnt[LC_LIGHT]);
    if (light) L.light[pos] = v;
    const bool heavy = pred && deg > heavy_degree;
    if (__ballot(heavy)) {
        if (heavy) {
            int64_t slot = (int64_t)atomicAdd(&L.cnt[LC_SLOTS], 1ull);
            int64_t nch = (deg + chunk_edges - 1) / chunk_edges;
            int64_t base = (int64_t)atomicAdd(&L.cnt[LC_CHUNKS], (unsigned long long)nch);
            L.hv[slot] = v;
            L.hn[slot] = (int32_t)nch;
            atomicAdd(&L.cnt[LC_HDEG], (unsigned long long)deg);
            for (int64_t c = 0; c < nch; ++c) L.hc[base + c] = (slot << 32) | c;
        }
    }
    wave_add(&L.cnt[LC_DEG], pred ? deg : 0);
}

// Block-level appends: one global atomic per block per call instead of one
// per wave (a single counter hit by every wave serialises at the L2).
// Every thread of the block must call, with block-uniform control flow.
constexpr int kMaxWavesPerBlock = 16;
struct BlockAppendLds {
    int32_t wc[kMaxWavesPerBlock];
    int64_t wd[kMaxWavesPerBlock];
    unsigned long long base;
};

// Reserve my_n consecutive slots of a global list: returns this thread's
// first slot.  One atomic per block per call (plus one for `sum_counter`,
// which receives the block's sum of my_sum).  Block-uniform calls only.
__device__ __forceinline__ int64_t block_reserve(int32_t my_n, unsigned long long *counter,
                                                 BlockAppendLds &lds, int64_t my_sum = 0,
                                                 unsigned long long *sum_counter = nullptr) {
    const int lane = lane_id(), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    int32_t incl = my_n;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int32_t y = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += y;
    }
    const int64_t wsum = sum_counter ? wave_sum(my_sum) : 0;
    if (lane == kWave - 1) { lds.wc[w] = incl; lds.wd[w] = wsum; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t run = 0;
        int64_t dt = 0;
        for (int i = 0; i < nw; ++i) {
            const int32_t c = lds.wc[i];
            lds.wc[i] = run;
            run += c;
            dt += lds.wd[i];
        }
        lds.base = run ? atomicAdd(counter, (unsigned long long)run) : 0ull;
        if (sum_counter && dt) atomicAdd(sum_counter, (unsigned long long)dt);
    }
    __syncthreads();
    const int64_t first = (int64_t)(lds.base + (unsigned long long)lds.wc[w]) + (incl - my_n);
    __syncthreads();
    return first;
}

// Block-level list_append.
__device__ __forceinline__ void block_list_append(const ListDev &L, bool pred, int32_t v,
                                                  int64_t deg, int heavy_degree, int chunk_edges,
                                                  BlockAppendLds &lds) {
    const bool light = pred && deg <= heavy_degree;
    const uint64_t m = __ballot(light);
    const int64_t dsum = wave_sum(pred ? deg : 0);
    const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    if (lane_id() == 0) { lds.wc[w] = __popcll(m); lds.wd[w] = dsum; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t run = 0;
        int64_t dtot = 0;
        for (int i = 0; i < nw; ++i) {
            const int32_t c = lds.wc[i];
            lds.wc[i] = run;
            run += c;
            dtot += lds.wd[i];
        }
        lds.base = run ? atomicAdd(&L.cnt[LC_LIGHT], (unsigned long long)run) : 0ull;
        if (dtot) atomicAdd(&L.cnt[LC_DEG], (unsigned long long)dtot);
    }
    __syncthreads();
    if (light) L.light[lds.base + (unsigned long long)lds.wc[w] + (unsigned long long)lanes_below(m)] = v;
    if (pred && deg > heavy_degree) {
        const int64_t slot = (int64_t)atomicAdd(&L.cnt[LC_SLOTS], 1ull);
        const int64_t nch = (deg + chunk_edges - 1) / chunk_edges;
        const int64_t base = (int64_t)atomicAdd(&L.cnt[LC_CHUNKS], (unsigned long long)nch);
        L.hv[slot] = v;
        L.hn[slot] = (int32_t)nch;
        atomicAdd(&L.cnt[LC_HDEG], (unsigned long long)deg);
        for (int64_t c = 0; c < nch; ++c) L.hc[base + c] = (slot << 32) | c;
    }
    __syncthreads();
}

// Bounds of one heavy chunk item.
__device__ __forceinline__ void chunk_bounds(const GraphDev &g, const ListDev &L, int64_t item,
                                             int chunk_edges, int32_t &slot, int32_t &v,
                                             int64_t &e0, int64_t &e1, int64_t &deg) {
    slot = (int32_t)(item >> 32);
    const int64_t chunk = item & 0xffffffffll;
    v = L.hv[slot];
    const int64_t s = g.off[v], t = g.off[v + 1];
    deg = t - s;
    e0 = s + chunk * chunk_edges;
    e1 = e0 + chunk_edges < t ? e0 + chunk_edges : t;
}

}  // namespace dmolp
