// halo.hip -- P ranks' shares of one graph on one GPU, each with an
// on-device halo of ghost labels, running phase 2 rank by rank.
//
// The reference partitions over MPI ranks: rank r owns the vertices
// v with v mod P == r (partition.cpp:55-57), holds their rows with
// neighbours renumbered as local ids (owned) or ghost ids after the local
// ones, in order of first appearance (partition.cpp:182-262), and gets
// each ghost's label from its owner (partition.cpp:264-318).  Every phase-2
// round each rank sums its partial statistics with the others'
// (phase2.cpp:55-117, 214-253), selects and decides on its own boundary,
// and the ranks exchange Delta{gid, label} lists so each refreshes its ghost
// labels (phase2.cpp:122-154, 349-381).
//
// Here every share lives in device memory and every per-vertex and
// per-entry step runs in kernels over the share: the ghost fetch, the
// counts, the boundary test, the selection (a stable radix sort by label
// and degree), the LP decision, the Delta list and the halo refresh (a
// binary search of each delta's global id in the share's sorted ghost ids).
// The collectives of the multi-rank run become what they compute on one
// GPU: the partial counts are summed (the Allreduce), the Delta lists are
// concatenated in rank order (the Allgatherv).  The host drives the round
// loop and computes the penalties, as the reference's host does
// (phase2.cpp:158-209).
//
// Selection scope 0 is the reference's: each rank takes each label's
// ceil(ratio * c) among its own boundary vertices, and its candidates are
// its previous boundary and their owned neighbours (ghost ids are never
// tested, gpu_lp_boundary.cu:251-252).  Scope 1 selects over all ranks'
// boundary vertices together, with candidates next to any rank's previous
// boundary (each rank learns which of its ghosts were in it through a
// second exchange of global ids), so that the run equals the one-rank run.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "halo.hpp"
#include "lp_rule.hpp"

namespace dmolp {
namespace {

constexpr int kHaloThreads = 256;
constexpr int kHaloWaves = kHaloThreads / kWave;

__global__ void k_sh_deg(const int64_t *off, int64_t n, int32_t P, int32_t r, int64_t nl,
                         int64_t *row_len) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nl;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = r + i * P;
        row_len[i] = off[v + 1] - off[v];
    }
}

// Owned neighbours become local ids (u / P); a ghost is written as -(u + 1)
// and its first position in the share's column array is kept in first[u].
__global__ void k_sh_cols(const int64_t *off, const int32_t *col, int32_t P, int32_t r, int64_t nl,
                          const int64_t *soff, int32_t *scol, unsigned long long *first) {
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int lane = lane_id();
    for (int64_t i = w0; i < nl; i += nw) {
        const int64_t v = r + i * P;
        const int64_t s = off[v], d = off[v + 1] - s, o = soff[i];
        for (int64_t j = lane; j < d; j += kWave) {
            const int32_t u = col[s + j];
            if (u % P == r) {
                scol[o + j] = u / P;
            } else {
                scol[o + j] = -(u + 1);
                atomicMin(&first[u], (unsigned long long)(o + j));
            }
        }
    }
}

__global__ void k_fill_u64(unsigned long long *a, int64_t n, unsigned long long x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = x;
}

__global__ void k_ghost_flags(const unsigned long long *first, int64_t n, uint8_t *flag) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (int64_t)gridDim.x * blockDim.x)
        flag[u] = first[u] != ~0ull;
}

__global__ void k_ghost_keys(const unsigned long long *first, const int32_t *sgid, int64_t ng,
                             unsigned long long *key) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < ng;
         j += (int64_t)gridDim.x * blockDim.x)
        key[j] = first[sgid[j]];
}

// slot_of[u] for the share's ghosts (ghost_gid in slot order).
__global__ void k_slot_map(const int32_t *ghost_gid, int64_t ng, int32_t *slot_of) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ng;
         s += (int64_t)gridDim.x * blockDim.x)
        slot_of[ghost_gid[s]] = (int32_t)s;
}

__global__ void k_sorted_slots(const int32_t *sgid, int64_t ng, const int32_t *slot_of,
                               int32_t *sslot, unsigned long long *first) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < ng;
         j += (int64_t)gridDim.x * blockDim.x) {
        sslot[j] = slot_of[sgid[j]];
        first[sgid[j]] = ~0ull;  // ready for the next share
    }
}

__global__ void k_ghost_cols(int32_t *scol, int64_t nnz, int64_t nl, const int32_t *slot_of) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = scol[e];
        if (c < 0) scol[e] = (int32_t)(nl + slot_of[-(c + 1)]);
    }
}

__global__ void k_own_labels(const int32_t *L, int32_t P, int32_t r, int64_t nl, int32_t *lab) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nl;
         i += (int64_t)gridDim.x * blockDim.x)
        lab[i] = L[r + i * P];
}

// The ghost fetch (partition.cpp:264-318): each ghost's label as its owner
// holds it, read from the owner's share.
__global__ void k_fetch_ghosts(const ShareDev *S, int32_t P, int32_t r) {
    const ShareDev s = S[r];
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < s.ng;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int32_t u = s.ghost_gid[g];
        s.lab[s.nl + g] = S[u % P].lab[u / P];
    }
}

// Partial counts of one share (phase2.cpp:55-98, 214-236 before the
// Allreduce): owned labels, intra and differing entries of its rows, -1
// labels met.  One wave per row.
__global__ __launch_bounds__(kHaloThreads) void k_sh_counts(ShareDev s, int32_t k,
                                                            unsigned long long *part) {
    // part: [0, k) vc, [k, 2k) intra entries, 2k differing entries, 2k+1 invalid
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int lane = lane_id();
    int64_t diffs = 0, invalid = 0;
    for (int64_t i = w0; i < s.nl; i += nw) {
        const int32_t li = s.lab[i];
        const bool ok = li >= 0 && li < k;
        int64_t intra = 0;
        for (int64_t e = s.off[i] + lane; e < s.off[i + 1]; e += kWave) {
            const int32_t lj = s.lab[s.col[e]];
            if (ok && lj == li) ++intra;
            if (li == -1 || lj == -1) { ++invalid; continue; }
            if (li != lj) ++diffs;
        }
        intra = wave_sum(intra);
        if (lane == 0) {
            if (ok) atomicAdd(&part[li], 1ull);
            if (intra) atomicAdd(&part[k + li], (unsigned long long)intra);
        }
    }
    diffs = wave_sum(diffs);
    invalid = wave_sum(invalid);
    if (lane == 0) {
        if (diffs) atomicAdd(&part[2 * k], (unsigned long long)diffs);
        if (invalid) atomicAdd(&part[2 * k + 1], (unsigned long long)invalid);
    }
}

// Candidates and the boundary test (gpu_lp_boundary.cu:197-269) for every
// owned vertex: key (label << 32 | degree key) for boundary vertices, else
// ~0.  The degree key orders a label's group by degree descending when its
// RE > 1.0, else ascending (gpu_lp_boundary.cu:277-294).
__global__ __launch_bounds__(kHaloThreads) void k_sh_bound(ShareDev s, int32_t scope,
                                                           const double *re,
                                                           unsigned long long *key) {
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int lane = lane_id();
    for (int64_t i = w0; i < s.nl; i += nw) {
        const int32_t li = s.lab[i];
        const int64_t e0 = s.off[i], e1 = s.off[i + 1];
        bool cand = s.inb[i] != 0, bd = false;
        for (int64_t e = e0 + lane; e < e1; e += kWave) {
            const int32_t j = s.col[e];
            if (j < s.nl) cand |= s.inb[j] != 0;
            else cand |= scope == 1 && s.ghost_inb[j - s.nl] != 0;
            bd |= s.lab[j] != li;
        }
        cand = __ballot(cand) != 0;
        bd = __ballot(bd) != 0;
        if (lane == 0) {
            unsigned long long k = ~0ull;
            if (cand && bd) {
                const uint32_t d = (uint32_t)(e1 - e0);
                const uint32_t dk = re[li] > 1.0 ? 0xffffffffu - d : d;
                k = (unsigned long long)(uint32_t)li << 32 | dk;
            }
            key[i] = k;
        }
    }
}

__global__ void k_key_flags(const unsigned long long *key, int64_t n, uint8_t *flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = key[i] != ~0ull;
}

// Scope 1: the shares' keys in global-id order.
__global__ void k_gather_keys(const ShareDev *S, int32_t P, int64_t n, unsigned long long *gkey) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        gkey[v] = S[v % P].key[v / P];
}

// Group sizes (labels of the sorted records).
__global__ void k_group_count(const unsigned long long *skey, int64_t m, unsigned long long *cnt) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m;
         p += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[skey[p] >> 32], 1ull);
}

// Records in sorted order: a record at position p of label l is taken when
// p - start[l] < take[l]; it marks its vertex in the owning share's new
// boundary (ids: local when P1 == 1, else global ids of a P1-rank split).
__global__ void k_take(const unsigned long long *skey, const int32_t *sval, int64_t m,
                       const int64_t *start, const int64_t *take, const ShareDev *S, int32_t P1,
                       int32_t r) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t l = (int32_t)(skey[p] >> 32);
        if (p - start[l] >= take[l]) continue;
        const int32_t x = sval[p];
        if (P1 == 1) S[r].nb[x] = 1;
        else S[x % P1].nb[x / P1] = 1;
    }
}

// LP decisions (gpu_lp_boundary.cu:354-421) of the share's new boundary
// on the unified labels, one wave per vertex; Jacobi: new labels go to
// `nxt`, changed vertices are flagged.
__global__ __launch_bounds__(kHaloThreads) void k_sh_lp(ShareDev s, int32_t k, const double *pen,
                                                        uint8_t *changed) {
    __shared__ int32_t h[kHaloWaves][kLpCap];
    const int w = threadIdx.x / kWave, lane = lane_id();
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int E = k < kLpCap ? k : kLpCap;
    for (int64_t i = w0; i < s.nl; i += nw) {
        if (lane == 0) changed[i] = 0;
        if (!s.inb[i]) continue;  // wave-uniform: B_t (k_b_list moved it here)
        if (lane < kLpCap) h[w][lane] = 0;
        wave_sync();
        for (int64_t e = s.off[i] + lane; e < s.off[i + 1]; e += kWave) {
            const int32_t l = s.lab[s.col[e]];
            if (l >= 0 && l < E) atomicAdd(&h[w][l], 1);
        }
        wave_sync();
        if (lane == 0) {
            const int32_t own = s.lab[i];
            const int32_t nl = lp_decide_by([&](int l) { return h[w][l]; }, E, pen, own);
            s.nxt[i] = nl;
            changed[i] = nl != own;
        }
        wave_sync();
    }
}

// Commit the share's decisions and list its deltas (gid, label) in local
// order (compacted ids in `ids`).
__global__ void k_sh_commit(ShareDev s, int32_t P, int32_t r, const int32_t *ids, int64_t m,
                            int32_t *d_gid, int32_t *d_lab) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int32_t i = ids[j];
        const int32_t l = s.nxt[i];
        s.lab[i] = l;
        d_gid[j] = r + i * P;
        d_lab[j] = l;
    }
}

// Position of global id u among the share's sorted ghost ids, or -1.
__device__ __forceinline__ int64_t ghost_pos(const ShareDev &s, int32_t u) {
    int64_t lo = 0, hi = s.ng;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (s.sgid[mid] < u) lo = mid + 1; else hi = mid;
    }
    return lo < s.ng && s.sgid[lo] == u ? lo : -1;
}

// The halo refresh (phase2.cpp:368-376): every delta whose vertex is one of
// the share's ghosts rewrites that ghost's label.
__global__ void k_halo_refresh(ShareDev s, const int32_t *d_gid, const int32_t *d_lab, int64_t m,
                               unsigned long long *updated) {
    int64_t mine = 0;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = ghost_pos(s, d_gid[j]);
        if (q >= 0) {
            s.lab[s.nl + s.sslot[q]] = d_lab[j];
            ++mine;
        }
    }
    mine = wave_sum(mine);
    if (lane_id() == 0 && mine) atomicAdd(updated, (unsigned long long)mine);
}

// Scope 1: which of the share's ghosts are in some rank's boundary.
__global__ void k_halo_bbits(ShareDev s, const int32_t *b_gid, int64_t m) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = ghost_pos(s, b_gid[j]);
        if (q >= 0) s.ghost_inb[s.sslot[q]] = 1;
    }
}

__global__ void k_b_list(ShareDev s, int32_t P, int32_t r, uint8_t *flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s.nl;
         i += (int64_t)gridDim.x * blockDim.x) {
        s.inb[i] = s.nb[i];
        flag[i] = s.nb[i];
        s.nb[i] = 0;
    }
}

__global__ void k_gid_of(const int32_t *ids, int64_t m, int32_t P, int32_t r, int32_t *gid) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x)
        gid[j] = r + ids[j] * P;
}

__global__ void k_scatter_labels(ShareDev s, int32_t P, int32_t r, int32_t *L) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s.nl;
         i += (int64_t)gridDim.x * blockDim.x)
        L[r + i * P] = s.lab[i];
}

__global__ void k_fill_u8(uint8_t *a, int64_t n, uint8_t x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = x;
}

int grid_for(int64_t items, int per_block = kHaloThreads) {
    const int64_t b = (items + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 4096));
}

}  // namespace

// ---------------------------------------------------------------- host side

namespace {

// Temporary storage of the hipCUB device calls, grown on demand.
void *cub_temp(dmolp_shares *S, size_t bytes) {
    S->temp.grow(bytes > 0 ? bytes : 1);
    return S->temp.p;
}

// ids (int32) of the flagged entries of [0, n), in order; returns the count.
int64_t select_flagged(dmolp_ctx *ctx, dmolp_shares *S, const uint8_t *flag, int64_t n,
                       int32_t *out) {
    hipStream_t st = ctx->stream;
    hipcub::CountingInputIterator<int32_t> it(0);
    size_t bytes = 0;
    DMOLP_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, bytes, it, flag, out,
                                                  S->count.as<int64_t>(), n, st));
    void *t = cub_temp(S, bytes);
    DMOLP_HIP_CHECK(hipcub::DeviceSelect::Flagged(t, bytes, it, flag, out, S->count.as<int64_t>(),
                                                  n, st));
    int64_t m = 0;
    DMOLP_HIP_CHECK(hipMemcpyAsync(&m, S->count.p, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    DMOLP_HIP_CHECK(hipStreamSynchronize(st));
    return m;
}

}  // namespace

dmolp_shares *shares_create(dmolp_ctx *ctx, dmolp_graph *g, int32_t P) {
    DMOLP_REQUIRE(P >= 1 && P <= 64 && P <= g->n, DMOLP_E_ARG, "P must be in [1, min(64, n)]");
    DMOLP_REQUIRE(g->labels_valid, DMOLP_E_STATE, "graph has no resident labels");
    hipStream_t st = ctx->stream;
    std::unique_ptr<dmolp_shares> S(new dmolp_shares());
    S->P = P;
    S->n = g->n;
    S->g = g;
    S->sh.resize((size_t)P);
    S->count.alloc(sizeof(int64_t) * 2);
    const int64_t n = g->n;
    DevMem first(sizeof(unsigned long long) * (size_t)n);
    DevMem slot_of(sizeof(int32_t) * (size_t)n);
    DevMem flags(std::max<int64_t>(n, 1));
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(kHaloThreads), 0, st,
                       first.as<unsigned long long>(), n, ~0ull);
    for (int32_t r = 0; r < P; ++r) {
        Share &s = S->sh[(size_t)r];
        const int64_t nl = (n - r + P - 1) / P;
        s.nl = nl;
        s.off.alloc(sizeof(int64_t) * (size_t)(nl + 1));
        DMOLP_HIP_CHECK(hipMemsetAsync(s.off.p, 0, sizeof(int64_t), st));
        hipLaunchKernelGGL(k_sh_deg, dim3(grid_for(nl)), dim3(kHaloThreads), 0, st,
                           g->off.as<int64_t>(), n, P, r, nl, s.off.as<int64_t>() + 1);
        size_t bytes = 0;
        DMOLP_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, s.off.as<int64_t>() + 1,
                                                         s.off.as<int64_t>() + 1, nl, st));
        DMOLP_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(cub_temp(S.get(), bytes), bytes,
                                                         s.off.as<int64_t>() + 1,
                                                         s.off.as<int64_t>() + 1, nl, st));
        DMOLP_HIP_CHECK(hipMemcpyAsync(&s.nnz, s.off.as<int64_t>() + nl, sizeof(int64_t),
                                       hipMemcpyDeviceToHost, st));
        DMOLP_HIP_CHECK(hipStreamSynchronize(st));
        s.col.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(s.nnz, 1));
        hipLaunchKernelGGL(k_sh_cols, dim3(grid_for(nl * kWave)), dim3(kHaloThreads), 0, st,
                           g->off.as<int64_t>(), g->col.as<int32_t>(), P, r, nl,
                           s.off.as<int64_t>(), s.col.as<int32_t>(), first.as<unsigned long long>());
        // Ghost ids ascending (compaction), then in order of first appearance
        // (their first positions sorted): the slot order.
        hipLaunchKernelGGL(k_ghost_flags, dim3(grid_for(n)), dim3(kHaloThreads), 0, st,
                           first.as<unsigned long long>(), n, flags.as<uint8_t>());
        DevMem ids(sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1));
        const int64_t ng = select_flagged(ctx, S.get(), flags.as<uint8_t>(), n, ids.as<int32_t>());
        s.ng = ng;
        s.sgid.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(ng, 1));
        s.sslot.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(ng, 1));
        s.ghost_gid.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(ng, 1));
        DMOLP_HIP_CHECK(hipMemcpyAsync(s.sgid.p, ids.p, sizeof(int32_t) * (size_t)ng,
                                       hipMemcpyDeviceToDevice, st));
        if (ng > 0) {
            DevMem keys(sizeof(unsigned long long) * (size_t)ng),
                keys2(sizeof(unsigned long long) * (size_t)ng);
            hipLaunchKernelGGL(k_ghost_keys, dim3(grid_for(ng)), dim3(kHaloThreads), 0, st,
                               first.as<unsigned long long>(), s.sgid.as<int32_t>(), ng,
                               keys.as<unsigned long long>());
            bytes = 0;
            DMOLP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
                nullptr, bytes, keys.as<unsigned long long>(), keys2.as<unsigned long long>(),
                s.sgid.as<int32_t>(), s.ghost_gid.as<int32_t>(), ng, 0, 64, st));
            DMOLP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
                cub_temp(S.get(), bytes), bytes, keys.as<unsigned long long>(),
                keys2.as<unsigned long long>(), s.sgid.as<int32_t>(), s.ghost_gid.as<int32_t>(),
                ng, 0, 64, st));
            hipLaunchKernelGGL(k_slot_map, dim3(grid_for(ng)), dim3(kHaloThreads), 0, st,
                               s.ghost_gid.as<int32_t>(), ng, slot_of.as<int32_t>());
            hipLaunchKernelGGL(k_sorted_slots, dim3(grid_for(ng)), dim3(kHaloThreads), 0, st,
                               s.sgid.as<int32_t>(), ng, slot_of.as<int32_t>(),
                               s.sslot.as<int32_t>(), first.as<unsigned long long>());
            hipLaunchKernelGGL(k_ghost_cols, dim3(grid_for(s.nnz)), dim3(kHaloThreads), 0, st,
                               s.col.as<int32_t>(), s.nnz, nl, slot_of.as<int32_t>());
            DMOLP_HIP_CHECK(hipStreamSynchronize(st));
        }
        s.lab.alloc(sizeof(int32_t) * (size_t)(nl + ng));
        s.nxt.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nl, 1));
        s.inb.alloc((size_t)std::max<int64_t>(nl, 1));
        s.nb.alloc((size_t)std::max<int64_t>(nl, 1));
        s.ghost_inb.alloc((size_t)std::max<int64_t>(ng, 1));
        s.key.alloc(sizeof(unsigned long long) * (size_t)std::max<int64_t>(nl, 1));
        hipLaunchKernelGGL(k_own_labels, dim3(grid_for(nl)), dim3(kHaloThreads), 0, st,
                           g->labels.as<int32_t>(), P, r, nl, s.lab.as<int32_t>());
    }
    S->dev.alloc(sizeof(ShareDev) * (size_t)P);
    S->upload_views(ctx);
    for (int32_t r = 0; r < P; ++r)
        hipLaunchKernelGGL(k_fetch_ghosts, dim3(grid_for(S->sh[(size_t)r].ng)), dim3(kHaloThreads),
                           0, st, S->dev.as<ShareDev>(), P, r);
    S->upd.alloc(sizeof(unsigned long long));
    S->dcount.assign((size_t)P, 0);
    std::vector<int64_t> off_h((size_t)n + 1);
    DMOLP_HIP_CHECK(hipMemcpyAsync(off_h.data(), g->off.p, sizeof(int64_t) * (size_t)(n + 1),
                                   hipMemcpyDeviceToHost, st));
    DMOLP_HIP_CHECK(hipGetLastError());
    DMOLP_HIP_CHECK(hipStreamSynchronize(st));
    S->deg_h.resize((size_t)n);
    for (int64_t v = 0; v < n; ++v) S->deg_h[(size_t)v] = off_h[(size_t)v + 1] - off_h[(size_t)v];
    return S.release();
}


int64_t halo_refresh(dmolp_ctx *ctx, dmolp_shares *S, int32_t r, const int32_t *d_gid,
                     const int32_t *d_lab, int64_t m) {
    hipStream_t st = ctx->stream;
    const Share &s = S->sh[(size_t)r];
    DMOLP_HIP_CHECK(hipMemsetAsync(S->upd.p, 0, sizeof(unsigned long long), st));
    if (m > 0)
        hipLaunchKernelGGL(k_halo_refresh, dim3(grid_for(m)), dim3(kHaloThreads), 0, st, s.view(),
                           d_gid, d_lab, m, S->upd.as<unsigned long long>());
    DMOLP_HIP_CHECK(hipGetLastError());
    unsigned long long u = 0;
    DMOLP_HIP_CHECK(hipMemcpyAsync(&u, S->upd.p, sizeof(u), hipMemcpyDeviceToHost, st));
    DMOLP_HIP_CHECK(hipStreamSynchronize(st));
    return (int64_t)u;
}

namespace {

// phase2.cpp:158-209 in source order on the host (the reference computes the
// penalties on its host and broadcasts them; glibc log).
void penalties_host(int k, const std::vector<int64_t> &vc, const std::vector<int64_t> &ic, int t,
                    std::vector<double> &pen, std::vector<double> &RE) {
    int64_t tv = 0, te = 0;
    for (int i = 0; i < k; ++i) { tv += vc[(size_t)i]; te += ic[(size_t)i]; }
    const double exp_v = static_cast<double>(tv) / k;
    const double exp_e = (te > 0) ? static_cast<double>(te) / k : 1.0;
    std::vector<double> RV((size_t)k);
    RE.assign((size_t)k, 0.0);
    for (int i = 0; i < k; ++i) {
        RV[(size_t)i] = (exp_v > 0) ? static_cast<double>(vc[(size_t)i]) / exp_v : 1.0;
        RE[(size_t)i] = (exp_e > 0) ? static_cast<double>(ic[(size_t)i]) / exp_e : 1.0;
    }
    double rv_mean = 0.0, re_mean = 0.0;
    for (int i = 0; i < k; ++i) { rv_mean += RV[(size_t)i]; re_mean += RE[(size_t)i]; }
    rv_mean /= k;
    re_mean /= k;
    double rv_var = 0.0, re_var = 0.0;
    for (int i = 0; i < k; ++i) {
        rv_var += (RV[(size_t)i] - rv_mean) * (RV[(size_t)i] - rv_mean);
        re_var += (RE[(size_t)i] - re_mean) * (RE[(size_t)i] - re_mean);
    }
    rv_var /= k;
    re_var /= k;
    const double total_var = rv_var + re_var;
    const double imb_rv = (total_var > 0) ? rv_var / total_var : 0.0;
    const double imb_re = (total_var > 0) ? re_var / total_var : 0.0;
    pen.assign((size_t)k, 0.0);
    for (int i = 0; i < k; ++i) {
        const double g_rv = (1.0 - RV[(size_t)i]) * (1.0 + log((double)t));
        const double g_re = (1.0 - RE[(size_t)i]) * (1.0 + log((double)t));
        pen[(size_t)i] = imb_rv * g_rv + imb_re * g_re;
    }
}

// Sum of every share's partial counts (the Allreduce): vc, ic (halved per
// label, phase2.cpp:110) and the cut.  Returns a DMOLP status.
int shares_counts(dmolp_ctx *ctx, dmolp_shares *S, int32_t k, std::vector<int64_t> &vc,
                  std::vector<int64_t> &ic, int64_t &cut) {
    hipStream_t st = ctx->stream;
    const size_t per = (size_t)(2 * k + 2);
    S->parts.grow(sizeof(unsigned long long) * per * (size_t)S->P);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for((int64_t)per * S->P)), dim3(kHaloThreads), 0, st,
                       S->parts.as<unsigned long long>(), (int64_t)per * S->P, 0ull);
    for (int32_t r = 0; r < S->P; ++r) {
        const Share &s = S->sh[(size_t)r];
        hipLaunchKernelGGL(k_sh_counts, dim3(grid_for(s.nl * kWave)), dim3(kHaloThreads), 0, st,
                           s.view(), k, S->parts.as<unsigned long long>() + per * (size_t)r);
    }
    DMOLP_HIP_CHECK(hipGetLastError());
    std::vector<unsigned long long> h(per * (size_t)S->P);
    DMOLP_HIP_CHECK(hipMemcpyAsync(h.data(), S->parts.p, sizeof(unsigned long long) * h.size(),
                                   hipMemcpyDeviceToHost, st));
    DMOLP_HIP_CHECK(hipStreamSynchronize(st));
    vc.assign((size_t)k, 0);
    ic.assign((size_t)k, 0);
    int64_t diffs = 0, invalid = 0;
    for (int32_t r = 0; r < S->P; ++r) {
        const unsigned long long *p = h.data() + per * (size_t)r;
        for (int i = 0; i < k; ++i) { vc[(size_t)i] += (int64_t)p[i]; ic[(size_t)i] += (int64_t)p[k + i]; }
        diffs += (int64_t)p[2 * k];
        invalid += (int64_t)p[2 * k + 1];
    }
    if (invalid) return DMOLP_E_INVALID_LABEL;
    if (diffs & 1) return DMOLP_E_ODD_CUT;
    for (auto &x : ic) x /= 2;
    cut = diffs / 2;
    return DMOLP_OK;
}

// Sort the m records (key, id) by key, stably, and mark each label's take
// in the shares' new boundaries.
void select_records(dmolp_ctx *ctx, dmolp_shares *S, int32_t k, double ratio, int64_t m,
                    unsigned long long *key, int32_t *val, int32_t P1, int32_t r) {
    if (m == 0) return;
    hipStream_t st = ctx->stream;
    DevMem key2(sizeof(unsigned long long) * (size_t)m), val2(sizeof(int32_t) * (size_t)m);
    int end_bit = 33;
    while ((1ll << (end_bit - 32)) < k) ++end_bit;
    size_t bytes = 0;
    DMOLP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key,
                                                       key2.as<unsigned long long>(), val,
                                                       val2.as<int32_t>(), m, 0, end_bit, st));
    DMOLP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub_temp(S, bytes), bytes, key,
                                                       key2.as<unsigned long long>(), val,
                                                       val2.as<int32_t>(), m, 0, end_bit, st));
    DevMem cnt(sizeof(unsigned long long) * (size_t)k);
    DMOLP_HIP_CHECK(hipMemsetAsync(cnt.p, 0, cnt.bytes, st));
    hipLaunchKernelGGL(k_group_count, dim3(grid_for(m)), dim3(kHaloThreads), 0, st,
                       key2.as<unsigned long long>(), m, cnt.as<unsigned long long>());
    std::vector<unsigned long long> c((size_t)k);
    DMOLP_HIP_CHECK(hipMemcpyAsync(c.data(), cnt.p, cnt.bytes, hipMemcpyDeviceToHost, st));
    DMOLP_HIP_CHECK(hipStreamSynchronize(st));
    // gpu_lp_boundary.cu:296-301: ceil(ratio * size), capped at the size.
    std::vector<int64_t> tab(2 * (size_t)k);
    int64_t run = 0;
    for (int l = 0; l < k; ++l) {
        tab[(size_t)l] = run;
        int take = static_cast<int>(std::ceil(ratio * static_cast<double>(c[(size_t)l])));
        if (take > (int)c[(size_t)l]) take = (int)c[(size_t)l];
        tab[(size_t)k + l] = take;
        run += (int64_t)c[(size_t)l];
    }
    DevMem dtab(sizeof(int64_t) * tab.size());
    DMOLP_HIP_CHECK(hipMemcpyAsync(dtab.p, tab.data(), dtab.bytes, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_take, dim3(grid_for(m)), dim3(kHaloThreads), 0, st,
                       key2.as<unsigned long long>(), val2.as<int32_t>(), m, dtab.as<int64_t>(),
                       dtab.as<int64_t>() + k, S->dev.as<ShareDev>(), P1, r);
    DMOLP_HIP_CHECK(hipGetLastError());
    DMOLP_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace

void shares_phase2(dmolp_ctx *ctx, dmolp_shares *S, int32_t k, int32_t max_rounds, double epsilon,
                   int32_t scope, int32_t *labels_out, int64_t *vc_out, int64_t *ic_out,
                   dmolp_phase2_metrics *m, dmolp_trace *tr, int64_t *ghost_updates) {
    DMOLP_REQUIRE(k >= 1 && k <= DMOLP_MAX_K, DMOLP_E_ARG, "k outside [1, 1024]");
    DMOLP_REQUIRE(max_rounds >= 0, DMOLP_E_ARG, "max_rounds must be >= 0");
    DMOLP_REQUIRE(scope == 0 || scope == 1, DMOLP_E_ARG, "scope must be 0 or 1");
    hipStream_t st = ctx->stream;
    const int32_t P = S->P;
    const int64_t n = S->n;
    hipEvent_t e0 = ctx->ev0, e1 = ctx->ev1;
    DMOLP_HIP_CHECK(hipEventRecord(e0, st));
    std::vector<int64_t> vc, ic;
    int64_t prev_cut = 0;
    int rc = shares_counts(ctx, S, k, vc, ic, prev_cut);
    DMOLP_REQUIRE(rc == DMOLP_OK, rc, "the labels failed the shares' counts");
    int64_t labelled = 0;
    for (int64_t x : vc) labelled += x;
    DMOLP_REQUIRE(labelled == n, DMOLP_E_INVALID_LABEL, "every label must lie in [0, k)");
    dmolp_phase2_metrics mm{};
    mm.initial_cut = prev_cut;
    for (int32_t r = 0; r < P; ++r) {  // phase2.cpp:285-287: every owned vertex
        const Share &s = S->sh[(size_t)r];
        hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(s.nl)), dim3(kHaloThreads), 0, st,
                           s.inb.as<uint8_t>(), s.nl, (uint8_t)1);
        hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(s.nl)), dim3(kHaloThreads), 0, st,
                           s.nb.as<uint8_t>(), s.nl, (uint8_t)0);
        hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(s.ng)), dim3(kHaloThreads), 0, st,
                           s.ghost_inb.as<uint8_t>(), s.ng, (uint8_t)1);
    }
    int64_t max_nl = 0;
    for (const Share &s : S->sh) max_nl = std::max(max_nl, s.nl);
    DevMem dpen(sizeof(double) * (size_t)k), dre(sizeof(double) * (size_t)k);
    DevMem flags(std::max<int64_t>(n, 1)), ids(sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1));
    DevMem gkey(sizeof(unsigned long long) * (size_t)std::max<int64_t>(n, 1));
    DevMem dgid(sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1));
    DevMem dlab(sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1));
    // |B_0| and its degree sum: every owned row (phase2.cpp:285-287).
    std::vector<int64_t> prev_n((size_t)P, 0), prev_d((size_t)P, 0);
    for (int32_t r = 0; r < P; ++r) {
        prev_n[(size_t)r] = S->sh[(size_t)r].nl;
        prev_d[(size_t)r] = S->sh[(size_t)r].nnz;
    }
    std::vector<double> pen, re;
    int cc = 0;
    for (int t = 1; t <= max_rounds; ++t) {
        mm.rounds = t;
        penalties_host(k, vc, ic, t, pen, re);
        DMOLP_HIP_CHECK(hipMemcpyAsync(dpen.p, pen.data(), dpen.bytes, hipMemcpyHostToDevice, st));
        DMOLP_HIP_CHECK(hipMemcpyAsync(dre.p, re.data(), dre.bytes, hipMemcpyHostToDevice, st));
        const double ratio = exp((-t) / (5.0));  // gpu_lp_boundary.cu:244
        // Boundary records and the selection.
        int64_t nbd = 0;
        for (int32_t r = 0; r < P; ++r) {
            const Share &s = S->sh[(size_t)r];
            hipLaunchKernelGGL(k_sh_bound, dim3(grid_for(s.nl * kWave)), dim3(kHaloThreads), 0, st,
                               s.view(), scope, dre.as<double>(), s.key.as<unsigned long long>());
        }
        if (scope == 0) {
            for (int32_t r = 0; r < P; ++r) {
                const Share &s = S->sh[(size_t)r];
                hipLaunchKernelGGL(k_key_flags, dim3(grid_for(s.nl)), dim3(kHaloThreads), 0, st,
                                   s.key.as<unsigned long long>(), s.nl, flags.as<uint8_t>());
                const int64_t mr = select_flagged(ctx, S, flags.as<uint8_t>(), s.nl,
                                                  ids.as<int32_t>());
                nbd += mr;
                DevMem rk(sizeof(unsigned long long) * (size_t)std::max<int64_t>(mr, 1));
                if (mr > 0) {
                    size_t bytes = 0;
                    DMOLP_HIP_CHECK(hipcub::DeviceSelect::Flagged(
                        nullptr, bytes, s.key.as<unsigned long long>(), flags.as<uint8_t>(),
                        rk.as<unsigned long long>(), S->count.as<int64_t>(), s.nl, st));
                    DMOLP_HIP_CHECK(hipcub::DeviceSelect::Flagged(
                        cub_temp(S, bytes), bytes, s.key.as<unsigned long long>(),
                        flags.as<uint8_t>(), rk.as<unsigned long long>(), S->count.as<int64_t>(),
                        s.nl, st));
                }
                select_records(ctx, S, k, ratio, mr, rk.as<unsigned long long>(), ids.as<int32_t>(),
                               1, r);
            }
        } else {
            hipLaunchKernelGGL(k_gather_keys, dim3(grid_for(n)), dim3(kHaloThreads), 0, st,
                               S->dev.as<ShareDev>(), P, n, gkey.as<unsigned long long>());
            hipLaunchKernelGGL(k_key_flags, dim3(grid_for(n)), dim3(kHaloThreads), 0, st,
                               gkey.as<unsigned long long>(), n, flags.as<uint8_t>());
            const int64_t mr = select_flagged(ctx, S, flags.as<uint8_t>(), n, ids.as<int32_t>());
            nbd = mr;
            DevMem rk(sizeof(unsigned long long) * (size_t)std::max<int64_t>(mr, 1));
            if (mr > 0) {
                size_t bytes = 0;
                DMOLP_HIP_CHECK(hipcub::DeviceSelect::Flagged(
                    nullptr, bytes, gkey.as<unsigned long long>(), flags.as<uint8_t>(),
                    rk.as<unsigned long long>(), S->count.as<int64_t>(), n, st));
                DMOLP_HIP_CHECK(hipcub::DeviceSelect::Flagged(
                    cub_temp(S, bytes), bytes, gkey.as<unsigned long long>(), flags.as<uint8_t>(),
                    rk.as<unsigned long long>(), S->count.as<int64_t>(), n, st));
            }
            select_records(ctx, S, k, ratio, mr, rk.as<unsigned long long>(), ids.as<int32_t>(), P,
                           0);
        }
        // New boundary lists: counts, degree sums, and (scope 1) the exchange
        // of their global ids.
        int64_t nb = 0, bdeg = 0, nprev = 0, dprev = 0;
        int64_t nball = 0;
        for (int32_t r = 0; r < P; ++r) {
            nprev += prev_n[(size_t)r];
            dprev += prev_d[(size_t)r];
            const Share &s = S->sh[(size_t)r];
            hipLaunchKernelGGL(k_b_list, dim3(grid_for(s.nl)), dim3(kHaloThreads), 0, st, s.view(),
                               P, r, flags.as<uint8_t>());
            const int64_t mr = select_flagged(ctx, S, flags.as<uint8_t>(), s.nl, ids.as<int32_t>());
            std::vector<int32_t> hid((size_t)mr);
            if (mr > 0)
                DMOLP_HIP_CHECK(hipMemcpyAsync(hid.data(), ids.p, sizeof(int32_t) * (size_t)mr,
                                               hipMemcpyDeviceToHost, st));
            hipLaunchKernelGGL(k_gid_of, dim3(grid_for(mr)), dim3(kHaloThreads), 0, st,
                               ids.as<int32_t>(), mr, P, r, dgid.as<int32_t>() + nball);
            DMOLP_HIP_CHECK(hipStreamSynchronize(st));
            nball += mr;
            prev_n[(size_t)r] = mr;
            int64_t d = 0;
            for (int32_t i : hid) d += S->sh_deg(r, i);
            prev_d[(size_t)r] = d;
            nb += mr;
            bdeg += d;
        }
        mm.lp_edges += bdeg;
        if (scope == 1) {
            for (int32_t r = 0; r < P; ++r) {
                const Share &s = S->sh[(size_t)r];
                hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(s.ng)), dim3(kHaloThreads), 0, st,
                                   s.ghost_inb.as<uint8_t>(), s.ng, (uint8_t)0);
                if (nball > 0)
                    hipLaunchKernelGGL(k_halo_bbits, dim3(grid_for(nball)), dim3(kHaloThreads), 0,
                                       st, s.view(), dgid.as<int32_t>(), nball);
            }
        }
        const int q = t - 1;
        auto put = [&](int64_t *a, int64_t v) { if (tr && a) a[q] = v; };
        if (tr) {
            put(tr->prev_count, nprev);
            put(tr->prev_deg, dprev);
            put(tr->cand_count, -1);
            put(tr->bd_count, nbd);
            put(tr->b_count, nb);
            put(tr->b_deg, bdeg);
            put(tr->chg_count, 0);
            put(tr->chg_deg, 0);
            put(tr->cut, -1);
            put(tr->b_light_count, -1);
            put(tr->b_light_deg, -1);
            put(tr->chg_light_count, -1);
            put(tr->chg_light_deg, -1);
        }
        if (ghost_updates) ghost_updates[q] = 0;
        if (nb == 0) { mm.exit_reason = DMOLP_EXIT_EMPTY_BOUNDARY; break; }
        // Decisions on every share's boundary (Jacobi: all read the labels of
        // the round's start), then each share commits its deltas.
        for (int32_t r = 0; r < P; ++r) {
            const Share &s = S->sh[(size_t)r];
            hipLaunchKernelGGL(k_sh_lp, dim3(grid_for(s.nl * kWave)), dim3(kHaloThreads), 0, st,
                               s.view(), k, dpen.as<double>(), flags.as<uint8_t>());
            const int64_t mr = select_flagged(ctx, S, flags.as<uint8_t>(), s.nl, ids.as<int32_t>());
            S->dcount[(size_t)r] = mr;
            S->deltas_of(r, mr);
            hipLaunchKernelGGL(k_sh_commit, dim3(grid_for(mr)), dim3(kHaloThreads), 0, st,
                               s.view(), P, r, ids.as<int32_t>(), mr,
                               S->sh[(size_t)r].d_gid.as<int32_t>(),
                               S->sh[(size_t)r].d_lab.as<int32_t>());
        }
        // The Allgatherv: every share's deltas in rank order.
        int64_t nchg = 0, cdeg = 0;
        for (int32_t r = 0; r < P; ++r) {
            const int64_t mr = S->dcount[(size_t)r];
            if (mr > 0) {
                DMOLP_HIP_CHECK(hipMemcpyAsync(dgid.as<int32_t>() + nchg, S->sh[(size_t)r].d_gid.p,
                                               sizeof(int32_t) * (size_t)mr, hipMemcpyDeviceToDevice,
                                               st));
                DMOLP_HIP_CHECK(hipMemcpyAsync(dlab.as<int32_t>() + nchg, S->sh[(size_t)r].d_lab.p,
                                               sizeof(int32_t) * (size_t)mr, hipMemcpyDeviceToDevice,
                                               st));
            }
            nchg += mr;
        }
        {
            std::vector<int32_t> g((size_t)nchg);
            if (nchg > 0)
                DMOLP_HIP_CHECK(hipMemcpyAsync(g.data(), dgid.p, sizeof(int32_t) * (size_t)nchg,
                                               hipMemcpyDeviceToHost, st));
            DMOLP_HIP_CHECK(hipStreamSynchronize(st));
            for (int32_t v : g) cdeg += S->global_deg(v);
        }
        // The halo refresh of every share.
        int64_t refreshed = 0;
        for (int32_t r = 0; r < P; ++r)
            refreshed += halo_refresh(ctx, S, r, dgid.as<int32_t>(), dlab.as<int32_t>(), nchg);
        if (ghost_updates) ghost_updates[q] = refreshed;
        put(tr ? tr->chg_count : nullptr, nchg);
        put(tr ? tr->chg_deg : nullptr, cdeg);
        if (nchg == 0) { mm.exit_reason = DMOLP_EXIT_NO_CHANGE; break; }  // phase2.cpp:384-390
        int64_t cur = 0;
        rc = shares_counts(ctx, S, k, vc, ic, cur);
        DMOLP_REQUIRE(rc == DMOLP_OK, rc, "the shares' counts failed");
        put(tr ? tr->cut : nullptr, cur);
        const double delta =
            (prev_cut > 0) ? std::fabs((double)(cur - prev_cut) / (double)prev_cut) : 1.0;
        if (delta < epsilon) ++cc; else cc = 0;
        prev_cut = cur;
        if (cc >= 10) { mm.exit_reason = DMOLP_EXIT_CONVERGED; break; }
        if (t == max_rounds) mm.exit_reason = DMOLP_EXIT_MAX_ROUNDS;
    }
    int64_t dummy = 0;
    rc = shares_counts(ctx, S, k, vc, ic, dummy);
    DMOLP_REQUIRE(rc == DMOLP_OK, rc, "the shares' counts failed");
    // The owned labels of every share are the graph's labels.
    for (int32_t r = 0; r < P; ++r)
        hipLaunchKernelGGL(k_scatter_labels, dim3(grid_for(S->sh[(size_t)r].nl)),
                           dim3(kHaloThreads), 0, st, S->sh[(size_t)r].view(), P, r,
                           S->g->labels.as<int32_t>());
    DMOLP_HIP_CHECK(hipEventRecord(e1, st));
    if (labels_out)
        DMOLP_HIP_CHECK(hipMemcpyAsync(labels_out, S->g->labels.p, sizeof(int32_t) * (size_t)n,
                                       hipMemcpyDeviceToHost, st));
    DMOLP_HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    DMOLP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    int64_t tv = 0, te = 0;
    for (int i = 0; i < k; ++i) { tv += vc[(size_t)i]; te += ic[(size_t)i]; }
    const double exp_v = static_cast<double>(tv) / k;
    const double exp_e = (te > 0) ? static_cast<double>(te) / k : 1.0;
    double max_v = 0.0, max_e = 0.0;
    for (int i = 0; i < k; ++i) {
        max_v = std::max(max_v, (exp_v > 0) ? static_cast<double>(vc[(size_t)i]) / exp_v : 1.0);
        max_e = std::max(max_e, (exp_e > 0) ? static_cast<double>(ic[(size_t)i]) / exp_e : 1.0);
    }
    mm.edge_cut = prev_cut;
    mm.vertex_balance = max_v;
    mm.edge_balance = max_e;
    mm.total_vertices = tv;
    mm.total_edges = te;
    mm.device_ms = ms;
    if (m) *m = mm;
    if (vc_out) std::copy(vc.begin(), vc.end(), vc_out);
    if (ic_out) std::copy(ic.begin(), ic.end(), ic_out);
}

}  // namespace dmolp

void dmolp_shares::upload_views(dmolp_ctx *ctx) {
    std::vector<dmolp::ShareDev> v((size_t)P);
    for (int32_t r = 0; r < P; ++r) v[(size_t)r] = sh[(size_t)r].view();
    DMOLP_HIP_CHECK(hipMemcpyAsync(dev.p, v.data(), sizeof(dmolp::ShareDev) * (size_t)P,
                                   hipMemcpyHostToDevice, ctx->stream));
}
