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
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "phase1.hpp"
#include "phase2.hpp"
#include "rows.hpp"
#include "select.hpp"

namespace dmolp {
namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr int32_t kNoClaim = 0x7f7f7f7f;  // claims are reset with a 0x7f byte fill

__device__ __forceinline__ bool p1_done(const P1Ctl *c) { return c->done != 0; }

__device__ __forceinline__ void mark_bit(uint32_t *bits, int32_t v) {
    const uint32_t m = 1u << (v & 31);
    if (!(bits[v >> 5] & m)) atomicOr(&bits[v >> 5], m);
}

// ---------------------------------------------------------------- seeds

// Level 0 and 1 of the ball of seeds[j] (seed.cpp:26-29, 39-95).
__global__ void k_ball1(GraphDev g, const int32_t *seeds, int j, uint32_t *covered, ListDev lvl1,
                        int heavy_degree, int chunk_edges, const P1Ctl *ctl) {
    if (ctl->seed_stop) return;
    const int32_t s = seeds[j];
    if (s < 0) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) mark_bit(covered, s);
    const int64_t e0 = g.off[s], e1 = g.off[s + 1];
    for (int64_t b = e0 + (global_wave() * kWave); b < e1; b += wave_count() * kWave) {
        const int64_t e = b + lane_id();
        const bool on = e < e1;
        int32_t u = 0;
        int64_t d = 0;
        if (on) {
            u = g.col[e];
            mark_bit(covered, u);
            d = g.off[u + 1] - g.off[u];
        }
        list_append(lvl1, on, u, d, heavy_degree, chunk_edges);
    }
}

// Level 2: every neighbour of a level-1 vertex.
__global__ __launch_bounds__(kThreads) void k_ball2_light(GraphDev g, ListDev lvl1,
                                                          uint32_t *covered, const P1Ctl *ctl) {
    if (ctl->seed_stop) return;
    __shared__ BatchLds s_b[kWavesPerBlock];
    const int w = threadIdx.x / kWave, lane = lane_id();
    const int64_t count = (int64_t)lvl1.cnt[LC_LIGHT];
    for (int64_t b0 = global_wave() * kWave; b0 < count; b0 += wave_count() * kWave) {
        int32_t v, deg;
        bool valid;
        const int32_t total = batch_load(g, lvl1.light, count, b0, s_b[w], v, deg, valid);
        for (int32_t q0 = 0; q0 < total; q0 += kWave) {
            const int32_t q = q0 + lane;
            if (q < total) mark_bit(covered, g.col[s_b[w].base[batch_slot(s_b[w], q)] + q]);
        }
        wave_sync();
    }
}

__global__ void k_ball2_heavy(GraphDev g, ListDev lvl1, int chunk_edges, uint32_t *covered,
                              const P1Ctl *ctl) {
    if (ctl->seed_stop) return;
    const int64_t nch = (int64_t)lvl1.cnt[LC_CHUNKS];
    for (int64_t j = global_wave(); j < nch; j += wave_count()) {
        int32_t slot, v;
        int64_t e0, e1, deg;
        chunk_bounds(g, lvl1, lvl1.hc[j], chunk_edges, slot, v, e0, e1, deg);
        for (int64_t e = e0 + lane_id(); e < e1; e += kWave) mark_bit(covered, g.col[e]);
    }
}

// seed.cpp:158-170: over unused, uncovered hubs, the largest degree, first
// in hub order; key = degree << 32 | ~index.
__global__ void k_next_seed(GraphDev g, const int32_t *hubs, int64_t nh, const uint8_t *used,
                            const uint32_t *covered, unsigned long long *best, const P1Ctl *ctl) {
    if (ctl->seed_stop) return;
    unsigned long long mine = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (used[i]) continue;
        const int32_t h = hubs[i];
        if (covered[h >> 5] & (1u << (h & 31))) continue;
        const unsigned long long d = (unsigned long long)(g.off[h + 1] - g.off[h]);
        const unsigned long long key = (d << 32) | (0xffffffffull - (unsigned long long)i);
        if (key > mine) mine = key;
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(mine, o, kWave);
        if (other > mine) mine = other;
    }
    if (lane_id() == 0 && mine) atomicMax(best, mine);
}

__global__ void k_take_seed(unsigned long long *best, const int32_t *hubs, uint8_t *used,
                            int32_t *seeds, int j, P1Ctl *ctl, ListDev lvl1) {
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

__global__ void k_expand_init(GraphDev g, const int32_t *seeds, int k, int32_t *L, RecDev f,
                              unsigned long long *sizes, P1Ctl *ctl) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < k) {
        const int32_t s = seeds[p];
        L[s] = p;  // partition.cpp:346-352
        f.x[p] = s;
        f.g[p] = p;
        f.d[p] = (int32_t)(g.off[s + 1] - g.off[s]);
        sizes[p] = 1;
    }
    if (p == 0) {
        *f.cnt = (unsigned long long)k;
        ctl->labelled = k;
        ctl->rounds = 0;
        ctl->fill_round = -1;
        ctl->done = (int64_t)k >= g.n ? 1 : 0;
    }
}

// Round start: stop when everything is labelled (partition.cpp:371); when
// every frontier is empty pick the fill partition (partition.cpp:376-387).
__global__ void k_p1_round_begin(P1Ctl *ctl, RecDev fcur, RecDev fnext, ListDev sel,
                                 const unsigned long long *sizes, int k, int64_t n,
                                 int64_t *touched_cnt) {
    if (p1_done(ctl)) return;
    if (ctl->labelled >= n) { ctl->done = 1; return; }
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
__global__ void k_p1_fill(GraphDev g, int32_t *L, RecDev fcur, unsigned long long *sizes,
                          P1Ctl *ctl) {
    if (p1_done(ctl)) return;
    const int b = ctl->fill_part;
    if (b < 0) return;
    int64_t added = 0;
    for (int64_t b0 = global_wave() * kWave; b0 < g.n; b0 += wave_count() * kWave) {
        const int64_t v = b0 + lane_id();
        const bool un = v < g.n && L[v] == -1;
        const int64_t pos = wave_append(un, fcur.cnt);
        if (un) {
            L[v] = b;
            fcur.x[pos] = (int32_t)v;
            fcur.g[pos] = b;
            fcur.d[pos] = (int32_t)(g.off[v + 1] - g.off[v]);
            ++added;
        }
    }
    added = wave_sum(added);
    if (lane_id() == 0 && added) {
        atomicAdd(&sizes[b], (unsigned long long)added);
        atomicAdd((unsigned long long *)&ctl->labelled, (unsigned long long)added);
    }
}

__device__ __forceinline__ void claim(int32_t v, int32_t p, int32_t *claims, int32_t *touched,
                                      int64_t *touched_cnt, const int32_t *L, bool pred) {
    bool first = false;
    if (pred && L[v] == -1) {
        if (claims[v] > p) first = atomicMin(&claims[v], p) == kNoClaim;
    }
    const int64_t pos = wave_append(first, (unsigned long long *)touched_cnt);
    if (first) touched[pos] = v;
}

// partition.cpp:451-476: claims by the selected frontier members.
__global__ __launch_bounds__(kThreads) void k_expand_light(GraphDev g, ListDev sel, const int32_t *L,
                                                           int32_t *claims, int32_t *touched,
                                                           int64_t *touched_cnt, const P1Ctl *ctl) {
    if (p1_done(ctl)) return;
    __shared__ BatchLds s_b[kWavesPerBlock];
    const int w = threadIdx.x / kWave, lane = lane_id();
    const int64_t count = (int64_t)sel.cnt[LC_LIGHT];
    for (int64_t b0 = global_wave() * kWave; b0 < count; b0 += wave_count() * kWave) {
        int32_t v, deg;
        bool valid;
        const int32_t total = batch_load(g, sel.light, count, b0, s_b[w], v, deg, valid);
        s_b[w].aux[lane] = valid ? L[v] : 0;
        wave_sync();
        for (int32_t q0 = 0; q0 < total; q0 += kWave) {
            const int32_t q = q0 + lane;
            const bool on = q < total;
            int32_t u = 0, p = 0;
            if (on) {
                const int slot = batch_slot(s_b[w], q);
                u = g.col[s_b[w].base[slot] + q];
                p = s_b[w].aux[slot];
            }
            claim(u, p, claims, touched, touched_cnt, L, on);
        }
        wave_sync();
    }
}

__global__ void k_expand_heavy(GraphDev g, ListDev sel, int chunk_edges, const int32_t *L,
                               int32_t *claims, int32_t *touched, int64_t *touched_cnt,
                               const P1Ctl *ctl) {
    if (p1_done(ctl)) return;
    const int64_t nch = (int64_t)sel.cnt[LC_CHUNKS];
    for (int64_t j = global_wave(); j < nch; j += wave_count()) {
        int32_t slot, v;
        int64_t e0, e1, deg;
        chunk_bounds(g, sel, sel.hc[j], chunk_edges, slot, v, e0, e1, deg);
        const int32_t p = L[v];
        for (int64_t eb = e0; eb < e1; eb += kWave) {
            const int64_t e = eb + lane_id();
            const bool on = e < e1;
            claim(on ? g.col[e] : 0, p, claims, touched, touched_cnt, L, on);
        }
    }
}

// partition.cpp:539-597: label the claimed vertices, extend the frontiers,
// recount sizes.
__global__ __launch_bounds__(kThreads) void k_p1_apply(GraphDev g, int32_t *L, int32_t *claims,
                                                       const int32_t *touched,
                                                       const int64_t *touched_cnt, RecDev fnext,
                                                       unsigned long long *sizes, int k,
                                                       P1Ctl *ctl) {
    if (p1_done(ctl)) return;
    __shared__ unsigned long long s_sz[DMOLP_MAX_K];
    for (int i = threadIdx.x; i < k; i += blockDim.x) s_sz[i] = 0;
    __syncthreads();
    const int64_t count = *touched_cnt;
    for (int64_t b0 = global_wave() * kWave; b0 < count; b0 += wave_count() * kWave) {
        const int64_t i = b0 + lane_id();
        const bool on = i < count;
        int32_t v = 0, p = 0;
        if (on) {
            v = touched[i];
            p = claims[v];
            claims[v] = kNoClaim;
            L[v] = p;
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
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd((unsigned long long *)&ctl->labelled, (unsigned long long)count);
}

__global__ void k_p1_round_end(P1Ctl *ctl) {
    if (p1_done(ctl)) return;
    ctl->rounds += 1;
}

// random.cpp:71-74: label = global id mod k.
__global__ void k_mod_labels(int64_t n, int k, int32_t *L) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (int64_t)gridDim.x * blockDim.x)
        L[v] = (int32_t)(v % k);
}

__global__ void k_degrees(GraphDev g, int32_t *deg) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
         v += (int64_t)gridDim.x * blockDim.x)
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

// init.cpp:139-177 at one rank: (id, degree) pairs in the degree map's
// iteration order (descending id), libstdc++ std::sort by degree, Lorenz
// tangent cut.  Double arithmetic in source order.
std::vector<int32_t> find_hubs_host(const std::vector<int32_t> &deg) {
    const int64_t n = (int64_t)deg.size();
    std::vector<std::pair<int, int>> sd;
    sd.reserve((size_t)n);
    for (int64_t v = n - 1; v >= 0; --v) sd.emplace_back((int)v, deg[(size_t)v]);
    std::sort(sd.begin(), sd.end(),
              [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.second < b.second; });
    std::vector<int32_t> hubs;
    const int N = (int)sd.size();
    if (N < 2) return hubs;
    std::vector<double> cum((size_t)N);
    double total = 0.0, run = 0.0;
    for (int i = 0; i < N; ++i) {
        const int d = sd[(size_t)i].second;
        total += d;
        run += d;
        cum[(size_t)i] = run;
    }
    for (int i = 0; i < N; ++i) cum[(size_t)i] /= total;
    const double inv_N = 1.0 / N;
    const double x1 = N * inv_N;
    const double y1 = cum[(size_t)N - 1];
    const double x0 = (N - 1) * inv_N;
    const double y0 = cum[(size_t)N - 2];
    const double slope = (y1 - y0) / (x1 - x0);
    const double xi = x1 - y1 / slope;
    for (int i = N - 1; i >= 0; --i) {
        const double x = (i + 1) * inv_N;
        if (x > xi) hubs.push_back(sd[(size_t)i].first);
        else break;
    }
    return hubs;
}

// init.cpp:179-187 over the degree map's order (descending id); the tie
// test compares an id with the running best degree.
std::pair<int32_t, int32_t> first_seed_host(const std::vector<int32_t> &deg) {
    int64_t node = -1, best = -1;
    for (int64_t v = (int64_t)deg.size() - 1; v >= 0; --v) {
        const int64_t d = deg[(size_t)v];
        if (d > best || (d == best && v < best)) { node = v; best = d; }
    }
    return {(int32_t)node, (int32_t)best};
}

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

void phase1_run(dmolp_ctx *ctx, dmolp_graph *g, int k, int theta, std::vector<int32_t> &seeds_out,
                dmolp_phase1_metrics *m) {
    DMOLP_REQUIRE(k >= 1 && k <= DMOLP_MAX_K, DMOLP_E_ARG, "k outside [1, 1024]");
    DMOLP_REQUIRE(theta > 0, DMOLP_E_ARG, "theta must be positive (theta <= 0 never terminates)");
    DMOLP_REQUIRE((int64_t)k <= g->n, DMOLP_E_SEEDS, "k larger than the vertex count");
    Phase1Work &w = phase1_work(ctx, g, k);
    hipStream_t s = ctx->stream;
    const GraphDev gd = g->dev();
    const int H = ctx->tune.heavy_degree, CH = ctx->tune.chunk_edges;
    const unsigned grid = (unsigned)ctx->grid_blocks();
    if (!g->labels.p) g->labels.alloc(sizeof(int32_t) * (size_t)g->n);
    int32_t *L = g->labels.as<int32_t>();
    P1Ctl *ctl = w.ctl.as<P1Ctl>();

    DMOLP_HIP_CHECK(hipEventRecord(ctx->ev0, s));
    // Degrees to the host for the hub sort.
    hipLaunchKernelGGL(k_degrees, dim3(ctx->cu_count * 4), dim3(256), 0, s, gd, w.deg.as<int32_t>());
    std::vector<int32_t> deg((size_t)g->n);
    DMOLP_HIP_CHECK(hipMemcpyAsync(deg.data(), w.deg.p, sizeof(int32_t) * (size_t)g->n,
                                   hipMemcpyDeviceToHost, s));
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    const auto th0 = std::chrono::steady_clock::now();
    const std::pair<int32_t, int32_t> fs = first_seed_host(deg);
    const std::vector<int32_t> hubs = find_hubs_host(deg);
    const auto th1 = std::chrono::steady_clock::now();
    const int64_t nh = (int64_t)hubs.size();

    // Seeds (seed.cpp:192-242).
    DMOLP_HIP_CHECK(hipMemsetAsync(w.ctl.p, 0, sizeof(P1Ctl), s));
    DMOLP_HIP_CHECK(hipMemsetAsync(w.covered.p, 0, w.covered.bytes, s));
    DMOLP_HIP_CHECK(hipMemsetAsync(w.best.p, 0, w.best.bytes, s));
    DMOLP_HIP_CHECK(hipMemsetAsync(w.seeds.p, 0xff, w.seeds.bytes, s));
    DMOLP_HIP_CHECK(hipMemcpyAsync(w.seeds.p, &fs.first, sizeof(int32_t), hipMemcpyHostToDevice, s));
    DevMem hubs_d(sizeof(int32_t) * (size_t)(nh > 0 ? nh : 1));
    DevMem used_d((size_t)(nh > 0 ? nh : 1));
    std::vector<uint8_t> used((size_t)(nh > 0 ? nh : 1), 0);
    for (int64_t i = 0; i < nh; ++i)
        if (hubs[(size_t)i] == fs.first) { used[(size_t)i] = 1; break; }
    if (nh > 0) {
        DMOLP_HIP_CHECK(hipMemcpyAsync(hubs_d.p, hubs.data(), sizeof(int32_t) * (size_t)nh,
                                       hipMemcpyHostToDevice, s));
        DMOLP_HIP_CHECK(hipMemcpyAsync(used_d.p, used.data(), (size_t)nh, hipMemcpyHostToDevice, s));
        DMOLP_HIP_CHECK(hipMemsetAsync(w.lvl1.cnt.p, 0, w.lvl1.cnt.bytes, s));
        for (int j = 1; j < k; ++j) {
            // Ball of seed j-1, then the next seed.
            hipLaunchKernelGGL(k_ball1, dim3(grid), dim3(kThreads), 0, s, gd, w.seeds.as<int32_t>(),
                               j - 1, w.covered.as<uint32_t>(), w.lvl1.dev, H, CH, ctl);
            hipLaunchKernelGGL(k_ball2_light, dim3(grid), dim3(kThreads), 0, s, gd, w.lvl1.dev,
                               w.covered.as<uint32_t>(), ctl);
            if (g->n_heavy > 0)
                hipLaunchKernelGGL(k_ball2_heavy, dim3(grid), dim3(kThreads), 0, s, gd, w.lvl1.dev,
                                   CH, w.covered.as<uint32_t>(), ctl);
            hipLaunchKernelGGL(k_next_seed, dim3(ctx->cu_count * 2), dim3(kThreads), 0, s, gd,
                               hubs_d.as<int32_t>(), nh, used_d.as<uint8_t>(),
                               w.covered.as<uint32_t>(), w.best.as<unsigned long long>(), ctl);
            hipLaunchKernelGGL(k_take_seed, dim3(1), dim3(1), 0, s, w.best.as<unsigned long long>(),
                               hubs_d.as<int32_t>(), used_d.as<uint8_t>(), w.seeds.as<int32_t>(), j,
                               ctl, w.lvl1.dev);
        }
        DMOLP_HIP_CHECK(hipGetLastError());
    }
    std::vector<int32_t> seeds((size_t)k, -1);
    P1Ctl hc{};
    DMOLP_HIP_CHECK(hipMemcpyAsync(seeds.data(), w.seeds.p, sizeof(int32_t) * (size_t)k,
                                   hipMemcpyDeviceToHost, s));
    DMOLP_HIP_CHECK(hipMemcpyAsync(&hc, w.ctl.p, sizeof(P1Ctl), hipMemcpyDeviceToHost, s));
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    int32_t found = (nh > 0) ? (hc.seed_stop ? hc.seeds_found : k) : 1;
    if (k == 1) found = 1;
    int32_t fallback = 0;
    if (found < k) {
        // Documented extension for seed.cpp:214-217 followed by the
        // out-of-range seeds[p] of partition.cpp:347: the missing seeds are
        // the remaining vertices by (degree descending, id ascending).
        std::vector<uint8_t> chosen((size_t)g->n, 0);
        for (int j = 0; j < found; ++j) chosen[(size_t)seeds[(size_t)j]] = 1;
        std::vector<int32_t> order;
        order.reserve((size_t)g->n);
        for (int64_t v = 0; v < g->n; ++v)
            if (!chosen[(size_t)v]) order.push_back((int32_t)v);
        const size_t need = (size_t)(k - found);
        auto cmp = [&](int32_t a, int32_t b) {
            if (deg[(size_t)a] != deg[(size_t)b]) return deg[(size_t)a] > deg[(size_t)b];
            return a < b;
        };
        std::partial_sort(order.begin(), order.begin() + (ptrdiff_t)need, order.end(), cmp);
        for (size_t i = 0; i < need; ++i) seeds[(size_t)found + i] = order[i];
        fallback = k - found;
        DMOLP_HIP_CHECK(hipMemcpyAsync(w.seeds.p, seeds.data(), sizeof(int32_t) * (size_t)k,
                                       hipMemcpyHostToDevice, s));
    }

    // Expansion (partition.cpp:323-604).
    DMOLP_HIP_CHECK(hipMemsetAsync(L, 0xff, sizeof(int32_t) * (size_t)g->n, s));
    DMOLP_HIP_CHECK(hipMemsetAsync(w.sizes.p, 0, w.sizes.bytes, s));
    DMOLP_HIP_CHECK(hipMemsetAsync(w.ctl.p, 0, sizeof(P1Ctl), s));
    hipLaunchKernelGGL(k_expand_init, dim3((k + 255) / 256), dim3(256), 0, s, gd,
                       w.seeds.as<int32_t>(), k, L, w.F[0].dev, w.sizes.as<unsigned long long>(), ctl);
    TakeRule rule{};
    rule.kind = 1;
    rule.theta = theta;
    int64_t enq = 0;
    const int batch = 16;
    for (;;) {
        for (int b = 0; b < batch; ++b, ++enq) {
            Recs &fcur = w.F[enq & 1];
            Recs &fnext = w.F[(enq + 1) & 1];
            hipLaunchKernelGGL(k_p1_round_begin, dim3(1), dim3(1), 0, s, ctl, fcur.dev, fnext.dev,
                               w.sel_list.dev, w.sizes.as<unsigned long long>(), k, g->n,
                               w.touched_cnt.as<int64_t>());
            hipLaunchKernelGGL(k_p1_fill, dim3(grid), dim3(kThreads), 0, s, gd, L, fcur.dev,
                               w.sizes.as<unsigned long long>(), ctl);
            select_enqueue(ctx, w.sel, gd, fcur.dev, rule, &ctl->done, w.sel_list.dev, &fnext.dev);
            hipLaunchKernelGGL(k_expand_light, dim3(grid), dim3(kThreads), 0, s, gd, w.sel_list.dev,
                               L, w.claims.as<int32_t>(), w.touched.as<int32_t>(),
                               w.touched_cnt.as<int64_t>(), ctl);
            if (g->n_heavy > 0)
                hipLaunchKernelGGL(k_expand_heavy, dim3(grid), dim3(kThreads), 0, s, gd,
                                   w.sel_list.dev, CH, L, w.claims.as<int32_t>(),
                                   w.touched.as<int32_t>(), w.touched_cnt.as<int64_t>(), ctl);
            hipLaunchKernelGGL(k_p1_apply, dim3(grid), dim3(kThreads), 0, s, gd, L,
                               w.claims.as<int32_t>(), w.touched.as<int32_t>(),
                               w.touched_cnt.as<int64_t>(), fnext.dev,
                               w.sizes.as<unsigned long long>(), k, ctl);
            hipLaunchKernelGGL(k_p1_round_end, dim3(1), dim3(1), 0, s, ctl);
        }
        DMOLP_HIP_CHECK(hipGetLastError());
        DMOLP_HIP_CHECK(hipMemcpyAsync(&hc, w.ctl.p, sizeof(P1Ctl), hipMemcpyDeviceToHost, s));
        DMOLP_HIP_CHECK(hipStreamSynchronize(s));
        if (hc.done) break;
        DMOLP_REQUIRE(enq < 64ll * (g->n + 16), DMOLP_E_HIP, "phase-1 expansion did not finish");
    }
    DMOLP_HIP_CHECK(hipEventRecord(ctx->ev1, s));
    DMOLP_HIP_CHECK(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    DMOLP_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    seeds_out = seeds;
    if (m) {
        std::memset(m, 0, sizeof(*m));
        m->first_seed = fs.first;
        m->first_seed_degree = fs.second;
        m->n_hubs = nh;
        m->n_seeds_found = found;
        m->n_seeds_fallback = fallback;
        m->rounds = hc.rounds;
        m->fill_round = hc.fill_round;
        m->device_ms = ms;
        m->host_ms = std::chrono::duration<double, std::milli>(th1 - th0).count();
    }
    g->labels_valid = true;
}

}  // namespace dmolp
