// graph.hip -- device-resident CSR: upload, validation, degree binning,
// download, labels.
#include "rows.hpp"

namespace dmolp {

int64_t bits_for(int64_t v) {
    int64_t b = 0;
    while (b < 63 && (v >> b) != 0) ++b;
    return b;
}

namespace {

__global__ void k_widen_offsets(const int32_t *src, int64_t *dst, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (int64_t)src[i];
}

constexpr int kValidateShort = 64;  // rows up to this length are checked by one thread

__device__ __forceinline__ void check_entry(const GraphDev &g, int64_t v, int64_t e, int64_t s,
                                            unsigned long long &bad_col,
                                            unsigned long long &bad_ord,
                                            unsigned long long &loops) {
    const int32_t u = g.col[e];
    if (u < 0 || u >= g.n) ++bad_col;
    if (e > s && u <= g.col[e - 1]) ++bad_ord;
    if (u == v) ++loops;
}

// Flags: [0] bad offsets, [1] column out of range, [2] unsorted or
// duplicate, [3] self-loop; [4] max degree; [5] long rows deferred.
__global__ void k_validate(GraphDev g, unsigned long long *flags, int32_t *longrows) {
    unsigned long long bad_off = 0, bad_col = 0, bad_ord = 0, loops = 0, maxd = 0;
    for (int64_t b0 = global_wave() * kWave; b0 < g.n; b0 += wave_count() * kWave) {
        const int64_t v = b0 + lane_id();
        bool defer = false;
        if (v < g.n) {
            const int64_t s = g.off[v], t = g.off[v + 1];
            if (s > t || s < 0 || t > g.nnz) {
                ++bad_off;
            } else {
                if ((unsigned long long)(t - s) > maxd) maxd = (unsigned long long)(t - s);
                if (t - s <= kValidateShort) {
                    for (int64_t e = s; e < t; ++e) check_entry(g, v, e, s, bad_col, bad_ord, loops);
                } else {
                    defer = true;
                }
            }
        }
        const int64_t pos = wave_append(defer, &flags[5]);
        if (defer) longrows[pos] = (int32_t)v;
    }
    if (bad_off) atomicAdd(&flags[0], bad_off);
    if (bad_col) atomicAdd(&flags[1], bad_col);
    if (bad_ord) atomicAdd(&flags[2], bad_ord);
    if (loops) atomicAdd(&flags[3], loops);
    if (maxd) atomicMax(&flags[4], maxd);
}

// Long rows: one wave per row.
__global__ void k_validate_long(GraphDev g, unsigned long long *flags, const int32_t *longrows) {
    unsigned long long bad_col = 0, bad_ord = 0, loops = 0;
    const int64_t nl = (int64_t)flags[5];
    for (int64_t j = global_wave(); j < nl; j += wave_count()) {
        const int64_t v = longrows[j];
        const int64_t s = g.off[v], t = g.off[v + 1];
        for (int64_t e = s + lane_id(); e < t; e += kWave) check_entry(g, v, e, s, bad_col, bad_ord, loops);
    }
    if (bad_col) atomicAdd(&flags[1], bad_col);
    if (bad_ord) atomicAdd(&flags[2], bad_ord);
    if (loops) atomicAdd(&flags[3], loops);
}

__global__ void k_max_degree(GraphDev g, unsigned long long *out) {
    unsigned long long m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long d = (unsigned long long)(g.off[v + 1] - g.off[v]);
        m = d > m ? d : m;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_down(m, o, kWave);
        m = x > m ? x : m;
    }
    if (lane_id() == 0 && m) atomicMax(out, m);
}

// Every vertex into the light list or a heavy slot.  Each block takes
// kBinItems vertices a thread per step and reserves its light entries with
// one atomic (a counter hit once per wave by every wave serialises).
constexpr int kBinItems = 8;
__global__ __launch_bounds__(256) void k_bin_all(GraphDev g, ListDev out, int heavy_degree,
                                                 int chunk_edges) {
    __shared__ BlockAppendLds s_app;
    const int64_t span = (int64_t)blockDim.x * kBinItems;
    for (int64_t b0 = (int64_t)blockIdx.x * span; b0 < g.n; b0 += (int64_t)gridDim.x * span) {
        int64_t d[kBinItems];
        int32_t nl = 0;
        int64_t dsum = 0;
#pragma unroll
        for (int j = 0; j < kBinItems; ++j) {
            const int64_t v = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            const int64_t vc = v < g.n ? v : g.n - 1;  // loads issued together
            const int64_t dv = g.off[vc + 1] - g.off[vc];
            d[j] = v < g.n ? dv : -1;
            nl += (d[j] >= 0 && d[j] <= heavy_degree) ? 1 : 0;
            dsum += d[j] >= 0 ? d[j] : 0;
        }
        int64_t pos = block_reserve(nl, &out.cnt[LC_LIGHT], s_app, dsum, &out.cnt[LC_DEG]);
#pragma unroll
        for (int j = 0; j < kBinItems; ++j) {
            const int64_t v = b0 + (int64_t)j * blockDim.x + threadIdx.x;
            if (d[j] < 0) continue;
            if (d[j] <= heavy_degree) {
                out.light[pos++] = (int32_t)v;
            } else {
                const int64_t slot = (int64_t)atomicAdd(&out.cnt[LC_SLOTS], 1ull);
                const int64_t nch = (d[j] + chunk_edges - 1) / chunk_edges;
                const int64_t base = (int64_t)atomicAdd(&out.cnt[LC_CHUNKS], (unsigned long long)nch);
                out.hv[slot] = (int32_t)v;
                out.hn[slot] = (int32_t)nch;
                atomicAdd(&out.cnt[LC_HDEG], (unsigned long long)d[j]);
                for (int64_t c = 0; c < nch; ++c) out.hc[base + c] = (slot << 32) | c;
            }
        }
    }
}

__global__ void k_heavy_ids(ListDev all, int32_t *hid, uint32_t *hbits) {
    const int64_t ns = (int64_t)all.cnt[LC_SLOTS];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ns;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = all.hv[i];
        hid[v] = (int32_t)i;
        atomicOr(&hbits[v >> 5], 1u << (v & 31));
    }
}

__global__ void k_first_nb(GraphDev g, int32_t *first, int32_t *deg) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = g.off[v], e = g.off[v + 1];
        first[v] = s < e ? g.col[s] : -1;
        deg[v] = (int32_t)(e - s);
    }
}

__global__ void k_count_heavy(GraphDev g, int heavy_degree, int chunk_edges,
                              unsigned long long *acc) {
    unsigned long long nh = 0, nc = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = g.off[v + 1] - g.off[v];
        if (d > heavy_degree) { ++nh; nc += (unsigned long long)((d + chunk_edges - 1) / chunk_edges); }
    }
    if (nh) { atomicAdd(&acc[0], nh); atomicAdd(&acc[1], nc); }
}

}  // namespace

void graph_validate(dmolp_ctx *ctx, dmolp_graph *g) {
    DevMem flags(sizeof(unsigned long long) * 8);
    DevMem longrows(sizeof(int32_t) * (size_t)(g->n > 0 ? g->n : 1));
    DMOLP_HIP_CHECK(hipMemsetAsync(flags.p, 0, flags.bytes, ctx->stream));
    const int blocks = ctx->cu_count * 4;
    if (g->n > 0) {
        hipLaunchKernelGGL(k_validate, dim3(blocks), dim3(256), 0, ctx->stream, g->dev(),
                           flags.as<unsigned long long>(), longrows.as<int32_t>());
        hipLaunchKernelGGL(k_validate_long, dim3(blocks), dim3(256), 0, ctx->stream, g->dev(),
                           flags.as<unsigned long long>(), longrows.as<int32_t>());
    }
    DMOLP_HIP_CHECK(hipGetLastError());
    unsigned long long h[8] = {0};
    DMOLP_HIP_CHECK(hipMemcpyAsync(h, flags.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    DMOLP_REQUIRE(h[0] == 0, DMOLP_E_GRAPH, "CSR offsets are not monotone within [0, nnz]");
    DMOLP_REQUIRE(h[1] == 0, DMOLP_E_GRAPH, "CSR column outside [0, n)");
    DMOLP_REQUIRE(h[2] == 0, DMOLP_E_GRAPH, "CSR row not strictly ascending (unsorted or duplicate)");
    DMOLP_REQUIRE(h[3] == 0, DMOLP_E_GRAPH, "CSR has a self-loop");
    g->max_degree = (int32_t)h[4];
}

void graph_max_degree(dmolp_ctx *ctx, dmolp_graph *g) {
    DevMem acc(sizeof(unsigned long long));
    DMOLP_HIP_CHECK(hipMemsetAsync(acc.p, 0, acc.bytes, ctx->stream));
    if (g->n > 0)
        hipLaunchKernelGGL(k_max_degree, dim3(ctx->cu_count * 4), dim3(256), 0, ctx->stream, g->dev(),
                           acc.as<unsigned long long>());
    DMOLP_HIP_CHECK(hipGetLastError());
    unsigned long long h = 0;
    DMOLP_HIP_CHECK(hipMemcpyAsync(&h, acc.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    g->max_degree = (int32_t)h;
}

void graph_prepare_bins(dmolp_ctx *ctx, dmolp_graph *g) {
    const int H = ctx->tune.heavy_degree, CH = ctx->tune.chunk_edges;
    if (g->all_heavy_degree == H && g->all_chunk_edges == CH) return;
    DevMem acc(sizeof(unsigned long long) * 2);
    DMOLP_HIP_CHECK(hipMemsetAsync(acc.p, 0, acc.bytes, ctx->stream));
    hipLaunchKernelGGL(k_count_heavy, dim3(ctx->cu_count * 4), dim3(256), 0, ctx->stream, g->dev(),
                       H, CH, acc.as<unsigned long long>());
    DMOLP_HIP_CHECK(hipGetLastError());
    unsigned long long h[2];
    DMOLP_HIP_CHECK(hipMemcpyAsync(h, acc.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    g->n_heavy = (int64_t)h[0];
    g->n_heavy_chunks = (int64_t)h[1];
    g->all.alloc(g->n, g->n_heavy, g->n_heavy_chunks);
    DMOLP_HIP_CHECK(hipMemsetAsync(g->all.cnt.p, 0, g->all.cnt.bytes, ctx->stream));
    hipLaunchKernelGGL(k_bin_all, dim3(ctx->grid_blocks()), dim3(256), 0, ctx->stream, g->dev(),
                       g->all.dev, H, CH);
    g->hid.grow(sizeof(int32_t) * (size_t)(g->n > 0 ? g->n : 1));
    DMOLP_HIP_CHECK(hipMemsetAsync(g->hid.p, 0xff, g->hid.bytes, ctx->stream));
    g->hbits.grow(sizeof(uint32_t) * (size_t)((g->n + 31) / 32 > 0 ? (g->n + 31) / 32 : 1));
    DMOLP_HIP_CHECK(hipMemsetAsync(g->hbits.p, 0, g->hbits.bytes, ctx->stream));
    hipLaunchKernelGGL(k_heavy_ids, dim3(ctx->cu_count), dim3(256), 0, ctx->stream, g->all.dev,
                       g->hid.as<int32_t>(), g->hbits.as<uint32_t>());
    DMOLP_HIP_CHECK(hipGetLastError());
    g->all_heavy_degree = H;
    g->all_chunk_edges = CH;
}

void graph_prepare_first(dmolp_ctx *ctx, dmolp_graph *g) {
    if (g->first_valid) return;
    g->first.grow(sizeof(int32_t) * (size_t)(g->n > 0 ? g->n : 1));
    g->deg.grow(sizeof(int32_t) * (size_t)(g->n > 0 ? g->n : 1));
    if (g->n > 0) {
        hipLaunchKernelGGL(k_first_nb, dim3(ctx->cu_count * 8), dim3(256), 0, ctx->stream, g->dev(),
                           g->first.as<int32_t>(), g->deg.as<int32_t>());
        DMOLP_HIP_CHECK(hipGetLastError());
    }
    g->first_valid = true;
}

void graph_widen_offsets(dmolp_ctx *ctx, const int32_t *d_src, int64_t *d_dst, int64_t count) {
    hipLaunchKernelGGL(k_widen_offsets, dim3(ctx->cu_count * 4), dim3(256), 0, ctx->stream, d_src,
                       d_dst, count);
    DMOLP_HIP_CHECK(hipGetLastError());
}

}  // namespace dmolp
