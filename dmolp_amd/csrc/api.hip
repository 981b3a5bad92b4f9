// api.hip -- the C ABI of include/dmolp_hip.h: contexts, graphs, phase 1,
// phase 2 and the test hooks.  Every entry point catches library errors
// and returns a status; nothing throws across the boundary.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gen_host.hpp"
#include "internal.hpp"
#include "phase1.hpp"
#include "phase2.hpp"
#include "batch.hpp"
#include "halo.hpp"
#include "noalloc.hpp"
#include "rows.hpp"

namespace dmolp {
void gen_device(dmolp_ctx *ctx, const dmolp_gen_spec *spec, dmolp_graph *g);
void graph_widen_offsets(dmolp_ctx *ctx, const int32_t *d_src, int64_t *d_dst, int64_t count);
void phase2_penalties_device(dmolp_ctx *ctx, int k, const unsigned long long *d_vc,
                             const unsigned long long *d_ic, int t, double *d_pen, double *d_re);
void phase2_select_enqueue(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w, int t);
void phase2_lp_enqueue(dmolp_ctx *ctx, dmolp_graph *g, Phase2Work &w);

namespace {

__global__ void k_bin_ids(GraphDev g, const int32_t *ids, int64_t count, ListDev out, int H,
                          int CH, unsigned long long *bad) {
    for (int64_t b0 = global_wave() * kWave; b0 < count; b0 += wave_count() * kWave) {
        const int64_t i = b0 + lane_id();
        bool valid = i < count;
        int32_t v = valid ? ids[i] : 0;
        if (valid && (v < 0 || v >= g.n)) { atomicAdd(bad, 1ull); valid = false; }
        const int64_t d = valid ? g.off[v + 1] - g.off[v] : 0;
        list_append(out, valid, v, d, H, CH);
    }
}

thread_local std::string g_tls_error;

int fail(dmolp_ctx *ctx, int status, const std::string &msg) {
    if (ctx) ctx->last_error = msg;
    g_tls_error = msg;
    return status;
}

template <typename F>
int guarded(dmolp_ctx *ctx, F &&f) {
    // Device allocations of the call come from (and go back to) ctx's pool.
    PoolScope scope(ctx ? &ctx->pool : tls_pool(), ctx ? ctx->stream : tls_stream());
    try {
        f();
        return DMOLP_OK;
    } catch (const Error &e) {
        return fail(ctx, e.status, e.what());
    } catch (const std::bad_alloc &) {
        return fail(ctx, DMOLP_E_NOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(ctx, DMOLP_E_HIP, e.what());
    }
}

// Upload an id list into a binned vertex list.
void upload_list(dmolp_ctx *ctx, dmolp_graph *g, const int32_t *ids, int64_t count, List &L) {
    hipStream_t s = ctx->stream;
    DMOLP_HIP_CHECK(hipMemsetAsync(L.cnt.p, 0, L.cnt.bytes, s));
    if (count == 0) return;
    DevMem d(sizeof(int32_t) * (size_t)count);
    DevMem bad(sizeof(unsigned long long));
    DMOLP_HIP_CHECK(hipMemsetAsync(bad.p, 0, bad.bytes, s));
    DMOLP_HIP_CHECK(hipMemcpyAsync(d.p, ids, d.bytes, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_bin_ids, dim3(ctx->grid_blocks()), dim3(256), 0, s, g->dev(),
                       d.as<int32_t>(), count, L.dev, ctx->tune.heavy_degree, ctx->tune.chunk_edges,
                       bad.as<unsigned long long>());
    DMOLP_HIP_CHECK(hipGetLastError());
    unsigned long long nb = 0;
    DMOLP_HIP_CHECK(hipMemcpyAsync(&nb, bad.p, sizeof(nb), hipMemcpyDeviceToHost, s));
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    DMOLP_REQUIRE(nb == 0, DMOLP_E_ARG, "vertex id outside [0, n) in list");
}

// Download a binned list as sorted ids.
std::vector<int32_t> download_list(dmolp_ctx *ctx, const List &L) {
    hipStream_t s = ctx->stream;
    unsigned long long c[8];
    DMOLP_HIP_CHECK(hipMemcpyAsync(c, L.cnt.p, sizeof(c), hipMemcpyDeviceToHost, s));
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<int32_t> out((size_t)(c[LC_LIGHT] + c[LC_SLOTS]));
    if (c[LC_LIGHT])
        DMOLP_HIP_CHECK(hipMemcpyAsync(out.data(), L.light.p, sizeof(int32_t) * c[LC_LIGHT],
                                       hipMemcpyDeviceToHost, s));
    if (c[LC_SLOTS])
        DMOLP_HIP_CHECK(hipMemcpyAsync(out.data() + c[LC_LIGHT], L.hv.p,
                                       sizeof(int32_t) * c[LC_SLOTS], hipMemcpyDeviceToHost, s));
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    std::sort(out.begin(), out.end());
    return out;
}

void set_labels(dmolp_ctx *ctx, dmolp_graph *g, const int32_t *labels) {
    if (!g->labels.p) g->labels.alloc(sizeof(int32_t) * (size_t)g->n);
    if (g->n > 0)
        DMOLP_HIP_CHECK(hipMemcpyAsync(g->labels.p, labels, sizeof(int32_t) * (size_t)g->n,
                                       hipMemcpyHostToDevice, ctx->stream));
    g->labels_valid = true;
}

void get_labels(dmolp_ctx *ctx, const dmolp_graph *g, int32_t *labels) {
    DMOLP_REQUIRE(g->labels_valid, DMOLP_E_STATE, "graph has no resident labels");
    if (g->n > 0)
        DMOLP_HIP_CHECK(hipMemcpyAsync(labels, g->labels.p, sizeof(int32_t) * (size_t)g->n,
                                       hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

// Full stats sweep on the resident labels: vc, ic, cut (phase2.cpp:55-117,
// 214-253).  Returns the device error code (0 when fine).
int stats_sweep(dmolp_ctx *ctx, dmolp_graph *g, int k, std::vector<int64_t> &vc,
                std::vector<int64_t> &ic, int64_t &cut) {
    Phase2Work &w = phase2_work(ctx, g, k);
    phase2_init_enqueue(ctx, phase2_single_set(ctx, g, w, 0, 0.0));
    Ctl c{};
    vc.assign((size_t)k, 0);
    ic.assign((size_t)k, 0);
    DMOLP_HIP_CHECK(hipMemcpyAsync(&c, w.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipMemcpyAsync(vc.data(), w.vc.p, sizeof(int64_t) * (size_t)k,
                                   hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipMemcpyAsync(ic.data(), w.ic.p, sizeof(int64_t) * (size_t)k,
                                   hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    cut = (int64_t)c.v[C_CUT];
    return c.err;
}

}  // namespace
}  // namespace dmolp

using namespace dmolp;

dmolp_graph::~dmolp_graph() = default;

hipEvent_t dmolp_ctx::prof_event() {
    if (ev_used == ev_pool.size()) {
        hipEvent_t e;
        DMOLP_HIP_CHECK(hipEventCreate(&e));
        ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
}

namespace dmolp {
bool pool_poison() {
    static const bool on = [] {
        const char *e = std::getenv("DMOLP_POOL_POISON");
        return e && e[0] == '1';
    }();
    return on;
}

bool graphs_enabled() {
    static const bool on = [] {
        const char *e = std::getenv("DMOLP_NO_GRAPHS");
        return !(e && e[0] == '1');
    }();
    return on;
}

bool round_poll(dmolp_ctx *ctx, const int32_t *d_flag, int t, bool stop_at_zero) {
    const int slot = t & 1;
    DMOLP_HIP_CHECK(hipMemcpyAsync(&ctx->poll_host[slot], d_flag, sizeof(int32_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
    DMOLP_HIP_CHECK(hipEventRecord(ctx->poll_ev[slot], ctx->stream));
    if (t < 2) return false;
    DMOLP_HIP_CHECK(hipEventSynchronize(ctx->poll_ev[slot ^ 1]));
    const int32_t v = ctx->poll_host[slot ^ 1];
    return stop_at_zero ? v <= 0 : v != 0;
}
}  // namespace dmolp

extern "C" {

int dmolp_abi_version(void) { return DMOLP_ABI_VERSION; }

const char *dmolp_status_string(int status) {
    switch (status) {
        case DMOLP_OK: return "ok";
        case DMOLP_E_ARG: return "invalid argument";
        case DMOLP_E_SEEDS: return "k larger than the vertex count";
        case DMOLP_E_INVALID_LABEL: return "a -1 label reached the cut";
        case DMOLP_E_ODD_CUT: return "odd edge cut (asymmetric CSR)";
        case DMOLP_E_HIP: return "HIP runtime failure";
        case DMOLP_E_NOMEM: return "out of memory";
        case DMOLP_E_UNSUPPORTED: return "unsupported request";
        case DMOLP_E_STATE: return "call out of order";
        case DMOLP_E_GRAPH: return "CSR failed validation";
        case DMOLP_E_INTERNAL: return "device consistency check failed";
        default: return "unknown status";
    }
}

const char *dmolp_last_error(const dmolp_ctx *ctx) {
    if (ctx) return ctx->last_error.c_str();
    return g_tls_error.c_str();
}

int dmolp_ctx_create(int device, dmolp_ctx **out) {
    if (!out) return fail(nullptr, DMOLP_E_ARG, "out is null");
    *out = nullptr;
    dmolp_ctx *ctx = new (std::nothrow) dmolp_ctx();
    if (!ctx) return fail(nullptr, DMOLP_E_NOMEM, "host allocation failed");
    int rc = guarded(ctx, [&] {
        int count = 0;
        DMOLP_HIP_CHECK(hipGetDeviceCount(&count));
        DMOLP_REQUIRE(device >= 0 && device < count, DMOLP_E_ARG, "device index out of range");
        DMOLP_HIP_CHECK(hipSetDevice(device));
        hipDeviceProp_t prop;
        DMOLP_HIP_CHECK(hipGetDeviceProperties(&prop, device));
        ctx->device = device;
        ctx->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        // Up to a quarter of the device's memory stays cached for reuse.
        ctx->pool = std::make_shared<DevPool>((size_t)prop.totalGlobalMem / 4);
        DMOLP_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        DMOLP_HIP_CHECK(hipEventCreate(&ctx->ev0));
        DMOLP_HIP_CHECK(hipEventCreate(&ctx->ev1));
        DMOLP_HIP_CHECK(hipHostMalloc((void **)&ctx->poll_host, 64, hipHostMallocDefault));
        for (hipEvent_t &e : ctx->poll_ev) DMOLP_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        noalloc_init_device(device);  // the no-allocation scratch's address on this device
    });
    if (rc != DMOLP_OK) {
        g_tls_error = ctx->last_error;
        delete ctx;
        return rc;
    }
    *out = ctx;
    return DMOLP_OK;
}

int dmolp_ctx_destroy(dmolp_ctx *ctx) {
    if (!ctx) return DMOLP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->pool) ctx->pool->close();  // memory still held elsewhere is freed when released
    ctx->scratch.release();
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (hipEvent_t e : ctx->poll_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->poll_host) (void)hipHostFree(ctx->poll_host);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    delete ctx;
    return DMOLP_OK;
}

void *dmolp_ctx_stream(dmolp_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }


int dmolp_ctx_set_profiling(dmolp_ctx *ctx, int on) {
    if (!ctx) return fail(nullptr, DMOLP_E_ARG, "null context");
    ctx->profiling = on != 0;
    return DMOLP_OK;
}

int dmolp_ctx_round_times(dmolp_ctx *ctx, int group, double *ms_out, int cap, int *n_out) {
    if (!ctx || group < 0 || group > DMOLP_PROF_INIT_LIGHT || cap < 0 || (cap > 0 && !ms_out))
        return fail(ctx, DMOLP_E_ARG, "bad arguments");
    const std::vector<double> empty;
    const std::vector<double> &v =
        (size_t)group < ctx->prof_ms.size() ? ctx->prof_ms[(size_t)group] : empty;
    const int n = (int)v.size();
    for (int i = 0; i < n && i < cap; ++i) ms_out[i] = v[(size_t)i];
    if (n_out) *n_out = n;
    return DMOLP_OK;
}

int dmolp_ctx_get_tuning(const dmolp_ctx *ctx, dmolp_tuning *out) {
    if (!ctx || !out) return fail(nullptr, DMOLP_E_ARG, "null argument");
    out->heavy_degree = ctx->tune.heavy_degree;
    out->chunk_edges = ctx->tune.chunk_edges;
    out->block_threads = ctx->tune.block_threads;
    out->grid_blocks = ctx->grid_blocks();
    out->pull_per_mille = (int32_t)(ctx->tune.pull_fraction * 1000.0 + 0.5);
    out->batch_work_per_block = (int32_t)ctx->tune.batch_work_per_block;
    out->hist_recount_per_mille = (int32_t)(ctx->tune.hist_recount_share * 1000.0 + 0.5);
    out->hist_push_per_mille = (int32_t)(ctx->tune.hist_push_share * 1000.0 + 0.5);
    out->light_hist = ctx->tune.light_hist ? 1 : 2;
    return DMOLP_OK;
}

int dmolp_ctx_set_tuning(dmolp_ctx *ctx, const dmolp_tuning *t) {
    if (!ctx || !t) return fail(ctx, DMOLP_E_ARG, "null argument");
    // Light rows keep 16-bit per-row counters: heavy_degree stays below 2^16.
    // A lane of the heavy initial sweep counts up to chunk_edges / 64 entries
    // in 16-bit counters: chunk_edges stays at most 2^21.
    if (t->heavy_degree < 0 || t->heavy_degree > 65535 || t->chunk_edges < 0 ||
        t->chunk_edges > (1 << 21) ||
        t->grid_blocks < 0 || t->pull_per_mille < 0 || t->pull_per_mille > 1000000 ||
        t->batch_work_per_block < 0 || t->hist_recount_per_mille < 0 ||
        t->hist_recount_per_mille > 1001 || t->hist_push_per_mille < 0 ||
        t->hist_push_per_mille > 1001 || t->light_hist < 0 || t->light_hist > 2 ||
        (t->block_threads != 0 && t->block_threads != 256))
        return fail(ctx, DMOLP_E_ARG, "tuning value out of range");
    if (t->heavy_degree) ctx->tune.heavy_degree = t->heavy_degree;
    if (t->chunk_edges) ctx->tune.chunk_edges = t->chunk_edges < 64 ? 64 : t->chunk_edges;
    if (t->grid_blocks) ctx->tune.grid_blocks = t->grid_blocks;
    if (t->pull_per_mille > 0) ctx->tune.pull_fraction = t->pull_per_mille / 1000.0;
    if (t->batch_work_per_block > 0) ctx->tune.batch_work_per_block = t->batch_work_per_block;
    // 0 keeps the current share; 1..1000 sets it; 1001 means "never" (above 1).
    if (t->hist_recount_per_mille > 0)
        ctx->tune.hist_recount_share = t->hist_recount_per_mille / 1000.0;
    if (t->hist_push_per_mille > 0) ctx->tune.hist_push_share = t->hist_push_per_mille / 1000.0;
    if (t->light_hist) ctx->tune.light_hist = t->light_hist == 1;
    return DMOLP_OK;
}

int dmolp_hubs_host(const int32_t *deg, int64_t n, int32_t *hubs_out, int64_t *n_hubs) {
    if (!n_hubs || n < 0 || n > 0x7fffffffll || (n > 0 && !deg)) return DMOLP_E_ARG;
    return guarded(nullptr, [&] {
        const std::vector<int32_t> d(deg, deg + n);
        const std::vector<int32_t> h = find_hubs_host(d, std::min(16, host_threads()));
        if (hubs_out) std::copy(h.begin(), h.end(), hubs_out);
        *n_hubs = (int64_t)h.size();
    });
}

int dmolp_first_seed_host(const int32_t *deg, int64_t n, int32_t *seed, int32_t *degree) {
    if (!seed || !degree || n < 0 || n > 0x7fffffffll || (n > 0 && !deg)) return DMOLP_E_ARG;
    return guarded(nullptr, [&] {
        const std::vector<int32_t> d(deg, deg + n);
        const std::pair<int32_t, int32_t> r = first_seed_host(d, std::min(16, host_threads()));
        *seed = r.first;
        *degree = r.second;
    });
}

int dmolp_ctx_set_hub_prefix(dmolp_ctx *ctx, int64_t prefix) {
    if (!ctx || prefix < 0) return fail(ctx, DMOLP_E_ARG, "bad arguments");
    ctx->hub_prefix = prefix;
    return DMOLP_OK;
}

int dmolp_hubs_device(dmolp_ctx *ctx, const int32_t *deg, int64_t n, int64_t limit, int64_t host_max,
                      int32_t *hubs_out, int64_t *n_hubs, int32_t *seed, int32_t *degree) {
    if (!ctx || !n_hubs || !seed || !degree || n < 0 || n > 0x7fffffffll || (n > 0 && !deg) ||
        host_max < 0)
        return fail(ctx, DMOLP_E_ARG, "bad arguments");
    return guarded(ctx, [&] {
        DevMem d(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        if (n > 0)
            DMOLP_HIP_CHECK(hipMemcpyAsync(d.p, deg, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice,
                                           ctx->stream));
        HubSearch hs;
        hs.n = n;
        hs.deg = d.as<int32_t>();
        hs.limit = limit < 0 ? -1 : limit;
        std::vector<HubSearch *> one{&hs};
        hub_search_device(ctx, one, host_max, std::min(16, host_threads()));
        if (hubs_out) std::copy(hs.hubs.begin(), hs.hubs.end(), hubs_out);
        *n_hubs = hs.n_hubs;
        *seed = hs.fs.first;
        *degree = hs.fs.second;
    });
}

int dmolp_top_degree_device(dmolp_ctx *ctx, const int32_t *deg, int64_t n, int64_t m, int64_t host_max,
                            int32_t *out) {
    if (!ctx || n < 0 || n > 0x7fffffffll || (n > 0 && (!deg || !out)) || m < 0 || host_max < 0)
        return fail(ctx, DMOLP_E_ARG, "bad arguments");
    return guarded(ctx, [&] {
        DevMem d(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        if (n > 0)
            DMOLP_HIP_CHECK(hipMemcpyAsync(d.p, deg, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice,
                                           ctx->stream));
        TopSearch ts;
        ts.n = n;
        ts.deg = d.as<int32_t>();
        ts.m = m;
        std::vector<TopSearch *> one{&ts};
        top_by_degree_device(ctx, one, host_max, std::min(16, host_threads()));
        std::copy(ts.top.begin(), ts.top.end(), out);
    });
}

int dmolp_ctx_set_hub_device(dmolp_ctx *ctx, int64_t dev_min, int64_t host_max) {
    if (!ctx || dev_min < 0 || host_max < 0) return fail(ctx, DMOLP_E_ARG, "bad arguments");
    ctx->hub_dev_min = dev_min;
    ctx->hub_host_max = host_max;
    return DMOLP_OK;
}

int dmolp_gen_host(const dmolp_gen_spec *spec, dmolp_csr *out) {
    return guarded(nullptr, [&] { gen_host(spec, out); });
}

void dmolp_csr_free(dmolp_csr *csr) {
    if (!csr) return;
    std::free(csr->off);
    std::free(csr->col);
    csr->off = nullptr;
    csr->col = nullptr;
    csr->n = csr->nnz = 0;
}

int dmolp_graph_upload(dmolp_ctx *ctx, int64_t n, int64_t nnz, const void *offsets,
                       int offset_bytes, const int32_t *cols, dmolp_graph **out) {
    if (!ctx || !out) return fail(ctx, DMOLP_E_ARG, "null argument");
    *out = nullptr;
    if (n <= 0 || n >= (1ll << 31) || nnz < 0 || !offsets || (nnz > 0 && !cols) ||
        (offset_bytes != 4 && offset_bytes != 8))
        return fail(ctx, DMOLP_E_ARG, "bad CSR arguments");
    if (offset_bytes == 4 && nnz >= (1ll << 31))
        return fail(ctx, DMOLP_E_ARG, "int32 offsets cannot address nnz >= 2^31");
    dmolp_graph *g = new (std::nothrow) dmolp_graph();
    if (!g) return fail(ctx, DMOLP_E_NOMEM, "host allocation failed");
    g->pool = ctx->pool.get();
    g->device = ctx->device;
    int rc = guarded(ctx, [&] {
        DMOLP_HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        g->n = n;
        g->nnz = nnz;
        g->off.alloc(sizeof(int64_t) * (size_t)(n + 1));
        g->col.alloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
        if (offset_bytes == 8) {
            DMOLP_HIP_CHECK(hipMemcpyAsync(g->off.p, offsets, sizeof(int64_t) * (size_t)(n + 1),
                                           hipMemcpyHostToDevice, s));
        } else {
            DevMem tmp(sizeof(int32_t) * (size_t)(n + 1));
            DMOLP_HIP_CHECK(hipMemcpyAsync(tmp.p, offsets, tmp.bytes, hipMemcpyHostToDevice, s));
            graph_widen_offsets(ctx, tmp.as<int32_t>(), g->off.as<int64_t>(), n + 1);
            DMOLP_HIP_CHECK(hipStreamSynchronize(s));
        }
        if (nnz > 0)
            DMOLP_HIP_CHECK(hipMemcpyAsync(g->col.p, cols, sizeof(int32_t) * (size_t)nnz,
                                           hipMemcpyHostToDevice, s));
        int64_t first = 0, last = 0;
        DMOLP_HIP_CHECK(hipMemcpyAsync(&first, g->off.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        DMOLP_HIP_CHECK(hipMemcpyAsync(&last, g->off.as<int64_t>() + n, sizeof(int64_t),
                                       hipMemcpyDeviceToHost, s));
        DMOLP_HIP_CHECK(hipStreamSynchronize(s));
        DMOLP_REQUIRE(first == 0 && last == nnz, DMOLP_E_GRAPH, "offsets must run from 0 to nnz");
        graph_validate(ctx, g);
    });
    if (rc != DMOLP_OK) { delete g; return rc; }
    *out = g;
    return DMOLP_OK;
}

int dmolp_graph_generate(dmolp_ctx *ctx, const dmolp_gen_spec *spec, dmolp_graph **out) {
    if (!ctx || !out || !spec) return fail(ctx, DMOLP_E_ARG, "null argument");
    *out = nullptr;
    dmolp_graph *g = new (std::nothrow) dmolp_graph();
    if (!g) return fail(ctx, DMOLP_E_NOMEM, "host allocation failed");
    g->pool = ctx->pool.get();
    g->device = ctx->device;
    int rc = guarded(ctx, [&] {
        DMOLP_HIP_CHECK(hipSetDevice(ctx->device));
        gen_device(ctx, spec, g);
        graph_max_degree(ctx, g);
    });
    if (rc != DMOLP_OK) { delete g; return rc; }
    *out = g;
    return DMOLP_OK;
}

int dmolp_graph_info(const dmolp_graph *g, int64_t *n, int64_t *nnz) {
    if (!g) return fail(nullptr, DMOLP_E_ARG, "null graph");
    if (n) *n = g->n;
    if (nnz) *nnz = g->nnz;
    return DMOLP_OK;
}

int dmolp_graph_bins(dmolp_ctx *ctx, dmolp_graph *g, int64_t *heavy_rows, int64_t *heavy_entries) {
    if (!ctx || !g) return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    return guarded(ctx, [&] {
        graph_prepare_bins(ctx, g);
        unsigned long long c[LC_N] = {};
        DMOLP_HIP_CHECK(hipMemcpyAsync(c, g->all.cnt.p, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
        DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (heavy_rows) *heavy_rows = (int64_t)c[LC_SLOTS];
        if (heavy_entries) *heavy_entries = (int64_t)c[LC_HDEG];
    });
}

int dmolp_graph_download(dmolp_ctx *ctx, const dmolp_graph *g, int64_t *off, int32_t *col) {
    if (!ctx || !g) return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    return guarded(ctx, [&] {
        if (off)
            DMOLP_HIP_CHECK(hipMemcpyAsync(off, g->off.p, sizeof(int64_t) * (size_t)(g->n + 1),
                                           hipMemcpyDeviceToHost, ctx->stream));
        if (col && g->nnz > 0)
            DMOLP_HIP_CHECK(hipMemcpyAsync(col, g->col.p, sizeof(int32_t) * (size_t)g->nnz,
                                           hipMemcpyDeviceToHost, ctx->stream));
        DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int dmolp_graph_free(dmolp_ctx *ctx, dmolp_graph *g) {
    if (!g) return DMOLP_OK;
    // The graph's memory goes back to its pool for reuse in stream order
    // (or is freed, if that pool's context is gone).  Work another context
    // queued on it must be finished first: then wait for the whole device.
    if (ctx && !g->mixed && g->pool == ctx->pool.get()) {
        (void)hipStreamSynchronize(ctx->stream);
    } else {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(g->device);
        (void)hipDeviceSynchronize();
        (void)hipSetDevice(cur);
    }
    delete g;
    return DMOLP_OK;
}

int dmolp_graph_set_labels(dmolp_ctx *ctx, dmolp_graph *g, const int32_t *labels) {
    if (!ctx || !g || !labels) return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    return guarded(ctx, [&] {
        set_labels(ctx, g, labels);
        DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int dmolp_graph_get_labels(dmolp_ctx *ctx, const dmolp_graph *g, int32_t *labels) {
    if (!ctx || !g || !labels) return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    return guarded(ctx, [&] { get_labels(ctx, g, labels); });
}

int dmolp_graph_labels_checkpoint(dmolp_ctx *ctx, dmolp_graph *g, int op) {
    if (!ctx || !g || (op != 0 && op != 1)) return fail(ctx, DMOLP_E_ARG, "bad arguments");
    note_use(ctx, g);
    return guarded(ctx, [&] {
        const size_t bytes = sizeof(int32_t) * (size_t)g->n;
        if (op == 0) {
            DMOLP_REQUIRE(g->labels_valid, DMOLP_E_STATE, "graph has no resident labels");
            if (!g->saved.p) g->saved.alloc(bytes);
            DMOLP_HIP_CHECK(hipMemcpyAsync(g->saved.p, g->labels.p, bytes, hipMemcpyDeviceToDevice,
                                           ctx->stream));
        } else {
            DMOLP_REQUIRE(g->saved.p, DMOLP_E_STATE, "no label checkpoint saved");
            DMOLP_HIP_CHECK(hipMemcpyAsync(g->labels.p, g->saved.p, bytes, hipMemcpyDeviceToDevice,
                                           ctx->stream));
            g->labels_valid = true;
        }
    });
}

int dmolp_phase1(dmolp_ctx *ctx, dmolp_graph *g, int32_t k, int32_t theta, int32_t *labels_out,
                 int32_t *seeds_out, int64_t *vc, int64_t *ec, dmolp_phase1_metrics *m) {
    if (!ctx || !g) return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    if (k <= 0 || k > DMOLP_MAX_K) return fail(ctx, DMOLP_E_ARG, "k outside [1, 1024]");
    if (theta <= 0) return fail(ctx, DMOLP_E_ARG, "theta must be positive");
    if ((int64_t)k > g->n) return fail(ctx, DMOLP_E_SEEDS, "k larger than the vertex count");
    return guarded(ctx, [&] {
        std::vector<int32_t> seeds;
        dmolp_phase1_metrics mm{};
        phase1_run(ctx, g, k, theta, seeds, &mm);
        // Metrics (phase1.cpp:268-339): counts from the stats sweep.
        std::vector<int64_t> cv, ce;
        int64_t cut = 0;
        const int err = stats_sweep(ctx, g, k, cv, ce, cut);
        DMOLP_REQUIRE(err == 0, err, "phase-1 labels failed the stats sweep");
        mm.initial_edge_cut = cut;
        const double expected_vertices = static_cast<double>(g->n) / k;
        double max_vr = 0;
        for (int p = 0; p < k; ++p) max_vr = std::max(max_vr, (double)cv[(size_t)p] / expected_vertices);
        double total_pe = 0;
        for (int p = 0; p < k; ++p) total_pe += (double)ce[(size_t)p];
        const double expected_edges = total_pe / k;
        double max_er = 0;
        for (int p = 0; p < k; ++p) max_er = std::max(max_er, (double)ce[(size_t)p] / expected_edges);
        mm.initial_vertex_balance = max_vr;
        mm.initial_edge_balance = max_er;
        mm.total_vertices = g->n;
        mm.total_edges = g->nnz / 2;
        if (m) *m = mm;
        if (vc) std::copy(cv.begin(), cv.end(), vc);
        if (ec) std::copy(ce.begin(), ce.end(), ec);
        if (seeds_out) std::copy(seeds.begin(), seeds.end(), seeds_out);
        if (labels_out) get_labels(ctx, g, labels_out);
    });
}

int dmolp_phase1_random(dmolp_ctx *ctx, dmolp_graph *g, int32_t k, int32_t *labels_out,
                        int64_t *vc, int64_t *ec, dmolp_phase1_metrics *m) {
    if (!ctx || !g) return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    if (k <= 0 || k > DMOLP_MAX_K) return fail(ctx, DMOLP_E_ARG, "k outside [1, 1024]");
    return guarded(ctx, [&] {
        DMOLP_HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
        phase1_random(ctx, g, k);
        DMOLP_HIP_CHECK(hipEventRecord(ctx->ev1, ctx->stream));
        std::vector<int64_t> cv, ce;
        int64_t cut = 0;
        const int err = stats_sweep(ctx, g, k, cv, ce, cut);
        DMOLP_REQUIRE(err == 0, err, "labels failed the stats sweep");
        float ms = 0.f;
        DMOLP_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        dmolp_phase1_metrics mm{};
        mm.initial_edge_cut = cut;
        const double expected_vertices = static_cast<double>(g->n) / k;
        double max_vr = 0, total_pe = 0, max_er = 0;
        for (int p = 0; p < k; ++p) max_vr = std::max(max_vr, (double)cv[(size_t)p] / expected_vertices);
        for (int p = 0; p < k; ++p) total_pe += (double)ce[(size_t)p];
        const double expected_edges = total_pe / k;
        for (int p = 0; p < k; ++p) max_er = std::max(max_er, (double)ce[(size_t)p] / expected_edges);
        mm.initial_vertex_balance = max_vr;
        mm.initial_edge_balance = max_er;
        mm.total_vertices = g->n;
        mm.total_edges = g->nnz / 2;
        mm.fill_round = -1;
        mm.device_ms = ms;
        if (m) *m = mm;
        if (vc) std::copy(cv.begin(), cv.end(), vc);
        if (ec) std::copy(ce.begin(), ce.end(), ec);
        if (labels_out) get_labels(ctx, g, labels_out);
    });
}

int dmolp_phase2(dmolp_ctx *ctx, dmolp_graph *g, int32_t k, int32_t max_rounds, double epsilon,
                 int64_t *vc, int64_t *ic, dmolp_phase2_metrics *m, dmolp_trace *trace) {
    if (!ctx || !g) return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    if (k <= 0 || k > DMOLP_MAX_K) return fail(ctx, DMOLP_E_ARG, "k outside [1, 1024]");
    if (max_rounds < 0) return fail(ctx, DMOLP_E_ARG, "max_rounds must be >= 0");
    if (!g->labels_valid) return fail(ctx, DMOLP_E_STATE, "graph has no resident labels");
    return guarded(ctx, [&] {
        Phase2Work &w = phase2_work(ctx, g, k);
        if (w.trace_cap < max_rounds) {
            w.trace.alloc(sizeof(int64_t) * (size_t)(max_rounds > 0 ? max_rounds : 1) * T_NCOL);
            w.trace_cap = max_rounds;
        }
        hipStream_t s = ctx->stream;
        ctx->ev_used = 0;
        ctx->prof_marks.clear();
        ctx->prof_ms.clear();
        phase2_round_table(ctx, w, max_rounds);
        const P2Set set = phase2_single_set(ctx, g, w, max_rounds, epsilon);
        const uint64_t key = phase2_single_key(ctx, g, w);
        DMOLP_HIP_CHECK(hipEventRecord(ctx->ev0, s));
        phase2_init_enqueue(ctx, set);
        Ctl c{};
        const int32_t *d_done = &w.ctl.as<Ctl>()->done;
        for (int t = 1; t <= max_rounds; ++t) {
            phase2_round_launch(ctx, set, w.round, key);
            if (t < max_rounds && round_poll(ctx, d_done, t, false)) break;
        }
        DMOLP_HIP_CHECK(hipEventRecord(ctx->ev1, s));
        std::vector<int64_t> cv((size_t)k), ci((size_t)k);
        DMOLP_HIP_CHECK(hipMemcpyAsync(&c, w.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
        DMOLP_HIP_CHECK(hipMemcpyAsync(cv.data(), w.vc.p, sizeof(int64_t) * (size_t)k,
                                       hipMemcpyDeviceToHost, s));
        DMOLP_HIP_CHECK(hipMemcpyAsync(ci.data(), w.ic.p, sizeof(int64_t) * (size_t)k,
                                       hipMemcpyDeviceToHost, s));
        std::vector<int64_t> tr((size_t)(max_rounds > 0 ? max_rounds : 1) * T_NCOL, 0);
        if (trace && max_rounds > 0)
            DMOLP_HIP_CHECK(hipMemcpyAsync(tr.data(), w.trace.p,
                                           sizeof(int64_t) * (size_t)max_rounds * T_NCOL,
                                           hipMemcpyDeviceToHost, s));
        DMOLP_HIP_CHECK(hipEventSynchronize(ctx->ev1));
        DMOLP_HIP_CHECK(hipStreamSynchronize(s));
        DMOLP_REQUIRE(c.err == 0, c.err,
                      c.err == DMOLP_E_INTERNAL ? "a phase-2 kernel found an impossible state"
                                                : "phase 2 rejected the labels (see status)");
        float ms = 0.f;
        DMOLP_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        if (ctx->profiling) {
            ctx->prof_ms.assign(ctx->prof_marks.size(), {});
            for (size_t gi = 0; gi < ctx->prof_marks.size(); ++gi) {
                for (int t = 0; t < c.rounds && t < (int)ctx->prof_marks[gi].size(); ++t) {
                    const auto &sp = ctx->prof_marks[gi][(size_t)t];
                    float e = 0.f;
                    DMOLP_HIP_CHECK(hipEventElapsedTime(&e, ctx->ev_pool[sp.first],
                                                        ctx->ev_pool[sp.second]));
                    ctx->prof_ms[gi].push_back(e);
                }
            }
        }
        // Final statistics (phase2.cpp:424-441).
        int64_t tv = 0, te = 0;
        for (int i = 0; i < k; ++i) { tv += cv[(size_t)i]; te += ci[(size_t)i]; }
        const double exp_v = static_cast<double>(tv) / k;
        const double exp_e = (te > 0) ? static_cast<double>(te) / k : 1.0;
        double max_v = 0.0, max_e = 0.0;
        for (int i = 0; i < k; ++i) {
            const double rv = (exp_v > 0) ? static_cast<double>(cv[(size_t)i]) / exp_v : 1.0;
            const double r_e = (exp_e > 0) ? static_cast<double>(ci[(size_t)i]) / exp_e : 1.0;
            max_v = std::max(max_v, rv);
            max_e = std::max(max_e, r_e);
        }
        if (m) {
            std::memset(m, 0, sizeof(*m));
            m->edge_cut = (int64_t)c.v[C_PREV_CUT];
            m->vertex_balance = max_v;
            m->edge_balance = max_e;
            m->total_vertices = tv;
            m->total_edges = te;
            m->rounds = c.rounds;
            m->exit_reason = max_rounds == 0 ? DMOLP_EXIT_MAX_ROUNDS : c.reason;
            m->lp_edges = (int64_t)c.v[C_LP_EDGES];
            m->initial_cut = (int64_t)c.v[C_INIT_CUT];
            m->device_ms = ms;
        }
        if (vc) std::copy(cv.begin(), cv.end(), vc);
        if (ic) std::copy(ci.begin(), ci.end(), ic);
        if (trace) {
            const int R = c.rounds;
            int64_t *cols[] = {trace->prev_count, trace->prev_deg, trace->cand_count,
                               trace->bd_count,   trace->b_count,  trace->b_deg,
                               trace->chg_count,  trace->chg_deg,  trace->cut,
                               trace->b_light_count, trace->b_light_deg,
                               trace->chg_light_count, trace->chg_light_deg};
            const int map[] = {T_PREV_N, T_PREV_DEG, T_CAND, T_BD, T_B, T_BDEG, T_CHG, T_CHGDEG, T_CUT,
                               T_BL, T_BLDEG, T_DL, T_DLDEG};
            for (int j = 0; j < 13; ++j) {
                if (!cols[j]) continue;
                for (int t = 0; t < R; ++t) cols[j][t] = tr[(size_t)t * T_NCOL + map[j]];
            }
        }
    });
}

int dmolp_stats(dmolp_ctx *ctx, dmolp_graph *g, int32_t k, const int32_t *labels, int64_t *vc,
                int64_t *ic, int64_t *cut) {
    if (!ctx || !g || !labels) return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    if (k <= 0 || k > DMOLP_MAX_K) return fail(ctx, DMOLP_E_ARG, "k outside [1, 1024]");
    return guarded(ctx, [&] {
        set_labels(ctx, g, labels);
        std::vector<int64_t> cv, ci;
        int64_t c = 0;
        const int err = stats_sweep(ctx, g, k, cv, ci, c);
        DMOLP_REQUIRE(err == 0, err, "labels failed the stats sweep");
        if (vc) std::copy(cv.begin(), cv.end(), vc);
        if (ic) std::copy(ci.begin(), ci.end(), ic);
        if (cut) *cut = c;
    });
}

int dmolp_stats_noalloc(dmolp_ctx *ctx, const dmolp_graph *g, int32_t k, const int32_t *labels,
                        int64_t *vc, int64_t *ic, int64_t *cut) {
    if (!ctx || !g || (!labels && g->n > 0)) return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    return guarded(ctx, [&] { stats_noalloc(ctx, g, k, labels, vc, ic, cut); });
}

int dmolp_penalties(dmolp_ctx *ctx, int32_t k, const int64_t *vc, const int64_t *ic, int32_t t,
                    double *pen, double *re) {
    if (!ctx || !vc || !ic || k <= 0 || k > DMOLP_MAX_K || t < 1)
        return fail(ctx, DMOLP_E_ARG, "bad arguments");
    return guarded(ctx, [&] {
        hipStream_t s = ctx->stream;
        DevMem dv(sizeof(int64_t) * (size_t)k), di(sizeof(int64_t) * (size_t)k);
        DevMem dp(sizeof(double) * (size_t)k), dr(sizeof(double) * (size_t)k);
        DMOLP_HIP_CHECK(hipMemcpyAsync(dv.p, vc, dv.bytes, hipMemcpyHostToDevice, s));
        DMOLP_HIP_CHECK(hipMemcpyAsync(di.p, ic, di.bytes, hipMemcpyHostToDevice, s));
        phase2_penalties_device(ctx, k, dv.as<unsigned long long>(), di.as<unsigned long long>(), t,
                                dp.as<double>(), dr.as<double>());
        if (pen) DMOLP_HIP_CHECK(hipMemcpyAsync(pen, dp.p, dp.bytes, hipMemcpyDeviceToHost, s));
        if (re) DMOLP_HIP_CHECK(hipMemcpyAsync(re, dr.p, dr.bytes, hipMemcpyDeviceToHost, s));
        DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int dmolp_select_boundary(dmolp_ctx *ctx, dmolp_graph *g, const int32_t *labels,
                          const int32_t *prev, int64_t n_prev, const double *re, int32_t k,
                          int32_t t, int32_t *out, int64_t *n_out) {
    if (!ctx || !g || !labels || !re || !out || !n_out || n_prev < 0 || (n_prev > 0 && !prev))
        return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    if (k <= 0 || k > DMOLP_MAX_K || t < 1) return fail(ctx, DMOLP_E_ARG, "bad k or t");
    return guarded(ctx, [&] {
        set_labels(ctx, g, labels);
        Phase2Work &w = phase2_work(ctx, g, k);
        phase2_init_enqueue(ctx, phase2_single_set(ctx, g, w, 0, 0.0));  // differing-neighbour counts
        Ctl c{};
        DMOLP_HIP_CHECK(hipMemcpyAsync(&c, w.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream));
        DMOLP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        DMOLP_REQUIRE(c.err == 0, c.err, "labels failed the stats sweep");
        upload_list(ctx, g, prev, n_prev, w.B);
        DMOLP_HIP_CHECK(hipMemcpyAsync(w.re.p, re, sizeof(double) * (size_t)k, hipMemcpyHostToDevice,
                                       ctx->stream));
        phase2_select_enqueue(ctx, g, w, t);
        const std::vector<int32_t> b = download_list(ctx, w.B);
        std::copy(b.begin(), b.end(), out);
        *n_out = (int64_t)b.size();
    });
}

int dmolp_lp_round(dmolp_ctx *ctx, dmolp_graph *g, const int32_t *labels, const int32_t *b,
                   int64_t nb, const double *pen, int32_t k, int32_t *labels_out) {
    if (!ctx || !g || !labels || !pen || !labels_out || nb < 0 || (nb > 0 && !b))
        return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    if (k <= 0 || k > DMOLP_MAX_K) return fail(ctx, DMOLP_E_ARG, "k outside [1, 1024]");
    return guarded(ctx, [&] {
        set_labels(ctx, g, labels);
        Phase2Work &w = phase2_work(ctx, g, k);
        std::vector<int32_t> ids(b, b + nb);
        std::sort(ids.begin(), ids.end());
        DMOLP_REQUIRE(std::adjacent_find(ids.begin(), ids.end()) == ids.end(), DMOLP_E_ARG,
                      "boundary list has duplicates");
        DMOLP_HIP_CHECK(hipMemsetAsync(w.ctl.p, 0, sizeof(Ctl), ctx->stream));
        upload_list(ctx, g, ids.data(), nb, w.B);
        DMOLP_HIP_CHECK(hipMemcpyAsync(w.pen.p, pen, sizeof(double) * (size_t)k,
                                       hipMemcpyHostToDevice, ctx->stream));
        phase2_lp_enqueue(ctx, g, w);
        get_labels(ctx, g, labels_out);
    });
}

int dmolp_batch_run(dmolp_ctx *ctx, int32_t n_jobs, const dmolp_job *jobs, dmolp_result *results,
                    double *device_ms) {
    if (!ctx || n_jobs < 0 || (n_jobs > 0 && (!jobs || !results)))
        return fail(ctx, DMOLP_E_ARG, "null argument");
    for (int32_t i = 0; i < n_jobs; ++i) note_use(ctx, jobs[i].graph);
    return guarded(ctx, [&] { batch_run(ctx, n_jobs, jobs, results, device_ms); });
}

int dmolp_batch_autotune(dmolp_ctx *ctx, int32_t n_jobs, const dmolp_job *jobs, int32_t max_trials,
                         dmolp_tuning *best) {
    if (!ctx || n_jobs <= 0 || !jobs) return fail(ctx, DMOLP_E_ARG, "null argument");
    for (int32_t i = 0; i < n_jobs; ++i) note_use(ctx, jobs[i].graph);
    return guarded(ctx, [&] { batch_autotune(ctx, n_jobs, jobs, max_trials, best); });
}

int dmolp_server_create(dmolp_ctx *ctx, int64_t budget, int32_t max_active, dmolp_server **out) {
    if (!ctx || !out) return fail(ctx, DMOLP_E_ARG, "null argument");
    *out = nullptr;
    return guarded(ctx, [&] { *out = server_create(ctx, budget, max_active); });
}

int dmolp_server_destroy(dmolp_server *srv) {
    if (srv) server_destroy(srv);
    return DMOLP_OK;
}

double dmolp_server_clock(const dmolp_server *srv) { return srv ? server_clock(srv) : 0.0; }

int dmolp_server_submit(dmolp_server *srv, const dmolp_request *req, int64_t *ticket) {
    if (!srv || !req || !ticket) return fail(nullptr, DMOLP_E_ARG, "null argument");
    return guarded(nullptr, [&] { *ticket = server_submit(srv, *req); });
}

int dmolp_server_step(dmolp_server *srv, int32_t *active, int32_t *queued) {
    if (!srv) return fail(nullptr, DMOLP_E_ARG, "null argument");
    return guarded(nullptr, [&] { server_step(srv, active, queued); });
}

int dmolp_server_take(dmolp_server *srv, int64_t ticket, dmolp_served *out, int32_t *labels_out) {
    if (!srv) return fail(nullptr, DMOLP_E_ARG, "null argument");
    int ready = 0;
    const int rc = guarded(nullptr, [&] { ready = server_take(srv, ticket, out, labels_out); });
    return rc != DMOLP_OK ? rc : ready;
}

// ---------------------------------------------------------------- rank shares

int dmolp_shares_create(dmolp_ctx *ctx, dmolp_graph *g, int32_t P, dmolp_shares **out) {
    if (!ctx || !g || !out) return fail(ctx, DMOLP_E_ARG, "null argument");
    note_use(ctx, g);
    *out = nullptr;
    return guarded(ctx, [&] { *out = shares_create(ctx, g, P); });
}

int dmolp_shares_destroy(dmolp_ctx *ctx, dmolp_shares *s) {
    if (!s) return DMOLP_OK;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    else (void)hipDeviceSynchronize();
    delete s;
    return DMOLP_OK;
}

int dmolp_shares_info(const dmolp_shares *s, int32_t r, int64_t *n_local, int64_t *n_ghost,
                      int64_t *nnz) {
    if (!s || r < 0 || r >= s->P) return DMOLP_E_ARG;
    const Share &x = s->sh[(size_t)r];
    if (n_local) *n_local = x.nl;
    if (n_ghost) *n_ghost = x.ng;
    if (nnz) *nnz = x.nnz;
    return DMOLP_OK;
}

int dmolp_shares_download(dmolp_ctx *ctx, const dmolp_shares *s, int32_t r, int64_t *off,
                          int32_t *col, int32_t *ghost_gid, int32_t *labels) {
    if (!ctx || !s || r < 0 || r >= s->P) return fail(ctx, DMOLP_E_ARG, "bad arguments");
    return guarded(ctx, [&] {
        const Share &x = s->sh[(size_t)r];
        hipStream_t st = ctx->stream;
        if (off)
            DMOLP_HIP_CHECK(hipMemcpyAsync(off, x.off.p, sizeof(int64_t) * (size_t)(x.nl + 1),
                                           hipMemcpyDeviceToHost, st));
        if (col && x.nnz > 0)
            DMOLP_HIP_CHECK(hipMemcpyAsync(col, x.col.p, sizeof(int32_t) * (size_t)x.nnz,
                                           hipMemcpyDeviceToHost, st));
        if (ghost_gid && x.ng > 0)
            DMOLP_HIP_CHECK(hipMemcpyAsync(ghost_gid, x.ghost_gid.p, sizeof(int32_t) * (size_t)x.ng,
                                           hipMemcpyDeviceToHost, st));
        if (labels && x.nl + x.ng > 0)
            DMOLP_HIP_CHECK(hipMemcpyAsync(labels, x.lab.p, sizeof(int32_t) * (size_t)(x.nl + x.ng),
                                           hipMemcpyDeviceToHost, st));
        DMOLP_HIP_CHECK(hipStreamSynchronize(st));
    });
}

int dmolp_halo_refresh(dmolp_ctx *ctx, dmolp_shares *s, int32_t r, const int32_t *gid,
                       const int32_t *label, int64_t count, int64_t *updated) {
    if (!ctx || !s || r < 0 || r >= s->P || count < 0 || (count > 0 && (!gid || !label)))
        return fail(ctx, DMOLP_E_ARG, "bad arguments");
    return guarded(ctx, [&] {
        DevMem dg(sizeof(int32_t) * (size_t)(count > 0 ? count : 1)),
            dl(sizeof(int32_t) * (size_t)(count > 0 ? count : 1));
        if (count > 0) {
            DMOLP_HIP_CHECK(hipMemcpyAsync(dg.p, gid, sizeof(int32_t) * (size_t)count,
                                           hipMemcpyHostToDevice, ctx->stream));
            DMOLP_HIP_CHECK(hipMemcpyAsync(dl.p, label, sizeof(int32_t) * (size_t)count,
                                           hipMemcpyHostToDevice, ctx->stream));
        }
        const int64_t u = halo_refresh(ctx, s, r, dg.as<int32_t>(), dl.as<int32_t>(), count);
        if (updated) *updated = u;
    });
}

int dmolp_phase2_shares(dmolp_ctx *ctx, dmolp_shares *s, int32_t k, int32_t max_rounds,
                        double epsilon, int32_t scope, int32_t *labels_out, int64_t *vc,
                        int64_t *ic, dmolp_phase2_metrics *m, dmolp_trace *trace,
                        int64_t *ghost_updates) {
    if (!ctx || !s) return fail(ctx, DMOLP_E_ARG, "null argument");
    return guarded(ctx, [&] {
        shares_phase2(ctx, s, k, max_rounds, epsilon, scope, labels_out, vc, ic, m, trace,
                      ghost_updates);
    });
}

}  // extern "C"
