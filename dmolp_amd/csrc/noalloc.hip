// noalloc.hip -- dmolp_stats_noalloc: the statistics and cut of a host
// labelling (computePartitionStats + computeEdgeCut, phase2.cpp:55-117,
// 214-253) with no device allocation, for callers that cannot allocate
// (code under HIP graph capture, tight serving loops).
//
// Its device memory beyond the graph and the caller's host buffers is one
// file-scope array, `scratch`: the labels copied in (at most
// kNoAllocMaxN vertices) and the counters.  Each GPU holds its own copy for
// the life of the process; its address on a device is taken once, when a
// context is created on that device, and kept per device.  Every call on a
// GPU reuses that GPU's array, so calls that overlap on one GPU race; the
// counters are cleared by a kernel at the start of each call.
#include <algorithm>
#include <mutex>

#include "internal.hpp"
#include "noalloc.hpp"
#include "rows.hpp"

namespace dmolp {
namespace {

constexpr int64_t kNoAllocMaxN = DMOLP_NOALLOC_MAX_N;
constexpr int kNoAllocCounters = 2 * DMOLP_MAX_K + 2;  // vc, intra entries, differing, invalid
constexpr size_t kScratchLabelBytes = sizeof(int32_t) * (size_t)kNoAllocMaxN;
constexpr size_t kScratchBytes =
    kScratchLabelBytes + sizeof(unsigned long long) * (size_t)kNoAllocCounters;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

constexpr int kNaThreads = 256;

__global__ void k_na_clear(unsigned long long *c, int n) {
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x))
        c[i] = 0ull;
}

// One wave per row: the row's label counts against its own label.  Labels
// below kLdsLabels count in LDS and are added once per block.
constexpr int kLdsLabels = DMOLP_MAX_K;
__global__ __launch_bounds__(kNaThreads) void k_na_stats(GraphDev g, const int32_t *L, int k,
                                                         unsigned long long *cnt) {
    __shared__ unsigned long long s_vc[kLdsLabels], s_ic[kLdsLabels];
    for (int i = threadIdx.x; i < k; i += blockDim.x) { s_vc[i] = 0; s_ic[i] = 0; }
    __syncthreads();
    const int lane = lane_id();
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
    int64_t diffs = 0, invalid = 0;
    for (int64_t u = w0; u < g.n; u += nw) {
        const int32_t lu = L[u];
        const bool ok = lu >= 0 && lu < k;
        int64_t intra = 0;
        for (int64_t e = g.off[u] + lane; e < g.off[u + 1]; e += kWave) {
            const int32_t lv = L[g.col[e]];
            if (ok && lv == lu) ++intra;
            if (lu == -1 || lv == -1) { ++invalid; continue; }
            if (lu != lv) ++diffs;
        }
        intra = wave_sum(intra);
        if (lane == 0) {
            if (ok) atomicAdd(&s_vc[lu], 1ull);
            if (intra) atomicAdd(&s_ic[lu], (unsigned long long)intra);
        }
    }
    diffs = wave_sum(diffs);
    invalid = wave_sum(invalid);
    if (lane == 0) {
        if (diffs) atomicAdd(&cnt[2 * k], (unsigned long long)diffs);
        if (invalid) atomicAdd(&cnt[2 * k + 1], (unsigned long long)invalid);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        if (s_vc[i]) atomicAdd(&cnt[i], s_vc[i]);
        if (s_ic[i]) atomicAdd(&cnt[k + i], s_ic[i]);
    }
}

}  // namespace

void noalloc_init_device(int device) {
    DMOLP_REQUIRE(device >= 0 && device < kMaxDevices, DMOLP_E_UNSUPPORTED,
                  "device index beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    DMOLP_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DMOLP_REQUIRE(p != nullptr, DMOLP_E_HIP, "no address for the no-allocation scratch array");
    g_scratch_ptr[device] = p;
}

void stats_noalloc(dmolp_ctx *ctx, const dmolp_graph *g, int32_t k, const int32_t *labels,
                   int64_t *vc, int64_t *ic, int64_t *cut) {
    DMOLP_REQUIRE(k >= 1 && k <= DMOLP_MAX_K, DMOLP_E_ARG, "k outside [1, 1024]");
    DMOLP_REQUIRE(g->n <= kNoAllocMaxN, DMOLP_E_UNSUPPORTED,
                  "graph larger than the no-allocation scratch (DMOLP_NOALLOC_MAX_N vertices)");
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
    }
    DMOLP_REQUIRE(base != nullptr, DMOLP_E_STATE, "no-allocation scratch not set up on this device");
    int32_t *d_labels = static_cast<int32_t *>(base);
    unsigned long long *d_cnt =
        reinterpret_cast<unsigned long long *>(static_cast<unsigned char *>(base) + kScratchLabelBytes);
    hipStream_t st = ctx->stream;
    const int nc = 2 * k + 2;
    hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(kNaThreads), 0, st, d_cnt, nc);
    if (g->n > 0)
        DMOLP_HIP_CHECK(hipMemcpyAsync(d_labels, labels, sizeof(int32_t) * (size_t)g->n,
                                       hipMemcpyHostToDevice, st));
    const int64_t waves = g->n > 0 ? g->n : 1;
    const int blocks = (int)std::min<int64_t>((waves + (kNaThreads / kWave) - 1) / (kNaThreads / kWave),
                                              (int64_t)ctx->cu_count * 8);
    hipLaunchKernelGGL(k_na_stats, dim3(blocks), dim3(kNaThreads), 0, st, g->dev(), d_labels, k, d_cnt);
    DMOLP_HIP_CHECK(hipGetLastError());
    unsigned long long h[kNoAllocCounters];  // host side on the stack: nothing allocated
    DMOLP_HIP_CHECK(hipMemcpyAsync(h, d_cnt, sizeof(unsigned long long) * (size_t)nc,
                                   hipMemcpyDeviceToHost, st));
    DMOLP_HIP_CHECK(hipStreamSynchronize(st));
    DMOLP_REQUIRE(h[(size_t)(2 * k + 1)] == 0, DMOLP_E_INVALID_LABEL,
                  "a -1 label reached the cut (phase2.cpp:239-244)");
    DMOLP_REQUIRE((h[(size_t)(2 * k)] & 1ull) == 0, DMOLP_E_ODD_CUT,
                  "odd cut: the CSR is not symmetric (phase2.cpp:246-250)");
    for (int i = 0; i < k; ++i) {
        if (vc) vc[i] = (int64_t)h[(size_t)i];
        if (ic) ic[i] = (int64_t)(h[(size_t)(k + i)] / 2);  // phase2.cpp:110
    }
    if (cut) *cut = (int64_t)(h[(size_t)(2 * k)] >> 1);
}

}  // namespace dmolp
