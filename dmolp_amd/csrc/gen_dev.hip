// gen_dev.hip -- device graph generator: same samples as gen_host.cpp
// (same counter-based stream), symmetrised into 64-bit (row, col) keys,
// radix-sorted and deduplicated with rocPRIM, then unpacked into CSR.
// This builds inputs; it is not part of the partitioning path.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "gen_common.hpp"
#include "gen_host.hpp"
#include "internal.hpp"

namespace dmolp {
namespace {

constexpr uint64_t kSentinel = ~0ull;

// An entry (u, v) is the key (u << sh) | v with sh bits per id, so the sort
// looks at 2 * sh bits only; self-loops become the sentinel, which sorts
// last (no entry key has all 2 * sh low bits set: that would be the loop
// (2^sh - 1, 2^sh - 1)).
__global__ void k_rmat_keys(uint64_t key, int scale, RmatThresholds th, int64_t m, uint64_t *out) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (int64_t)gridDim.x * blockDim.x) {
        uint32_t u, v;
        rmat_edge(key, scale, th, (uint64_t)e, &u, &v);
        const bool loop = u == v;
        out[2 * e] = loop ? kSentinel : (((uint64_t)u << scale) | v);
        out[2 * e + 1] = loop ? kSentinel : (((uint64_t)v << scale) | u);
    }
}

__global__ void k_chung_lu_keys(uint64_t key, const uint64_t *cdf, int64_t n, int64_t m, int sh,
                                uint64_t *out) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
         e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t u = chung_lu_pick(cdf, n, rand_at(key, 2 * (uint64_t)e));
        const uint32_t v = chung_lu_pick(cdf, n, rand_at(key, 2 * (uint64_t)e + 1));
        const bool loop = u == v;
        out[2 * e] = loop ? kSentinel : (((uint64_t)u << sh) | v);
        out[2 * e + 1] = loop ? kSentinel : (((uint64_t)v << sh) | u);
    }
}

// keys[0..nnz) sorted unique: col = low sh bits; off[r] = first index of
// row r.
__global__ void k_unpack(const uint64_t *keys, int64_t nnz, int64_t n, int sh, int32_t *col,
                         int64_t *off) {
    const uint64_t mask = (1ull << sh) - 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i < nnz ? (int64_t)(keys[i] >> sh) : n;
        const int64_t rp = i > 0 ? (int64_t)(keys[i - 1] >> sh) : -1;
        if (i < nnz) col[i] = (int32_t)(keys[i] & mask);
        for (int64_t x = rp + 1; x <= r; ++x) off[x] = i;
    }
}

}  // namespace

void gen_device(dmolp_ctx *ctx, const dmolp_gen_spec *spec, dmolp_graph *g) {
    DMOLP_REQUIRE(spec, DMOLP_E_ARG, "null spec");
    hipStream_t s = ctx->stream;
    const uint64_t key = stream_key(spec->seed);
    int64_t n = 0, m = 0;
    int sh = 0;  // bits per vertex id in a key
    DevMem keys_a, keys_b, cdf_d;
    const unsigned grid = (unsigned)ctx->cu_count * 16;
    if (spec->kind == DMOLP_GEN_RMAT) {
        DMOLP_REQUIRE(spec->scale >= 1 && spec->scale <= 30 && spec->edge_factor >= 1, DMOLP_E_ARG,
                      "bad R-MAT spec");
        n = 1ll << spec->scale;
        m = (int64_t)spec->edge_factor << spec->scale;
        sh = spec->scale;
        keys_a.alloc(sizeof(uint64_t) * 2 * (size_t)m);
        hipLaunchKernelGGL(k_rmat_keys, dim3(grid), dim3(256), 0, s, key, spec->scale,
                           rmat_thresholds(spec->a, spec->b, spec->c), m, keys_a.as<uint64_t>());
    } else if (spec->kind == DMOLP_GEN_CHUNG_LU) {
        n = spec->n;
        m = spec->m;
        DMOLP_REQUIRE(n >= 2 && n < (1ll << 31) && m >= 1, DMOLP_E_ARG, "bad Chung-Lu size");
        sh = (int)bits_for(n - 1);
        const std::vector<uint64_t> cdf = chung_lu_cdf(n, m, spec->gamma, spec->dmax);
        cdf_d.alloc(sizeof(uint64_t) * cdf.size());
        DMOLP_HIP_CHECK(hipMemcpyAsync(cdf_d.p, cdf.data(), cdf_d.bytes, hipMemcpyHostToDevice, s));
        keys_a.alloc(sizeof(uint64_t) * 2 * (size_t)m);
        hipLaunchKernelGGL(k_chung_lu_keys, dim3(grid), dim3(256), 0, s, key, cdf_d.as<uint64_t>(),
                           n, m, sh, keys_a.as<uint64_t>());
        DMOLP_HIP_CHECK(hipStreamSynchronize(s));  // cdf (host vector) must outlive the copy
        cdf_d.release();
    } else {
        // Grids have no randomness: build on the host and upload.
        dmolp_csr h{};
        gen_host(spec, &h);
        g->n = h.n;
        g->nnz = h.nnz;
        g->off.alloc(sizeof(int64_t) * (size_t)(h.n + 1));
        g->col.alloc(sizeof(int32_t) * (size_t)(h.nnz > 0 ? h.nnz : 1));
        DMOLP_HIP_CHECK(hipMemcpyAsync(g->off.p, h.off, sizeof(int64_t) * (size_t)(h.n + 1),
                                       hipMemcpyHostToDevice, s));
        if (h.nnz > 0)
            DMOLP_HIP_CHECK(hipMemcpyAsync(g->col.p, h.col, sizeof(int32_t) * (size_t)h.nnz,
                                           hipMemcpyHostToDevice, s));
        DMOLP_HIP_CHECK(hipStreamSynchronize(s));
        std::free(h.off);
        std::free(h.col);
        return;
    }
    DMOLP_HIP_CHECK(hipGetLastError());
    const size_t nk = 2 * (size_t)m;
    keys_b.alloc(sizeof(uint64_t) * nk);
    // Sort (row, col) keys; the sentinel sorts last.
    rocprim::double_buffer<uint64_t> db(keys_a.as<uint64_t>(), keys_b.as<uint64_t>());
    size_t tmp_bytes = 0;
    const unsigned key_bits = 2u * (unsigned)sh;
    DMOLP_HIP_CHECK(rocprim::radix_sort_keys(nullptr, tmp_bytes, db, nk, 0u, key_bits, s));
    DevMem tmp(tmp_bytes);
    DMOLP_HIP_CHECK(rocprim::radix_sort_keys(tmp.p, tmp_bytes, db, nk, 0u, key_bits, s));
    tmp.release();
    uint64_t *sorted = db.current();
    uint64_t *uniq = sorted == keys_a.as<uint64_t>() ? keys_b.as<uint64_t>() : keys_a.as<uint64_t>();
    DevMem ucount(sizeof(size_t));
    size_t utmp = 0;
    DMOLP_HIP_CHECK(rocprim::unique(nullptr, utmp, sorted, uniq, ucount.as<size_t>(), nk,
                                    rocprim::equal_to<uint64_t>(), s));
    DevMem tmp2(utmp);
    DMOLP_HIP_CHECK(rocprim::unique(tmp2.p, utmp, sorted, uniq, ucount.as<size_t>(), nk,
                                    rocprim::equal_to<uint64_t>(), s));
    size_t nu = 0;
    DMOLP_HIP_CHECK(hipMemcpyAsync(&nu, ucount.p, sizeof(size_t), hipMemcpyDeviceToHost, s));
    uint64_t last = 0;
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
    if (nu > 0) {
        DMOLP_HIP_CHECK(hipMemcpy(&last, uniq + nu - 1, sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
    const int64_t nnz = (int64_t)nu - ((nu > 0 && last == kSentinel) ? 1 : 0);
    g->n = n;
    g->nnz = nnz;
    g->off.alloc(sizeof(int64_t) * (size_t)(n + 1));
    g->col.alloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    hipLaunchKernelGGL(k_unpack, dim3(grid), dim3(256), 0, s, uniq, nnz, n, sh, g->col.as<int32_t>(),
                       g->off.as<int64_t>());
    DMOLP_HIP_CHECK(hipGetLastError());
    DMOLP_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace dmolp
