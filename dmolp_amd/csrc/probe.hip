// probe.hip -- dmolp_probe_atomics: the device's rate of scattered 32-bit
// atomic adds into 64-byte rows, the ceiling that bounds the Delta pass
// (k_delta_light moves counts in the kept light rows of random neighbours).
// bench.py reports it beside the kernel's HBM roofline.
#include <algorithm>

#include "internal.hpp"

namespace dmolp {
namespace {

__device__ __forceinline__ uint64_t probe_mix(uint64_t z) {
    z ^= z >> 30;
    z *= 0xbf58476d1ce4e5b9ull;
    z ^= z >> 27;
    z *= 0x94d049bb133111ebull;
    z ^= z >> 31;
    return z;
}

// Groups of `share` consecutive lanes pick one random row (a 64-byte row, 16
// words) and each lane adds 1 to a random word of it, `reps` times.  The
// row index is below rows, the word below 16: every add is in bounds.
__global__ void k_probe_atomics(uint32_t *rows16, uint64_t rows, int share, int reps, uint64_t seed) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t grp = t / (uint64_t)share;
    for (int i = 0; i < reps; ++i) {
        const uint64_t r = probe_mix(seed + grp * 0x9e3779b97f4a7c15ull + (uint64_t)i);
        const uint64_t row = r % rows;
        const uint32_t word = (uint32_t)(probe_mix(r ^ t) & 15u);
        atomicAdd(rows16 + row * 16 + word, 1u);
    }
}

}  // namespace
}  // namespace dmolp

using namespace dmolp;

extern "C" int dmolp_probe_atomics(dmolp_ctx *ctx, int64_t rows, int32_t share, int32_t reps, double *ms_out,
                                   int64_t *adds_out) {
    if (!ctx || rows < 1 || rows > ((int64_t)1 << 32) || share < 1 || share > 64 || (64 % share) != 0 ||
        reps < 1 || reps > 4096 || !ms_out)
        return DMOLP_E_ARG;
    try {
        DMOLP_HIP_CHECK(hipSetDevice(ctx->device));
        DevMem buf(sizeof(uint32_t) * 16 * (size_t)rows);
        hipStream_t s = ctx->stream;
        DMOLP_HIP_CHECK(hipMemsetAsync(buf.p, 0, buf.bytes, s));
        const int blocks = ctx->cu_count * 8, threads = 256;
        float best = 0.f;
        for (int it = 0; it < 3; ++it) {  // the first launch warms up; the fastest of the last two
            DMOLP_HIP_CHECK(hipEventRecord(ctx->ev0, s));
            hipLaunchKernelGGL(k_probe_atomics, dim3(blocks), dim3(threads), 0, s, buf.as<uint32_t>(),
                               (uint64_t)rows, (int)share, (int)reps, 77ull + (uint64_t)it);
            DMOLP_HIP_CHECK(hipGetLastError());
            DMOLP_HIP_CHECK(hipEventRecord(ctx->ev1, s));
            DMOLP_HIP_CHECK(hipEventSynchronize(ctx->ev1));
            float ms = 0.f;
            DMOLP_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            if (it > 0 && (best == 0.f || ms < best)) best = ms;
        }
        DMOLP_HIP_CHECK(hipStreamSynchronize(s));
        *ms_out = best;
        if (adds_out) *adds_out = (int64_t)blocks * threads * reps;
        return DMOLP_OK;
    } catch (const Error &e) {
        ctx->last_error = e.what();
        return e.status;
    }
}
