// dev_common.hpp -- device helpers shared by the gfx950 kernels.
//
// Wave size is 64 (CDNA4).  Work lists are appended with one atomic per
// wave (ballot + mbcnt), never one per lane.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace dmolp {

constexpr int kWave = 64;
constexpr int kLpCap = 32;          // gpu_lp_boundary.cu:360
constexpr int kSelBins = 4096;      // level-1 selection bins: min(deg, 4095)
constexpr int kDigitBits = 8;       // radix-select digit width (histograms fit LDS)
constexpr int kDigitBins = 1 << kDigitBits;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of active lanes below this one in `mask`.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <typename T>
__device__ __forceinline__ T wave_bcast(T v, int src) {
    return __shfl(v, src, kWave);
}

// Inclusive scan over the 64 lanes (all lanes must participate).
__device__ __forceinline__ int64_t wave_incl_scan(int64_t x) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int64_t y = __shfl_up(x, o, kWave);
        if (lane_id() >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ int64_t wave_sum(int64_t x) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}

// Segmented inclusive sum over the 64 lanes (all lanes must participate).
// Each lane holds key << 24 | count; lanes whose keys are equal must be
// consecutive, and every run's count sum must stay below 2^24.  A lane ends
// with its key and the sum of the counts of its run's lanes up to itself.
// Hillis-Steele steps of 1, 2, 4 and 8 within each row of 16 lanes (DPP
// row_shr), then the end of each row is carried into the next rows (DPP
// row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2 and 3).  A lane
// with no source keeps kNone, a key no lane holds.
__device__ __forceinline__ uint32_t wave_seg_incl(uint32_t x) {
    constexpr int kNone = (int)0xfe000000u;
    auto step = [](uint32_t x, int p) {
        const uint32_t y = (uint32_t)p;
        return (y >> 24) == (x >> 24) ? x + (y & 0xffffffu) : x;
    };
    x = step(x, __builtin_amdgcn_update_dpp(kNone, (int)x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = step(x, __builtin_amdgcn_update_dpp(kNone, (int)x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = step(x, __builtin_amdgcn_update_dpp(kNone, (int)x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = step(x, __builtin_amdgcn_update_dpp(kNone, (int)x, 0x118, 0xf, 0xf, false));  // row_shr:8
    x = step(x, __builtin_amdgcn_update_dpp(kNone, (int)x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    x = step(x, __builtin_amdgcn_update_dpp(kNone, (int)x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return x;
}

// Inclusive max-scan over the 64 lanes for values >= -1 (all lanes must
// participate): the same DPP steps as wave_seg_incl, with -1 for lanes that
// have no source.
__device__ __forceinline__ int32_t wave_max_incl(int32_t x) {
    const auto step = [](int32_t x, int32_t y) { return y > x ? y : x; };
    x = step(x, __builtin_amdgcn_update_dpp(-1, x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = step(x, __builtin_amdgcn_update_dpp(-1, x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = step(x, __builtin_amdgcn_update_dpp(-1, x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = step(x, __builtin_amdgcn_update_dpp(-1, x, 0x118, 0xf, 0xf, false));  // row_shr:8
    x = step(x, __builtin_amdgcn_update_dpp(-1, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    x = step(x, __builtin_amdgcn_update_dpp(-1, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return x;
}

// Append the lanes with `pred` to list[*count ...]: one atomic per wave.
// Returns this lane's slot (or -1).  All lanes must call it.
__device__ __forceinline__ int64_t wave_append(bool pred, unsigned long long *count) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return -1;
    int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane_id() == leader) base = atomicAdd(count, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader, kWave);
    return pred ? (int64_t)(base + (unsigned long long)lanes_below(mask)) : -1;
}

// Add a wave-reduced value to a global counter with one atomic.
__device__ __forceinline__ void wave_add(unsigned long long *dst, int64_t v) {
    int64_t s = wave_sum(v);
    if (lane_id() == 0 && s != 0) atomicAdd(dst, (unsigned long long)s);
}

__device__ __forceinline__ int bitrev5(int x) { return (int)(__brev((unsigned)x) >> 27); }

// Global wave index and wave count of a 1-D launch.
__device__ __forceinline__ int64_t global_wave() {
    return ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
}
__device__ __forceinline__ int64_t wave_count() {
    return ((int64_t)gridDim.x * blockDim.x) / kWave;
}

// Job records live in device memory, so the pointers in them are loaded at
// run time and the compiler cannot tell which address space they point to:
// it would emit flat memory instructions, which serialise against LDS
// traffic.  gload reads a stored pointer through a global-address-space
// view, so every access derived from it compiles to global instructions.
template <class T>
using as1_ptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ T *gload(T *const &field) {
    return (T *)(*reinterpret_cast<const as1_ptr<T> *>(&field));
}

// One launch serves a set of jobs (graphs): job j owns the blocks
// [bstart[j], bstart[j+1]).  A kernel body sees its job's share of the
// launch as a virtual grid, so the same body runs one graph or many.
struct VGrid {
    int64_t block;   // block index within the job
    int64_t blocks;  // blocks of the job
};

// The job owning this block (largest j with bstart[j] <= blockIdx.x),
// made wave-uniform so the job's fields load through the scalar unit.
__device__ __forceinline__ int job_of_block(const int32_t *bstart, int jobs) {
    int lo = 0, hi = jobs - 1;
    const int b = (int)blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}
__device__ __forceinline__ VGrid vgrid_of(const int32_t *bstart, int j) {
    const int32_t b0 = bstart[j], b1 = bstart[j + 1];
    return VGrid{(int64_t)blockIdx.x - b0, (int64_t)(b1 - b0)};
}
__device__ __forceinline__ int64_t vwave(const VGrid &v) {
    return v.block * (int64_t)(blockDim.x / kWave) + (int64_t)(threadIdx.x / kWave);
}
__device__ __forceinline__ int64_t vwaves(const VGrid &v) {
    return v.blocks * (int64_t)(blockDim.x / kWave);
}
__device__ __forceinline__ int64_t vthread(const VGrid &v) {
    return v.block * (int64_t)blockDim.x + (int64_t)threadIdx.x;
}
__device__ __forceinline__ int64_t vthreads(const VGrid &v) {
    return v.blocks * (int64_t)blockDim.x;
}

}  // namespace dmolp
