// gen_common.hpp -- graph generator primitives shared by the host generator
// and the device generator, so both produce the same graph from a seed.
//
// The reference reads WebGraph-converted text files (tool/convert.sh:76-79,
// 106-116: symmetrise, drop self-loops, canonicalise, sort -u).  No such data
// exists offline, so graphs come from these generators and go through the
// same post-processing: symmetrise, drop self-loops, dedupe, sort rows.
#pragma once
#include <cstdint>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define DMOLP_HD __host__ __device__ __forceinline__
#else
#define DMOLP_HD inline
#endif

namespace dmolp {

// SplitMix64 output function (Steele, Lea, Flood 2014).
DMOLP_HD uint64_t mix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xbf58476d1ce4e5b9ULL;
    z ^= z >> 27;
    z *= 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return z;
}

// Counter-based stream: the ctr-th output of SplitMix64 seeded by `key`.
DMOLP_HD uint64_t rand_at(uint64_t key, uint64_t ctr) {
    return mix64(key + (ctr + 1) * 0x9E3779B97F4A7C15ULL);
}

DMOLP_HD uint64_t stream_key(uint64_t seed) { return mix64(seed ^ 0x6A09E667F3BCC909ULL); }

// High 64 bits of a 64x64 product: maps a uniform 64-bit word to [0, range).
DMOLP_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// R-MAT quadrant thresholds as 32-bit integers, so host and device draw
// identical edges: quadrant = 0 if r < t0, 1 if r < t1, 2 if r < t2, else 3.
struct RmatThresholds {
    uint32_t t0, t1, t2;
};

// One directed R-MAT sample (no noise, no permutation).  Level l draws 32
// bits; two levels share one 64-bit word.
DMOLP_HD void rmat_edge(uint64_t key, int scale, RmatThresholds th, uint64_t e,
                        uint32_t *u_out, uint32_t *v_out) {
    uint32_t u = 0, v = 0;
    const uint64_t words = (uint64_t)((scale + 1) / 2);
    uint64_t w = 0;
    for (int l = 0; l < scale; ++l) {
        if ((l & 1) == 0) w = rand_at(key, e * words + (uint64_t)(l >> 1));
        uint32_t r = (l & 1) ? (uint32_t)w : (uint32_t)(w >> 32);
        uint32_t q = r < th.t0 ? 0u : (r < th.t1 ? 1u : (r < th.t2 ? 2u : 3u));
        const int bit = scale - 1 - l;
        u |= (q >> 1) << bit;
        v |= (q & 1u) << bit;
    }
    *u_out = u;
    *v_out = v;
}

// Chung-Lu endpoint: the vertex whose integer-weight interval contains a
// uniform draw.  cdf has n+1 entries, cdf[0] = 0, cdf[n] = total.
DMOLP_HD uint32_t chung_lu_pick(const uint64_t *cdf, int64_t n, uint64_t r) {
    const uint64_t target = mulhi64(r, cdf[n]);
    int64_t lo = 0, hi = n;  // invariant: cdf[lo] <= target < cdf[hi]
    while (hi - lo > 1) {
        int64_t mid = lo + ((hi - lo) >> 1);
        if (cdf[mid] <= target) lo = mid; else hi = mid;
    }
    return (uint32_t)lo;
}

}  // namespace dmolp
