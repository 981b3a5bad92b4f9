// hubs.hpp -- hub list and first seed of phase 1 (find_hub_nodes and
// find_first_seed, init.cpp:139-187): on the host for short degree
// sequences, and with the partition steps on the device for long ones.
#pragma once
#include <cstdint>
#include <thread>
#include <utility>
#include <vector>

#include "internal.hpp"

namespace dmolp {

// (vertex id, degree), the element type of the reference's sort; trivial,
// so large arrays start uninitialised.
struct DegPair {
    int first, second;
};
struct ByDegree {
    __host__ __device__ bool operator()(const DegPair &a, const DegPair &b) const {
        return a.second < b.second;
    }
};

// body(t) for t in [0, T) on T threads (the caller runs t = 0).
template <class F>
void run_threads(int T, F &&body) {
    if (T <= 1) { body(0); return; }
    std::vector<std::thread> th;
    th.reserve((size_t)T - 1);
    for (int t = 1; t < T; ++t) th.emplace_back([&body, t] { body(t); });
    body(0);
    for (auto &x : th) x.join();
}

// The number of hubs M of init.cpp:139-177 (the Lorenz tangent cut) from
// the vertex count, the degree total and the largest degree.
int64_t hub_count(int64_t n, int64_t degree_sum, int32_t degree_max);

// threads: host threads for the long partition steps and passes (1: all
// on the calling thread); the result does not depend on it.
// limit >= 0: only the first min(limit, M) hubs of the order; m_out gets M,
// the number of hubs.
std::vector<int32_t> find_hubs_host(const std::vector<int32_t> &deg, int threads = 1,
                                    int64_t limit = -1, int64_t *m_out = nullptr);
std::pair<int32_t, int32_t> first_seed_host(const std::vector<int32_t> &deg, int threads = 1);

// One job of a device hub search: the degrees are on the device.
struct HubSearch {
    int64_t n = 0;
    const int32_t *deg = nullptr;  // device, n entries
    int64_t limit = -1;            // as find_hubs_host
    std::vector<int32_t> hubs;     // the first min(limit, M) hubs, in order
    std::vector<int32_t> hub_deg;  // their degrees
    int64_t n_hubs = 0;            // M
    std::pair<int32_t, int32_t> fs{-1, -1};  // first seed and its degree
};

// The same results as find_hubs_host and first_seed_host for every job,
// several jobs together on the context's stream.  Each partition step of a
// range of at least `host_max` entries runs on the device (the ranges of
// all jobs of one step in one set of launches); shorter ranges are copied
// to the host and finished there by the same code as find_hubs_host.
void hub_search_device(dmolp_ctx *ctx, const std::vector<HubSearch *> &jobs, int64_t host_max,
                       int threads);

// The first m vertices by (degree descending, id ascending), with the same
// partition steps under that total order (the device does the long ranges,
// the host sorts the short ones).
struct TopSearch {
    int64_t n = 0;
    const int32_t *deg = nullptr;  // device, n entries
    int64_t m = 0;
    std::vector<int32_t> top;      // min(m, n) ids in that order
};
void top_by_degree_device(dmolp_ctx *ctx, const std::vector<TopSearch *> &jobs, int64_t host_max,
                          int threads);

}  // namespace dmolp
