// internal.hpp -- context, graph handle and device-view structures behind
// the C ABI in include/dmolp_hip.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dmolp_hip.h"

namespace dmolp {

// Error carried from deep inside the library to the ABI boundary, where it
// becomes a status code and the context's last-error message.
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};

#define DMOLP_HIP_CHECK(expr)                                                              \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            throw ::dmolp::Error(e_ == hipErrorOutOfMemory ? DMOLP_E_NOMEM : DMOLP_E_HIP,  \
                                 std::string(#expr) + ": " + hipGetErrorString(e_) +      \
                                     " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
        }                                                                                  \
    } while (0)

#define DMOLP_REQUIRE(cond, status, msg)                              \
    do {                                                              \
        if (!(cond)) throw ::dmolp::Error((status), std::string(msg)); \
    } while (0)

// Device memory recycling for one context.  A released block is kept and
// handed out again to a later allocation of the same context, so that a
// steady stream of graphs and calls makes no hipMalloc or hipFree (hipFree
// waits for the whole device, which would serialise served jobs).  Reuse is
// safe in stream order: every kernel and copy that touches a context's
// memory is queued on that context's stream, so a block's next owner's work
// runs after its previous owner's.  Graphs used with a second context are
// freed behind a device-wide synchronisation (dmolp_graph_free).  Blocks are
// rounded to size classes (powers of two below 1 MiB, then whole MiB) and a
// cached block serves a request it exceeds by at most a quarter.  At most
// `limit` bytes stay cached; past it a released block is freed at once.
class DevPool {
  public:
    explicit DevPool(size_t limit) : limit_(limit) {}
    DevPool(const DevPool &) = delete;
    DevPool &operator=(const DevPool &) = delete;
    ~DevPool() { trim(); }
    static size_t round(size_t b) {
        if (b < (1u << 20)) {
            size_t r = 512;
            while (r < b) r <<= 1;
            return r;
        }
        return (b + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
    }
    // A cached block of at least `want` bytes (a rounded size), or nullptr.
    void *take(size_t want, size_t &cap) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.lower_bound(want);
        const size_t slack = want < (1u << 20) ? 0 : want / 4;
        if (it == free_.end() || it->first > want + slack) return nullptr;
        void *p = it->second;
        cap = it->first;
        cached_ -= cap;
        free_.erase(it);
        return p;
    }
    void give(void *p, size_t cap) {
        std::lock_guard<std::mutex> lk(mu_);
        if (closed_ || cached_ + cap > limit_) {
            (void)hipFree(p);
            return;
        }
        free_.emplace(cap, p);
        cached_ += cap;
    }
    // Frees every cached block (hipFree waits for the device first).
    void trim() {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto &kv : free_) (void)hipFree(kv.second);
        free_.clear();
        cached_ = 0;
    }
    // The owning context is going away: free the cache, and from now on free
    // blocks as they come back.
    void close() {
        trim();
        std::lock_guard<std::mutex> lk(mu_);
        closed_ = true;
    }

  private:
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    size_t cached_ = 0;
    size_t limit_;
    bool closed_ = false;
};

// The pool that allocations on this thread draw from, and its context's
// stream: set for the duration of an ABI call (PoolScope); none outside
// calls.
inline const std::shared_ptr<DevPool> *&tls_pool() {
    static thread_local const std::shared_ptr<DevPool> *p = nullptr;
    return p;
}
inline hipStream_t &tls_stream() {
    static thread_local hipStream_t s = nullptr;
    return s;
}
struct PoolScope {
    const std::shared_ptr<DevPool> *prev;
    hipStream_t prev_stream;
    PoolScope(const std::shared_ptr<DevPool> *p, hipStream_t s)
        : prev(tls_pool()), prev_stream(tls_stream()) {
        tls_pool() = p;
        tls_stream() = s;
    }
    ~PoolScope() {
        tls_pool() = prev;
        tls_stream() = prev_stream;
    }
    PoolScope(const PoolScope &) = delete;
    PoolScope &operator=(const PoolScope &) = delete;
};
// DMOLP_POOL_POISON=1 (read once): every block an allocation hands out,
// fresh or reused, is first filled with 0xa5 bytes on the call's stream, so
// that code reading memory it never wrote fails in tests instead of passing
// on the zeros of fresh pages.
bool pool_poison();

// Owning device allocation: drawn from the calling context's pool when
// there is one, and given back to that pool when released.
struct DevMem {
    void *p = nullptr;
    size_t bytes = 0;  // requested size
    size_t cap = 0;    // size of the block behind p
    std::shared_ptr<DevPool> pool;
    DevMem() = default;
    explicit DevMem(size_t b) { alloc(b); }
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    DevMem(DevMem &&o) noexcept : p(o.p), bytes(o.bytes), cap(o.cap), pool(std::move(o.pool)) {
        o.p = nullptr;
        o.bytes = o.cap = 0;
    }
    DevMem &operator=(DevMem &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            cap = o.cap;
            pool = std::move(o.pool);
            o.p = nullptr;
            o.bytes = o.cap = 0;
        }
        return *this;
    }
    ~DevMem() { release(); }
    void alloc(size_t b) {
        release();
        if (b == 0) b = 256;
        const size_t r = DevPool::round(b);
        const std::shared_ptr<DevPool> *tp = tls_pool();
        if (tp && *tp) {
            size_t c = 0;
            if (void *q = (*tp)->take(r, c)) {
                p = q;
                cap = c;
                bytes = b;
                pool = *tp;
                if (pool_poison()) DMOLP_HIP_CHECK(hipMemsetAsync(p, 0xa5, cap, tls_stream()));
                return;
            }
        }
        hipError_t e = hipMalloc(&p, r);
        if (e == hipErrorOutOfMemory && tp && *tp) {  // give the cache back and retry once
            (void)hipGetLastError();
            (*tp)->trim();
            e = hipMalloc(&p, r);
        }
        if (e != hipSuccess) {
            p = nullptr;
            DMOLP_HIP_CHECK(e);
        }
        bytes = b;
        cap = r;
        if (tp) pool = *tp;
        if (pool_poison()) DMOLP_HIP_CHECK(hipMemsetAsync(p, 0xa5, cap, tls_stream()));
    }
    void grow(size_t b) {
        if (!p || bytes < b) alloc(b);
    }
    void release() {
        if (!p) return;
        if (pool) pool->give(p, cap);
        else (void)hipFree(p);
        p = nullptr;
        bytes = cap = 0;
        pool.reset();
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Device view of a vertex work list.  Rows of at most `heavy_degree` edges
// are "light" and processed 64 list entries per wave; longer rows are
// "heavy": they get a slot and are split into chunks of `chunk_edges`
// edges, one chunk per wave.
// LC_FILL / LC_HFILL: fill cursors of light ids / heavy slots rebuilt from
// a bitmap (k_b_list).
enum ListCount {
    LC_LIGHT = 0, LC_SLOTS = 1, LC_CHUNKS = 2, LC_DEG = 3, LC_HDEG = 4, LC_FILL = 5, LC_HFILL = 6,
    LC_N = 7
};
struct ListDev {
    int32_t *light;            // light vertex ids
    int32_t *hv;               // heavy slot -> vertex
    int32_t *hn;               // heavy slot -> chunk count
    int64_t *hc;               // chunk items: (slot << 32) | chunk index
    unsigned long long *cnt;   // ListCount counters (device)
};

struct List {
    DevMem light, hv, hn, hc, cnt;
    ListDev dev{};
    void alloc(int64_t cap_light, int64_t cap_slots, int64_t cap_chunks) {
        light.alloc(sizeof(int32_t) * (size_t)(cap_light > 0 ? cap_light : 1));
        hv.alloc(sizeof(int32_t) * (size_t)(cap_slots > 0 ? cap_slots : 1));
        hn.alloc(sizeof(int32_t) * (size_t)(cap_slots > 0 ? cap_slots : 1));
        hc.alloc(sizeof(int64_t) * (size_t)(cap_chunks > 0 ? cap_chunks : 1));
        cnt.alloc(sizeof(unsigned long long) * 8);
        dev = ListDev{light.as<int32_t>(), hv.as<int32_t>(), hn.as<int32_t>(), hc.as<int64_t>(),
                      cnt.as<unsigned long long>()};
    }
};

struct Tuning {
    int heavy_degree = 2048;  // rows longer than this are heavy (kept int32 histograms, chunked)
    int chunk_edges = 4096;
    int block_threads = 256;
    int grid_blocks = 0;  // 0: derived from the CU count
    double pull_fraction = 0.05;  // pull candidates when B_{t-1} covers this share of entries
    int64_t batch_work_per_block = 32768;  // batched launches: vertices + nnz/4 per block
    double hist_recount_share = 0.2;  // heavy histograms: recount when changes hold >= this share
    double hist_push_share = 0.02;    // ... push the changes when they hold <= this share
    bool light_hist = true;  // light rows keep their label counts across rounds
};

struct GraphDev {
    int64_t n;
    int64_t nnz;
    const int64_t *off;
    const int32_t *col;
};

// Per-graph phase-2 state, allocated on first use for a given k.
struct Phase2Work;
struct Phase1Work;

// How many blocks of a multi-job launch each job gets.
enum LaunchPolicy {
    P_GRID = 0,  // grid-stride kernels over vertices or edges
    P_HIST1,     // level-1 selection histogram
    P_GROUPS,    // one block per selection group
    P_ONE,       // one block per job
    P_HEAVY,     // one thread per heavy row
    P_NPOL
};

// Device block-start arrays of a job set, one per launch policy.
struct LaunchSet {
    int jobs = 0;
    int total[P_NPOL] = {};
    DevMem bstart;  // int32 [P_NPOL][jobs + 1]
    const int32_t *bs(int pol) const { return bstart.as<int32_t>() + (size_t)pol * (jobs + 1); }
};

// A captured phase-2 round (or batch of phase-1 rounds): the graph, its
// executable and the launch shape (`key`) it was captured for.  The graph is
// kept as long as the executable.  Owners release it only after the stream
// has drained (graph free and context teardown synchronise first; a
// re-capture waits for the stream).
struct RoundGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t key = 0;
    RoundGraph() = default;
    RoundGraph(const RoundGraph &) = delete;
    RoundGraph &operator=(const RoundGraph &) = delete;
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key = 0;
    }
    ~RoundGraph() { release(); }
};

// Device records of a phase-1 job set, and its captured batch of expansion
// rounds: captured the second time the same launch shape is enqueued (a
// shape used once is launched directly, so job sets that change every call
// pay no capture).
struct P1SetDev {
    int G = 0;
    int passes = 0;
    bool any_heavy = false;
    int64_t max_n = 0;
    LaunchSet ls;
    DevMem jobs, sel, done, remaining;
    RoundGraph rounds;
    uint64_t seen_key = 0;
    int kmax = 0;
    DevMem run_out;  // results of p1_rounds_wg
    P1SetDev() = default;
    P1SetDev(const P1SetDev &) = delete;
    P1SetDev &operator=(const P1SetDev &) = delete;
};

// Device records and the captured round of the last phase-2 job set.
struct BatchState {
    DevMem jobs, sel, remaining, done;
    LaunchSet ls;
    RoundGraph round;
    BatchState() = default;
    BatchState(const BatchState &) = delete;
    BatchState &operator=(const BatchState &) = delete;
};

}  // namespace dmolp

struct dmolp_ctx {
    int device = 0;
    int cu_count = 256;
    hipStream_t stream = nullptr;
    std::string last_error;
    dmolp::Tuning tune;
    std::shared_ptr<dmolp::DevPool> pool;  // recycled device memory of this context
    dmolp::DevMem scratch;  // small per-call scratch (host<->device staging)
    dmolp::DevMem ops_dev;  // staging of OpBatch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Kernel-group timing of the last phase-2 run (profiling on): events
    // recorded around each group every round.
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::vector<std::pair<size_t, size_t>>> prof_marks;  // [group] -> (ev_a, ev_b) per round
    std::vector<std::vector<double>> prof_ms;                         // [group] -> ms per round
    hipEvent_t prof_event();
    int grid_blocks() const { return tune.grid_blocks > 0 ? tune.grid_blocks : cu_count * 8; }
    dmolp::P1SetDev p1_set;   // phase-1 job set of the plain and batched paths
    // A second stream (default priority) and the set it runs: a batch's
    // long one-workgroup phase-1 launch beside the main stream's launches.
    dmolp::P1SetDev p1_side;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    void ensure_side() {
        if (side) return;
        DMOLP_HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        DMOLP_HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        DMOLP_HIP_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    dmolp::BatchState batch;  // phase-2 job set of dmolp_batch_run
    // Stop polling of round loops: a device flag copied to pinned memory
    // behind each round, read one round later (see round_poll).
    int32_t *poll_host = nullptr;
    hipEvent_t poll_ev[2] = {nullptr, nullptr};
    // Pinned host staging for device-to-host copies of vertex arrays (grows,
    // freed with the context).
    int32_t *pin = nullptr;
    size_t pin_count = 0;
    int64_t hub_prefix = int64_t(1) << 16;  // hubs a phase-1 host search orders first (0: all)
    // Phase-1 hub searches of graphs with at least hub_dev_min vertices run
    // their partition steps on the device; ranges shorter than hub_host_max
    // are finished on the host.
    int64_t hub_dev_min = int64_t(1) << 18;
    int64_t hub_host_max = int64_t(1) << 15;
    int32_t *pinned_i32(size_t count) {
        if (count > pin_count) {
            if (pin) (void)hipHostFree(pin);
            pin = nullptr;
            pin_count = 0;
            DMOLP_HIP_CHECK(hipHostMalloc((void **)&pin, sizeof(int32_t) * (count > 0 ? count : 1),
                                          hipHostMallocDefault));
            pin_count = count;
        }
        return pin;
    }
};

struct dmolp_graph {
    // The pool of the context that created the graph; `mixed` is set when a
    // call of another context has queued work on it (freeing it then waits
    // for the whole device, not one stream).
    const dmolp::DevPool *pool = nullptr;
    mutable bool mixed = false;
    int device = 0;
    int64_t n = 0;
    int64_t nnz = 0;
    int32_t max_degree = 0;
    dmolp::DevMem off;     // int64[n + 1]
    dmolp::DevMem col;     // int32[nnz]
    dmolp::DevMem labels;  // int32[n], resident labels
    dmolp::DevMem saved;   // int32[n], label checkpoint
    bool labels_valid = false;
    // All vertices binned by degree (for full sweeps), built for one
    // (heavy_degree, chunk_edges) pair.
    dmolp::List all;
    int all_heavy_degree = -1, all_chunk_edges = -1;
    int64_t n_heavy = 0, n_heavy_chunks = 0;
    dmolp::DevMem hid;    // int32[n]: heavy slot of the all list, -1 for light rows
    dmolp::DevMem hbits;  // uint32[(n + 31) / 32]: bit v set for the heavy rows
    dmolp::DevMem first;  // int32[n]: v's first (smallest) neighbour, -1 for an empty row
    dmolp::DevMem deg;    // int32[n]: v's degree (rows shorter than 2^31 entries)
    bool first_valid = false;  // first and deg
    std::unique_ptr<dmolp::Phase2Work> p2;
    std::unique_ptr<dmolp::Phase1Work> p1;
    dmolp::GraphDev dev() const {
        return dmolp::GraphDev{n, nnz, off.as<int64_t>(), col.as<int32_t>()};
    }
    ~dmolp_graph();
};

namespace dmolp {
// Implemented in the .hip files.
void graph_prepare_bins(dmolp_ctx *ctx, dmolp_graph *g);  // builds g->all for ctx->tune
void graph_prepare_first(dmolp_ctx *ctx, dmolp_graph *g);  // g->first, g->deg
void graph_validate(dmolp_ctx *ctx, dmolp_graph *g);       // offsets, sorted rows, no self-loops
// The largest degree only, for graphs the library generated itself (sorted,
// deduplicated and loop-free by construction).
void graph_max_degree(dmolp_ctx *ctx, dmolp_graph *g);
int64_t bits_for(int64_t v);                                // bits needed to hold 0..v
// After enqueueing round t of a loop: queue a copy of the device flag behind
// it and read the copy queued after round t-1 (the host stays one round
// ahead, so at most one round runs past the stop).  True when that earlier
// flag says stop: nonzero (stop_at_zero false) or <= 0 (stop_at_zero true).
bool round_poll(dmolp_ctx *ctx, const int32_t *d_flag, int t, bool stop_at_zero);
// False when DMOLP_NO_GRAPHS=1 is set: the phase-1 and phase-2 rounds are
// then launched kernel by kernel instead of replayed as captured graphs
// (same kernels, same results).  For profilers whose graph-launch
// interception misbehaves: rocprofv3 --kernel-trace has ended processes
// with a host SIGSEGV inside its callback under hipGraphLaunch.
bool graphs_enabled();
// Records that ctx queues work on g: a graph made by another context is then
// freed behind a device-wide synchronisation.
inline void note_use(const dmolp_ctx *ctx, const dmolp_graph *g) {
    if (ctx && g && g->pool != ctx->pool.get()) g->mixed = true;
}
}  // namespace dmolp
