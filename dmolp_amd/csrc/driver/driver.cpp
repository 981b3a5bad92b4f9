// driver.cpp -- run_phase1 / run_phase2 with the reference's signatures
// (include/phase1/phase1.h:7-16, include/phase2/phase2.h:7-13) over the C
// ABI.  One process, one GPU: mpi_rank must be 0 and mpi_size 1; the ghost
// set of a single rank is empty.  Failures the reference turns into
// MPI_Abort (phase2.cpp:239-250) or exit(1) throw std::runtime_error here.
#include <chrono>
#include <cstdio>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/dmolp_hip.h"
#include "../../../include/graph_types.h"
#include "../../../include/phase1/phase1.h"
#include "../../../include/phase2/phase2.h"

namespace {

dmolp_ctx *context_for(int device) {
    static std::mutex mu;
    static std::map<int, dmolp_ctx *> ctxs;
    std::lock_guard<std::mutex> lock(mu);
    auto it = ctxs.find(device);
    if (it != ctxs.end()) return it->second;
    dmolp_ctx *c = nullptr;
    const int rc = dmolp_ctx_create(device, &c);
    if (rc != DMOLP_OK)
        throw std::runtime_error(std::string("dmolp_ctx_create: ") + dmolp_last_error(nullptr));
    ctxs[device] = c;
    return c;
}

void check(int rc, dmolp_ctx *ctx, const char *what) {
    if (rc != DMOLP_OK)
        throw std::runtime_error(std::string(what) + ": " + dmolp_status_string(rc) + ": " +
                                 dmolp_last_error(ctx));
}

void require_single_rank(int mpi_rank, int mpi_size) {
    if (mpi_rank != 0 || mpi_size != 1)
        throw std::runtime_error("the library partitions on one GPU in one process (mpi_size 1)");
}

// Offsets of a graph as the ABI takes them (offset_bytes 4 or 8).
inline const void *offsets_of(const Graph &g) { return g.row_ptr.data(); }
inline const void *offsets_of(const GraphWide &g) { return g.row_ptr.data(); }
inline int offset_bytes(const Graph &) { return 4; }
inline int offset_bytes(const GraphWide &) { return 8; }

// Upload a graph (ghost ids would be >= num_vertices; a single rank has
// none).
template <typename G>
dmolp_graph *upload(dmolp_ctx *ctx, const G &g) {
    dmolp_graph *dg = nullptr;
    if ((int64_t)g.row_ptr.size() != (int64_t)g.num_vertices + 1)
        throw std::runtime_error("row_ptr must hold num_vertices + 1 entries");
    check(dmolp_graph_upload(ctx, g.num_vertices, (int64_t)g.col_indices.size(), offsets_of(g),
                             offset_bytes(g), g.col_indices.data(), &dg),
          ctx, "dmolp_graph_upload");
    return dg;
}

// Whether a CSR of n vertices and nnz entries fits the graph type.
inline bool fits(const Graph *, int64_t n, int64_t nnz) {
    return n < (1ll << 31) && nnz < (1ll << 31);
}
inline bool fits(const GraphWide *, int64_t n, int64_t) { return n < (1ll << 31); }

template <typename G>
Phase1Metrics phase1_impl(int mpi_rank, int mpi_size, const char *graph_file, int num_parts,
                          int theta, bool mode, bool verbose, G &local_graph,
                          GhostNodes &ghost_nodes) {
    require_single_rank(mpi_rank, mpi_size);
    dmolp_ctx *ctx = context_for(0);
    const auto t0 = std::chrono::high_resolution_clock::now();
    dmolp_csr csr{};
    int32_t cached = 0;
    check(dmolp_load_adjacency_cached(graph_file, &csr, &cached), ctx, "dmolp_load_adjacency_cached");
    const auto t1 = std::chrono::high_resolution_clock::now();
    if (!fits(&local_graph, csr.n, csr.nnz)) {
        dmolp_csr_free(&csr);
        throw std::runtime_error("graph has 2^31 or more CSR entries: the int32 Graph cannot hold "
                                 "it; call the GraphWide overload of run_phase1");
    }
    dmolp_graph *g = nullptr;
    int rc = dmolp_graph_upload(ctx, csr.n, csr.nnz, csr.off, 8, csr.col, &g);
    if (rc != DMOLP_OK) { dmolp_csr_free(&csr); check(rc, ctx, "dmolp_graph_upload"); }
    const int k = num_parts;
    std::vector<int32_t> labels((size_t)csr.n);
    std::vector<int64_t> vc((size_t)k), ec((size_t)k);
    dmolp_phase1_metrics m{};
    rc = mode ? dmolp_phase1_random(ctx, g, k, labels.data(), vc.data(), ec.data(), &m)
              : dmolp_phase1(ctx, g, k, theta, labels.data(), nullptr, vc.data(), ec.data(), &m);
    dmolp_graph_free(ctx, g);
    if (rc != DMOLP_OK) { dmolp_csr_free(&csr); check(rc, ctx, "dmolp_phase1"); }
    const auto t2 = std::chrono::high_resolution_clock::now();

    local_graph = G{};
    ghost_nodes.clear();
    local_graph.num_vertices = (decltype(local_graph.num_vertices))csr.n;
    local_graph.num_edges = (decltype(local_graph.num_edges))csr.nnz;
    local_graph.row_ptr.assign(csr.off, csr.off + csr.n + 1);
    local_graph.col_indices.assign(csr.col, csr.col + csr.nnz);
    local_graph.global_ids.resize((size_t)csr.n);
    for (int64_t v = 0; v < csr.n; ++v) local_graph.global_ids[(size_t)v] = (int)v;
    local_graph.vertex_labels.assign(labels.begin(), labels.end());
    dmolp_csr_free(&csr);

    Phase1Metrics out;
    out.initial_edge_cut = (int)m.initial_edge_cut;
    out.initial_vertex_balance = m.initial_vertex_balance;
    out.initial_edge_balance = m.initial_edge_balance;
    out.loading_time_ms = std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
    out.distribution_time_ms = std::chrono::duration_cast<std::chrono::milliseconds>(t2 - t1).count();
    out.total_vertices = (int)m.total_vertices;
    out.total_edges = (uint64_t)m.total_edges;
    out.partition_vertex_counts.assign(vc.begin(), vc.end());
    out.partition_edge_counts.assign(ec.begin(), ec.end());
    if (verbose)
        std::printf("[phase1] graph %s (%s), seeds found %d, fallback %d, rounds %lld\n",
                    graph_file, cached ? "binary cache" : "text", m.n_seeds_found,
                    m.n_seeds_fallback, (long long)m.rounds);
    return out;
}

template <typename G>
PartitioningMetrics phase2_impl(int mpi_rank, int mpi_size, int num_partitions, G &local_graph,
                                GhostNodes &ghost_nodes, int gpu_id, int max_rounds,
                                double epsilon) {
    require_single_rank(mpi_rank, mpi_size);
    if (!ghost_nodes.global_ids.empty())
        throw std::runtime_error("a single rank has no ghost vertices");
    if ((int64_t)local_graph.vertex_labels.size() != (int64_t)local_graph.num_vertices)
        throw std::runtime_error("vertex_labels must hold num_vertices entries");
    dmolp_ctx *ctx = context_for(gpu_id);
    const auto t0 = std::chrono::high_resolution_clock::now();
    dmolp_graph *g = upload(ctx, local_graph);
    int rc = dmolp_graph_set_labels(ctx, g, local_graph.vertex_labels.data());
    dmolp_phase2_metrics m{};
    if (rc == DMOLP_OK)
        rc = dmolp_phase2(ctx, g, num_partitions, max_rounds, epsilon, nullptr, nullptr, &m, nullptr);
    if (rc == DMOLP_OK) rc = dmolp_graph_get_labels(ctx, g, local_graph.vertex_labels.data());
    dmolp_graph_free(ctx, g);
    check(rc, ctx, "dmolp_phase2");
    const auto t1 = std::chrono::high_resolution_clock::now();
    PartitioningMetrics out;
    out.edge_cut = (int)m.edge_cut;
    out.vertex_balance = m.vertex_balance;
    out.edge_balance = m.edge_balance;
    out.loading_time_ms = std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
    out.distribution_time_ms = 0;
    out.num_partitions = num_partitions;
    out.total_vertices = (int)m.total_vertices;
    out.total_edges = (uint64_t)m.total_edges;
    return out;
}

}  // namespace

Phase1Metrics run_phase1(int mpi_rank, int mpi_size, const char *graph_file, int num_parts,
                         int theta, bool mode, bool verbose, Graph &local_graph,
                         GhostNodes &ghost_nodes) {
    return phase1_impl(mpi_rank, mpi_size, graph_file, num_parts, theta, mode, verbose, local_graph,
                       ghost_nodes);
}

Phase1Metrics run_phase1(int mpi_rank, int mpi_size, const char *graph_file, int num_parts,
                         int theta, bool mode, bool verbose, GraphWide &local_graph,
                         GhostNodes &ghost_nodes) {
    return phase1_impl(mpi_rank, mpi_size, graph_file, num_parts, theta, mode, verbose, local_graph,
                       ghost_nodes);
}

PartitioningMetrics run_phase2_ex(int mpi_rank, int mpi_size, int num_partitions,
                                  Graph &local_graph, GhostNodes &ghost_nodes, int gpu_id,
                                  int max_rounds, double epsilon) {
    return phase2_impl(mpi_rank, mpi_size, num_partitions, local_graph, ghost_nodes, gpu_id,
                       max_rounds, epsilon);
}

PartitioningMetrics run_phase2_ex(int mpi_rank, int mpi_size, int num_partitions,
                                  GraphWide &local_graph, GhostNodes &ghost_nodes, int gpu_id,
                                  int max_rounds, double epsilon) {
    return phase2_impl(mpi_rank, mpi_size, num_partitions, local_graph, ghost_nodes, gpu_id,
                       max_rounds, epsilon);
}

// The reference's constants (phase2.cpp:263-264): 500 rounds, epsilon 0.03.
PartitioningMetrics run_phase2(int mpi_rank, int mpi_size, int num_partitions, Graph &local_graph,
                               GhostNodes &ghost_nodes, int gpu_id) {
    return run_phase2_ex(mpi_rank, mpi_size, num_partitions, local_graph, ghost_nodes, gpu_id, 500,
                         0.03);
}

PartitioningMetrics run_phase2(int mpi_rank, int mpi_size, int num_partitions,
                               GraphWide &local_graph, GhostNodes &ghost_nodes, int gpu_id) {
    return run_phase2_ex(mpi_rank, mpi_size, num_partitions, local_graph, ghost_nodes, gpu_id, 500,
                         0.03);
}
