// graph_types.h -- the data types of DMOLP's partitioner path
// (graph_types.h:47-57 Delta, :126-190 Graph, :196-229 GhostNodes,
// :232-273 PartitionStats, :276-309 Phase1Metrics, :312-346
// PartitioningMetrics), with the same field names and types so code written
// against the reference compiles unchanged.  GraphWide is the int64-offset
// form for graphs with 2^31 or more CSR entries (the reference's int
// row_ptr overflows there).
#pragma once
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <unordered_map>
#include <vector>

struct Delta {
    int gid = -1;
    int new_label = -1;
    Delta() = default;
    Delta(int g, int l) : gid(g), new_label(l) {}
};

// Per-rank CSR: local ids 0..num_vertices-1, ghost ids num_vertices.. in
// col_indices; labels of owned vertices only.
struct Graph {
    int num_vertices = 0;
    int num_edges = 0;
    std::vector<int> row_ptr;
    std::vector<int> col_indices;
    std::vector<int> global_ids;
    std::vector<int> vertex_labels;

    void clear() {
        num_vertices = 0;
        num_edges = 0;
        row_ptr.clear();
        col_indices.clear();
        global_ids.clear();
        vertex_labels.clear();
    }

    void print() const {
        std::cout << "=== Graph ===\n"
                  << "Number of vertices: " << num_vertices << "\n"
                  << "Number of edges: " << num_edges << "\n";
        for (int i = 0; i < num_vertices && i + 1 < (int)row_ptr.size(); ++i) {
            std::cout << "Vertex " << i << " (global_id=" << global_ids[i]
                      << ", label=" << vertex_labels[i] << "):";
            for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) std::cout << ' ' << col_indices[e];
            std::cout << '\n';
        }
        std::cout << std::endl;
    }
};

// The same graph with 64-bit offsets.
struct GraphWide {
    int64_t num_vertices = 0;
    int64_t num_edges = 0;
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> col_indices;
    std::vector<int32_t> global_ids;
    std::vector<int32_t> vertex_labels;
};

struct GhostNodes {
    std::vector<int> global_ids;
    std::vector<int> ghost_labels;
    std::unordered_map<int, int> global_to_local;

    void clear() {
        global_ids.clear();
        ghost_labels.clear();
        global_to_local.clear();
    }

    void print() const {
        std::cout << "=== Ghost Nodes ===\n"
                  << "Number of ghost nodes: " << global_ids.size() << "\n" << std::endl;
    }
};

struct PartitionStats {
    std::vector<int> local_vertex_counts;
    std::vector<int> local_edge_counts;
    std::vector<int> global_vertex_counts;
    std::vector<int> global_edge_counts;
    int total_vertices = 0;
    uint64_t total_edges = 0;
    double expected_vertices = 0.0;
    double expected_edges = 0.0;
};

struct Phase1Metrics {
    int initial_edge_cut = 0;
    double initial_vertex_balance = 0.0;
    double initial_edge_balance = 0.0;
    long loading_time_ms = 0;
    long distribution_time_ms = 0;
    int total_vertices = 0;
    uint64_t total_edges = 0;
    std::vector<int> partition_vertex_counts;
    std::vector<int> partition_edge_counts;

    void print() const {
        std::cout << "=== Phase 1 Metrics ===\n"
                  << "Initial edge cut: " << initial_edge_cut << "\n"
                  << "Initial vertex balance: " << std::fixed << std::setprecision(4)
                  << initial_vertex_balance << "\n"
                  << "Initial edge balance: " << std::fixed << std::setprecision(4)
                  << initial_edge_balance << "\n"
                  << "Loading time: " << loading_time_ms << " ms\n"
                  << "Distribution time: " << distribution_time_ms << " ms\n"
                  << "Total vertices: " << total_vertices << "\n"
                  << "Total edges: " << total_edges << "\n"
                  << "Partition vertex counts: ";
        for (int c : partition_vertex_counts) std::cout << c << ' ';
        std::cout << "\nPartition edge counts: ";
        for (int c : partition_edge_counts) std::cout << c << ' ';
        std::cout << "\n" << std::endl;
    }
};

struct PartitioningMetrics {
    int edge_cut = 0;
    double vertex_balance = 0.0;
    double edge_balance = 0.0;
    long loading_time_ms = 0;
    long distribution_time_ms = 0;
    int num_partitions = 0;
    int total_vertices = 0;
    uint64_t total_edges = 0;

    PartitioningMetrics() = default;
    PartitioningMetrics(const Phase1Metrics &m, int num_parts)
        : edge_cut(m.initial_edge_cut),
          vertex_balance(m.initial_vertex_balance),
          edge_balance(m.initial_edge_balance),
          loading_time_ms(m.loading_time_ms),
          distribution_time_ms(m.distribution_time_ms),
          num_partitions(num_parts),
          total_vertices(m.total_vertices),
          total_edges(m.total_edges) {}

    void print() const {
        std::cout << "=== Partitioning Metrics ===\n"
                  << "Edge cut: " << edge_cut << "\n"
                  << "Vertex balance: " << std::fixed << std::setprecision(4) << vertex_balance << "\n"
                  << "Edge balance: " << std::fixed << std::setprecision(4) << edge_balance << "\n"
                  << "Loading time: " << loading_time_ms << " ms\n"
                  << "Distribution time: " << distribution_time_ms << " ms\n"
                  << "Number of partitions: " << num_partitions << "\n"
                  << "Total vertices: " << total_vertices << "\n"
                  << "Total edges: " << total_edges << "\n" << std::endl;
    }
};
