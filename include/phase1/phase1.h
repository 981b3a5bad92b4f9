// phase1.h -- run_phase1 with the reference's signature
// (include/phase1/phase1.h:7-16), implemented over the C ABI.
#pragma once
#include "../graph_types.h"

// mpi_rank must be 0 and mpi_size 1 (one GPU, no MPI).  mode false = hub
// seeds (bfs), true = random labels gid % k (random.cpp:24-127).
// graph_file uses the reference's text adjacency format.
// The int32 Graph holds fewer than 2^31 CSR entries; a larger file fails
// with std::runtime_error naming the GraphWide overload.
Phase1Metrics run_phase1(int mpi_rank, int mpi_size, const char *graph_file, int num_parts,
                         int theta, bool mode, bool verbose, Graph &local_graph,
                         GhostNodes &ghost_nodes);

// The same for any size (int64 offsets).  No counterpart in the reference,
// whose int row_ptr overflows at 2^31 entries (graph_types.h:127-130).
Phase1Metrics run_phase1(int mpi_rank, int mpi_size, const char *graph_file, int num_parts,
                         int theta, bool mode, bool verbose, GraphWide &local_graph,
                         GhostNodes &ghost_nodes);
