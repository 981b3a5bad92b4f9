// phase1.h -- run_phase1 with the reference's signature
// (include/phase1/phase1.h:7-16), implemented over the C ABI.
#pragma once
#include "../graph_types.h"

// mpi_rank must be 0 and mpi_size 1 (one GPU, no MPI).  mode false = hub
// seeds (bfs), true = random labels gid % k (random.cpp:24-127).
// graph_file uses the reference's text adjacency format.
Phase1Metrics run_phase1(int mpi_rank, int mpi_size, const char *graph_file, int num_parts,
                         int theta, bool mode, bool verbose, Graph &local_graph,
                         GhostNodes &ghost_nodes);
