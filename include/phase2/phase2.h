// phase2.h -- run_phase2 with the reference's signature
// (include/phase2/phase2.h:7-13), implemented over the C ABI.
#pragma once
#include "../graph_types.h"

// The reference's constants (phase2.cpp:263-264): 500 rounds, epsilon 0.03.
PartitioningMetrics run_phase2(int mpi_rank, int mpi_size, int num_partitions, Graph &local_graph,
                               GhostNodes &ghost_nodes, int gpu_id);

// Same with the round limit and convergence epsilon as parameters.
PartitioningMetrics run_phase2_ex(int mpi_rank, int mpi_size, int num_partitions,
                                  Graph &local_graph, GhostNodes &ghost_nodes, int gpu_id,
                                  int max_rounds, double epsilon);

// The same on a GraphWide (int64 offsets, any size).
PartitioningMetrics run_phase2(int mpi_rank, int mpi_size, int num_partitions,
                               GraphWide &local_graph, GhostNodes &ghost_nodes, int gpu_id);
PartitioningMetrics run_phase2_ex(int mpi_rank, int mpi_size, int num_partitions,
                                  GraphWide &local_graph, GhostNodes &ghost_nodes, int gpu_id,
                                  int max_rounds, double epsilon);

// printComparisonReport (utils.cpp:337-394, declared in utils.hpp): the
// phase-1 versus phase-2 table the reference prints at the end of a run,
// line for line in the reference's format so its log parsers keep working.
void printComparisonReport(const PartitioningMetrics &m1, const PartitioningMetrics &m2);
