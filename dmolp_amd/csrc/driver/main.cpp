// main.cpp -- the reference's command line (main.cpp:123-205) over the
// drop-in run_phase1 / run_phase2:
//   hpc_partitioning [--verbose] -m <bfs|random> <graph_file> <k> <theta>
// One process on GPU 0.  Output as the reference's main.cpp:280-283:
// Phase1Metrics::print, then printComparisonReport.  Extra options:
//   --rounds R, --epsilon E  the round limit and convergence epsilon (the
//                            reference fixes 500 and 0.03, phase2.cpp:263-264)
//   --labels-out FILE        write the final labels (int32, native order)
//   --layout graph|wide      run on the reference's int32 Graph or on
//                            GraphWide (the default; any size)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iostream>
#include <string>
#include <vector>

#include "../../../include/graph_types.h"
#include "../../../include/phase1/phase1.h"
#include "../../../include/phase2/phase2.h"

namespace {

struct Options {
    bool verbose = false;
    bool random_mode = false;
    const char *graph_file = nullptr;
    int num_partitions = 0;
    int theta = 0;
    int rounds = 500;
    double epsilon = 0.03;
    const char *labels_out = nullptr;
    bool int32_graph = false;
};

bool parse(int argc, char **argv, Options &o) {
    std::vector<const char *> pos;
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        if (!std::strcmp(a, "--")) {
            for (int j = i + 1; j < argc; ++j) pos.push_back(argv[j]);
            break;
        }
        if (!std::strcmp(a, "--verbose") || !std::strcmp(a, "-v")) { o.verbose = true; continue; }
        const char *mode = nullptr;
        if (!std::strcmp(a, "-m")) {
            if (i + 1 >= argc) { std::cerr << "error: -m needs a value\n"; return false; }
            mode = argv[++i];
        } else if (!std::strncmp(a, "--mode=", 7)) {
            mode = a + 7;
        }
        if (mode) {
            if (!std::strcmp(mode, "bfs")) o.random_mode = false;
            else if (!std::strcmp(mode, "random")) o.random_mode = true;
            else { std::cerr << "error: mode must be bfs or random: " << mode << "\n"; return false; }
            continue;
        }
        if (!std::strcmp(a, "--rounds") && i + 1 < argc) { o.rounds = std::atoi(argv[++i]); continue; }
        if (!std::strcmp(a, "--epsilon") && i + 1 < argc) { o.epsilon = std::atof(argv[++i]); continue; }
        if (!std::strcmp(a, "--labels-out") && i + 1 < argc) { o.labels_out = argv[++i]; continue; }
        if (!std::strcmp(a, "--layout") && i + 1 < argc) {
            const char *l = argv[++i];
            if (!std::strcmp(l, "graph")) o.int32_graph = true;
            else if (!std::strcmp(l, "wide")) o.int32_graph = false;
            else { std::cerr << "error: layout must be graph or wide: " << l << "\n"; return false; }
            continue;
        }
        if (a[0] == '-') { std::cerr << "error: unknown option: " << a << "\n"; return false; }
        pos.push_back(a);
    }
    if (pos.size() != 3) { std::cerr << "error: expected exactly 3 positional arguments\n"; return false; }
    o.graph_file = pos[0];
    o.num_partitions = std::atoi(pos[1]);
    o.theta = std::atoi(pos[2]);
    // The reference accepts theta = 0 (main.cpp:188), which never finishes
    // (partition.cpp:427); it is rejected here.
    if (o.num_partitions <= 0 || o.theta <= 0) {
        std::cerr << "error: num_partitions > 0 and theta > 0 required\n";
        return false;
    }
    return true;
}

template <typename G>
int run(const Options &o) {
    G g;
    GhostNodes ghosts;
    std::cout << "\n=== Phase 1: Initial Partitioning & Distribution ===\n";
    Phase1Metrics m1 = run_phase1(0, 1, o.graph_file, o.num_partitions, o.theta, o.random_mode,
                                  o.verbose, g, ghosts);
    PartitioningMetrics p1(m1, o.num_partitions);
    std::cout << "\n=== Phase 2: Label Propagation ===\n";
    PartitioningMetrics p2 = run_phase2_ex(0, 1, o.num_partitions, g, ghosts, 0, o.rounds, o.epsilon);
    m1.print();
    printComparisonReport(p1, p2);
    if (o.labels_out) {
        FILE *fp = std::fopen(o.labels_out, "wb");
        const size_t n = g.vertex_labels.size();
        const bool ok = fp && std::fwrite(g.vertex_labels.data(), sizeof(int32_t), n, fp) == n;
        if (fp) std::fclose(fp);
        if (!ok) { std::cerr << "error: cannot write " << o.labels_out << "\n"; return 2; }
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    Options o;
    if (!parse(argc, argv, o)) {
        std::cout << "Usage: hpc_partitioning [--verbose] -m <bfs|random> <graph_file> "
                     "<num_partitions> <theta> [--rounds R] [--epsilon E] "
                     "[--labels-out FILE] [--layout graph|wide]\n";
        return 1;
    }
    try {
        return o.int32_graph ? run<Graph>(o) : run<GraphWide>(o);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 2;
    }
    return 0;
}
