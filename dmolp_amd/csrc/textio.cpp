// textio.cpp -- the reference's text adjacency format.
//
// Format (init.cpp:25-100, written by tool/convert.sh:132-176): the first
// line holds V; line i+1 lists the neighbours of vertex i.  As in the
// reference, the vertex count is the number of adjacency lines: the header
// is read but its value is replaced (init.cpp:96 overwrites V with the
// summed line count).  Differences from the reference's loader
// (init.cpp:15-23,53-82), on malformed input only:
//  - a last line without '\n' is kept (the reference drops it);
//  - a line longer than the 16 MB read buffer is kept whole (the reference
//    keeps only its tail, init.cpp:81);
//  - a line starting with a non-digit does not yield a phantom neighbour 0;
//  - a neighbour id above 2^31-1 (which the reference's int parser wraps)
//    or outside [0, V) fails the load with DMOLP_E_GRAPH.
// Rows are sorted and deduplicated as build_local_graph_and_ghosts does
// (partition.cpp:231-232) and self-loops are dropped (the device path
// requires none; random.cpp:90 drops them too).  tool/convert.sh writes
// files with neither (convert.sh:76-79,106-116); for a file that holds
// duplicate neighbours or self-loops, the reference's phase 1 takes degrees
// from the raw line lengths (init.cpp:86-90) and keeps self-loops in its
// CSR, whereas here every degree is the cleaned row's length.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

using dmolp::Error;

extern "C" int dmolp_read_adjacency_text(const char *path, dmolp_csr *out) {
    if (!path || !out) return DMOLP_E_ARG;
    std::memset(out, 0, sizeof(*out));
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return DMOLP_E_ARG;
    std::vector<char> buf(1 << 24);
    std::vector<int64_t> off{0};
    std::vector<int32_t> col;
    std::vector<int32_t> row;
    int64_t declared = -1, line = 0;
    int64_t cur = 0;
    bool in_num = false, have_header = false, too_big = false;
    constexpr int64_t kIdMax = 0x7fffffffll;
    auto end_line = [&]() {
        if (in_num) {
            if (!have_header) declared = cur;
            else if (cur > kIdMax) too_big = true;
            else row.push_back((int32_t)cur);
        }
        in_num = false;
        cur = 0;
        if (!have_header) { have_header = true; return; }
        const int64_t v = line++;
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        for (int32_t u : row)
            if (u != v) col.push_back(u);
        off.push_back((int64_t)col.size());
        row.clear();
    };
    size_t len;
    bool pending = false;
    while ((len = std::fread(buf.data(), 1, buf.size(), fp)) > 0) {
        for (size_t i = 0; i < len; ++i) {
            const char ch = buf[i];
            if (ch == '\n') { end_line(); pending = false; continue; }
            pending = true;
            if (ch >= '0' && ch <= '9') {
                // Saturate above the id range (the header may be larger; a
                // saturated neighbour id fails the load below).
                cur = cur > kIdMax ? cur : cur * 10 + (ch - '0');
                in_num = true;
            } else if (in_num) {
                if (!have_header) declared = cur;
                else if (cur > kIdMax) too_big = true;
                else row.push_back((int32_t)cur);
                in_num = false;
                cur = 0;
            }
        }
    }
    std::fclose(fp);
    if (pending) end_line();
    if (!have_header || declared < 0 || too_big) return DMOLP_E_GRAPH;
    const int64_t n = line;  // the header's V is replaced by the line count (init.cpp:96)
    if (n > kIdMax) return DMOLP_E_GRAPH;
    for (int32_t u : col)
        if (u < 0 || u >= n) return DMOLP_E_GRAPH;
    out->n = n;
    out->nnz = (int64_t)col.size();
    out->off = (int64_t *)std::malloc(sizeof(int64_t) * off.size());
    out->col = (int32_t *)std::malloc(sizeof(int32_t) * (col.empty() ? 1 : col.size()));
    if (!out->off || !out->col) {
        std::free(out->off);
        std::free(out->col);
        std::memset(out, 0, sizeof(*out));
        return DMOLP_E_NOMEM;
    }
    std::memcpy(out->off, off.data(), sizeof(int64_t) * off.size());
    if (!col.empty()) std::memcpy(out->col, col.data(), sizeof(int32_t) * col.size());
    return DMOLP_OK;
}

extern "C" int dmolp_write_adjacency_text(const char *path, int64_t n, const int64_t *off,
                                          const int32_t *col) {
    if (!path || n < 0 || !off || (off[n] > 0 && !col)) return DMOLP_E_ARG;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return DMOLP_E_ARG;
    std::fprintf(fp, "%lld\n", (long long)n);
    std::string line;
    for (int64_t v = 0; v < n; ++v) {
        line.clear();
        for (int64_t e = off[v]; e < off[v + 1]; ++e) {
            if (e > off[v]) line.push_back(' ');
            line += std::to_string(col[e]);
        }
        line.push_back('\n');
        std::fwrite(line.data(), 1, line.size(), fp);
    }
    return std::fclose(fp) == 0 ? DMOLP_OK : DMOLP_E_ARG;
}
