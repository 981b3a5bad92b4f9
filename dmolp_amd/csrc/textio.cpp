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
#include <sys/stat.h>

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

// ---------------------------------------------------------------- binary CSR cache

namespace {

constexpr char kCsrMagic[8] = {'D', 'M', 'O', 'L', 'P', 'C', 'S', 'R'};
constexpr uint32_t kCsrVersion = 1;

struct CsrHeader {
    char magic[8];
    uint32_t version;
    uint32_t offset_bytes;
    int64_t n;
    int64_t nnz;
};

bool write_all(FILE *fp, const void *p, size_t bytes) {
    return bytes == 0 || std::fwrite(p, 1, bytes, fp) == bytes;
}

bool read_all(FILE *fp, void *p, size_t bytes) {
    return bytes == 0 || std::fread(p, 1, bytes, fp) == bytes;
}

}  // namespace

extern "C" int dmolp_write_csr_binary(const char *path, int64_t n, const int64_t *off,
                                      const int32_t *col) {
    if (!path || n < 0 || !off || (off[n] > 0 && !col)) return DMOLP_E_ARG;
    // Written to a temporary name and renamed, so a reader never sees a
    // half-written cache.
    const std::string tmp = std::string(path) + ".tmp";
    FILE *fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return DMOLP_E_ARG;
    CsrHeader h{};
    std::memcpy(h.magic, kCsrMagic, sizeof(h.magic));
    h.version = kCsrVersion;
    h.offset_bytes = 8;
    h.n = n;
    h.nnz = off[n];
    bool ok = write_all(fp, &h, sizeof(h)) &&
              write_all(fp, off, sizeof(int64_t) * (size_t)(n + 1)) &&
              write_all(fp, col, sizeof(int32_t) * (size_t)h.nnz);
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return DMOLP_E_ARG;
    }
    return DMOLP_OK;
}

extern "C" int dmolp_read_csr_binary(const char *path, dmolp_csr *out) {
    if (!path || !out) return DMOLP_E_ARG;
    std::memset(out, 0, sizeof(*out));
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return DMOLP_E_ARG;
    CsrHeader h{};
    int rc = DMOLP_OK;
    long long length = -1;
    if (std::fseek(fp, 0, SEEK_END) == 0) length = std::ftell(fp);
    std::rewind(fp);
    if (!read_all(fp, &h, sizeof(h)) || std::memcmp(h.magic, kCsrMagic, sizeof(h.magic)) != 0 ||
        h.version != kCsrVersion || h.offset_bytes != 8 || h.n < 0 || h.n > 0x7fffffffll ||
        h.nnz < 0 || h.nnz > (1ll << 40) ||
        (long long)(sizeof(h) + 8 * (uint64_t)(h.n + 1) + 4 * (uint64_t)h.nnz) != length) {
        std::fclose(fp);
        return DMOLP_E_GRAPH;
    }
    out->off = (int64_t *)std::malloc(sizeof(int64_t) * (size_t)(h.n + 1));
    out->col = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(h.nnz > 0 ? h.nnz : 1));
    if (!out->off || !out->col) rc = DMOLP_E_NOMEM;
    if (rc == DMOLP_OK && (!read_all(fp, out->off, sizeof(int64_t) * (size_t)(h.n + 1)) ||
                           !read_all(fp, out->col, sizeof(int32_t) * (size_t)h.nnz)))
        rc = DMOLP_E_GRAPH;
    std::fclose(fp);
    if (rc == DMOLP_OK) {
        // Offsets must run from 0 to nnz without decreasing, columns stay in
        // [0, n) (row order and self-loops are checked on upload).
        bool good = out->off[0] == 0 && out->off[h.n] == h.nnz;
        for (int64_t v = 0; good && v < h.n; ++v) good = out->off[v] <= out->off[v + 1];
        for (int64_t e = 0; good && e < h.nnz; ++e) good = out->col[e] >= 0 && out->col[e] < h.n;
        if (!good) rc = DMOLP_E_GRAPH;
    }
    if (rc != DMOLP_OK) {
        std::free(out->off);
        std::free(out->col);
        std::memset(out, 0, sizeof(*out));
        return rc;
    }
    out->n = h.n;
    out->nnz = h.nnz;
    return DMOLP_OK;
}

extern "C" int dmolp_load_adjacency_cached(const char *path, dmolp_csr *out, int32_t *from_cache) {
    if (!path || !out) return DMOLP_E_ARG;
    if (from_cache) *from_cache = 0;
    const std::string cache = std::string(path) + ".csr";
    struct stat st_text {}, st_cache {};
    const bool have_text = ::stat(path, &st_text) == 0;
    const bool have_cache = ::stat(cache.c_str(), &st_cache) == 0;
    if (have_text && have_cache &&
        (st_cache.st_mtim.tv_sec > st_text.st_mtim.tv_sec ||
         (st_cache.st_mtim.tv_sec == st_text.st_mtim.tv_sec &&
          st_cache.st_mtim.tv_nsec >= st_text.st_mtim.tv_nsec)) &&
        dmolp_read_csr_binary(cache.c_str(), out) == DMOLP_OK) {
        if (from_cache) *from_cache = 1;
        return DMOLP_OK;
    }
    const int rc = dmolp_read_adjacency_text(path, out);
    if (rc != DMOLP_OK) return rc;
    (void)dmolp_write_csr_binary(cache.c_str(), out->n, out->off, out->col);
    return DMOLP_OK;
}
