// gen_host.cpp -- host graph generators and the CSR builder.
//
// Directed samples are symmetrised, self-loops dropped, rows sorted and
// deduplicated: the same post-processing tool/convert.sh:76-79,106-116
// applies to the reference's inputs.  The device generator (gen_dev.hip)
// draws the same samples from the same counter-based stream.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <new>
#include <thread>
#include <vector>

#include "gen_common.hpp"
#include "gen_host.hpp"
#include "internal.hpp"

namespace dmolp {

int host_threads() {
    const char *e = std::getenv("DMOLP_HOST_THREADS");
    int t = e ? std::atoi(e) : 0;
    if (t <= 0) t = (int)std::thread::hardware_concurrency();
    if (t <= 0) t = 1;
    return t > 64 ? 64 : t;
}

// Ranges of [0, n) handed out in chunks of about n / (64 T) to T threads,
// so that skewed work (the long rows of a power-law graph sit at low ids)
// spreads over all of them.
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &body, int64_t grain) {
    const int T = host_threads();
    if (n <= 0) return;
    if (T == 1 || (grain <= 0 && n < 4096)) { body(0, n); return; }
    const int64_t chunk = grain > 0 ? grain : std::max<int64_t>(1024, n / ((int64_t)T * 64));
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&] {
            for (;;) {
                const int64_t a = next.fetch_add(chunk, std::memory_order_relaxed);
                if (a >= n) return;
                body(a, std::min(n, a + chunk));
            }
        });
    for (auto &x : th) x.join();
}

namespace {

// Search guide for Chung-Lu draws: guide[b] is the vertex whose weight
// interval holds b << shift, so a draw's binary search over the whole
// cumulative-weight table shrinks to the few vertices between guide[b] and
// guide[b+1] (the same vertex either way: the interval holding the draw is
// unique).
struct ClGuide {
    int shift = 0;
    std::vector<int64_t> at;
};

ClGuide chung_lu_guide(const std::vector<uint64_t> &cdf, int64_t n) {
    ClGuide g;
    const uint64_t total = cdf[(size_t)n];
    const int bits = 22;
    int top = 63;
    while (top > 0 && !((total >> top) & 1)) --top;
    g.shift = top + 1 > bits ? top + 1 - bits : 0;
    const int64_t nb = (int64_t)((total >> g.shift) + 1);
    g.at.resize((size_t)nb);
    parallel_for(nb, [&](int64_t a, int64_t b) {
        int64_t v = std::upper_bound(cdf.begin(), cdf.begin() + n + 1, (uint64_t)a << g.shift) -
                    cdf.begin() - 1;
        v = std::min<int64_t>(v, n - 1);
        for (int64_t i = a; i < b; ++i) {
            const uint64_t x = (uint64_t)i << g.shift;
            while (v + 1 < n && cdf[(size_t)v + 1] <= x) ++v;
            g.at[(size_t)i] = v;
        }
    });
    return g;
}

// chung_lu_pick with the guide: the vertex v with cdf[v] <= target < cdf[v+1].
inline uint32_t chung_lu_pick_guided(const uint64_t *cdf, int64_t n, const ClGuide &g,
                                     uint64_t r) {
    const uint64_t target = mulhi64(r, cdf[n]);
    const int64_t b = (int64_t)(target >> g.shift);
    int64_t lo = g.at[(size_t)b];
    int64_t hi = b + 1 < (int64_t)g.at.size() ? g.at[(size_t)b + 1] + 1 : n;
    while (hi - lo > 1) {  // invariant: cdf[lo] <= target < cdf[hi]
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (cdf[mid] <= target) lo = mid; else hi = mid;
    }
    return (uint32_t)lo;
}

}  // namespace

RmatThresholds rmat_thresholds(double a, double b, double c) {
    auto q = [](double p) -> uint32_t {
        double x = std::floor(p * 4294967296.0);
        if (x < 0) x = 0;
        if (x > 4294967295.0) x = 4294967295.0;
        return (uint32_t)x;
    };
    return RmatThresholds{q(a), q(a + b), q(a + b + c)};
}

// psi(x) for x > 0 (recurrence to x >= 6, then the asymptotic series).
static double digamma(double x) {
    double r = 0.0;
    while (x < 6.0) { r -= 1.0 / x; x += 1.0; }
    const double f = 1.0 / (x * x);
    return r + std::log(x) - 0.5 / x -
           f * (1.0 / 12 - f * (1.0 / 120 - f * (1.0 / 252 - f * (1.0 / 240 - f / 132))));
}

// Chung-Lu integer weight table.  Expected degrees follow
// w_i = dmax * (i0 / (i + i0))^(1/(gamma-1)); i0 is fitted so that
// sum w_i = 2m (gamma = 2 uses the digamma closed form of the harmonic
// sum).  Integer weights W_i = floor(1024 w_i) make host and device draws
// identical.
std::vector<uint64_t> chung_lu_cdf(int64_t n, int64_t m, double gamma, double dmax) {
    DMOLP_REQUIRE(n > 0 && m > 0 && gamma > 1.0 && dmax >= 1.0, DMOLP_E_ARG, "bad Chung-Lu spec");
    const double target = 2.0 * (double)m;
    const double e = 1.0 / (gamma - 1.0);
    auto total = [&](double i0) -> double {
        if (gamma == 2.0) return dmax * i0 * (digamma((double)n + i0) - digamma(i0));
        double s = 0.0;  // general gamma: direct sum (fitting only)
        for (int64_t i = 0; i < n; ++i) s += dmax * std::pow(i0 / ((double)i + i0), e);
        return s;
    };
    double lo = 1e-9, hi = 1.0;
    while (total(hi) < target && hi < 1e18) hi *= 2.0;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (total(mid) < target) lo = mid; else hi = mid;
    }
    const double i0 = 0.5 * (lo + hi);
    std::vector<uint64_t> cdf((size_t)n + 1);
    cdf[0] = 0;
    std::vector<uint64_t> w((size_t)n);
    parallel_for(n, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i) {
            const double x = i0 / ((double)i + i0);
            const double wi = gamma == 2.0 ? dmax * x : dmax * std::pow(x, e);
            w[(size_t)i] = (uint64_t)std::floor(wi * 1024.0);
        }
    });
    for (int64_t i = 0; i < n; ++i) cdf[(size_t)i + 1] = cdf[(size_t)i] + w[(size_t)i];
    return cdf;
}

// Symmetrise, drop self-loops, sort and dedupe rows.  The samples are cut
// into C slices; slice c counts its entries per row in a private array and
// then writes them from its own offset inside each row, so no counter is
// shared between threads (the long rows of a power-law graph would
// serialise shared ones).  Rows are then sorted in ranges of about equal
// entry counts.
void build_csr(int64_t n, int64_t m, const uint32_t *src, const uint32_t *dst, dmolp_csr *out) {
    int64_t C = std::min<int64_t>(host_threads(), 16);
    while (C > 1 && C * n * (int64_t)sizeof(int32_t) > (3ll << 30)) --C;
    const int64_t per = (m + C - 1) / C;
    std::vector<std::vector<int32_t>> cnt((size_t)C);
    auto slices = [&](const std::function<void(int64_t, int64_t, int64_t)> &body) {
        std::vector<std::thread> th;
        for (int64_t c = 0; c < C; ++c)
            th.emplace_back([&, c] { body(c, std::min(m, c * per), std::min(m, (c + 1) * per)); });
        for (auto &x : th) x.join();
    };
    slices([&](int64_t c, int64_t a, int64_t b) {
        std::vector<int32_t> &k = cnt[(size_t)c];
        k.assign((size_t)n, 0);
        for (int64_t e = a; e < b; ++e) {
            if (src[e] == dst[e]) continue;
            ++k[src[e]];
            ++k[dst[e]];
        }
    });
    // Row lengths, then each slice's start inside each row (cnt becomes the
    // slice's running write offset relative to the row start).
    std::vector<int64_t> off((size_t)n + 1, 0);
    parallel_for(n, [&](int64_t a, int64_t b) {
        for (int64_t v = a; v < b; ++v) {
            int64_t run = 0;
            for (int64_t c = 0; c < C; ++c) {
                const int32_t x = cnt[(size_t)c][(size_t)v];
                cnt[(size_t)c][(size_t)v] = (int32_t)run;
                run += x;
            }
            off[(size_t)v + 1] = run;
        }
    });
    for (int64_t v = 0; v < n; ++v) off[(size_t)v + 1] += off[(size_t)v];
    std::vector<int32_t> col((size_t)off[(size_t)n]);
    slices([&](int64_t c, int64_t a, int64_t b) {
        std::vector<int32_t> &k = cnt[(size_t)c];
        for (int64_t e = a; e < b; ++e) {
            const uint32_t u = src[e], v = dst[e];
            if (u == v) continue;
            col[(size_t)(off[u] + k[u]++)] = (int32_t)v;
            col[(size_t)(off[v] + k[v]++)] = (int32_t)u;
        }
    });
    cnt.clear();
    cnt.shrink_to_fit();
    // Sort and dedupe rows, in ranges of about nnz / (64 T) entries.
    std::vector<int64_t> cuts{0};
    {
        const int64_t target = std::max<int64_t>(4096, off[(size_t)n] / (64ll * host_threads()));
        int64_t last = 0;
        for (int64_t v = 1; v <= n; ++v)
            if (off[(size_t)v] - off[(size_t)last] >= target || v == n) { cuts.push_back(v); last = v; }
    }
    std::vector<int64_t> ndeg((size_t)n, 0);
    parallel_for((int64_t)cuts.size() - 1, [&](int64_t a, int64_t b) {
        for (int64_t r = a; r < b; ++r)
            for (int64_t v = cuts[(size_t)r]; v < cuts[(size_t)r + 1]; ++v) {
                int32_t *r0 = col.data() + off[(size_t)v], *r1 = col.data() + off[(size_t)v + 1];
                std::sort(r0, r1);
                ndeg[(size_t)v] = std::unique(r0, r1) - r0;
            }
    }, 1);
    out->n = n;
    out->off = (int64_t *)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
    DMOLP_REQUIRE(out->off, DMOLP_E_NOMEM, "host allocation failed");
    out->off[0] = 0;
    for (int64_t v = 0; v < n; ++v) out->off[v + 1] = out->off[v] + ndeg[(size_t)v];
    out->nnz = out->off[n];
    out->col = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(out->nnz > 0 ? out->nnz : 1));
    DMOLP_REQUIRE(out->col, DMOLP_E_NOMEM, "host allocation failed");
    parallel_for(n, [&](int64_t a, int64_t b) {
        for (int64_t v = a; v < b; ++v)
            std::memcpy(out->col + out->off[v], col.data() + off[(size_t)v],
                        sizeof(int32_t) * (size_t)ndeg[(size_t)v]);
    });
}

void gen_host(const dmolp_gen_spec *spec, dmolp_csr *out) {
    DMOLP_REQUIRE(spec && out, DMOLP_E_ARG, "null argument");
    std::memset(out, 0, sizeof(*out));
    const uint64_t key = stream_key(spec->seed);
    if (spec->kind == DMOLP_GEN_RMAT) {
        DMOLP_REQUIRE(spec->scale >= 1 && spec->scale <= 30 && spec->edge_factor >= 1, DMOLP_E_ARG,
                      "bad R-MAT spec");
        const int64_t n = 1ll << spec->scale;
        const int64_t m = (int64_t)spec->edge_factor << spec->scale;
        const RmatThresholds th = rmat_thresholds(spec->a, spec->b, spec->c);
        std::vector<uint32_t> src((size_t)m), dst((size_t)m);
        parallel_for(m, [&](int64_t a, int64_t b) {
            for (int64_t e = a; e < b; ++e)
                rmat_edge(key, spec->scale, th, (uint64_t)e, &src[(size_t)e], &dst[(size_t)e]);
        });
        build_csr(n, m, src.data(), dst.data(), out);
    } else if (spec->kind == DMOLP_GEN_CHUNG_LU) {
        const int64_t n = spec->n, m = spec->m;
        DMOLP_REQUIRE(n >= 2 && n < (1ll << 31) && m >= 1, DMOLP_E_ARG, "bad Chung-Lu size");
        const std::vector<uint64_t> cdf = chung_lu_cdf(n, m, spec->gamma, spec->dmax);
        const ClGuide guide = chung_lu_guide(cdf, n);
        std::vector<uint32_t> src((size_t)m), dst((size_t)m);
        parallel_for(m, [&](int64_t a, int64_t b) {
            for (int64_t e = a; e < b; ++e) {
                src[(size_t)e] = chung_lu_pick_guided(cdf.data(), n, guide,
                                                      rand_at(key, 2 * (uint64_t)e));
                dst[(size_t)e] = chung_lu_pick_guided(cdf.data(), n, guide,
                                                      rand_at(key, 2 * (uint64_t)e + 1));
            }
        });
        build_csr(n, m, src.data(), dst.data(), out);
    } else if (spec->kind == DMOLP_GEN_GRID) {
        const int64_t R = spec->rows, C = spec->cols;
        DMOLP_REQUIRE(R >= 1 && C >= 1 && R * C < (1ll << 31), DMOLP_E_ARG, "bad grid spec");
        const int64_t n = R * C;
        out->n = n;
        out->off = (int64_t *)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
        DMOLP_REQUIRE(out->off, DMOLP_E_NOMEM, "host allocation failed");
        out->off[0] = 0;
        for (int64_t v = 0; v < n; ++v) {
            const int64_t r = v / C, c = v % C;
            out->off[v + 1] = out->off[v] + (r > 0) + (c > 0) + (c + 1 < C) + (r + 1 < R);
        }
        out->nnz = out->off[n];
        out->col = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(out->nnz > 0 ? out->nnz : 1));
        DMOLP_REQUIRE(out->col, DMOLP_E_NOMEM, "host allocation failed");
        parallel_for(n, [&](int64_t a, int64_t b) {
            for (int64_t v = a; v < b; ++v) {
                const int64_t r = v / C, c = v % C;
                int32_t *p = out->col + out->off[v];
                if (r > 0) *p++ = (int32_t)(v - C);
                if (c > 0) *p++ = (int32_t)(v - 1);
                if (c + 1 < C) *p++ = (int32_t)(v + 1);
                if (r + 1 < R) *p++ = (int32_t)(v + C);
            }
        });
    } else {
        throw Error(DMOLP_E_ARG, "unknown generator kind");
    }
}

}  // namespace dmolp
