// ref_cpu.cpp -- CPU oracle for the DMOLP partitioner path at one rank.
//
// TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it.  It is built with -O3 -fno-fast-math -ffp-contract=off so every double
// operation is a correctly rounded IEEE operation in source order.
//
// OpenMP: the integer sweeps (statistics, cut, candidate marking, boundary
// test, LP decisions, expansion claims, depth-2 balls) run on
// ref_set_threads() threads (default: OMP_NUM_THREADS).  Every parallel
// step either sums integers or produces a set that is sorted or consumed as
// a set afterwards, and the claim step keeps the lowest claiming part with
// an atomic minimum, so results do not depend on the thread count.  The
// floating-point steps (hub tangent, penalties, selection ratio, balances)
// stay sequential in source order.
//
// Each function names the reference lines it restates.  The restatement is
// written for one rank: owner(v) = 0, no ghosts, every MPI collective is the
// identity (partition.cpp:55-60, utils.cpp:159-166).
//
// Pinning (what checks each function beyond this file):
//   ref_probe_degree_map_order : itself a probe of libstdc++; tests assert
//                                it yields descending ids.
//   ref_first_seed             : closed form (min id < D among max-degree
//                                vertices, else max id) on crafted graphs.
//   ref_hubs                   : constant-degree graph (all hubs), stars,
//                                the |H| ~ N*avg/max bound, NaN total.
//   ref_ball2 / ref_seeds      : networkx shortest-path distance <= 2,
//                                brute-force seed rule in Python.
//   ref_expand                 : hand traces (path, star, 3x3 grid) for
//                                theta in {1,2,inf}; theta=inf equals
//                                lowest-p multi-source BFS (networkx);
//                                invariants.
//   ref_stats / metrics        : numpy recount, cut + sum ic = nnz/2.
//   ref_penalties              : special cases t=1, balanced, two-part
//                                closed forms computed by hand.
//   ref_select_boundary        : Python brute force with full sorts.
//   ref_lp_round               : simulation of the 32-lane shuffle tree of
//                                gpu_lp_boundary.cu:393-421 in Python.
#include "ref_cpu.h"

#include <omp.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

inline int64_t deg_of(const int64_t *off, int64_t v) { return off[v + 1] - off[v]; }

// Labels invisible to the LP: gpu_lp_boundary.cu:360 caps at 32.
constexpr int kLpLabelCap = 32;
// Consecutive small-change rounds before stopping: phase2.cpp:265.
constexpr int kConvergeRounds = 10;

// 5-bit bit reversal: the lane priority the shuffle tree of
// gpu_lp_boundary.cu:406-414 produces among equal scores.
inline int bitrev5(int x) {
    int r = 0;
    for (int i = 0; i < 5; ++i) r |= ((x >> i) & 1) << (4 - i);
    return r;
}

// Breadth-first ball of depth <= 2 (seed.cpp:18 MAX_LEVELS = 3 keeps levels
// 0, 1, 2; seed.cpp:33-98), OR-ed into `mark`.  The ball of s is computed on
// its own (compute_bfs starts from empty bitsets, seed.cpp:21-24), so level
// 1 is every neighbour of s whatever `mark` already holds.
// Concurrent writers only ever store 1, so the marks are relaxed atomic
// byte stores.
inline void set_mark(uint8_t *m) { __atomic_store_n(m, (uint8_t)1, __ATOMIC_RELAXED); }

void ball2(int64_t n, const int64_t *off, const int32_t *col, int32_t s,
           std::vector<uint8_t> &mark) {
    (void)n;
    uint8_t *mk = mark.data();
    mk[s] = 1;
    const int64_t e0 = off[s], e1 = off[s + 1];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t e = e0; e < e1; ++e) {
        const int32_t u = col[e];
        set_mark(mk + u);
        for (int64_t f = off[u]; f < off[u + 1]; ++f) set_mark(mk + col[f]);
    }
}

// Ids v in [0, n) with mark[v] set, ascending; the marks are cleared.
void collect_marked(int64_t n, std::vector<uint8_t> &mark, std::vector<int32_t> &out) {
    const int T = omp_get_max_threads();
    std::vector<int64_t> cnt((size_t)T + 1, 0);
    uint8_t *mk = mark.data();
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
        const int64_t a = n * t / nt, b = n * (t + 1) / nt;
        int64_t c = 0;
        for (int64_t v = a; v < b; ++v) c += mk[v] != 0;
        cnt[(size_t)t + 1] = c;
#pragma omp barrier
#pragma omp single
        for (int i = 0; i < nt; ++i) cnt[(size_t)i + 1] += cnt[(size_t)i];
        int64_t w = cnt[(size_t)t];
        if (t == 0) out.resize((size_t)cnt[(size_t)nt]);
#pragma omp barrier
        for (int64_t v = a; v < b; ++v)
            if (mk[v]) { out[(size_t)w++] = (int32_t)v; mk[v] = 0; }
    }
}

}  // namespace

extern "C" {

void ref_set_threads(int32_t t) { omp_set_num_threads(t > 0 ? t : 1); }
int32_t ref_get_threads(void) { return omp_get_max_threads(); }

// init.cpp:43-44,68,90-93 (adj and local_degree built with reserve(2V) in
// ascending source order) and init.cpp:123-135 (global_degree with
// reserve(V) filled in local_degree's iteration order), single rank.
int ref_probe_degree_map_order(int32_t n, int32_t *order_out) {
    if (n < 0 || (n > 0 && !order_out)) return REF_E_ARG;
    std::unordered_map<int, std::vector<int>> adj;
    std::unordered_map<int, int> local_degree;
    adj.reserve((size_t)n * 2);
    local_degree.reserve((size_t)n * 2);
    for (int s = 0; s < n; ++s) (void)adj[s];
    for (auto &kv : adj) local_degree[kv.first] = (int)kv.second.size();
    std::vector<int> flat;
    for (auto &kv : local_degree) { flat.push_back(kv.first); flat.push_back(kv.second); }
    std::unordered_map<int, int> global_degree;
    global_degree.reserve(flat.size() / 2);
    for (size_t i = 0; i < flat.size(); i += 2) global_degree[flat[i]] = flat[i + 1];
    int32_t i = 0;
    for (auto &kv : global_degree) order_out[i++] = kv.first;
    return REF_OK;
}

// init.cpp:179-187.  The degree map iterates in descending id at one rank;
// the tie test compares a vertex id with the running best *degree*.
int ref_first_seed(int64_t n, const int64_t *off, int32_t *seed_out, int32_t *deg_out) {
    if (n <= 0 || !off) return REF_E_ARG;
    int64_t best_node = -1, best_deg = -1;
    for (int64_t v = n - 1; v >= 0; --v) {
        int64_t d = deg_of(off, v);
        if (d > best_deg || (d == best_deg && v < best_deg)) { best_node = v; best_deg = d; }
    }
    if (seed_out) *seed_out = (int32_t)best_node;
    if (deg_out) *deg_out = (int32_t)best_deg;
    return REF_OK;
}

// init.cpp:139-177.  The (id, degree) pairs enter std::sort in the degree
// map's iteration order (descending id); the sort is libstdc++ introsort,
// not stable, so the same call on the same sequence is required.
int ref_hubs(int64_t n, const int64_t *off, int32_t *hubs_out, int64_t *n_hubs_out) {
    if (n <= 0 || !off || !n_hubs_out) return REF_E_ARG;
    std::vector<std::pair<int, int>> sd;
    sd.reserve((size_t)n);
    for (int64_t v = n - 1; v >= 0; --v) sd.emplace_back((int)v, (int)deg_of(off, v));
    std::sort(sd.begin(), sd.end(),
              [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.second < b.second; });
    const int N = (int)sd.size();
    int64_t nh = 0;
    if (N >= 2) {
        std::vector<double> cum((size_t)N);
        double total = 0.0, run = 0.0;
        for (int i = 0; i < N; ++i) {
            int d = sd[(size_t)i].second;
            total += d;
            run += d;
            cum[(size_t)i] = run;
        }
        for (int i = 0; i < N; ++i) cum[(size_t)i] /= total;
        double inv_N = 1.0 / N;
        double x1 = N * inv_N;
        double y1 = cum[(size_t)N - 1];
        double x0 = (N - 1) * inv_N;
        double y0 = cum[(size_t)N - 2];
        double slope = (y1 - y0) / (x1 - x0);
        double xi = x1 - y1 / slope;
        for (int i = N - 1; i >= 0; --i) {
            double x = (i + 1) * inv_N;
            if (x > xi) {
                if (hubs_out) hubs_out[nh] = sd[(size_t)i].first;
                ++nh;
            } else {
                break;
            }
        }
    }
    *n_hubs_out = nh;
    return REF_OK;
}

int ref_ball2(int64_t n, const int64_t *off, const int32_t *col, int32_t s, uint8_t *ball_out) {
    if (n <= 0 || s < 0 || s >= n || !ball_out) return REF_E_ARG;
    std::vector<uint8_t> mark((size_t)n, 0);
    ball2(n, off, col, s, mark);
    std::memcpy(ball_out, mark.data(), (size_t)n);
    return REF_OK;
}

// seed.cpp:138-242 at one rank.  Next seed = first hub in hub order that is
// unused, outside the union of the chosen seeds' depth-2 balls, with the
// strictly largest degree (seed.cpp:158-170).  The search stops at the first
// failure (seed.cpp:214-217).  Documented extension for the reference's
// undefined behaviour when fewer than k seeds exist (partition.cpp:347 reads
// seeds[p] past the end): the missing seeds are the remaining vertices in
// (degree descending, id ascending) order, skipping chosen seeds.
int ref_seeds(int64_t n, const int64_t *off, const int32_t *col, int32_t k,
              int32_t *seeds_out, int32_t *n_found_out, int32_t *n_fallback_out) {
    if (n <= 0 || k <= 0 || !seeds_out) return REF_E_ARG;
    if ((int64_t)k > n) return REF_E_SEEDS;
    int32_t s0, d0;
    ref_first_seed(n, off, &s0, &d0);
    std::vector<int32_t> hubs((size_t)n);
    int64_t nh = 0;
    ref_hubs(n, off, hubs.data(), &nh);
    hubs.resize((size_t)nh);

    std::vector<int32_t> seeds{s0};
    if (nh > 0) {
        std::vector<uint8_t> used((size_t)nh, 0);
        std::vector<uint8_t> covered((size_t)n, 0);
        auto mark_used = [&](int32_t s) {
            for (int64_t i = 0; i < nh; ++i)
                if (hubs[(size_t)i] == s) { used[(size_t)i] = 1; break; }
        };
        mark_used(s0);
        ball2(n, off, col, s0, covered);
        for (int j = 1; j < k; ++j) {
            int64_t best_deg = -1;
            int32_t best = -1;
            for (int64_t i = 0; i < nh; ++i) {
                if (used[(size_t)i]) continue;
                int32_t h = hubs[(size_t)i];
                if (covered[(size_t)h]) continue;
                int64_t d = deg_of(off, h);
                if (d > best_deg) { best_deg = d; best = h; }
            }
            if (best == -1) break;
            seeds.push_back(best);
            mark_used(best);
            ball2(n, off, col, best, covered);
        }
    }
    int32_t found = (int32_t)seeds.size();
    if (found < k) {
        std::vector<uint8_t> chosen((size_t)n, 0);
        for (int32_t s : seeds) chosen[(size_t)s] = 1;
        std::vector<int32_t> order((size_t)n);
        for (int64_t v = 0; v < n; ++v) order[(size_t)v] = (int32_t)v;
        std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            int64_t da = deg_of(off, a), db = deg_of(off, b);
            if (da != db) return da > db;
            return a < b;
        });
        for (size_t i = 0; i < order.size() && (int32_t)seeds.size() < k; ++i)
            if (!chosen[(size_t)order[i]]) seeds.push_back(order[i]);
    }
    for (int32_t p = 0; p < k; ++p) seeds_out[p] = seeds[(size_t)p];
    if (n_found_out) *n_found_out = found < k ? found : k;
    if (n_fallback_out) *n_fallback_out = found < k ? k - found : 0;
    return REF_OK;
}

// partition.cpp:323-604 at one rank.  Per round: the fill step when every
// frontier is empty (partition.cpp:376-394); per partition the min(theta,|F|)
// frontier members smallest in (degree, id) are expanded
// (partition.cpp:400-431); an unlabelled neighbour is claimed by the lowest
// partition that reaches it, because partitions are scanned in ascending
// order against one shared visited set (partition.cpp:451-476) and at one
// rank each vertex then has a single claimant (partition.cpp:530-538);
// frontiers become carry + newly labelled (partition.cpp:588-597); sizes are
// recounted (partition.cpp:586).  A frontier is used only as a set (the
// selection takes its theta smallest keys, which are unique), so the order
// in which claims join it does not matter.
int ref_expand(int64_t n, const int64_t *off, const int32_t *col, int32_t k, int32_t theta,
               const int32_t *seeds, int64_t *rounds_out, int32_t *fill_round_out,
               int32_t *labels_out) {
    if (n <= 0 || k <= 0 || theta <= 0 || !seeds || !labels_out) return REF_E_ARG;
    int32_t *L = labels_out;
    for (int64_t v = 0; v < n; ++v) L[v] = -1;
    std::vector<std::vector<uint64_t>> F((size_t)k);  // keys (deg << 32) | id
    auto key = [&](int32_t v) { return ((uint64_t)deg_of(off, v) << 32) | (uint32_t)v; };
    for (int32_t p = 0; p < k; ++p) {
        int32_t s = seeds[p];
        if (s < 0 || s >= n) return REF_E_ARG;
        L[s] = p;
        F[(size_t)p].push_back(key(s));
    }
    std::vector<int64_t> size((size_t)k, 0);
    int64_t labelled = 0;
    auto recount = [&]() {
        std::fill(size.begin(), size.end(), 0);
        labelled = 0;
#pragma omp parallel
        {
            std::vector<int64_t> mine((size_t)k, 0);
            int64_t lab = 0;
#pragma omp for schedule(static) nowait
            for (int64_t v = 0; v < n; ++v)
                if (L[v] >= 0) { ++mine[(size_t)L[v]]; ++lab; }
#pragma omp critical
            {
                for (int32_t p = 0; p < k; ++p) size[(size_t)p] += mine[(size_t)p];
                labelled += lab;
            }
        }
    };
    recount();
    // claim[v]: lowest part whose selected vertices reach v this round
    // (INT32_MAX: none).
    std::vector<int32_t> claim((size_t)n, INT32_MAX);
    int64_t round = 0;
    int32_t fill_round = -1;
    std::vector<std::vector<uint64_t>> sel((size_t)k);
    while (labelled < n) {
        bool all_empty = true;
        for (int32_t p = 0; p < k; ++p)
            if (!F[(size_t)p].empty()) { all_empty = false; break; }
        if (all_empty) {
            int32_t b = 0;
            for (int32_t p = 0; p < k; ++p)
                if (size[(size_t)p] < size[(size_t)b] || (size[(size_t)p] == size[(size_t)b] && p < b)) b = p;
            for (int64_t v = 0; v < n; ++v)
                if (L[v] < 0) { L[v] = b; F[(size_t)b].push_back(key((int32_t)v)); }
            if (fill_round < 0) fill_round = (int32_t)round;
        }
        // Selection: the theta smallest keys of each frontier.
#pragma omp parallel for schedule(dynamic, 1)
        for (int32_t p = 0; p < k; ++p) {
            auto &f = F[(size_t)p];
            sel[(size_t)p].clear();
            if ((int64_t)f.size() > theta) {
                std::nth_element(f.begin(), f.begin() + theta, f.end());
                sel[(size_t)p].assign(f.begin(), f.begin() + theta);
                f.erase(f.begin(), f.begin() + theta);
            } else {
                sel[(size_t)p].swap(f);
                f.clear();
            }
        }
        // Claims: the lowest part reaching v wins; the label test reads labels
        // as they were before this round's claims.
        std::vector<std::pair<int32_t, int64_t>> work;  // (part, index into sel[part])
        for (int32_t p = 0; p < k; ++p)
            for (size_t i = 0; i < sel[(size_t)p].size(); ++i) work.emplace_back(p, (int64_t)i);
        std::vector<int32_t> touched;
#pragma omp parallel
        {
            std::vector<int32_t> mine;
#pragma omp for schedule(dynamic, 16) nowait
            for (int64_t w = 0; w < (int64_t)work.size(); ++w) {
                const int32_t p = work[(size_t)w].first;
                const int32_t u = (int32_t)(sel[(size_t)p][(size_t)work[(size_t)w].second] & 0xffffffffu);
                for (int64_t e = off[u]; e < off[u + 1]; ++e) {
                    const int32_t v = col[e];
                    if (L[v] != -1) continue;
                    int32_t cur = __atomic_load_n(&claim[(size_t)v], __ATOMIC_RELAXED);
                    while (p < cur) {
                        if (__atomic_compare_exchange_n(&claim[(size_t)v], &cur, p, false,
                                                        __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                            if (cur == INT32_MAX) mine.push_back(v);
                            break;
                        }
                    }
                }
            }
#pragma omp critical
            touched.insert(touched.end(), mine.begin(), mine.end());
        }
        for (int32_t v : touched) {
            const int32_t p = claim[(size_t)v];
            L[v] = p;
            F[(size_t)p].push_back(key(v));
            claim[(size_t)v] = INT32_MAX;
        }
        recount();
        ++round;
    }
    if (rounds_out) *rounds_out = round;
    if (fill_round_out) *fill_round_out = fill_round;
    return REF_OK;
}

// phase1.cpp:268-339 at one rank.
int ref_phase1_metrics(int64_t n, const int64_t *off, const int32_t *col, int32_t k,
                       const int32_t *labels, int64_t *vc, int64_t *ec, ref_p1_metrics *m) {
    if (n <= 0 || k <= 0 || !labels || !m) return REF_E_ARG;
    std::vector<int64_t> cv((size_t)k, 0), ce((size_t)k, 0);
    int64_t cut2 = 0;
#pragma omp parallel
    {
        std::vector<int64_t> mv((size_t)k, 0), me((size_t)k, 0);
        int64_t mc = 0;
#pragma omp for schedule(dynamic, 1024) nowait
        for (int64_t u = 0; u < n; ++u) {
            int32_t lu = labels[u];
            if (lu >= 0) ++mv[(size_t)lu];
            for (int64_t e = off[u]; e < off[u + 1]; ++e) {
                int32_t lv = labels[col[e]];
                if (lu != -1 && lv != -1 && lu != lv) ++mc;
                if (lu >= 0 && lv == lu) ++me[(size_t)lu];
            }
        }
#pragma omp critical
        {
            for (int32_t p = 0; p < k; ++p) { cv[(size_t)p] += mv[(size_t)p]; ce[(size_t)p] += me[(size_t)p]; }
            cut2 += mc;
        }
    }
    for (int32_t p = 0; p < k; ++p) ce[(size_t)p] /= 2;
    m->initial_edge_cut = cut2 / 2;
    double expected_vertices = static_cast<double>(n) / k;
    double max_vr = 0;
    for (int32_t p = 0; p < k; ++p) max_vr = std::max(max_vr, (double)cv[(size_t)p] / expected_vertices);
    m->initial_vertex_balance = max_vr;
    double total_pe = 0;
    for (int32_t p = 0; p < k; ++p) total_pe += (double)ce[(size_t)p];
    double expected_edges = total_pe / k;
    double max_er = 0;
    for (int32_t p = 0; p < k; ++p) max_er = std::max(max_er, (double)ce[(size_t)p] / expected_edges);
    m->initial_edge_balance = max_er;
    m->total_vertices = n;
    m->total_edges = off[n] / 2;
    if (vc) std::copy(cv.begin(), cv.end(), vc);
    if (ec) std::copy(ce.begin(), ce.end(), ec);
    return REF_OK;
}

// phase1.cpp:18-161 (load excluded: the graph arrives as CSR).
int ref_phase1(int64_t n, const int64_t *off, const int32_t *col, int32_t k, int32_t theta,
               int32_t *labels_out, int32_t *seeds_out, int64_t *vc, int64_t *ec,
               ref_p1_metrics *m) {
    if (n <= 0 || k <= 0 || theta <= 0 || !labels_out || !m) return REF_E_ARG;
    std::memset(m, 0, sizeof(*m));
    std::vector<int32_t> seeds((size_t)k);
    int32_t found = 0, fb = 0;
    int rc = ref_seeds(n, off, col, k, seeds.data(), &found, &fb);
    if (rc != REF_OK) return rc;
    int64_t rounds = 0;
    int32_t fill_round = -1;
    rc = ref_expand(n, off, col, k, theta, seeds.data(), &rounds, &fill_round, labels_out);
    if (rc != REF_OK) return rc;
    rc = ref_phase1_metrics(n, off, col, k, labels_out, vc, ec, m);
    if (rc != REF_OK) return rc;
    int32_t s0, d0;
    ref_first_seed(n, off, &s0, &d0);
    int64_t nh = 0;
    ref_hubs(n, off, nullptr, &nh);
    m->first_seed = s0;
    m->first_seed_degree = d0;
    m->n_hubs = nh;
    m->n_seeds_found = found;
    m->n_seeds_fallback = fb;
    m->rounds = rounds;
    m->fill_round = fill_round;
    if (seeds_out) std::copy(seeds.begin(), seeds.end(), seeds_out);
    return REF_OK;
}

// phase2.cpp:55-117 (vertex and intra-partition counts, intra entries / 2
// per partition) and phase2.cpp:214-253 (cut = differing entries / 2; abort
// on a -1 label or an odd count).
int ref_stats(int64_t n, const int64_t *off, const int32_t *col, int32_t k,
              const int32_t *labels, int64_t *vc, int64_t *ic, int64_t *cut_out) {
    if (n <= 0 || k <= 0 || !labels) return REF_E_ARG;
    std::vector<int64_t> cv((size_t)k, 0), ce((size_t)k, 0);
    int64_t cut2 = 0, invalid = 0;
#pragma omp parallel
    {
        std::vector<int64_t> mv((size_t)k, 0), me((size_t)k, 0);
        int64_t mc = 0, mi = 0;
#pragma omp for schedule(dynamic, 1024) nowait
        for (int64_t u = 0; u < n; ++u) {
            int32_t lu = labels[u];
            bool lu_ok = lu >= 0 && lu < k;
            if (lu_ok) ++mv[(size_t)lu];
            for (int64_t e = off[u]; e < off[u + 1]; ++e) {
                int32_t lv = labels[col[e]];
                if (lu_ok && lv == lu) ++me[(size_t)lu];
                if (lu == -1 || lv == -1) { ++mi; continue; }
                if (lu != lv) ++mc;
            }
        }
#pragma omp critical
        {
            for (int32_t p = 0; p < k; ++p) { cv[(size_t)p] += mv[(size_t)p]; ce[(size_t)p] += me[(size_t)p]; }
            cut2 += mc;
            invalid += mi;
        }
    }
    if (invalid) return REF_E_INVALID_LABEL;
    if (cut2 & 1) return REF_E_ODD_CUT;
    for (int32_t p = 0; p < k; ++p) ce[(size_t)p] /= 2;
    if (vc) std::copy(cv.begin(), cv.end(), vc);
    if (ic) std::copy(ce.begin(), ce.end(), ic);
    if (cut_out) *cut_out = cut2 >> 1;
    return REF_OK;
}

// phase2.cpp:158-209 and phase2.cpp:112-115 (expected counts).  Source
// order, no contraction; glibc log.
void ref_penalties(int32_t k, const int64_t *vc, const int64_t *ic, int32_t t,
                   double *pen, double *re) {
    int64_t tv = 0, te = 0;
    for (int32_t i = 0; i < k; ++i) { tv += vc[i]; te += ic[i]; }
    double exp_v = static_cast<double>(tv) / k;
    double exp_e = (te > 0) ? static_cast<double>(te) / k : 1.0;
    std::vector<double> RV((size_t)k), RE((size_t)k);
    for (int32_t i = 0; i < k; ++i) {
        RV[(size_t)i] = (exp_v > 0) ? static_cast<double>(vc[i]) / exp_v : 1.0;
        RE[(size_t)i] = (exp_e > 0) ? static_cast<double>(ic[i]) / exp_e : 1.0;
    }
    double rv_mean = 0.0, re_mean = 0.0;
    for (int32_t i = 0; i < k; ++i) { rv_mean += RV[(size_t)i]; re_mean += RE[(size_t)i]; }
    rv_mean /= k;
    re_mean /= k;
    double rv_var = 0.0, re_var = 0.0;
    for (int32_t i = 0; i < k; ++i) {
        rv_var += (RV[(size_t)i] - rv_mean) * (RV[(size_t)i] - rv_mean);
        re_var += (RE[(size_t)i] - re_mean) * (RE[(size_t)i] - re_mean);
    }
    rv_var /= k;
    re_var /= k;
    double total_var = rv_var + re_var;
    double imb_rv = (total_var > 0) ? rv_var / total_var : 0.0;
    double imb_re = (total_var > 0) ? re_var / total_var : 0.0;
    for (int32_t i = 0; i < k; ++i) {
        double g_rv = (1.0 - RV[(size_t)i]) * (1.0 + log((double)t));
        double g_re = (1.0 - RE[(size_t)i]) * (1.0 + log((double)t));
        if (pen) pen[i] = imb_rv * g_rv + imb_re * g_re;
        if (re) re[i] = RE[(size_t)i];
    }
}

namespace {

struct SelectCounts {
    int64_t prev_deg = 0, cand = 0, scan = 0, bd = 0;
};

// gpu_lp_boundary.cu:185-305.
void select_boundary(int64_t n, const int64_t *off, const int32_t *col, const int32_t *L,
                     const std::vector<int32_t> &prev, const double *re, int32_t k, int32_t t,
                     std::vector<int32_t> &out, std::vector<uint8_t> &mark, SelectCounts &sc) {
    // Candidates: previous boundary and its neighbours, unique and sorted
    // (gpu_lp_boundary.cu:207-240): marked, then collected in id order.
    std::vector<int32_t> cand;
    uint8_t *mk = mark.data();
    int64_t prev_deg = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : prev_deg)
    for (int64_t i = 0; i < (int64_t)prev.size(); ++i) {
        const int32_t v = prev[(size_t)i];
        if (v < 0 || v >= n) continue;
        prev_deg += deg_of(off, v);
        set_mark(mk + v);
        for (int64_t e = off[v]; e < off[v + 1]; ++e) set_mark(mk + col[e]);
    }
    sc.prev_deg += prev_deg;
    collect_marked(n, mark, cand);
    sc.cand = (int64_t)cand.size();
    // Boundary test with early exit (gpu_lp_boundary.cu:251-264), grouped by
    // label (gpu_lp_boundary.cu:266-269).  Each group is sorted by a total
    // order below, so the order in which threads fill it does not matter.
    std::vector<std::vector<std::pair<int32_t, int64_t>>> groups((size_t)k);
    int64_t scan = 0, nbd = 0;
#pragma omp parallel reduction(+ : scan, nbd)
    {
        std::vector<std::vector<std::pair<int32_t, int64_t>>> mine((size_t)k);
#pragma omp for schedule(dynamic, 1024) nowait
        for (int64_t i = 0; i < (int64_t)cand.size(); ++i) {
            const int32_t x = cand[(size_t)i];
            int32_t lx = L[x];
            bool bd = false;
            int64_t scanned = 0;
            for (int64_t e = off[x]; e < off[x + 1]; ++e) {
                ++scanned;
                if (L[col[e]] != lx) { bd = true; break; }
            }
            scan += scanned;
            if (bd) {
                mine[(size_t)lx].emplace_back(x, deg_of(off, x));
                ++nbd;
            }
        }
#pragma omp critical
        for (int32_t p = 0; p < k; ++p)
            groups[(size_t)p].insert(groups[(size_t)p].end(), mine[(size_t)p].begin(), mine[(size_t)p].end());
    }
    sc.scan += scan;
    sc.bd += nbd;
    // Per label, the first ceil(ratio * size) by (degree desc, id) when
    // RE > 1.0, else (degree asc, id) (gpu_lp_boundary.cu:244,277-302).
    double ratio = exp((-t) / (5.0));
    std::vector<int64_t> take_of((size_t)k, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t p = 0; p < k; ++p) {
        auto &g = groups[(size_t)p];
        if (g.empty()) continue;
        if (re[p] > 1.0) {
            std::sort(g.begin(), g.end(), [](const auto &a, const auto &b) {
                if (a.second != b.second) return a.second > b.second;
                return a.first < b.first;
            });
        } else {
            std::sort(g.begin(), g.end(), [](const auto &a, const auto &b) {
                if (a.second != b.second) return a.second < b.second;
                return a.first < b.first;
            });
        }
        int take = static_cast<int>(std::ceil(ratio * static_cast<double>(g.size())));
        if (take > (int)g.size()) take = (int)g.size();
        take_of[(size_t)p] = take;
    }
    out.clear();
    for (int32_t p = 0; p < k; ++p)
        for (int64_t i = 0; i < take_of[(size_t)p]; ++i) set_mark(mk + groups[(size_t)p][(size_t)i].first);
    collect_marked(n, mark, out);
}

// gpu_lp_boundary.cu:354-426 for one boundary vertex: histogram over labels
// below min(k,32); score = count * (1 + pen) for non-zero counts; the
// shuffle-tree argmax prefers the smallest 5-bit bit-reversed label among
// equal scores; the winner must beat the raw own-label count.
inline int32_t lp_decide(const int64_t *off, const int32_t *col, const int32_t *L, int32_t b,
                         const double *pen, int32_t k, int64_t *hist) {
    const int E = std::min(k, kLpLabelCap);
    for (int l = 0; l < E; ++l) hist[l] = 0;
    for (int64_t e = off[b]; e < off[b + 1]; ++e) {
        int32_t l = L[col[e]];
        if (l >= 0 && l < E) ++hist[l];
    }
    int32_t own = L[b];
    double base = (own >= 0 && own < E) ? (double)hist[own] : -1e300;
    int best_l = -1;
    double best_s = 0.0;
    for (int l = 0; l < E; ++l) {
        double s = (double)hist[l];
        if (s > 0.0) s *= (1.0 + pen[l]);
        if (best_l < 0 || s > best_s || (s == best_s && bitrev5(l) < bitrev5(best_l))) {
            best_l = l;
            best_s = s;
        }
    }
    return (best_s > base) ? best_l : own;
}

}  // namespace

int ref_select_boundary(int64_t n, const int64_t *off, const int32_t *col,
                        const int32_t *labels, const int32_t *prev, int64_t n_prev,
                        const double *re, int32_t k, int32_t t, int32_t *out, int64_t *n_out) {
    if (n <= 0 || k <= 0 || !labels || !re || !out || !n_out || n_prev < 0) return REF_E_ARG;
    std::vector<int32_t> pv(prev, prev + n_prev), o;
    std::vector<uint8_t> mark((size_t)n, 0);
    SelectCounts sc;
    select_boundary(n, off, col, labels, pv, re, k, t, o, mark, sc);
    std::copy(o.begin(), o.end(), out);
    *n_out = (int64_t)o.size();
    return REF_OK;
}

int ref_lp_round(int64_t n, const int64_t *off, const int32_t *col, const int32_t *labels,
                 const int32_t *b, int64_t nb, const double *pen, int32_t k, int32_t *labels_out) {
    if (n <= 0 || k <= 0 || !labels || !pen || !labels_out || nb < 0) return REF_E_ARG;
    for (int64_t i = 0; i < nb; ++i)
        if (b[i] < 0 || b[i] >= n) return REF_E_ARG;
#pragma omp parallel
    {
        std::vector<int64_t> hist((size_t)kLpLabelCap);
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < nb; ++i)
            labels_out[b[i]] = lp_decide(off, col, labels, b[i], pen, k, hist.data());
    }
    return REF_OK;
}

// phase2.cpp:256-449 at one rank.
int ref_phase2(int64_t n, const int64_t *off, const int32_t *col, int32_t k,
               int32_t max_rounds, double epsilon, int32_t *labels,
               int64_t *vc_out, int64_t *ic_out, ref_p2_metrics *m, ref_p2_trace *tr) {
    if (n <= 0 || k <= 0 || max_rounds < 0 || !labels || !m) return REF_E_ARG;
    std::memset(m, 0, sizeof(*m));
    std::vector<int64_t> vc((size_t)k), ic((size_t)k);
    int64_t prev_cut = 0;
    int rc = ref_stats(n, off, col, k, labels, vc.data(), ic.data(), &prev_cut);
    if (rc != REF_OK) return rc;
    m->initial_cut = prev_cut;
    int cc = 0;
    std::vector<int32_t> B((size_t)n);
    for (int64_t v = 0; v < n; ++v) B[(size_t)v] = (int32_t)v;
    std::vector<int32_t> nextB;
    std::vector<uint8_t> mark((size_t)n, 0);
    std::vector<double> pen((size_t)k), re((size_t)k);
    std::vector<int32_t> newL(labels, labels + n);
    std::vector<int32_t> changed;
    int32_t rounds = 0, reason = 0;
    int64_t lp_edges = 0;
    for (int32_t t = 1; t <= max_rounds; ++t) {
        rounds = t;
        ref_penalties(k, vc.data(), ic.data(), t, pen.data(), re.data());
        SelectCounts sc;
        int64_t nprev = (int64_t)B.size();
        select_boundary(n, off, col, labels, B, re.data(), k, t, nextB, mark, sc);
        B.swap(nextB);
        int64_t bdeg = 0;
        for (int32_t v : B) bdeg += deg_of(off, v);
        lp_edges += bdeg;
        if (tr) {
            int i = t - 1;
            if (tr->prev_count) tr->prev_count[i] = nprev;
            if (tr->prev_deg) tr->prev_deg[i] = sc.prev_deg;
            if (tr->cand_count) tr->cand_count[i] = sc.cand;
            if (tr->scan_len) tr->scan_len[i] = sc.scan;
            if (tr->bd_count) tr->bd_count[i] = sc.bd;
            if (tr->b_count) tr->b_count[i] = (int64_t)B.size();
            if (tr->b_deg) tr->b_deg[i] = bdeg;
            if (tr->chg_count) tr->chg_count[i] = 0;
            if (tr->chg_deg) tr->chg_deg[i] = 0;
            if (tr->cut) tr->cut[i] = -1;
        }
        if (B.empty()) { reason = 1; break; }
        // Jacobi update: every decision reads the labels of round t-1.
        changed.clear();
#pragma omp parallel
        {
            std::vector<int64_t> h((size_t)kLpLabelCap);
            std::vector<int32_t> mine;
#pragma omp for schedule(dynamic, 256) nowait
            for (int64_t i = 0; i < (int64_t)B.size(); ++i) {
                const int32_t b = B[(size_t)i];
                int32_t nl = lp_decide(off, col, labels, b, pen.data(), k, h.data());
                if (nl != labels[b]) { newL[(size_t)b] = nl; mine.push_back(b); }
            }
#pragma omp critical
            changed.insert(changed.end(), mine.begin(), mine.end());
        }
        int64_t cdeg = 0;
        for (int32_t v : changed) { labels[v] = newL[(size_t)v]; cdeg += deg_of(off, v); }
        if (tr) {
            if (tr->chg_count) tr->chg_count[t - 1] = (int64_t)changed.size();
            if (tr->chg_deg) tr->chg_deg[t - 1] = cdeg;
        }
        if (changed.empty()) { reason = 2; break; }
        int64_t cur = 0;
        rc = ref_stats(n, off, col, k, labels, vc.data(), ic.data(), &cur);
        if (rc != REF_OK) return rc;
        if (tr && tr->cut) tr->cut[t - 1] = cur;
        double delta = (prev_cut > 0) ? std::abs((double)(cur - prev_cut) / (double)prev_cut) : 1.0;
        if (delta < epsilon) ++cc; else cc = 0;
        prev_cut = cur;
        if (cc >= kConvergeRounds) { reason = 3; break; }
    }
    // Final statistics (phase2.cpp:424-441).
    int64_t dummy_cut = 0;
    rc = ref_stats(n, off, col, k, labels, vc.data(), ic.data(), &dummy_cut);
    if (rc != REF_OK) return rc;
    int64_t tv = 0, te = 0;
    for (int32_t i = 0; i < k; ++i) { tv += vc[(size_t)i]; te += ic[(size_t)i]; }
    double exp_v = static_cast<double>(tv) / k;
    double exp_e = (te > 0) ? static_cast<double>(te) / k : 1.0;
    double max_v = 0.0, max_e = 0.0;
    for (int32_t i = 0; i < k; ++i) {
        double rv = (exp_v > 0) ? static_cast<double>(vc[(size_t)i]) / exp_v : 1.0;
        double r_e = (exp_e > 0) ? static_cast<double>(ic[(size_t)i]) / exp_e : 1.0;
        max_v = std::max(max_v, rv);
        max_e = std::max(max_e, r_e);
    }
    m->edge_cut = prev_cut;
    m->vertex_balance = max_v;
    m->edge_balance = max_e;
    m->total_vertices = tv;
    m->total_edges = te;
    m->rounds = rounds;
    m->exit_reason = reason;
    m->lp_edges = lp_edges;
    if (vc_out) std::copy(vc.begin(), vc.end(), vc_out);
    if (ic_out) std::copy(ic.begin(), ic.end(), ic_out);
    return REF_OK;
}


// ---- Phase 2 as P ranks (the reference's multi-rank path at one process) ----

namespace {

// One rank's share: owned rows in ascending global id (partition.cpp:193-196),
// columns as local ids for owned neighbours and nl + slot for ghosts, slots
// given in order of first appearance over the rows (partition.cpp:225-257).
struct RankShare {
    int64_t nl = 0;
    std::vector<int32_t> gid;        // local id -> global id
    std::vector<int64_t> off;        // nl + 1
    std::vector<int32_t> col;        // unified ids
    std::vector<int32_t> ghost_gid;  // slot -> global id
    std::unordered_map<int32_t, int32_t> slot_of;  // ghost global id -> slot
    std::vector<int32_t> lab;        // unified labels: owned, then ghosts
    std::vector<int32_t> B;          // this rank's boundary (local ids)
    std::vector<uint8_t> ghost_in_b; // ghosts in some rank's boundary (global scope)
};

inline int32_t owner_of(int64_t v, int32_t P) { return (int32_t)(v % P); }

void build_share(int64_t n, const int64_t *off, const int32_t *col, const int32_t *labels,
                 int32_t P, int32_t r, RankShare &s) {
    for (int64_t v = r; v < n; v += P) s.gid.push_back((int32_t)v);
    s.nl = (int64_t)s.gid.size();
    s.off.assign(1, 0);
    for (int64_t i = 0; i < s.nl; ++i) {
        const int32_t v = s.gid[(size_t)i];
        for (int64_t e = off[v]; e < off[v + 1]; ++e) {
            const int32_t u = col[e];
            if (owner_of(u, P) == r) {
                s.col.push_back((int32_t)(u / P));  // rank of u among the owned ids
            } else {
                auto it = s.slot_of.find(u);
                int32_t slot;
                if (it == s.slot_of.end()) {
                    slot = (int32_t)s.ghost_gid.size();
                    s.slot_of.emplace(u, slot);
                    s.ghost_gid.push_back(u);
                } else {
                    slot = it->second;
                }
                s.col.push_back((int32_t)(s.nl + slot));
            }
        }
        s.off.push_back((int64_t)s.col.size());
    }
    // Owned labels, then each ghost's label as its owner reports it
    // (partition.cpp:264-318).
    s.lab.resize((size_t)(s.nl + (int64_t)s.ghost_gid.size()));
    for (int64_t i = 0; i < s.nl; ++i) s.lab[(size_t)i] = labels[s.gid[(size_t)i]];
    for (size_t g = 0; g < s.ghost_gid.size(); ++g) s.lab[(size_t)s.nl + g] = labels[s.ghost_gid[g]];
    s.ghost_in_b.assign(s.ghost_gid.size(), 0);
}

// Partial counts of one rank (phase2.cpp:55-98, 214-236 before the
// Allreduce): owned labels, intra entries and differing entries of its rows.
int share_counts(const RankShare &s, int32_t k, std::vector<int64_t> &vc, std::vector<int64_t> &ic,
                 int64_t &cut2) {
    for (int64_t i = 0; i < s.nl; ++i) {
        const int32_t li = s.lab[(size_t)i];
        const bool ok = li >= 0 && li < k;
        if (ok) ++vc[(size_t)li];
        for (int64_t e = s.off[(size_t)i]; e < s.off[(size_t)i + 1]; ++e) {
            const int32_t lj = s.lab[(size_t)s.col[(size_t)e]];
            if (ok && lj == li) ++ic[(size_t)li];
            if (li == -1 || lj == -1) return REF_E_INVALID_LABEL;
            if (li != lj) ++cut2;
        }
    }
    return REF_OK;
}

// The Allreduce of every rank's partial counts.
int global_counts(const std::vector<RankShare> &R, int32_t k, std::vector<int64_t> &vc,
                  std::vector<int64_t> &ic, int64_t &cut) {
    std::fill(vc.begin(), vc.end(), 0);
    std::fill(ic.begin(), ic.end(), 0);
    int64_t cut2 = 0;
    for (const RankShare &s : R) {
        const int rc = share_counts(s, k, vc, ic, cut2);
        if (rc != REF_OK) return rc;
    }
    if (cut2 & 1) return REF_E_ODD_CUT;
    for (auto &x : ic) x /= 2;  // phase2.cpp:110
    cut = cut2 >> 1;
    return REF_OK;
}

// One boundary vertex of a rank: local id, label, degree, global id.
struct BdRec {
    int32_t i, label;
    int64_t deg;
    int32_t gid;
};

}  // namespace

// phase2.cpp:256-449 run by P ranks, each on its share of the graph
// (owner(v) = v mod P, partition.cpp:55-57), with ghost labels
// (partition.cpp:182-318) refreshed every round from the allgathered deltas
// (phase2.cpp:122-154, 349-381).  Statistics and cut are the ranks' partial
// counts summed (the Allreduce of phase2.cpp:100-117, 229-236).
//
// Selection scope 0 follows the reference rank by rank: candidates are the
// rank's previous boundary and its neighbours, of which only owned vertices
// are tested (gpu_lp_boundary.cu:251-252 skips ids past the owned rows), and
// each label's ceil(ratio * c) is taken over the rank's own boundary vertices
// (gpu_lp_boundary.cu:274-302 on the rank's vectors).  The boundary test
// reads a ghost neighbour's label from the ghost labels (the reference
// indexes its owned-label vector with the ghost id there, out of bounds,
// gpu_lp_boundary.cu:259-260; the LP kernel reads the same ghost labels).
// Scope 1 selects over all ranks together (candidates from every rank's
// previous boundary, each label's take over all boundary vertices, ties by
// global id), so the run equals the one-rank run for any P.
//
// labels: global, in/out.  ghost_updates (nullable, max_rounds entries):
// ghost labels rewritten by each round's refresh, summed over ranks.
int ref_phase2_ranks(int64_t n, const int64_t *off, const int32_t *col, int32_t k, int32_t P,
                     int32_t scope, int32_t max_rounds, double epsilon, int32_t *labels,
                     int64_t *vc_out, int64_t *ic_out, ref_p2_metrics *m, ref_p2_trace *tr,
                     int64_t *ghost_updates) {
    if (n <= 0 || k <= 0 || P <= 0 || P > n || (scope != 0 && scope != 1) || max_rounds < 0 ||
        !labels || !m)
        return REF_E_ARG;
    std::memset(m, 0, sizeof(*m));
    for (int64_t v = 0; v < n; ++v)
        if (labels[v] < 0 || labels[v] >= k) return REF_E_INVALID_LABEL;
    std::vector<RankShare> R((size_t)P);
    for (int32_t r = 0; r < P; ++r) build_share(n, off, col, labels, P, r, R[(size_t)r]);
    std::vector<int64_t> vc((size_t)k), ic((size_t)k);
    int64_t prev_cut = 0;
    int rc = global_counts(R, k, vc, ic, prev_cut);
    if (rc != REF_OK) return rc;
    m->initial_cut = prev_cut;
    for (RankShare &s : R) {  // phase2.cpp:285-287: every owned vertex
        s.B.resize((size_t)s.nl);
        for (int64_t i = 0; i < s.nl; ++i) s.B[(size_t)i] = (int32_t)i;
        std::fill(s.ghost_in_b.begin(), s.ghost_in_b.end(), 1);
    }
    std::vector<double> pen((size_t)k), re((size_t)k);
    int cc = 0;
    int32_t rounds = 0, reason = 0;
    int64_t lp_edges = 0;
    for (int32_t t = 1; t <= max_rounds; ++t) {
        rounds = t;
        ref_penalties(k, vc.data(), ic.data(), t, pen.data(), re.data());
        const double ratio = exp((-t) / (5.0));
        // Boundary records of every rank (candidates, then the test).
        std::vector<std::vector<BdRec>> bd((size_t)P);
        int64_t nprev = 0, dprev = 0;
        for (int32_t r = 0; r < P; ++r) {
            RankShare &s = R[(size_t)r];
            std::vector<uint8_t> cand((size_t)s.nl, 0), in_b((size_t)s.nl, 0);
            for (int32_t i : s.B) {
                in_b[(size_t)i] = 1;
                ++nprev;
                dprev += s.off[(size_t)i + 1] - s.off[(size_t)i];
            }
            for (int64_t i = 0; i < s.nl; ++i) {
                bool c = in_b[(size_t)i] != 0;
                for (int64_t e = s.off[(size_t)i]; !c && e < s.off[(size_t)i + 1]; ++e) {
                    const int32_t j = s.col[(size_t)e];
                    c = j < s.nl ? in_b[(size_t)j] != 0
                                 : (scope == 1 && s.ghost_in_b[(size_t)(j - s.nl)] != 0);
                }
                cand[(size_t)i] = c;
            }
            for (int64_t i = 0; i < s.nl; ++i) {
                if (!cand[(size_t)i]) continue;
                const int32_t li = s.lab[(size_t)i];
                bool b = false;
                for (int64_t e = s.off[(size_t)i]; !b && e < s.off[(size_t)i + 1]; ++e)
                    b = s.lab[(size_t)s.col[(size_t)e]] != li;
                if (b)
                    bd[(size_t)r].push_back({(int32_t)i, li, s.off[(size_t)i + 1] - s.off[(size_t)i],
                                             s.gid[(size_t)i]});
            }
        }
        // Per label, the first ceil(ratio * size) by (degree desc if RE > 1
        // else asc, id): per rank (scope 0) or over all ranks (scope 1).
        auto take_groups = [&](std::vector<BdRec> &recs, std::vector<uint8_t> &pick) {
            std::vector<std::vector<size_t>> groups((size_t)k);
            for (size_t x = 0; x < recs.size(); ++x) groups[(size_t)recs[x].label].push_back(x);
            pick.assign(recs.size(), 0);
            for (int32_t p = 0; p < k; ++p) {
                auto &g = groups[(size_t)p];
                if (g.empty()) continue;
                const bool desc = re[p] > 1.0;
                std::sort(g.begin(), g.end(), [&](size_t a, size_t b) {
                    const BdRec &x = recs[a], &y = recs[b];
                    if (x.deg != y.deg) return desc ? x.deg > y.deg : x.deg < y.deg;
                    return x.gid < y.gid;
                });
                int take = static_cast<int>(std::ceil(ratio * static_cast<double>(g.size())));
                if (take > (int)g.size()) take = (int)g.size();
                for (int q = 0; q < take; ++q) pick[g[(size_t)q]] = 1;
            }
        };
        int64_t nbd = 0, nb = 0, bdeg = 0;
        if (scope == 0) {
            for (int32_t r = 0; r < P; ++r) {
                std::vector<uint8_t> pick;
                take_groups(bd[(size_t)r], pick);
                RankShare &s = R[(size_t)r];
                s.B.clear();
                for (size_t x = 0; x < pick.size(); ++x)
                    if (pick[x]) s.B.push_back(bd[(size_t)r][x].i);
                nbd += (int64_t)bd[(size_t)r].size();
            }
        } else {
            std::vector<BdRec> all;
            std::vector<int32_t> rank_of;
            for (int32_t r = 0; r < P; ++r)
                for (const BdRec &x : bd[(size_t)r]) { all.push_back(x); rank_of.push_back(r); }
            std::vector<uint8_t> pick;
            take_groups(all, pick);
            for (RankShare &s : R) s.B.clear();
            for (size_t x = 0; x < pick.size(); ++x)
                if (pick[x]) R[(size_t)rank_of[x]].B.push_back(all[x].i);
            for (RankShare &s : R) std::sort(s.B.begin(), s.B.end());
            nbd = (int64_t)all.size();
        }
        for (const RankShare &s : R) {
            nb += (int64_t)s.B.size();
            for (int32_t i : s.B) bdeg += s.off[(size_t)i + 1] - s.off[(size_t)i];
        }
        lp_edges += bdeg;
        if (tr) {
            const int q = t - 1;
            if (tr->prev_count) tr->prev_count[q] = nprev;
            if (tr->prev_deg) tr->prev_deg[q] = dprev;
            if (tr->cand_count) tr->cand_count[q] = -1;
            if (tr->scan_len) tr->scan_len[q] = -1;
            if (tr->bd_count) tr->bd_count[q] = nbd;
            if (tr->b_count) tr->b_count[q] = nb;
            if (tr->b_deg) tr->b_deg[q] = bdeg;
            if (tr->chg_count) tr->chg_count[q] = 0;
            if (tr->chg_deg) tr->chg_deg[q] = 0;
            if (tr->cut) tr->cut[q] = -1;
        }
        if (ghost_updates) ghost_updates[t - 1] = 0;
        if (nb == 0) { reason = 1; break; }  // every rank's boundary is empty
        // Each rank decides on its boundary from its unified labels (Jacobi),
        // then builds its delta list (phase2.cpp:349-366).
        std::vector<std::pair<int32_t, int32_t>> deltas;  // (global id, new label), rank order
        int64_t cdeg = 0;
        std::vector<int64_t> hist((size_t)kLpLabelCap);
        for (RankShare &s : R) {
            std::vector<std::pair<int32_t, int32_t>> mine;
            for (int32_t i : s.B) {
                const int32_t nl = lp_decide(s.off.data(), s.col.data(), s.lab.data(), i, pen.data(),
                                             k, hist.data());
                if (nl != s.lab[(size_t)i]) mine.emplace_back(i, nl);
            }
            for (const auto &d : mine) {
                s.lab[(size_t)d.first] = d.second;
                deltas.emplace_back(s.gid[(size_t)d.first], d.second);
                cdeg += s.off[(size_t)d.first + 1] - s.off[(size_t)d.first];
            }
        }
        // Allgather of the deltas and the ghost refresh (phase2.cpp:122-154,
        // 368-376); in scope 1 the ranks also share their new boundary.
        int64_t refreshed = 0;
        for (RankShare &s : R) {
            for (const auto &d : deltas) {
                auto it = s.slot_of.find(d.first);
                if (it != s.slot_of.end()) {
                    s.lab[(size_t)(s.nl + it->second)] = d.second;
                    ++refreshed;
                }
            }
            if (scope == 1) {
                std::fill(s.ghost_in_b.begin(), s.ghost_in_b.end(), 0);
                for (const RankShare &o : R)
                    for (int32_t i : o.B) {
                        auto it = s.slot_of.find(o.gid[(size_t)i]);
                        if (it != s.slot_of.end()) s.ghost_in_b[(size_t)it->second] = 1;
                    }
            }
        }
        if (ghost_updates) ghost_updates[t - 1] = refreshed;
        if (tr) {
            if (tr->chg_count) tr->chg_count[t - 1] = (int64_t)deltas.size();
            if (tr->chg_deg) tr->chg_deg[t - 1] = cdeg;
        }
        if (deltas.empty()) { reason = 2; break; }  // phase2.cpp:384-390
        int64_t cur = 0;
        rc = global_counts(R, k, vc, ic, cur);
        if (rc != REF_OK) return rc;
        if (tr && tr->cut) tr->cut[t - 1] = cur;
        double delta = (prev_cut > 0) ? std::abs((double)(cur - prev_cut) / (double)prev_cut) : 1.0;
        if (delta < epsilon) ++cc; else cc = 0;
        prev_cut = cur;
        if (cc >= kConvergeRounds) { reason = 3; break; }
    }
    int64_t dummy = 0;
    rc = global_counts(R, k, vc, ic, dummy);
    if (rc != REF_OK) return rc;
    int64_t tv = 0, te = 0;
    for (int32_t i = 0; i < k; ++i) { tv += vc[(size_t)i]; te += ic[(size_t)i]; }
    double exp_v = static_cast<double>(tv) / k;
    double exp_e = (te > 0) ? static_cast<double>(te) / k : 1.0;
    double max_v = 0.0, max_e = 0.0;
    for (int32_t i = 0; i < k; ++i) {
        max_v = std::max(max_v, (exp_v > 0) ? static_cast<double>(vc[(size_t)i]) / exp_v : 1.0);
        max_e = std::max(max_e, (exp_e > 0) ? static_cast<double>(ic[(size_t)i]) / exp_e : 1.0);
    }
    for (const RankShare &s : R)
        for (int64_t i = 0; i < s.nl; ++i) labels[s.gid[(size_t)i]] = s.lab[(size_t)i];
    m->edge_cut = prev_cut;
    m->vertex_balance = max_v;
    m->edge_balance = max_e;
    m->total_vertices = tv;
    m->total_edges = te;
    m->rounds = rounds;
    m->exit_reason = reason;
    m->lp_edges = lp_edges;
    if (vc_out) std::copy(vc.begin(), vc.end(), vc_out);
    if (ic_out) std::copy(ic.begin(), ic.end(), ic_out);
    return REF_OK;
}

}  // extern "C"
