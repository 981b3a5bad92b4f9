// hubs.hip -- find_hub_nodes and find_first_seed (init.cpp:139-187) at one
// rank.  The hub order is the tie order libstdc++'s std::sort leaves on the
// (id, degree) sequence, so the partition steps of its introsort are
// reproduced one by one on the ranges that reach the top of the order: on
// host threads for short sequences, on the device for long ones (the step's
// passes over the range as kernels, several ranges and jobs per launch),
// with the short ranges finished on the host in either case.
#include <algorithm>
#include <atomic>
#include <memory>
#include <thread>
#include <utility>
#include <vector>

#include "dev_common.hpp"
#include "hubs.hpp"
#include "opbatch.hpp"

namespace dmolp {
namespace {

// Ranges at least this long are partitioned by several threads.
constexpr int64_t kParPartitionMin = int64_t(1) << 20;


// The partition step of libstdc++'s __introsort_loop on a[first, last)
// after the median has been moved to a[first] (__unguarded_partition of
// [first + 1, last) around a[first]), sequential.  Returns the cut.
int64_t partition_seq(DegPair *a, int64_t first, int64_t last) {
    const ByDegree less;
    int64_t lo = first + 1, hi = last;
    const DegPair *pivot = a + first;
    for (;;) {
        while (less(a[lo], *pivot)) ++lo;
        --hi;
        while (less(*pivot, a[hi])) --hi;
        if (!(lo < hi)) break;
        std::swap(a[lo], a[hi]);
        ++lo;
    }
    return lo;
}

// The same step on T threads, with the same result.  The sequential scans
// stop, from the left, at elements not below the pivot (L_1 < L_2 < ...)
// and, from the right, at elements not above it (R_1 > R_2 > ...), and
// swap the j-th pair while L_j < R_j.  A swap changes only the two
// positions it names, and the scans have already passed them, so L and R
// can be read off the array before any swap; the pivot at `first` is never
// swapped (the right scan can stop there only once the scans have crossed).
// With m swaps (L_j < R_j holds exactly for j <= m: L rises, R falls), the
// left scan then stops at min(L_{m+1}, R_m): the next original candidate,
// or position R_m, which now holds the old L_m, not below the pivot; with
// no swap it stops at L_1.  Every L_j with j <= m lies below every such
// R_j, so the swaps touch disjoint positions and the threads' reads during
// the swaps (their own slice's left candidates, their own run of right
// candidates) never meet another thread's writes.
int64_t partition_par(DegPair *a, int64_t first, int64_t last, int T) {
    const int P = a[first].second;
    const int64_t len = last - first;
    // Fine slices, so that locating the j-th candidate scans a short range.
    const int S = (int)std::min<int64_t>(4096, std::max<int64_t>(1, len / 4096));
    const auto lo_of = [&](int64_t s) { return first + len * s / S; };
    const auto is_l = [&](int64_t p) { return p != first && a[p].second >= P; };
    const auto is_r = [&](int64_t p) { return a[p].second <= P; };
    // Candidates per slice, the slices dealt to the threads in blocks.
    std::vector<int64_t> cl((size_t)S), cr((size_t)S);
    run_threads(T, [&](int t) {
        for (int64_t s = (int64_t)S * t / T; s < (int64_t)S * (t + 1) / T; ++s) {
            int64_t l = 0, r = 0;
            for (int64_t p = lo_of(s); p < lo_of(s + 1); ++p) {
                const int d = a[p].second;
                l += d >= P ? 1 : 0;
                r += d <= P ? 1 : 0;
            }
            cl[(size_t)s] = l - (s == 0 ? 1 : 0);  // the pivot's position is not a left candidate
            cr[(size_t)s] = r;
        }
    });
    // bl[s]: left candidates (0-based index) before slice s; br[s]: right
    // candidates in slices s and above (R is numbered from the right end).
    std::vector<int64_t> bl((size_t)S + 1, 0), br((size_t)S + 1, 0);
    for (int s = 0; s < S; ++s) bl[(size_t)s + 1] = bl[(size_t)s] + cl[(size_t)s];
    for (int s = S - 1; s >= 0; --s) br[(size_t)s] = br[(size_t)s + 1] + cr[(size_t)s];
    const int64_t nl = bl[(size_t)S], nr = br[0];
    if (nl == 0) return partition_seq(a, first, last);  // not reached: the median's sample holds one
    // Slices holding L_j and R_j, then their positions by a scan of that
    // slice, on the array as it is before any swap.
    const auto slice_l = [&](int64_t j) {  // bl[s] <= j < bl[s + 1]
        return (int)(std::upper_bound(bl.begin(), bl.end(), j) - bl.begin()) - 1;
    };
    const auto slice_r = [&](int64_t j) {  // br[s + 1] <= j < br[s]; br falls with s
        return (int)(std::lower_bound(br.begin(), br.end(), j, std::greater<int64_t>()) - br.begin()) - 1;
    };
    const auto pos_l = [&](int64_t j) {
        const int s = slice_l(j);
        int64_t c = bl[(size_t)s];
        for (int64_t p = lo_of(s);; ++p)
            if (is_l(p) && c++ == j) return p;
    };
    const auto pos_r = [&](int64_t j) {
        const int s = slice_r(j);
        int64_t c = br[(size_t)s + 1];
        for (int64_t p = lo_of(s + 1) - 1;; --p)
            if (is_r(p) && c++ == j) return p;
    };
    // m: the first j with L_j >= R_j.
    int64_t lo = 0, hi = std::min(nl, nr);
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        const int sl = slice_l(mid), sr = slice_r(mid);
        const bool lt = sl != sr ? sl < sr : pos_l(mid) < pos_r(mid);
        if (lt) lo = mid + 1;
        else hi = mid;
    }
    const int64_t m = lo;
    if (m == 0) return pos_l(0);
    const int64_t cut_l = m < nl ? pos_l(m) : last, cut_r = pos_r(m - 1);
    // Swaps: thread t takes the pairs j in [m t / T, m (t + 1) / T), walking
    // up from L at the first and down from R at the first.
    std::vector<int64_t> ls((size_t)T, -1), rs((size_t)T, -1);
    run_threads(T, [&](int t) {
        const int64_t j0 = m * t / T;
        if (j0 < m * (t + 1) / T) {
            ls[(size_t)t] = pos_l(j0);
            rs[(size_t)t] = pos_r(j0);
        }
    });
    run_threads(T, [&](int t) {
        int64_t j = m * t / T;
        const int64_t j1 = m * (t + 1) / T;
        if (j >= j1) return;
        int64_t l = ls[(size_t)t], r = rs[(size_t)t];
        for (;;) {
            std::swap(a[l], a[r]);
            if (++j == j1) break;
            for (++l; !is_l(l); ++l) {}
            for (--r; !is_r(r); --r) {}
        }
    });
    return std::min(cut_l, cut_r);
}

// The positions [keep, last0) of a[first0, last0) exactly as the
// introsort loop of std::sort(a, a + n, ByDegree) started on that range with
// depth limit depth0 (2 lg n for the whole array), followed by the final
// insertion sort, leaves them; introsort_top is the whole array.  For the
// whole array: the positions [keep, n) of a[0, n) exactly as std::sort(a, a + n, ByDegree)
// leaves them with libstdc++ (bits/stl_algo.h: __introsort_loop with
// median-of-three pivots and unguarded partitions, depth limit 2 lg n, heap
// sort past it, then one insertion sort).  The comparison is not a total
// order (ties in degree), so the tie order at the top is that algorithm's,
// and it is reproduced step for step on every range that reaches into
// [keep, n); ranges wholly below `keep` are left as they are, since no step
// moves an element across a partition cut and the final insertion sort
// only reorders within the ranges the partitions leave (every element of a
// range is no greater than every element of the next).  Cost: the
// partitions along the path to the top, about 2n, instead of n lg n; the
// long ones run on `threads` threads (partition_par).
void introsort_from(DegPair *a, int64_t first0, int64_t last0, int depth0, int64_t keep,
                    int threads) {
    const ByDegree less;
    std::vector<std::pair<int64_t, int64_t>> leaves;  // ranges the final insertion sort orders
    // __introsort_loop(first, last, depth), right parts first (recursion),
    // then the loop over the left part; ranges below `keep` are skipped.
    struct Frame { int64_t first, last; int depth; };
    std::vector<Frame> stack{{first0, last0, depth0}};
    while (!stack.empty()) {
        Frame f = stack.back();
        stack.pop_back();
        int64_t first = f.first, last = f.last;
        int depth = f.depth;
        while (last - first > 16) {
            if (last <= keep) { first = last; break; }  // no effect on [keep, n)
            if (depth == 0) {  // __partial_sort(first, last, last): heap sort
                std::make_heap(a + first, a + last, less);
                std::sort_heap(a + first, a + last, less);
                first = last;  // sorted: the insertion sort leaves it as it is
                break;
            }
            --depth;
            // __unguarded_partition_pivot
            const int64_t mid = first + (last - first) / 2;
            {
                DegPair *r = a + first, *x = a + first + 1, *y = a + mid, *z = a + last - 1;
                if (less(*x, *y)) {
                    if (less(*y, *z)) std::iter_swap(r, y);
                    else if (less(*x, *z)) std::iter_swap(r, z);
                    else std::iter_swap(r, x);
                } else if (less(*x, *z)) std::iter_swap(r, x);
                else if (less(*y, *z)) std::iter_swap(r, z);
                else std::iter_swap(r, y);
            }
            int64_t cut;
            if (threads > 1 && last - first >= kParPartitionMin) {
                cut = partition_par(a, first, last, threads);
            } else {
                cut = partition_seq(a, first, last);
            }
            // Recursion on [cut, last) runs before the loop goes on with
            // [first, cut): push the left part, process the right part now.
            stack.push_back({first, cut, depth});
            first = cut;
        }
        if (last - first > 0 && last > keep && last - first <= 16) leaves.push_back({first, last});
    }
    // __final_insertion_sort restricted to the ranges that reach [keep, n).
    for (const auto &lf : leaves) {
        for (int64_t i = lf.first + 1; i < lf.second; ++i) {
            const DegPair val = a[i];
            int64_t j = i;
            while (j > lf.first && less(val, a[j - 1])) { a[j] = a[j - 1]; --j; }
            a[j] = val;
        }
    }
}

int depth_limit(int64_t n) {  // 2 lg n, __introsort_loop's initial depth
    int lg = 0;
    while ((int64_t(1) << (lg + 1)) <= n) ++lg;
    return 2 * lg;
}

void introsort_top(DegPair *a, int64_t n, int64_t keep, int threads) {
    introsort_from(a, 0, n, depth_limit(n), keep, threads);
}

// find_first_seed's result from the largest degree D and the smallest and
// largest ids of degree D (first_seed_host derives this form).
std::pair<int32_t, int32_t> first_seed_from(int32_t D, int64_t vmin, int64_t vmax) {
    const int64_t node = (vmin < vmax && vmin < (int64_t)D) ? vmin : vmax;
    return {(int32_t)node, D};
}

}  // namespace

// The Lorenz tangent of init.cpp:139-177.  The tangent needs only the last
// two cumulative shares, which are exact integer sums divided by the total
// whatever the order (every partial sum is an integer below 2^53, so the
// reference's double accumulation gives these exact sums), so M follows
// from the vertex count, the degree total and the largest degree.  Double
// arithmetic in source order.
int64_t hub_count(int64_t n, int64_t isum, int32_t dmax) {
    const int N = (int)n;
    if (N < 2) return 0;
    const double total = (double)isum;
    const double inv_N = 1.0 / N;
    const double x1 = N * inv_N;
    const double y1 = total / total;            // cum[N-1]
    const double x0 = (N - 1) * inv_N;
    const double y0 = (total - dmax) / total;   // cum[N-2]: all but the largest degree
    const double slope = (y1 - y0) / (x1 - x0);
    const double xi = x1 - y1 / slope;
    int64_t M = 0;
    for (int i = N - 1; i >= 0; --i) {
        const double x = (i + 1) * inv_N;
        if (x > xi) ++M;
        else break;
    }
    return M;
}

// init.cpp:139-177 at one rank: (id, degree) pairs in the degree map's
// iteration order (descending id), libstdc++ std::sort by degree, Lorenz
// tangent cut.  The number of hubs is known before sorting (hub_count), so
// only the top of the sorted order is produced (introsort_top).
std::vector<int32_t> find_hubs_host(const std::vector<int32_t> &deg, int threads, int64_t limit,
                                    int64_t *m_out) {
    const int64_t n = (int64_t)deg.size();
    std::vector<int32_t> hubs;
    if (m_out) *m_out = 0;
    if (n < 2) return hubs;
    const int T = n >= kParPartitionMin ? std::max(1, threads) : 1;
    // Degree total (every partial sum is an integer below 2^53, so the
    // reference's double accumulation is this exact sum) and maximum.
    std::vector<int64_t> tsum((size_t)T, 0);
    std::vector<int> tmax((size_t)T, 0);
    run_threads(T, [&](int t) {
        int64_t s = 0;
        int mx = 0;
        for (int64_t v = n * t / T; v < n * (t + 1) / T; ++v) {
            s += deg[(size_t)v];
            mx = std::max(mx, deg[(size_t)v]);
        }
        tsum[(size_t)t] = s;
        tmax[(size_t)t] = mx;
    });
    int64_t isum = 0;
    int dmax = 0;
    for (int t = 0; t < T; ++t) {
        isum += tsum[(size_t)t];
        dmax = std::max(dmax, tmax[(size_t)t]);
    }
    int64_t M = hub_count(n, isum, dmax);
    if (m_out) *m_out = M;
    if (M == 0) return hubs;
    if (limit >= 0 && limit < M) M = limit;  // only the first `limit` hubs of the order
    std::unique_ptr<DegPair[]> sd(new DegPair[(size_t)n]);  // (id, degree), descending id
    run_threads(T, [&](int t) {
        for (int64_t i = n * t / T; i < n * (t + 1) / T; ++i) {
            const int64_t v = n - 1 - i;
            sd[(size_t)i] = DegPair{(int)v, deg[(size_t)v]};
        }
    });
    introsort_top(sd.get(), n, n - M, T);
    hubs.reserve((size_t)M);
    for (int64_t i = n - 1; i >= n - M; --i) hubs.push_back(sd[(size_t)i].first);
    return hubs;
}

// init.cpp:179-187 over the degree map's order (descending id): the scan
// keeps (node, best) and takes v when deg v > best, or when deg v == best
// and v < best (the tie test compares an id with the running best degree).
// Once best reaches the largest degree D (at the largest id v0 of degree D)
// only the tie rule can move node, to ids of degree D below both v0 and D,
// and the last such move is the smallest one: the result is the smallest id
// of degree D when it lies below D and is not v0 itself, else v0.  Three
// reductions (D, then the smallest and largest ids of degree D) on
// `threads` threads instead of a sequential scan.
std::pair<int32_t, int32_t> first_seed_host(const std::vector<int32_t> &deg, int threads) {
    const int64_t n = (int64_t)deg.size();
    if (n == 0) return {-1, -1};
    const int T = n >= kParPartitionMin ? std::max(1, threads) : 1;
    std::vector<int32_t> tmax((size_t)T, -1);
    run_threads(T, [&](int t) {
        int32_t m = -1;
        for (int64_t v = n * t / T; v < n * (t + 1) / T; ++v) m = std::max(m, deg[(size_t)v]);
        tmax[(size_t)t] = m;
    });
    const int32_t D = *std::max_element(tmax.begin(), tmax.end());
    std::vector<int64_t> tlo((size_t)T, n), thi((size_t)T, -1);
    run_threads(T, [&](int t) {
        int64_t lo = n, hi = -1;
        for (int64_t v = n * t / T; v < n * (t + 1) / T; ++v)
            if (deg[(size_t)v] == D) {
                if (lo == n) lo = v;
                hi = v;
            }
        tlo[(size_t)t] = lo;
        thi[(size_t)t] = hi;
    });
    const int64_t vmin = *std::min_element(tlo.begin(), tlo.end());
    const int64_t vmax = *std::max_element(thi.begin(), thi.end());
    return first_seed_from(D, vmin, vmax);
}

// ---------------------------------------------------------------- device search
//
// The partition step of a range (partition_par's formulation) as five
// passes over the range: count, per chunk of kHsChunk entries, the left
// candidates (not below the pivot, the pivot's own position excluded) and
// the right candidates (not above it); scan the counts; write every left
// candidate's position to lpos[j] (the j-th from the left) and every right
// candidate's to rpos[j] (the j-th from the right); find m, the first j with
// lpos[j] >= rpos[j]; swap the pairs j < m.  The cut is min(lpos[m],
// rpos[m - 1]) (lpos[0] without swaps).  Every range of a step, of every
// job, is one entry of the step's frame table and its chunks are blocks of
// the same launches.
//
// The same steps under a total order (degree, then id descending) select
// the first vertices by (degree descending, id ascending): the seed
// fallback's candidates.  There only the set at the top matters, so the
// host sorts each short range outright.

namespace {

constexpr int kHsThreads = 256;
constexpr int kHsItems = 16;
constexpr int64_t kHsChunk = (int64_t)kHsThreads * kHsItems;

struct HsFrame {
    DegPair *a;            // the job's (id, degree) sequence
    int32_t *lpos, *rpos;  // the job's candidate positions, indexed from `first`
    int64_t first, last;
    int64_t chunk0;        // the frame's first chunk in the step's chunk table
    int64_t nchunks;
};
struct HsOut {
    DegPair P;  // the pivot
    int64_t nl, nr;
    unsigned long long m;
    int64_t cut;  // -1: no left candidate (the median-of-three sample rules it out)
};
// A job's sequence set-up: the (id, degree) pairs in descending id order and
// the reductions behind the Lorenz tangent and the first seed.
struct HsInit {
    const int32_t *deg;
    DegPair *a;
    int64_t n;
    int64_t chunk0, nchunks;
    unsigned long long *stat;  // [0] max (deg, ~id), [1] max (deg, id), [2] degree total
};

// ByDegree, or the total order (degree, then id descending).
template <bool Total>
__device__ __forceinline__ bool hs_less(const DegPair &x, const DegPair &y) {
    return x.second < y.second || (Total && x.second == y.second && x.first > y.first);
}
struct ByDegreeThenId {  // the total order on the host
    bool operator()(const DegPair &x, const DegPair &y) const {
        return x.second < y.second || (x.second == y.second && x.first > y.first);
    }
};

template <class T>
__device__ __forceinline__ int frame_of(const T *F, int nf, int64_t b) {
    int lo = 0, hi = nf - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (F[mid].chunk0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Sum of x over the block; every thread gets it.
__device__ __forceinline__ int64_t block_sum(int64_t x, int64_t *s) {
    x = wave_sum(x);
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if (lane_id() == 0) s[w] = x;
    __syncthreads();
    int64_t t = 0;
    for (int q = 0; q < kHsThreads / kWave; ++q) t += s[q];
    return t;
}

__global__ __launch_bounds__(kHsThreads) void k_hs_init(const HsInit *J, int nj) {
    const int j = frame_of(J, nj, blockIdx.x);
    const HsInit s = J[j];
    const int64_t lo = ((int64_t)blockIdx.x - s.chunk0) * kHsChunk;
    const int64_t hi = lo + kHsChunk < s.n ? lo + kHsChunk : s.n;
    unsigned long long kmin = 0, kmax = 0;
    int64_t sum = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kHsThreads) {
        const int64_t v = s.n - 1 - i;
        const int32_t d = s.deg[v];
        s.a[i] = DegPair{(int)v, d};
        const unsigned long long hi32 = (unsigned long long)(uint32_t)d << 32;
        const unsigned long long a0 = hi32 | (0xffffffffull - (unsigned long long)v);
        const unsigned long long a1 = hi32 | (unsigned long long)v;
        kmin = a0 > kmin ? a0 : kmin;
        kmax = a1 > kmax ? a1 : kmax;
        sum += d;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const unsigned long long b0 = __shfl_xor(kmin, o, kWave), b1 = __shfl_xor(kmax, o, kWave);
        kmin = b0 > kmin ? b0 : kmin;
        kmax = b1 > kmax ? b1 : kmax;
    }
    sum = wave_sum(sum);
    if (lane_id() == 0) {
        atomicMax(&s.stat[0], kmin);
        atomicMax(&s.stat[1], kmax);
        if (sum) atomicAdd(&s.stat[2], (unsigned long long)sum);
    }
}

// __move_median_to_first(first, first + 1, mid, last - 1): one thread a frame.
template <bool Total>
__global__ void k_hs_pivot(const HsFrame *F, HsOut *O) {
    if (threadIdx.x != 0) return;
    const HsFrame f = F[blockIdx.x];
    DegPair *a = f.a;
    const int64_t r = f.first, x = f.first + 1, y = f.first + (f.last - f.first) / 2, z = f.last - 1;
    int64_t pick;
    if (hs_less<Total>(a[x], a[y])) {
        if (hs_less<Total>(a[y], a[z])) pick = y;
        else if (hs_less<Total>(a[x], a[z])) pick = z;
        else pick = x;
    } else if (hs_less<Total>(a[x], a[z])) pick = x;
    else if (hs_less<Total>(a[y], a[z])) pick = z;
    else pick = y;
    const DegPair t = a[r];
    a[r] = a[pick];
    a[pick] = t;
    O[blockIdx.x].P = a[r];
}

template <bool Total>
__global__ __launch_bounds__(kHsThreads) void k_hs_count(const HsFrame *F, int nf, const HsOut *O,
                                                        int32_t *cl, int32_t *cr) {
    __shared__ int64_t s[kHsThreads / kWave];
    const int fi = frame_of(F, nf, blockIdx.x);
    const HsFrame f = F[fi];
    const DegPair P = O[fi].P;
    const int64_t lo = f.first + ((int64_t)blockIdx.x - f.chunk0) * kHsChunk;
    const int64_t hi = lo + kHsChunk < f.last ? lo + kHsChunk : f.last;
    int64_t l = 0, r = 0;
    for (int64_t p = lo + threadIdx.x; p < hi; p += kHsThreads) {
        const DegPair e = f.a[p];
        l += (p != f.first && !hs_less<Total>(e, P)) ? 1 : 0;
        r += !hs_less<Total>(P, e) ? 1 : 0;
    }
    l = block_sum(l, s);
    r = block_sum(r, s);
    if (threadIdx.x == 0) {
        cl[blockIdx.x] = (int32_t)l;
        cr[blockIdx.x] = (int32_t)r;
    }
}

// Per frame (one block): bl = left candidates in the chunks before each
// chunk, br = right candidates in the chunks after it; nl, nr; m starts at
// min(nl, nr).
__global__ __launch_bounds__(kHsThreads) void k_hs_scan(const HsFrame *F, const int32_t *cl,
                                                       const int32_t *cr, int64_t *bl, int64_t *br,
                                                       HsOut *O) {
    __shared__ int64_t sl[kHsThreads / kWave], sr[kHsThreads / kWave];
    const HsFrame f = F[blockIdx.x];
    const int w = threadIdx.x / kWave;
    int64_t runl = 0, runr = 0;
    for (int64_t t0 = 0; t0 < f.nchunks; t0 += kHsThreads) {
        const int64_t i = t0 + threadIdx.x;
        const int64_t l = i < f.nchunks ? cl[f.chunk0 + i] : 0;
        const int64_t r = i < f.nchunks ? cr[f.chunk0 + i] : 0;
        const int64_t il = wave_incl_scan(l), ir = wave_incl_scan(r);
        __syncthreads();
        if (lane_id() == kWave - 1) { sl[w] = il; sr[w] = ir; }
        __syncthreads();
        int64_t pl = 0, pr = 0, tl = 0, tr = 0;
        for (int q = 0; q < kHsThreads / kWave; ++q) {
            if (q < w) { pl += sl[q]; pr += sr[q]; }
            tl += sl[q];
            tr += sr[q];
        }
        if (i < f.nchunks) {
            bl[f.chunk0 + i] = runl + pl + il - l;
            br[f.chunk0 + i] = runr + pr + ir;  // inclusive for now
        }
        runl += tl;
        runr += tr;
    }
    for (int64_t i = threadIdx.x; i < f.nchunks; i += kHsThreads)  // the same thread wrote it
        br[f.chunk0 + i] = runr - br[f.chunk0 + i];
    if (threadIdx.x == 0) {
        O[blockIdx.x].nl = runl;
        O[blockIdx.x].nr = runr;
        O[blockIdx.x].m = (unsigned long long)(runl < runr ? runl : runr);
        O[blockIdx.x].cut = -1;
    }
}

template <bool Total>
__global__ __launch_bounds__(kHsThreads) void k_hs_scatter(const HsFrame *F, int nf, const HsOut *O,
                                                          const int32_t *cr, const int64_t *bl,
                                                          const int64_t *br) {
    __shared__ int32_t sl[kHsThreads / kWave], sr[kHsThreads / kWave];
    const int fi = frame_of(F, nf, blockIdx.x);
    const HsFrame f = F[fi];
    const DegPair P = O[fi].P;
    const int64_t lo = f.first + ((int64_t)blockIdx.x - f.chunk0) * kHsChunk;
    const int64_t hi = lo + kHsChunk < f.last ? lo + kHsChunk : f.last;
    const int64_t basel = bl[blockIdx.x], baser = br[blockIdx.x];
    const int64_t totr = cr[blockIdx.x];
    const int w = threadIdx.x / kWave;
    int64_t offl = 0, offr = 0;
    for (int64_t s0 = lo; s0 < hi; s0 += kHsThreads) {
        const int64_t p = s0 + threadIdx.x;
        const bool valid = p < hi;
        const DegPair e = f.a[valid ? p : lo];
        const bool isl = valid && p != f.first && !hs_less<Total>(e, P);
        const bool isr = valid && !hs_less<Total>(P, e);
        const uint64_t ml = __ballot(isl), mr = __ballot(isr);
        if (lane_id() == 0) {
            sl[w] = __popcll(ml);
            sr[w] = __popcll(mr);
        }
        __syncthreads();
        int64_t pl = 0, pr = 0, tl = 0, tr = 0;
        for (int q = 0; q < kHsThreads / kWave; ++q) {
            if (q < w) { pl += sl[q]; pr += sr[q]; }
            tl += sl[q];
            tr += sr[q];
        }
        if (isl) f.lpos[f.first + basel + offl + pl + lanes_below(ml)] = (int32_t)p;
        if (isr) {
            const int64_t incl = offr + pr + lanes_below(mr) + 1;  // right candidates of the chunk up to p
            f.rpos[f.first + baser + totr - incl] = (int32_t)p;
        }
        offl += tl;
        offr += tr;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kHsThreads) void k_hs_find_m(const HsFrame *F, int nf, HsOut *O) {
    __shared__ unsigned long long s[kHsThreads / kWave];
    const int fi = frame_of(F, nf, blockIdx.x);
    const HsFrame f = F[fi];
    const int64_t mn = O[fi].nl < O[fi].nr ? O[fi].nl : O[fi].nr;
    const int64_t j0 = ((int64_t)blockIdx.x - f.chunk0) * kHsChunk;
    const int64_t j1 = j0 + kHsChunk < mn ? j0 + kHsChunk : mn;
    unsigned long long best = ~0ull;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += kHsThreads)
        if (f.lpos[f.first + j] >= f.rpos[f.first + j]) {  // holds for every later j too
            best = (unsigned long long)j;
            break;
        }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(best, o, kWave);
        best = x < best ? x : best;
    }
    if (lane_id() == 0) s[threadIdx.x / kWave] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kHsThreads / kWave; ++q) best = s[q] < best ? s[q] : best;
        if (best != ~0ull) atomicMin(&O[fi].m, best);
    }
}

__global__ __launch_bounds__(kHsThreads) void k_hs_swap(const HsFrame *F, int nf, HsOut *O) {
    const int fi = frame_of(F, nf, blockIdx.x);
    const HsFrame f = F[fi];
    const int64_t m = (int64_t)O[fi].m;
    const int64_t c = (int64_t)blockIdx.x - f.chunk0;
    const int64_t j0 = c * kHsChunk;
    const int64_t j1 = j0 + kHsChunk < m ? j0 + kHsChunk : m;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += kHsThreads) {
        const int64_t l = f.lpos[f.first + j], r = f.rpos[f.first + j];
        const DegPair t = f.a[l];
        f.a[l] = f.a[r];
        f.a[r] = t;
    }
    if (c == 0 && threadIdx.x == 0) {
        const int64_t nl = O[fi].nl;
        int64_t cut = -1;
        if (nl > 0) {
            if (m == 0) {
                cut = f.lpos[f.first];
            } else {
                const int64_t cl = m < nl ? (int64_t)f.lpos[f.first + m] : f.last;
                const int64_t cr = f.rpos[f.first + m - 1];
                cut = cl < cr ? cl : cr;
            }
        }
        O[fi].cut = cut;
    }
}

struct HsRange {
    int64_t first, last;
    int depth;
};

// One job's sequence and ranges.  keep: positions [keep, n) are the top the
// caller wants.
struct HsJob {
    int64_t n = 0;
    const int32_t *deg = nullptr;
    DevMem a, lpos, rpos;
    int64_t keep = 0;
    std::vector<HsRange> dev, host;
    std::vector<std::vector<DegPair>> host_data;
    std::vector<DegPair> top;  // positions n - 1 down to keep
};

// The sequences of every job (jobs with n > 0) and their reductions:
// stats[3 j ..] = (max (deg, ~id), max (deg, id), degree total).
void hs_setup(dmolp_ctx *ctx, std::vector<HsJob> &J, std::vector<unsigned long long> &stats) {
    hipStream_t st = ctx->stream;
    DevMem stat(sizeof(unsigned long long) * 3 * (J.size() ? J.size() : 1));
    DMOLP_HIP_CHECK(hipMemsetAsync(stat.p, 0, stat.bytes, st));
    std::vector<HsInit> init;
    int64_t chunks = 0;
    for (size_t i = 0; i < J.size(); ++i) {
        HsJob &j = J[i];
        if (j.n == 0) continue;
        j.a.alloc(sizeof(DegPair) * (size_t)j.n);
        HsInit h{j.deg, j.a.as<DegPair>(), j.n, chunks, (j.n + kHsChunk - 1) / kHsChunk,
                 stat.as<unsigned long long>() + 3 * i};
        chunks += h.nchunks;
        init.push_back(h);
    }
    DevMem tab;
    if (!init.empty()) {
        tab.alloc(sizeof(HsInit) * init.size());
        DMOLP_HIP_CHECK(hipMemcpyAsync(tab.p, init.data(), tab.bytes, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_hs_init, dim3((unsigned)chunks), dim3(kHsThreads), 0, st, tab.as<HsInit>(),
                           (int)init.size());
        DMOLP_HIP_CHECK(hipGetLastError());
    }
    stats.assign(3 * J.size(), 0);
    if (!J.empty())
        DMOLP_HIP_CHECK(hipMemcpyAsync(stats.data(), stat.p, sizeof(unsigned long long) * stats.size(),
                                       hipMemcpyDeviceToHost, st));
    DMOLP_HIP_CHECK(hipStreamSynchronize(st));
}

// Partition steps from the range [0, n) of every job with keep < n, until
// every range that reaches [keep, n) is shorter than host_max (or a leaf,
// or at its depth limit); then those ranges are copied out and finished on
// the host (Total: sorted outright), and j.top holds positions n - 1 down
// to keep.
template <bool Total>
void hs_run(dmolp_ctx *ctx, std::vector<HsJob> &J, int64_t host_max, int threads) {
    hipStream_t st = ctx->stream;
    const auto place = [&](HsJob &j, const HsRange &r) {
        if (r.last <= j.keep || r.last <= r.first) return;  // no effect on [keep, n)
        if (r.last - r.first > 16 && r.depth > 0 && r.last - r.first >= host_max) j.dev.push_back(r);
        else j.host.push_back(r);
    };
    for (HsJob &j : J) {
        j.dev.clear();
        j.host.clear();
        j.top.clear();
        if (j.n == 0 || j.keep >= j.n) continue;
        place(j, HsRange{0, j.n, depth_limit(j.n)});
        if (!j.dev.empty()) {
            j.lpos.alloc(sizeof(int32_t) * (size_t)j.n);
            j.rpos.alloc(sizeof(int32_t) * (size_t)j.n);
        }
    }
    DevMem ftab, otab, cl, cr, bl, br;
    std::vector<HsFrame> fr;
    std::vector<std::pair<size_t, HsRange>> owner;
    std::vector<HsOut> out;
    for (;;) {
        fr.clear();
        owner.clear();
        int64_t chunks = 0;
        for (size_t i = 0; i < J.size(); ++i) {
            for (const HsRange &r : J[i].dev) {
                const int64_t nc = (r.last - r.first + kHsChunk - 1) / kHsChunk;
                fr.push_back(HsFrame{J[i].a.as<DegPair>(), J[i].lpos.as<int32_t>(), J[i].rpos.as<int32_t>(),
                                     r.first, r.last, chunks, nc});
                owner.push_back({i, r});
                chunks += nc;
            }
            J[i].dev.clear();
        }
        if (fr.empty()) break;
        const int nf = (int)fr.size();
        ftab.grow(sizeof(HsFrame) * fr.size());
        otab.grow(sizeof(HsOut) * fr.size());
        cl.grow(sizeof(int32_t) * (size_t)chunks);
        cr.grow(sizeof(int32_t) * (size_t)chunks);
        bl.grow(sizeof(int64_t) * (size_t)chunks);
        br.grow(sizeof(int64_t) * (size_t)chunks);
        DMOLP_HIP_CHECK(hipMemcpyAsync(ftab.p, fr.data(), sizeof(HsFrame) * fr.size(), hipMemcpyHostToDevice, st));
        const HsFrame *F = ftab.as<HsFrame>();
        HsOut *O = otab.as<HsOut>();
        const dim3 g((unsigned)chunks), b(kHsThreads);
        hipLaunchKernelGGL(k_hs_pivot<Total>, dim3((unsigned)nf), dim3(kWave), 0, st, F, O);
        hipLaunchKernelGGL(k_hs_count<Total>, g, b, 0, st, F, nf, (const HsOut *)O, cl.as<int32_t>(),
                           cr.as<int32_t>());
        hipLaunchKernelGGL(k_hs_scan, dim3((unsigned)nf), b, 0, st, F, (const int32_t *)cl.as<int32_t>(),
                           (const int32_t *)cr.as<int32_t>(), bl.as<int64_t>(), br.as<int64_t>(), O);
        hipLaunchKernelGGL(k_hs_scatter<Total>, g, b, 0, st, F, nf, (const HsOut *)O,
                           (const int32_t *)cr.as<int32_t>(), (const int64_t *)bl.as<int64_t>(),
                           (const int64_t *)br.as<int64_t>());
        hipLaunchKernelGGL(k_hs_find_m, g, b, 0, st, F, nf, O);
        hipLaunchKernelGGL(k_hs_swap, g, b, 0, st, F, nf, O);
        DMOLP_HIP_CHECK(hipGetLastError());
        out.resize(fr.size());
        DMOLP_HIP_CHECK(hipMemcpyAsync(out.data(), O, sizeof(HsOut) * fr.size(), hipMemcpyDeviceToHost, st));
        DMOLP_HIP_CHECK(hipStreamSynchronize(st));
        for (size_t f = 0; f < fr.size(); ++f) {
            HsJob &j = J[owner[f].first];
            const HsRange r = owner[f].second;
            const int64_t cut = out[f].cut;
            DMOLP_REQUIRE(cut >= r.first && cut <= r.last, DMOLP_E_INTERNAL, "hub partition step out of range");
            // __introsort_loop: [cut, last) and [first, cut), both one level deeper.
            place(j, HsRange{cut, r.last, r.depth - 1});
            place(j, HsRange{r.first, cut, r.depth - 1});
        }
    }
    // The short ranges: copied out, finished by the host loop (with its
    // depth limit, heap sort and final insertion sort) or, for the total
    // order, sorted; then read off from the top.
    OpBatch ob;
    std::vector<std::pair<size_t, size_t>> items;
    for (size_t i = 0; i < J.size(); ++i) {
        HsJob &j = J[i];
        j.host_data.resize(j.host.size());
        for (size_t h = 0; h < j.host.size(); ++h) {
            const HsRange &r = j.host[h];
            j.host_data[h].resize((size_t)(r.last - r.first));
            ob.download(j.host_data[h].data(), j.a.as<DegPair>() + r.first,
                        sizeof(DegPair) * (size_t)(r.last - r.first));
            items.push_back({i, h});
        }
    }
    ob.run(ctx);
    std::atomic<size_t> next{0};
    const int T = (int)std::min<size_t>((size_t)std::max(1, threads), items.size());
    run_threads(T, [&](int) {
        for (size_t t = next.fetch_add(1); t < items.size(); t = next.fetch_add(1)) {
            HsJob &j = J[items[t].first];
            const HsRange &r = j.host[items[t].second];
            std::vector<DegPair> &d = j.host_data[items[t].second];
            if (Total) {
                std::sort(d.begin(), d.end(), ByDegreeThenId());
            } else {
                const int64_t keep = j.keep > r.first ? j.keep - r.first : 0;
                introsort_from(d.data(), 0, r.last - r.first, r.depth, keep, 1);
            }
        }
    });
    for (HsJob &j : J) {
        if (j.host.empty()) continue;
        std::vector<size_t> ord(j.host.size());
        for (size_t h = 0; h < ord.size(); ++h) ord[h] = h;
        std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return j.host[x].first > j.host[y].first; });
        j.top.reserve((size_t)(j.n - j.keep));
        int64_t at = j.n;  // the ranges tile [keep, n) from the top down
        for (size_t h : ord) {
            const HsRange &r = j.host[h];
            DMOLP_REQUIRE(r.last == at, DMOLP_E_INTERNAL, "hub ranges do not tile the top");
            const std::vector<DegPair> &d = j.host_data[h];
            for (int64_t p = r.last - 1; p >= r.first && p >= j.keep; --p) j.top.push_back(d[(size_t)(p - r.first)]);
            at = r.first;
        }
        DMOLP_REQUIRE(at <= j.keep && (int64_t)j.top.size() == j.n - j.keep, DMOLP_E_INTERNAL,
                      "hub ranges do not cover the top");
        j.host_data.clear();
    }
}

}  // namespace

void hub_search_device(dmolp_ctx *ctx, const std::vector<HubSearch *> &jobs, int64_t host_max,
                       int threads) {
    std::vector<HsJob> J(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) {
        HubSearch &s = *jobs[i];
        s.hubs.clear();
        s.hub_deg.clear();
        s.n_hubs = 0;
        s.fs = {-1, -1};
        J[i].n = s.n;
        J[i].deg = s.deg;
    }
    std::vector<unsigned long long> q;
    hs_setup(ctx, J, q);
    for (size_t i = 0; i < J.size(); ++i) {
        HsJob &j = J[i];
        HubSearch &s = *jobs[i];
        j.keep = j.n;
        if (j.n == 0) continue;
        const int32_t D = (int32_t)(q[3 * i] >> 32);
        const int64_t vmin = (int64_t)(0xffffffffull - (q[3 * i] & 0xffffffffull));
        const int64_t vmax = (int64_t)(q[3 * i + 1] & 0xffffffffull);
        s.fs = first_seed_from(D, vmin, vmax);
        int64_t M = hub_count(j.n, (int64_t)q[3 * i + 2], D);
        s.n_hubs = M;
        if (s.limit >= 0 && s.limit < M) M = s.limit;
        j.keep = j.n - M;
    }
    hs_run<false>(ctx, J, host_max, threads);
    for (size_t i = 0; i < J.size(); ++i) {
        HubSearch &s = *jobs[i];
        s.hubs.reserve(J[i].top.size());
        s.hub_deg.reserve(J[i].top.size());
        for (const DegPair &e : J[i].top) {
            s.hubs.push_back(e.first);
            s.hub_deg.push_back(e.second);
        }
    }
}

void top_by_degree_device(dmolp_ctx *ctx, const std::vector<TopSearch *> &jobs, int64_t host_max,
                          int threads) {
    std::vector<HsJob> J(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) {
        J[i].n = jobs[i]->n;
        J[i].deg = jobs[i]->deg;
        J[i].keep = J[i].n - std::min<int64_t>(std::max<int64_t>(jobs[i]->m, 0), J[i].n);
    }
    std::vector<unsigned long long> q;
    hs_setup(ctx, J, q);
    hs_run<true>(ctx, J, host_max, threads);
    for (size_t i = 0; i < J.size(); ++i) {
        jobs[i]->top.clear();
        for (const DegPair &e : J[i].top) jobs[i]->top.push_back(e.first);
    }
}

}  // namespace dmolp
