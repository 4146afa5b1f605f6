// ORACLE — test infrastructure only. A CPU (C++17/OpenMP) restatement of the reference's
// triangle-counting, set-intersection and sorted-lookup path, used by tests/ as the checker,
// by __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg as the timed
// CPU baseline. Nothing in the product links or calls this library.
//
// The reference path cannot be compiled here (it is CUDA + Thrust and there is no nvcc), so each
// function below restates one reference function and cites the file:line it follows.
//
// Pinning (tests/test_oracle.py): closed forms (K_n, wheels, friendship graphs, trees,
// complete bipartite), brute force O(n^3) on small graphs, scipy trace(A^3)/6 and networkx
// triangle counts on seeded random graphs, agreement of the three counting formulations
// (1A2, 2A1, per-edge intersection), invariance of the count under id permutation and pair
// reversal, the set-intersection generator invariant of the reference
// (set_intersection/utils/util.cpp:39-51 with data_generator/data.cpp:45-58), and numpy
// searchsorted for the first-occurrence lookup.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include <omp.h>
#include <parallel/algorithm>

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

// BinaryCheck, tc/bs_tc/bs_tc.cu:14-25: is val in data[l..r] (inclusive)?
inline bool binary_check(const int32_t *data, int64_t l, int64_t r, int32_t val) {
    while (l <= r) {
        int64_t mid = (l + r) >> 1;
        if (data[mid] == val) return true;
        if (data[mid] < val) l = mid + 1;
        else r = mid - 1;
    }
    return false;
}

inline uint64_t merge_count(const int32_t *a, int64_t na, const int32_t *b, int64_t nb) {
    uint64_t c = 0;
    int64_t i = 0, j = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (a[i] > b[j]) ++j;
        else { ++c; ++i; ++j; }
    }
    return c;
}

inline void set_threads(int threads) {
    if (threads > 0) omp_set_num_threads(threads);
}

}  // namespace

ORC_API int orc_version() { return 1; }

ORC_API int orc_max_threads() { return omp_get_max_threads(); }

// Graph::ReadBinaryEdges relabel, tc/graph/graph.cpp:40-62, with its intended semantics (the
// reference's memset at :48 has swapped arguments and leaves `befores` uninitialised): ids
// start at 0 if any endpoint is 0, else at 1; the ids present are renumbered densely in
// increasing order. Relabels `pairs` in place and returns the number of vertices.
ORC_API int32_t orc_relabel(int32_t *pairs, int64_t npairs) {
    int32_t max_id = 0;
    int start_index = 1;
    for (int64_t i = 0; i < 2 * npairs; ++i) {
        if (pairs[i] == 0) start_index = 0;
        max_id = std::max(max_id, pairs[i]);
    }
    int64_t nodes = (start_index == 0) ? (int64_t)max_id + 1 : max_id;
    std::vector<int32_t> befores((size_t)(nodes + start_index), 0);
    for (int64_t i = 0; i < 2 * npairs; ++i) befores[pairs[i]] = 1;
    for (int64_t i = start_index + 1; i < start_index + nodes; ++i) befores[i] += befores[i - 1];
    for (int64_t i = 0; i < 2 * npairs; ++i) pairs[i] = befores[pairs[i]] - 1;
    if (npairs == 0) return 0;
    return befores[nodes + start_index - 1];
}

// Graph::TCPreprocessing + Graph::ConvertToCSR, tc/graph/graph.cpp:104-177 and :71-102, on
// pairs already relabelled to [0, nodes):
//   raw degree counts every pair endpoint, duplicates and self-loops included (:109-114);
//   orient src->dst iff deg[src] < deg[dst] or (equal and src < dst) (:116-124);
//   vertices with raw degree <= 1 are removed and the rest renumbered x - removes[x], edges
//   touching a removed vertex dropped (:126-152);
//   the remaining pairs are sorted by (src, dst) (ThrustSortEdgesInBatch,
//   tc/graph/thrust_helper.cu:16-42 — ten src buckets each sorted, i.e. a global sort);
//   self-loops and repeats of the previous pair are dropped (:159-173);
//   CSR rows are the surviving pairs in order, offsets the exclusive scan of row counts.
// Writes offsets[n+1] and adjs[m] (adjs must hold npairs ints); returns m, *n_out = n.
ORC_API int64_t orc_tc_preprocess(const int32_t *pairs, int64_t npairs, int32_t nodes,
                                  int32_t *offsets, int32_t *adjs, int32_t *n_out) {
    std::vector<int32_t> degree((size_t)nodes, 0);
    for (int64_t i = 0; i < npairs; ++i) {
        degree[pairs[2 * i]] += 1;
        degree[pairs[2 * i + 1]] += 1;
    }
    std::vector<int32_t> removes((size_t)nodes, 0);
    if (nodes > 0) {
        removes[0] = degree[0] <= 1;
        for (int32_t i = 1; i < nodes; ++i) removes[i] = removes[i - 1] + (degree[i] <= 1);
    }
    const int32_t n = nodes > 0 ? nodes - removes[nodes - 1] : 0;
    std::vector<uint64_t> keys;
    keys.reserve((size_t)npairs);
    for (int64_t i = 0; i < npairs; ++i) {
        int32_t s = pairs[2 * i], d = pairs[2 * i + 1];
        if (degree[s] > degree[d] || (degree[s] == degree[d] && s > d)) std::swap(s, d);
        if (degree[s] <= 1 || degree[d] <= 1) continue;
        s -= removes[s];
        d -= removes[d];
        keys.push_back(((uint64_t)(uint32_t)s << 32) | (uint32_t)d);
    }
    __gnu_parallel::sort(keys.begin(), keys.end());
    std::vector<int64_t> count((size_t)n + 1, 0);
    int64_t m = 0;
    uint64_t last = ~0ull;
    for (size_t i = 0; i < keys.size(); ++i) {
        uint64_t k = keys[i];
        int32_t s = (int32_t)(k >> 32), d = (int32_t)(k & 0xFFFFFFFFu);
        bool drop = (s == d) || (k == last);
        last = k;
        if (drop) continue;
        adjs[m++] = d;
        count[s] += 1;
    }
    int64_t acc = 0;
    for (int32_t i = 0; i < n; ++i) {
        offsets[i] = (int32_t)acc;
        acc += count[i];
    }
    offsets[n] = (int32_t)acc;
    *n_out = n;
    return m;
}

// Symmetric simple CSR (both directions, rows sorted, duplicates and self-loops dropped) on
// relabelled pairs: the adjacency the serving queries |N(u) & N(v)| run on. No vertex is
// removed. Same I/O convention as orc_tc_preprocess (adjs must hold 2*npairs ints).
ORC_API int64_t orc_symmetric_csr(const int32_t *pairs, int64_t npairs, int32_t nodes,
                                  int32_t *offsets, int32_t *adjs) {
    std::vector<uint64_t> keys;
    keys.reserve((size_t)(2 * npairs));
    for (int64_t i = 0; i < npairs; ++i) {
        uint32_t s = (uint32_t)pairs[2 * i], d = (uint32_t)pairs[2 * i + 1];
        if (s == d) continue;
        keys.push_back(((uint64_t)s << 32) | d);
        keys.push_back(((uint64_t)d << 32) | s);
    }
    __gnu_parallel::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<int64_t> count((size_t)nodes + 1, 0);
    for (size_t i = 0; i < keys.size(); ++i) {
        adjs[i] = (int32_t)(keys[i] & 0xFFFFFFFFu);
        count[keys[i] >> 32] += 1;
    }
    int64_t acc = 0;
    for (int32_t i = 0; i < nodes; ++i) {
        offsets[i] = (int32_t)acc;
        acc += count[i];
    }
    offsets[nodes] = (int32_t)acc;
    return (int64_t)keys.size();
}

namespace {
// Row of every edge index in [e_begin, e_end), stored at row[e - e_begin] (the reference finds
// it by binary search over offsets per edge, bs_tc.cu:126-134; the result is the same row).
std::vector<int32_t> edge_rows(const int32_t *offsets, int32_t n, int64_t e_begin = 0, int64_t e_end = -1) {
    if (e_end < 0) e_end = offsets[n];
    std::vector<int32_t> row((size_t)std::max<int64_t>(e_end - e_begin, 0));
    if (e_end <= e_begin) return row;
    // first row whose range contains e_begin / last containing e_end - 1
    const int32_t u0 = (int32_t)(std::upper_bound(offsets, offsets + n + 1, (int32_t)e_begin) - offsets) - 1;
    const int32_t u1 = (int32_t)(std::upper_bound(offsets, offsets + n + 1, (int32_t)(e_end - 1)) - offsets) - 1;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t u = u0; u <= u1; ++u) {
        const int64_t b = std::max<int64_t>(offsets[u], e_begin), e = std::min<int64_t>(offsets[u + 1], e_end);
        for (int64_t k = b; k < e; ++k) row[k - e_begin] = u;
    }
    return row;
}
}  // namespace

// ListIntersectionKernel (1A2), tc/bs_tc/bs_tc.cu:120-148: for every oriented edge e=(u->v)
// in [e_begin, e_end), for every w in N+(u) with d+(w) > 0, count BinaryCheck(N+(w), v).
// The reference accumulates into a 32-bit uint (bs_tc.cu:145, 240-242); this sums in uint64,
// whose low 32 bits are the reference's printed value. OpenMP schedule(dynamic, 4096).
ORC_API uint64_t orc_tc_count_1a2(const int32_t *offsets, const int32_t *adjs, int32_t n,
                                  int64_t e_begin, int64_t e_end, int threads) {
    set_threads(threads);
    const int64_t m = offsets[n];
    if (e_end <= 0 || e_end > m) e_end = m;
    if (e_begin < 0) e_begin = 0;
    if (e_begin >= e_end) return 0;
    std::vector<int32_t> row = edge_rows(offsets, n, e_begin, e_end);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : total)
    for (int64_t e = e_begin; e < e_end; ++e) {
        const int32_t u = row[e - e_begin], v = adjs[e];
        uint64_t cnt = 0;
        for (int64_t i = offsets[u]; i < offsets[u + 1]; ++i) {
            const int32_t w = adjs[i];
            const int64_t len = (int64_t)offsets[w + 1] - offsets[w];
            if (len == 0) continue;
            cnt += binary_check(adjs + offsets[w], 0, len - 1, v);
        }
        total += cnt;
    }
    return total;
}

// The per-edge intersection form: sum over oriented edges (u->v) of |N+(u) & N+(v)| by a
// linear merge. Same total as 1A2 (triangle a<b<c counted at edge a->b with common c).
ORC_API uint64_t orc_tc_count_intersect(const int32_t *offsets, const int32_t *adjs, int32_t n,
                                        int64_t e_begin, int64_t e_end, int threads) {
    set_threads(threads);
    const int64_t m = offsets[n];
    if (e_end <= 0 || e_end > m) e_end = m;
    if (e_begin < 0) e_begin = 0;
    if (e_begin >= e_end) return 0;
    std::vector<int32_t> row = edge_rows(offsets, n, e_begin, e_end);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : total)
    for (int64_t e = e_begin; e < e_end; ++e) {
        const int32_t u = row[e - e_begin], v = adjs[e];
        total += merge_count(adjs + offsets[u], offsets[u + 1] - offsets[u], adjs + offsets[v],
                             offsets[v + 1] - offsets[v]);
    }
    return total;
}

// Per-edge counts. form 0: |N+(u) & N+(v)| at edge u->v. form 1: the 1A2 attribution of
// ListIntersectionKernel, |{w in N+(u) : v in N+(w)}| at edge u->v (bs_tc.cu:135-145).
ORC_API void orc_tc_per_edge(const int32_t *offsets, const int32_t *adjs, int32_t n, int form,
                             uint32_t *out) {
    const int64_t m = offsets[n];
    std::vector<int32_t> row = edge_rows(offsets, n);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t e = 0; e < m; ++e) {
        const int32_t u = row[e], v = adjs[e];
        if (form == 0) {
            out[e] = (uint32_t)merge_count(adjs + offsets[u], offsets[u + 1] - offsets[u],
                                           adjs + offsets[v], offsets[v + 1] - offsets[v]);
        } else {
            uint32_t cnt = 0;
            for (int64_t i = offsets[u]; i < offsets[u + 1]; ++i) {
                const int32_t w = adjs[i];
                const int64_t len = (int64_t)offsets[w + 1] - offsets[w];
                if (len) cnt += binary_check(adjs + offsets[w], 0, len - 1, v);
            }
            out[e] = cnt;
        }
    }
}

// Preprocessing + HashmapKernel (2A1), tc/bs_tc/bs_tc.cu:153-229: for every u, the multiset
// of N+(w) over w in N+(u) is sorted and run-length encoded into (u, x, mult) triples
// (:172-205); the kernel adds mult for each triple with x in N+(u) (:213-229). The
// reference's outer loop is serial (:168); this one runs vertices in parallel.
ORC_API uint64_t orc_tc_count_2a1(const int32_t *offsets, const int32_t *adjs, int32_t n,
                                  int threads) {
    set_threads(threads);
    uint64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
        std::vector<int32_t> two_hop;
#pragma omp for schedule(dynamic, 256)
        for (int32_t u = 0; u < n; ++u) {
            two_hop.clear();
            for (int64_t j = offsets[u]; j < offsets[u + 1]; ++j) {
                const int32_t w = adjs[j];
                two_hop.insert(two_hop.end(), adjs + offsets[w], adjs + offsets[w + 1]);
            }
            if (two_hop.empty()) continue;
            std::sort(two_hop.begin(), two_hop.end());
            const int64_t du = (int64_t)offsets[u + 1] - offsets[u];
            size_t i = 0;
            while (i < two_hop.size()) {
                size_t k = i;
                while (k < two_hop.size() && two_hop[k] == two_hop[i]) ++k;
                const uint64_t mult = k - i;
                if (du > 0 && binary_check(adjs + offsets[u], 0, du - 1, two_hop[i])) total += mult;
                i = k;
            }
        }
    }
    return total;
}

// Wedge statistics of an oriented CSR: sum over edges of d+(v), of d+(u), and the algorithmic
// byte model sum over edges of 24 + 4*min(s+l, s*(1+ceil(log2(l+1)))), s/l = min/max of
// d+(u), d+(v). out[0..3] = {wedges_v, wedges_u, alg_bytes, merge_bytes}.
ORC_API void orc_tc_edge_model(const int32_t *offsets, const int32_t *adjs, int32_t n,
                               uint64_t *out) {
    const int64_t m = offsets[n];
    std::vector<int32_t> row = edge_rows(offsets, n);
    uint64_t wv = 0, wu = 0, alg = 0, mrg = 0;
#pragma omp parallel for schedule(static) reduction(+ : wv, wu, alg, mrg)
    for (int64_t e = 0; e < m; ++e) {
        const int32_t u = row[e], v = adjs[e];
        const uint64_t du = offsets[u + 1] - offsets[u], dv = offsets[v + 1] - offsets[v];
        const uint64_t s = std::min(du, dv), l = std::max(du, dv);
        uint64_t lg = 0;
        while ((1ull << lg) < l + 1) ++lg;  // ceil(log2(l+1))
        wv += dv;
        wu += du;
        alg += 24 + 4 * std::min(s + l, s * (1 + lg));
        mrg += 24 + 4 * (du + dv);
    }
    out[0] = wv;
    out[1] = wu;
    out[2] = alg;
    out[3] = mrg;
}

// Pairwise intersection sizes (std::set_intersection semantics on sorted unique lists).
ORC_API void orc_si_pairs(const uint32_t *offsets, const uint32_t *elements,
                          const uint32_t *pairs, int64_t npairs, uint32_t *out, int threads) {
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t k = 0; k < npairs; ++k) {
        const uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        const uint32_t *pa = elements + offsets[a], *pb = elements + offsets[b];
        const int64_t na = offsets[a + 1] - offsets[a], nb = offsets[b + 1] - offsets[b];
        uint32_t c = 0;
        int64_t i = 0, j = 0;
        while (i < na && j < nb) {
            if (pa[i] < pb[j]) ++i;
            else if (pa[i] > pb[j]) ++j;
            else { ++c; ++i; ++j; }
        }
        out[k] = c;
    }
}

// BinarySearchBasedIntersection / BlockBasedOBS, set_intersection/baseline/bs.cu:91-135 and
// :138-209: counts[(a)*num_b + b] = number of elements of B set b found in A set a, for A sets
// [0, num_a) and B sets [num_a, num_a+num_b). Intended semantics: the reference's phase-2
// range may read one element past set a (bs.cu:191, right = (right+1)*|A|/blockDim), which
// only matters when that element equals a target; the intended count is |A_a & B_b|, the
// value its Check (util.cpp:39-51) expects.
ORC_API void orc_si_tile(const uint32_t *set_offsets, const uint32_t *elements, uint32_t num_a,
                         uint32_t num_b, uint32_t *counts, int threads) {
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int64_t a = 0; a < num_a; ++a) {
        for (int64_t b = 0; b < num_b; ++b) {
            const uint32_t *pa = elements + set_offsets[a];
            const int64_t na = set_offsets[a + 1] - set_offsets[a];
            const uint32_t *pb = elements + set_offsets[num_a + b];
            const int64_t nb = set_offsets[num_a + b + 1] - set_offsets[num_a + b];
            uint32_t c = 0;
            for (int64_t j = 0; j < nb; ++j) {
                if (std::binary_search(pa, pa + na, pb[j])) ++c;
            }
            counts[a * num_b + b] = c;
        }
    }
}

// BSAccessCount, rt_search/simple_bs.cu:35-53: l = lower_bound clamped to n-1 (the loop
// starts at r = n-1); result l if data[l] == key else -1. Equivalent to the index of the
// first element equal to the key, or -1. n == 0 gives -1 (the reference would read data[0]).
template <class T>
static void lookup(const T *data, int64_t n, const T *keys, int64_t q, int32_t *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < q; ++i) {
        if (n == 0) { out[i] = -1; continue; }
        int64_t l = 0, r = n - 1;
        const T t = keys[i];
        while (l < r) {
            int64_t mid = l + (r - l) / 2;
            if (data[mid] < t) l = mid + 1;
            else r = mid;
        }
        out[i] = (data[l] == t) ? (int32_t)l : -1;
    }
}

ORC_API void orc_lookup_u32(const uint32_t *data, int64_t n, const uint32_t *keys, int64_t q,
                            int32_t *out) {
    lookup<uint32_t>(data, n, keys, q, out);
}

ORC_API void orc_lookup_i32(const int32_t *data, int64_t n, const int32_t *keys, int64_t q,
                            int32_t *out) {
    lookup<int32_t>(data, n, keys, q, out);
}
