// Seeded synthetic-input generators (host side).
//
// The reference's generators are unseeded networkx scripts (tools/barabasi_albert_graph.py:5-7,
// tools/powerlaw_cluster_graph.py:6-17, tools/powerlaw_configuration.py:6-17) and a
// default-seeded std::default_random_engine (rt_search/measure.cpp:24-58), so none of their
// outputs can be reproduced. These generators draw every number from rng.hpp's counter-based
// stream instead, so each output is a pure function of its parameters and seed.
#include "rtg_internal.hpp"
#include "rng.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

using namespace rtg;

namespace {

// Sorted, duplicate-free sample of `len` values from [0, range) for list `list_id`.
void sample_sorted_unique(uint64_t seed, uint64_t list_id, uint32_t len, uint32_t range,
                          uint32_t *out) {
    std::vector<uint32_t> v;
    v.reserve(len + 16);
    uint64_t ctr = 0;
    const uint64_t stream = kStreamElem ^ (list_id << 24);
    while (v.size() < len) {
        size_t have = v.size();
        for (size_t i = have; i < len; ++i) v.push_back(bounded32(crand(seed, stream, ctr++), range));
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    std::memcpy(out, v.data(), sizeof(uint32_t) * len);
}

// Zipf(s) on k >= 1, clipped to kmax: inverse CDF over an exact table of P(k), k < kmax, with
// the remaining mass (k >= kmax) assigned to kmax.
struct ZipfTable {
    std::vector<double> cdf;  // cdf[k-1] = P(K <= k) for k in [1, kmax-1]
    uint32_t kmax;
    ZipfTable(double s, uint32_t kmax_) : kmax(kmax_) {
        // zeta(s) = sum_{j<K} j^-s + tail, tail by Euler-Maclaurin at K (error << 1e-15).
        const double K = (double)kmax;
        double head = 0.0;
        std::vector<double> terms(kmax);
        for (uint32_t j = 1; j < kmax; ++j) terms[j] = std::pow((double)j, -s);
        for (uint32_t j = kmax - 1; j >= 1; --j) head += terms[j];  // small terms first
        double tail = std::pow(K, 1.0 - s) / (s - 1.0) + 0.5 * std::pow(K, -s) +
                      s * std::pow(K, -s - 1.0) / 12.0;
        double zeta = head + tail;
        cdf.resize(kmax - 1);
        double acc = 0.0;
        for (uint32_t j = 1; j < kmax; ++j) {
            acc += terms[j];
            cdf[j - 1] = acc / zeta;
        }
    }
    uint32_t sample(double u) const {
        auto it = std::upper_bound(cdf.begin(), cdf.end(), u);
        if (it == cdf.end()) return kmax;
        return (uint32_t)(it - cdf.begin()) + 1;
    }
};

}  // namespace

extern "C" {

// Barabasi-Albert preferential attachment, as networkx.barabasi_albert_graph defines the
// process: a star on nodes 0..m, then node s = m+1..n-1 attaches to m distinct targets picked
// uniformly from the degree-weighted endpoint list. Writes m*(n-m) pairs.
RTG_API int rtg_gen_ba(int32_t n, int32_t m, uint64_t seed, int32_t *pairs, int64_t *npairs) {
    if (n < 0 || m < 1 || (n > 0 && n <= m)) return rtg_fail(RTG_EINVAL, "rtg_gen_ba: need n > m >= 1");
    std::vector<int32_t> repeated;
    repeated.reserve((size_t)2 * m * (size_t)(n - m) + 2);
    int64_t e = 0;
    for (int32_t i = 1; i <= m; ++i) {
        pairs[2 * e] = 0;
        pairs[2 * e + 1] = i;
        ++e;
    }
    for (int32_t i = 0; i < m; ++i) repeated.push_back(0);
    for (int32_t i = 1; i <= m; ++i) repeated.push_back(i);
    uint64_t ctr = 0;
    std::vector<int32_t> targets;
    targets.reserve(m);
    for (int32_t s = m + 1; s < n; ++s) {
        targets.clear();
        while ((int32_t)targets.size() < m) {
            uint64_t r = crand(seed, kStreamBA, ctr++);
            int32_t t = repeated[bounded32(r, (uint32_t)repeated.size())];
            if (std::find(targets.begin(), targets.end(), t) == targets.end()) targets.push_back(t);
        }
        for (int32_t t : targets) {
            pairs[2 * e] = s;
            pairs[2 * e + 1] = t;
            ++e;
        }
        for (int32_t t : targets) repeated.push_back(t);
        for (int32_t i = 0; i < m; ++i) repeated.push_back(s);
    }
    *npairs = e;
    return RTG_OK;
}

// R-MAT / Kronecker sampling without noise: sample i descends `scale` levels, picking quadrant
// (0,0),(0,1),(1,0),(1,1) with probabilities a,b,c,1-a-b-c from draw crand(seed, RMAT, i*64+l);
// both endpoints are then relabelled by the keyed bijection permute_bits (Graph500 style).
RTG_API int rtg_gen_rmat(uint32_t scale, int64_t nsamples, double a, double b, double c,
                         uint64_t seed, int32_t *pairs) {
    if (scale < 1 || scale > 30 || nsamples < 0) return rtg_fail(RTG_EINVAL, "rtg_gen_rmat: bad scale");
    const uint64_t t1 = prob53(a), t2 = prob53(a + b), t3 = prob53(a + b + c);
    parallel_for(nsamples, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            uint32_t src = 0, dst = 0;
            for (uint32_t l = 0; l < scale; ++l) {
                uint64_t r = top53(crand(seed, kStreamRmat, (uint64_t)i * 64 + l));
                uint32_t sb = (r >= t2) ? 1u : 0u;
                uint32_t db = (r >= t1 && r < t2) || (r >= t3) ? 1u : 0u;
                src = (src << 1) | sb;
                dst = (dst << 1) | db;
            }
            pairs[2 * i] = (int32_t)permute_bits(src, scale, seed ^ kStreamRmatPerm);
            pairs[2 * i + 1] = (int32_t)permute_bits(dst, scale, seed ^ kStreamRmatPerm);
        }
    });
    return RTG_OK;
}

// Degree sequence of the power-law configuration model: continuous power law with exponent
// gamma on [kmin, kmax] by inverse CDF, floored; kmin is chosen by bisection so that the
// expected degree equals `mean`. The sum is made even by adding one to the last vertex.
RTG_API int rtg_gen_powerlaw_degrees(int32_t n, double gamma, double kmax, double mean,
                                     uint64_t seed, int32_t *deg, int64_t *total) {
    if (n < 2 || gamma <= 1.0 || mean < 1.0 || kmax <= mean) return rtg_fail(RTG_EINVAL, "rtg_gen_powerlaw_degrees: bad parameters");
    const double g1 = 1.0 - gamma;
    auto cont_mean = [&](double kmin) {
        // E[X] for density ~ x^-gamma on [kmin, kmax]
        double r = kmax / kmin;
        if (std::fabs(gamma - 2.0) < 1e-12) return kmin * std::log(r) / (1.0 - 1.0 / r);
        return kmin * (g1 / (2.0 - gamma)) * (std::pow(r, 2.0 - gamma) - 1.0) / (std::pow(r, g1) - 1.0);
    };
    double lo = 1.0, hi = mean;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        // flooring loses ~0.5 on average
        if (cont_mean(mid) - 0.5 < mean) lo = mid; else hi = mid;
    }
    const double kmin = lo;
    const double rr = std::pow(kmax / kmin, g1);
    std::vector<int64_t> part(64, 0);
    parallel_for(n, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            double u = u01(crand(seed, kStreamDegree, (uint64_t)i));
            double x = kmin * std::pow(1.0 - u * (1.0 - rr), 1.0 / g1);
            int64_t d = (int64_t)std::floor(x);
            if (d < 1) d = 1;
            if (d > (int64_t)kmax) d = (int64_t)kmax;
            deg[i] = (int32_t)d;
        }
    });
    int64_t s = 0;
    for (int32_t i = 0; i < n; ++i) s += deg[i];
    if (s & 1) {
        deg[n - 1] += 1;
        s += 1;
    }
    *total = s;
    return RTG_OK;
}

// Configuration-model stub matching plus triangle closing, from a degree sequence.
// Stub s (owned by vertex owner[s], stubs of a vertex contiguous) gets the 64-bit sort key
// (hi32(crand(seed, STUB, s)) << 32) | s; stubs sorted by key are paired (2k, 2k+1) into the
// base edges. Then for each vertex u and each of its stubs s after the first, with probability
// p_close (draw crand(seed, CLOSE, s)) the closing edge (partner(s-1), partner(s)) is added,
// closing the wedge partner(s-1) - u - partner(s). Output: base edges in pair order, then
// closing edges in stub order. Self-loops and multi-edges are left for preprocessing to drop.
// `pairs` must hold 2 * (total/2 + total) ints; *npairs receives the count written.
RTG_API int rtg_gen_config_model(int32_t n, const int32_t *deg, int64_t total, double p_close,
                                 uint64_t seed, int32_t *pairs, int64_t *npairs) {
    if (total & 1) return rtg_fail(RTG_EINVAL, "rtg_gen_config_model: odd stub total");
    std::vector<int32_t> owner(total);
    std::vector<int64_t> start(n + 1, 0);
    for (int32_t i = 0; i < n; ++i) start[i + 1] = start[i] + deg[i];
    if (start[n] != total) return rtg_fail(RTG_EINVAL, "rtg_gen_config_model: degree sum != total");
    parallel_for(n, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i)
            for (int64_t s = start[i]; s < start[i + 1]; ++s) owner[s] = (int32_t)i;
    });
    std::vector<uint64_t> key(total);
    parallel_for(total, [&](int64_t b, int64_t e) {
        for (int64_t s = b; s < e; ++s)
            key[s] = (crand(seed, kStreamStub, (uint64_t)s) & 0xFFFFFFFF00000000ull) | (uint64_t)s;
    });
    std::sort(key.begin(), key.end());
    std::vector<int32_t> partner(total);
    const int64_t half = total / 2;
    for (int64_t k = 0; k < half; ++k) {
        int64_t s0 = (int64_t)(key[2 * k] & 0xFFFFFFFFull), s1 = (int64_t)(key[2 * k + 1] & 0xFFFFFFFFull);
        pairs[2 * k] = owner[s0];
        pairs[2 * k + 1] = owner[s1];
        partner[s0] = owner[s1];
        partner[s1] = owner[s0];
    }
    const uint64_t tp = prob53(p_close);
    int64_t e = half;
    for (int32_t u = 0; u < n; ++u) {
        for (int64_t s = start[u] + 1; s < start[u + 1]; ++s) {
            if (top53(crand(seed, kStreamClose, (uint64_t)s)) < tp) {
                pairs[2 * e] = partner[s - 1];
                pairs[2 * e + 1] = partner[s];
                ++e;
            }
        }
    }
    *npairs = e;
    return RTG_OK;
}

// Lengths of `nlists` lists, each min(Zipf(s), kmax) with k >= 1 (draw crand(seed, ZIPF, i)).
RTG_API int rtg_gen_zipf_lengths(int64_t nlists, double s, uint32_t kmax, uint64_t seed,
                                 uint32_t *len) {
    if (s <= 1.0 || kmax < 1) return rtg_fail(RTG_EINVAL, "rtg_gen_zipf_lengths: need s > 1");
    if (kmax == 1) {
        for (int64_t i = 0; i < nlists; ++i) len[i] = 1;
        return RTG_OK;
    }
    ZipfTable table(s, kmax);
    parallel_for(nlists, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) len[i] = table.sample(u01(crand(seed, kStreamZipf, (uint64_t)i)));
    });
    return RTG_OK;
}

// Fills list i (elements[offsets[i] .. offsets[i+1])) with sorted distinct uniform values in
// [0, range), independently per list.
RTG_API int rtg_gen_sorted_lists(int64_t nlists, const uint32_t *offsets, uint32_t range,
                                 uint64_t seed, uint32_t *elements) {
    for (int64_t i = 0; i < nlists; ++i)
        if (offsets[i + 1] - offsets[i] > range) return rtg_fail(RTG_EINVAL, "rtg_gen_sorted_lists: list longer than value range");
    parallel_for(nlists, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i)
            sample_sorted_unique(seed, (uint64_t)i, offsets[i + 1] - offsets[i], range, elements + offsets[i]);
    }, 64);
    return RTG_OK;
}

// Uniform uint32 values (duplicates allowed), ascending: value i = hi32(crand(seed, ARRAY, i)),
// then sorted.
RTG_API int rtg_gen_sorted_array(int64_t n, uint64_t seed, uint32_t *out) {
    parallel_for(n, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) out[i] = (uint32_t)(crand(seed, kStreamArray, (uint64_t)i) >> 32);
    });
    std::sort(out, out + n);
    return RTG_OK;
}

// Uniform uint32 keys: key i = hi32(crand(seed, stream, i)).
RTG_API int rtg_gen_uniform_u32(int64_t n, uint64_t seed, uint64_t stream, uint32_t *out) {
    parallel_for(n, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) out[i] = (uint32_t)(crand(seed, stream, (uint64_t)i) >> 32);
    });
    return RTG_OK;
}

}  // extern "C"
