// Sorted-set intersection sizes for lists of sorted, duplicate-free uint32 values.
//
// Pair k intersects list pairs[2k] with list pairs[2k+1] (pairs mode) or, in the reference's
// A x B tile mode (set_intersection/baseline/bs.cu:91-135), A set k / num_b with B set
// num_a + k % num_b. With s = min and l = max of the two lengths:
//
//   SI_PLAIN      one wave per pair, lanes binary-search the shorter list in the longer.
//   SI_REFERENCE  one block per pair with an LDS sample of the longer list (element
//                 t*|A|/blockDim, the BlockBasedOBS layout, bs.cu:148-152): each probe is a
//                 binary search over the samples in LDS, then over one sample interval in
//                 HBM. The interval is clamped to the list (the reference's upper bound can
//                 reach one element past the set, bs.cu:191).
//   SI_BATCHED    a classify pass gives every pair whose shorter list is longer than 64
//                 ceil(s/64) items of 64 of its elements (item counts scanned into offsets).
//                 One launch then runs the items (a wave per item, a lane per element, one
//                 search each, partial counts added into the result) and flat groups of 64
//                 consecutive pairs, whose shorter lists of 0..64 elements are concatenated
//                 over the lanes so that every lane does one search per round whatever the
//                 pairs' lengths. Every search in the launch is one search deep (a four-way
//                 lifting search of the longer list): neither a long pair nor a run of
//                 length-1 pairs serialises a chain of searches or leaves lanes idle. Work is assigned statically; no
//                 counter is shared between waves. Unit offsets are 64-bit, so any tile fits.
//   SI_MERGE, SI_HASH, SI_BITMAP
//                 the same launch, with the long pairs where merging is cheaper than searching
//                 run as merge-path segments, LDS hash chunks or LDS bitmap windows (see
//                 kLongMerge and its siblings below).
//   SI_AUTO       times the four batched strategies and SI_REFERENCE on the input once and
//                 keeps the fastest.
#include "device.hpp"
#include "si.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace {

constexpr int kThreads = 256;
constexpr int kSamples = 1024;

struct PairSrc {
    const uint32_t *pairs;  // null: tile mode
    uint32_t num_a, num_b;
    __device__ __forceinline__ void get(int64_t k, uint32_t &a, uint32_t &b) const {
        if (pairs) {
            const uint2 p = reinterpret_cast<const uint2 *>(pairs)[k];
            a = p.x;
            b = p.y;
        } else {
            a = (uint32_t)(k / num_b);
            b = num_a + (uint32_t)(k % num_b);
        }
    }
};

// First index in data[0..len) with data[i] >= v.
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *__restrict__ data, uint32_t len, uint32_t v) {
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (data[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t contains(const uint32_t *__restrict__ data, uint32_t len, uint32_t v) {
    const uint32_t p = lower_bound_u32(data, len, v);
    return (p < len && data[p] == v) ? 1u : 0u;
}

struct Lists {
    const uint32_t *off;
    const uint32_t *el;
    uint32_t n;     // lists in off; an id at or above n reads as an empty list
    uint32_t *bad;  // set to 1 when a pair names such an id (null: not reported)
    // Start (in el) and length of list id.
    __device__ __forceinline__ void range(uint32_t id, uint32_t &b, uint32_t &len) const {
        if (id >= n) {
            if (bad) *bad = 1u;
            b = 0;
            len = 0;
            return;
        }
        b = off[id];
        len = off[id + 1] - b;
    }
    __device__ __forceinline__ void span(uint32_t id, const uint32_t *&p, uint32_t &len) const {
        uint32_t b;
        range(id, b, len);
        p = el + b;
    }
};

// Orders (a, b) so that `s` is the shorter list.
__device__ __forceinline__ void short_long(const Lists &L, uint32_t a, uint32_t b, const uint32_t *&s, uint32_t &ns,
                                           const uint32_t *&l, uint32_t &nl) {
    const uint32_t *pa, *pb;
    uint32_t la, lb;
    L.span(a, pa, la);
    L.span(b, pb, lb);
    if (la <= lb) { s = pa; ns = la; l = pb; nl = lb; }
    else { s = pb; ns = lb; l = pa; nl = la; }
}

__device__ __forceinline__ uint32_t wave_pair(const Lists &L, uint32_t a, uint32_t b) {
    const uint32_t *s, *l;
    uint32_t ns, nl;
    short_long(L, a, b, s, ns, l, nl);
    uint32_t c = 0;
    for (uint32_t i = __lane_id(); i < ns; i += 64) c += contains(l, nl, s[i]);
    return rtgd::wave_sum_u32(c);
}

__device__ __forceinline__ uint32_t thread_pair(const Lists &L, uint32_t a, uint32_t b) {
    const uint32_t *s, *l;
    uint32_t ns, nl;
    short_long(L, a, b, s, ns, l, nl);
    uint32_t c = 0;
    // Successive elements of the short list only move the lower bound forward.
    uint32_t base = 0;
    for (uint32_t i = 0; i < ns && base < nl; ++i) {
        const uint32_t v = s[i];
        base += lower_bound_u32(l + base, nl - base, v);
        if (base < nl && l[base] == v) {
            ++c;
            ++base;
        }
    }
    return c;
}

// Block per pair: samples of the longer list in LDS, then one interval in HBM.
__device__ __forceinline__ uint32_t block_pair(const Lists &L, uint32_t a, uint32_t b, uint32_t *samp,
                                               uint32_t *red) {
    const uint32_t *s, *l;
    uint32_t ns, nl;
    short_long(L, a, b, s, ns, l, nl);
    const uint32_t ks = nl < (uint32_t)kSamples ? nl : (uint32_t)kSamples;
    for (uint32_t k = threadIdx.x; k < ks; k += blockDim.x) samp[k] = l[((uint64_t)k * nl) / ks];
    __syncthreads();
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
        const uint32_t v = s[i];
        // number of samples <= v, minus one: the interval that can hold v
        uint32_t lo = 0, hi = ks;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (samp[mid] <= v) lo = mid + 1;
            else hi = mid;
        }
        if (lo == 0) continue;  // v < l[0]
        const uint32_t k = lo - 1;
        const uint32_t b0 = (uint32_t)(((uint64_t)k * nl) / ks);
        const uint32_t b1 = (k + 1 < ks) ? (uint32_t)(((uint64_t)(k + 1) * nl) / ks) : nl;
        c += contains(l + b0, b1 - b0, v);
    }
    c = rtgd::wave_sum_u32(c);
    if (__lane_id() == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    uint32_t tot = 0;
    if (threadIdx.x == 0)
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) tot += red[w];
    __syncthreads();  // samp and red are reused by the next pair
    return tot;
}

__global__ __launch_bounds__(kThreads) void si_plain(Lists L, PairSrc P, int64_t npairs, uint32_t *__restrict__ out) {
    const int64_t nw = (int64_t)gridDim.x * (kThreads / 64);
    for (int64_t k = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); k < npairs; k += nw) {
        uint32_t a, b;
        P.get(k, a, b);
        const uint32_t c = wave_pair(L, a, b);
        if (__lane_id() == 0) out[k] = c;
    }
}

__global__ __launch_bounds__(kThreads) void si_reference(Lists L, PairSrc P, int64_t npairs, uint32_t *__restrict__ out) {
    __shared__ uint32_t samp[kSamples];
    __shared__ uint32_t red[kThreads / 64];
    for (int64_t k = blockIdx.x; k < npairs; k += gridDim.x) {
        uint32_t a, b;
        P.get(k, a, b);
        const uint32_t c = block_pair(L, a, b, samp, red);
        if (threadIdx.x == 0) out[k] = c;
    }
}

// ---- batched ----
constexpr uint32_t kFlatMax = 64;  // longest shorter-list handled in flat groups

// Strategies for the pairs whose shorter list is longer than kFlatMax. kLongSearch splits the
// shorter list into 64-element items searched in the longer list. The others serve "balanced"
// pairs, where a merge reads fewer bytes than the searches (s + l <= s * (1 + ceil(log2(l + 1)))
// with s, l the shorter and longer length); skewed pairs keep the search items:
//   kLongMerge   merge path (set_intersection/baseline/ip.cu:138-287): the merge of the two
//                lists is cut into segments of kMergeSpan steps; a wave stages the segment's
//                windows of both lists in LDS and each lane walks kMergeV steps of it.
//   kLongHash    the hash scheme (hi.cu:134-195): chunks of kHashChunk elements of the shorter
//                list go into an LDS hash table; the longer list's elements in the chunk's value
//                range probe it.
//   kLongBitmap  the bitmap scheme (bi_dynamic.cu:134-169) over windows of kBitmapSpan values
//                kept in LDS, for pairs whose shorter list is dense enough (at least kBitmapFill
//                elements per window on average); sparser balanced pairs take merge segments.
enum : int { kLongSearch = 0, kLongMerge = 1, kLongHash = 2, kLongBitmap = 3, kLongKinds = 4 };
constexpr uint32_t kMergeV = 16;
constexpr uint32_t kMergeSpan = 64 * kMergeV;
constexpr uint32_t kHashChunk = 1024;
constexpr uint32_t kHashSlots = 2 * kHashChunk;  // load factor at most 1/2
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kBitmapWords = 2048;
constexpr uint32_t kBitmapSpan = 32 * kBitmapWords;
constexpr uint32_t kBitmapFill = 512;

// The unit kind of a pair whose shorter list s (length ns > kFlatMax) meets a longer list of
// length nl, and its number of units.
template <int kLong>
__device__ __forceinline__ int long_kind(const uint32_t *__restrict__ s, uint32_t ns, uint32_t nl, uint64_t &units) {
    units = (ns + 63u) / 64u;
    if (kLong == kLongSearch) return kLongSearch;
    const uint32_t bits = 32 - __builtin_clz(nl);  // nl >= ns > 0
    if ((uint64_t)ns + nl > (uint64_t)ns * (1 + bits)) return kLongSearch;
    if (kLong == kLongHash) {
        units = (ns + kHashChunk - 1) / kHashChunk;
        return kLongHash;
    }
    if (kLong == kLongBitmap) {
        const uint64_t windows = (uint64_t)((s[ns - 1] - s[0]) / kBitmapSpan) + 1;
        if (windows * kBitmapFill <= ns) {
            units = windows;
            return kLongBitmap;
        }
    }
    units = ((uint64_t)ns + nl + kMergeSpan - 1) / kMergeSpan;
    return kLongMerge;
}

// Per pair: its unit count when its shorter list is longer than kFlatMax (those pairs' results
// are cleared here and accumulated by the units), else 0.
template <int kLong>
__global__ __launch_bounds__(kThreads) void si_classify(Lists L, PairSrc P, int64_t npairs, uint64_t *__restrict__ units,
                                                        uint32_t *__restrict__ out) {
    for (int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x; k < npairs; k += (int64_t)gridDim.x * kThreads) {
        uint32_t a, b;
        P.get(k, a, b);
        const uint32_t *s, *l;
        uint32_t ns, nl;
        short_long(L, a, b, s, ns, l, nl);
        uint64_t u = 0;
        if (ns > kFlatMax) {
            (void)long_kind<kLong>(s, ns, nl, u);
            out[k] = 0u;
        }
        units[k] = u;
    }
}

// Number of elements of l[0..nl) below v, nl >= 1, by K-way lifting (K = 4 or 8): from the
// largest power of K not above nl down to 1, each level issues K - 1 independent probes and
// advances by the number of them below v. The trip count depends only on nl, every probe
// address is clamped into the list, and the chain of dependent loads is log_K(nl) long
// instead of binary lifting's log_2(nl), for (K - 1) / log2(K) times the loads. Four-way suits
// the throughput launches (more loads compete for the memory system); eight-way the latency
// path, whose launches leave the memory system idle.
template <uint32_t K>
__device__ __forceinline__ uint32_t count_below_kway(const uint32_t *__restrict__ l, uint32_t nl, uint32_t v) {
    static_assert(K == 4 || K == 8, "K-way lifting with K = 4 or 8");
    constexpr uint32_t kBits = K == 4 ? 2 : 3;
    const uint32_t lg = 31 - __builtin_clz(nl);
    uint32_t base = 0;
    for (uint32_t step = 1u << (lg - lg % kBits); step > 0; step >>= kBits) {  // from the largest power of K <= nl
        const uint32_t rem = nl - base;
        uint32_t x[K - 1];
        bool in[K - 1];
#pragma unroll
        for (uint32_t j = 0; j < K - 1; ++j) {
            // (j + 1) * step fits 32 bits for K = 4 (3 * 2^30); for K = 8 only when it is <= rem
            in[j] = K == 4 ? (j + 1) * step <= rem : (uint64_t)(j + 1) * step <= rem;
            x[j] = l[base + (in[j] ? (j + 1) * step : rem) - 1];
        }
        uint32_t add = 0;
#pragma unroll
        for (uint32_t j = 0; j < K - 1; ++j) add += (in[j] && x[j] < v) ? step : 0u;
        base += add;
    }
    return base;
}

template <uint32_t K>
__device__ __forceinline__ uint32_t member(const uint32_t *__restrict__ l, uint32_t nl, uint32_t v) {
    const uint32_t b = count_below_kway<K>(l, nl, v);
    return (b < nl && l[b] == v) ? 1u : 0u;
}

// Wave-cooperative search for the first i in [lo, hi) where pred(i) is false (hi if none),
// pred being true on a prefix: each round tests 64 evenly spaced candidates, one per lane, and
// keeps the gap between the last true and the first false one. The range shrinks 64-fold per
// round, so a search over n positions takes ceil(log64 n) dependent loads instead of log2 n.
template <class Pred>
__device__ __forceinline__ uint64_t wave_partition_point(uint64_t lo, uint64_t hi, Pred pred) {
    const uint64_t lane = (uint64_t)__lane_id();
    while (lo < hi) {
        const uint64_t step = (hi - lo + 63) / 64;
        const uint64_t c = lo + lane * step;
        const bool t = c < hi && pred(c);
        const uint64_t m = (uint64_t)__popcll(__ballot(t));
        const uint64_t nlo = m == 0 ? lo : lo + (m - 1) * step + 1;
        const uint64_t nhi = min(hi, lo + m * step);
        lo = rtgd::uniform_u64(nlo);
        hi = rtgd::uniform_u64(nhi);
    }
    return lo;
}

// First index in x[from..n) holding a value >= v (v may be 2^32: then n).
__device__ __forceinline__ uint32_t wave_lower_bound(const uint32_t *__restrict__ x, uint32_t from, uint32_t n, uint64_t v) {
    return (uint32_t)wave_partition_point(from, n, [&](uint64_t i) { return (uint64_t)x[i] < v; });
}

// The same for one value in two arrays at once: both searches' loads are issued in the same
// rounds, so the two chains of dependent loads overlap.
__device__ __forceinline__ void wave_lower_bound2(const uint32_t *__restrict__ x, uint32_t &xlo, uint32_t xn,
                                                  const uint32_t *__restrict__ y, uint32_t &ylo, uint32_t yn, uint64_t v) {
    const uint32_t lane = __lane_id();
    uint32_t xhi = xn, yhi = yn;
    while (xlo < xhi || ylo < yhi) {
        const uint32_t xs = (uint32_t)(((uint64_t)(xhi - xlo) + 63) / 64), ys = (uint32_t)(((uint64_t)(yhi - ylo) + 63) / 64);
        const uint64_t cx = (uint64_t)xlo + (uint64_t)lane * xs, cy = (uint64_t)ylo + (uint64_t)lane * ys;
        const bool inx = cx < xhi, iny = cy < yhi;
        const uint32_t vx = inx ? x[cx] : 0u, vy = iny ? y[cy] : 0u;
        const uint32_t mx = (uint32_t)__popcll(__ballot(inx && (uint64_t)vx < v));
        const uint32_t my = (uint32_t)__popcll(__ballot(iny && (uint64_t)vy < v));
        const uint32_t nxlo = mx == 0 ? xlo : xlo + (mx - 1) * xs + 1, nylo = my == 0 ? ylo : ylo + (my - 1) * ys + 1;
        const uint32_t nxhi = (uint32_t)min<uint64_t>(xhi, (uint64_t)xlo + (uint64_t)mx * xs);
        const uint32_t nyhi = (uint32_t)min<uint64_t>(yhi, (uint64_t)ylo + (uint64_t)my * ys);
        xlo = rtgd::uniform_u32(nxlo);
        xhi = rtgd::uniform_u32(nxhi);
        ylo = rtgd::uniform_u32(nylo);
        yhi = rtgd::uniform_u32(nyhi);
    }
}

// Number of elements of A among the first d elements of the merge of A and B that takes A
// first on ties (the merge-path co-rank).
__device__ __forceinline__ uint32_t wave_corank(const uint32_t *__restrict__ A, uint32_t na, const uint32_t *__restrict__ B,
                                                uint32_t nb, uint64_t d) {
    const uint64_t lo = d > nb ? d - nb : 0, hi = min<uint64_t>(d, na);
    return (uint32_t)wave_partition_point(lo, hi, [&](uint64_t i) { return A[i] <= B[d - i - 1]; });
}

// Calls f(x[t]) for t in [t0, t1) over the wave's lanes, kU loads per lane in flight per round.
// Offsets are taken relative to the round's base, so no index passes t1 (or wraps).
template <int kU, class F>
__device__ __forceinline__ void wave_stream(const uint32_t *__restrict__ x, uint32_t t0, uint32_t t1, F f) {
    const uint32_t lane = __lane_id();
    for (uint32_t base = t0; base < t1;) {
        const uint32_t lim = t1 - base;
        uint32_t v[kU];
#pragma unroll
        for (uint32_t q = 0; q < kU; ++q) {
            const uint32_t o = lane + 64 * q;
            v[q] = o < lim ? x[base + o] : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < kU; ++q)
            if (lane + 64 * q < lim) f(v[q]);
        if (lim <= 64 * kU) break;
        base += 64 * kU;
    }
}

// Carried from one unit of a wave's run to the next unit of the same pair: where the previous
// unit stopped in each list, so that a continuing unit needs no search to find its start.
struct Carry {
    uint32_t i = 0, j = 0;
};

// Merge windows in LDS: one padding slot per 8 elements, so that lanes reading about 8 elements
// apart (each lane walks kMergeV steps, about half in each list) hit distinct banks.
__device__ __forceinline__ uint32_t mpad(uint32_t t) { return t + (t >> 3); }
constexpr uint32_t kMergeWin = kMergeSpan + kMergeSpan / 8;

// Merge segment r: path steps [r * kMergeSpan, +kMergeSpan). Each element a of A taken by the
// merge is counted when it equals the first unconsumed element of B, which is then the first
// element of B not below a; lists are duplicate-free, so every common value counts once.
__device__ uint32_t merge_unit(const uint32_t *__restrict__ A, uint32_t na, const uint32_t *__restrict__ B, uint32_t nb,
                               uint64_t r, uint32_t *win, Carry &cr, bool carried) {
    const uint32_t lane = __lane_id();
    const uint64_t d0 = r * kMergeSpan;
    const uint32_t len = (uint32_t)min<uint64_t>(kMergeSpan, (uint64_t)na + nb - d0);
    uint32_t i0 = cr.i, j0 = cr.j;
    if (!carried) {
        i0 = wave_corank(A, na, B, nb, d0);
        j0 = (uint32_t)(d0 - i0);
    }
    const uint32_t naw = min(kMergeSpan, na - i0), nbw = min(kMergeSpan, nb - j0);
    uint32_t *sA = win, *sB = win + kMergeWin;
    {
        uint32_t va[kMergeV], vb[kMergeV];
#pragma unroll
        for (uint32_t q = 0; q < kMergeV; ++q) {
            const uint32_t t = lane + 64 * q;
            va[q] = t < naw ? A[i0 + t] : 0u;
            vb[q] = t < nbw ? B[j0 + t] : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < kMergeV; ++q) {
            const uint32_t t = lane + 64 * q;
            if (t < naw) sA[mpad(t)] = va[q];
            if (t < nbw) sB[mpad(t)] = vb[q];
        }
    }
    rtgd::wave_sync_lds();
    // The path of this segment stays in the windows: a step that takes A happens at position
    // i + j < len <= kMergeSpan, so j < kMergeSpan, and j == nbw there means B has ended.
    const uint32_t t = lane * kMergeV;
    uint32_t c = 0, i = 0, j = 0;
    if (t < len) {
        uint32_t lo = t > nbw ? t - nbw : 0, hi = min(t, naw);
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sA[mpad(mid)] <= sB[mpad(t - mid - 1)]) lo = mid + 1;
            else hi = mid;
        }
        i = lo;
        j = t - lo;
        const uint32_t steps = min(kMergeV, len - t);
        uint32_t a = i < naw ? sA[mpad(i)] : 0u, b = j < nbw ? sB[mpad(j)] : 0u;
        for (uint32_t q = 0; q < steps; ++q) {
            if (i < naw && (j >= nbw || a <= b)) {
                c += (j < nbw && a == b) ? 1u : 0u;
                ++i;
                a = i < naw ? sA[mpad(i)] : 0u;
            } else {
                ++j;
                b = j < nbw ? sB[mpad(j)] : 0u;
            }
        }
    }
    // lane 63 ends where a full segment ends: the next segment's start
    cr.i = i0 + (uint32_t)__builtin_amdgcn_readlane((int)i, 63);
    cr.j = j0 + (uint32_t)__builtin_amdgcn_readlane((int)j, 63);
    rtgd::wave_sync_lds();  // the windows are rewritten by the next unit
    return c;
}

__device__ __forceinline__ uint32_t hash_slot(uint32_t v) { return (v * 0x9E3779B1u) >> (32 - 11); }
static_assert(kHashSlots == 2048, "hash_slot yields 11 bits");

// Hash chunk r: A[r * kHashChunk, +kHashChunk) in an LDS table (linear probing, empty slots
// hold kEmpty; a chunk holding the value kEmpty itself, necessarily A's last element, flags
// it instead). The chunk's B range runs from the first element of B not below the chunk's
// first element to the next chunk's, or past A's last element for the last chunk; consecutive
// chunks' ranges meet, so a continuing chunk starts where the previous one stopped.
__device__ uint32_t hash_unit(const uint32_t *__restrict__ A, uint32_t na, const uint32_t *__restrict__ B, uint32_t nb,
                              uint64_t r, uint32_t *tab, Carry &cr, bool carried) {
    const uint32_t lane = __lane_id();
    const uint32_t a0 = (uint32_t)(r * kHashChunk), a1 = min(na, a0 + kHashChunk);
    const uint32_t b0 = carried ? cr.j : wave_lower_bound(B, 0, nb, A[a0]);
    const uint64_t stop = a1 < na ? (uint64_t)A[a1] : (uint64_t)A[na - 1] + 1;
    const uint32_t b1 = wave_lower_bound(B, b0, nb, stop);
    cr.j = b1;
#pragma unroll
    for (uint32_t q = 0; q < kHashSlots / 64; ++q) tab[lane + 64 * q] = kEmpty;
    rtgd::wave_sync_lds();
    bool has_max = false;
    wave_stream<16>(A, a0, a1, [&](uint32_t v) {
        if (v == kEmpty) {
            has_max = true;
            return;
        }
        uint32_t h = hash_slot(v);
        while (atomicCAS(&tab[h], kEmpty, v) != kEmpty) h = (h + 1) & (kHashSlots - 1);
    });
    has_max = __ballot(has_max) != 0;
    rtgd::wave_sync_lds();
    uint32_t c = 0;
    wave_stream<16>(B, b0, b1, [&](uint32_t v) {
        if (v == kEmpty) {
            c += has_max ? 1u : 0u;
            return;
        }
        // Slots h .. h+3 are read at once: with linear probing and no deletions, every slot
        // from a value's home to the value itself is occupied, so v is present iff it shows up
        // before the first empty slot. Only when all four hold other values (rare at a load
        // factor of 1/2) does the walk go on slot by slot.
        uint32_t h = hash_slot(v);
        uint32_t x[4];
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) x[i] = tab[(h + i) & (kHashSlots - 1)];
        bool hit = false, open = false;
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            hit |= x[i] == v;
            open |= x[i] == kEmpty;
        }
        if (!hit && !open) {
            for (h = (h + 4) & (kHashSlots - 1);; h = (h + 1) & (kHashSlots - 1)) {
                const uint32_t y = tab[h];
                if (y == v || y == kEmpty) {
                    hit = y == v;
                    break;
                }
            }
        }
        c += hit ? 1u : 0u;
    });
    rtgd::wave_sync_lds();  // the table is rewritten by the next unit
    return c;
}

// Bitmap window r: values [A[0] + r * kBitmapSpan, +kBitmapSpan), one bit each in LDS. A's
// elements in the window set their bits, B's elements in the window test them; both ranges
// continue from the previous window's ends.
__device__ uint32_t bitmap_unit(const uint32_t *__restrict__ A, uint32_t na, const uint32_t *__restrict__ B, uint32_t nb,
                                uint64_t r, uint32_t *bits, Carry &cr, bool carried) {
    const uint32_t lane = __lane_id();
    const uint64_t lo = (uint64_t)rtgd::uniform_u32(A[0]) + r * kBitmapSpan, hi = lo + kBitmapSpan;
    uint32_t a0 = cr.i, b0 = cr.j;
    if (!carried) {
        a0 = 0;
        b0 = 0;
        wave_lower_bound2(A, a0, na, B, b0, nb, lo);
    }
    uint32_t a1 = a0, b1 = b0;
    wave_lower_bound2(A, a1, na, B, b1, nb, hi);
    cr.i = a1;
    cr.j = b1;
    if (a0 == a1 || b0 == b1) return 0u;
#pragma unroll
    for (uint32_t q = 0; q < kBitmapWords / 64; ++q) bits[lane + 64 * q] = 0u;
    rtgd::wave_sync_lds();
    wave_stream<16>(A, a0, a1, [&](uint32_t v) {
        const uint32_t o = (uint32_t)(v - lo);
        atomicOr(&bits[o >> 5], 1u << (o & 31));
    });
    rtgd::wave_sync_lds();
    uint32_t c = 0;
    wave_stream<16>(B, b0, b1, [&](uint32_t v) {
        const uint32_t o = (uint32_t)(v - lo);
        c += (bits[o >> 5] >> (o & 31)) & 1u;
    });
    rtgd::wave_sync_lds();  // the bitmap is rewritten by the next unit
    return c;
}

// Work is assigned statically (no shared queue: a single-address atomic per grab costs more
// than the grab's work here). Units: wave w takes a contiguous run of units, so it locates
// its first pair by one search over the unit offsets and walks forward, keeping the pair's
// lists and the unit's end position while the run stays on one pair. Flat groups: wave w
// takes pairs 64g .. 64g+63 for g = w, w + waves, ...; pairs whose shorter list is 1..64 long
// concatenate their elements over the lanes (a lane per element per round, one search
// each), pairs with an empty list get 0, longer pairs are left to the units.
template <int kLong>
__global__ __launch_bounds__(kThreads) void si_drain(Lists L, PairSrc P, const uint64_t *__restrict__ unit_off,
                                                     const uint64_t *__restrict__ total_units, int64_t npairs,
                                                     uint32_t *__restrict__ out) {
    constexpr int kWaves = kThreads / 64;
    // One LDS region per wave, used by the units (merge windows, hash table or bitmap) and
    // afterwards by the flat groups: exclusive element offsets of the group's pairs (65), the
    // shorter lists' starts, the longer lists' starts and lengths, and per-pair counts (64 each).
    constexpr uint32_t kFlatLds = 65 + 4 * 64;
    constexpr uint32_t kUnitLds = kLong == kLongSearch ? 0
                                  : kLong == kLongHash  ? kHashSlots
                                  : kLong == kLongMerge ? 2 * kMergeWin
                                                        // bitmap pairs may take merge units
                                                        : (2 * kMergeWin > kBitmapWords ? 2 * kMergeWin : kBitmapWords);
    static_assert(kLong != kLongBitmap || kUnitLds >= kBitmapWords, "bitmap window fits the wave's region");
    constexpr uint32_t kRegion = kUnitLds > kFlatLds ? kUnitLds : kFlatLds;
    __shared__ uint32_t region_all[kWaves][kRegion];
    const int lane = __lane_id();
    const int wave = (int)rtgd::uniform_u32(threadIdx.x >> 6);
    uint32_t *ex = region_all[wave], *sh = ex + 65, *lo = sh + 64, *ln = lo + 64, *cnt = ln + 64;
    const int64_t nwaves = (int64_t)gridDim.x * kWaves;
    const int64_t gw = (int64_t)blockIdx.x * kWaves + wave;

    // Units of the long pairs: unit i belongs to the last pair k with unit_off[k] <= i and is
    // that pair's unit i - unit_off[k].
    const uint64_t nunits = rtgd::uniform_u64(*total_units);
    const uint64_t per = (nunits + nwaves - 1) / nwaves;
    const uint64_t i0 = (uint64_t)gw * per, i1 = min(nunits, i0 + per);
    int64_t k = -1;
    uint64_t k_off = 0, k_end = 0;  // the pair's units are [k_off, k_end)
    const uint32_t *s = nullptr, *l = nullptr;
    uint32_t ns = 0, nl = 0;
    int kind = kLongSearch;
    Carry cr;
    for (uint64_t i = i0; i < i1; ++i) {
        const bool carried = k >= 0 && i < k_end;
        if (!carried) {
            int64_t a0 = k + 1, a1 = npairs;  // upper_bound(unit_off, i) - 1; unit_off[0] = 0
            while (a0 < a1) {
                const int64_t mid = (a0 + a1) >> 1;
                if (rtgd::uniform_u64(unit_off[mid]) <= i) a0 = mid + 1;
                else a1 = mid;
            }
            k = a0 - 1;
            k_off = rtgd::uniform_u64(unit_off[k]);
            k_end = k + 1 < npairs ? rtgd::uniform_u64(unit_off[k + 1]) : nunits;
            uint32_t a, b;
            P.get(k, a, b);
            short_long(L, a, b, s, ns, l, nl);
            uint64_t nu;
            if (kLong != kLongSearch) kind = long_kind<kLong>(s, ns, nl, nu);
        }
        const uint64_t r = i - k_off;
        uint32_t c = 0;
        if (kind == kLongSearch) {
            const uint32_t e = 64u * (uint32_t)r + (uint32_t)lane;
            c = e < ns ? member<4>(l, nl, s[e]) : 0u;
        } else if (kLong != kLongSearch) {
            uint32_t *win = region_all[wave];
            if (kind == kLongMerge) c = merge_unit(s, ns, l, nl, r, win, cr, carried);
            else if (kind == kLongHash) c = hash_unit(s, ns, l, nl, r, win, cr, carried);
            else c = bitmap_unit(s, ns, l, nl, r, win, cr, carried);
        }
        c = rtgd::wave_sum_u32(c);
        if (lane == 0 && c) atomicAdd(&out[k], c);
    }

    // Flat groups.
    const int64_t ngroups = (npairs + 63) / 64;
    for (int64_t grp = gw; grp < ngroups; grp += nwaves) {
        const int64_t kk = grp * 64 + lane;
        const bool in = kk < npairs;
        uint32_t ns = 0, s0 = 0, l0 = 0, nl = 0;
        bool mine = false;  // this lane's pair is answered by the group
        if (in) {
            uint32_t a, b;
            P.get(kk, a, b);
            uint32_t ba, la, bb, lb;
            L.range(a, ba, la);
            L.range(b, bb, lb);
            if (la <= lb) { s0 = ba; ns = la; l0 = bb; nl = lb; }
            else { s0 = bb; ns = lb; l0 = ba; nl = la; }
            mine = ns <= kFlatMax;
            if (!mine) ns = 0;
        }
        const uint32_t incl = (uint32_t)rtgd::wave_inclusive_scan((int32_t)ns);
        const uint32_t total = rtgd::uniform_u32(__builtin_amdgcn_readlane(incl, 63));
        ex[lane] = incl - ns;
        sh[lane] = s0;
        lo[lane] = l0;
        ln[lane] = nl;
        cnt[lane] = 0u;
        if (lane == 0) ex[64] = total;
        rtgd::wave_sync_lds();
        for (uint32_t e0 = 0; e0 < total; e0 += 64) {
            const uint32_t e = e0 + (uint32_t)lane;
            if (e < total) {
                uint32_t p = 0;  // last pair whose first element is at or before e
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1)
                    if (ex[p + step] <= e) p += step;
                const uint32_t v = L.el[sh[p] + (e - ex[p])];
                if (member<4>(L.el + lo[p], ln[p], v)) atomicAdd(&cnt[p], 1u);
            }
        }
        rtgd::wave_sync_lds();
        if (mine) out[kk] = cnt[lane];
        rtgd::wave_sync_lds();  // the group's LDS slots are rewritten by the next group
    }
}

// ---- probe items (latency path) ----
// items[i] = {k, first}: one wave takes elements first .. first+63 of pair k's shorter list,
// one lane per element and one eight-way lifting search of the longer list each, and adds its
// count to out[k] (cleared beforehand). Splitting every pair into 64-element items puts all of
// a batch's probes in flight at once: the launch lasts about one search, however skewed the
// pairs are.
__global__ __launch_bounds__(kThreads) void si_items(Lists L, const uint2 *__restrict__ pairs,
                                                     const uint2 *__restrict__ items, int64_t nitems,
                                                     uint32_t *__restrict__ out) {
    const int64_t nw = (int64_t)gridDim.x * (kThreads / 64);
    for (int64_t it = rtgd::uniform_u64((uint64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)); it < nitems;
         it += nw) {
        const uint2 item = items[it];
        const uint2 p = pairs[item.x];
        const uint32_t *s, *l;
        uint32_t ns, nl;
        short_long(L, p.x, p.y, s, ns, l, nl);
        const uint32_t i = item.y + __lane_id();
        uint32_t c = 0;
        if (i < ns) {
            const uint32_t v = s[i];
            c = member<8>(l, nl, v);
        }
        c = rtgd::wave_sum_u32(c);
        if (__lane_id() == 0 && c) atomicAdd(&out[item.x], c);
    }
}

// Statistics pass: algorithmic bytes (sum over pairs of 20 + list_bytes(s, l)) and the units
// of each long-pair kind the batched launch runs.
template <int kLong>
__global__ __launch_bounds__(kThreads) void si_stats_pass(Lists L, PairSrc P, int64_t npairs,
                                                          unsigned long long *acc) {
    uint64_t v[1 + kLongKinds] = {};
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < npairs; k += (int64_t)gridDim.x * blockDim.x) {
        uint32_t a, b;
        P.get(k, a, b);
        const uint32_t *s, *l;
        uint32_t ns, nl;
        short_long(L, a, b, s, ns, l, nl);
        const uint64_t bits = nl ? 32 - __builtin_clz(nl) : 0;
        const uint64_t merge = (uint64_t)ns + nl, bs = (uint64_t)ns * (1 + bits);
        v[0] += 20 + 4 * (merge < bs ? merge : bs);
        if (kLong >= 0 && ns > kFlatMax) {
            uint64_t u;
            const int kind = long_kind<kLong < 0 ? 0 : kLong>(s, ns, nl, u);
            v[1 + kind] += u;
        }
    }
    // one atomic per block and counter: same-address atomics from every wave would serialise
    __shared__ uint64_t part[kThreads / 64][1 + kLongKinds];
#pragma unroll
    for (int q = 0; q <= kLongKinds; ++q) {
        const uint64_t w = rtgd::wave_sum_u64(v[q]);
        if (__lane_id() == 0) part[threadIdx.x >> 6][q] = w;
    }
    __syncthreads();
    if (threadIdx.x <= kLongKinds) {
        uint64_t t = 0;
        for (int w = 0; w < kThreads / 64; ++w) t += part[w][threadIdx.x];
        if (t) atomicAdd(&acc[threadIdx.x], (unsigned long long)t);
    }
}

unsigned grid_for(rtg_ctx *ctx, int64_t n, int per_cu) {
    const int64_t want = rtg::ceil_div(std::max<int64_t>(n, 1), kThreads);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)ctx->cu_count * per_cu));
}

struct Free {
    void *p = nullptr;
    ~Free() {
        if (p) (void)hipFree(p);
    }
};

// What identifies a call for the autotuner's memory.
struct SiKey {
    const void *pairs;
    uint32_t num_a, num_b;
};

int long_of(int32_t method) {
    switch (method) {
        case RTG_SI_BATCHED: return kLongSearch;
        case RTG_SI_MERGE: return kLongMerge;
        case RTG_SI_HASH: return kLongHash;
        case RTG_SI_BITMAP: return kLongBitmap;
        default: return -1;
    }
}

struct BatchBufs {
    uint64_t *unit_off;
    uint64_t *total;
    void *scan_work;
};

template <int kLong>
int launch_batched_t(rtg_ctx *ctx, Lists L, PairSrc P, int64_t npairs, const BatchBufs &B, uint32_t *d_out) {
    hipStream_t st = ctx->stream;
    si_classify<kLong><<<grid_for(ctx, npairs, 16), kThreads, 0, st>>>(L, P, npairs, B.unit_off, d_out);
    HIP_TRY(hipGetLastError());
    RTG_TRY(rtg::scan_exclusive_u64_w(ctx, B.unit_off, B.unit_off, npairs, B.total, B.scan_work));
    // The search drain runs two generations of blocks; the LDS-staging drains one (four blocks
    // per CU fit), so each wave's run of units is as long as it can be.
    const int per_cu = kLong == kLongSearch ? 8 : 4;
    si_drain<kLong><<<(unsigned)((int64_t)ctx->cu_count * per_cu), kThreads, 0, st>>>(L, P, B.unit_off, B.total, npairs,
                                                                                      d_out);
    HIP_TRY(hipGetLastError());
    return RTG_OK;
}

int launch_batched(rtg_ctx *ctx, int kind, Lists L, PairSrc P, int64_t npairs, const BatchBufs &B, uint32_t *d_out) {
    switch (kind) {
        case kLongSearch: return launch_batched_t<kLongSearch>(ctx, L, P, npairs, B, d_out);
        case kLongMerge: return launch_batched_t<kLongMerge>(ctx, L, P, npairs, B, d_out);
        case kLongHash: return launch_batched_t<kLongHash>(ctx, L, P, npairs, B, d_out);
        default: return launch_batched_t<kLongBitmap>(ctx, L, P, npairs, B, d_out);
    }
}

// Times every long-pair strategy once on this input (each run writes the full, exact result)
// and returns the fastest; the launch after it is the one the caller's timing sees.
// The block-per-pair form with LDS samples (RTG_SI_REFERENCE), also an autotuner candidate:
// with few, long pairs whose longer list is shared (a tile with one B set) it can win.
constexpr int kSampled = kLongKinds;

int launch_sampled(rtg_ctx *ctx, Lists L, PairSrc P, int64_t npairs, uint32_t *d_out) {
    si_reference<<<(unsigned)std::min<int64_t>(npairs, (int64_t)ctx->cu_count * 8), kThreads, 0, ctx->stream>>>(L, P, npairs,
                                                                                                              d_out);
    HIP_TRY(hipGetLastError());
    return RTG_OK;
}

int autotune(rtg_ctx *ctx, Lists L, PairSrc P, int64_t npairs, const BatchBufs &B, uint32_t *d_out, int *best) {
    hipStream_t st = ctx->stream;
    float best_ms = 0.f;
    *best = kLongSearch;
    for (int kind = 0; kind <= kSampled; ++kind) {
        HIP_TRY(hipEventRecord(ctx->ev[2], st));
        if (kind == kSampled) RTG_TRY(launch_sampled(ctx, L, P, npairs, d_out));
        else RTG_TRY(launch_batched(ctx, kind, L, P, npairs, B, d_out));
        HIP_TRY(hipEventRecord(ctx->ev[3], st));
        HIP_TRY(hipEventSynchronize(ctx->ev[3]));
        const float ms = rtg::elapsed_ms(ctx->ev[2], ctx->ev[3]);
        if (kind == 0 || ms < best_ms) {
            best_ms = ms;
            *best = kind;
        }
    }
    return RTG_OK;
}

// The counting launch of a method (kind: its long-pair strategy, or kSampled when AUTO picked
// the block-per-pair form); B is unused by PLAIN and REFERENCE.
int launch_method(rtg_ctx *ctx, Lists L, PairSrc P, int64_t npairs, int32_t method, int kind, const BatchBufs &B,
                  uint32_t *d_out) {
    if (method == RTG_SI_PLAIN) {
        si_plain<<<grid_for(ctx, npairs * 64, 32), kThreads, 0, ctx->stream>>>(L, P, npairs, d_out);
        HIP_TRY(hipGetLastError());
        return RTG_OK;
    }
    if (method == RTG_SI_REFERENCE || kind == kSampled) return launch_sampled(ctx, L, P, npairs, d_out);
    return launch_batched(ctx, kind, L, P, npairs, B, d_out);
}

// nlists: lists in d_off; pairs naming an id at or above it count 0 and make the call fail
// with RTG_EINVAL after it ran (the kernels never read past the offsets).
int si_run(rtg_ctx *ctx, const uint32_t *d_off, const uint32_t *d_el, uint32_t nlists, PairSrc P, int64_t npairs,
           int32_t method, uint32_t *d_out, rtg_si_stats *stats) {
    using namespace rtg;
    hipStream_t st = ctx->stream;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (npairs <= 0) return RTG_OK;
    // Context scratch: bytes [0, 64) hold the statistics pass's sums, the word at byte 192 the
    // out-of-range flag.
    void *scr = nullptr;
    RTG_TRY(ctx_scratch(ctx, 256, &scr));
    uint32_t *d_bad = reinterpret_cast<uint32_t *>(static_cast<char *>(scr) + 192);
    HIP_TRY(hipMemsetAsync(d_bad, 0, sizeof(uint32_t), st));
    Lists L{d_off, d_el, nlists, d_bad};
    if (npairs >= ((int64_t)1 << 32)) return set_error(RTG_EINVAL, "rtg_si: more than 2^32 pairs");
    const bool autotuned = method == RTG_SI_AUTO;
    int kind = autotuned ? kLongSearch : long_of(method);
    if (kind < 0 && method != RTG_SI_PLAIN && method != RTG_SI_REFERENCE)
        return set_error(RTG_EINVAL, "rtg_si: unknown method");
    Free fb;
    BatchBufs B{};
    if (kind >= 0) {
        const size_t sw = scan_work_bytes(npairs);
        const size_t bytes = Arena::round(sizeof(uint64_t) * (size_t)npairs) + Arena::round(sw) + 1024;
        HIP_TRY(hipMalloc(&fb.p, bytes));
        Arena ar{(char *)fb.p, 0, bytes};
        B.unit_off = ar.take<uint64_t>((size_t)npairs);
        B.total = ar.take<uint64_t>(4);
        B.scan_work = ar.take<char>(sw);
    }
    double t_tune = 0;
    if (autotuned) {
        auto &memo = ctx->si_tuned;
        const SiKey key{P.pairs, P.num_a, P.num_b};
        auto hit = std::find_if(memo.begin(), memo.end(), [&](const rtg_ctx::SiTuned &e) {
            return e.off == d_off && e.el == d_el && e.pairs == key.pairs && e.npairs == npairs && e.num_a == key.num_a &&
                   e.num_b == key.num_b;
        });
        if (hit != memo.end()) {
            kind = hit->kind;
        } else {
            const auto t0 = std::chrono::steady_clock::now();
            RTG_TRY(autotune(ctx, L, P, npairs, B, d_out, &kind));
            t_tune = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (memo.size() >= 64) memo.erase(memo.begin());
            memo.push_back({d_off, d_el, key.pairs, npairs, key.num_a, key.num_b, kind});
        }
    }
    HIP_TRY(hipEventRecord(ctx->ev[0], st));
    RTG_TRY(launch_method(ctx, L, P, npairs, method, kind, B, d_out));
    HIP_TRY(hipEventRecord(ctx->ev[1], st));
    uint32_t bad = 0;
    if (stats) {
        constexpr size_t kAcc = sizeof(unsigned long long) * (1 + kLongKinds);
        static_assert(kAcc <= 192, "statistics sums overlap the out-of-range flag");
        HIP_TRY(hipMemsetAsync(scr, 0, kAcc, st));
        auto *acc = (unsigned long long *)scr;
        const unsigned g = grid_for(ctx, npairs, 2);
        switch (kind) {
            case kLongSearch: si_stats_pass<kLongSearch><<<g, kThreads, 0, st>>>(L, P, npairs, acc); break;
            case kLongMerge: si_stats_pass<kLongMerge><<<g, kThreads, 0, st>>>(L, P, npairs, acc); break;
            case kLongHash: si_stats_pass<kLongHash><<<g, kThreads, 0, st>>>(L, P, npairs, acc); break;
            case kLongBitmap: si_stats_pass<kLongBitmap><<<g, kThreads, 0, st>>>(L, P, npairs, acc); break;
            default: si_stats_pass<-1><<<g, kThreads, 0, st>>>(L, P, npairs, acc); break;
        }
        HIP_TRY(hipGetLastError());
        unsigned long long v[1 + kLongKinds] = {};
        HIP_TRY(hipMemcpyAsync(v, scr, kAcc, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        stats->t_kernel_ms = elapsed_ms(ctx->ev[0], ctx->ev[1]);
        stats->t_plan_ms = t_tune;
        stats->alg_bytes = v[0];
        stats->strategy = kind;
        if (kind >= 0 && kind < kLongKinds) {  // units of each long-pair kind: search items, merge segments, hash chunks, bitmap windows
            stats->nbuckets = kLongKinds;
            for (int q = 0; q < kLongKinds; ++q) stats->bucket_items[q] = (int64_t)v[1 + q];
        }
    } else {
        HIP_TRY(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (bad) return set_error(RTG_EINVAL, "rtg_si: a pair names a list id at or above the number of lists (counted as 0)");
    return RTG_OK;
}

}  // namespace

// ---- allocation-free entry point ---------------------------------------------------------
// rtg_si_pairs_noalloc keeps its scratch — the out-of-range flag, the unit total, the per-pair
// unit offsets and the scan's tile sums — in this file-scope device array instead of memory
// allocated per call. Every GPU holds its own copy (loaded with the code object, held until the
// process ends); host code reaches it only through the address hipGetSymbolAddress returns for
// the current device, never through the array's name, which in host code is only a stand-in.
// Layout: [0, 4) flag, [64, 96) unit total, [256, 256 + 8 * RTG_SI_NOALLOC_MAX_PAIRS) unit
// offsets, then the scan work.
namespace {
constexpr size_t kNoallocOffsets = 256;
constexpr size_t kNoallocScanWork = 64 * 1024;
constexpr size_t kNoallocBytes = kNoallocOffsets + 8 * (size_t)RTG_SI_NOALLOC_MAX_PAIRS + kNoallocScanWork;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {

constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoallocMaxDevices];      // device address of `scratch`, per device
int g_noalloc_status[kNoallocMaxDevices];     // 0: not looked up yet, 1: ok, 2: lookup failed

__global__ void noalloc_clear_flag(uint32_t *flag) {
    if (threadIdx.x == 0) *flag = 0u;
}

}  // namespace

namespace rtg {

// Called by rtg_create with `device` current: looks up this device's copy of the scratch array
// once per device and keeps its address.
int si_noalloc_init(int device) {
    if (device < 0 || device >= kNoallocMaxDevices) return RTG_OK;  // the entry point will refuse
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_status[device] != 0) return RTG_OK;
    static_assert(kNoallocScanWork >= 4096, "scan work");
    if (scan_work_bytes(RTG_SI_NOALLOC_MAX_PAIRS) > kNoallocScanWork) {
        g_noalloc_status[device] = 2;
        return set_error(RTG_EHIP, "set-intersection scratch: scan work does not fit");
    }
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        g_noalloc_status[device] = 2;
        return set_error(RTG_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_noalloc_ptr[device] = ptr;
    g_noalloc_status[device] = 1;
    return RTG_OK;
}

}  // namespace rtg

namespace {

int si_run_noalloc(rtg_ctx *ctx, const uint32_t *d_off, const uint32_t *d_el, uint32_t nlists, PairSrc P,
                   int64_t npairs, int32_t method, uint32_t *d_out, double *t_kernel_ms) {
    using namespace rtg;
    hipStream_t st = ctx->stream;
    char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (ctx->device < kNoallocMaxDevices && g_noalloc_status[ctx->device] == 1)
            base = static_cast<char *>(g_noalloc_ptr[ctx->device]);
    }
    if (!base) return set_error(RTG_EHIP, "rtg_si_pairs_noalloc: the scratch array of this device is unavailable");
    uint32_t *d_bad = reinterpret_cast<uint32_t *>(base);
    BatchBufs B{reinterpret_cast<uint64_t *>(base + kNoallocOffsets), reinterpret_cast<uint64_t *>(base + 64),
                base + kNoallocOffsets + 8 * (size_t)RTG_SI_NOALLOC_MAX_PAIRS};
    const int kind = long_of(method);
    Lists L{d_off, d_el, nlists, d_bad};
    // The flag keeps the previous call's value: cleared by a kernel (no memset node, so the
    // sequence can also be recorded into a graph by a caller).
    noalloc_clear_flag<<<1, 64, 0, st>>>(d_bad);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[0], st));
    RTG_TRY(launch_method(ctx, L, P, npairs, method, kind, B, d_out));
    HIP_TRY(hipEventRecord(ctx->ev[1], st));
    uint32_t bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (t_kernel_ms) *t_kernel_ms = elapsed_ms(ctx->ev[0], ctx->ev[1]);
    if (bad) return set_error(RTG_EINVAL, "rtg_si_pairs_noalloc: a pair names a list id at or above the number of lists (counted as 0)");
    return RTG_OK;
}

}  // namespace

namespace rtg {

int si_launch_items(rtg_ctx *ctx, hipStream_t st, const uint32_t *d_off, const uint32_t *d_el, const uint32_t *d_pairs,
                    const uint32_t *d_items, int64_t nitems, uint32_t *d_out) {
    if (nitems <= 0) return RTG_OK;
    Lists L{d_off, d_el, 0xFFFFFFFFu, nullptr};  // vertex ids were checked when the requests came in
    const int64_t blocks = ceil_div(nitems, (int64_t)(kThreads / 64));
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks, (int64_t)ctx->cu_count * 16));
    si_items<<<grid, kThreads, 0, st>>>(L, reinterpret_cast<const uint2 *>(d_pairs), reinterpret_cast<const uint2 *>(d_items),
                                        nitems, d_out);
    HIP_TRY(hipGetLastError());
    return RTG_OK;
}

}  // namespace rtg

extern "C" {

RTG_API int rtg_si_pairs(rtg_ctx *ctx, const uint32_t *d_offsets, const uint32_t *d_elements, int64_t nlists,
                         const uint32_t *d_pairs, int64_t npairs, int32_t method, uint32_t *d_out,
                         rtg_si_stats *stats) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (nlists < 0 || npairs < 0 || !d_offsets || (npairs > 0 && (!d_pairs || !d_out)))
        return rtg_fail(RTG_EINVAL, "rtg_si_pairs: bad arguments");
    if (((uintptr_t)d_pairs & 7) != 0) return rtg_fail(RTG_EINVAL, "rtg_si_pairs: d_pairs must be 8-byte aligned");
    if (nlists > (int64_t)0xFFFFFFFFu) return rtg_fail(RTG_EINVAL, "rtg_si_pairs: list ids are 32-bit (at most 2^32 - 1 lists)");
    std::lock_guard<std::mutex> lock(ctx->mu);
    PairSrc P{d_pairs, 0, 1};
    return si_run(ctx, d_offsets, d_elements, (uint32_t)nlists, P, npairs, method, d_out, stats);
}

RTG_API int rtg_si_pairs_noalloc(rtg_ctx *ctx, const uint32_t *d_offsets, const uint32_t *d_elements, int64_t nlists,
                                 const uint32_t *d_pairs, int64_t npairs, int32_t method, uint32_t *d_out,
                                 double *t_kernel_ms) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (t_kernel_ms) *t_kernel_ms = 0.0;
    if (nlists < 0 || npairs < 0 || !d_offsets || (npairs > 0 && (!d_pairs || !d_out)))
        return rtg_fail(RTG_EINVAL, "rtg_si_pairs_noalloc: bad arguments");
    if (((uintptr_t)d_pairs & 7) != 0) return rtg_fail(RTG_EINVAL, "rtg_si_pairs_noalloc: d_pairs must be 8-byte aligned");
    if (nlists > (int64_t)0xFFFFFFFFu)
        return rtg_fail(RTG_EINVAL, "rtg_si_pairs_noalloc: list ids are 32-bit (at most 2^32 - 1 lists)");
    if (npairs > RTG_SI_NOALLOC_MAX_PAIRS)
        return rtg_fail(RTG_EINVAL, "rtg_si_pairs_noalloc: more than RTG_SI_NOALLOC_MAX_PAIRS (" +
                                        std::to_string((long long)RTG_SI_NOALLOC_MAX_PAIRS) + ") pairs");
    if (method != RTG_SI_PLAIN && method != RTG_SI_REFERENCE && long_of(method) < 0)
        return rtg_fail(RTG_EINVAL, "rtg_si_pairs_noalloc: method must be PLAIN, BATCHED, REFERENCE, MERGE, HASH or BITMAP");
    if (npairs == 0) return RTG_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    PairSrc P{d_pairs, 0, 1};
    return si_run_noalloc(ctx, d_offsets, d_elements, (uint32_t)nlists, P, npairs, method, d_out, t_kernel_ms);
}

RTG_API int rtg_si_tile(rtg_ctx *ctx, const uint32_t *d_offsets, const uint32_t *d_elements, uint32_t num_a,
                        uint32_t num_b, int32_t method, uint32_t *d_counts, rtg_si_stats *stats) {
    RTG_TRY(rtg::ctx_enter(ctx));
    const int64_t np = (int64_t)num_a * num_b;
    if (!d_offsets || (np > 0 && !d_counts)) return rtg_fail(RTG_EINVAL, "rtg_si_tile: bad arguments");
    if ((uint64_t)num_a + num_b > 0xFFFFFFFFull) return rtg_fail(RTG_EINVAL, "rtg_si_tile: num_a + num_b must fit 32 bits");
    std::lock_guard<std::mutex> lock(ctx->mu);
    PairSrc P{nullptr, num_a, num_b ? num_b : 1};
    return si_run(ctx, d_offsets, d_elements, num_a + num_b, P, np, method, d_counts, stats);
}

}  // extern "C"
