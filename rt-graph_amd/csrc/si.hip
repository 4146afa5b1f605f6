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
//                 binary search each, partial counts added into the result) and flat groups
//                 of 64 consecutive pairs, whose shorter lists of 0..64 elements are
//                 concatenated over the lanes so that every lane does one binary search per
//                 round whatever the pairs' lengths. Every search in the launch is one binary
//                 search deep: neither a long pair nor a run of length-1 pairs serialises a
//                 chain of searches or leaves lanes idle. Work is assigned statically; no
//                 counter is shared between waves.
#include "device.hpp"
#include "si.hpp"

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
    __device__ __forceinline__ void span(uint32_t id, const uint32_t *&p, uint32_t &len) const {
        const uint32_t b = off[id];
        p = el + b;
        len = off[id + 1] - b;
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

struct Batch {
    uint32_t *item_off;     // per pair: item count (s > kFlatMax), scanned in place
    uint32_t *total_items;  // written by the scan
};

// Per pair: the number of 64-element items of its shorter list when it is longer than
// kFlatMax (those pairs' results are cleared here and accumulated by the items), else 0.
__global__ __launch_bounds__(kThreads) void si_classify(Lists L, PairSrc P, int64_t npairs, Batch B,
                                                        uint32_t *__restrict__ out) {
    for (int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x; k < npairs; k += (int64_t)gridDim.x * kThreads) {
        uint32_t a, b;
        P.get(k, a, b);
        const uint32_t la = L.off[a + 1] - L.off[a], lb = L.off[b + 1] - L.off[b];
        const uint32_t sl = la < lb ? la : lb;
        const uint32_t items = sl > kFlatMax ? (sl + 63u) / 64u : 0u;
        if (items) out[k] = 0u;
        B.item_off[k] = items;
    }
}

// Number of elements of l[0..nl) below v, nl >= 1, by binary lifting: the trip count depends
// only on nl, and every probe address is clamped into the list.
__device__ __forceinline__ uint32_t count_below_lift(const uint32_t *__restrict__ l, uint32_t nl, uint32_t v) {
    uint32_t base = 0;
    for (uint32_t step = 1u << (31 - __builtin_clz(nl)); step > 0; step >>= 1) {
        const uint32_t x = l[min(base + step, nl) - 1];
        if (base + step <= nl && x < v) base += step;
    }
    return base;
}

__device__ __forceinline__ uint32_t member(const uint32_t *__restrict__ l, uint32_t nl, uint32_t v) {
    const uint32_t b = count_below_lift(l, nl, v);
    return (b < nl && l[b] == v) ? 1u : 0u;
}

// Work is assigned statically (no shared queue: a single-address atomic per grab costs more
// than the grab's work here). Items: wave w takes a contiguous run of items, so it locates
// its first pair by one binary search over item_off and walks forward. Flat groups: wave w
// takes pairs 64g .. 64g+63 for g = w, w + waves, ...; pairs whose shorter list is 1..64 long
// concatenate their elements over the lanes (a lane per element per round, one binary search
// each), pairs with an empty list get 0, longer pairs are left to the items.
__global__ __launch_bounds__(kThreads) void si_drain(Lists L, PairSrc P, Batch B, int64_t npairs,
                                                     uint32_t *__restrict__ out) {
    constexpr int kWaves = kThreads / 64;
    __shared__ uint32_t ex_all[kWaves][65];   // exclusive element offsets of the group's pairs
    __shared__ uint32_t sh_all[kWaves][64];   // shorter list: start in elements
    __shared__ uint32_t lo_all[kWaves][64];   // longer list: start in elements
    __shared__ uint32_t ln_all[kWaves][64];   // longer list: length
    __shared__ uint32_t cnt_all[kWaves][64];  // per-pair counts
    const int lane = __lane_id();
    const int wave = (int)rtgd::uniform_u32(threadIdx.x >> 6);
    uint32_t *ex = ex_all[wave], *sh = sh_all[wave], *lo = lo_all[wave], *ln = ln_all[wave], *cnt = cnt_all[wave];
    const int64_t nwaves = (int64_t)gridDim.x * kWaves;
    const int64_t gw = (int64_t)blockIdx.x * kWaves + wave;

    // Items of the long pairs: item i belongs to the last pair k with item_off[k] <= i and
    // covers elements 64*(i - item_off[k]) .. +63 of its shorter list.
    const int64_t nitems = rtgd::uniform_u32(*B.total_items);
    const int64_t per = (nitems + nwaves - 1) / nwaves;
    const int64_t i0 = gw * per, i1 = min(nitems, i0 + per);
    int64_t k = -1;
    for (int64_t i = i0; i < i1; ++i) {
        if (k < 0 || (k + 1 < npairs && rtgd::uniform_u32(B.item_off[k + 1]) <= (uint64_t)i)) {
            int64_t a0 = k + 1, a1 = npairs;  // upper_bound(item_off, i) - 1; item_off[0] = 0
            while (a0 < a1) {
                const int64_t mid = (a0 + a1) >> 1;
                if (rtgd::uniform_u32(B.item_off[mid]) <= (uint64_t)i) a0 = mid + 1;
                else a1 = mid;
            }
            k = a0 - 1;
        }
        uint32_t a, b;
        P.get(k, a, b);
        const uint32_t *s, *l;
        uint32_t ns, nl;
        short_long(L, a, b, s, ns, l, nl);
        const uint32_t e = 64u * (uint32_t)(i - B.item_off[k]) + (uint32_t)lane;
        uint32_t c = e < ns ? member(l, nl, s[e]) : 0u;
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
            const uint32_t ba = L.off[a], la = L.off[a + 1] - ba, bb = L.off[b], lb = L.off[b + 1] - bb;
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
                if (member(L.el + lo[p], ln[p], v)) atomicAdd(&cnt[p], 1u);
            }
        }
        rtgd::wave_sync_lds();
        if (mine) out[kk] = cnt[lane];
        rtgd::wave_sync_lds();  // the group's LDS slots are rewritten by the next group
    }
}

// ---- probe items (latency path) ----
// items[i] = {k, first}: one wave takes elements first .. first+63 of pair k's shorter list,
// one lane per element and one binary search of the longer list each, and adds its count to
// out[k] (cleared beforehand). Splitting every pair into 64-element items puts all of a
// batch's probes in flight at once: the launch lasts about one binary search, however
// skewed the pairs are.
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
            c = member(l, nl, v);
        }
        c = rtgd::wave_sum_u32(c);
        if (__lane_id() == 0 && c) atomicAdd(&out[item.x], c);
    }
}

__global__ __launch_bounds__(kThreads) void si_alg_bytes(Lists L, PairSrc P, int64_t npairs, unsigned long long *acc) {
    uint64_t s = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < npairs; k += (int64_t)gridDim.x * blockDim.x) {
        uint32_t a, b;
        P.get(k, a, b);
        const uint32_t la = L.off[a + 1] - L.off[a], lb = L.off[b + 1] - L.off[b];
        const uint64_t sm = la < lb ? la : lb, lg = la < lb ? lb : la;
        const uint64_t bits = lg ? 64 - __clzll(lg) : 0;
        const uint64_t merge = sm + lg, bs = sm * (1 + bits);
        s += 20 + 4 * (merge < bs ? merge : bs);
    }
    // one atomic per block: same-address atomics from every wave would serialise
    __shared__ uint64_t part[kThreads / 64];
    s = rtgd::wave_sum_u64(s);
    if (__lane_id() == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kThreads / 64; ++w) t += part[w];
        if (t) atomicAdd(acc, (unsigned long long)t);
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

int si_run(rtg_ctx *ctx, const uint32_t *d_off, const uint32_t *d_el, PairSrc P, int64_t npairs, int32_t method,
           uint32_t *d_out, rtg_si_stats *stats) {
    using namespace rtg;
    hipStream_t st = ctx->stream;
    Lists L{d_off, d_el};
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (npairs <= 0) return RTG_OK;
    if (npairs >= ((int64_t)1 << 32)) return set_error(RTG_EINVAL, "rtg_si: more than 2^32 pairs");
    Free fb;
    Batch B{};
    void *scan_work = nullptr;
    if (method == RTG_SI_BATCHED) {
        const size_t sw = scan_work_bytes(npairs);
        const size_t bytes = Arena::round(sizeof(uint32_t) * (size_t)npairs) + Arena::round(sw) + 1024;
        HIP_TRY(hipMalloc(&fb.p, bytes));
        Arena ar{(char *)fb.p, 0, bytes};
        B.item_off = ar.take<uint32_t>((size_t)npairs);
        B.total_items = ar.take<uint32_t>(4);
        scan_work = ar.take<char>(sw);
    }
    HIP_TRY(hipEventRecord(ctx->ev[0], st));
    switch (method) {
        case RTG_SI_PLAIN:
            si_plain<<<grid_for(ctx, npairs * 64, 32), kThreads, 0, st>>>(L, P, npairs, d_out);
            break;
        case RTG_SI_REFERENCE:
            si_reference<<<(unsigned)std::min<int64_t>(npairs, (int64_t)ctx->cu_count * 8), kThreads, 0, st>>>(L, P, npairs, d_out);
            break;
        case RTG_SI_BATCHED: {
            si_classify<<<grid_for(ctx, npairs, 16), kThreads, 0, st>>>(L, P, npairs, B, d_out);
            RTG_TRY(scan_exclusive_u32_w(ctx, B.item_off, B.item_off, npairs, B.total_items, scan_work));
            si_drain<<<(unsigned)((int64_t)ctx->cu_count * 8), kThreads, 0, st>>>(L, P, B, npairs, d_out);
            break;
        }
        default:
            return set_error(RTG_EINVAL, "rtg_si: unknown method");
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[1], st));
    if (stats) {
        void *s = nullptr;
        RTG_TRY(ctx_scratch(ctx, 256, &s));
        HIP_TRY(hipMemsetAsync(s, 0, sizeof(unsigned long long), st));
        si_alg_bytes<<<grid_for(ctx, npairs, 2), kThreads, 0, st>>>(L, P, npairs, (unsigned long long *)s);
        HIP_TRY(hipGetLastError());
        unsigned long long ab = 0;
        uint32_t nitems = 0;
        HIP_TRY(hipMemcpyAsync(&ab, s, sizeof(ab), hipMemcpyDeviceToHost, st));
        if (method == RTG_SI_BATCHED) HIP_TRY(hipMemcpyAsync(&nitems, B.total_items, sizeof(nitems), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        stats->t_kernel_ms = elapsed_ms(ctx->ev[0], ctx->ev[1]);
        stats->alg_bytes = ab;
        if (method == RTG_SI_BATCHED) {  // 64-element items of the pairs answered by items
            stats->nbuckets = 1;
            stats->bucket_items[0] = (int64_t)nitems;
        }
    } else {
        HIP_TRY(hipStreamSynchronize(st));
    }
    return RTG_OK;
}

}  // namespace

namespace rtg {

int si_launch_items(rtg_ctx *ctx, hipStream_t st, const uint32_t *d_off, const uint32_t *d_el, const uint32_t *d_pairs,
                    const uint32_t *d_items, int64_t nitems, uint32_t *d_out) {
    if (nitems <= 0) return RTG_OK;
    Lists L{d_off, d_el};
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
    std::lock_guard<std::mutex> lock(ctx->mu);
    PairSrc P{d_pairs, 0, 1};
    return si_run(ctx, d_offsets, d_elements, P, npairs, method, d_out, stats);
}

RTG_API int rtg_si_tile(rtg_ctx *ctx, const uint32_t *d_offsets, const uint32_t *d_elements, uint32_t num_a,
                        uint32_t num_b, int32_t method, uint32_t *d_counts, rtg_si_stats *stats) {
    RTG_TRY(rtg::ctx_enter(ctx));
    const int64_t np = (int64_t)num_a * num_b;
    if (!d_offsets || (np > 0 && !d_counts)) return rtg_fail(RTG_EINVAL, "rtg_si_tile: bad arguments");
    std::lock_guard<std::mutex> lock(ctx->mu);
    PairSrc P{nullptr, num_a, num_b ? num_b : 1};
    return si_run(ctx, d_offsets, d_elements, P, np, method, d_counts, stats);
}

}  // extern "C"
