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
//   SI_BATCHED    a classify pass writes s = 0 pairs' results directly and appends the other
//                 pairs to three size buckets; one persistent launch then drains the block
//                 bucket (block per pair, sampled search as above), the wave bucket (wave per
//                 pair) and the thread bucket (thread per pair), each through an atomic
//                 cursor, so that no launch parameter depends on the bucket sizes.
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
struct Buckets {
    uint32_t *list[3];                // pair indices: 0 thread, 1 wave, 2 block
    unsigned long long *count;        // [0..2] bucket sizes, [3..5] drain cursors
};

__device__ __forceinline__ void bucket_append(bool valid, uint32_t k, uint32_t *list, unsigned long long *counter) {
    const uint64_t mask = __ballot(valid);
    if (mask == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (valid) list[base + __popcll(mask & rtgd::lanemask_lt())] = k;
}

__global__ void si_reset(unsigned long long *c) {
    if (threadIdx.x < 6) c[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(kThreads) void si_classify(Lists L, PairSrc P, int64_t npairs, Buckets B,
                                                        uint32_t thread_max, uint32_t wave_max, uint32_t *__restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    const int64_t start = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    for (int64_t base = start - (start & 63); base < npairs; base += stride) {
        const int64_t k = base + (start & 63);
        int bucket = -1;
        if (k < npairs) {
            uint32_t a, b;
            P.get(k, a, b);
            const uint32_t la = L.off[a + 1] - L.off[a], lb = L.off[b + 1] - L.off[b];
            const uint32_t s = la < lb ? la : lb;
            const uint32_t work = s;  // probes
            if (s == 0) out[k] = 0u;
            else bucket = (work <= thread_max) ? 0 : (work <= wave_max ? 1 : 2);
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) bucket_append(bucket == q, (uint32_t)k, B.list[q], &B.count[q]);
    }
}

__global__ __launch_bounds__(kThreads) void si_drain(Lists L, PairSrc P, Buckets B, uint32_t *__restrict__ out) {
    __shared__ uint32_t samp[kSamples];
    __shared__ uint32_t red[kThreads / 64];
    __shared__ unsigned long long slot;
    // Block bucket: block per pair. Every grab is published through LDS and read back as a
    // wave-uniform (scalar) value, so the loop below is uniform.
    const uint64_t n2 = rtgd::uniform_u64(B.count[2]);
    if (threadIdx.x == 0) slot = atomicAdd(&B.count[5], 1ull);
    __syncthreads();
    uint64_t i2 = rtgd::uniform_u64(slot);
    while (i2 < n2) {
        const uint32_t k = B.list[2][i2];
        uint32_t a, b;
        P.get(k, a, b);
        const uint32_t c = block_pair(L, a, b, samp, red);  // ends with a barrier
        if (threadIdx.x == 0) {
            out[k] = c;
            slot = atomicAdd(&B.count[5], 1ull);
        }
        __syncthreads();
        i2 = rtgd::uniform_u64(slot);
    }
    // Wave bucket: wave per pair, 4 pairs per grab.
    const uint64_t n1 = rtgd::uniform_u64(B.count[1]);
    while (true) {
        uint64_t g = 0;
        if (__lane_id() == 0) g = atomicAdd(&B.count[4], 4ull);
        const uint64_t i = rtgd::uniform_u64(g);
        if (i >= n1) break;
        const uint64_t e = (i + 4 < n1) ? i + 4 : n1;
        for (uint64_t j = i; j < e; ++j) {
            const uint32_t k = B.list[1][j];
            uint32_t a, b;
            P.get(k, a, b);
            const uint32_t c = wave_pair(L, a, b);
            if (__lane_id() == 0) out[k] = c;
        }
    }
    // Thread bucket: thread per pair, a wave's worth per grab.
    const uint64_t n0 = rtgd::uniform_u64(B.count[0]);
    while (true) {
        uint64_t g = 0;
        if (__lane_id() == 0) g = atomicAdd(&B.count[3], 64ull);
        const uint64_t base = rtgd::uniform_u64(g);
        if (base >= n0) break;
        const uint64_t i = base + __lane_id();
        if (i < n0) {
            const uint32_t k = B.list[0][i];
            uint32_t a, b;
            P.get(k, a, b);
            out[k] = thread_pair(L, a, b);
        }
    }
}

// ---- probe items (latency path) ----
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
            const uint32_t b = count_below_lift(l, nl, v);
            c = (b < nl && l[b] == v) ? 1u : 0u;
        }
        c = rtgd::wave_sum_u32(c);
        if (__lane_id() == 0 && c) atomicAdd(&out[item.x], c);
    }
}

__global__ void si_alg_bytes(Lists L, PairSrc P, int64_t npairs, unsigned long long *acc) {
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
    s = rtgd::wave_sum_u64(s);
    if (__lane_id() == 0 && s) atomicAdd(acc, (unsigned long long)s);
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
    Buckets B{};
    if (method == RTG_SI_BATCHED) {
        const size_t bytes = Arena::round(sizeof(uint32_t) * (size_t)npairs) * 3 + 1024;
        HIP_TRY(hipMalloc(&fb.p, bytes));
        Arena ar{(char *)fb.p, 0, bytes};
        for (auto &l : B.list) l = ar.take<uint32_t>((size_t)npairs);
        B.count = ar.take<unsigned long long>(8);
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
            si_reset<<<1, 64, 0, st>>>(B.count);
            si_classify<<<grid_for(ctx, npairs, 16), kThreads, 0, st>>>(L, P, npairs, B, 16u, 2048u, d_out);
            si_drain<<<(unsigned)((int64_t)ctx->cu_count * 8), kThreads, 0, st>>>(L, P, B, d_out);
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
        si_alg_bytes<<<grid_for(ctx, npairs, 8), kThreads, 0, st>>>(L, P, npairs, (unsigned long long *)s);
        HIP_TRY(hipGetLastError());
        unsigned long long ab = 0, cnt[3] = {0, 0, 0};
        HIP_TRY(hipMemcpyAsync(&ab, s, sizeof(ab), hipMemcpyDeviceToHost, st));
        if (method == RTG_SI_BATCHED) HIP_TRY(hipMemcpyAsync(cnt, B.count, sizeof(cnt), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        stats->t_kernel_ms = elapsed_ms(ctx->ev[0], ctx->ev[1]);
        stats->alg_bytes = ab;
        if (method == RTG_SI_BATCHED) {
            stats->nbuckets = 3;
            for (int q = 0; q < 3; ++q) stats->bucket_items[q] = (int64_t)cnt[q];
        }
    } else {
        HIP_TRY(hipStreamSynchronize(st));
    }
    return RTG_OK;
}

}  // namespace

namespace rtg {

size_t si_batched_work_bytes(int64_t npairs) {
    return Arena::round(sizeof(uint32_t) * (size_t)std::max<int64_t>(npairs, 1)) * 3 + 1024;
}

int si_launch_batched(rtg_ctx *ctx, hipStream_t st, const uint32_t *d_off, const uint32_t *d_el,
                      const uint32_t *d_pairs, int64_t npairs, void *work, uint32_t *d_out) {
    if (npairs <= 0) return RTG_OK;
    Lists L{d_off, d_el};
    PairSrc P{d_pairs, 0, 1};
    Arena ar{(char *)work, 0, si_batched_work_bytes(npairs)};
    Buckets B{};
    for (auto &l : B.list) l = ar.take<uint32_t>((size_t)npairs);
    B.count = ar.take<unsigned long long>(8);
    si_reset<<<1, 64, 0, st>>>(B.count);
    si_classify<<<grid_for(ctx, npairs, 16), kThreads, 0, st>>>(L, P, npairs, B, 16u, 2048u, d_out);
    const unsigned drain = (unsigned)std::min<int64_t>((int64_t)ctx->cu_count * 8, std::max<int64_t>(npairs, 1));
    si_drain<<<drain, kThreads, 0, st>>>(L, P, B, d_out);
    HIP_TRY(hipGetLastError());
    return RTG_OK;
}

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
