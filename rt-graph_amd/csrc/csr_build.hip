// Device build of the degree-oriented CSR from an edge-pair list: the reference's
// Graph::ReadBinaryEdges relabel + Graph::TCPreprocessing + Graph::ConvertToCSR
// (tc/graph/graph.cpp:40-62, 104-177, 71-102) as data-parallel passes over HBM:
//
//   degree    raw degree of every input id (duplicates and self-loops count,
//             graph.cpp:109-114) by a partitioned LDS histogram: endpoints grouped by id range,
//             each range counted in LDS, one global add per non-zero bin and block
//   renumber  the relabel (present = degree > 0, dense and order-preserving, graph.cpp:40-62)
//             and the degree<=1 removal (x - removes[x], graph.cpp:126-152) are both
//             order-preserving renumberings, so the final id of a kept vertex is one exclusive
//             scan of (degree > 1) over the input ids
//   orient    src->dst iff deg[src] < deg[dst] or (equal and src < dst) (graph.cpp:116-124),
//             on the input ids (same order as the relabelled ones), kept pairs packed with
//             their final ids into 64-bit keys (src << b | dst), b = bits(n')
//   sort      stable LSD radix sort on 2b bits (replaces ThrustSortEdgesInBatch)
//   csr       drop self-loops and repeats of the previous key (graph.cpp:159-173), compact
//             by exclusive scan, write adjs and the row offsets (graph.cpp:71-102)
//
// The output arrays are byte-identical to the reference semantics (checked against the
// CPU oracle). Without RTG_PAIRS_ORIENT the same passes build the symmetric simple CSR
// (both directions of every non-loop pair, sorted, deduplicated, no vertex removed).
#include "graph.hpp"

#include <cstring>

namespace {

constexpr int kT = 256;

struct MinMax {
    int32_t mn;
    int32_t mx;
};

// Smallest and largest id: four ids per 16-byte load (the pair buffer is 16-byte aligned and
// holds an even count), the odd tail of n % 4 ids by the first threads.
__global__ void minmax_kernel(const int32_t *__restrict__ v, int64_t n, MinMax *out) {
    int32_t mn = INT32_MAX, mx = INT32_MIN;
    const int64_t n4 = n / 4;
    const int4 *v4 = reinterpret_cast<const int4 *>(v);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {  // four independent 16-byte loads in flight
        int4 q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = v4[i + k * stride];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            mn = min(mn, min(min(q[k].x, q[k].y), min(q[k].z, q[k].w)));
            mx = max(mx, max(max(q[k].x, q[k].y), max(q[k].z, q[k].w)));
        }
    }
    for (; i < n4; i += stride) {
        const int4 q = v4[i];
        mn = min(mn, min(min(q.x, q.y), min(q.z, q.w)));
        mx = max(mx, max(max(q.x, q.y), max(q.z, q.w)));
    }
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n - 4 * n4) {
        const int32_t x = v[4 * n4 + t];
        mn = min(mn, x);
        mx = max(mx, x);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, __shfl_xor(mn, off, 64));
        mx = max(mx, __shfl_xor(mx, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&out->mn, mn);
        atomicMax(&out->mx, mx);
    }
}

// Raw degree by one device-scope atomic per endpoint (id domains too large to partition).
__global__ void raw_degree(const int32_t *__restrict__ v, int64_t n, uint32_t *__restrict__ deg) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&deg[v[i]], 1u);
}

// ---- raw degree by a partitioned LDS histogram ----
// The endpoints are first grouped by id >> kPartBits (one counting-sort pass: per-tile bin
// histograms in LDS, a scan, a scatter through LDS cursors), then each block counts a
// contiguous run of one partition's endpoints into an LDS histogram of the partition's
// 2^kPartBits ids and adds its non-zero bins to the global degrees. One device-scope atomic per
// (block, bin) instead of one per endpoint: R-MAT hubs, whose same-address atomics serialise,
// are counted in LDS.
constexpr int kPartBits = 14;                 // 16,384 ids (64 KiB of LDS counters) per partition
constexpr int64_t kMaxPartitions = 1024;      // domains up to 2^24 ids; larger ones use raw_degree
constexpr int kPartItems = 32;                // endpoints per thread in the partitioning tiles
constexpr int64_t kPartTile = (int64_t)kT * kPartItems;
constexpr int64_t kCountChunk = 1 << 17;      // endpoints per block in the counting pass

__global__ __launch_bounds__(kT) void part_hist(const int32_t *__restrict__ v, int64_t n, int32_t nparts,
                                                uint32_t *__restrict__ hist, uint32_t ntiles) {
    __shared__ uint32_t h[kMaxPartitions];
    for (int p = threadIdx.x; p < nparts; p += kT) h[p] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kPartTile;
#pragma unroll 8
    for (int i = 0; i < kPartItems; ++i) {
        const int64_t idx = base + (int64_t)i * kT + threadIdx.x;
        if (idx < n) atomicAdd(&h[(uint32_t)v[idx] >> kPartBits], 1u);
    }
    __syncthreads();
    for (int p = threadIdx.x; p < nparts; p += kT) hist[(size_t)p * ntiles + blockIdx.x] = h[p];
}

// Scatters the tile's endpoints to their partitions' runs (starts = scanned histogram): the tile
// is first grouped by partition in LDS (ranks from LDS atomics, then the tile's own bin offsets),
// so that each partition's share of the tile is written out as one contiguous run. The order
// inside a run is arbitrary, which a histogram does not mind.
__global__ __launch_bounds__(kT) void part_scatter(const int32_t *__restrict__ v, int64_t n, int32_t nparts,
                                                   const uint32_t *__restrict__ starts, uint32_t ntiles,
                                                   int32_t *__restrict__ out) {
    __shared__ uint32_t cnt[kMaxPartitions];     // the tile's count per partition, then its offset
    __shared__ uint32_t gstart[kMaxPartitions];  // where the tile's run of each partition starts in out
    __shared__ uint32_t wsum[kT / 64];
    __shared__ int32_t keys[kPartTile];
    constexpr int kBinsPerThread = kMaxPartitions / kT;
    for (int p = threadIdx.x; p < kMaxPartitions; p += kT) cnt[p] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kPartTile;
    int32_t x[kPartItems];
    uint32_t rank[kPartItems];
#pragma unroll
    for (int i = 0; i < kPartItems; ++i) {
        const int64_t idx = base + (int64_t)i * kT + threadIdx.x;
        x[i] = idx < n ? v[idx] : -1;
        rank[i] = x[i] >= 0 ? atomicAdd(&cnt[(uint32_t)x[i] >> kPartBits], 1u) : 0u;
    }
    __syncthreads();
    // exclusive scan of the tile's counts: kBinsPerThread consecutive bins per thread
    uint32_t c[kBinsPerThread], sum = 0;
#pragma unroll
    for (int k = 0; k < kBinsPerThread; ++k) {
        c[k] = cnt[threadIdx.x * kBinsPerThread + k];
        sum += c[k];
    }
    const uint32_t incl = rtgd::wave_inclusive_scan(sum);
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t off = incl - sum;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) off += wsum[w];
#pragma unroll
    for (int k = 0; k < kBinsPerThread; ++k) {
        const int p = threadIdx.x * kBinsPerThread + k;
        cnt[p] = off;
        if (p < nparts) gstart[p] = starts[(size_t)p * ntiles + blockIdx.x];
        off += c[k];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPartItems; ++i)
        if (x[i] >= 0) keys[cnt[(uint32_t)x[i] >> kPartBits] + rank[i]] = x[i];
    __syncthreads();
    const int32_t count = (int32_t)min<int64_t>(kPartTile, n - base);
    for (int32_t j = threadIdx.x; j < count; j += kT) {
        const int32_t y = keys[j];
        const uint32_t p = (uint32_t)y >> kPartBits;
        out[gstart[p] + ((uint32_t)j - cnt[p])] = y;
    }
}

// part_start[p] = first position of partition p in the grouped array, part_start[nparts] = n.
__global__ void part_starts(const uint32_t *__restrict__ scanned, uint32_t ntiles, int32_t nparts, int64_t n,
                            uint32_t *__restrict__ part_start) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p <= nparts; p += gridDim.x * blockDim.x)
        part_start[p] = p < nparts ? scanned[(size_t)p * ntiles] : (uint32_t)n;
}

// Block b counts the grouped endpoints [b * kCountChunk, (b + 1) * kCountChunk), one partition
// run at a time, in an LDS histogram, and adds each non-zero bin to deg (sized nparts << kPartBits).
__global__ __launch_bounds__(kT) void part_count(const int32_t *__restrict__ grouped, int64_t n,
                                                 const uint32_t *__restrict__ part_start, int32_t nparts,
                                                 uint32_t *__restrict__ deg) {
    __shared__ uint32_t h[1 << kPartBits];
    const int64_t c0 = (int64_t)blockIdx.x * kCountChunk;
    const int64_t c1 = min(n, c0 + kCountChunk);
    // first partition whose run reaches past c0 (uniform: every thread computes the same)
    int32_t lo = 0, hi = nparts;  // invariant: part_start[lo] <= c0 < part_start[hi]
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if ((int64_t)part_start[mid] <= c0) lo = mid;
        else hi = mid;
    }
    for (int32_t p = (int32_t)rtgd::uniform_u32((uint32_t)lo); p < nparts; ++p) {
        const int64_t s0 = part_start[p], s1 = part_start[p + 1];
        const int64_t r0 = rtgd::uniform_u64((uint64_t)max(c0, s0)), r1 = rtgd::uniform_u64((uint64_t)min(c1, s1));
        if (s0 >= c1) break;
        if (r1 <= r0) continue;
        for (int j = threadIdx.x; j < (1 << kPartBits); j += kT) h[j] = 0u;
        __syncthreads();
        for (int64_t i = r0 + threadIdx.x; i < r1; i += kT) atomicAdd(&h[(uint32_t)grouped[i] & ((1u << kPartBits) - 1u)], 1u);
        __syncthreads();
        uint32_t *dst = deg + ((size_t)p << kPartBits);
        for (int j = threadIdx.x; j < (1 << kPartBits); j += kT) {
            const uint32_t c = h[j];
            if (c) atomicAdd(&dst[j], c);
        }
        __syncthreads();  // h is cleared for the next run
    }
}

// Vertex map of the original id domain: vinfo[x] = {raw degree, new id}, new id = exclusive
// scan of keep; keep = degree > 1 when orienting (the relabel and the degree<=1 removal
// of graph.cpp:40-62, 126-152 in one renumbering: both are order-preserving), else degree > 0
// (relabel only). present counts ids of degree > 0 (the relabelled vertex count).
__global__ void keep_flags(const uint32_t *__restrict__ deg, int64_t n, uint32_t min_deg, uint32_t *__restrict__ keep,
                           unsigned long long *present) {
    uint64_t cnt = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t d = deg[i];
        keep[i] = d >= min_deg ? 1u : 0u;
        cnt += d > 0 ? 1u : 0u;
    }
    cnt = rtgd::wave_sum_u64(cnt);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(present, (unsigned long long)cnt);
}

__global__ void vertex_info(const uint32_t *__restrict__ deg, const uint32_t *__restrict__ newid, int64_t n,
                            uint2 *__restrict__ vinfo) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        vinfo[i] = make_uint2(deg[i], newid[i]);
}

// Block-aggregated append: each thread holds up to kPer keys; one exclusive scan over the block
// and one atomic per block place them contiguously in out[] (order across blocks is arbitrary;
// the keys are sorted afterwards).
constexpr int kPer = 8;
constexpr int kPackThreads = 256;

template <int kMax>
__device__ __forceinline__ void block_append(const uint64_t (&keys)[kMax], uint32_t mask, uint64_t *out,
                                             unsigned long long *counter) {
    const int cnt = __popc(mask);
    __shared__ uint32_t wsum[kPackThreads / 64];
    __shared__ unsigned long long base_s;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t incl = rtgd::wave_inclusive_scan((uint32_t)cnt);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kPackThreads / 64; ++w) {
        const uint32_t v = wsum[w];
        woff += (w < wave) ? v : 0u;
        tot += v;
    }
    if (threadIdx.x == 0) base_s = tot ? atomicAdd(counter, (unsigned long long)tot) : 0ull;
    __syncthreads();
    const unsigned long long base = base_s + woff + (incl - (uint32_t)cnt);
#pragma unroll
    for (int i = 0; i < kMax; ++i)
        if ((mask >> i) & 1u) out[base + __popc(mask & ((1u << i) - 1u))] = keys[i];
    __syncthreads();  // wsum and base_s are reused by the next tile
}

// Orientation, removal and key packing (TCPreprocessing) on the input's own ids: vinfo[x] =
// {raw degree, final id}. The orientation compares (degree, id) — the relabel preserves order,
// so the input ids order ties as the relabelled ones do (graph.cpp:116-124) — and the kept
// endpoints (degree > 1) are renumbered by their final ids (graph.cpp:126-152).
__global__ __launch_bounds__(kPackThreads) void orient_pack(const int32_t *__restrict__ pairs, int64_t np,
                                                            const uint2 *__restrict__ vinfo, int bits,
                                                            uint64_t *__restrict__ keys, unsigned long long *counter) {
    const int64_t tile = (int64_t)kPackThreads * kPer;
    for (int64_t t0 = (int64_t)blockIdx.x * tile; t0 < np; t0 += (int64_t)gridDim.x * tile) {
        uint64_t k[kPer];
        uint32_t mask = 0;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int64_t idx = t0 + (int64_t)i * kPackThreads + threadIdx.x;  // coalesced across the block
            k[i] = 0;
            if (idx < np) {
                const int2 pr = reinterpret_cast<const int2 *>(pairs)[idx];
                int32_t s = pr.x, d = pr.y;
                uint2 is = vinfo[s], id = vinfo[d];
                if (is.x > id.x || (is.x == id.x && s > d)) {
                    const uint2 t = is;
                    is = id;
                    id = t;
                }
                if (is.x > 1u && id.x > 1u) {
                    k[i] = ((uint64_t)is.y << bits) | (uint64_t)id.y;
                    mask |= 1u << i;
                }
            }
        }
        block_append<kPer>(k, mask, keys, counter);
    }
}

// Symmetric mode: both directions of every non-loop pair, ids mapped through newid when given
// (the relabel), else kept.
__global__ __launch_bounds__(kPackThreads) void symmetric_pack(const int32_t *__restrict__ pairs, int64_t np,
                                                               const uint32_t *__restrict__ newid, int bits,
                                                               uint64_t *__restrict__ keys, unsigned long long *counter) {
    const int64_t tile = (int64_t)kPackThreads * kPer;
    for (int64_t t0 = (int64_t)blockIdx.x * tile; t0 < np; t0 += (int64_t)gridDim.x * tile) {
        uint64_t k[2 * kPer];
        uint32_t mask = 0;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int64_t idx = t0 + (int64_t)i * kPackThreads + threadIdx.x;  // coalesced across the block
            k[2 * i] = k[2 * i + 1] = 0;
            if (idx < np) {
                const int2 pr = reinterpret_cast<const int2 *>(pairs)[idx];
                if (pr.x != pr.y) {
                    const uint64_t s = newid ? newid[pr.x] : (uint32_t)pr.x, d = newid ? newid[pr.y] : (uint32_t)pr.y;
                    k[2 * i] = (s << bits) | d;
                    k[2 * i + 1] = (d << bits) | s;
                    mask |= 3u << (2 * i);
                }
            }
        }
        block_append<2 * kPer>(k, mask, keys, counter);
    }
}

__global__ void dedup_flags(const uint64_t *__restrict__ keys, int64_t n, int bits,
                            uint32_t *__restrict__ flag) {
    const uint64_t lo = (1ull << bits) - 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        const bool loop = (k >> bits) == (k & lo);
        const bool repeat = i > 0 && keys[i - 1] == k;
        flag[i] = (!loop && !repeat) ? 1u : 0u;
    }
}

__global__ void dedup_scatter(const uint64_t *__restrict__ keys, int64_t n, int bits,
                              const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
                              int32_t *__restrict__ src_c, int32_t *__restrict__ adjs) {
    const uint64_t lo = (1ull << bits) - 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (flag[i]) {
            const uint64_t k = keys[i];
            const uint32_t p = pos[i];
            src_c[p] = (int32_t)(k >> bits);
            adjs[p] = (int32_t)(k & lo);
        }
    }
}

// offsets[u] = first compacted index whose source is >= u (rows are sorted by source).
__global__ void fill_offsets(const int32_t *__restrict__ src_c, int64_t m, int32_t n,
                             int32_t *__restrict__ offsets) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = src_c[j];
        const int32_t p = j == 0 ? -1 : src_c[j - 1];
        for (int32_t u = p + 1; u <= s; ++u) offsets[u] = (int32_t)j;
        if (j == m - 1)
            for (int32_t u = s + 1; u <= n; ++u) offsets[u] = (int32_t)m;
    }
}

__global__ void fill_i32(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void max_row_kernel(const int32_t *__restrict__ offsets, int32_t n, int32_t *out) {
    int32_t mx = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        mx = max(mx, offsets[i + 1] - offsets[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, mx);
}

// CSR validation for rtg_graph_from_csr: err bit 1 = offsets not monotone / bad bounds,
// bit 2 = column out of range, bit 4 = row not strictly increasing (duplicate or unsorted),
// bit 8 = self-loop.
__global__ void validate_csr(const int32_t *__restrict__ offsets, const int32_t *__restrict__ adjs,
                             int32_t n, int64_t m, int32_t *err) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = offsets[u], e = offsets[u + 1];
        int32_t bad = 0;
        if (b > e || b < 0 || (int64_t)e > m) bad |= 1;
        else {
            int32_t prev = -1;
            for (int32_t j = b; j < e; ++j) {
                const int32_t v = adjs[j];
                if (v < 0 || v >= n) bad |= 2;
                if (v <= prev) bad |= 4;
                if (v == (int32_t)u) bad |= 8;
                prev = v;
            }
        }
        if (bad) atomicOr(err, bad);
    }
}

unsigned grid_for(rtg_ctx *ctx, int64_t n) {
    int64_t want = rtg::ceil_div(n, kT);
    int64_t cap = (int64_t)ctx->cu_count * 16;
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (unsigned)want;
}

int bits_for(int64_t n) {
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) < n) ++b;
    return b;
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

namespace rtg {

int csr_max_row(rtg_ctx *ctx, const int32_t *d_offsets, int32_t n, int32_t *out) {
    *out = 0;
    if (n <= 0) return RTG_OK;
    void *s = nullptr;
    RTG_TRY(ctx_scratch(ctx, 256, &s));
    int32_t *d = (int32_t *)s;
    HIP_TRY(hipMemsetAsync(d, 0, sizeof(int32_t), ctx->stream));
    max_row_kernel<<<grid_for(ctx, n), kT, 0, ctx->stream>>>(d_offsets, n, d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, d, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTG_OK;
}

}  // namespace rtg

namespace rtg {
bool graph_oriented(const rtg_graph *g) { return g && g->oriented; }
void graph_set_oriented(rtg_graph *g, bool oriented) {
    if (g) g->oriented = oriented;
}
}  // namespace rtg

extern "C" {

RTG_API int rtg_graph_from_pairs(rtg_ctx *ctx, const int32_t *pairs, int64_t npairs, uint32_t flags,
                                 rtg_graph **out, rtg_build_stats *stats) {
    using namespace rtg;
    RTG_TRY(ctx_enter(ctx));
    if (!out) return rtg_fail(RTG_EINVAL, "rtg_graph_from_pairs: null out");
    *out = nullptr;
    if (npairs < 0 || (!pairs && npairs > 0)) return rtg_fail(RTG_EINVAL, "rtg_graph_from_pairs: bad pairs");
    const bool orient = flags & RTG_PAIRS_ORIENT;
    const int64_t max_keys = orient ? npairs : 2 * npairs;
    if (max_keys >= ((int64_t)1 << 31)) return rtg_fail(RTG_EINVAL, "rtg_graph_from_pairs: too many pairs (limit 2^31 keys)");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t st = ctx->stream;
    hipEvent_t *ev = ctx->ev;
    const int64_t nv = 2 * npairs;  // endpoint count

    // Working copy of the pairs.
    DevBuf pbuf;
    HIP_TRY(hipMalloc(&pbuf.p, sizeof(int32_t) * (size_t)std::max<int64_t>(nv, 1)));
    int32_t *d_pairs = (int32_t *)pbuf.p;
    if (npairs > 0)
        HIP_TRY(hipMemcpyAsync(d_pairs, pairs, sizeof(int32_t) * (size_t)nv,
                               (flags & RTG_PAIRS_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(ev[0], st));

    // ---- domain, raw degree, renumbering (graph.cpp:40-62, 109-114, 126-152) ----
    MinMax mm{0, -1};
    if (npairs > 0) {
        void *s = nullptr;
        RTG_TRY(ctx_scratch(ctx, 256, &s));
        MinMax init{INT32_MAX, INT32_MIN};
        HIP_TRY(hipMemcpyAsync(s, &init, sizeof(init), hipMemcpyHostToDevice, st));
        minmax_kernel<<<grid_for(ctx, nv), kT, 0, st>>>(d_pairs, nv, (MinMax *)s);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&mm, s, sizeof(mm), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (mm.mn < 0) return rtg_fail(RTG_EINVAL, "rtg_graph_from_pairs: negative vertex id");
    }
    const int64_t domain_ids = npairs > 0 ? (int64_t)mm.mx + 1 : 0;  // the input's id range
    if (domain_ids >= ((int64_t)1 << 31) - 1) return rtg_fail(RTG_EINVAL, "rtg_graph_from_pairs: vertex ids too large");
    const int64_t domain = std::max<int64_t>(domain_ids, 1);
    const bool relabel = (flags & RTG_PAIRS_RELABEL) != 0;
    const bool need_degree = orient || relabel;  // a symmetric CSR without relabel keeps its ids
    const int64_t nparts = ceil_div(domain, (int64_t)1 << kPartBits);
    const bool partitioned = need_degree && npairs > 0 && nparts <= kMaxPartitions;
    const int64_t deg_len = partitioned ? (nparts << kPartBits) : domain;
    const uint32_t ntiles = (uint32_t)std::max<int64_t>(1, ceil_div(nv, kPartTile));
    const int64_t nhist = partitioned ? nparts * (int64_t)ntiles : 1;
    DevBuf vbuf;  // vertex-sized scratch: degree, keep flags / new ids, vertex info; endpoint grouping
    const size_t vbytes = Arena::round(4 * (size_t)deg_len) + Arena::round(4 * (size_t)(domain + 1)) * 2 +
                          Arena::round(8 * (size_t)domain) + scan_work_bytes(std::max<int64_t>(domain + 1, nhist)) +
                          (partitioned ? Arena::round(4 * (size_t)nv) + Arena::round(4 * (size_t)nhist) +
                                             Arena::round(4 * (size_t)(nparts + 1))
                                       : 0) +
                          8192;
    HIP_TRY(hipMalloc(&vbuf.p, vbytes));
    Arena va{(char *)vbuf.p, 0, vbytes};
    uint32_t *deg = va.take<uint32_t>((size_t)deg_len);
    uint32_t *keep = va.take<uint32_t>((size_t)domain + 1);
    uint32_t *newid = va.take<uint32_t>((size_t)domain + 1);
    uint2 *vinfo = va.take<uint2>((size_t)domain);
    unsigned long long *d_small = va.take<unsigned long long>(8);  // [0] present ids, [1] kept (u32)
    void *vscan = va.take<char>(scan_work_bytes(std::max<int64_t>(domain + 1, nhist)));
    int64_t nodes = domain_ids, n_out = domain_ids;
    if (need_degree && npairs > 0) {
        HIP_TRY(hipMemsetAsync(deg, 0, 4 * (size_t)deg_len, st));
        HIP_TRY(hipMemsetAsync(d_small, 0, sizeof(unsigned long long) * 8, st));
        if (partitioned) {
            int32_t *grouped = va.take<int32_t>((size_t)nv);
            uint32_t *hist = va.take<uint32_t>((size_t)nhist);
            uint32_t *pstart = va.take<uint32_t>((size_t)nparts + 1);
            part_hist<<<ntiles, kT, 0, st>>>(d_pairs, nv, (int32_t)nparts, hist, ntiles);
            HIP_TRY(hipGetLastError());
            RTG_TRY(scan_exclusive_u32_w(ctx, hist, hist, nhist, nullptr, vscan));
            part_scatter<<<ntiles, kT, 0, st>>>(d_pairs, nv, (int32_t)nparts, hist, ntiles, grouped);
            HIP_TRY(hipGetLastError());
            part_starts<<<(unsigned)ceil_div(nparts + 1, 256), 256, 0, st>>>(hist, ntiles, (int32_t)nparts, nv, pstart);
            HIP_TRY(hipGetLastError());
            part_count<<<(unsigned)ceil_div(nv, kCountChunk), kT, 0, st>>>(grouped, nv, pstart, (int32_t)nparts, deg);
            HIP_TRY(hipGetLastError());
        } else {
            raw_degree<<<grid_for(ctx, nv), kT, 0, st>>>(d_pairs, nv, deg);
            HIP_TRY(hipGetLastError());
        }
        keep_flags<<<grid_for(ctx, domain), kT, 0, st>>>(deg, domain, orient ? 2u : 1u, keep, d_small);
        HIP_TRY(hipGetLastError());
        RTG_TRY(scan_exclusive_u32_w(ctx, keep, newid, domain, reinterpret_cast<uint32_t *>(d_small + 1), vscan));
        if (orient) {
            vertex_info<<<grid_for(ctx, domain), kT, 0, st>>>(deg, newid, domain, vinfo);
            HIP_TRY(hipGetLastError());
        }
        unsigned long long hs[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(hs, d_small, sizeof(hs), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const int64_t present = (int64_t)hs[0], kept = (int64_t)(uint32_t)hs[1];
        nodes = relabel ? present : domain_ids;
        n_out = orient ? kept : present;
    }
    HIP_TRY(hipEventRecord(ev[1], st));

    // ---- orientation and key packing (graph.cpp:116-124, 126-152), or both directions ----
    int bits = bits_for(std::max<int64_t>(n_out, 2));
    DevBuf kbuf;
    const int64_t kcap = std::max<int64_t>(max_keys, 1);
    const size_t kbytes = Arena::round(8 * (size_t)kcap) * 2 + radix_sort_work_bytes(kcap) +
                          Arena::round(4 * (size_t)kcap) * 4 + scan_work_bytes(kcap) + 4096;
    HIP_TRY(hipMalloc(&kbuf.p, kbytes));
    Arena ka{(char *)kbuf.p, 0, kbytes};
    uint64_t *keys0 = ka.take<uint64_t>(kcap);
    uint64_t *keys1 = ka.take<uint64_t>(kcap);
    void *sort_work = ka.take<char>(radix_sort_work_bytes(kcap));
    uint32_t *flag = ka.take<uint32_t>(kcap);
    uint32_t *pos = ka.take<uint32_t>(kcap);
    int32_t *src_c = ka.take<int32_t>(kcap);
    unsigned long long *counter = ka.take<unsigned long long>(4);
    uint32_t *d_m = ka.take<uint32_t>(4);
    void *kscan = ka.take<char>(scan_work_bytes(kcap));

    HIP_TRY(hipMemsetAsync(counter, 0, sizeof(unsigned long long), st));
    if (npairs > 0) {
        const unsigned pgrid = (unsigned)std::min<int64_t>(ceil_div(npairs, (int64_t)kPackThreads * kPer), (int64_t)ctx->cu_count * 8);
        if (orient)
            orient_pack<<<pgrid, kPackThreads, 0, st>>>(d_pairs, npairs, vinfo, bits, keys0, counter);
        else
            symmetric_pack<<<pgrid, kPackThreads, 0, st>>>(d_pairs, npairs, relabel ? newid : nullptr, bits, keys0, counter);
        HIP_TRY(hipGetLastError());
    }
    unsigned long long nkeys = 0;
    HIP_TRY(hipMemcpyAsync(&nkeys, counter, sizeof(nkeys), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipEventRecord(ev[2], st));

    // ---- sort (thrust_helper.cu:16-42) ----
    uint64_t *kk[2] = {keys0, keys1};
    int which = 0, passes = 0;
    RTG_TRY(radix_sort_u64(ctx, kk, (int64_t)nkeys, 2 * bits, &which, &passes, sort_work));
    uint64_t *sorted = kk[which];
    HIP_TRY(hipEventRecord(ev[3], st));

    // ---- dedup + CSR (graph.cpp:159-173, 71-102) ----
    auto *g = new rtg_graph();
    g->ctx = ctx;
    g->n = (int32_t)n_out;
    g->oriented = orient;
    uint32_t m = 0;
    if (nkeys > 0) {
        dedup_flags<<<grid_for(ctx, nkeys), kT, 0, st>>>(sorted, nkeys, bits, flag);
        HIP_TRY(hipGetLastError());
        RTG_TRY(scan_exclusive_u32_w(ctx, flag, pos, nkeys, d_m, kscan));
        HIP_TRY(hipMemcpyAsync(&m, d_m, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    g->m = m;
    if (hipMalloc(&g->d_offsets, sizeof(int32_t) * (size_t)(n_out + 1)) != hipSuccess ||
        hipMalloc(&g->d_adjs, sizeof(int32_t) * (size_t)std::max<uint32_t>(m, 1)) != hipSuccess) {
        rtg_graph_destroy(g);
        return rtg_fail(RTG_ENOMEM, "rtg_graph_from_pairs: CSR allocation failed");
    }
    if (m > 0) {
        dedup_scatter<<<grid_for(ctx, nkeys), kT, 0, st>>>(sorted, nkeys, bits, flag, pos, src_c, g->d_adjs);
        fill_offsets<<<grid_for(ctx, m), kT, 0, st>>>(src_c, m, (int32_t)n_out, g->d_offsets);
    } else {
        fill_i32<<<grid_for(ctx, n_out + 1), kT, 0, st>>>(g->d_offsets, n_out + 1, 0);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[4], st));
    HIP_TRY(hipStreamSynchronize(st));
    int st2 = csr_max_row(ctx, g->d_offsets, g->n, &g->max_row);
    if (st2 != RTG_OK) {
        rtg_graph_destroy(g);
        return st2;
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->t_relabel_ms = elapsed_ms(ev[0], ev[1]);
        stats->t_orient_ms = elapsed_ms(ev[1], ev[2]);
        stats->t_sort_ms = elapsed_ms(ev[2], ev[3]);
        stats->t_csr_ms = elapsed_ms(ev[3], ev[4]);
        stats->t_total_ms = elapsed_ms(ev[0], ev[4]);
        stats->raw_pairs = npairs;
        stats->valid_pairs = (int64_t)nkeys;
        stats->nodes_in = (int32_t)nodes;
        stats->sort_passes = passes;
    }
    *out = g;
    return RTG_OK;
}

RTG_API int rtg_graph_from_csr(rtg_ctx *ctx, const int32_t *offsets, const int32_t *adjs, int32_t n,
                               int64_t m, rtg_graph **out) {
    using namespace rtg;
    RTG_TRY(ctx_enter(ctx));
    if (!out || !offsets || n < 0 || m < 0 || (m > 0 && !adjs))
        return rtg_fail(RTG_EINVAL, "rtg_graph_from_csr: bad arguments");
    if (offsets[0] != 0 || offsets[n] != m) return rtg_fail(RTG_EINVAL, "rtg_graph_from_csr: offsets[0] != 0 or offsets[n] != m");
    if (m >= ((int64_t)1 << 31)) return rtg_fail(RTG_EINVAL, "rtg_graph_from_csr: m >= 2^31");
    std::lock_guard<std::mutex> lock(ctx->mu);
    auto *g = new rtg_graph();
    g->ctx = ctx;
    g->n = n;
    g->m = m;
    if (hipMalloc(&g->d_offsets, sizeof(int32_t) * (size_t)(n + 1)) != hipSuccess ||
        hipMalloc(&g->d_adjs, sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1)) != hipSuccess) {
        rtg_graph_destroy(g);
        return rtg_fail(RTG_ENOMEM, "rtg_graph_from_csr: allocation failed");
    }
    hipStream_t st = ctx->stream;
    int rc = RTG_OK;
    do {
        if (hipMemcpyAsync(g->d_offsets, offsets, sizeof(int32_t) * (size_t)(n + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
            (m > 0 && hipMemcpyAsync(g->d_adjs, adjs, sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice, st) != hipSuccess)) {
            rc = rtg_fail(RTG_EHIP, "rtg_graph_from_csr: copy failed");
            break;
        }
        void *s = nullptr;
        if ((rc = ctx_scratch(ctx, 256, &s)) != RTG_OK) break;
        int32_t *err = (int32_t *)s, herr = 0;
        if (hipMemsetAsync(err, 0, 4, st) != hipSuccess) { rc = rtg_fail(RTG_EHIP, "memset failed"); break; }
        if (n > 0) validate_csr<<<grid_for(ctx, n), kT, 0, st>>>(g->d_offsets, g->d_adjs, n, m, err);
        if (hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = rtg_fail(RTG_EHIP, "rtg_graph_from_csr: validation failed to run");
            break;
        }
        if (herr) {
            rc = rtg_fail(RTG_EINVAL, std::string("rtg_graph_from_csr: invalid CSR (") +
                                          ((herr & 1) ? "offsets " : "") + ((herr & 2) ? "column-range " : "") +
                                          ((herr & 4) ? "unsorted-or-duplicate " : "") + ((herr & 8) ? "self-loop" : "") + ")");
            break;
        }
        rc = csr_max_row(ctx, g->d_offsets, n, &g->max_row);
    } while (0);
    if (rc != RTG_OK) {
        rtg_graph_destroy(g);
        return rc;
    }
    *out = g;
    return RTG_OK;
}

RTG_API int rtg_graph_info(const rtg_graph *g, int32_t *n, int64_t *m, int32_t *max_row) {
    if (!g) return rtg_fail(RTG_EINVAL, "null graph");
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (max_row) *max_row = g->max_row;
    return RTG_OK;
}

RTG_API int rtg_graph_download(const rtg_graph *g, int32_t *offsets, int32_t *adjs) {
    if (!g) return rtg_fail(RTG_EINVAL, "null graph");
    RTG_TRY(rtg::ctx_enter(g->ctx));
    hipStream_t st = g->ctx->stream;
    if (offsets) HIP_TRY(hipMemcpyAsync(offsets, g->d_offsets, sizeof(int32_t) * (size_t)(g->n + 1), hipMemcpyDeviceToHost, st));
    if (adjs && g->m > 0) HIP_TRY(hipMemcpyAsync(adjs, g->d_adjs, sizeof(int32_t) * (size_t)g->m, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return RTG_OK;
}

RTG_API int rtg_graph_device_ptrs(const rtg_graph *g, const int32_t **d_offsets, const int32_t **d_adjs) {
    if (!g) return rtg_fail(RTG_EINVAL, "null graph");
    if (d_offsets) *d_offsets = g->d_offsets;
    if (d_adjs) *d_adjs = g->d_adjs;
    return RTG_OK;
}

RTG_API int rtg_graph_destroy(rtg_graph *g) {
    if (!g) return RTG_OK;
    if (g->ctx) (void)hipSetDevice(g->ctx->device);
    if (g->d_offsets) (void)hipFree(g->d_offsets);
    if (g->d_adjs) (void)hipFree(g->d_adjs);
    delete g;
    return RTG_OK;
}

}  // extern "C"
