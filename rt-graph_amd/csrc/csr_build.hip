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
//             on the input ids (same order as the relabelled ones), kept non-loop pairs packed
//             with their final ids into 64-bit keys (src << b | dst), b = bits(n')
//   sort      stable LSD radix sort on 2b bits (replaces ThrustSortEdgesInBatch)
//   csr       drop repeats of the previous key (graph.cpp:159-173): kept keys per 4096-key
//             tile, a scan of the tile counts, then one pass that ranks the kept keys inside
//             each tile and writes adjs and the row offsets (graph.cpp:71-102)
//
// The output arrays are byte-identical to the reference semantics (checked against the
// CPU oracle). Without RTG_PAIRS_ORIENT the same passes build the symmetric simple CSR
// (both directions of every non-loop pair, sorted, deduplicated, no vertex removed).
#include "graph.hpp"

#include <cstring>
#include <memory>

namespace {

constexpr int kT = 256;

struct MinMax {
    int32_t mn;
    int32_t mx;
};

// Smallest and largest id: four ids per 16-byte load (the pair buffer is 16-byte aligned and
// holds an even count), the odd tail of n % 4 ids by the first threads; one partial per block
// (no same-address atomics: thousands of waves updating one word serialise), reduced by
// minmax_final.
__global__ __launch_bounds__(kT) void minmax_kernel(const int32_t *__restrict__ v, int64_t n, MinMax *__restrict__ part) {
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
    __shared__ MinMax w[kT / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = MinMax{mn, mx};
    __syncthreads();
    if (threadIdx.x == 0) {
        MinMax r = w[0];
#pragma unroll
        for (int k = 1; k < kT / 64; ++k) {
            r.mn = min(r.mn, w[k].mn);
            r.mx = max(r.mx, w[k].mx);
        }
        part[blockIdx.x] = r;
    }
}

// One block: out = the min and max over the nb partials.
__global__ __launch_bounds__(kT) void minmax_final(const MinMax *__restrict__ part, int nb, MinMax *__restrict__ out) {
    int32_t mn = INT32_MAX, mx = INT32_MIN;
    for (int b = threadIdx.x; b < nb; b += kT) {
        mn = min(mn, part[b].mn);
        mx = max(mx, part[b].mx);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, __shfl_xor(mn, off, 64));
        mx = max(mx, __shfl_xor(mx, off, 64));
    }
    __shared__ MinMax w[kT / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = MinMax{mn, mx};
    __syncthreads();
    if (threadIdx.x == 0) {
        MinMax r = w[0];
#pragma unroll
        for (int k = 1; k < kT / 64; ++k) {
            r.mn = min(r.mn, w[k].mn);
            r.mx = max(r.mx, w[k].mx);
        }
        *out = r;
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
#pragma unroll
    for (int r = 0; r < kPartItems; r += 8) {  // eight loads in flight, then their LDS atomics
        int32_t y[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int64_t idx = base + (int64_t)(r + u) * kT + threadIdx.x;
            y[u] = idx < n ? v[idx] : -1;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (y[u] >= 0) atomicAdd(&h[(uint32_t)y[u] >> kPartBits], 1u);
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
    for (int i = 0; i < kPartItems; ++i) {  // every load issued before the first LDS atomic
        const int64_t idx = base + (int64_t)i * kT + threadIdx.x;
        x[i] = idx < n ? v[idx] : -1;
    }
#pragma unroll
    for (int i = 0; i < kPartItems; ++i) rank[i] = x[i] >= 0 ? atomicAdd(&cnt[(uint32_t)x[i] >> kPartBits], 1u) : 0u;
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
        // eight independent loads per round, then their LDS atomics (a load per trip followed
        // by its atomic would keep one load in flight per thread)
        constexpr int kU = 8;
        int64_t i = r0 + threadIdx.x;
        for (; i + (kU - 1) * kT < r1; i += kU * kT) {
            int32_t y[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) y[u] = grouped[i + u * kT];
#pragma unroll
            for (int u = 0; u < kU; ++u) atomicAdd(&h[(uint32_t)y[u] & ((1u << kPartBits) - 1u)], 1u);
        }
        for (; i < r1; i += kT) atomicAdd(&h[(uint32_t)grouped[i] & ((1u << kPartBits) - 1u)], 1u);
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
__global__ __launch_bounds__(kT) void keep_flags(const uint32_t *__restrict__ deg, int64_t n, uint32_t min_deg, uint32_t *__restrict__ keep,
                           unsigned long long *present) {
    uint64_t cnt = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t d = deg[i];
        keep[i] = d >= min_deg ? 1u : 0u;
        cnt += d > 0 ? 1u : 0u;
    }
    cnt = rtgd::wave_sum_u64(cnt);
    __shared__ unsigned long long w[kT / 64];  // one device atomic per block, not per wave
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = 0;
#pragma unroll
        for (int k = 0; k < kT / 64; ++k) c += w[k];
        if (c) atomicAdd(present, c);
    }
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
// endpoints (degree > 1) are renumbered by their final ids (graph.cpp:126-152). Self-loops,
// which the reference drops with the duplicates (graph.cpp:159-173), are dropped here, so the
// sorted keys hold none.
__global__ __launch_bounds__(kPackThreads) void orient_pack(const int32_t *__restrict__ pairs, int64_t np,
                                                            const uint2 *__restrict__ vinfo, int bits,
                                                            uint64_t *__restrict__ keys, unsigned long long *counter) {
    const int64_t tile = (int64_t)kPackThreads * kPer;
    for (int64_t t0 = (int64_t)blockIdx.x * tile; t0 < np; t0 += (int64_t)gridDim.x * tile) {
        // Three phases, so that the loads of all kPer items are in flight together: the pairs,
        // then the 2 * kPer vertex gathers, then the orientation (an item past the end reads
        // pair 0 and is masked off).
        int2 pr[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int64_t idx = t0 + (int64_t)i * kPackThreads + threadIdx.x;  // coalesced across the block
            pr[i] = reinterpret_cast<const int2 *>(pairs)[idx < np ? idx : 0];
        }
        uint2 vs[kPer], vd[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            vs[i] = vinfo[pr[i].x];
            vd[i] = vinfo[pr[i].y];
        }
        uint64_t k[kPer];
        uint32_t mask = 0;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int64_t idx = t0 + (int64_t)i * kPackThreads + threadIdx.x;
            const int32_t s = pr[i].x, d = pr[i].y;
            uint2 is = vs[i], id = vd[i];
            if (is.x > id.x || (is.x == id.x && s > d)) {
                const uint2 t = is;
                is = id;
                id = t;
            }
            k[i] = ((uint64_t)is.y << bits) | (uint64_t)id.y;
            if (idx < np && s != d && is.x > 1u && id.x > 1u)  // a self-loop counts in the degree only
                mask |= 1u << i;
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
        // pairs first, then the id gathers of all items (an item past the end reads pair 0 and is
        // masked off), so that each phase's loads are in flight together
        int2 pr[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int64_t idx = t0 + (int64_t)i * kPackThreads + threadIdx.x;  // coalesced across the block
            pr[i] = reinterpret_cast<const int2 *>(pairs)[idx < np ? idx : 0];
        }
        uint64_t k[2 * kPer];
        uint32_t mask = 0;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int64_t idx = t0 + (int64_t)i * kPackThreads + threadIdx.x;
            const uint64_t s = newid ? newid[pr[i].x] : (uint32_t)pr[i].x, d = newid ? newid[pr[i].y] : (uint32_t)pr[i].y;
            k[2 * i] = (s << bits) | d;
            k[2 * i + 1] = (d << bits) | s;
            if (idx < np && pr[i].x != pr[i].y) mask |= 3u << (2 * i);
        }
        block_append<2 * kPer>(k, mask, keys, counter);
    }
}

// ---- dedup + CSR (graph.cpp:159-173, 71-102): the keys are sorted and hold no self-loops, so
// key i is kept iff it differs from key i - 1 ----
constexpr int kDedupItems = 16;
constexpr int64_t kDedupTile = (int64_t)kT * kDedupItems;  // 4096 keys: rounds of kT consecutive keys
static_assert(kDedupItems * (kT / 64) == 64, "one wave scans the tile's (round, wave) counts");

// Kept keys per tile.
__global__ __launch_bounds__(kT) void dedup_count(const uint64_t *__restrict__ keys, int64_t n,
                                                  uint32_t *__restrict__ tile_cnt) {
    const int64_t base = (int64_t)blockIdx.x * kDedupTile;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < kDedupItems; ++j) {
        const int64_t i = base + (int64_t)j * kT + threadIdx.x;
        if (i < n) c += (i == 0 || keys[i - 1] != keys[i]) ? 1u : 0u;
    }
    c = rtgd::wave_sum_u32(c);
    __shared__ uint32_t w[kT / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int k = 0; k < kT / 64; ++k) t += w[k];
        tile_cnt[blockIdx.x] = t;
    }
}

// Writes each kept key's column to adjs[p], p = tile_off[tile] + its rank among the tile's kept
// keys (a ballot per wave and round, one wave scan over the tile's 64 (round, wave) counts), and
// the row offsets: a kept key at p in row s, whose predecessor key is in row ps < s, sets
// offsets[ps + 1 .. s] = p (the first key: offsets[0 .. s] = 0); the last key, in row s, sets
// offsets[s + 1 .. nrows] = m, the kept total.
__global__ __launch_bounds__(kT) void dedup_write(const uint64_t *__restrict__ keys, int64_t n, int bits,
                                                  const uint32_t *__restrict__ tile_off,
                                                  const uint32_t *__restrict__ d_m, int32_t nrows,
                                                  int32_t *__restrict__ adjs, int32_t *__restrict__ offsets) {
    __shared__ uint32_t cell[64];  // kept keys per (round, wave), then their exclusive prefix
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kDedupTile;
    uint64_t k[kDedupItems];
    int32_t prow[kDedupItems];  // row of the preceding key (-1 before the first)
    uint64_t bal[kDedupItems];
#pragma unroll
    for (int j = 0; j < kDedupItems; ++j) {
        const int64_t i = base + (int64_t)j * kT + threadIdx.x;
        bool keep = false;
        k[j] = 0;
        prow[j] = -1;
        if (i < n) {
            k[j] = keys[i];
            if (i > 0) {
                const uint64_t q = keys[i - 1];
                keep = q != k[j];
                prow[j] = (int32_t)(q >> bits);
            } else {
                keep = true;
            }
        }
        bal[j] = __ballot(keep);
        if (lane == 0) cell[j * (kT / 64) + wave] = (uint32_t)__popcll(bal[j]);
    }
    __syncthreads();
    if (wave == 0) {  // cells in key order: round-major, then wave
        const uint32_t v = cell[lane];
        const uint32_t inc = rtgd::wave_inclusive_scan(v);
        cell[lane] = inc - v;
    }
    __syncthreads();
    const uint32_t t0 = tile_off[blockIdx.x];
    const uint64_t lt = rtgd::lanemask_lt();
    const uint64_t lo = (1ull << bits) - 1;
#pragma unroll
    for (int j = 0; j < kDedupItems; ++j) {
        const int64_t i = base + (int64_t)j * kT + threadIdx.x;
        const int32_t row = (int32_t)(k[j] >> bits);
        if ((bal[j] >> lane) & 1ull) {
            const uint32_t p = t0 + cell[j * (kT / 64) + wave] + (uint32_t)__popcll(bal[j] & lt);
            adjs[p] = (int32_t)(k[j] & lo);
            for (int32_t u = prow[j] + 1; u <= row; ++u) offsets[u] = (int32_t)p;
        }
        if (i == n - 1) {
            const int32_t m = (int32_t)*d_m;
            for (int32_t u = row + 1; u <= nrows; ++u) offsets[u] = m;
        }
    }
}

__global__ void fill_i32(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ __launch_bounds__(kT) void max_row_kernel(const int32_t *__restrict__ offsets, int32_t n, int32_t *out) {
    int32_t mx = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        mx = max(mx, offsets[i + 1] - offsets[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off, 64));
    __shared__ int32_t w[kT / 64];  // one device atomic per block, not per wave
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 1; k < kT / 64; ++k) mx = max(mx, w[k]);
        atomicMax(out, max(mx, w[0]));
    }
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

// Blocks of kT threads for n items, at most per_cu per CU (grid-stride loops take the rest).
unsigned grid_for(rtg_ctx *ctx, int64_t n, int per_cu = 16) {
    int64_t want = rtg::ceil_div(n, kT);
    int64_t cap = (int64_t)ctx->cu_count * per_cu;
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
    max_row_kernel<<<grid_for(ctx, n, 4), kT, 0, ctx->stream>>>(d_offsets, n, d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, d, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTG_OK;
}

}  // namespace rtg

namespace rtg {

int id_histogram(rtg_ctx *ctx, const int32_t *ids, int64_t n, int64_t domain, uint32_t *deg) {
    if (n <= 0) return RTG_OK;
    hipStream_t st = ctx->stream;
    const int64_t nparts = ceil_div(std::max<int64_t>(domain, 1), (int64_t)1 << kPartBits);
    if (nparts > kMaxPartitions) {
        raw_degree<<<grid_for(ctx, n), kT, 0, st>>>(ids, n, deg);
        HIP_TRY(hipGetLastError());
        return RTG_OK;
    }
    const uint32_t ntiles = (uint32_t)std::max<int64_t>(1, ceil_div(n, kPartTile));
    const int64_t nhist = nparts * (int64_t)ntiles;
    const size_t bytes = Arena::round(4 * (size_t)n) + Arena::round(4 * (size_t)nhist) +
                         Arena::round(4 * (size_t)(nparts + 1)) + Arena::round(scan_work_bytes(nhist)) + 1024;
    DevBuf buf;
    HIP_TRY(hipMalloc(&buf.p, bytes));
    Arena a{(char *)buf.p, 0, bytes};
    int32_t *grouped = a.take<int32_t>((size_t)n);
    uint32_t *hist = a.take<uint32_t>((size_t)nhist);
    uint32_t *pstart = a.take<uint32_t>((size_t)nparts + 1);
    void *scan_work = a.take<char>(scan_work_bytes(nhist));
    part_hist<<<ntiles, kT, 0, st>>>(ids, n, (int32_t)nparts, hist, ntiles);
    HIP_TRY(hipGetLastError());
    RTG_TRY(scan_exclusive_u32_w(ctx, hist, hist, nhist, nullptr, scan_work));
    part_scatter<<<ntiles, kT, 0, st>>>(ids, n, (int32_t)nparts, hist, ntiles, grouped);
    HIP_TRY(hipGetLastError());
    part_starts<<<(unsigned)ceil_div(nparts + 1, 256), 256, 0, st>>>(hist, ntiles, (int32_t)nparts, n, pstart);
    HIP_TRY(hipGetLastError());
    part_count<<<(unsigned)ceil_div(n, kCountChunk), kT, 0, st>>>(grouped, n, pstart, (int32_t)nparts, deg);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));  // the temporaries are freed on return
    return RTG_OK;
}

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

    // Working copy of the pairs, and the key-sized buffers (sized by the pair count, allocated
    // before the first synchronisation so that no allocation waits on the device mid-build).
    DevBuf pbuf, kbuf;
    HIP_TRY(hipMalloc(&pbuf.p, sizeof(int32_t) * (size_t)std::max<int64_t>(nv, 1)));
    int32_t *d_pairs = (int32_t *)pbuf.p;
    const int64_t kcap = std::max<int64_t>(max_keys, 1);
    const int64_t ktiles = ceil_div(kcap, kDedupTile);
    const int mm_blocks = (int)grid_for(ctx, std::max<int64_t>(nv, 1));
    const size_t kbytes = Arena::round(8 * (size_t)kcap) * 2 + Arena::round(radix_sort_work_bytes(kcap)) +
                          Arena::round(4 * (size_t)ktiles) + Arena::round(scan_work_bytes(ktiles)) +
                          Arena::round(sizeof(MinMax) * (size_t)(mm_blocks + 1)) + 4096;
    HIP_TRY(hipMalloc(&kbuf.p, kbytes));
    Arena ka{(char *)kbuf.p, 0, kbytes};
    uint64_t *keys0 = ka.take<uint64_t>(kcap);
    uint64_t *keys1 = ka.take<uint64_t>(kcap);
    void *sort_work = ka.take<char>(radix_sort_work_bytes(kcap));
    uint32_t *tile_off = ka.take<uint32_t>(ktiles);
    void *kscan = ka.take<char>(scan_work_bytes(ktiles));
    MinMax *mm_part = ka.take<MinMax>((size_t)mm_blocks + 1);
    unsigned long long *counter = ka.take<unsigned long long>(4);
    uint32_t *d_m = ka.take<uint32_t>(4);
    if (npairs > 0)
        HIP_TRY(hipMemcpyAsync(d_pairs, pairs, sizeof(int32_t) * (size_t)nv,
                               (flags & RTG_PAIRS_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(ev[0], st));

    // ---- domain, raw degree, renumbering (graph.cpp:40-62, 109-114, 126-152) ----
    MinMax mm{0, -1};
    if (npairs > 0) {
        minmax_kernel<<<mm_blocks, kT, 0, st>>>(d_pairs, nv, mm_part);
        HIP_TRY(hipGetLastError());
        minmax_final<<<1, kT, 0, st>>>(mm_part, mm_blocks, mm_part + mm_blocks);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&mm, mm_part + mm_blocks, sizeof(mm), hipMemcpyDeviceToHost, st));
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
        keep_flags<<<grid_for(ctx, domain, 4), kT, 0, st>>>(deg, domain, orient ? 2u : 1u, keep, d_small);
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
    auto *g = new rtg_graph();
    g->ctx = ctx;
    g->n = (int32_t)n_out;
    g->oriented = orient;
    if (hipMalloc(&g->d_offsets, sizeof(int32_t) * (size_t)(n_out + 1)) != hipSuccess) {
        (void)hipGetLastError();
        rtg_graph_destroy(g);
        return rtg_fail(RTG_ENOMEM, "rtg_graph_from_pairs: CSR allocation failed");
    }
    std::unique_ptr<rtg_graph, int (*)(rtg_graph *)> gguard(g, rtg_graph_destroy);

    // ---- orientation and key packing (graph.cpp:116-124, 126-152), or both directions ----
    int bits = bits_for(std::max<int64_t>(n_out, 2));
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
    // The columns, sized by the keys before deduplication (an upper bound of m: the count of
    // distinct keys is only known after the sort, and sizing to it would stall the build on a
    // synchronisation and an allocation).
    if (hipMalloc(&g->d_adjs, sizeof(int32_t) * (size_t)std::max<unsigned long long>(nkeys, 1)) != hipSuccess) {
        (void)hipGetLastError();
        return rtg_fail(RTG_ENOMEM, "rtg_graph_from_pairs: CSR allocation failed");
    }

    // ---- sort (thrust_helper.cu:16-42) ----
    uint64_t *kk[2] = {keys0, keys1};
    int which = 0, passes = 0;
    RTG_TRY(radix_sort_u64(ctx, kk, (int64_t)nkeys, 2 * bits, &which, &passes, sort_work));
    uint64_t *sorted = kk[which];
    HIP_TRY(hipEventRecord(ev[3], st));

    // ---- dedup + CSR (graph.cpp:159-173, 71-102) ----
    uint32_t m = 0;
    if (nkeys > 0) {
        const int64_t tiles = ceil_div((int64_t)nkeys, kDedupTile);
        dedup_count<<<(unsigned)tiles, kT, 0, st>>>(sorted, (int64_t)nkeys, tile_off);
        HIP_TRY(hipGetLastError());
        RTG_TRY(scan_exclusive_u32_w(ctx, tile_off, tile_off, tiles, d_m, kscan));
        dedup_write<<<(unsigned)tiles, kT, 0, st>>>(sorted, (int64_t)nkeys, bits, tile_off, d_m, (int32_t)n_out,
                                                    g->d_adjs, g->d_offsets);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&m, d_m, 4, hipMemcpyDeviceToHost, st));
    } else {
        fill_i32<<<grid_for(ctx, n_out + 1), kT, 0, st>>>(g->d_offsets, n_out + 1, 0);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ev[4], st));
    HIP_TRY(hipStreamSynchronize(st));
    g->m = m;
    RTG_TRY(csr_max_row(ctx, g->d_offsets, g->n, &g->max_row));
    gguard.release();
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
    rtg::renum_free(g->renum);
    delete g;
    return RTG_OK;
}

}  // extern "C"

namespace rtg {
// Loads this file's kernels onto the current device (an attribute query of one of them).
int preload_csr_build() {
    hipFuncAttributes attr;
    return hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&minmax_kernel)) == hipSuccess ? RTG_OK : RTG_EHIP;
}
}  // namespace rtg
