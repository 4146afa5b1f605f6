// Degree-ordered copy of an oriented CSR for batched counting plans.
//
// A triangle count does not depend on vertex ids, so a plan may count on a renumbered copy of
// the graph. In descending order of total degree (out + in in the oriented CSR; ties by id)
// the out-neighbours of a vertex — the endpoints of higher degree, by the orientation of
// graph.cpp:116-124 — get small ids. On R-MAT scale 21 (1.27M vertices) that puts 81% of the
// streamed partner elements in lists whose entries are all below 2^16, so they can be streamed
// as 16-bit values, and 92% of them meet owners whose whole list lies below 2^15, so a bitmap
// of the owner's range is a few KB of LDS and a probe is one LDS read with no hashing.
//
//   degrees   in-degree by the partitioned id histogram over adjs, out-degree from the offsets
//   order     stable radix sort of ((2^db - 1 - total degree) << vb | id), db + vb bits: new id
//             of v = its position
//   offsets   out-degrees scattered to the new ids and scanned
//   entries   each row's entries renamed and sorted where its new row starts: rows of up to 64
//             entries by a bitonic network across one wave's lanes (shuffles), up to 512 in a
//             wave's LDS buffer, up to 4096 in a block's; the 32- and 16-bit rows written
//             together. (A graph with a longer row takes one key (new src << b | new dst) per
//             edge and a radix sort on 2b bits.)
//   16 bits   the low 16 bits of every entry, each row's largest (last) entry, every 16th and
//             256th entry sampled (searches in a row)
//   edge map  for per-edge plans (built on their first use): each edge of the copy's index in
//             the graph's own CSR, by a binary search in the graph's row
#include "graph.hpp"

#include <chrono>
#include <cstdlib>

namespace {

constexpr int kT = 256;

unsigned grid_of(rtg_ctx *ctx, int64_t n, int per_cu = 16) {
    const int64_t want = std::max<int64_t>(1, rtg::ceil_div(n, kT));
    return (unsigned)std::min<int64_t>(want, (int64_t)ctx->cu_count * per_cu);
}

int bits_of(int64_t n) {
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) < n) ++b;
    return b;
}

// key[v] = ((2^db - 1 - (out + in)) << vb) | v (a total degree is at most m < 2^db, an id below
// 2^vb): an ascending sort puts the highest total degree first and breaks ties by id.
__global__ void order_keys(const int32_t *__restrict__ off, const uint32_t *__restrict__ deg_in, int32_t n, int db,
                           int vb, uint64_t *__restrict__ keys) {
    const uint64_t top = ((uint64_t)1 << db) - 1;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t d = (uint64_t)(uint32_t)(off[v + 1] - off[v]) + deg_in[v];
        keys[v] = ((top - d) << vb) | (uint64_t)(uint32_t)v;
    }
}

// newid[v] = position of v in the sorted order, inv[i] = v; cnt[newid[v]] = out-degree of v.
__global__ void order_scatter(const uint64_t *__restrict__ sorted, const int32_t *__restrict__ off, int32_t n, int vb,
                              uint32_t *__restrict__ newid, int32_t *__restrict__ inv, uint32_t *__restrict__ cnt) {
    const uint64_t vmask = ((uint64_t)1 << vb) - 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)(sorted[i] & vmask);
        newid[v] = (uint32_t)i;
        inv[i] = (int32_t)v;
        cnt[i] = (uint32_t)(off[v + 1] - off[v]);
    }
}

// emap[i] = the graph's index of the copy's edge i = (u' -> v'): the position of inv[v'] in the
// graph's sorted row inv[u'] (a binary search; the edge is there, the copy being the same DAG
// renamed). Walks each block's range of 256 copy rows like edge_keys.
__global__ __launch_bounds__(kT) void edge_map(const int32_t *__restrict__ coff, const int32_t *__restrict__ cadj,
                                               int32_t n, const int32_t *__restrict__ inv,
                                               const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                               int32_t *__restrict__ emap) {
    __shared__ int32_t so[kT + 1];
    __shared__ int32_t rb[kT];  // the graph's row start of each copy row
    __shared__ int32_t rl[kT];  // and its length
    for (int64_t v0 = (int64_t)blockIdx.x * kT; v0 < n; v0 += (int64_t)gridDim.x * kT) {
        const int32_t nv = (int32_t)min<int64_t>(kT, n - v0);
        for (int i = threadIdx.x; i <= nv; i += kT) so[i] = coff[v0 + i];
        if ((int)threadIdx.x < nv) {
            const int32_t u = inv[v0 + threadIdx.x];
            rb[threadIdx.x] = off[u];
            rl[threadIdx.x] = off[u + 1] - off[u];
        }
        __syncthreads();
        const int32_t e1 = so[nv];
        int32_t v = 0;  // so[v] <= e < so[v + 1]
        for (int32_t e = so[0] + threadIdx.x; e < e1; e += kT) {
            while (so[v + 1] <= e) ++v;
            const int32_t x = inv[cadj[e]], b = rb[v];
            int32_t lo = 0, hi = rl[v];  // first position of the row holding a value >= x
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (adj[b + mid] < x) lo = mid + 1;
                else hi = mid;
            }
            emap[e] = b + lo;
        }
        __syncthreads();  // so, rb and rl are refilled for the next rows
    }
}

// keys[e] = (newid[u] << bits) | newid[adj[e]] for every edge e of row u. A block takes 256
// consecutive rows, whose edges form one contiguous range, and walks it with coalesced loads
// and stores, each thread advancing a cursor over the block's row offsets in LDS.
__global__ __launch_bounds__(kT) void edge_keys(const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                int32_t n, const uint32_t *__restrict__ newid, int bits,
                                                uint64_t *__restrict__ keys) {
    __shared__ int32_t so[kT + 1];
    __shared__ uint32_t sid[kT];
    for (int64_t v0 = (int64_t)blockIdx.x * kT; v0 < n; v0 += (int64_t)gridDim.x * kT) {
        const int32_t nv = (int32_t)min<int64_t>(kT, n - v0);
        for (int i = threadIdx.x; i <= nv; i += kT) so[i] = off[v0 + i];
        if ((int)threadIdx.x < nv) sid[threadIdx.x] = newid[v0 + threadIdx.x];
        __syncthreads();
        const int32_t e1 = so[nv];
        int32_t v = 0;  // so[v] <= e < so[v + 1]
        for (int32_t e = so[0] + threadIdx.x; e < e1; e += kT) {
            while (so[v + 1] <= e) ++v;
            keys[e] = ((uint64_t)sid[v] << bits) | (uint64_t)newid[adj[e]];
        }
        __syncthreads();  // so and sid are refilled for the next rows
    }
}

// Rows of the copy sorted in place of the key sort: row v of the graph becomes row newid[v],
// its entries renamed by newid and sorted ascending, written as 32- and 16-bit entries.
constexpr int32_t kSortSmall = 64, kSortMid = 512, kSortLarge = 4096;

// Ascending bitonic sort of one value per lane across the wave (lane i ends with the i-th
// smallest): 21 compare-exchange stages, the partner's value through a lane shuffle.
__device__ __forceinline__ uint32_t wave_bitonic_sort(uint32_t y, int lane) {
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t o = (uint32_t)__shfl_xor((int)y, j);
            const bool asc = (lane & k) == 0, lower = (lane & j) == 0;
            y = (lower == asc) ? min(y, o) : max(y, o);
        }
    }
    return y;
}

// Ascending bitonic sort of buf[0 .. P) (P a power of two, at least 128) by one wave: each stage
// loads all of a lane's P / 128 pairs before it writes any (one LDS round trip per stage, not
// one per pair), then a wave fence.
template <int P>
__device__ __forceinline__ void wave_lds_sort(uint32_t *buf, int lane) {
    constexpr int kPer = P / 128;
    for (int32_t k = 2; k <= P; k <<= 1) {
        for (int32_t j = k >> 1; j > 0; j >>= 1) {
            uint32_t a[kPer], c[kPer];
            int32_t idx[kPer];
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const int32_t t = lane + 64 * q;
                idx[q] = ((t & ~(j - 1)) << 1) | (t & (j - 1));  // pair (i, i + j), bit j of i clear
                a[q] = buf[idx[q]];
                c[q] = buf[idx[q] + j];
            }
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const bool asc = (idx[q] & k) == 0;
                const uint32_t lo = min(a[q], c[q]), hi = max(a[q], c[q]);
                buf[idx[q]] = asc ? lo : hi;
                buf[idx[q] + j] = asc ? hi : lo;
            }
            rtgd::wave_sync_lds();
        }
    }
}

__device__ __forceinline__ int32_t pow2_at_least(int32_t x) { return x <= 1 ? 1 : (int32_t)(1u << (32 - __clz((uint32_t)(x - 1)))); }

// Rows of 1 .. 64 entries, one wave per row.
__global__ __launch_bounds__(kT) void sort_rows_small(const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                      int32_t n, const uint32_t *__restrict__ newid,
                                                      const int32_t *__restrict__ coff, int32_t *__restrict__ cadj,
                                                      uint16_t *__restrict__ cadj16) {
    const int lane = threadIdx.x & 63;
    const int64_t w0 = (int64_t)blockIdx.x * (kT / 64) + rtgd::uniform_u32(threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * (kT / 64);
    for (int64_t v = w0; v < n; v += nw) {
        const int32_t b = off[v], len = off[v + 1] - b;
        if (len == 0 || len > kSortSmall) continue;
        uint32_t y = lane < len ? newid[adj[b + lane]] : 0xFFFFFFFFu;
        y = wave_bitonic_sort(y, lane);
        const int32_t dst = coff[newid[v]];
        if (lane < len) {
            cadj[dst + lane] = (int32_t)y;
            cadj16[dst + lane] = (uint16_t)y;
        }
    }
}

// Row lists for the two LDS sorts: the rows of 65 .. 512 entries and of 513 .. 4096 entries
// (any order: a row's place in the copy is its new offset). Each block takes a contiguous range
// of rows, counts its rows of each list, reserves their places with one device atomic per list
// (atomics on one address serialize), then writes their ids.
__global__ __launch_bounds__(kT) void classify_rows(const int32_t *__restrict__ off, int32_t n,
                                                    int32_t *__restrict__ rows_mid, int32_t *__restrict__ rows_large,
                                                    uint32_t *__restrict__ counts) {
    __shared__ uint32_t wc[2][kT / 64];
    __shared__ uint32_t base[2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t lt = rtgd::lanemask_lt();
    const int64_t chunk = (((int64_t)n + gridDim.x - 1) / gridDim.x + kT - 1) / kT * kT;
    const int64_t v0 = (int64_t)blockIdx.x * chunk, v1 = min<int64_t>((int64_t)n, v0 + chunk);
    const auto kind = [&](int64_t v) -> int {  // 1 mid, 2 large, 0 neither
        if (v >= v1) return 0;
        const int32_t len = off[v + 1] - off[v];
        return (len > kSortSmall && len <= kSortMid) ? 1 : (len > kSortMid && len <= kSortLarge) ? 2 : 0;
    };
    uint32_t cm = 0, cl = 0;
    for (int64_t v = v0 + threadIdx.x; v < v1; v += kT) {
        const int k = kind(v);
        cm += k == 1;
        cl += k == 2;
    }
    cm = rtgd::wave_sum_u32(cm);
    cl = rtgd::wave_sum_u32(cl);
    if (lane == 0) {
        wc[0][wv] = cm;
        wc[1][wv] = cl;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t t = 0;
        for (int w = 0; w < kT / 64; ++w) t += wc[threadIdx.x][w];
        base[threadIdx.x] = t ? atomicAdd(&counts[threadIdx.x], t) : 0u;
    }
    __syncthreads();
    uint32_t pm = base[0], pl = base[1];  // the block's next places, the same in every thread
    for (int64_t vb = v0; vb < v1; vb += kT) {  // tiles of kT rows, in order
        const int k = kind(vb + threadIdx.x);
        const uint64_t bm = __ballot(k == 1), bl = __ballot(k == 2);
        __syncthreads();  // wc is rewritten below
        if (lane == 0) {
            wc[0][wv] = (uint32_t)__popcll(bm);
            wc[1][wv] = (uint32_t)__popcll(bl);
        }
        __syncthreads();
        uint32_t om = pm, ol = pl;
        for (int w = 0; w < kT / 64; ++w) {
            if (w < wv) {
                om += wc[0][w];
                ol += wc[1][w];
            }
            pm += wc[0][w];
            pl += wc[1][w];
        }
        if (k == 1) rows_mid[om + __popcll(bm & lt)] = (int32_t)(vb + threadIdx.x);
        if (k == 2) rows_large[ol + __popcll(bl & lt)] = (int32_t)(vb + threadIdx.x);
    }
}

// One wave per listed row, sorted in the wave's LDS buffer (rows of up to kBuf entries).
template <int kBuf>
__device__ __forceinline__ void sort_listed_row(const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                const uint32_t *__restrict__ newid, const int32_t *__restrict__ coff,
                                                int32_t *__restrict__ cadj, uint16_t *__restrict__ cadj16,
                                                uint32_t *buf, int32_t v, int lane) {
    const int32_t b = off[v], len = off[v + 1] - b;
    const int32_t P = max(pow2_at_least(len), 128);
    for (int32_t i = lane; i < P; i += 64) buf[i] = i < len ? newid[adj[b + i]] : 0xFFFFFFFFu;
    rtgd::wave_sync_lds();
    if constexpr (kBuf <= kSortMid) {
        if (P == 128) wave_lds_sort<128>(buf, lane);
        else if (P == 256) wave_lds_sort<256>(buf, lane);
        else wave_lds_sort<512>(buf, lane);
    } else {
        if (P == 1024) wave_lds_sort<1024>(buf, lane);
        else if (P == 2048) wave_lds_sort<2048>(buf, lane);
        else wave_lds_sort<4096>(buf, lane);
    }
    const int32_t dst = coff[newid[v]];
    for (int32_t i = lane; i < len; i += 64) {
        const uint32_t y = buf[i];
        cadj[dst + i] = (int32_t)y;
        cadj16[dst + i] = (uint16_t)y;
    }
    rtgd::wave_sync_lds();  // buf is refilled for the next row
}

// Rows of 65 .. 512 entries (rows_mid[0 .. count)), one wave per row.
__global__ __launch_bounds__(kT) void sort_rows_mid(const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                    const int32_t *__restrict__ rows, const uint32_t *__restrict__ count,
                                                    const uint32_t *__restrict__ newid,
                                                    const int32_t *__restrict__ coff, int32_t *__restrict__ cadj,
                                                    uint16_t *__restrict__ cadj16) {
    __shared__ uint32_t buf_all[kT / 64][kSortMid];
    const int lane = threadIdx.x & 63;
    const int wv = (int)rtgd::uniform_u32(threadIdx.x >> 6);
    const int64_t nr = (int64_t)*count;
    for (int64_t r = (int64_t)blockIdx.x * (kT / 64) + wv; r < nr; r += (int64_t)gridDim.x * (kT / 64))
        sort_listed_row<kSortMid>(off, adj, newid, coff, cadj, cadj16, buf_all[wv], rows[r], lane);
}

// Rows of 513 .. 4096 entries (rows_large[0 .. count)), one single-wave block per row: a 16 KB
// buffer each, and no block barriers between the sort's stages.
__global__ __launch_bounds__(64) void sort_rows_large(const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                      const int32_t *__restrict__ rows, const uint32_t *__restrict__ count,
                                                      const uint32_t *__restrict__ newid,
                                                      const int32_t *__restrict__ coff, int32_t *__restrict__ cadj,
                                                      uint16_t *__restrict__ cadj16) {
    __shared__ uint32_t buf[kSortLarge];
    const int64_t nr = (int64_t)*count;
    for (int64_t r = blockIdx.x; r < nr; r += gridDim.x)
        sort_listed_row<kSortLarge>(off, adj, newid, coff, cadj, cadj16, buf, rows[r], (int)threadIdx.x);
}

__global__ void split_entries(const uint64_t *__restrict__ sorted, int64_t m, int bits, int32_t *__restrict__ adj,
                              uint16_t *__restrict__ adj16) {
    const uint64_t mask = (1ull << bits) - 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(sorted[i] & mask);
        adj[i] = (int32_t)x;
        adj16[i] = (uint16_t)x;
    }
}

// s16[i] = adj[16 i] for i < 16 * ceil(m / 256) + 16 (INT32_MAX past the entries), and
// s256[j] = adj[256 j].
__global__ void sample_rows(const int32_t *__restrict__ adj, int64_t m, int32_t *__restrict__ s16,
                            int32_t *__restrict__ s256) {
    const int64_t n16 = 16 * ((m + 255) / 256) + 16;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = 16 * i < m ? adj[16 * i] : INT32_MAX;
        s16[i] = x;
        if ((i & 15) == 0 && 16 * i < m) s256[i >> 4] = x;
    }
}

// vinfo[v] = {row v's start, its end, its largest entry (-1 when empty), 0}: what planning
// gathers for a random row, in one 16-byte load; counts[0] = non-empty rows whose entries all lie
// below 2^16, counts[1] = their entries (one device atomic per block).
__global__ __launch_bounds__(kT) void row_max(const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                              int32_t n, int4 *__restrict__ vinfo,
                                              unsigned long long *__restrict__ counts) {
    unsigned long long r = 0, e = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = off[v], d = off[v + 1] - b;
        const int32_t x = d > 0 ? adj[b + d - 1] : -1;
        vinfo[v] = make_int4(b, b + d, x, 0);
        const bool f = d > 0 && x < 65536;
        r += f ? 1 : 0;
        e += f ? (unsigned long long)d : 0ull;
    }
    r = rtgd::wave_sum_u64(r);
    e = rtgd::wave_sum_u64(e);
    __shared__ unsigned long long w[2][kT / 64];
    if ((threadIdx.x & 63) == 0) {
        w[0][threadIdx.x >> 6] = r;
        w[1][threadIdx.x >> 6] = e;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned long long t = 0;
        for (int k = 0; k < kT / 64; ++k) t += w[threadIdx.x][k];
        if (t) atomicAdd(&counts[threadIdx.x], t);
    }
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

int build(rtg_ctx *ctx, const rtg_graph *g, rtg_renum *r) {
    using namespace rtg;
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = ctx->stream;
    const int32_t n = g->n;
    const int64_t m = g->m;
    r->n = n;
    r->m = m;
    HIP_TRY(hipMalloc(&r->d_off, sizeof(int32_t) * ((size_t)n + 1)));
    HIP_TRY(hipMalloc(&r->d_adj, sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1)));
    HIP_TRY(hipMalloc(&r->d_adj16, sizeof(uint16_t) * (size_t)std::max<int64_t>(m, 1)));
    HIP_TRY(hipMalloc(&r->d_vinfo, 4 * sizeof(int32_t) * (size_t)std::max<int32_t>(n, 1)));
    HIP_TRY(hipMalloc(&r->d_inv, sizeof(int32_t) * (size_t)std::max<int32_t>(n, 1)));
    const int64_t n256 = (m + 255) / 256;
    HIP_TRY(hipMalloc(&r->d_s16, sizeof(int32_t) * (size_t)(16 * n256 + 16)));
    HIP_TRY(hipMalloc(&r->d_s256, sizeof(int32_t) * (size_t)std::max<int64_t>(n256, 1)));
    const int bits = bits_of(std::max<int64_t>(n, 2));
    // rows sorted in place, else (longer rows, or RTG_RENUMBER_KEY_SORT=1) the key sort over all edges
    const char *ks = std::getenv("RTG_RENUMBER_KEY_SORT");
    const bool row_sort = g->max_row <= kSortLarge && !(ks && ks[0] == '1');
    const int64_t kn = row_sort ? (int64_t)n + 1 : std::max<int64_t>(m, (int64_t)n + 1);
    const size_t bytes = Arena::round(8 * (size_t)kn) * 2 + Arena::round(radix_sort_work_bytes(kn)) +
                         Arena::round(4 * ((size_t)n + 1)) * 5 + Arena::round(scan_work_bytes((int64_t)n + 1)) + 4096;
    DevBuf buf;
    HIP_TRY(hipMalloc(&buf.p, bytes));
    Arena a{(char *)buf.p, 0, bytes};
    uint64_t *k0 = a.take<uint64_t>((size_t)kn);
    uint64_t *k1 = a.take<uint64_t>((size_t)kn);
    void *sort_work = a.take<char>(radix_sort_work_bytes(kn));
    uint32_t *deg_in = a.take<uint32_t>((size_t)n + 1);
    uint32_t *newid = a.take<uint32_t>((size_t)n + 1);
    uint32_t *cnt = a.take<uint32_t>((size_t)n + 1);
    void *scan_work = a.take<char>(scan_work_bytes((int64_t)n + 1));
    unsigned long long *counts = a.take<unsigned long long>(2);
    int32_t *rows_mid = row_sort ? a.take<int32_t>((size_t)n) : nullptr;
    int32_t *rows_large = row_sort ? a.take<int32_t>((size_t)n) : nullptr;
    uint32_t *row_counts = a.take<uint32_t>(2);

    // degrees and the order
    HIP_TRY(hipMemsetAsync(deg_in, 0, sizeof(uint32_t) * ((size_t)n + 1), st));
    RTG_TRY(id_histogram(ctx, g->d_adjs, m, n, deg_in));
    const int db = bits_of(m + 1);
    order_keys<<<grid_of(ctx, n), kT, 0, st>>>(g->d_offsets, deg_in, n, db, bits, k0);
    HIP_TRY(hipGetLastError());
    uint64_t *kk[2] = {k0, k1};
    int which = 0, passes = 0;
    RTG_TRY(radix_sort_u64(ctx, kk, n, db + bits, &which, &passes, sort_work));
    HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * ((size_t)n + 1), st));
    order_scatter<<<grid_of(ctx, n), kT, 0, st>>>(kk[which], g->d_offsets, n, bits, newid, r->d_inv, cnt);
    HIP_TRY(hipGetLastError());
    RTG_TRY(scan_exclusive_u32_w(ctx, cnt, reinterpret_cast<uint32_t *>(r->d_off), (int64_t)n + 1, nullptr, scan_work));

    // entries: every row renamed and sorted where its new row starts
    if (row_sort) {
        HIP_TRY(hipMemsetAsync(row_counts, 0, 2 * sizeof(uint32_t), st));
        classify_rows<<<grid_of(ctx, n, 4), kT, 0, st>>>(g->d_offsets, n, rows_mid, rows_large, row_counts);
        HIP_TRY(hipGetLastError());
        if (g->max_row > kSortMid) {  // the long rows first: fewest and slowest
            sort_rows_large<<<(unsigned)(ctx->cu_count * 10), 64, 0, st>>>(g->d_offsets, g->d_adjs, rows_large,
                                                                           row_counts + 1, newid, r->d_off, r->d_adj,
                                                                           r->d_adj16);
            HIP_TRY(hipGetLastError());
        }
        if (g->max_row > kSortSmall) {
            sort_rows_mid<<<(unsigned)(ctx->cu_count * 8), kT, 0, st>>>(g->d_offsets, g->d_adjs, rows_mid, row_counts,
                                                                        newid, r->d_off, r->d_adj, r->d_adj16);
            HIP_TRY(hipGetLastError());
        }
        sort_rows_small<<<grid_of(ctx, (int64_t)n * 64), kT, 0, st>>>(g->d_offsets, g->d_adjs, n, newid, r->d_off,
                                                                      r->d_adj, r->d_adj16);
        HIP_TRY(hipGetLastError());
    } else {
        edge_keys<<<grid_of(ctx, n), kT, 0, st>>>(g->d_offsets, g->d_adjs, n, newid, bits, k0);
        HIP_TRY(hipGetLastError());
        kk[0] = k0;
        kk[1] = k1;
        RTG_TRY(radix_sort_u64(ctx, kk, m, 2 * bits, &which, &passes, sort_work));
        split_entries<<<grid_of(ctx, m), kT, 0, st>>>(kk[which], m, bits, r->d_adj, r->d_adj16);
        HIP_TRY(hipGetLastError());
    }
    sample_rows<<<grid_of(ctx, 16 * n256 + 16), kT, 0, st>>>(r->d_adj, m, r->d_s16, r->d_s256);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(counts, 0, sizeof(unsigned long long) * 2, st));
    row_max<<<grid_of(ctx, n, 4), kT, 0, st>>>(r->d_off, r->d_adj, n, reinterpret_cast<int4 *>(r->d_vinfo), counts);
    HIP_TRY(hipGetLastError());
    unsigned long long hc[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(hc, counts, sizeof(hc), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));  // the temporaries are freed on return
    r->rows16 = (int64_t)hc[0];
    r->entries16 = (int64_t)hc[1];
    r->t_build_ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RTG_OK;
}

}  // namespace

namespace rtg {

void renum_free(rtg_renum *r) {
    if (!r) return;
    void *p[] = {r->d_off, r->d_adj, r->d_adj16, r->d_vinfo, r->d_inv, r->d_emap, r->d_s16, r->d_s256};
    for (void *x : p)
        if (x) (void)hipFree(x);
    delete r;
}

int graph_renumbered(rtg_ctx *ctx, const rtg_graph *g, bool with_emap, const rtg_renum **out) {
    *out = nullptr;
    rtg_renum *r = g->renum;
    if (!r) {
        if (g->n <= 0 || g->m <= 0) return set_error(RTG_EINVAL, "renumbered CSR of an empty graph");
        if (g->m >= ((int64_t)1 << 31)) return set_error(RTG_EINVAL, "renumbered CSR: too many edges");
        r = new rtg_renum();
        const int rc = build(ctx, g, r);
        if (rc != RTG_OK) {
            (void)hipGetLastError();
            renum_free(r);
            return rc;
        }
        const_cast<rtg_graph *>(g)->renum = r;  // a cache on the graph; plan creation holds ctx->mu
    }
    if (with_emap && !r->d_emap) {
        const auto t0 = std::chrono::steady_clock::now();
        int32_t *emap = nullptr;
        HIP_TRY(hipMalloc(&emap, sizeof(int32_t) * (size_t)r->m));
        edge_map<<<grid_of(ctx, r->n), kT, 0, ctx->stream>>>(r->d_off, r->d_adj, r->n, r->d_inv, g->d_offsets, g->d_adjs,
                                                             emap);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
            (void)hipFree(emap);
            return set_error(RTG_EHIP, "renumbered CSR: edge map failed");
        }
        r->d_emap = emap;
        r->t_emap_ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    *out = r;
    return RTG_OK;
}

}  // namespace rtg

extern "C" {

RTG_API int rtg_graph_ordered_csr(rtg_ctx *ctx, const rtg_graph *g, int32_t *offsets, int32_t *adjs, int32_t *inv) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!g) return rtg_fail(RTG_EINVAL, "rtg_graph_ordered_csr: null graph");
    if (!g->oriented) return rtg_fail(RTG_EINVAL, "rtg_graph_ordered_csr: graph is not degree-oriented");
    std::lock_guard<std::mutex> lock(ctx->mu);
    const rtg_renum *r = nullptr;
    RTG_TRY(rtg::graph_renumbered(ctx, g, false, &r));
    hipStream_t st = ctx->stream;
    if (offsets) HIP_TRY(hipMemcpyAsync(offsets, r->d_off, sizeof(int32_t) * ((size_t)r->n + 1), hipMemcpyDeviceToHost, st));
    if (adjs) HIP_TRY(hipMemcpyAsync(adjs, r->d_adj, sizeof(int32_t) * (size_t)r->m, hipMemcpyDeviceToHost, st));
    if (inv) HIP_TRY(hipMemcpyAsync(inv, r->d_inv, sizeof(int32_t) * (size_t)r->n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return RTG_OK;
}

}  // extern "C"

namespace rtg {
// Loads this file's kernels onto the current device (an attribute query of one of them).
int preload_renumber() {
    hipFuncAttributes attr;
    return hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&order_keys)) == hipSuccess ? RTG_OK : RTG_EHIP;
}
}  // namespace rtg
