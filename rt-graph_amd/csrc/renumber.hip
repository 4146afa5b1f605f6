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
//   order     stable radix sort of (~total degree << 32 | id): new id of v = its position
//   offsets   out-degrees scattered to the new ids and scanned
//   entries   one key (new src << b | new dst) per edge, written while walking each block's
//             edge range with a cursor over its 256 row offsets; radix sort on 2b bits; the
//             sorted keys' low bits are the new rows, in order and sorted
//   16 bits   the low 16 bits of every entry, and a flag per row whose largest (last) entry
//             is below 2^16; every 32nd entry sampled (two-level searches in a row)
//   edge map  for per-edge plans (built on their first use): each edge of the copy's index in
//             the graph's own CSR, by a binary search in the graph's row
#include "graph.hpp"

#include <chrono>

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

// key[v] = (~(out + in) << 32) | v: an ascending sort puts the highest total degree first and
// breaks ties by id.
__global__ void order_keys(const int32_t *__restrict__ off, const uint32_t *__restrict__ deg_in, int32_t n,
                           uint64_t *__restrict__ keys) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t d = (uint32_t)(off[v + 1] - off[v]) + deg_in[v];
        keys[v] = ((uint64_t)(~d) << 32) | (uint64_t)(uint32_t)v;
    }
}

// newid[v] = position of v in the sorted order, inv[i] = v; cnt[newid[v]] = out-degree of v.
__global__ void order_scatter(const uint64_t *__restrict__ sorted, const int32_t *__restrict__ off, int32_t n,
                              uint32_t *__restrict__ newid, int32_t *__restrict__ inv, uint32_t *__restrict__ cnt) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)sorted[i];
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

// vmax[v] = row v's largest entry (-1 when empty); counts[0] = non-empty rows whose entries all
// lie below 2^16, counts[1] = their entries (one device atomic per block).
__global__ __launch_bounds__(kT) void row_max(const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                              int32_t n, int32_t *__restrict__ vmax,
                                              unsigned long long *__restrict__ counts) {
    unsigned long long r = 0, e = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = off[v], d = off[v + 1] - b;
        const int32_t x = d > 0 ? adj[b + d - 1] : -1;
        vmax[v] = x;
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
    HIP_TRY(hipMalloc(&r->d_vmax, sizeof(int32_t) * (size_t)std::max<int32_t>(n, 1)));
    HIP_TRY(hipMalloc(&r->d_inv, sizeof(int32_t) * (size_t)std::max<int32_t>(n, 1)));
    const int64_t n256 = (m + 255) / 256;
    HIP_TRY(hipMalloc(&r->d_s16, sizeof(int32_t) * (size_t)(16 * n256 + 16)));
    HIP_TRY(hipMalloc(&r->d_s256, sizeof(int32_t) * (size_t)std::max<int64_t>(n256, 1)));
    const int bits = bits_of(std::max<int64_t>(n, 2));
    const int64_t kn = std::max<int64_t>(m, (int64_t)n + 1);
    const size_t bytes = Arena::round(8 * (size_t)kn) * 2 + Arena::round(radix_sort_work_bytes(kn)) +
                         Arena::round(4 * ((size_t)n + 1)) * 3 + Arena::round(scan_work_bytes((int64_t)n + 1)) + 4096;
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

    // degrees and the order
    HIP_TRY(hipMemsetAsync(deg_in, 0, sizeof(uint32_t) * ((size_t)n + 1), st));
    RTG_TRY(id_histogram(ctx, g->d_adjs, m, n, deg_in));
    order_keys<<<grid_of(ctx, n), kT, 0, st>>>(g->d_offsets, deg_in, n, k0);
    HIP_TRY(hipGetLastError());
    uint64_t *kk[2] = {k0, k1};
    int which = 0, passes = 0;
    RTG_TRY(radix_sort_u64(ctx, kk, n, 64, &which, &passes, sort_work));
    HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * ((size_t)n + 1), st));
    order_scatter<<<grid_of(ctx, n), kT, 0, st>>>(kk[which], g->d_offsets, n, newid, r->d_inv, cnt);
    HIP_TRY(hipGetLastError());
    RTG_TRY(scan_exclusive_u32_w(ctx, cnt, reinterpret_cast<uint32_t *>(r->d_off), (int64_t)n + 1, nullptr, scan_work));

    // entries: keys of the renumbered edges, sorted, split into the 32- and 16-bit rows
    edge_keys<<<grid_of(ctx, n), kT, 0, st>>>(g->d_offsets, g->d_adjs, n, newid, bits, k0);
    HIP_TRY(hipGetLastError());
    kk[0] = k0;
    kk[1] = k1;
    RTG_TRY(radix_sort_u64(ctx, kk, m, 2 * bits, &which, &passes, sort_work));
    split_entries<<<grid_of(ctx, m), kT, 0, st>>>(kk[which], m, bits, r->d_adj, r->d_adj16);
    HIP_TRY(hipGetLastError());
    sample_rows<<<grid_of(ctx, 16 * n256 + 16), kT, 0, st>>>(r->d_adj, m, r->d_s16, r->d_s256);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(counts, 0, sizeof(unsigned long long) * 2, st));
    row_max<<<grid_of(ctx, n, 4), kT, 0, st>>>(r->d_off, r->d_adj, n, r->d_vmax, counts);
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
    void *p[] = {r->d_off, r->d_adj, r->d_adj16, r->d_vmax, r->d_inv, r->d_emap, r->d_s16, r->d_s256};
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
