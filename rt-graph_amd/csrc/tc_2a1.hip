// Two-hop (2A1) triangle count on the device: the reference's default method
// (tc/rt_tc/include/config.h:3-4, RTTC_METHOD 1), whose Preprocessing (tc/bs_tc/bs_tc.cu:153-211)
// builds, per vertex u, the multiset of N+(w) over w in N+(u), sorts it and run-length
// encodes it into (u, x, mult) triples on the host, and whose HashmapKernel (:213-229) adds
// mult for every triple with x in N+(u).
//
// Here the two-hop lists are gathered on the device in chunks of consecutive vertices whose
// wedge total fits a fixed budget, packed as (u - u0) << 32 | x keys, radix-sorted, and each
// run of equal keys is checked against N+(u) once with weight = run length.
#include "tc_kernels.hpp"

#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

__global__ void wedge_counts(const int32_t *__restrict__ off, const int32_t *__restrict__ adj, int32_t n,
                             uint64_t *__restrict__ wc) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
         u += (int64_t)gridDim.x * blockDim.x) {
        uint64_t s = 0;
        for (int32_t i = off[u]; i < off[u + 1]; ++i) {
            const int32_t w = adj[i];
            s += (uint64_t)(off[w + 1] - off[w]);
        }
        wc[u] = s;
    }
}

// One wave per vertex of the chunk: copies N+(w) for each w in N+(u) into the key buffer.
__global__ void gather_two_hop(const int32_t *__restrict__ off, const int32_t *__restrict__ adj, int32_t u0,
                               int32_t u1, const uint64_t *__restrict__ wofs, uint64_t base,
                               uint64_t *__restrict__ keys) {
    const int lane = __lane_id();
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t u = u0 + (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); u < u1; u += nw) {
        uint64_t pos = wofs[u] - base;
        const uint64_t hi = (uint64_t)(u - u0) << 32;
        for (int32_t i = off[u]; i < off[u + 1]; ++i) {
            const int32_t w = adj[i];
            const int32_t wb = off[w], wl = off[w + 1] - wb;
            for (int32_t j = lane; j < wl; j += 64) keys[pos + j] = hi | (uint32_t)adj[wb + j];
            pos += (uint64_t)wl;
        }
    }
}

__device__ __forceinline__ bool member(const int32_t *__restrict__ data, int32_t len, int32_t val) {
    int32_t l = 0, r = len - 1;
    while (l <= r) {
        const int32_t mid = (l + r) >> 1;
        const int32_t x = data[mid];
        if (x == val) return true;
        if (x < val) l = mid + 1;
        else r = mid - 1;
    }
    return false;
}

// For every run start i of equal keys: mult = run length; add mult if x in N+(u).
__global__ void runs_check(const uint64_t *__restrict__ keys, int64_t nk, int32_t u0,
                           const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                           unsigned long long *count) {
    uint64_t local = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        if (i > 0 && keys[i - 1] == k) continue;
        int64_t j = i + 1;
        while (j < nk && keys[j] == k) ++j;
        const int32_t u = u0 + (int32_t)(k >> 32);
        const int32_t x = (int32_t)(k & 0xFFFFFFFFu);
        const int32_t b = off[u];
        if (member(adj + b, off[u + 1] - b, x)) local += (uint64_t)(j - i);
    }
    local = rtgd::wave_sum_u64(local);
    if (__lane_id() == 0 && local) atomicAdd(count, (unsigned long long)local);
}

struct Buf {
    void *p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int rtg_tc_count_2a1_impl(rtg_ctx *ctx, const rtg_graph *g, uint64_t *count, rtg_tc_stats *stats) {
    using namespace rtg;
    const int32_t n = g->n;
    hipStream_t st = ctx->stream;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    *count = 0;
    if (n == 0 || g->m == 0) return RTG_OK;
    // Budget of two-hop entries per chunk (8 bytes each, double-buffered for the sort).
    const char *env = std::getenv("RTG_2A1_CHUNK");
    const int64_t budget = env ? std::max<int64_t>(1024, std::atoll(env)) : ((int64_t)1 << 26);
    Buf wbuf;
    HIP_TRY(hipMalloc(&wbuf.p, sizeof(uint64_t) * (size_t)(n + 1) * 2 + scan_work_bytes(n + 1) + 4096));
    Arena wa{(char *)wbuf.p, 0, 0};
    uint64_t *wc = wa.take<uint64_t>((size_t)n + 1);
    uint64_t *wofs = wa.take<uint64_t>((size_t)n + 1);
    void *sw = wa.take<char>(scan_work_bytes(n + 1));
    HIP_TRY(hipMemsetAsync(wc, 0, sizeof(uint64_t) * (size_t)(n + 1), st));
    wedge_counts<<<(unsigned)std::min<int64_t>(ceil_div(n, 256), (int64_t)ctx->cu_count * 16), 256, 0, st>>>(
        g->d_offsets, g->d_adjs, n, wc);
    HIP_TRY(hipGetLastError());
    RTG_TRY(scan_exclusive_u64_w(ctx, wc, wofs, n + 1, nullptr, sw));
    std::vector<uint64_t> h((size_t)n + 1);
    HIP_TRY(hipMemcpyAsync(h.data(), wofs, sizeof(uint64_t) * (size_t)(n + 1), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // Largest single-vertex wedge list bounds the chunk size.
    int64_t maxw = 0;
    for (int32_t u = 0; u < n; ++u) maxw = std::max<int64_t>(maxw, (int64_t)(h[u + 1] - h[u]));
    const int64_t cap = std::max<int64_t>(budget, maxw);
    if (cap >= ((int64_t)1 << 32)) return set_error(RTG_ENOMEM, "2A1: a single vertex has >= 2^32 wedges");
    Buf kbuf, cbuf;
    HIP_TRY(hipMalloc(&kbuf.p, sizeof(uint64_t) * (size_t)cap * 2 + radix_sort_work_bytes(cap) + 4096));
    Arena ka{(char *)kbuf.p, 0, 0};
    uint64_t *k0 = ka.take<uint64_t>((size_t)cap);
    uint64_t *k1 = ka.take<uint64_t>((size_t)cap);
    void *sortw = ka.take<char>(radix_sort_work_bytes(cap));
    HIP_TRY(hipMalloc(&cbuf.p, sizeof(unsigned long long)));
    unsigned long long *d_count = (unsigned long long *)cbuf.p;
    HIP_TRY(hipMemsetAsync(d_count, 0, sizeof(unsigned long long), st));
    HIP_TRY(hipEventRecord(ctx->ev[0], st));
    int32_t u0 = 0;
    while (u0 < n) {
        int32_t u1 = u0 + 1;
        while (u1 < n && (int64_t)(h[u1 + 1] - h[u0]) <= cap) ++u1;
        const int64_t nk = (int64_t)(h[u1] - h[u0]);
        if (nk > 0) {
            const int64_t nv = u1 - u0;
            gather_two_hop<<<(unsigned)std::min<int64_t>(ceil_div(nv * 64, 256), (int64_t)ctx->cu_count * 16), 256, 0, st>>>(
                g->d_offsets, g->d_adjs, u0, u1, wofs, h[u0], k0);
            HIP_TRY(hipGetLastError());
            int hb = 1;
            while (hb < 31 && ((int64_t)1 << hb) < nv) ++hb;
            uint64_t *kk[2] = {k0, k1};
            int which = 0, passes = 0;
            RTG_TRY(radix_sort_u64(ctx, kk, nk, 32 + hb, &which, &passes, sortw));
            runs_check<<<(unsigned)std::min<int64_t>(ceil_div(nk, 256), (int64_t)ctx->cu_count * 16), 256, 0, st>>>(
                kk[which], nk, u0, g->d_offsets, g->d_adjs, d_count);
            HIP_TRY(hipGetLastError());
        }
        u0 = u1;
    }
    HIP_TRY(hipEventRecord(ctx->ev[1], st));
    unsigned long long c = 0;
    HIP_TRY(hipMemcpyAsync(&c, d_count, sizeof(c), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *count = c;
    if (stats) {
        stats->t_count_ms = elapsed_ms(ctx->ev[0], ctx->ev[1]);
        stats->edges = g->m;
        stats->kernel_launches = 0;
    }
    return RTG_OK;
}

namespace rtg {
// Loads this file's kernels onto the current device (an attribute query of one of them).
int preload_tc_2a1() {
    hipFuncAttributes attr;
    return hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&wedge_counts)) == hipSuccess ? RTG_OK : RTG_EHIP;
}
}  // namespace rtg
