// Device-side synthetic inputs. Each value is the same pure function of (seed, index) as in
// gen_host.cpp, so the device and host generators produce identical arrays.
#include "device.hpp"
#include "rng.hpp"

#include <vector>

namespace {

__global__ void rmat_kernel(uint32_t scale, int64_t nsamples, uint64_t t1, uint64_t t2, uint64_t t3,
                            uint64_t seed, int32_t *__restrict__ pairs) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nsamples;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t src = 0, dst = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            const uint64_t r = rtg::top53(rtg::crand(seed, rtg::kStreamRmat, (uint64_t)i * 64 + l));
            const uint32_t sb = (r >= t2) ? 1u : 0u;
            const uint32_t db = ((r >= t1 && r < t2) || (r >= t3)) ? 1u : 0u;
            src = (src << 1) | sb;
            dst = (dst << 1) | db;
        }
        int2 p;
        p.x = (int32_t)rtg::permute_bits(src, scale, seed ^ rtg::kStreamRmatPerm);
        p.y = (int32_t)rtg::permute_bits(dst, scale, seed ^ rtg::kStreamRmatPerm);
        reinterpret_cast<int2 *>(pairs)[i] = p;
    }
}

__global__ void uniform_u32_kernel(int64_t n, uint64_t seed, uint64_t stream, uint32_t *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)(rtg::crand(seed, stream, (uint64_t)i) >> 32);
}

__global__ void array_keys_kernel(int64_t n, uint64_t seed, uint64_t *__restrict__ keys) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        keys[i] = rtg::crand(seed, rtg::kStreamArray, (uint64_t)i) >> 32;
}

__global__ void low32_kernel(const uint64_t *__restrict__ keys, int64_t n, uint32_t *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)keys[i];
}

// Stub s of vertex u (stubs of a vertex contiguous): owner[s] = u.
__global__ void stub_owner_kernel(const int64_t *__restrict__ start, int32_t n, int32_t *__restrict__ owner) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (int64_t)gridDim.x * blockDim.x)
        for (int64_t s = start[u]; s < start[u + 1]; ++s) owner[s] = (int32_t)u;
}

__global__ void stub_key_kernel(int64_t total, uint64_t seed, uint64_t *__restrict__ keys) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total; s += (int64_t)gridDim.x * blockDim.x)
        keys[s] = (rtg::crand(seed, rtg::kStreamStub, (uint64_t)s) & 0xFFFFFFFF00000000ull) | (uint64_t)s;
}

// Base edge k pairs the stubs at sorted positions 2k and 2k+1.
__global__ void match_kernel(const uint64_t *__restrict__ sorted, int64_t half, const int32_t *__restrict__ owner,
                             int32_t *__restrict__ pairs, int32_t *__restrict__ partner) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < half; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s0 = (int64_t)(sorted[2 * k] & 0xFFFFFFFFull), s1 = (int64_t)(sorted[2 * k + 1] & 0xFFFFFFFFull);
        const int32_t o0 = owner[s0], o1 = owner[s1];
        pairs[2 * k] = o0;
        pairs[2 * k + 1] = o1;
        partner[s0] = o1;
        partner[s1] = o0;
    }
}

// Closing flag of stub s: not the first stub of its vertex, and the draw is below p.
__global__ void close_flag_kernel(const int32_t *__restrict__ owner, const int64_t *__restrict__ start, int64_t total,
                                  uint64_t seed, uint64_t tp, uint32_t *__restrict__ flag) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total; s += (int64_t)gridDim.x * blockDim.x) {
        const bool first = s == start[owner[s]];
        flag[s] = (!first && rtg::top53(rtg::crand(seed, rtg::kStreamClose, (uint64_t)s)) < tp) ? 1u : 0u;
    }
}

__global__ void close_emit_kernel(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos, int64_t total,
                                  const int32_t *__restrict__ partner, int64_t base, int32_t *__restrict__ pairs) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total; s += (int64_t)gridDim.x * blockDim.x)
        if (flag[s]) {
            const int64_t e = base + pos[s];
            pairs[2 * e] = partner[s - 1];
            pairs[2 * e + 1] = partner[s];
        }
}

struct Buf {
    void *p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
};

unsigned grid_of(rtg_ctx *ctx, int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(rtg::ceil_div(n, 256), (int64_t)ctx->cu_count * 32));
}

}  // namespace

extern "C" {

// Sorted uniform uint32 array on the device: the same values as rtg_gen_sorted_array.
RTG_API int rtg_gen_sorted_array_device(rtg_ctx *ctx, int64_t n, uint64_t seed, uint32_t *d_out) {
    using namespace rtg;
    RTG_TRY(ctx_enter(ctx));
    if (n < 0 || (n > 0 && !d_out) || n >= ((int64_t)1 << 32)) return rtg_fail(RTG_EINVAL, "rtg_gen_sorted_array_device: bad arguments");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (n == 0) return RTG_OK;
    hipStream_t st = ctx->stream;
    Buf kb;
    const size_t bytes = Arena::round(8 * (size_t)n) * 2 + radix_sort_work_bytes(n) + 1024;
    HIP_TRY(hipMalloc(&kb.p, bytes));
    Arena a{(char *)kb.p, 0, bytes};
    uint64_t *k0 = a.take<uint64_t>((size_t)n), *k1 = a.take<uint64_t>((size_t)n);
    void *w = a.take<char>(radix_sort_work_bytes(n));
    array_keys_kernel<<<grid_of(ctx, n), 256, 0, st>>>(n, seed, k0);
    uint64_t *kk[2] = {k0, k1};
    int which = 0, passes = 0;
    RTG_TRY(radix_sort_u64(ctx, kk, n, 32, &which, &passes, w));
    low32_kernel<<<grid_of(ctx, n), 256, 0, st>>>(kk[which], n, d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return RTG_OK;
}

// Configuration model with triangle closing on the device: the same pairs, in the same order,
// as rtg_gen_config_model for the same degree sequence (deg is a host array). d_pairs holds
// 2 * (total/2 + total) ints; *npairs receives the count written.
RTG_API int rtg_gen_config_model_device(rtg_ctx *ctx, int32_t n, const int32_t *deg, int64_t total, double p_close,
                                        uint64_t seed, int32_t *d_pairs, int64_t *npairs) {
    using namespace rtg;
    RTG_TRY(ctx_enter(ctx));
    if (n < 1 || !deg || !d_pairs || !npairs || (total & 1) || total >= ((int64_t)1 << 32))
        return rtg_fail(RTG_EINVAL, "rtg_gen_config_model_device: bad arguments");
    std::vector<int64_t> start((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) start[i + 1] = start[i] + deg[i];
    if (start[n] != total) return rtg_fail(RTG_EINVAL, "rtg_gen_config_model_device: degree sum != total");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t st = ctx->stream;
    Buf buf;
    const size_t bytes = Arena::round(8 * (size_t)(n + 1)) + Arena::round(4 * (size_t)total) * 4 +
                         Arena::round(8 * (size_t)total) * 2 + radix_sort_work_bytes(total) + scan_work_bytes(total) + 4096;
    HIP_TRY(hipMalloc(&buf.p, bytes));
    Arena a{(char *)buf.p, 0, bytes};
    int64_t *d_start = a.take<int64_t>((size_t)n + 1);
    int32_t *owner = a.take<int32_t>((size_t)total);
    int32_t *partner = a.take<int32_t>((size_t)total);
    uint32_t *flag = a.take<uint32_t>((size_t)total);
    uint32_t *pos = a.take<uint32_t>((size_t)total);
    uint64_t *k0 = a.take<uint64_t>((size_t)total), *k1 = a.take<uint64_t>((size_t)total);
    void *sw = a.take<char>(radix_sort_work_bytes(total));
    void *cw = a.take<char>(scan_work_bytes(total));
    uint32_t *d_cnt = a.take<uint32_t>(4);
    HIP_TRY(hipMemcpyAsync(d_start, start.data(), 8 * (size_t)(n + 1), hipMemcpyHostToDevice, st));
    stub_owner_kernel<<<grid_of(ctx, n), 256, 0, st>>>(d_start, n, owner);
    stub_key_kernel<<<grid_of(ctx, total), 256, 0, st>>>(total, seed, k0);
    uint64_t *kk[2] = {k0, k1};
    int which = 0, passes = 0;
    RTG_TRY(radix_sort_u64(ctx, kk, total, 64, &which, &passes, sw));
    const int64_t half = total / 2;
    match_kernel<<<grid_of(ctx, half), 256, 0, st>>>(kk[which], half, owner, d_pairs, partner);
    close_flag_kernel<<<grid_of(ctx, total), 256, 0, st>>>(owner, d_start, total, seed, prob53(p_close), flag);
    RTG_TRY(scan_exclusive_u32_w(ctx, flag, pos, total, d_cnt, cw));
    close_emit_kernel<<<grid_of(ctx, total), 256, 0, st>>>(flag, pos, total, partner, half, d_pairs);
    HIP_TRY(hipGetLastError());
    uint32_t nclose = 0;
    HIP_TRY(hipMemcpyAsync(&nclose, d_cnt, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *npairs = half + (int64_t)nclose;
    return RTG_OK;
}


RTG_API int rtg_gen_rmat_device(rtg_ctx *ctx, uint32_t scale, int64_t nsamples, double a, double b, double c,
                                uint64_t seed, int32_t *d_pairs) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (scale < 1 || scale > 30 || nsamples < 0 || (!d_pairs && nsamples > 0))
        return rtg_fail(RTG_EINVAL, "rtg_gen_rmat_device: bad arguments");
    if (((uintptr_t)d_pairs & 7) != 0) return rtg_fail(RTG_EINVAL, "rtg_gen_rmat_device: d_pairs must be 8-byte aligned");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (nsamples == 0) return RTG_OK;
    const unsigned grid = (unsigned)std::min<int64_t>(rtg::ceil_div(nsamples, 256), (int64_t)ctx->cu_count * 32);
    rmat_kernel<<<grid, 256, 0, ctx->stream>>>(scale, nsamples, rtg::prob53(a), rtg::prob53(a + b),
                                               rtg::prob53(a + b + c), seed, d_pairs);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTG_OK;
}

RTG_API int rtg_gen_uniform_u32_device(rtg_ctx *ctx, int64_t n, uint64_t seed, uint64_t stream, uint32_t *d_out) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (n < 0 || (!d_out && n > 0)) return rtg_fail(RTG_EINVAL, "rtg_gen_uniform_u32_device: bad arguments");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (n == 0) return RTG_OK;
    const unsigned grid = (unsigned)std::min<int64_t>(rtg::ceil_div(n, 256), (int64_t)ctx->cu_count * 32);
    uniform_u32_kernel<<<grid, 256, 0, ctx->stream>>>(n, seed, stream, d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTG_OK;
}

}  // extern "C"

namespace rtg {
// Loads this file's kernels onto the current device (an attribute query of one of them).
int preload_gen_device() {
    hipFuncAttributes attr;
    return hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&rmat_kernel)) == hipSuccess ? RTG_OK : RTG_EHIP;
}
}  // namespace rtg
