// Device-side synthetic inputs. Each value is the same pure function of (seed, index) as in
// gen_host.cpp, so the device and host generators produce identical arrays.
#include "device.hpp"
#include "rng.hpp"

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

}  // namespace

extern "C" {

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
