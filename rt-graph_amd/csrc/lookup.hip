// Sorted-array lookup: out[i] = index of the first element equal to keys[i], or -1 — the
// semantics of BSAccessCount (rt_search/simple_bs.cu:35-53), whose lower_bound starts with
// r = n - 1 so the candidate index is clamped to n - 1 before the equality test.
//
// MI355X form: every block first copies an evenly spaced sample of the array (kSamples
// entries, the top levels of the search) into LDS; a query is located between two samples by
// a binary search in LDS and finished by a lower_bound over one sample interval in HBM, so
// log2(kSamples) of the ~log2(n) dependent probes never leave the CU. Each thread carries
// kQ queries in lockstep so that several dependent HBM loads are in flight per lane.
#include "device.hpp"

#include "lookup.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kSamples = 4096;
constexpr int kQ = 4;

template <class T>
__global__ __launch_bounds__(kThreads) void lookup_sampled(const T *__restrict__ data, int64_t n,
                                                           const T *__restrict__ keys, int64_t q,
                                                           int32_t *__restrict__ out) {
    __shared__ T samp[kSamples];
    const bool use_samples = n >= 4 * (int64_t)kSamples;
    if (use_samples) {
        for (int k = threadIdx.x; k < kSamples; k += kThreads) samp[k] = data[((int64_t)k * n) / kSamples];
        __syncthreads();
    }
    const int64_t stride = (int64_t)gridDim.x * kThreads * kQ;
    for (int64_t base = ((int64_t)blockIdx.x * kThreads + threadIdx.x) * kQ; base < q; base += stride) {
        T t[kQ];
        int64_t lo[kQ], hi[kQ];
        bool act[kQ];
#pragma unroll
        for (int j = 0; j < kQ; ++j) {
            const int64_t i = base + j;
            act[j] = i < q;
            t[j] = act[j] ? keys[i] : (T)0;
            lo[j] = 0;
            hi[j] = n;
            if (act[j] && use_samples) {
                // number of samples < t
                int a = 0, b = kSamples;
                while (a < b) {
                    const int mid = (a + b) >> 1;
                    if (samp[mid] < t[j]) a = mid + 1;
                    else b = mid;
                }
                const int ks = a - 1;  // last sample < t, or -1
                lo[j] = ks < 0 ? 0 : ((int64_t)ks * n) / kSamples + 1;
                hi[j] = (ks + 1 < kSamples) ? ((int64_t)(ks + 1) * n) / kSamples : n;
            }
        }
        // lower_bound in [lo, hi) for all kQ queries in lockstep
        bool any = true;
        while (any) {
            any = false;
            T v[kQ];
            int64_t mid[kQ];
#pragma unroll
            for (int j = 0; j < kQ; ++j) {
                mid[j] = lo[j] + ((hi[j] - lo[j]) >> 1);
                if (act[j] && lo[j] < hi[j]) v[j] = data[mid[j]];
            }
#pragma unroll
            for (int j = 0; j < kQ; ++j) {
                if (act[j] && lo[j] < hi[j]) {
                    if (v[j] < t[j]) lo[j] = mid[j] + 1;
                    else hi[j] = mid[j];
                    any |= lo[j] < hi[j];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kQ; ++j) {
            if (!act[j]) continue;
            int32_t r = -1;
            if (n > 0) {
                const int64_t l = lo[j] < n ? lo[j] : n - 1;
                if (data[l] == t[j]) r = (int32_t)l;
            }
            out[base + j] = r;
        }
    }
}

}  // namespace

namespace rtg {

template <class T>
int lookup_launch(rtg_ctx *ctx, hipStream_t st, const T *d_data, int64_t n, const T *d_keys, int64_t q,
                  int32_t *d_out) {
    if (q <= 0) return RTG_OK;
    int64_t want = ceil_div(q, (int64_t)kThreads * kQ);
    // Enough blocks to fill the device; each block amortises its sample load over many keys.
    const int64_t cap = (int64_t)ctx->cu_count * 8;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min(want, cap));
    lookup_sampled<T><<<grid, kThreads, 0, st>>>(d_data, n, d_keys, q, d_out);
    HIP_TRY(hipGetLastError());
    return RTG_OK;
}

template int lookup_launch<uint32_t>(rtg_ctx *, hipStream_t, const uint32_t *, int64_t, const uint32_t *, int64_t, int32_t *);
template int lookup_launch<int32_t>(rtg_ctx *, hipStream_t, const int32_t *, int64_t, const int32_t *, int64_t, int32_t *);

}  // namespace rtg

namespace {
template <class T>
int lookup_api(rtg_ctx *ctx, const T *d_data, int64_t n, const T *d_keys, int64_t q, int32_t *d_out,
               double *t_kernel_ms) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (n < 0 || q < 0 || (n > 0 && !d_data) || (q > 0 && (!d_keys || !d_out)))
        return rtg_fail(RTG_EINVAL, "rtg_lookup: bad arguments");
    if (n > INT32_MAX) return rtg_fail(RTG_EINVAL, "rtg_lookup: n > 2^31-1 (results are int32 indices)");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
    RTG_TRY(rtg::lookup_launch<T>(ctx, ctx->stream, d_data, n, d_keys, q, d_out));
    HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (t_kernel_ms) *t_kernel_ms = rtg::elapsed_ms(ctx->ev[0], ctx->ev[1]);
    return RTG_OK;
}
}  // namespace

extern "C" {

RTG_API int rtg_lookup_u32(rtg_ctx *ctx, const uint32_t *d_data, int64_t n, const uint32_t *d_keys, int64_t q,
                           int32_t *d_out, double *t_kernel_ms) {
    return lookup_api<uint32_t>(ctx, d_data, n, d_keys, q, d_out, t_kernel_ms);
}

RTG_API int rtg_lookup_i32(rtg_ctx *ctx, const int32_t *d_data, int64_t n, const int32_t *d_keys, int64_t q,
                           int32_t *d_out, double *t_kernel_ms) {
    return lookup_api<int32_t>(ctx, d_data, n, d_keys, q, d_out, t_kernel_ms);
}

}  // extern "C"
