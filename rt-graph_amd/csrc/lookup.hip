// Sorted-array lookup: out[i] = index of the first element equal to keys[i], or -1 — the
// semantics of BSAccessCount (rt_search/simple_bs.cu:35-53), whose lower_bound starts with
// r = n - 1 so the candidate index is clamped to n - 1 before the equality test.
//
// Two MI355X forms:
//  - lookup_sampled (stateless calls): every block copies an evenly spaced sample of the
//    array (kSamples entries, the top levels of the search) into LDS; a query is located
//    between two samples by a binary search in LDS and finished by a lower_bound over one
//    sample interval in HBM, so log2(kSamples) of the ~log2(n) dependent probes never leave
//    the CU. Each thread carries kQ queries in lockstep so that several dependent HBM loads
//    are in flight per lane.
//  - lookup_indexed (rtg_lookup_index, built once per array): a 32-ary implicit B+-tree whose
//    nodes are single 128-byte lines. The top level (<= 8192 entries) is searched in LDS;
//    every level below costs one line read per query (8 independent 16-byte loads, counted
//    with 32 compares), so a 2^28-entry array takes three dependent global reads per key —
//    an L2-resident level, an infinity-cache-resident level and the data line in HBM —
//    instead of about sixteen.
#include "device.hpp"

#include <limits>
#include <type_traits>

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

template <class T>
__device__ __forceinline__ T type_max() {
    return std::numeric_limits<T>::max();
}

template <class T>
struct IndexArgs {
    const T *lev[rtg_lookup_index::kMaxLevels + 1];  // lev[0] = the data
    int64_t len[rtg_lookup_index::kMaxLevels + 1];
    const T *tail;  // the data's last, partial run of 32, padded with the type's maximum
    int32_t top;
};

// Every level and the tail in one launch: entry i of level L >= 1 is the largest value of its
// subtree, data[min((i + 1) * 32^L - 1, n - 1)] (the last element of the subtree's run of the
// data), or the type's maximum past the real entries; the tail is the data's last, partial run
// of 32, padded. One launch for every level (the levels are not built from each other).
template <class T>
struct LevelLayout {
    int64_t off[rtg_lookup_index::kMaxLevels + 1];
    int64_t len[rtg_lookup_index::kMaxLevels + 1];
    int64_t padded[rtg_lookup_index::kMaxLevels + 1];
    int64_t span[rtg_lookup_index::kMaxLevels + 1];  // 32^L
    int64_t off_tail;
    int32_t top;
};

template <class T>
__global__ __launch_bounds__(256) void index_build_all(const T *__restrict__ data, int64_t n, LevelLayout<T> lay,
                                                       T *__restrict__ levels) {
    const int64_t total = lay.off_tail + rtg::kIndexFan;
    for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < total; g += (int64_t)gridDim.x * 256) {
        if (g >= lay.off_tail) {
            const int64_t b = n / rtg::kIndexFan * rtg::kIndexFan, k = g - lay.off_tail;
            levels[g] = b + k < n ? data[b + k] : type_max<T>();
            continue;
        }
        int L = 1;
        while (L < lay.top && g >= lay.off[L + 1]) ++L;
        const int64_t i = g - lay.off[L];
        levels[g] = i < lay.len[L] ? data[min((i + 1) * lay.span[L] - 1, n - 1)] : type_max<T>();
    }
}

// One node: 32 consecutive entries from a 16-byte aligned address, as 8 independent loads.
template <class T>
__device__ __forceinline__ void load_node(const T *p, T (&x)[rtg::kIndexFan]) {
    const uint4 *v = reinterpret_cast<const uint4 *>(p);
#pragma unroll
    for (int k = 0; k < rtg::kIndexFan / 4; ++k) {
        const uint4 w = v[k];
        x[4 * k + 0] = __builtin_bit_cast(T, w.x);
        x[4 * k + 1] = __builtin_bit_cast(T, w.y);
        x[4 * k + 2] = __builtin_bit_cast(T, w.z);
        x[4 * k + 3] = __builtin_bit_cast(T, w.w);
    }
}

template <class T>
__device__ __forceinline__ int count_below(const T (&x)[rtg::kIndexFan], T t) {
    int c = 0;
#pragma unroll
    for (int k = 0; k < rtg::kIndexFan; ++k) c += x[k] < t ? 1 : 0;
    return c;
}

template <class T>
__device__ __forceinline__ int count_not_above(const T (&x)[rtg::kIndexFan], T t) {
    int c = 0;
#pragma unroll
    for (int k = 0; k < rtg::kIndexFan; ++k) c += x[k] <= t ? 1 : 0;
    return c;
}

// Descent for Q keys per thread in lockstep. c = number of top-level entries below the key
// is the subtree holding the key's lower bound (entries are subtree maxima); at each level
// below, c becomes 32*c + (entries of node c below the key). A key above every top entry is
// above every element: -1. At level 0 the node is a run of the data itself; the key is present
// iff some entry equals it, and then its first occurrence is at 32*c + (entries below it).
// Loads are issued for every slot whether or not its key is still live (a finished slot reads
// node 0), so that each level is one straight-line batch of 8*Q independent line loads.
template <class T, int Q>
__global__ __launch_bounds__(kThreads) void lookup_indexed(IndexArgs<T> a, const T *__restrict__ keys, int64_t q,
                                                           int32_t *__restrict__ out) {
    __shared__ T top[rtg::kIndexTop];
    const int ntop = (int)a.len[a.top];
    if (a.top > 0) {  // an index level: 128-byte aligned, padded to a multiple of 32 entries
        const uint4 *g = reinterpret_cast<const uint4 *>(a.lev[a.top]);
        uint4 *l = reinterpret_cast<uint4 *>(top);
        const int nv = (ntop + rtg::kIndexFan - 1) / rtg::kIndexFan * (rtg::kIndexFan / 4);
        for (int k = threadIdx.x; k < nv; k += kThreads) l[k] = g[k];
    } else {
        for (int k = threadIdx.x; k < ntop; k += kThreads) top[k] = a.lev[0][k];
    }
    // pad to a power of two with the type's maximum, so every search step reads in bounds
    int pw = 1;
    while (pw < ntop) pw *= 2;
    const int filled = a.top > 0 ? (ntop + rtg::kIndexFan - 1) / rtg::kIndexFan * rtg::kIndexFan : ntop;
    for (int k = filled + threadIdx.x; k < pw; k += kThreads) top[k] = type_max<T>();
    __syncthreads();
    const int64_t n = a.len[0];
    const int64_t stride = (int64_t)gridDim.x * kThreads * Q;
    for (int64_t base = ((int64_t)blockIdx.x * kThreads + threadIdx.x) * Q; base < q; base += stride) {
        T t[Q];
        uint32_t c[Q];
        bool act[Q];
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            t[j] = base + j < q ? keys[base + j] : T(0);
            c[j] = 0;
        }
        // number of top entries below each key: binary lifting over the pw padded entries
        // (steps pw/2 .. 1 reach at most pw-1), then one last test of entry c
        for (int step = pw >> 1; step > 0; step >>= 1) {
#pragma unroll
            for (int j = 0; j < Q; ++j) {
                const T v = top[c[j] + step - 1];
                c[j] = v < t[j] ? c[j] + step : c[j];
            }
        }
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            c[j] += top[c[j]] < t[j] ? 1u : 0u;
            act[j] = base + j < q && (int)c[j] < ntop;
            if (!act[j]) c[j] = 0;
        }
        for (int L = a.top - 1; L >= 1; --L) {
            const T *lev = a.lev[L];
            T x[Q][rtg::kIndexFan];
#pragma unroll
            for (int j = 0; j < Q; ++j) load_node(lev + (uint64_t)c[j] * rtg::kIndexFan, x[j]);
#pragma unroll
            for (int j = 0; j < Q; ++j) {
                const uint32_t nc = c[j] * rtg::kIndexFan + (uint32_t)count_below(x[j], t[j]);
                c[j] = act[j] ? nc : 0u;
            }
        }
        int32_t r[Q];
        if (a.top == 0) {
#pragma unroll
            for (int j = 0; j < Q; ++j) r[j] = act[j] && top[c[j]] == t[j] ? (int32_t)c[j] : -1;
        } else {
            T x[Q][rtg::kIndexFan];
#pragma unroll
            for (int j = 0; j < Q; ++j) {
                // the last run, when partial, is read from its padded copy; its padding (the
                // type's maximum) is never below the key and equals it only if the last element does
                const uint64_t b = (uint64_t)c[j] * rtg::kIndexFan;
                load_node((int64_t)b + rtg::kIndexFan <= n ? a.lev[0] + b : a.tail, x[j]);
            }
#pragma unroll
            for (int j = 0; j < Q; ++j) {
                const int lt = count_below(x[j], t[j]);
                const bool hit = count_not_above(x[j], t[j]) > lt;
                r[j] = act[j] && hit ? (int32_t)(c[j] * rtg::kIndexFan + (uint32_t)lt) : -1;
            }
        }
#pragma unroll
        for (int j = 0; j < Q; ++j)
            if (base + j < q) out[base + j] = r[j];
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

template <class T>
int lookup_index_launch(const rtg_lookup_index *ix, hipStream_t st, const T *d_keys, int64_t q, int32_t *d_out) {
    if (q <= 0) return RTG_OK;
    IndexArgs<T> a{};
    a.lev[0] = static_cast<const T *>(ix->d_data);
    for (int L = 0; L <= ix->top; ++L) {
        a.len[L] = ix->len[L];
        if (L > 0) a.lev[L] = static_cast<const T *>(ix->d_levels) + ix->off[L];
    }
    a.top = ix->top;
    a.tail = static_cast<const T *>(ix->d_levels) + ix->off_tail;
    constexpr int Q = 2;
    const int64_t want = ceil_div(q, (int64_t)kThreads * Q);
    // Each block stages the top level (up to 32 KB) once and then strides over its keys.
    const int64_t cap = (int64_t)ix->ctx->cu_count * 4;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min(want, cap));
    lookup_indexed<T, Q><<<grid, kThreads, 0, st>>>(a, d_keys, q, d_out);
    HIP_TRY(hipGetLastError());
    return RTG_OK;
}

template int lookup_index_launch<uint32_t>(const rtg_lookup_index *, hipStream_t, const uint32_t *, int64_t, int32_t *);
template int lookup_index_launch<int32_t>(const rtg_lookup_index *, hipStream_t, const int32_t *, int64_t, int32_t *);

namespace {
template <class T>
int build_levels(rtg_lookup_index *ix) {
    if (ix->top == 0) return RTG_OK;
    LevelLayout<T> lay{};
    int64_t span = 1;
    for (int L = 1; L <= ix->top; ++L) {
        span *= kIndexFan;
        lay.off[L] = ix->off[L];
        lay.len[L] = ix->len[L];
        lay.padded[L] = ceil_div(ix->len[L], kIndexFan) * kIndexFan;
        lay.span[L] = span;
    }
    lay.off_tail = ix->off_tail;
    lay.top = ix->top;
    const int64_t total = ix->off_tail + kIndexFan;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, 256), (int64_t)ix->ctx->cu_count * 16));
    index_build_all<T><<<grid, 256, 0, ix->ctx->stream>>>(static_cast<const T *>(ix->d_data), ix->n, lay,
                                                            static_cast<T *>(ix->d_levels));
    HIP_TRY(hipGetLastError());
    return RTG_OK;
}
}  // namespace

namespace {
// Level sizes and offsets of the index over n entries; returns its element count.
int64_t index_layout(rtg_lookup_index *ix, rtg_ctx *ctx, const void *d_data, int64_t n, int is_signed) {
    ix->ctx = ctx;
    ix->device = ctx->device;
    ix->is_signed = is_signed;
    ix->d_data = d_data;
    ix->n = n;
    ix->len[0] = n;
    int64_t total = 0;
    int L = 0;
    while (ix->len[L] > kIndexTop) {
        ix->len[L + 1] = ceil_div(ix->len[L], kIndexFan);
        ix->off[L + 1] = total;
        total += ceil_div(ix->len[L + 1], kIndexFan) * kIndexFan;
        ++L;
    }
    ix->top = L;
    if (L > 0) {  // room for the padded copy of the data's last run
        ix->off_tail = total;
        total += kIndexFan;
    }
    ix->bytes = 4 * total;
    return total;
}
}  // namespace

// A transient index for one stateless call: its levels go into the context's grow-only scratch
// (no allocation once the scratch is large enough), the build is only enqueued, and nothing is
// freed afterwards (ix->d_levels is not owned).
int lookup_index_transient(rtg_ctx *ctx, const void *d_data, int64_t n, int is_signed, rtg_lookup_index *ix) {
    const int64_t total = index_layout(ix, ctx, d_data, n, is_signed);
    if (total == 0) return RTG_OK;
    RTG_TRY(ctx_scratch(ctx, (size_t)ix->bytes, &ix->d_levels));
    return is_signed ? build_levels<int32_t>(ix) : build_levels<uint32_t>(ix);
}

int lookup_index_build(rtg_ctx *ctx, const void *d_data, int64_t n, int is_signed, rtg_lookup_index **out) {
    auto *ix = new rtg_lookup_index();
    const int64_t total = index_layout(ix, ctx, d_data, n, is_signed);
    int rc = RTG_OK;
    if (total > 0) {
        if (hipMalloc(&ix->d_levels, (size_t)ix->bytes) != hipSuccess) {
            (void)hipGetLastError();
            delete ix;
            return rtg_fail(RTG_ENOMEM, "rtg_lookup_index: out of device memory");
        }
        rc = is_signed ? build_levels<int32_t>(ix) : build_levels<uint32_t>(ix);
        if (rc == RTG_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = rtg_fail(RTG_EHIP, "rtg_lookup_index: build failed");
    }
    if (rc != RTG_OK) {
        lookup_index_free(ix);
        return rc;
    }
    *out = ix;
    return RTG_OK;
}

void lookup_index_free(rtg_lookup_index *ix) {
    if (!ix) return;
    if (ix->d_levels) (void)hipFree(ix->d_levels);
    delete ix;
}

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
    // A transient search index pays off when the array fits the LDS top level (no levels to
    // build) or when the keys are many enough to amortise one pass over the array (measured on
    // 2^28 uint32: build about 0.26 ms; the two forms break even near q = n/256); the build is
    // inside the timed region. Its levels live in the context's scratch, so repeated calls
    // allocate nothing.
    const bool aligned = ((uintptr_t)d_data & 15) == 0;
    if (aligned && q > 0 && (n <= rtg::kIndexTop || q >= n / 128)) {
        rtg_lookup_index ix;
        RTG_TRY(rtg::lookup_index_transient(ctx, d_data, n, std::is_signed<T>::value ? 1 : 0, &ix));
        RTG_TRY(rtg::lookup_index_launch<T>(&ix, ctx->stream, d_keys, q, d_out));
    } else {
        RTG_TRY(rtg::lookup_launch<T>(ctx, ctx->stream, d_data, n, d_keys, q, d_out));
    }
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

namespace {
// Sorts n 32-bit keys in place on the device: LSD radix sort, four 8-bit passes (even, so the
// result ends in the caller's buffer); *t_ms gets the device time of the passes.
int sort32_api(rtg_ctx *ctx, uint32_t *d_data, int64_t n, bool is_signed, double *t_ms) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (t_ms) *t_ms = 0.0;
    if (n < 0 || (n > 0 && !d_data)) return rtg_fail(RTG_EINVAL, "rtg_sort: bad arguments");
    if (n >= ((int64_t)1 << 32)) return rtg_fail(RTG_EINVAL, "rtg_sort: more than 2^32 - 1 keys");
    if (n <= 1) return RTG_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    const size_t wb = rtg::radix_sort_work_bytes(n);
    const size_t bytes = rtg::Arena::round(sizeof(uint32_t) * (size_t)n) + rtg::Arena::round(wb) + 256;
    void *buf = nullptr;
    if (hipMalloc(&buf, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return rtg_fail(RTG_ENOMEM, "rtg_sort: temporary of " + std::to_string(bytes) + " bytes failed");
    }
    rtg::Arena ar{(char *)buf, 0, bytes};
    uint32_t *keys[2] = {d_data, ar.take<uint32_t>((size_t)n)};
    void *work = ar.take<char>(wb);
    int which = 0, passes = 0, rc = RTG_OK;
    hipStream_t st = ctx->stream;
    if (hipEventRecord(ctx->ev[0], st) != hipSuccess) rc = rtg_fail(RTG_EHIP, "rtg_sort: event record failed");
    if (rc == RTG_OK) rc = rtg::radix_sort_32(ctx, keys, n, is_signed, &which, &passes, work);
    if (rc == RTG_OK && which != 0 &&
        hipMemcpyAsync(d_data, keys[1], sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, st) != hipSuccess)
        rc = rtg_fail(RTG_EHIP, "rtg_sort: copy back failed");
    if (rc == RTG_OK && (hipEventRecord(ctx->ev[1], st) != hipSuccess || hipEventSynchronize(ctx->ev[1]) != hipSuccess))
        rc = rtg_fail(RTG_EHIP, "rtg_sort: synchronisation failed");
    if (rc == RTG_OK && t_ms) *t_ms = rtg::elapsed_ms(ctx->ev[0], ctx->ev[1]);
    (void)hipStreamSynchronize(st);  // the passes may still read the temporary after a failure
    (void)hipFree(buf);
    return rc;
}
}  // namespace

RTG_API int rtg_sort_u32(rtg_ctx *ctx, uint32_t *d_data, int64_t n, double *t_ms) {
    return sort32_api(ctx, d_data, n, false, t_ms);
}

RTG_API int rtg_sort_i32(rtg_ctx *ctx, int32_t *d_data, int64_t n, double *t_ms) {
    return sort32_api(ctx, reinterpret_cast<uint32_t *>(d_data), n, true, t_ms);
}

RTG_API int rtg_lookup_index_create_u32(rtg_ctx *ctx, const uint32_t *d_data, int64_t n, rtg_lookup_index **out) {
    if (!out) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_create: null out");
    *out = nullptr;
    RTG_TRY(rtg::ctx_enter(ctx));
    if (n < 0 || (n > 0 && !d_data)) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_create: bad arguments");
    if (n > INT32_MAX) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_create: n > 2^31-1 (results are int32 indices)");
    if ((uintptr_t)d_data & 15) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_create: d_data must be 16-byte aligned");
    std::lock_guard<std::mutex> lock(ctx->mu);
    return rtg::lookup_index_build(ctx, d_data, n, 0, out);
}

RTG_API int rtg_lookup_index_create_i32(rtg_ctx *ctx, const int32_t *d_data, int64_t n, rtg_lookup_index **out) {
    if (!out) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_create: null out");
    *out = nullptr;
    RTG_TRY(rtg::ctx_enter(ctx));
    if (n < 0 || (n > 0 && !d_data)) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_create: bad arguments");
    if (n > INT32_MAX) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_create: n > 2^31-1 (results are int32 indices)");
    if ((uintptr_t)d_data & 15) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_create: d_data must be 16-byte aligned");
    std::lock_guard<std::mutex> lock(ctx->mu);
    return rtg::lookup_index_build(ctx, d_data, n, 1, out);
}

RTG_API int rtg_lookup_index_info(const rtg_lookup_index *ix, int32_t *levels, int64_t *bytes) {
    if (!ix) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_info: null index");
    if (levels) *levels = ix->top;
    if (bytes) *bytes = ix->bytes;
    return RTG_OK;
}

RTG_API int rtg_lookup_index_run(rtg_ctx *ctx, const rtg_lookup_index *ix, const void *d_keys, int64_t q, int32_t *d_out,
                                 double *t_kernel_ms) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!ix || ix->ctx != ctx) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_run: index belongs to another context");
    if (q < 0 || (q > 0 && (!d_keys || !d_out))) return rtg_fail(RTG_EINVAL, "rtg_lookup_index_run: bad arguments");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
    if (ix->is_signed) RTG_TRY(rtg::lookup_index_launch<int32_t>(ix, ctx->stream, static_cast<const int32_t *>(d_keys), q, d_out));
    else RTG_TRY(rtg::lookup_index_launch<uint32_t>(ix, ctx->stream, static_cast<const uint32_t *>(d_keys), q, d_out));
    HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (t_kernel_ms) *t_kernel_ms = rtg::elapsed_ms(ctx->ev[0], ctx->ev[1]);
    return RTG_OK;
}

RTG_API int rtg_lookup_index_destroy(rtg_lookup_index *ix) {
    if (!ix) return RTG_OK;
    (void)hipSetDevice(ix->device);  // the context may already be gone
    rtg::lookup_index_free(ix);
    return RTG_OK;
}

}  // extern "C"

namespace rtg {
// Loads this file's kernels onto the current device (an attribute query of one of them).
int preload_lookup() {
    hipFuncAttributes attr;
    return hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&lookup_sampled<uint32_t>)) == hipSuccess ? RTG_OK : RTG_EHIP;
}
}  // namespace rtg
