// Device-wide exclusive prefix scan (reduce-then-scan) for uint32 and uint64 values.
//
// Level 0 splits the input into tiles of 256 threads x 16 items; a reduce kernel writes one
// sum per tile, the tile sums are scanned recursively with the same kernels (or, for at most
// 4096 tiles, added up by each block of the final kernel itself), and a scan kernel rescans
// each tile locally and adds its tile offset. Traffic: 2 reads + 1 write of
// the input size, plus the (1/4096-sized) tile-sum levels.
#include "device.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;

template <class T>
__device__ __forceinline__ T block_exclusive_sum(T v, T *lds_wave, T *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T inc = rtgd::wave_inclusive_scan(v);
    if (lane == 63) lds_wave[wave] = inc;
    __syncthreads();
    T wave_off = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        T s = lds_wave[w];
        if (w < wave) wave_off += s;
        all += s;
    }
    *total = all;
    return wave_off + inc - v;
}

template <class T>
__global__ __launch_bounds__(kThreads) void tile_reduce(const T *__restrict__ in, int64_t n,
                                                        T *__restrict__ partial) {
    __shared__ T lds[kThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    T s = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        int64_t idx = base + (int64_t)i * kThreads + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    if constexpr (sizeof(T) == 8) s = rtgd::wave_sum_u64(s);
    else s = rtgd::wave_sum_u32(s);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        T t = 0;
        for (int w = 0; w < kThreads / 64; ++w) t += lds[w];
        partial[blockIdx.x] = t;
    }
}

// partial == nullptr: single tile, offset 0. kPrefixed: partial holds the unscanned tile sums
// (at most kTile of them) and each block adds up those before its own tile, which saves the
// separate scan of the tile sums and the launch it would need.
template <class T, bool kPrefixed>
__global__ __launch_bounds__(kThreads) void tile_scan(const T *__restrict__ in, T *__restrict__ out,
                                                      int64_t n, const T *__restrict__ partial,
                                                      T *__restrict__ d_total) {
    __shared__ T lds[kThreads / 64];
    __shared__ T off_s;
    if constexpr (kPrefixed) {
        T s = 0;
        for (int64_t t = threadIdx.x; t < (int64_t)blockIdx.x; t += kThreads) s += partial[t];
        if constexpr (sizeof(T) == 8) s = rtgd::wave_sum_u64(s);
        else s = rtgd::wave_sum_u32(s);
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            T t = 0;
            for (int w = 0; w < kThreads / 64; ++w) t += lds[w];
            off_s = t;
        }
        __syncthreads();  // lds is reused by the tile's own scan
    }
    // Coalesced (striped) loads into LDS, each thread scans kItems consecutive values read back
    // from LDS, writes the results there, and the tile leaves with coalesced stores. One pad
    // element per 16 keeps the consecutive-item reads on distinct banks.
    __shared__ T tile[kTile + kTile / 16];
    const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int j = i * kThreads + threadIdx.x;
        const int64_t idx = base + j;
        tile[j + (j >> 4)] = idx < n ? in[idx] : (T)0;
    }
    __syncthreads();
    T v[kItems];
    T local = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int j = threadIdx.x * kItems + i;
        v[i] = tile[j + (j >> 4)];
        local += v[i];
    }
    T tile_total;
    T run = block_exclusive_sum<T>(local, lds, &tile_total);
    T off = 0;
    if constexpr (kPrefixed) off = off_s;
    else if (partial) off = partial[blockIdx.x];
    run += off;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int j = threadIdx.x * kItems + i;
        tile[j + (j >> 4)] = run;
        run += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int j = i * kThreads + threadIdx.x;
        const int64_t idx = base + j;
        if (idx < n) out[idx] = tile[j + (j >> 4)];
    }
    if (d_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *d_total = off + tile_total;
}

template <class T>
size_t work_bytes(int64_t n) {
    size_t total = 0;
    int64_t m = n;
    while (m > kTile) {
        m = rtg::ceil_div(m, kTile);
        total += rtg::Arena::round(sizeof(T) * (size_t)m);
    }
    return total + 256;
}

template <class T>
int scan_impl(rtg_ctx *ctx, const T *in, T *out, int64_t n, T *d_total, char *work) {
    if (n <= 0) {
        if (d_total) HIP_TRY(hipMemsetAsync(d_total, 0, sizeof(T), ctx->stream));
        return RTG_OK;
    }
    const int64_t tiles = rtg::ceil_div(n, kTile);
    if (tiles == 1) {
        tile_scan<T, false><<<1, kThreads, 0, ctx->stream>>>(in, out, n, nullptr, d_total);
        HIP_TRY(hipGetLastError());
        return RTG_OK;
    }
    T *partial = reinterpret_cast<T *>(work);
    char *next = work + rtg::Arena::round(sizeof(T) * (size_t)tiles);
    tile_reduce<T><<<(unsigned)tiles, kThreads, 0, ctx->stream>>>(in, n, partial);
    HIP_TRY(hipGetLastError());
    if (tiles <= kTile) {  // each block adds up the tile sums before it (at most kTile - 1)
        tile_scan<T, true><<<(unsigned)tiles, kThreads, 0, ctx->stream>>>(in, out, n, partial, d_total);
        HIP_TRY(hipGetLastError());
        return RTG_OK;
    }
    RTG_TRY(scan_impl<T>(ctx, partial, partial, tiles, (T *)nullptr, next));
    tile_scan<T, false><<<(unsigned)tiles, kThreads, 0, ctx->stream>>>(in, out, n, partial, d_total);
    HIP_TRY(hipGetLastError());
    return RTG_OK;
}

}  // namespace

namespace rtg {

size_t scan_work_bytes(int64_t n) { return work_bytes<uint64_t>(n); }

int scan_exclusive_u32_w(rtg_ctx *ctx, const uint32_t *in, uint32_t *out, int64_t n,
                         uint32_t *d_total, void *work) {
    return scan_impl<uint32_t>(ctx, in, out, n, d_total, (char *)work);
}

int scan_exclusive_u64_w(rtg_ctx *ctx, const uint64_t *in, uint64_t *out, int64_t n,
                         uint64_t *d_total, void *work) {
    return scan_impl<uint64_t>(ctx, in, out, n, d_total, (char *)work);
}

}  // namespace rtg

namespace rtg {
// Loads this file's kernels onto the current device (an attribute query of one of them).
int preload_scan() {
    hipFuncAttributes attr;
    return hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&tile_reduce<uint32_t>)) == hipSuccess ? RTG_OK : RTG_EHIP;
}
}  // namespace rtg
