// Stable LSD radix sort of uint64 keys for the CSR build (sorting the packed (src << b | dst)
// edge keys, the role thrust::sort plays in ThrustSortEdgesInBatch,
// tc/graph/thrust_helper.cu:16-42), and of 32-bit keys, unsigned or signed (the device sort of
// the lookup array, thrust::sort in BSAccessCountFunction, rt_search/simple_bs.cu:99-108; a
// signed key's sign bit is flipped in the pass that holds it, so negative values sort first).
//
// Digits are 8 or 9 bits wide: a sort on B bits takes the fewest passes that 9-bit digits
// allow, as many of them 9 bits wide as needed and the rest 8 (42 bits: 9+9+8+8+8, five passes
// instead of six), since every pass costs a full read-and-write of the keys.
//
// Per pass: (1) every 4096-key tile writes its digit histogram, digit-major; (2) one exclusive
// scan over the histograms gives each (digit, tile) its global start; (3) every tile ranks its
// keys stably — each wave64 owns 1024 consecutive keys, peers with the same digit are found
// with one ballot per digit bit, and per-wave LDS counters carry the running count — then
// reorders the tile in LDS by digit and writes it out in digit runs, so that stores are mostly
// contiguous. Traffic per pass: 2 reads + 1 write of the key size per key.
#include "device.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;  // 4096 keys
constexpr int kMaxDigitBits = 9;

// dxor: XORed into every digit of the pass (the sign bit's position when the pass holds a
// signed key's sign bit, else 0).
template <class K, int kDB>
__global__ __launch_bounds__(kThreads) void digit_histogram(const K *__restrict__ in,
                                                            int64_t n, int shift, uint32_t dxor,
                                                            uint32_t *__restrict__ hist,
                                                            uint32_t ntiles) {
    constexpr int kBins = 1 << kDB;
    constexpr uint32_t kMask = kBins - 1;
    __shared__ uint32_t h[kBins];
#pragma unroll
    for (int b = 0; b < kBins / kThreads; ++b) h[b * kThreads + threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kTile;
    K k[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int64_t idx = base + (int64_t)i * kThreads + threadIdx.x;
        k[i] = idx < n ? in[idx] : (K)0;
    }
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int64_t idx = base + (int64_t)i * kThreads + threadIdx.x;
        if (idx < n) atomicAdd(&h[((uint32_t)(k[i] >> shift) & kMask) ^ dxor], 1u);
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kBins / kThreads; ++b) {
        const int d = b * kThreads + threadIdx.x;
        hist[(size_t)d * ntiles + blockIdx.x] = h[d];
    }
}

template <class K, int kDB>
__global__ __launch_bounds__(kThreads) void rank_scatter(const K *__restrict__ in,
                                                         K *__restrict__ out, int64_t n,
                                                         int shift, uint32_t dxor,
                                                         const uint32_t *__restrict__ starts,
                                                         uint32_t ntiles) {
    constexpr int kBins = 1 << kDB;
    constexpr uint32_t kMask = kBins - 1;
    constexpr int kPer = kBins / kThreads;  // digits per thread in the block-wide digit scan
    __shared__ uint32_t wave_cnt[kWaves][kBins];
    __shared__ uint32_t digit_start[kBins];
    __shared__ uint32_t global_minus_local[kBins];
    __shared__ uint32_t wsum[kWaves];
    __shared__ K tile_keys[kTile];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int b = 0; b < kBins / 64; ++b) wave_cnt[wave][b * 64 + lane] = 0;
    __syncthreads();

    const int64_t tile_base = (int64_t)blockIdx.x * kTile;
    const int64_t wave_base = tile_base + (int64_t)wave * (64 * kItems);
    const uint64_t lt = rtgd::lanemask_lt();
    K key[kItems];
    uint32_t rank[kItems];
    // All of the wave's loads first: the ranking below has wave barriers between items, which
    // would otherwise hold each load back until the previous item is ranked.
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int64_t idx = wave_base + (int64_t)i * 64 + lane;
        key[i] = idx < n ? in[idx] : (K)0;
    }
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int64_t idx = wave_base + (int64_t)i * 64 + lane;
        const bool valid = idx < n;
        const K k = key[i];
        const uint32_t dg = ((uint32_t)(k >> shift) & kMask) ^ dxor;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kDB; ++b) {
            const bool bit = (dg >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t before = 0;
        if (valid) before = wave_cnt[wave][dg];
        __builtin_amdgcn_wave_barrier();
        const bool leader = valid && ((peers >> lane) == 1ull);
        if (leader) wave_cnt[wave][dg] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[i] = before + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();

    // Digits [kPer * t, kPer * t + kPer) of thread t: per-wave exclusive prefix (in place) and
    // the block total, then an exclusive scan of the totals over all digits -> digit_start.
    uint32_t tot[kPer], tsum = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int d = threadIdx.x * kPer + k;
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = wave_cnt[w][d];
            wave_cnt[w][d] = s;
            s += c;
        }
        tot[k] = s;
        tsum += s;
    }
    const uint32_t inc = rtgd::wave_inclusive_scan(tsum);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t ds = inc - tsum;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
        if (w < wave) ds += wsum[w];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int d = threadIdx.x * kPer + k;
        digit_start[d] = ds;
        global_minus_local[d] = starts[(size_t)d * ntiles + blockIdx.x] - ds;
        ds += tot[k];
    }
    __syncthreads();

#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int64_t idx = wave_base + (int64_t)i * 64 + lane;
        if (idx < n) {
            const uint32_t dg = ((uint32_t)(key[i] >> shift) & kMask) ^ dxor;
            tile_keys[digit_start[dg] + wave_cnt[wave][dg] + rank[i]] = key[i];
        }
    }
    __syncthreads();
    const int64_t count = (n - tile_base) < kTile ? (n - tile_base) : kTile;
    for (int j = threadIdx.x; j < count; j += kThreads) {
        const K k = tile_keys[j];
        const uint32_t dg = ((uint32_t)(k >> shift) & kMask) ^ dxor;
        out[(size_t)(global_minus_local[dg] + (uint32_t)j)] = k;
    }
}

// One pass of digit width kDB at `shift`.
template <class K, int kDB>
int sort_pass(rtg_ctx *ctx, const K *in, K *out, int64_t n, int shift, uint32_t dxor, uint32_t *hist,
              void *scan_work) {
    const int64_t tiles = rtg::ceil_div(n, kTile);
    digit_histogram<K, kDB><<<(unsigned)tiles, kThreads, 0, ctx->stream>>>(in, n, shift, dxor, hist, (uint32_t)tiles);
    HIP_TRY(hipGetLastError());
    RTG_TRY(rtg::scan_exclusive_u32_w(ctx, hist, hist, tiles << kDB, nullptr, scan_work));
    rank_scatter<K, kDB><<<(unsigned)tiles, kThreads, 0, ctx->stream>>>(in, out, n, shift, dxor, hist, (uint32_t)tiles);
    HIP_TRY(hipGetLastError());
    return RTG_OK;
}

}  // namespace

namespace rtg {

size_t radix_sort_work_bytes(int64_t n) {
    const int64_t tiles = std::max<int64_t>(1, ceil_div(n, kTile));
    const int64_t h = tiles << kMaxDigitBits;
    return Arena::round(sizeof(uint32_t) * (size_t)h) + scan_work_bytes(h) + 256;
}

// Digit widths of a sort on `bits` bits: the pass count 9-bit digits need, the first
// (bits - 8 * passes) of them 9 bits wide (when positive), the others 8.
int radix_digit_plan(int bits, int widths[16]) {
    if (bits <= 0) return 0;
    const int passes = (bits + kMaxDigitBits - 1) / kMaxDigitBits;
    const int nine = std::max(0, bits - 8 * passes);
    for (int p = 0; p < passes; ++p) widths[p] = p < nine ? 9 : 8;
    return passes;
}

template <class K>
int radix_sort_t(rtg_ctx *ctx, K *keys[2], int64_t n, int bits, bool sign_top, int *which, int *passes,
                 void *work) {
    *which = 0;
    if (passes) *passes = 0;
    if (n <= 1 || bits <= 0) return RTG_OK;
    if (n >= (int64_t)1 << 32) return set_error(RTG_EINVAL, "radix sort: more than 2^32 keys");
    const int64_t tiles = ceil_div(n, kTile);
    uint32_t *hist = reinterpret_cast<uint32_t *>(work);
    char *scan_work = reinterpret_cast<char *>(work) + Arena::round(sizeof(uint32_t) * (size_t)(tiles << kMaxDigitBits));
    int widths[16];
    const int np = radix_digit_plan(bits, widths);
    int cur = 0, shift = 0;
    for (int p = 0; p < np; ++p) {
        const int w = widths[p];
        // the sign bit (bits - 1) lies in this pass: flip it inside the digit
        const uint32_t dxor = (sign_top && shift + w >= bits) ? (1u << (bits - 1 - shift)) : 0u;
        if (w == 9)
            RTG_TRY((sort_pass<K, 9>(ctx, keys[cur], keys[cur ^ 1], n, shift, dxor, hist, scan_work)));
        else
            RTG_TRY((sort_pass<K, 8>(ctx, keys[cur], keys[cur ^ 1], n, shift, dxor, hist, scan_work)));
        shift += w;
        cur ^= 1;
    }
    *which = cur;
    if (passes) *passes = np;
    return RTG_OK;
}

int radix_sort_u64(rtg_ctx *ctx, uint64_t *keys[2], int64_t n, int bits, int *which,
                   int *passes, void *work) {
    return radix_sort_t<uint64_t>(ctx, keys, n, bits, false, which, passes, work);
}

int radix_sort_32(rtg_ctx *ctx, uint32_t *keys[2], int64_t n, bool is_signed, int *which, int *passes,
                  void *work) {
    return radix_sort_t<uint32_t>(ctx, keys, n, 32, is_signed, which, passes, work);
}

}  // namespace rtg

namespace rtg {
// Loads this file's kernels onto the current device (an attribute query of one of them).
int preload_radix_sort() {
    hipFuncAttributes attr;
    return hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&digit_histogram<uint64_t, 8>)) == hipSuccess ? RTG_OK : RTG_EHIP;
}
}  // namespace rtg
