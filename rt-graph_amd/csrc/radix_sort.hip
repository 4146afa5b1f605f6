// Stable LSD radix sort of uint64 keys, 8-bit digits, for the CSR build (sorting the packed
// (src << b | dst) edge keys, the role thrust::sort plays in ThrustSortEdgesInBatch,
// tc/graph/thrust_helper.cu:16-42), and of 32-bit keys, unsigned or signed (the device sort of
// the lookup array, thrust::sort in BSAccessCountFunction, rt_search/simple_bs.cu:99-108; a
// signed key's top digit has its sign bit flipped, so negative values sort first).
//
// Per pass: (1) every 4096-key tile writes its 256-bin digit histogram, digit-major;
// (2) one exclusive scan over the histograms gives each (digit, tile) its global start;
// (3) every tile ranks its keys stably — each wave64 owns 1024 consecutive keys, peers with
// the same digit are found with eight ballots, and per-wave LDS counters carry the running
// count — then reorders the tile in LDS by digit and writes it out in digit runs, so that
// stores are mostly contiguous. Traffic per pass: 2 reads + 1 write of the key size per key.
#include "device.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;  // 4096 keys
constexpr int kBins = 256;

// dxor: XORed into every digit of the pass (0x80 on a signed key's top digit, else 0).
template <class K>
__global__ __launch_bounds__(kThreads) void digit_histogram(const K *__restrict__ in,
                                                            int64_t n, int shift, uint32_t dxor,
                                                            uint32_t *__restrict__ hist,
                                                            uint32_t ntiles) {
    __shared__ uint32_t h[kBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        int64_t idx = base + (int64_t)i * kThreads + threadIdx.x;
        if (idx < n) atomicAdd(&h[((uint32_t)(in[idx] >> shift) & 0xFFu) ^ dxor], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

template <class K>
__global__ __launch_bounds__(kThreads) void rank_scatter(const K *__restrict__ in,
                                                         K *__restrict__ out, int64_t n,
                                                         int shift, uint32_t dxor,
                                                         const uint32_t *__restrict__ starts,
                                                         uint32_t ntiles) {
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
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int64_t idx = wave_base + (int64_t)i * 64 + lane;
        const bool valid = idx < n;
        const K k = valid ? in[idx] : (K)0;
        const uint32_t dg = ((uint32_t)(k >> shift) & 0xFFu) ^ dxor;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dg >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        key[i] = k;
        uint32_t before = 0;
        if (valid) before = wave_cnt[wave][dg];
        __builtin_amdgcn_wave_barrier();
        const bool leader = valid && ((peers >> lane) == 1ull);
        if (leader) wave_cnt[wave][dg] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[i] = before + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();

    // Digit d: per-wave exclusive prefix (in place) and the block total.
    const int d = threadIdx.x;
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        uint32_t c = wave_cnt[w][d];
        wave_cnt[w][d] = tot;
        tot += c;
    }
    // Exclusive scan of block totals over digits -> digit_start.
    uint32_t inc = rtgd::wave_inclusive_scan(tot);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t woff = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
        if (w < wave) woff += wsum[w];
    const uint32_t ds = woff + inc - tot;
    digit_start[d] = ds;
    global_minus_local[d] = starts[(size_t)d * ntiles + blockIdx.x] - ds;
    __syncthreads();

#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int64_t idx = wave_base + (int64_t)i * 64 + lane;
        if (idx < n) {
            const uint32_t dg = ((uint32_t)(key[i] >> shift) & 0xFFu) ^ dxor;
            tile_keys[digit_start[dg] + wave_cnt[wave][dg] + rank[i]] = key[i];
        }
    }
    __syncthreads();
    const int64_t count = (n - tile_base) < kTile ? (n - tile_base) : kTile;
    for (int j = threadIdx.x; j < count; j += kThreads) {
        const K k = tile_keys[j];
        const uint32_t dg = ((uint32_t)(k >> shift) & 0xFFu) ^ dxor;
        out[(size_t)(global_minus_local[dg] + (uint32_t)j)] = k;
    }
}

}  // namespace

namespace rtg {

size_t radix_sort_work_bytes(int64_t n) {
    const int64_t tiles = std::max<int64_t>(1, ceil_div(n, kTile));
    const int64_t h = tiles * kBins;
    return Arena::round(sizeof(uint32_t) * (size_t)h) + scan_work_bytes(h) + 256;
}

template <class K>
int radix_sort_t(rtg_ctx *ctx, K *keys[2], int64_t n, int bits, bool sign_top, int *which, int *passes,
                 void *work) {
    *which = 0;
    if (passes) *passes = 0;
    if (n <= 1 || bits <= 0) return RTG_OK;
    if (n >= (int64_t)1 << 32) return set_error(RTG_EINVAL, "radix sort: more than 2^32 keys");
    const int64_t tiles = ceil_div(n, kTile);
    const int64_t h = tiles * kBins;
    uint32_t *hist = reinterpret_cast<uint32_t *>(work);
    char *scan_work = reinterpret_cast<char *>(work) + Arena::round(sizeof(uint32_t) * (size_t)h);
    int cur = 0, np = 0;
    for (int shift = 0; shift < bits; shift += 8, ++np) {
        const uint32_t dxor = (sign_top && shift + 8 >= bits) ? 0x80u : 0u;
        digit_histogram<K><<<(unsigned)tiles, kThreads, 0, ctx->stream>>>(keys[cur], n, shift, dxor, hist,
                                                                          (uint32_t)tiles);
        HIP_TRY(hipGetLastError());
        RTG_TRY(scan_exclusive_u32_w(ctx, hist, hist, h, nullptr, scan_work));
        rank_scatter<K><<<(unsigned)tiles, kThreads, 0, ctx->stream>>>(keys[cur], keys[cur ^ 1], n, shift, dxor,
                                                                       hist, (uint32_t)tiles);
        HIP_TRY(hipGetLastError());
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
