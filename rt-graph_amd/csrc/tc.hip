// Triangle counting on the degree-oriented CSR.
//
// Plain kernels (single launch, no bucketing):
//   tc_plain_thread  thread per oriented edge, the 1A2 form of ListIntersectionKernel
//                    (tc/bs_tc/bs_tc.cu:120-148): for w in N+(u), BinaryCheck(N+(w), v); the
//                    per-edge row comes from a precomputed row[] instead of the reference's
//                    binary search over offsets (:126-134), and counts are reduced per wave
//                    before one 64-bit atomic instead of one 32-bit atomic per edge (:145).
//   tc_plain_wave    wave64 per oriented edge: every lane binary-searches elements of the
//                    shorter of N+(u), N+(v) in the longer.
//
// Batched kernel (tc_batched): every edge (u,v) with both lists non-empty is assigned to its
// owner — the endpoint with the longer out-list — and grouped by owner at plan time (a stable
// radix sort on the owner, so each owner's partners stay in edge order). A
// persistent grid pulls items from a queue ordered heaviest first. For an item the owner's
// out-list is inserted into an LDS open-addressing hash table (block-wide for large owners,
// wave-private for small ones), then the partners' out-lists — always the shorter side — are
// streamed with coalesced loads, flattened across the 64 lanes of a wave, and probed. Work per
// edge is min(d+(u), d+(v)) LDS probes and 4*min bytes streamed; the longer list is read once
// per item instead of once per edge. The per-edge instantiation (kPerEdge) also stores every
// partner span's count at its edge.
#include "tc_kernels.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <cstring>
#include <string>
#include <numeric>

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
// Fourth word of a raw buffer resource on gfx9-family GPUs (gfx950 included): a 32-bit data
// format, which enables the range check that makes loads at offsets >= num_records return 0.
constexpr int kBufferWord3 = 0x00020000;
constexpr int kPlainThreads = 256;

// row[e] = source vertex of edge e. A block takes 256 consecutive vertices, whose edges form one
// contiguous range, and writes that range with coalesced stores: each thread walks its strided
// edges in order and advances a cursor over the block's row offsets (in LDS).
__global__ __launch_bounds__(256) void edge_rows_kernel(const int32_t *__restrict__ off, int32_t n,
                                                        int32_t *__restrict__ row) {
    __shared__ int32_t so[257];
    for (int64_t v0 = (int64_t)blockIdx.x * 256; v0 < n; v0 += (int64_t)gridDim.x * 256) {
        const int32_t nv = (int32_t)min<int64_t>(256, n - v0);
        for (int i = threadIdx.x; i <= nv; i += 256) so[i] = off[v0 + i];
        __syncthreads();
        const int32_t e1 = so[nv];
        int32_t v = 0;  // so[v] <= e < so[v + 1]
        for (int32_t e = so[0] + threadIdx.x; e < e1; e += 256) {
            while (so[v + 1] <= e) ++v;
            row[e] = (int32_t)v0 + v;
        }
        __syncthreads();  // so is refilled for the next vertex range
    }
}

// [0] count, [1] the general queue's head, [2] the bitmap queue's, [3] the wave items' queue's.
__global__ void reset_counters(unsigned long long *c) {
    if (threadIdx.x < 4) c[threadIdx.x] = 0ull;
}

// BinaryCheck, tc/bs_tc/bs_tc.cu:14-25, on data[0 .. len-1].
__device__ __forceinline__ uint32_t binary_check(const int32_t *__restrict__ data, int32_t len, int32_t val) {
    int32_t l = 0, r = len - 1;
    while (l <= r) {
        const int32_t mid = (l + r) >> 1;
        const int32_t x = data[mid];
        if (x == val) return 1u;
        if (x < val) l = mid + 1;
        else r = mid - 1;
    }
    return 0u;
}

__global__ __launch_bounds__(kPlainThreads) void tc_plain_thread(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj, const int32_t *__restrict__ row,
    int64_t e_begin, int64_t e_end, unsigned long long *count) {
    uint64_t local = 0;
    for (int64_t e = e_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < e_end;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t u = row[e], v = adj[e];
        const int32_t b = off[u], en = off[u + 1];
        uint32_t cnt = 0;
        for (int32_t i = b; i < en; ++i) {
            const int32_t w = adj[i];
            const int32_t wb = off[w], wl = off[w + 1] - wb;
            if (wl) cnt += binary_check(adj + wb, wl, v);
        }
        local += cnt;
    }
    local = rtgd::wave_sum_u64(local);
    if (__lane_id() == 0 && local) atomicAdd(count, (unsigned long long)local);
}

__global__ __launch_bounds__(kPlainThreads) void tc_plain_wave(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj, const int32_t *__restrict__ row,
    int64_t e_begin, int64_t e_end, unsigned long long *count) {
    const int lane = __lane_id();
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    uint64_t local = 0;
    for (int64_t e = e_begin + (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); e < e_end; e += nw) {
        const int32_t u = row[e], v = adj[e];
        int32_t a0 = off[u], na = off[u + 1] - a0;
        int32_t b0 = off[v], nb = off[v + 1] - b0;
        if (na > nb) {
            int32_t t = a0; a0 = b0; b0 = t;
            t = na; na = nb; nb = t;
        }
        if (nb == 0) continue;
        for (int32_t i = lane; i < na; i += 64) local += binary_check(adj + b0, nb, adj[a0 + i]);
    }
    local = rtgd::wave_sum_u64(local);
    if (lane == 0 && local) atomicAdd(count, (unsigned long long)local);
}

// Per-edge counts, intersection attribution: out[e] = |N+(u) & N+(v)| (wave per edge).
__global__ __launch_bounds__(kPlainThreads) void tc_per_edge_intersect(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj, const int32_t *__restrict__ row,
    int64_t m, uint32_t *__restrict__ out) {
    const int lane = __lane_id();
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t e = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); e < m; e += nw) {
        const int32_t u = row[e], v = adj[e];
        int32_t a0 = off[u], na = off[u + 1] - a0;
        int32_t b0 = off[v], nb = off[v + 1] - b0;
        if (na > nb) {
            int32_t t = a0; a0 = b0; b0 = t;
            t = na; na = nb; nb = t;
        }
        uint32_t c = 0;
        if (nb > 0)
            for (int32_t i = lane; i < na; i += 64) c += binary_check(adj + b0, nb, adj[a0 + i]);
        c = rtgd::wave_sum_u32(c);
        if (lane == 0) out[e] = c;
    }
}

// Per-edge counts, 1A2 attribution of ListIntersectionKernel (bs_tc.cu:135-145).
__global__ __launch_bounds__(kPlainThreads) void tc_per_edge_1a2(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj, const int32_t *__restrict__ row,
    int64_t m, uint32_t *__restrict__ out) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t u = row[e], v = adj[e];
        uint32_t cnt = 0;
        for (int32_t i = off[u]; i < off[u + 1]; ++i) {
            const int32_t w = adj[i];
            const int32_t wb = off[w], wl = off[w + 1] - wb;
            if (wl) cnt += binary_check(adj + wb, wl, v);
        }
        out[e] = cnt;
    }
}

// ---------------------------------------------------------------------------------------------
// Batched kernel.
//
// The owner's out-list goes into a two-choice cuckoo table in LDS: 2*T slots, x stored at
// slot h1(x) of the first half or h2(x) of the second, so that a lookup is exactly two LDS
// reads with no probe loop (a linear-probing loop runs until the longest chain among the 64
// lanes ends, which dominated the instruction count). T is about 1.25 * d, so the load is at
// most 0.4; the parallel insertion (atomic exchanges, bounded eviction chains) is retried with
// new hash multipliers on failure and, after kRetries failures, the item falls back to linear
// probing over the same slots.

constexpr int kRetries = 4;
constexpr int32_t kMaxTableSlots = 150 * 1024 / 4;  // LDS table slots one workgroup may use
constexpr int kEvict = 64;

// Slots per half for an out-list of length d.
__host__ __device__ __forceinline__ uint32_t cuckoo_half(int32_t d) {
    uint32_t t = (uint32_t)d + ((uint32_t)d >> 2) + 1u;
    return (t + 31u) & ~31u;
}

__device__ __forceinline__ uint32_t hmul(uint32_t x, uint32_t k, uint32_t t) { return __umulhi(x * k, t); }

// Probe modes of an owner table: two-choice cuckoo hash, linear probing (the hash's fallback),
// a bitmap over the owner's value range, the owner's sorted list, and (degree-ordered copies
// only, in place of kBitmap) a bitmap from 0 that covers at least every 16-bit value.
constexpr uint32_t kCuckoo = 0, kLinearProbe = 1, kBitmap = 2, kSorted = 3, kBitmap16 = 4;

struct Table {
    uint32_t *tab;
    uint32_t t;      // cuckoo: slots per half; linear: half the slots; bitmap: value range
                     // (bits); sorted: list length
    uint32_t k1, k2; // hash multipliers; bitmap: k1 = the owner's smallest value
    uint32_t mode;   // kCuckoo, kLinearProbe, kBitmap, kSorted
};

__device__ __forceinline__ bool cuckoo_insert(const Table &c, uint32_t x) {
    uint32_t cur = x;
    for (int it = 0; it < kEvict; ++it) {
        uint32_t old = atomicExch(&c.tab[hmul(cur, c.k1, c.t)], cur);
        if (old == kEmpty) return true;
        cur = old;
        old = atomicExch(&c.tab[c.t + hmul(cur, c.k2, c.t)], cur);
        if (old == kEmpty) return true;
        cur = old;
    }
    return false;
}

__device__ __forceinline__ void linear_insert(const Table &c, uint32_t x) {
    const uint32_t size = 2u * c.t;
    uint32_t h = hmul(x, c.k1, size);
    while (true) {
        const uint32_t old = atomicCAS(&c.tab[h], kEmpty, x);
        if (old == kEmpty) return;
        h = (h + 1u == size) ? 0u : h + 1u;
    }
}

__device__ __forceinline__ uint32_t cuckoo_has(const Table &c, uint32_t x) {
    const uint32_t a = c.tab[hmul(x, c.k1, c.t)];
    const uint32_t b = c.tab[c.t + hmul(x, c.k2, c.t)];
    return (uint32_t)(a == x) | (uint32_t)(b == x);
}

// Words of a bitmap over [lo, hi]: bits 0 .. t = hi - lo + 1, bit t (one past the range) always
// clear, so that a probe can clamp instead of branching.
__host__ __device__ __forceinline__ uint32_t bitmap_words(uint32_t lo, uint32_t hi) { return ((hi - lo + 1u) >> 5) + 1u; }

// Values outside [lo, hi] (below lo they wrap to large offsets) are clamped to bit t, which is
// clear: a branch-free probe whose LDS reads all issue back to back.
__device__ __forceinline__ uint32_t bitmap_has(const Table &c, uint32_t x) {
    const uint32_t o = min(x - c.k1, c.t);
    // bit o of the table: word o >> 5, then the bit field extract's offset is o's low 5 bits
    return __builtin_amdgcn_ubfe(c.tab[__builtin_amdgcn_ubfe(o, 5u, 27u)], o, 1u);
}

// A bitmap from 0 over [0, t) with bit t clear, t = max(hi + 1, 2^16) for the owner's largest
// value hi: a 16-bit value needs neither an offset nor a clamp.
__host__ __device__ __forceinline__ uint32_t bitmap16_range(uint32_t hi) { return hi + 1u > 65536u ? hi + 1u : 65536u; }
__host__ __device__ __forceinline__ uint32_t bitmap16_words(uint32_t hi) { return (bitmap16_range(hi) >> 5) + 1u; }

// Bit o of a bitmap at LDS address `base`: the word's address in one shift-add (the compiler
// otherwise spends a shift, a mask and an add), one LDS read, one bit field extract.
__device__ __forceinline__ uint32_t bitmap_bit(uint32_t base, uint32_t o) {
    uint32_t addr;
    asm("v_lshl_add_u32 %0, %1, 2, %2" : "=v"(addr) : "v"(o >> 5), "v"(base));
    const uint32_t w = *reinterpret_cast<const __attribute__((address_space(3))) uint32_t *>((uintptr_t)addr);
    return __builtin_amdgcn_ubfe(w, o, 1u);
}

__device__ __forceinline__ uint32_t lds_address(const uint32_t *p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) uint32_t *)p;
}

// kBitmap16 probes: any value (clamped to the clear bit t), and a 16-bit value (no clamp).
__device__ __forceinline__ uint32_t bitmap16_has(const Table &c, uint32_t x) { return bitmap_bit(lds_address(c.tab), min(x, c.t)); }
__device__ __forceinline__ uint32_t bitmap16_has16(const Table &c, uint32_t x) { return bitmap_bit(lds_address(c.tab), x); }

// First index of tab[lo, lo + len) whose value is not below x (lo + len if none): binary lifting
// whose trip count depends on len only, so the lanes of a wave stay together.
__device__ __forceinline__ uint32_t lds_lower_bound(const uint32_t *tab, uint32_t lo, uint32_t len, uint32_t x) {
    uint32_t pos = lo;
    const uint32_t end = lo + len;
    for (uint32_t step = len ? 1u << (31 - __clz(len)) : 0u; step; step >>= 1)
        if (pos + step <= end && tab[pos + step - 1] < x) pos += step;
    return pos;
}

// x in the sorted list tab[0, t) (t >= 1).
__device__ __forceinline__ uint32_t sorted_has(const Table &c, uint32_t x) {
    const uint32_t p = lds_lower_bound(c.tab, 0, c.t, x);
    return (p < c.t && c.tab[min(p, c.t - 1)] == x) ? 1u : 0u;
}

__device__ __forceinline__ uint32_t linear_has(const Table &c, uint32_t x) {
    const uint32_t size = 2u * c.t;
    uint32_t h = hmul(x, c.k1, size);
    while (true) {
        const uint32_t t = c.tab[h];
        if (t == x) return 1u;
        if (t == kEmpty) return 0u;
        h = (h + 1u == size) ? 0u : h + 1u;
    }
}

__device__ __forceinline__ void table_seed(Table &c, int r) {
    c.k1 = 0x9E3779B1u + 0x7F4A7C16u * (uint32_t)r;
    c.k2 = (0x85EBCA77u ^ (0xC2B2AE3Du * (uint32_t)(r + 1))) | 1u;
    c.k1 |= 1u;
}

// Block-wide build of the owner table from adj[xb .. xb+d). Returns the table (uniform).
template <int kThreads>
__device__ Table build_block_table(uint32_t *tab, const int32_t *__restrict__ adj, int32_t xb, int32_t d,
                                   uint32_t *fail_flag) {
    Table c{tab, cuckoo_half(d), 0, 0, kCuckoo};
    for (int r = 0; r < kRetries; ++r) {
        table_seed(c, r);
        for (uint32_t i = threadIdx.x; i < 2u * c.t; i += kThreads) tab[i] = kEmpty;
        if (threadIdx.x == 0) *fail_flag = 0u;
        __syncthreads();
        bool ok = true;
        for (int32_t i = threadIdx.x; i < d; i += kThreads) ok &= cuckoo_insert(c, (uint32_t)adj[xb + i]);
        if (!ok) *fail_flag = 1u;
        __syncthreads();
        if (rtgd::uniform_u32(*fail_flag) == 0u) return c;
        __syncthreads();  // everyone has read the flag before it is reset
    }
    c.mode = kLinearProbe;
    for (uint32_t i = threadIdx.x; i < 2u * c.t; i += kThreads) tab[i] = kEmpty;
    __syncthreads();
    for (int32_t i = threadIdx.x; i < d; i += kThreads) linear_insert(c, (uint32_t)adj[xb + i]);
    __syncthreads();
    return c;
}

// Block-wide bitmap of adj[xb .. xb+d) over its value range [adj[xb], adj[xb+d-1]] (the caller
// has checked that the range's words fit the table).
template <int kThreads>
__device__ Table build_block_bitmap(uint32_t *tab, const int32_t *__restrict__ adj, int32_t xb, int32_t d) {
    const uint32_t lo = (uint32_t)adj[xb], hi = (uint32_t)adj[xb + d - 1];
    const Table c{tab, hi - lo + 1u, lo, 0u, kBitmap};
    const uint32_t words = bitmap_words(lo, hi);
    for (uint32_t i = threadIdx.x; i < words; i += kThreads) tab[i] = 0u;
    __syncthreads();
    for (int32_t i = threadIdx.x; i < d; i += kThreads) {
        const uint32_t o = (uint32_t)adj[xb + i] - lo;
        atomicOr(&tab[o >> 5], 1u << (o & 31u));
    }
    __syncthreads();
    return c;
}

// Block-wide bitmap from 0 of adj[xb .. xb+d) (kBitmap16; the caller has checked that its
// bitmap16_words fit the table).
template <int kThreads>
__device__ Table build_block_bitmap16(uint32_t *tab, const int32_t *__restrict__ adj, int32_t xb, int32_t d) {
    const uint32_t hi = (uint32_t)adj[xb + d - 1];
    const uint32_t words = bitmap16_words(hi);
    for (uint32_t i = threadIdx.x; i < words; i += kThreads) tab[i] = 0u;
    __syncthreads();
    for (int32_t i = threadIdx.x; i < d; i += kThreads) {
        const uint32_t o = (uint32_t)adj[xb + i];
        atomicOr(&tab[o >> 5], 1u << (o & 31u));
    }
    __syncthreads();
    return Table{tab, bitmap16_range(hi), 0u, 0u, kBitmap16};
}

// Block-wide copy of the sorted list adj[xb .. xb+d) (d fits the table).
template <int kThreads>
__device__ Table build_block_sorted(uint32_t *tab, const int32_t *__restrict__ adj, int32_t xb, int32_t d) {
    for (int32_t i = threadIdx.x; i < d; i += kThreads) tab[i] = (uint32_t)adj[xb + i];
    __syncthreads();
    return Table{tab, (uint32_t)d, 0u, 0u, kSorted};
}

// Wave-private build (small owners).
__device__ Table build_wave_table(uint32_t *tab, const int32_t *__restrict__ adj, int32_t xb, int32_t d) {
    const int lane = __lane_id();
    Table c{tab, cuckoo_half(d), 0, 0, kCuckoo};
    for (int r = 0; r < kRetries; ++r) {
        table_seed(c, r);
        for (uint32_t i = lane; i < 2u * c.t; i += 64) tab[i] = kEmpty;
        rtgd::wave_sync_lds();
        bool ok = true;
        for (int32_t i = lane; i < d; i += 64) ok &= cuckoo_insert(c, (uint32_t)adj[xb + i]);
        rtgd::wave_sync_lds();
        if (__ballot(!ok) == 0ull) return c;
    }
    c.mode = kLinearProbe;
    for (uint32_t i = lane; i < 2u * c.t; i += 64) tab[i] = kEmpty;
    rtgd::wave_sync_lds();
    for (int32_t i = lane; i < d; i += 64) linear_insert(c, (uint32_t)adj[xb + i]);
    rtgd::wave_sync_lds();
    return c;
}

// Wave-private bitmap or sorted copy (the caller has checked that it fits `cap` words; else it
// builds the hash table). kB16: bitmaps from 0 (kBitmap16).
template <bool kB16 = false>
__device__ Table build_wave_other(uint32_t *tab, const int32_t *__restrict__ adj, int32_t xb, int32_t d, uint32_t mode,
                                  uint32_t cap) {
    const int lane = __lane_id();
    if (kB16 && mode == kBitmap) {
        const uint32_t hi = (uint32_t)adj[xb + d - 1];
        const uint32_t words = bitmap16_words(hi);
        if (words <= cap) {
            for (uint32_t i = lane; i < words; i += 64) tab[i] = 0u;
            rtgd::wave_sync_lds();
            for (int32_t i = lane; i < d; i += 64) {
                const uint32_t o = (uint32_t)adj[xb + i];
                atomicOr(&tab[o >> 5], 1u << (o & 31u));
            }
            rtgd::wave_sync_lds();
            return Table{tab, bitmap16_range(hi), 0u, 0u, kBitmap16};
        }
    } else if (mode == kBitmap) {
        const uint32_t lo = (uint32_t)adj[xb], hi = (uint32_t)adj[xb + d - 1];
        const uint32_t words = bitmap_words(lo, hi);
        if (words <= cap) {
            for (uint32_t i = lane; i < words; i += 64) tab[i] = 0u;
            rtgd::wave_sync_lds();
            for (int32_t i = lane; i < d; i += 64) {
                const uint32_t o = (uint32_t)adj[xb + i] - lo;
                atomicOr(&tab[o >> 5], 1u << (o & 31u));
            }
            rtgd::wave_sync_lds();
            return Table{tab, hi - lo + 1u, lo, 0u, kBitmap};
        }
    } else if (mode == kSorted && (uint32_t)d <= cap) {
        for (int32_t i = lane; i < d; i += 64) tab[i] = (uint32_t)adj[xb + i];
        rtgd::wave_sync_lds();
        return Table{tab, (uint32_t)d, 0u, 0u, kSorted};
    }
    return build_wave_table(tab, adj, xb, d);
}

template <uint32_t kMode>
__device__ __forceinline__ uint32_t probe(const Table &c, uint32_t x) {
    if constexpr (kMode == kLinearProbe) return linear_has(c, x);
    else if constexpr (kMode == kBitmap) return bitmap_has(c, x);
    else if constexpr (kMode == kBitmap16) return bitmap16_has(c, x);
    else if constexpr (kMode == kSorted) return sorted_has(c, x);
    else return cuckoo_has(c, x);
}

// Owners too long for any LDS table: every element of the partner lists wspan[p0 .. p1) is
// binary-searched in the owner's sorted list own[0 .. d) in HBM. Kept out of line and compact
// (no unrolling): it is rare, and a second copy of the streaming loop would only crowd the
// instruction cache of the common path. With eout set, each span's count is also stored at
// eout[pi] (per-edge counts, by span position).
template <bool kPerEdge>
__device__ __noinline__ uint32_t stream_search(const int32_t *__restrict__ adj, const int2 *__restrict__ wspan,
                                               int64_t p0, int64_t p1, const int32_t *__restrict__ own, int32_t d,
                                               uint32_t *__restrict__ eout) {
    uint32_t hits = 0;
    for (int64_t pi = p0; pi < p1; ++pi) {
        const int2 sp = wspan[pi];
        uint32_t sh = 0;
        for (int32_t j = __lane_id(); j < sp.y; j += 64) {
            const int32_t x = adj[(sp.x & 0x7FFFFFFF) + j];
            int32_t lo = 0, hi = d;
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (own[mid] < x) lo = mid + 1;
                else hi = mid;
            }
            sh += (lo < d && own[lo] == x) ? 1u : 0u;
        }
        hits += sh;
        if constexpr (kPerEdge) {
            sh = rtgd::wave_sum_u32(sh);
            if (__lane_id() == 0) eout[pi] = sh;
        }
    }
    return hits;
}

// Streams the out-lists of partner spans wspan[p0 .. p1) ({start, length} in adj) through the
// 64 lanes of one wave and probes every element against the owner table. Partners come 64 at
// a time. Lists of at least `long_min` elements are streamed one after another with lane j
// reading element j (coalesced buffer loads bounded by the list); shorter ones are
// concatenated: their lengths are scanned into the wave's LDS slots and each lane walks a
// cursor, so that lane j of a round handles element j of the concatenation.
// kPerEdge: every span's count is also stored at eout[pi], by span position (coalesced; a gather
// puts them in edge order afterwards) — a long span's by a DPP wave
// sum (the lanes' hits and pads summed once per span), a concatenated short span's through its
// LDS counter pe[ci].
// kSorted (the owner's sorted list in LDS) merges a long span with the owner's list: both
// ascend, so a round of 64*U consecutive elements can only match the owner-list window between
// the lower bound of its first element and the upper bound of its last (two searches the wave
// shares), each lane searches only that window, and the next round starts where it ended.
// One group of W loads of kW-byte elements per lane from buffer resource rs (lane offsets vo + u
// * 64 * kW), then their W probes: the loads all issue before the first probe waits.
template <uint32_t kMode, int W, typename T>
__device__ __forceinline__ uint32_t probe_group(const __amdgpu_buffer_rsrc_t &rs, int vo, const Table &c) {
    constexpr int kW = (int)sizeof(T);
    uint32_t x[W];
#pragma unroll
    for (int u = 0; u < W; ++u) {
        if constexpr (kW == 4) x[u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, vo + u * 64 * kW, 0, 0);
        else x[u] = (uint32_t)(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(rs, vo + u * 64 * kW, 0, 0);
    }
    uint32_t h = 0;
#pragma unroll
    for (int u = 0; u < W; ++u) {
        if constexpr (kMode == kBitmap16 && kW == 2) h += bitmap16_has16(c, x[u]);
        else h += probe<kMode>(c, x[u]);
    }
    return h;
}

// The last, partial round of a span: nu (1 .. W) groups of 64 elements are left; the group of
// the smallest power-of-two width >= nu runs. Returns its width in *w.
template <uint32_t kMode, int W, typename T>
__device__ __forceinline__ uint32_t tail_group(const __amdgpu_buffer_rsrc_t &rs, int vo, const Table &c, int32_t nu,
                                               int32_t *w) {
    if constexpr (W > 1) {
        if (nu <= W / 2) return tail_group<kMode, W / 2, T>(rs, vo, c, nu, w);
    }
    *w = W;
    return probe_group<kMode, W, T>(rs, vo, c);
}

// The rounds of one long span base[0 .. l0): U loads of kW-byte elements per lane through a raw
// buffer resource rebased on the round's first element (scalar work per round; a lane's offsets
// are loop-invariant plus an immediate) whose range ends with the span, so a lane past the end
// reads 0 without a memory access and the rounds need no per-element bounds test or branch.
// Full rounds of 64*U elements come first; the last, partial round runs a group of 1, 2, 4 ..
// U loads, the fewest that cover it (a span of 150 elements takes 256 probes per wave instead
// of 512 at U = 8). Returns the span's hits, including the probes of 0 of the loads past the
// end; *slots is the number of probes each lane made.
template <uint32_t kMode, int U, typename T>
__device__ __forceinline__ uint32_t long_rounds(const T *base, int32_t l0, const Table &c, int32_t *slots) {
    constexpr int kW = (int)sizeof(T);
    int vo = __lane_id() * kW;
    uint32_t h = 0;
    int32_t jb = 0;
    for (; l0 - jb >= 64 * U; jb += 64 * U) {
        // An empty asm that "redefines" vo each round: without it the U offsets vo + u * 64 * kW
        // are hoisted out of every loop into U registers kept live across the kernel (spilled
        // at its register budget); with it they fold into the loads' immediate offsets.
        asm volatile("" : "+v"(vo));
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(base + jb), (short)0,
                                                                            (l0 - jb) * kW, kBufferWord3);
        h += probe_group<kMode, U, T>(rs, vo, c);
    }
    int32_t w = 0;
    if (jb < l0) {
        asm volatile("" : "+v"(vo));
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(base + jb), (short)0,
                                                                            (l0 - jb) * kW, kBufferWord3);
        h += tail_group<kMode, U, T>(rs, vo, c, (l0 - jb + 63) >> 6, &w);
    }
    *slots = (jb >> 6) + w;
    return h;
}

// The owner's sorted list in LDS merged with one long span base[0 .. l0) (kSorted), in rounds
// of 64*U elements loaded as in long_rounds: both ascend, so a round can only match the
// owner-list window between the lower bound of its first element and the upper bound of its
// last (two searches the wave shares); each lane searches only that window (loads past the end
// are not counted), and the next round starts where it ended.
template <int U, typename T>
__device__ __forceinline__ uint32_t merge_rounds(const T *base, int32_t l0, const Table &c) {
    constexpr int kW = (int)sizeof(T);
    const int lane = __lane_id();
    int vo = lane * kW;
    uint32_t h = 0;
    uint32_t w0 = 0;  // owner-list position the span has reached
    for (int32_t jb = 0; jb < l0; jb += 64 * U) {
        asm volatile("" : "+v"(vo));
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(base + jb), (short)0,
                                                                            (l0 - jb) * kW, kBufferWord3);
        uint32_t x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (kW == 4)
                x[u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, vo + u * 64 * kW, 0, 0);
            else
                x[u] = (uint32_t)(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(rs, vo + u * 64 * kW, 0, 0);
        }
        const int32_t last = min(64 * U, l0 - jb) - 1;  // the round's last element
        uint32_t xl = x[0];
#pragma unroll
        for (int u = 1; u < U; ++u)
            if ((last >> 6) == u) xl = x[u];
        const uint32_t first_v = __builtin_amdgcn_readfirstlane(x[0]);
        const uint32_t last_v = __builtin_amdgcn_readlane(xl, last & 63);
        const uint32_t wa = rtgd::uniform_u32(lds_lower_bound(c.tab, w0, c.t - w0, first_v));
        const uint32_t wb = rtgd::uniform_u32(lds_lower_bound(c.tab, wa, c.t - wa, last_v + 1u));
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t q = lds_lower_bound(c.tab, wa, wb - wa, x[u]);
            h += (jb + u * 64 + lane < l0 && q < wb && c.tab[min(q, c.t - 1)] == x[u]) ? 1u : 0u;
        }
        w0 = wb;
    }
    return h;
}

// kP16: a span whose start has bit 31 set is a row whose entries all fit 16 bits; it is streamed
// from adj16 (the same positions, 2 bytes each), half the bytes of the 32-bit rows.
template <uint32_t kMode, int U, bool kPerEdge, bool kP16 = false>
__device__ __forceinline__ uint32_t stream_spans(const int32_t *__restrict__ adj, const int2 *__restrict__ wspan,
                                                 int64_t p0, int64_t p1, const Table &c, int32_t long_min,
                                                 int32_t *ex, int32_t *st, uint32_t *__restrict__ eout, uint32_t *pe,
                                                 const uint16_t *__restrict__ adj16 = nullptr) {
    const int lane = __lane_id();
    const uint64_t lt = rtgd::lanemask_lt();
    uint32_t hits = 0;
    const uint32_t has0 = probe<kMode>(c, 0u);  // what a load past a long span's end counts
    for (int64_t bs = p0; bs < p1; bs += 64) {
        const int64_t pi = bs + lane;
        int2 sp = make_int2(0, 0);
        if (pi < p1) sp = wspan[pi];
        const bool is_long = sp.y >= long_min;
        uint64_t lm = __ballot(is_long);
        while (lm) {
            const int p = __builtin_ctzll(lm);
            lm &= lm - 1;
            const int32_t s0r = __builtin_amdgcn_readlane(sp.x, p);
            const int32_t s0 = s0r & 0x7FFFFFFF;
            const int32_t l0 = __builtin_amdgcn_readlane(sp.y, p);
            uint32_t h;
            if constexpr (kMode == kSorted) {
                if (kP16 && s0r < 0) h = merge_rounds<U>(adj16 + s0, l0, c);
                else h = merge_rounds<U>(adj + s0, l0, c);
                if constexpr (kPerEdge) h = rtgd::wave_sum_dpp(h);
            } else if constexpr (kPerEdge) {
                // the span's hits over the wave, uniform; every lane made `slots` probes, the
                // 64 * slots - l0 past the end probed 0 and hit iff the owner's list holds 0
                int32_t slots = 0;
                if (kP16 && s0r < 0) h = long_rounds<kMode, U>(adj16 + s0, l0, c, &slots);
                else h = long_rounds<kMode, U>(adj + s0, l0, c, &slots);
                h = rtgd::wave_sum_dpp(h) - has0 * (uint32_t)(64 * slots - l0);
            } else {
                int32_t slots = 0;
                if (kP16 && s0r < 0) h = long_rounds<kMode, U>(adj16 + s0, l0, c, &slots);
                else h = long_rounds<kMode, U>(adj + s0, l0, c, &slots);
                // This lane's loads past the end each probed 0 and hit iff the owner's list holds
                // 0 (has0): take them back. The lane's elements: positions lane, lane + 64, ...
                const int32_t in = lane < l0 ? (l0 - lane + 63) >> 6 : 0;
                h -= has0 * (uint32_t)(slots - in);
            }
            if constexpr (kPerEdge) {
                if (lane == 0) {  // the span's count, once per wave
                    hits += h;
                    eout[bs + p] = h;
                }
            } else {
                hits += h;
            }
        }
        const bool is_short = !is_long && sp.y > 0;
        const uint64_t sm = __ballot(is_short);
        if (sm) {
            const int32_t sl = is_short ? sp.y : 0;
            const int32_t incl = rtgd::wave_inclusive_scan(sl);
            const int32_t total = __builtin_amdgcn_readlane(incl, 63);
            const int ns = __popcll(sm);
            if (is_short) {
                const int ci = __popcll(sm & lt);
                ex[ci] = incl - sl;
                st[ci] = sp.x & 0x7FFFFFFF;
                if constexpr (kPerEdge) pe[ci] = 0u;
            }
            if (lane == 0) ex[ns] = total;
            rtgd::wave_sync_lds();
            int k = 0;
            for (int32_t jb = 0; jb < total; jb += 64 * U) {
                uint32_t x[U];
                int ku[kPerEdge ? U : 1];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int32_t j = jb + u * 64 + lane;
                    x[u] = kEmpty;
                    if (j < total) {
                        while (ex[k + 1] <= j) ++k;
                        x[u] = (uint32_t)adj[st[k] + (j - ex[k])];
                    }
                    if constexpr (kPerEdge) ku[u] = k;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (x[u] != kEmpty) {
                        const uint32_t h = probe<kMode>(c, x[u]);
                        hits += h;
                        if constexpr (kPerEdge)
                            if (h) atomicAdd(&pe[ku[u]], 1u);
                    }
                }
            }
            rtgd::wave_sync_lds();
            if constexpr (kPerEdge) {
                if (is_short) eout[pi] = pe[__popcll(sm & lt)];
                rtgd::wave_sync_lds();  // pe is rewritten by the next group of partners
            }
        }
    }
    return hits;
}

// The bitmap and sorted-list strategies out of line: the hash path's code stays as compact as
// before (instruction cache), and these run only for the buckets the planner gave them.
template <uint32_t kMode, int U, bool kPerEdge>
__device__ __noinline__ uint32_t stream_spans_ool(const int32_t *__restrict__ adj, const int2 *__restrict__ wspan,
                                                  int64_t p0, int64_t p1, Table c, int32_t long_min, int32_t *ex,
                                                  int32_t *st, uint32_t *__restrict__ eout, uint32_t *pe) {
    return stream_spans<kMode, U, kPerEdge>(adj, wspan, p0, p1, c, long_min, ex, st, eout, pe);
}

// With 16-bit spans: linear probing (the hash's fallback) and the sorted list, out of line.
template <bool kP16, int U, bool kPerEdge>
__device__ __noinline__ uint32_t stream_spans_ool(const int32_t *__restrict__ adj, const int2 *__restrict__ wspan,
                                                  int64_t p0, int64_t p1, Table c, int32_t long_min, int32_t *ex,
                                                  int32_t *st, uint32_t *__restrict__ eout, uint32_t *pe,
                                                  const uint16_t *__restrict__ adj16, uint32_t mode) {
    if (mode == kLinearProbe)
        return stream_spans<kLinearProbe, U, kPerEdge, true>(adj, wspan, p0, p1, c, long_min, ex, st, eout, pe,
                                                              adj16);
    return stream_spans<kSorted, U, kPerEdge, true>(adj, wspan, p0, p1, c, long_min, ex, st, eout, pe, adj16);
}

// Partners [p0, p1) against the owner table c, in the probe mode the table was built in. kAll:
// the bitmap and sorted-list modes are compiled in (else only the hash modes exist). kP16:
// plans on a degree-ordered copy of the graph (16-bit spans), where the bitmap is the common
// block strategy: it is compiled inline beside the hash, the others stay out of line.
template <int U, bool kPerEdge, bool kAll, bool kP16 = false>
__device__ __forceinline__ uint32_t stream_table(const int32_t *__restrict__ adj, const int2 *__restrict__ wspan,
                                                 int64_t p0, int64_t p1, const Table &c, int32_t long_min, int32_t *ex,
                                                 int32_t *st, uint32_t *__restrict__ eout, uint32_t *pe,
                                                 const uint16_t *__restrict__ adj16 = nullptr) {
    const uint32_t mode = rtgd::uniform_u32(c.mode);
    if constexpr (kP16) {
        static_assert(kAll, "16-bit spans come with every strategy");
        if (mode == kBitmap16)
            return stream_spans<kBitmap16, U, kPerEdge, true>(adj, wspan, p0, p1, c, long_min, ex, st, eout, pe, adj16);
        if (mode == kCuckoo)
            return stream_spans<kCuckoo, U, kPerEdge, true>(adj, wspan, p0, p1, c, long_min, ex, st, eout, pe, adj16);
        return stream_spans_ool<kP16, U, kPerEdge>(adj, wspan, p0, p1, c, long_min, ex, st, eout, pe, adj16, mode);
    } else {
        if (mode == kCuckoo) return stream_spans<kCuckoo, U, kPerEdge>(adj, wspan, p0, p1, c, long_min, ex, st, eout, pe);
        if (!kAll || mode == kLinearProbe)
            return stream_spans<kLinearProbe, U, kPerEdge>(adj, wspan, p0, p1, c, long_min, ex, st, eout, pe);
        if constexpr (kAll) {
            if (mode == kBitmap)
                return stream_spans_ool<kBitmap, U, kPerEdge>(adj, wspan, p0, p1, c, long_min, ex, st, eout, pe);
            return stream_spans_ool<kSorted, U, kPerEdge>(adj, wspan, p0, p1, c, long_min, ex, st, eout, pe);
        }
        return 0u;
    }
}

struct BatchArgs {
    const int32_t *off;
    const int32_t *adj;
    const int32_t *wofs;
    const int2 *wspan;
    const int32_t *small_list;
    const TcItem *items;
    int64_t n_items;
    unsigned long long *count;  // [0] total, [1] queue head
    int32_t max_table;          // block table slots
    int32_t wave_table;         // wave-private table slots
    int32_t long_min;           // direct-stream threshold
    int32_t grab;               // partners a wave takes per grab in a block item (<= 64)
    uint32_t *eout;             // per-edge counts: each span's count by position (else null)
    const uint16_t *adj16;      // 16-bit copy of adj (degree-ordered plans, else null)
};

// The 16-bit instantiation (degree-ordered copies) is held to 7 waves per SIMD (72 VGPRs, a few
// dozen spilled outside the streaming loops): on R-MAT scale 21 it takes 10.07 ms, at 6 waves
// 10.39, at 8 10.52. The launch waits on memory (two thirds of wave time), so the seventh wave's
// loads in flight pay for its spills; at 8 the spills cost more.
// The per-edge instantiation is held to 8 waves per SIMD (at most 64 VGPRs, a few spilled to
// scratch outside the streaming loops): left alone it takes 74 and runs 6 waves, which costs
// more than the spills (R-MAT scale 21: 23.7 against 25.3 ms).
// kAll: every strategy compiled in. Plans whose items are all hash, search or hash wave items
// run the instantiation without the bitmap and sorted-list code (the out-of-line calls would
// hold more registers across the hash path and cost it occupancy); there, bitmap and sorted
// items would be searched and wave items hashed, which gives the same counts.
template <int kThreads, int U, bool kPerEdge, bool kAll, bool kP16 = false>
__global__ __launch_bounds__(kThreads, kP16 ? 7 : (kPerEdge || kAll) ? 8 : 1) void tc_batched(BatchArgs a) {
    constexpr int kWaves = kThreads / 64;
    extern __shared__ uint32_t table[];
    __shared__ int32_t ex_all[kWaves][65];
    __shared__ int32_t st_all[kWaves][64];
    __shared__ uint32_t pe_all[kPerEdge ? kWaves : 1][kPerEdge ? 64 : 1];
    __shared__ uint64_t item_slot;
    __shared__ int32_t next_batch;
    __shared__ uint32_t fail_flag;

    const int lane = __lane_id();
    const int wave = (int)rtgd::uniform_u32(threadIdx.x >> 6);
    int32_t *ex = ex_all[wave];
    int32_t *st = st_all[wave];
    uint32_t *pe = pe_all[kPerEdge ? wave : 0];
    uint64_t hits = 0;

    // Queue grabs are published through LDS and read back as wave-uniform (scalar) values so
    // that every loop below is uniform.
    if (threadIdx.x == 0) {
        item_slot = atomicAdd(&a.count[1], 1ull);
        next_batch = 0;
    }
    __syncthreads();
    uint64_t it_idx = rtgd::uniform_u64(item_slot);
    while (it_idx < (uint64_t)a.n_items) {
        const TcItem it = a.items[it_idx];
        uint32_t kind = rtgd::uniform_u32((uint32_t)it.kind);
        const int32_t x = it.a;
        if (kind == kItemBitmap || kind == kItemSorted) {
            // The planner only gives these strategies to owners whose bitmap or list fits the
            // table; anything else is searched in HBM rather than written past it.
            const int32_t xb = a.off[x], d = a.off[x + 1] - xb;
            const uint32_t need = kind == kItemSorted ? (uint32_t)d
                                  : kP16 ? bitmap16_words((uint32_t)a.adj[xb + d - 1])
                                         : bitmap_words((uint32_t)a.adj[xb], (uint32_t)a.adj[xb + d - 1]);
            if (!kAll || rtgd::uniform_u32(need) > (uint32_t)a.max_table) kind = kItemSearch;
        }
        if (kind == kItemHash || kind == kItemBitmap || kind == kItemSorted) {
            // Block item: owner it.a staged in the block's table, partners [it.b, it.c).
            const int32_t xb = a.off[x], d = a.off[x + 1] - xb;
            Table c;
            if (!kAll || kind == kItemHash) c = build_block_table<kThreads>(table, a.adj, xb, d, &fail_flag);
            else if (kind == kItemBitmap) c = kP16 ? build_block_bitmap16<kThreads>(table, a.adj, xb, d)
                                                   : build_block_bitmap<kThreads>(table, a.adj, xb, d);
            else c = build_block_sorted<kThreads>(table, a.adj, xb, d);
            while (true) {
                int32_t g = 0;
                if (lane == 0) g = atomicAdd(&next_batch, a.grab);
                const int32_t b = (int32_t)rtgd::uniform_u32((uint32_t)g);
                const int64_t p0 = (int64_t)it.b + b;
                if (p0 >= it.c) break;
                const int64_t p1 = min((int64_t)it.c, p0 + a.grab);
                hits += stream_table<U, kPerEdge, kAll, kP16>(a.adj, a.wspan, p0, p1, c, a.long_min, ex, st, 
                                                              a.eout, pe, a.adj16);
            }
        } else if (kind == kItemSearch) {
            // Search item: partners [it.b, it.c) are binary-searched in the owner's sorted
            // list in HBM (owners too long for an LDS table, or with too little reuse).
            const int32_t xb = a.off[x], d = a.off[x + 1] - xb;
            while (true) {
                int32_t g = 0;
                if (lane == 0) g = atomicAdd(&next_batch, a.grab);
                const int32_t b = (int32_t)rtgd::uniform_u32((uint32_t)g);
                const int64_t p0 = (int64_t)it.b + b;
                if (p0 >= it.c) break;
                const int64_t p1 = min((int64_t)it.c, p0 + a.grab);
                hits += stream_search<kPerEdge>(a.adj, a.wspan, p0, p1, a.adj + xb, d, a.eout);
            }
        } else {
            // Wave items: small owners small_list[it.b, it.c), one per wave at a time, each in a
            // wave-private table of the group's strategy it.a (kCuckoo, kBitmap or kSorted).
            uint32_t *tab = table + wave * a.wave_table;
            const uint32_t mode = kAll ? rtgd::uniform_u32((uint32_t)it.a) : kCuckoo;
            for (int32_t s = it.b + wave; s < it.c; s += kWaves) {
                const int32_t xs = a.small_list[s];
                const int32_t xb = a.off[xs], d = a.off[xs + 1] - xb;
                const Table c = mode == kCuckoo ? build_wave_table(tab, a.adj, xb, d)
                                                : build_wave_other<kP16>(tab, a.adj, xb, d, mode, (uint32_t)a.wave_table);
                hits += stream_table<U, kPerEdge, kAll, kP16>(a.adj, a.wspan, a.wofs[xs], a.wofs[xs + 1], c, a.long_min,
                                                              ex, st, a.eout, pe, a.adj16);
                rtgd::wave_sync_lds();
            }
        }
        __syncthreads();  // every wave is done with the table, item_slot and next_batch
        if (threadIdx.x == 0) {
            item_slot = atomicAdd(&a.count[1], 1ull);
            next_batch = 0;
        }
        __syncthreads();
        it_idx = rtgd::uniform_u64(item_slot);
    }
    hits = rtgd::wave_sum_u64(hits);
    if (lane == 0 && hits) atomicAdd(&a.count[0], (unsigned long long)hits);
}

// Block bitmap items of a degree-ordered copy only (kItemBitmap, kBitmap16 tables), from the
// queue a.items[0 .. n_items) with its own head a.count[2]: the bitmap path without the
// registers the other strategies hold across it.
template <int kThreads, int U, bool kPerEdge>
__global__ __launch_bounds__(kThreads, 8) void tc_bitmap_items(BatchArgs a) {
    constexpr int kWaves = kThreads / 64;
    extern __shared__ uint32_t table[];
    __shared__ int32_t ex_all[kWaves][65];
    __shared__ int32_t st_all[kWaves][64];
    __shared__ uint32_t pe_all[kPerEdge ? kWaves : 1][kPerEdge ? 64 : 1];
    __shared__ uint64_t item_slot;
    __shared__ int32_t next_batch;
    const int lane = __lane_id();
    const int wave = (int)rtgd::uniform_u32(threadIdx.x >> 6);
    int32_t *ex = ex_all[wave];
    int32_t *st = st_all[wave];
    uint32_t *pe = pe_all[kPerEdge ? wave : 0];
    uint64_t hits = 0;
    if (threadIdx.x == 0) {
        item_slot = atomicAdd(&a.count[2], 1ull);
        next_batch = 0;
    }
    __syncthreads();
    uint64_t it_idx = rtgd::uniform_u64(item_slot);
    while (it_idx < (uint64_t)a.n_items) {
        const TcItem it = a.items[it_idx];
        const int32_t x = it.a;
        const int32_t xb = a.off[x], d = a.off[x + 1] - xb;
        const bool fits = rtgd::uniform_u32(bitmap16_words((uint32_t)a.adj[xb + d - 1])) <= (uint32_t)a.max_table;
        Table c{};
        if (fits) c = build_block_bitmap16<kThreads>(table, a.adj, xb, d);
        while (true) {
            int32_t g = 0;
            if (lane == 0) g = atomicAdd(&next_batch, a.grab);
            const int32_t b = (int32_t)rtgd::uniform_u32((uint32_t)g);
            const int64_t p0 = (int64_t)it.b + b;
            if (p0 >= it.c) break;
            const int64_t p1 = min((int64_t)it.c, p0 + a.grab);
            if (fits)
                hits += stream_spans<kBitmap16, U, kPerEdge, true>(a.adj, a.wspan, p0, p1, c, a.long_min, ex, st, 
                                                                   a.eout, pe, a.adj16);
            else
                hits += stream_search<kPerEdge>(a.adj, a.wspan, p0, p1, a.adj + xb, d, a.eout);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            item_slot = atomicAdd(&a.count[2], 1ull);
            next_batch = 0;
        }
        __syncthreads();
        it_idx = rtgd::uniform_u64(item_slot);
    }
    hits = rtgd::wave_sum_u64(hits);
    if (lane == 0 && hits) atomicAdd(&a.count[0], (unsigned long long)hits);
}

using BatchKernel = void (*)(BatchArgs);

// Instantiated (block size, unroll) pairs of the batched kernel, counting or per-edge, with the
// hash strategies only or with every strategy.
template <bool kPerEdge, bool kAll>
BatchKernel batched_kernel_t(int threads, int unroll) {
    if (threads == 256 && unroll == 2) return tc_batched<256, 2, kPerEdge, kAll>;
    if (threads == 256 && unroll == 4) return tc_batched<256, 4, kPerEdge, kAll>;
    if (threads == 256 && unroll == 8) return tc_batched<256, 8, kPerEdge, kAll>;
    if (threads == 512 && unroll == 2) return tc_batched<512, 2, kPerEdge, kAll>;
    if (threads == 512 && unroll == 4) return tc_batched<512, 4, kPerEdge, kAll>;
    if (threads == 512 && unroll == 8) return tc_batched<512, 8, kPerEdge, kAll>;
    return nullptr;
}

// On a degree-ordered copy (16-bit spans, every strategy), counting or per-edge.
template <bool kPerEdge>
BatchKernel batched_kernel_p16(int threads, int unroll) {
    if (threads == 256 && unroll == 2) return tc_batched<256, 2, kPerEdge, true, true>;
    if (threads == 256 && unroll == 4) return tc_batched<256, 4, kPerEdge, true, true>;
    if (threads == 256 && unroll == 8) return tc_batched<256, 8, kPerEdge, true, true>;
    if (threads == 512 && unroll == 2) return tc_batched<512, 2, kPerEdge, true, true>;
    if (threads == 512 && unroll == 4) return tc_batched<512, 4, kPerEdge, true, true>;
    if (threads == 512 && unroll == 8) return tc_batched<512, 8, kPerEdge, true, true>;
    return nullptr;
}

// Wave items only (small owners in wave-private cuckoo tables, linear probing as their
// fallback) of a degree-ordered plan whose small owners are hashed, from the queue a.items[0 ..
// n_items) with its own head a.count[3].
template <int kThreads, int U, bool kPerEdge>
__global__ __launch_bounds__(kThreads, 8) void tc_wave_items(BatchArgs a) {
    constexpr int kWaves = kThreads / 64;
    extern __shared__ uint32_t table[];
    __shared__ int32_t ex_all[kWaves][65];
    __shared__ int32_t st_all[kWaves][64];
    __shared__ uint32_t pe_all[kPerEdge ? kWaves : 1][kPerEdge ? 64 : 1];
    __shared__ uint64_t item_slot;
    const int wave = (int)rtgd::uniform_u32(threadIdx.x >> 6);
    int32_t *ex = ex_all[wave];
    int32_t *st = st_all[wave];
    uint32_t *pe = pe_all[kPerEdge ? wave : 0];
    uint32_t *tab = table + wave * a.wave_table;
    uint64_t hits = 0;
    if (threadIdx.x == 0) item_slot = atomicAdd(&a.count[3], 1ull);
    __syncthreads();
    uint64_t it_idx = rtgd::uniform_u64(item_slot);
    while (it_idx < (uint64_t)a.n_items) {
        const TcItem it = a.items[it_idx];
        for (int32_t s = it.b + wave; s < it.c; s += kWaves) {
            const int32_t xs = a.small_list[s];
            const int32_t xb = a.off[xs], d = a.off[xs + 1] - xb;
            const Table c = build_wave_table(tab, a.adj, xb, d);
            const int64_t p0 = a.wofs[xs], p1 = a.wofs[xs + 1];
            if (rtgd::uniform_u32(c.mode) == kCuckoo)
                hits += stream_spans<kCuckoo, U, kPerEdge, true>(a.adj, a.wspan, p0, p1, c, a.long_min, ex, st, 
                                                                 a.eout, pe, a.adj16);
            else
                hits += stream_spans<kLinearProbe, U, kPerEdge, true>(a.adj, a.wspan, p0, p1, c, a.long_min, ex, st,
                                                                       a.eout, pe, a.adj16);
            rtgd::wave_sync_lds();
        }
        __syncthreads();  // every wave is done with item_slot
        if (threadIdx.x == 0) item_slot = atomicAdd(&a.count[3], 1ull);
        __syncthreads();
        it_idx = rtgd::uniform_u64(item_slot);
    }
    hits = rtgd::wave_sum_u64(hits);
    if (__lane_id() == 0 && hits) atomicAdd(&a.count[0], (unsigned long long)hits);
}

// The block-bitmap-only kernel of a degree-ordered plan, counting or per-edge.
template <bool kPerEdge>
BatchKernel bitmap_kernel_t(int threads, int unroll) {
    if (threads == 256 && unroll == 2) return tc_bitmap_items<256, 2, kPerEdge>;
    if (threads == 256 && unroll == 4) return tc_bitmap_items<256, 4, kPerEdge>;
    if (threads == 256 && unroll == 8) return tc_bitmap_items<256, 8, kPerEdge>;
    if (threads == 512 && unroll == 2) return tc_bitmap_items<512, 2, kPerEdge>;
    if (threads == 512 && unroll == 4) return tc_bitmap_items<512, 4, kPerEdge>;
    if (threads == 512 && unroll == 8) return tc_bitmap_items<512, 8, kPerEdge>;
    return nullptr;
}

BatchKernel bitmap_kernel(int threads, int unroll, bool per_edge) {
    return per_edge ? bitmap_kernel_t<true>(threads, unroll) : bitmap_kernel_t<false>(threads, unroll);
}

// The wave-items-only kernel of a degree-ordered plan whose small owners are hashed, counting or
// per-edge.
template <bool kPerEdge>
BatchKernel wave_kernel_t(int threads, int unroll) {
    if (threads == 256 && unroll == 2) return tc_wave_items<256, 2, kPerEdge>;
    if (threads == 256 && unroll == 4) return tc_wave_items<256, 4, kPerEdge>;
    if (threads == 256 && unroll == 8) return tc_wave_items<256, 8, kPerEdge>;
    if (threads == 512 && unroll == 2) return tc_wave_items<512, 2, kPerEdge>;
    if (threads == 512 && unroll == 4) return tc_wave_items<512, 4, kPerEdge>;
    if (threads == 512 && unroll == 8) return tc_wave_items<512, 8, kPerEdge>;
    return nullptr;
}

BatchKernel wave_kernel(int threads, int unroll, bool per_edge) {
    return per_edge ? wave_kernel_t<true>(threads, unroll) : wave_kernel_t<false>(threads, unroll);
}

BatchKernel batched_kernel(int threads, int unroll, bool per_edge, bool all_strategies, bool p16 = false) {
    if (p16) return per_edge ? batched_kernel_p16<true>(threads, unroll) : batched_kernel_p16<false>(threads, unroll);
    if (per_edge) return all_strategies ? batched_kernel_t<true, true>(threads, unroll) : batched_kernel_t<true, false>(threads, unroll);
    return all_strategies ? batched_kernel_t<false, true>(threads, unroll) : batched_kernel_t<false, false>(threads, unroll);
}

void launch_batched(BatchKernel k, int grid, int threads, size_t lds, hipStream_t st, const BatchArgs &a) {
    hipLaunchKernelGGL(k, dim3(grid), dim3(threads), lds, st, a);
}

// ---- planning kernels ----

// Per edge of the range [e_begin, e_end): the byte models, and a grouping key — the owner
// (the endpoint with the longer out-list) in the low `bits`, above them the partner (the other
// endpoint) or, for per-edge plans (with_edge), the edge's index in the range. `bits` is a whole number of bytes: the radix sort orders by 8-bit digits, so the
// owner field must fill its last digit (edge bits inside it would split an owner's edges).
// Edges with an empty list get owner n, so that a radix sort on the low bits (stable) lists
// every owner's edges together, in edge order, and those edges last.
constexpr int kPlanU = 4;  // independent items per thread and round in the plan kernels

__global__ __launch_bounds__(256) void plan_keys(const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                 const int32_t *__restrict__ row, int64_t e_begin, int64_t e_end,
                                                 int bits, int32_t n, bool with_edge, uint64_t *__restrict__ keys,
                                                 unsigned long long *bytes /* [0] alg, [1] merge, [2] edges */) {
    uint64_t alg = 0, mrg = 0, ne = 0;
    // kPlanU edges per thread and round, loads in phases (endpoints, then their row offsets), so
    // that a round keeps kPlanU independent chains in flight; a slot past the end reads edge
    // e_begin and is dropped.
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e0 = e_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e0 < e_end; e0 += stride * kPlanU) {
        int32_t u[kPlanU], v[kPlanU];
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) {
            const int64_t e = e0 + k * stride;
            const int64_t ec = e < e_end ? e : e_begin;
            u[k] = row[ec];
            v[k] = adj[ec];
        }
        int32_t u0[kPlanU], u1[kPlanU], v0[kPlanU], v1[kPlanU];
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) {
            u0[k] = off[u[k]];
            u1[k] = off[u[k] + 1];
            v0[k] = off[v[k]];
            v1[k] = off[v[k] + 1];
        }
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) {
            const int64_t e = e0 + k * stride;
            if (e >= e_end) break;
            const uint32_t du = (uint32_t)(u1[k] - u0[k]), dv = (uint32_t)(v1[k] - v0[k]);
            const uint32_t s = min(du, dv), l = max(du, dv);
            const uint32_t lg = l ? 32u - __clz(l) : 0u;  // ceil(log2(l+1))
            alg += 24ull + 4ull * min((uint64_t)s + l, (uint64_t)s * (1u + lg));
            mrg += 24ull + 4ull * ((uint64_t)du + dv);
            int32_t owner = n;
            if (s > 0) {
                owner = (du > dv || (du == dv && u[k] < v[k])) ? u[k] : v[k];
                ne += 1;
            }
            const uint64_t payload = with_edge ? (uint64_t)(e - e_begin) : (uint64_t)(uint32_t)(owner == u[k] ? v[k] : u[k]);
            keys[e - e_begin] = (payload << bits) | (uint64_t)(uint32_t)owner;
        }
    }
    // one atomic per block and counter
    __shared__ uint64_t part[4][3];
    alg = rtgd::wave_sum_u64(alg);
    mrg = rtgd::wave_sum_u64(mrg);
    ne = rtgd::wave_sum_u64(ne);
    if (__lane_id() == 0) {
        part[threadIdx.x >> 6][0] = alg;
        part[threadIdx.x >> 6][1] = mrg;
        part[threadIdx.x >> 6][2] = ne;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint64_t t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w][threadIdx.x];
        if (t) atomicAdd(&bytes[threadIdx.x], (unsigned long long)t);
    }
}

// Position i of the owner-sorted keys (the first nw are work edges): the partner's out-list
// span {start, length} (and the edge, for per-edge plans), its length for the work sums, and
// each owner's first position. Keys of counting plans carry the partner itself, so only the
// partner's row offsets are gathered; per-edge plans' keys carry the edge, whose endpoints are
// read back.
// vinfo, s16 and s256 (degree-ordered copies, else null: each row's {start, end, largest entry,
// 0} and the samples of every 16th and 256th entry): a span is cut at a multiple of 16 past which
// every element exceeds the owner's largest entry, a span whose elements fit 16 bits gets bit 31
// in its start, and their elements are summed into *streamed16.
__global__ void plan_spans(const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
                           const int32_t *__restrict__ row, int64_t e_begin, const uint64_t *__restrict__ sorted,
                           int64_t nw, int bits, int2 *__restrict__ wspan, uint32_t *__restrict__ wedge,
                           uint64_t *__restrict__ lens, uint32_t *__restrict__ first, const int4 *__restrict__ vinfo,
                           const int32_t *__restrict__ s16, const int32_t *__restrict__ s256,
                           unsigned long long *__restrict__ streamed16, const int32_t *__restrict__ emap) {
    uint64_t s16sum = 0;
    const uint64_t mask = (1ull << bits) - 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // kPlanU positions per thread and round, loads in phases (keys, then the edges' endpoints
    // for per-edge plans, then the partners' row offsets); a slot past the end reads position 0.
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < nw; i0 += stride * kPlanU) {
        uint64_t key[kPlanU], prev[kPlanU];
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) {
            const int64_t i = i0 + k * stride;
            const int64_t ic = i < nw ? i : 0;
            key[k] = sorted[ic];
            prev[k] = ic > 0 ? sorted[ic - 1] : ~0ull;
        }
        int32_t owner[kPlanU], partner[kPlanU];
        uint32_t firsts = 0;  // bit k: position k starts its owner's run
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) {
            owner[k] = (int32_t)(key[k] & mask);
            firsts |= (i0 + k * stride == 0 || (int32_t)(prev[k] & mask) != owner[k]) ? 1u << k : 0u;
        }
        uint32_t wed[kPlanU];  // per-edge plans: the graph's edge of each position
        if (wedge) {
            int64_t edge[kPlanU];
            int32_t eu[kPlanU], ev[kPlanU];
#pragma unroll
            for (int k = 0; k < kPlanU; ++k) {
                edge[k] = e_begin + (int64_t)(key[k] >> bits);
                eu[k] = row[edge[k]];
                ev[k] = adj[edge[k]];
                wed[k] = emap ? (uint32_t)emap[edge[k]] : (uint32_t)edge[k];
            }
#pragma unroll
            for (int k = 0; k < kPlanU; ++k) partner[k] = owner[k] == eu[k] ? ev[k] : eu[k];
        } else {
#pragma unroll
            for (int k = 0; k < kPlanU; ++k) {
                partner[k] = (int32_t)(key[k] >> bits);
                wed[k] = 0;
            }
        }
        int32_t p0[kPlanU], p1[kPlanU], pmax[kPlanU], hi[kPlanU];
        if (vinfo) {  // one 16-byte gather per partner: its row and largest entry
#pragma unroll
            for (int k = 0; k < kPlanU; ++k) {
                const int4 pi = vinfo[partner[k]];
                p0[k] = pi.x;
                p1[k] = pi.y;
                pmax[k] = pi.z;
                hi[k] = vinfo[owner[k]].z;
            }
        } else {
#pragma unroll
            for (int k = 0; k < kPlanU; ++k) {
                p0[k] = off[partner[k]];
                p1[k] = off[partner[k] + 1];
                pmax[k] = INT32_MAX;
                hi[k] = INT32_MAX;
            }
        }
        uint32_t f16[kPlanU];
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) f16[k] = pmax[k] < 65536 ? 1u : 0u;
        if (vinfo) {
            // Degree-ordered copies: an element above the owner's largest entry hi matches
            // nothing, and rows are sorted, so a span whose last element exceeds hi (and that
            // holds a sample) is cut before its first sample of s16 above hi. The owner's
            // smallest entry is near 0 in this order, so the head is kept. Level 1 finds by
            // binary lifting (all kPlanU searches together) the first s256 sample of the row
            // above hi; the 16 s16 samples of the chunk before it (one aligned 64-byte line,
            // four independent loads) give the first s16 sample above hi. A span becomes 16-bit
            // when that sample is at most 2^16: every kept element lies below it.
            bool need[kPlanU];
            int32_t jf[kPlanU], jl[kPlanU], nmax = 0;
#pragma unroll
            for (int k = 0; k < kPlanU; ++k) {
                need[k] = p1[k] - p0[k] > 16 && pmax[k] > hi[k];
                jf[k] = (p0[k] + 255) >> 8;  // s256 samples inside the row: jf .. jl
                jl[k] = (p1[k] - 1) >> 8;
                nmax = max(nmax, need[k] ? jl[k] - jf[k] + 1 : 0);
            }
            for (int32_t step = nmax > 0 ? (int32_t)(1u << (31 - __clz(nmax))) : 0; step; step >>= 1) {
                int32_t v[kPlanU];
#pragma unroll
                for (int k = 0; k < kPlanU; ++k) v[k] = (need[k] && jf[k] + step <= jl[k] + 1) ? s256[jf[k] + step - 1] : INT32_MAX;
#pragma unroll
                for (int k = 0; k < kPlanU; ++k)
                    if (v[k] <= hi[k]) jf[k] += step;
            }
            // jf: the row's first s256 sample above hi (jl + 1 if none); the chunk c before it
            // (the last one when none) holds the first s16 sample above hi, unless it is sample
            // 16 jf itself. Two positions' lines are in flight at a time (all four would double
            // the registers).
#pragma unroll
            for (int h = 0; h < kPlanU; h += 2) {
                int4 q[2][4];
                int32_t c[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int k = h + u;
                    c[u] = jf[k] <= jl[k] ? jf[k] - 1 : jl[k];
                    const int4 *line = reinterpret_cast<const int4 *>(s16 + 16 * (int64_t)max(c[u], 0));
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        q[u][t] = (need[k] && c[u] >= 0) ? line[t] : make_int4(INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX);
                }
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int k = h + u;
                    if (!need[k]) continue;
                    const int32_t i0s = (p0[k] + 15) >> 4, i1s = (p1[k] - 1) >> 4;  // s16 samples inside the row
                    const int32_t base = 16 * c[u];
                    const int32_t lo = max(i0s, base);
                    int32_t cnt = 0, nextv = INT32_MAX;
#pragma unroll
                    for (int t = 0; t < 16; ++t) {
                        const int32_t x = (&q[u][t >> 2].x)[t & 3];
                        const int32_t idx = base + t;
                        const bool in = idx >= lo && idx <= i1s;
                        cnt += (in && x <= hi[k]) ? 1 : 0;
                        nextv = (in && x > hi[k] && nextv == INT32_MAX) ? x : nextv;
                    }
                    const int32_t is = lo + cnt;  // first s16 sample of the row above hi
                    if (is > i1s) continue;       // none: the span is kept whole
                    if (is == base + 16) nextv = s256[jf[k]];
                    p1[k] = 16 * is;
                    f16[k] |= nextv <= 65536 ? 1u : 0u;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) {
            const int64_t i = i0 + k * stride;
            if (i >= nw) break;
            const int32_t len = p1[k] - p0[k];
            if (wedge) wedge[i] = wed[k];
            wspan[i] = make_int2((int32_t)((uint32_t)p0[k] | (f16[k] << 31)), len);
            lens[i] = (uint64_t)len;
            s16sum += f16[k] ? (uint64_t)len : 0ull;
            if ((firsts >> k) & 1u) first[owner[k]] = (uint32_t)i;
        }
    }
    if (vinfo) {  // one device atomic per wave
        s16sum = rtgd::wave_sum_u64(s16sum);
        if (__lane_id() == 0 && s16sum) atomicAdd(streamed16, (unsigned long long)s16sum);
    }
}

// Per-edge plans: epos[e - e_begin] = i for span i of positive length and its oriented edge e =
// wedge[i] (the fill leaves kNoSpan for edges without one, whose count is 0): where each edge's
// count lies among the spans' counts, which the counting kernels store by span position.
constexpr uint32_t kNoSpan = 0xFFFFFFFFu;

__global__ void edge_positions(const uint32_t *__restrict__ wedge, const int2 *__restrict__ wspan, int64_t nw,
                               int64_t e_begin, uint32_t *__restrict__ epos) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x)
        if (wspan[i].y > 0) epos[(int64_t)wedge[i] - e_begin] = (uint32_t)i;
}

// out[j] = the count of the range's edge j: ecnt at its span position, 0 without one. The
// per-edge output is written in edge order with whole-line stores; storing each span's count
// straight at its edge (random 4-byte stores, a partial line each) cost R-MAT scale 21 2.4 ms
// of a 12.3 ms per-edge run. kPlanU edges per thread and round keep the random reads in flight.
__global__ __launch_bounds__(256) void gather_edge_counts(const uint32_t *__restrict__ epos, int64_t ne,
                                                          const uint32_t *__restrict__ ecnt,
                                                          uint32_t *__restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j0 < ne; j0 += stride * kPlanU) {
        uint32_t q[kPlanU], v[kPlanU];
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) {
            const int64_t j = j0 + k * stride;
            q[k] = j < ne ? epos[j] : kNoSpan;
        }
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) v[k] = q[k] != kNoSpan ? ecnt[q[k]] : 0u;
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) {
            const int64_t j = j0 + k * stride;
            if (j < ne) out[j] = v[k];
        }
    }
}

// At each owner's last position: its partner count and its streamed elements (from the
// exclusive prefix sums of the span lengths).
__global__ void plan_runs(const uint64_t *__restrict__ sorted, int64_t nw, int bits, const uint32_t *__restrict__ first,
                          const uint64_t *__restrict__ lens, const uint64_t *__restrict__ lens_ex, uint32_t *__restrict__ cnt,
                          unsigned long long *__restrict__ work) {
    const uint64_t mask = (1ull << bits) - 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < nw; i0 += stride * kPlanU) {
        uint64_t key[kPlanU], next[kPlanU];  // kPlanU positions' loads in flight together
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) {
            const int64_t i = i0 + k * stride;
            const int64_t ic = i < nw ? i : 0;
            key[k] = sorted[ic];
            next[k] = ic + 1 < nw ? sorted[ic + 1] : ~0ull;
        }
#pragma unroll
        for (int k = 0; k < kPlanU; ++k) {
            const int64_t i = i0 + k * stride;
            const int32_t owner = (int32_t)(key[k] & mask);
            if (i >= nw || (i + 1 < nw && (int32_t)(next[k] & mask) == owner)) continue;
            const uint32_t f = first[owner];
            cnt[owner] = (uint32_t)(i + 1 - f);
            work[owner] = (unsigned long long)(lens_ex[i] + lens[i] - lens_ex[f]);
        }
    }
}

unsigned grid_cap(rtg_ctx *ctx, int64_t n, int threads, int per_cu) {
    int64_t want = rtg::ceil_div(std::max<int64_t>(n, 1), threads);
    int64_t cap = (int64_t)ctx->cu_count * per_cu;
    return (unsigned)std::max<int64_t>(1, std::min(want, cap));
}

struct DevFree {
    void *p = nullptr;
    ~DevFree() {
        if (p) (void)hipFree(p);
    }
};

template <class T>
int dev_alloc(T **p, size_t count) {
    if (hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return rtg::set_error(RTG_ENOMEM, "device allocation of " + std::to_string(sizeof(T) * count) + " bytes failed");
    }
    return RTG_OK;
}

// Tunables of the batched plan, written as a comma list of key=value (keys threads, unroll,
// item, wave_item, wave_max, long, grab, table_cap, strat); chosen by the autotuner, read from
// the tuning cache or set by the RTG_TC_TUNE environment variable. `strat` holds one letter per
// owner bucket (h hash, b bitmap, m merge with the sorted list, s search in HBM).
constexpr int kBuckets = 4;
constexpr int kStratHash = 0, kStratSearch = 1, kStratBitmap = 2, kStratMerge = 3;
constexpr char kStratLetter[] = {'h', 's', 'b', 'm'};
const char *const kStratName[] = {"hash", "search", "bitmap", "merge"};

struct BatchTune {
    int32_t block_threads = 256;
    int32_t unroll = 4;
    int64_t item_work = 1 << 21;   // streamed elements per block item
    int64_t wave_item_work = 1 << 12;
    int32_t wave_owner_max = -1;   // -1: as large as the wave's table share allows
    int32_t long_min = 128;
    int32_t grab = 64;
    int32_t table_cap = 0;         // LDS table slots per block (0: what a workgroup can hold)
    int32_t strat[kBuckets] = {kStratHash, kStratHash, kStratHash, kStratSearch};
};

std::string strat_string(const int32_t *strat) {
    std::string s;
    for (int b = 0; b < kBuckets; ++b) s.push_back(kStratLetter[strat[b]]);
    return s;
}

std::string tune_string(const BatchTune &t) {
    return "threads=" + std::to_string(t.block_threads) + ",unroll=" + std::to_string(t.unroll) +
           ",item=" + std::to_string(t.item_work) + ",wave_item=" + std::to_string(t.wave_item_work) +
           ",wave_max=" + std::to_string(t.wave_owner_max) + ",long=" + std::to_string(t.long_min) +
           ",grab=" + std::to_string(t.grab) + (t.table_cap > 0 ? ",table_cap=" + std::to_string(t.table_cap) : "") +
           ",strat=" + strat_string(t.strat);
}

void tune_from_string(BatchTune &t, const std::string &s) {
    size_t pos = 0;
    while (pos < s.size()) {
        size_t e = s.find(',', pos);
        if (e == std::string::npos) e = s.size();
        const std::string kv = s.substr(pos, e - pos);
        const size_t eq = kv.find('=');
        if (eq != std::string::npos) {
            const std::string k = kv.substr(0, eq);
            const std::string sv = kv.substr(eq + 1);
            const long long v = std::atoll(sv.c_str());
            if (k == "threads") t.block_threads = (int32_t)v;
            else if (k == "unroll") t.unroll = (int32_t)v;
            else if (k == "item") t.item_work = std::max<long long>(1, v);
            else if (k == "wave_item") t.wave_item_work = std::max<long long>(1, v);
            else if (k == "wave_max") t.wave_owner_max = (int32_t)v;
            else if (k == "long") t.long_min = (int32_t)std::max<long long>(1, v);
            else if (k == "grab") t.grab = (int32_t)std::max<long long>(1, std::min<long long>(64, v));
            else if (k == "table_cap") t.table_cap = (int32_t)std::max<long long>(0, std::min<long long>(kMaxTableSlots, v));
            else if (k == "strat")
                for (int b = 0; b < kBuckets && b < (int)sv.size(); ++b)
                    for (int c = 0; c < 4; ++c)
                        if (sv[(size_t)b] == kStratLetter[c]) t.strat[b] = c;
        }
        pos = e + 1;
    }
}

// ---- device planning: partner offsets, validation, fingerprint ----

// flag[0] = 1 unless wofs[0..n] is a partition of the nw partner spans (starts at 0, ends at nw,
// never decreases): the counting kernel reads spans through these offsets.
__global__ void plan_check_offsets(const int32_t *__restrict__ wofs, int32_t n, int64_t nw, uint32_t *flag) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        const bool bad = wofs[x] > wofs[x + 1] || (x == 0 && wofs[0] != 0) || (x == n - 1 && (int64_t)wofs[n] != nw);
        if (bad) atomicOr(flag, 1u);
    }
}

// A fingerprint of the row offsets (index-mixed, so order matters) for the tuning cache key.
__global__ void offsets_fingerprint(const int32_t *__restrict__ off, int32_t n, unsigned long long *h) {
    uint64_t acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = ((uint64_t)(uint32_t)off[i] << 32 | (uint64_t)i) + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        acc += z ^ (z >> 31);
    }
    acc = rtgd::wave_sum_u64(acc);
    if (__lane_id() == 0) atomicAdd(h, (unsigned long long)acc);
}

// Graph-level part of the batched plan, on the device: each work edge goes to its owner (the
// endpoint with the longer out-list); the edges are grouped by owner with a stable radix sort on
// the owner bits (each owner's partners stay in edge order, so plans are deterministic); then
// partner spans, run boundaries, per-owner partner counts and streamed work, the partner
// offsets (scanned counts), their validation and the offsets' fingerprint. Only the sums come
// back to the host.
int plan_graph(rtg_ctx *ctx, rtg_tc_plan *p) {
    using namespace rtg;
    const rtg_graph *g = p->g;
    const int32_t n = g->n;
    hipStream_t st = ctx->stream;
    const int64_t ne = p->e_end - p->e_begin;
    // owner ids 0 .. n, in whole bytes (the radix sort's digits)
    const int bits = (std::max(1, 32 - __builtin_clz((uint32_t)n + 1u)) + 7) / 8 * 8;
    if (ne >= ((int64_t)1 << (63 - bits)) || bits > 32)
        return set_error(RTG_EINVAL, "batched TC: edge range too large for the plan keys");
    const size_t kn = (size_t)std::max<int64_t>(ne, 1);
    const size_t bytes = Arena::round(sizeof(int32_t) * (size_t)std::max<int64_t>(g->m, 1)) * (p->opts.per_edge ? 2 : 1) +
                         Arena::round(sizeof(uint32_t) * ((size_t)n + 1)) * 2 + Arena::round(64) +
                         Arena::round(sizeof(uint64_t) * kn) * 4 + Arena::round(radix_sort_work_bytes((int64_t)kn)) +
                         Arena::round(std::max(scan_work_bytes((int64_t)kn), scan_work_bytes(n + 1))) + 4096;
    // The planning temporaries (each edge's row, the keys, spans' lengths, sort and scan work)
    // live in the context's grow-only scratch: a plan of the same graph allocates nothing for
    // them (a fresh hipMalloc and hipFree of 2.7 GB cost about 0.8 ms on R-MAT scale 21).
    void *sbuf = nullptr;
    RTG_TRY(ctx_scratch(ctx, bytes, &sbuf));
    Arena ar{static_cast<char *>(sbuf), 0, bytes};
    int32_t *row = ar.take<int32_t>((size_t)std::max<int64_t>(g->m, 1));
    // per-edge plans: each span's oriented edge (at most m spans)
    uint32_t *wedge = p->opts.per_edge ? ar.take<uint32_t>((size_t)std::max<int64_t>(g->m, 1)) : nullptr;
    uint32_t *d_cnt = ar.take<uint32_t>((size_t)n + 1);
    uint32_t *d_first = ar.take<uint32_t>((size_t)n + 1);
    unsigned long long *d_bytes = ar.take<unsigned long long>(8);  // alg, merge, edges, flag, fingerprint
    uint64_t *keys[2] = {ar.take<uint64_t>(kn), ar.take<uint64_t>(kn)};
    uint64_t *lens = ar.take<uint64_t>(kn);
    uint64_t *lens_ex = ar.take<uint64_t>(kn);
    void *sort_work = ar.take<char>(radix_sort_work_bytes((int64_t)kn));
    void *scan_work = ar.take<char>(std::max(scan_work_bytes((int64_t)kn), scan_work_bytes(n + 1)));
    RTG_TRY(dev_alloc(&p->d_work, (size_t)n + 1));
    HIP_TRY(hipMemsetAsync(d_cnt, 0, sizeof(uint32_t) * (size_t)(n + 1), st));
    HIP_TRY(hipMemsetAsync(p->d_work, 0, sizeof(unsigned long long) * (size_t)(n + 1), st));
    HIP_TRY(hipMemsetAsync(d_bytes, 0, sizeof(unsigned long long) * 8, st));
    int which = 0;
    if (n > 0) {
        edge_rows_kernel<<<grid_cap(ctx, n, 256, 16), 256, 0, st>>>(p->v_off, n, row);
        HIP_TRY(hipGetLastError());
    }
    if (ne > 0) {
        plan_keys<<<grid_cap(ctx, ne, 256, 16), 256, 0, st>>>(p->v_off, p->v_adj, row, p->e_begin, p->e_end, bits,
                                                             n, p->opts.per_edge != 0, keys[0], d_bytes);
        HIP_TRY(hipGetLastError());
        int passes = 0;
        RTG_TRY(radix_sort_u64(ctx, keys, ne, bits, &which, &passes, sort_work));
    }
    offsets_fingerprint<<<grid_cap(ctx, (int64_t)n + 1, 256, 4), 256, 0, st>>>(g->d_offsets, n, d_bytes + 4);
    HIP_TRY(hipGetLastError());
    unsigned long long hb[8] = {};
    HIP_TRY(hipMemcpyAsync(hb, d_bytes, sizeof(hb), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    p->alg_bytes = hb[0];
    p->merge_bytes = hb[1];
    p->n_work_edges = (int64_t)hb[2];
    p->offsets_fingerprint = hb[4];
    const int64_t nw = p->n_work_edges;
    RTG_TRY(dev_alloc(&p->d_wspan, (size_t)std::max<int64_t>(nw, 1)));
    if (p->opts.per_edge) {
        RTG_TRY(dev_alloc(&p->d_epos, (size_t)std::max<int64_t>(ne, 1)));
        RTG_TRY(dev_alloc(&p->d_ecnt, (size_t)std::max<int64_t>(nw, 1)));
        HIP_TRY(hipMemsetAsync(p->d_epos, 0xFF, sizeof(uint32_t) * (size_t)std::max<int64_t>(ne, 1), st));
    }
    if (nw > 0) {
        plan_spans<<<grid_cap(ctx, nw, 256, 16), 256, 0, st>>>(p->v_off, p->v_adj, row, p->e_begin, keys[which],
                                                              nw, bits, p->d_wspan, wedge, lens, d_first,
                                                              reinterpret_cast<const int4 *>(p->v_vinfo),
                                                              p->v_s16, p->v_s256, d_bytes + 5, p->v_emap);
        HIP_TRY(hipGetLastError());
        RTG_TRY(scan_exclusive_u64_w(ctx, lens, lens_ex, nw, nullptr, scan_work));
        plan_runs<<<grid_cap(ctx, nw, 256, 16), 256, 0, st>>>(keys[which], nw, bits, d_first, lens, lens_ex, d_cnt,
                                                             p->d_work);
        HIP_TRY(hipGetLastError());
        if (wedge) {
            edge_positions<<<grid_cap(ctx, nw, 256, 16), 256, 0, st>>>(wedge, p->d_wspan, nw, p->e_begin, p->d_epos);
            HIP_TRY(hipGetLastError());
        }
    }
    RTG_TRY(dev_alloc(&p->d_wofs, (size_t)n + 1));
    RTG_TRY(scan_exclusive_u32_w(ctx, d_cnt, (uint32_t *)p->d_wofs, n + 1, nullptr, scan_work));
    uint32_t *d_flag = reinterpret_cast<uint32_t *>(d_bytes + 3);
    if (n > 0) {
        plan_check_offsets<<<grid_cap(ctx, n, 256, 4), 256, 0, st>>>(p->d_wofs, n, nw, d_flag);
        HIP_TRY(hipGetLastError());
    }
    uint32_t flag = 0;
    unsigned long long s16 = 0;
    uint64_t last[2] = {0, 0};  // the last span's length and exclusive prefix: all streamed elements
    HIP_TRY(hipMemcpyAsync(&flag, d_flag, sizeof(flag), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&s16, d_bytes + 5, sizeof(s16), hipMemcpyDeviceToHost, st));
    if (nw > 0) {
        HIP_TRY(hipMemcpyAsync(&last[0], lens + (nw - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(&last[1], lens_ex + (nw - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (flag) return set_error(RTG_EHIP, "batched TC: inconsistent partner offsets in the plan");
    p->streamed16 = (int64_t)s16;
    p->streamed = (int64_t)(last[0] + last[1]);
    return RTG_OK;
}

// ---- device planning: the work queue ----

// Owner buckets: 0 wave owners (small), 1 block owners of out-degree <= 1024, 2 larger block
// owners whose hash table fits the LDS, 3 owners too long for an LDS hash table.
__host__ __device__ __forceinline__ int owner_bucket(int32_t d, int32_t wave_owner_max, int32_t fit_block) {
    return d <= wave_owner_max ? 0 : d > fit_block ? 3 : d <= 1024 ? 1 : 2;
}

// Order key of a cost: 16 steps per octave, heavier first once sorted ascending.
__device__ __forceinline__ uint32_t cost_key(uint64_t k) {
    if (k == 0) return 0xFFFFu;
    const int e = 63 - __clzll((long long)k);
    const uint32_t frac = e >= 4 ? (uint32_t)((k >> (e - 4)) & 15u) : (uint32_t)((k << (4 - e)) & 15u);
    return 0xFFFFu - (1u + (uint32_t)e * 16u + frac);
}

struct QueueParams {
    const int32_t *off;
    const int32_t *adj;
    const int32_t *wofs;
    const unsigned long long *work;  // streamed partner elements per owner
    int32_t n;
    int32_t wave_owner_max, fit_block, max_table, wave_table;
    int32_t only_bucket;             // -1: every owner; else only this bucket's (autotuning)
    int32_t bitmap16;                // degree-ordered copy: bitmaps run from 0 (kBitmap16)
    int32_t small_tables;            // small owners get bitmap or sorted-list wave tables
    int64_t item_work;
    int64_t wave_item_work;          // a small owner is heavy when its work reaches this
    int32_t strat[kBuckets];
};

// What the queue does with owner x: its bucket, item kind and number of block items (0 for a
// small owner or an owner without partners), the wave mode of a small owner.
struct OwnerPlan {
    int32_t bucket, d, c;
    uint32_t kind;
    uint32_t words;  // the owner's bitmap words (its table if a bitmap item)
    int64_t chunks;
    uint64_t w;
    bool small;
};

__device__ __forceinline__ OwnerPlan owner_plan(const QueueParams &q, int32_t x) {
    OwnerPlan o{};
    o.c = q.wofs[x + 1] - q.wofs[x];
    if (o.c == 0) return o;
    o.d = q.off[x + 1] - q.off[x];
    o.w = q.work[x];
    o.bucket = owner_bucket(o.d, q.wave_owner_max, q.fit_block);
    // A small owner with much work (an owner of many edges) would keep one wave busy long after
    // the others finish (on R-MAT scale 21, 208 of 501,708 owners of out-degree <= 256 have
    // 65,536 to 282,000 elements of work): above 16 wave items' work it becomes block items, in
    // chunks of that work, shared by the blocks' waves.
    const uint64_t heavy = 16ull * (uint64_t)q.wave_item_work;
    const bool promoted = o.bucket == 0 && q.wave_item_work > 0 && o.w > heavy;
    if (promoted) o.bucket = 1;
    if (q.only_bucket >= 0 && o.bucket != q.only_bucket) {
        o.c = 0;
        return o;
    }
    if (o.bucket == 0) {
        o.small = true;
        return o;
    }
    const int32_t s = q.strat[o.bucket];
    const uint32_t hi = (uint32_t)q.adj[q.off[x] + o.d - 1];
    const uint32_t words = q.bitmap16 ? bitmap16_words(hi) : bitmap_words((uint32_t)q.adj[q.off[x]], hi);
    o.words = words;
    if (s == kStratBitmap && words <= (uint32_t)q.max_table) o.kind = kItemBitmap;
    else if (s == kStratMerge && o.d <= q.max_table) o.kind = kItemSorted;
    else if (s == kStratSearch || o.d > q.fit_block) o.kind = kItemSearch;
    else o.kind = kItemHash;
    int64_t chunks = promoted ? (int64_t)((o.w + heavy - 1) / heavy)
                              : (int64_t)((o.w + (uint64_t)q.item_work / 2) / (uint64_t)q.item_work);
    o.chunks = std::max<int64_t>(1, std::min<int64_t>(chunks, o.c));
    return o;
}

// Per owner: block items (nblk) or small-owner slot (nsmall), and the sums acc[0] staged bytes,
// acc[1 + b] items of bucket b (owners for bucket 0), acc[5] search items, acc[6] items with a
// bitmap or sorted-list table (including small owners whose wave mode is one of them), acc[7]
// heavy small owners (order key at or above that of wave_item_work: the first acc[7] of the
// sorted small owners), acc[9] block bitmap items (acc[8] holds the scans' totals), acc[10]
// the most bitmap words of one (atomic max), acc[11] the small owners' streamed elements.
__global__ __launch_bounds__(256) void queue_classify(QueueParams q, uint32_t *__restrict__ nblk,
                                                      uint32_t *__restrict__ nsmall, unsigned long long *acc) {
    uint64_t v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t words_max = 0;
    const uint32_t heavy_key = cost_key((uint64_t)q.wave_item_work);
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < q.n; x += (int64_t)gridDim.x * blockDim.x) {
        const OwnerPlan o = owner_plan(q, (int32_t)x);
        uint32_t nb = 0, ns = 0;
        if (o.c > 0) {
            v[0] += 4ull * o.w + 8ull * (uint64_t)o.c + 8ull;  // partner elements and spans, owner row offsets
            if (o.small) {
                ns = 1;
                v[0] += 4ull + 8ull + 4ull * (uint64_t)o.d;  // small-list entry, partner offsets, owner list
                v[1] += 1;
                v[6] += q.small_tables ? 1 : 0;
                v[7] += cost_key(o.w + (uint64_t)o.d) <= heavy_key ? 1 : 0;
                v[9] += o.w;
            } else {
                nb = (uint32_t)o.chunks;
                v[0] += (uint64_t)o.chunks * (16ull + 4ull * (uint64_t)o.d);  // queue entry, owner list per item
                v[1 + o.bucket] += (uint64_t)o.chunks;
                v[5] += o.kind == kItemSearch ? (uint64_t)o.chunks : 0;
                v[6] += (o.kind == kItemBitmap || o.kind == kItemSorted) ? (uint64_t)o.chunks : 0;
                v[8] += o.kind == kItemBitmap ? (uint64_t)o.chunks : 0;
                if (o.kind == kItemBitmap) words_max = max(words_max, o.words);
            }
        }
        nblk[x] = nb;
        nsmall[x] = ns;
    }
    words_max = rtgd::wave_max_u32(words_max);
    if (__lane_id() == 0 && words_max) atomicMax(&acc[10], (unsigned long long)words_max);
    __shared__ uint64_t part[4][10];
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        const uint64_t t = rtgd::wave_sum_u64(v[k]);
        if (__lane_id() == 0) part[threadIdx.x >> 6][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < 10) {
        uint64_t t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w][threadIdx.x];
        // v[0..7] -> acc[0..7], v[8] -> acc[9], v[9] -> acc[11] (acc[8] scan totals, acc[10] a max)
        const int slot = threadIdx.x < 8 ? (int)threadIdx.x : threadIdx.x == 8 ? 9 : 11;
        if (t) atomicAdd(&acc[slot], (unsigned long long)t);
    }
}

// Writes each owner's block items (partners split evenly into o.chunks runs) and small-owner
// entry, with their order keys: a small owner's (position << 16 | heaviness key), a block item's
// (position << 17 | other << 16 | heaviness key), other = 1 for the items that are not block
// bitmaps on a degree-ordered copy, so that those bitmaps come first (their own kernel's queue).
__global__ __launch_bounds__(256) void queue_emit(QueueParams q, const uint32_t *__restrict__ blk_off,
                                                  const uint32_t *__restrict__ small_off, TcItem *__restrict__ items,
                                                  uint64_t *__restrict__ ikeys, int32_t *__restrict__ small,
                                                  uint64_t *__restrict__ swork, uint64_t *__restrict__ skeys) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < q.n; x += (int64_t)gridDim.x * blockDim.x) {
        const OwnerPlan o = owner_plan(q, (int32_t)x);
        if (o.c == 0) continue;
        if (o.small) {
            const uint32_t j = small_off[x];
            small[j] = (int32_t)x;
            swork[j] = o.w + (uint64_t)o.d;
            skeys[j] = ((uint64_t)j << 16) | cost_key(o.w + (uint64_t)o.d);
            continue;
        }
        const int32_t base = q.wofs[x];
        const uint64_t per = o.kind == kItemSearch ? (uint64_t)(32 - __clz((uint32_t)o.d)) / 2 : 1;  // ~log2(d) reads per element
        for (int64_t k = 0; k < o.chunks; ++k) {
            const uint32_t idx = blk_off[x] + (uint32_t)k;
            const int32_t b = base + (int32_t)((int64_t)o.c * k / o.chunks);
            const int32_t e = base + (int32_t)((int64_t)o.c * (k + 1) / o.chunks);
            items[idx] = TcItem{(int32_t)o.kind, (int32_t)x, b, e};
            const uint64_t other = q.bitmap16 && o.kind != kItemBitmap ? 1ull : 0ull;
            ikeys[idx] = ((uint64_t)idx << 17) | (other << 16) | cost_key(o.w / (uint64_t)o.chunks * per + (uint64_t)o.d);
        }
    }
}

// Heaviest first: dst[i] = src[key_i >> 17] over the keys sorted on their low 17 bits (stable,
// so equal heaviness keeps owner order).
__global__ void queue_gather_items(const uint64_t *__restrict__ keys, int64_t n, const TcItem *__restrict__ src,
                                   TcItem *__restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[keys[i] >> 17];
}

__global__ void queue_gather_small(const uint64_t *__restrict__ keys, int64_t n, const int32_t *__restrict__ src,
                                   const uint64_t *__restrict__ wsrc, int32_t *__restrict__ dst, uint64_t *__restrict__ wdst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i] >> 16;
        dst[i] = src[k];
        wdst[i] = wsrc[k];
    }
}

// Wave items: consecutive small owners (heaviest first) in groups with at least `waves` owners
// and about `target` = waves * wave_item_work of work, as a greedy pass would make them, but
// in parallel: the first i0 owners (the heavy ones, each with at least wave_item_work) go
// `waves` to a group; the light ones after them are cut where the prefix of their work,
// counted from i0, crosses a multiple of target (so each light group also has more than `waves`
// owners). P is the exclusive prefix of the work. flag[i] = 1 where a group starts.
__device__ __forceinline__ uint64_t wave_group_of(const uint64_t *__restrict__ pw, int64_t i, int64_t i0, uint64_t target,
                                                  int waves) {
    if (i < i0) return (uint64_t)i / (uint64_t)waves;
    return ((uint64_t)i0 + (uint64_t)waves - 1) / (uint64_t)waves + (pw[i] - pw[i0]) / target;
}

__global__ void queue_group_flags(const uint64_t *__restrict__ pw, int64_t n, int64_t i0, uint64_t target, int waves,
                                  uint32_t *__restrict__ flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = (i == 0 || wave_group_of(pw, i, i0, target, waves) != wave_group_of(pw, i - 1, i0, target, waves)) ? 1u : 0u;
}

__global__ void queue_group_starts(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ gid, int64_t n,
                                   uint32_t *__restrict__ starts) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (flag[i]) starts[gid[i]] = (uint32_t)i;
}

__global__ void queue_wave_items(const uint32_t *__restrict__ starts, const uint32_t *__restrict__ ngroups_p, int64_t nsmall,
                                 int32_t mode, TcItem *__restrict__ queue) {
    const int64_t ng = *ngroups_p;
    for (int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gi < ng; gi += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = (int32_t)starts[gi];
        const int32_t e = gi + 1 < ng ? (int32_t)starts[gi + 1] : (int32_t)nsmall;
        queue[gi] = TcItem{(int32_t)kItemWave, mode, b, e};
    }
}

// Table sizes of a tune: LDS words per block, per wave, the largest owner whose hash table fits
// a block, the largest wave owner.
struct TableSizes {
    int32_t max_table, wave_table, fit_block, wave_owner_max;
};

TableSizes table_sizes(const BatchTune &tune, int32_t max_d, bool bitmap16) {
    const int waves = tune.block_threads / 64;
    // Block table: both cuckoo halves of the largest owner that fits, at least 2048 slots (with
    // bitmap16, a degree-ordered plan with bitmap owners: at least one bitmap from 0, 2,049
    // words — at 2,048 slots no owner of R-MAT scale 17 or 18 got a bitmap, and the block
    // owners ran as hash items in the general kernel), a multiple of 64 * waves so that the
    // wave-private tables split it evenly, at most the cap. Owners whose table would exceed the
    // cap are searched (or given a bitmap or list).
    const int32_t quantum = 64 * waves;
    const int32_t cap = (tune.table_cap > 0 ? tune.table_cap : kMaxTableSlots) / quantum * quantum;
    int32_t fit_block = 0;
    {
        int32_t lo = 0, hi = std::max<int32_t>(max_d, 0);
        while (lo < hi) {
            const int32_t mid = lo + (hi - lo + 1) / 2;
            if (2 * (int64_t)cuckoo_half(mid) <= cap) lo = mid;
            else hi = mid - 1;
        }
        fit_block = lo;
    }
    const int32_t floor = bitmap16 ? std::max<int32_t>(2048, (int32_t)bitmap16_words(0)) : 2048;
    int32_t max_table = std::max<int32_t>(std::min<int32_t>(floor, cap), 2 * (int32_t)cuckoo_half(fit_block));
    max_table = std::min<int32_t>(cap, (max_table + quantum - 1) / quantum * quantum);
    const int32_t wave_table = max_table / waves;
    int32_t fit = 0;  // largest out-degree whose table fits a wave's share
    while (2 * (int32_t)cuckoo_half(fit + 1) <= wave_table) ++fit;
    const int32_t wave_owner_max = tune.wave_owner_max >= 0 ? std::min(tune.wave_owner_max, fit) : fit;
    return {max_table, wave_table, fit_block, wave_owner_max};
}

// Tune-dependent part, on the device: owners classified into buckets with their strategies, the
// block items and small owners ordered heaviest first (radix sort on a log-scale key), small
// owners grouped into wave items, and the kernel instance. only_bucket >= 0 builds the queue of
// that bucket alone (autotuning). Two small downloads: the sizes, then the number of wave items.
int plan_items(rtg_ctx *ctx, rtg_tc_plan *p, const BatchTune &tune, int only_bucket = -1) {
    using namespace rtg;
    const rtg_graph *g = p->g;
    const int32_t n = g->n;
    hipStream_t st = ctx->stream;
    const int threads = tune.block_threads;
    if (threads != 256 && threads != 512) return set_error(RTG_EINVAL, "batched TC: threads must be 256 or 512");
    const int waves = threads / 64;
    const TableSizes ts = table_sizes(tune, g->max_row,
                                      p->renumbered && (tune.strat[1] == kStratBitmap || tune.strat[2] == kStratBitmap));
    // Wave items of at most 1/4 per CU of all the streamed elements: on a small graph the tuned
    // size leaves most CUs idle (BA(100,000, 5): all 91,184 owners small, 156 wave items of
    // 4,096 elements per wave took 1.23 ms, 1,226 items of 512 0.23 ms). Graphs of more than
    // 2^23 streamed elements keep the tuned size at 2,048.
    const int64_t wave_item = std::min<int64_t>(
        tune.wave_item_work, std::max<int64_t>(256, p->streamed / ((int64_t)waves * ctx->cu_count * 4)));
    p->wave_item_used = wave_item;
    // Small owners get hash tables: by their strategy, or on a degree-ordered copy whose wave
    // share cannot hold a bitmap from 0 (bitmap16_words is at least 2,049), where a bitmap
    // would fall back to the hash table anyway.
    const bool small_hashed = tune.strat[0] == kStratHash ||
                              (p->renumbered && tune.strat[0] == kStratBitmap && (uint32_t)ts.wave_table < bitmap16_words(0));
    QueueParams q{p->v_off, p->v_adj, p->d_wofs, p->d_work, n, ts.wave_owner_max, ts.fit_block, ts.max_table,
                  ts.wave_table, only_bucket, p->renumbered ? 1 : 0,
                  !small_hashed && (tune.strat[0] == kStratBitmap || tune.strat[0] == kStratMerge) ? 1 : 0,
                  tune.item_work, wave_item, {}};
    for (int b = 0; b < kBuckets; ++b) q.strat[b] = tune.strat[b];
    if (p->d_items) (void)hipFree(p->d_items);
    if (p->d_small) (void)hipFree(p->d_small);
    p->d_items = nullptr;
    p->d_small = nullptr;

    // pass 1: classification, counts, sums
    const size_t nn = (size_t)n + 1;
    const size_t b1 = Arena::round(sizeof(uint32_t) * nn) * 4 + Arena::round(sizeof(unsigned long long) * 16) +
                      Arena::round(scan_work_bytes((int64_t)nn)) + 4096;
    // pass 1's buffers in the context's scratch (after the graph-level plan, already at least the
    // size of that plan's temporaries), pass 2's behind them when they fit there
    void *s1 = nullptr;
    RTG_TRY(ctx_scratch(ctx, b1, &s1));
    char *buf1 = static_cast<char *>(s1);
    Arena a1{buf1, 0, b1};
    uint32_t *nblk = a1.take<uint32_t>(nn), *nsmall = a1.take<uint32_t>(nn);
    uint32_t *blk_off = a1.take<uint32_t>(nn), *small_off = a1.take<uint32_t>(nn);
    unsigned long long *acc = a1.take<unsigned long long>(16);  // [0..7] sums, [8] totals (two u32)
    void *scan1 = a1.take<char>(scan_work_bytes((int64_t)nn));
    HIP_TRY(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * 16, st));
    uint32_t *totals = reinterpret_cast<uint32_t *>(acc + 8);
    if (n > 0) {
        queue_classify<<<grid_cap(ctx, n, 256, 8), 256, 0, st>>>(q, nblk, nsmall, acc);
        HIP_TRY(hipGetLastError());
        RTG_TRY(scan_exclusive_u32_w(ctx, nblk, blk_off, n, totals, scan1));
        RTG_TRY(scan_exclusive_u32_w(ctx, nsmall, small_off, n, totals + 1, scan1));
    }
    unsigned long long h[12] = {};
    HIP_TRY(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const int64_t nb = (int64_t)(uint32_t)(h[8] & 0xFFFFFFFFull), ns = (int64_t)(uint32_t)(h[8] >> 32);
    const int64_t heavy_small = std::min<int64_t>((int64_t)h[7], ns);
    // a wave item per small owner at most; 16-bit spans read 2 of the 4 bytes per element
    p->staged_bytes = h[0] + 16ull * (uint64_t)ns - 2ull * (uint64_t)p->streamed16;
    for (int b = 0; b < kBuckets; ++b) p->bucket_items[b] = (int64_t)h[1 + b];
    p->n_search_items = (int64_t)h[5];
    p->all_strategies = h[6] > 0;
    p->n_block_items = nb;
    p->n_small = ns;
    p->n_bitmap_items = p->renumbered ? (int64_t)h[9] : 0;
    p->bitmap_table = (int32_t)std::min<uint64_t>(h[10], (uint64_t)ts.max_table);  // the bitmap kernel's table
    p->small_work = (int64_t)h[11];

    // pass 2: items, small owners, order, wave groups
    RTG_TRY(dev_alloc(&p->d_items, (size_t)std::max<int64_t>(nb + ns, 1)));
    RTG_TRY(dev_alloc(&p->d_small, (size_t)std::max<int64_t>(ns, 1)));
    const int64_t mx = std::max<int64_t>(std::max(nb, ns), 1);
    const size_t b2 = Arena::round(sizeof(TcItem) * (size_t)mx) + Arena::round(sizeof(uint64_t) * (size_t)mx) * 6 +
                      Arena::round(sizeof(int32_t) * (size_t)mx) + Arena::round(sizeof(uint32_t) * (size_t)mx) * 3 +
                      Arena::round(radix_sort_work_bytes(mx)) + Arena::round(scan_work_bytes(mx)) + 4096;
    char *buf2 = nullptr;
    DevFree f2;
    if (ctx->scratch_bytes >= Arena::round(b1) + b2) {
        buf2 = buf1 + Arena::round(b1);
    } else {
        RTG_TRY(dev_alloc(&buf2, b2));
        f2.p = buf2;
    }
    Arena a2{buf2, 0, b2};
    TcItem *items_tmp = a2.take<TcItem>((size_t)mx);
    uint64_t *k0 = a2.take<uint64_t>((size_t)mx), *k1 = a2.take<uint64_t>((size_t)mx);
    uint64_t *sk0 = a2.take<uint64_t>((size_t)mx), *sk1 = a2.take<uint64_t>((size_t)mx);
    uint64_t *swork = a2.take<uint64_t>((size_t)mx), *sw = a2.take<uint64_t>((size_t)mx);
    int32_t *small_tmp = a2.take<int32_t>((size_t)mx);
    uint32_t *flag = a2.take<uint32_t>((size_t)mx), *gid = a2.take<uint32_t>((size_t)mx);
    uint32_t *starts = a2.take<uint32_t>((size_t)mx);
    void *sort2 = a2.take<char>(radix_sort_work_bytes(mx));
    void *scan2 = a2.take<char>(scan_work_bytes(mx));
    uint32_t *ngroups = totals + 2;
    HIP_TRY(hipMemsetAsync(ngroups, 0, sizeof(uint32_t), st));
    if (nb + ns > 0) {
        queue_emit<<<grid_cap(ctx, n, 256, 8), 256, 0, st>>>(q, blk_off, small_off, items_tmp, k0, small_tmp, swork, sk0);
        HIP_TRY(hipGetLastError());
    }
    if (nb > 0) {
        uint64_t *kk[2] = {k0, k1};
        int which = 0, passes = 0;
        RTG_TRY(radix_sort_u64(ctx, kk, nb, 17, &which, &passes, sort2));
        queue_gather_items<<<grid_cap(ctx, nb, 256, 8), 256, 0, st>>>(kk[which], nb, items_tmp, p->d_items);
        HIP_TRY(hipGetLastError());
    }
    if (ns > 0) {
        uint64_t *kk[2] = {sk0, sk1};
        int which = 0, passes = 0;
        RTG_TRY(radix_sort_u64(ctx, kk, ns, 16, &which, &passes, sort2));
        queue_gather_small<<<grid_cap(ctx, ns, 256, 8), 256, 0, st>>>(kk[which], ns, small_tmp, swork, p->d_small, sw);
        HIP_TRY(hipGetLastError());
        RTG_TRY(scan_exclusive_u64_w(ctx, sw, swork, ns, nullptr, scan2));  // exclusive work prefix, into swork
        const uint64_t target = (uint64_t)std::max<int64_t>(1, wave_item * waves);
        queue_group_flags<<<grid_cap(ctx, ns, 256, 8), 256, 0, st>>>(swork, ns, heavy_small, target, waves, flag);
        HIP_TRY(hipGetLastError());
        RTG_TRY(scan_exclusive_u32_w(ctx, flag, gid, ns, ngroups, scan2));
        queue_group_starts<<<grid_cap(ctx, ns, 256, 8), 256, 0, st>>>(flag, gid, ns, starts);
        HIP_TRY(hipGetLastError());
        const int32_t mode = small_hashed ? (int32_t)kCuckoo
                             : tune.strat[0] == kStratBitmap ? (int32_t)kBitmap
                             : tune.strat[0] == kStratMerge  ? (int32_t)kSorted
                                                             : (int32_t)kCuckoo;
        queue_wave_items<<<grid_cap(ctx, ns, 256, 8), 256, 0, st>>>(starts, ngroups, ns, mode, p->d_items + nb);
        HIP_TRY(hipGetLastError());
    }
    uint32_t ng = 0;
    HIP_TRY(hipMemcpyAsync(&ng, ngroups, sizeof(ng), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    p->n_items = nb + (int64_t)ng;
    for (int b = 0; b < kBuckets; ++b) p->strat[b] = tune.strat[b];

    // Persistent grid: as many blocks as fit on the device at once.
    const size_t lds = sizeof(uint32_t) * (size_t)ts.max_table;
    BatchKernel k = batched_kernel(threads, tune.unroll, false, p->all_strategies, p->renumbered);
    if (!k) return set_error(RTG_EINVAL, "batched TC: unsupported (threads, unroll) = (" + std::to_string(threads) + ", " +
                                             std::to_string(tune.unroll) + ")");
    int per_cu = 0;
    HIP_TRY(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, threads, lds));
    if (per_cu < 1) per_cu = 1;
    p->max_table = ts.max_table;
    p->block_threads = threads;
    p->wave_owner_max = ts.wave_owner_max;
    p->fit_block = ts.fit_block;
    p->kernel = (void *)k;
    // items for the general kernel: after the block bitmaps, and before the wave items when those
    // have their own kernel
    const int64_t n_general = p->n_items - p->n_bitmap_items - (p->renumbered && small_hashed ? (int64_t)ng : 0);
    // occupancy-sized persistent grid of kernel kk for `items` queue entries; `side` (the wave
    // kernel beside the bitmap kernel): blocks per CU in proportion to the wave items' share of
    // the count's time
    const auto grid_of_kernel = [&](BatchKernel kk, int64_t items, int32_t *grid, size_t kl = 0, bool side = false) -> int {
        int pc = 0;
        if (!kl) kl = lds;
        HIP_TRY(hipFuncSetAttribute((const void *)kk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kl));
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, kk, threads, kl));
        if (side) {
            pc = std::max(1, std::min(pc, (int)std::ceil(2.0 * pc * p->wave_share)));
            p->side_blocks = pc;
        }
        *grid = (int32_t)std::min<int64_t>((int64_t)std::max(pc, 1) * ctx->cu_count, std::max<int64_t>(items, 1));
        return RTG_OK;
    };
    // degree-ordered plans whose small owners are hashed run the wave items in their own kernel
    p->n_wave_items = p->renumbered && small_hashed ? (int64_t)ng : 0;
    // The wave items run beside the block bitmaps, on the context's side stream. The bitmap
    // kernel keeps its full grid: its blocks beyond what fits start as the wave kernel's retire.
    // The wave kernel gets blocks per CU in proportion to its share of the count's time (a wave
    // item's element costs about 4 block bitmap elements), doubled so that it does not become the
    // tail (the bitmap kernel, bound by HBM, loses little to it). Measured on R-MAT scale 21 (the
    // small owners 1.6% of the streamed elements: 1 block per CU) 7.16 ms, 7.20-7.40 with 2-8
    // blocks; on the n = 3M configuration model (16%: 7 blocks) 6.11-6.15 ms with 4-8 blocks,
    // 6.43 with 3, 15.7 with 1, 6.47 with the two kernels one after the other.
    p->side = p->n_bitmap_items > 0 && p->n_wave_items > 0;
    p->side_blocks = 0;
    {
        const double sw = 4.0 * (double)p->small_work;
        const double rest = (double)std::max<int64_t>(p->streamed - p->small_work, 0);
        p->wave_share = sw / std::max(sw + rest, 1.0);
    }
    // its wave tables: both cuckoo halves of the largest small owner (not the block table's share)
    p->wave_table_wv = std::min<int32_t>(ts.wave_table, 2 * (int32_t)cuckoo_half(ts.wave_owner_max));
    const size_t lds_wv = sizeof(uint32_t) * (size_t)p->wave_table_wv * (size_t)waves;
    if (p->n_wave_items > 0) {
        BatchKernel kw = wave_kernel(threads, tune.unroll, false);
        p->kernel_wv = (void *)kw;
        RTG_TRY(grid_of_kernel(kw, p->n_wave_items, &p->grid_wv, lds_wv, p->side));
    }
    if (p->n_bitmap_items > 0) {
        BatchKernel kb = bitmap_kernel(threads, tune.unroll, false);
        p->kernel_bm = (void *)kb;
        RTG_TRY(grid_of_kernel(kb, p->n_bitmap_items, &p->grid_bm, sizeof(uint32_t) * (size_t)p->bitmap_table));
    }
    if (p->opts.per_edge) {
        BatchKernel ke = batched_kernel(threads, tune.unroll, true, p->all_strategies, p->renumbered);
        p->kernel_pe = (void *)ke;
        RTG_TRY(grid_of_kernel(ke, n_general, &p->grid_pe));
        if (p->n_bitmap_items > 0) {
            BatchKernel kbe = bitmap_kernel(threads, tune.unroll, true);
            p->kernel_pe_bm = (void *)kbe;
            RTG_TRY(grid_of_kernel(kbe, p->n_bitmap_items, &p->grid_pe_bm, sizeof(uint32_t) * (size_t)p->bitmap_table));
        }
        if (p->n_wave_items > 0) {
            BatchKernel kwe = wave_kernel(threads, tune.unroll, true);
            p->kernel_pe_wv = (void *)kwe;
            RTG_TRY(grid_of_kernel(kwe, p->n_wave_items, &p->grid_pe_wv, lds_wv, p->side));
        }
    }
    p->unroll = tune.unroll;
    p->long_min = tune.long_min;
    p->grab = tune.grab;
    p->tune = tune_string(tune);
    p->grid = (int32_t)std::min<int64_t>((int64_t)per_cu * ctx->cu_count, std::max<int64_t>(n_general, 1));
    return RTG_OK;
}

// ---- autotuning ----------------------------------------------------------------------------
// Candidate configurations, timed on the plan's own graph. The first is the default.
const char *kCandidates[] = {
    "threads=256,unroll=4,item=2097152,wave_item=4096,wave_max=-1,long=128,grab=64",
    "threads=256,unroll=4,item=2097152,wave_item=4096,wave_max=64,long=128,grab=64",
    "threads=256,unroll=4,item=524288,wave_item=4096,wave_max=-1,long=128,grab=64",
    "threads=256,unroll=4,item=2097152,wave_item=4096,wave_max=-1,long=64,grab=64",
    "threads=256,unroll=4,item=2097152,wave_item=4096,wave_max=-1,long=128,grab=32",
    "threads=256,unroll=4,item=2097152,wave_item=16384,wave_max=0,long=128,grab=64",
    "threads=256,unroll=8,item=2097152,wave_item=4096,wave_max=-1,long=128,grab=64",
    "threads=256,unroll=8,item=2097152,wave_item=4096,wave_max=64,long=128,grab=64",
    "threads=256,unroll=2,item=131072,wave_item=4096,wave_max=-1,long=32,grab=16",
    "threads=512,unroll=4,item=2097152,wave_item=4096,wave_max=-1,long=128,grab=64",
    "threads=512,unroll=8,item=2097152,wave_item=4096,wave_max=-1,long=128,grab=64",
    // smaller wave items (more, lighter groups of small owners) with finer queue grabs
    "threads=256,unroll=4,item=2097152,wave_item=2048,wave_max=64,long=128,grab=32",
    "threads=256,unroll=4,item=2097152,wave_item=2048,wave_max=128,long=128,grab=32",
    "threads=256,unroll=4,item=2097152,wave_item=1024,wave_max=64,long=128,grab=32",
    "threads=256,unroll=8,item=2097152,wave_item=2048,wave_max=64,long=128,grab=32",
    // small owners up to out-degree 256 (degree-ordered copies: the owners above it are block
    // bitmaps, and the wave-items kernel's smaller tables fit more blocks per CU)
    "threads=256,unroll=8,item=2097152,wave_item=2048,wave_max=256,long=128,grab=32",
    "threads=256,unroll=8,item=2097152,wave_item=1024,wave_max=256,long=128,grab=32",
    "threads=256,unroll=8,item=2097152,wave_item=2048,wave_max=256,long=128,grab=64",
};

// Strategies the autotuner times per bucket (the first is the default).
const int32_t kBucketCandidates[kBuckets][4] = {
    {kStratHash, kStratBitmap, kStratMerge, -1},
    {kStratHash, kStratBitmap, kStratMerge, kStratSearch},
    {kStratHash, kStratBitmap, kStratMerge, kStratSearch},
    {kStratSearch, kStratMerge, kStratBitmap, -1},
};

// Revision of the batched kernel's code and candidates, part of the cache key: raised when a
// change to the kernel or to the candidate list can change which candidate the autotuner keeps,
// so that entries tuned earlier are re-tuned.
constexpr int kTuneRevision = 19;

// Cache key: device architecture and CU count, kernel revision, graph shape, the row offsets'
// fingerprint, the edge range and whether the plan is for per-edge counts or counts on the
// degree-ordered copy.
std::string tune_key(rtg_ctx *ctx, const rtg_tc_plan *p) {
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s/cu%d/r%d/n%d/m%lld/%016llx/e%lld-%lld%s", ctx->prop.gcnArchName, ctx->cu_count,
                  kTuneRevision, p->g->n, (long long)p->g->m, (unsigned long long)p->offsets_fingerprint,
                  (long long)p->e_begin, (long long)p->e_end, p->opts.per_edge ? (p->renumbered ? "/pe/rn" : "/pe")
                                                                              : (p->renumbered ? "/rn" : ""));
    return buf;
}

// The cache file is a flat JSON object of "key": "tune string" entries.
std::vector<std::pair<std::string, std::string>> read_tune_cache(const char *path) {
    std::vector<std::pair<std::string, std::string>> out;
    if (!path) return out;
    std::ifstream f(path);
    if (!f) return out;
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<std::string> strs;
    size_t i = 0;
    while (i < text.size()) {
        if (text[i] == '"') {
            std::string cur;
            ++i;
            while (i < text.size() && text[i] != '"') {
                if (text[i] == '\\' && i + 1 < text.size()) ++i;
                cur.push_back(text[i++]);
            }
            strs.push_back(cur);
        }
        ++i;
    }
    for (size_t k = 0; k + 1 < strs.size(); k += 2) out.push_back({strs[k], strs[k + 1]});
    return out;
}

int write_tune_cache(const char *path, const std::vector<std::pair<std::string, std::string>> &entries) {
    std::ofstream f(path, std::ios::out | std::ios::trunc);
    if (!f) return rtg::set_error(RTG_EIO, std::string("cannot write tuning cache ") + path);
    f << "{\n";
    for (size_t i = 0; i < entries.size(); ++i)
        f << "  \"" << entries[i].first << "\": \"" << entries[i].second << "\"" << (i + 1 < entries.size() ? "," : "") << "\n";
    f << "}\n";
    return RTG_OK;
}

int launch_count(rtg_ctx *ctx, rtg_tc_plan *p);
int launch_per_edge(rtg_ctx *ctx, rtg_tc_plan *p, uint32_t *eout);

// Median of three timed launches after one warm-up, of the plan as it stands (the kernel the
// plan will run: the per-edge instantiation into a scratch output for per-edge plans).
int time_plan(rtg_ctx *ctx, rtg_tc_plan *p, uint32_t *pe_out, float *median_ms) {
    float ms[3];
    for (int r = -1; r < 3; ++r) {
        HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));
        if (pe_out) RTG_TRY(launch_per_edge(ctx, p, pe_out));
        else RTG_TRY(launch_count(ctx, p));
        HIP_TRY(hipEventRecord(ctx->ev[3], ctx->stream));
        HIP_TRY(hipEventSynchronize(ctx->ev[3]));
        if (r >= 0) ms[r] = rtg::elapsed_ms(ctx->ev[2], ctx->ev[3]);
    }
    std::sort(ms, ms + 3);
    *median_ms = ms[1];
    return RTG_OK;
}

// Two stages: every tile-shape candidate with the default strategies, timed on the whole plan;
// then, for the fastest shape, each bucket's strategy candidates timed on a queue of that bucket
// alone (a strategy no owner of the bucket can use is skipped). The combination is kept if it
// is at least as fast as the default strategies on the whole plan.
// The starting configuration: hash tables for every bucket but the last, except on a
// degree-ordered copy, where block owners' lists lie in a short id range and their bitmaps (one
// LDS read per element, no hashing) are the better default, with the shape the autotuner keeps
// on both benchmark graphs (small owners up to out-degree 256, wave items of about 2,048
// elements per wave, 8 loads per lane and round, grabs of 32): a plan without a cache entry
// (the drop-in CountOnGPU, one-off counts) then counts R-MAT scale 21 in 7.17 ms instead of 8.51
// and the n = 3M configuration model in 6.29 instead of 6.64 (6.15 tuned, with grabs of 64).
// Every candidate sets these fields, so the autotuner's measurements do not depend on them.
BatchTune default_tune(const rtg_tc_plan *p) {
    BatchTune t;
    if (p->renumbered) {
        t.strat[1] = t.strat[2] = kStratBitmap;
        t.unroll = 8;
        t.wave_item_work = 2048;
        t.wave_owner_max = 256;
        t.grab = 32;
    }
    return t;
}

int autotune(rtg_ctx *ctx, rtg_tc_plan *p, BatchTune *best) {
    uint32_t *pe_out = nullptr;
    DevFree fpe;
    if (p->opts.per_edge) {
        RTG_TRY(dev_alloc(&pe_out, (size_t)std::max<int64_t>(p->g->m, 1)));
        fpe.p = pe_out;
    }
    float best_ms = 1e30f;
    std::string best_s;
    for (const char *cand : kCandidates) {
        BatchTune t = default_tune(p);
        tune_from_string(t, cand);
        float ms = 0.f;
        if (plan_items(ctx, p, t) != RTG_OK || time_plan(ctx, p, pe_out, &ms) != RTG_OK) continue;
        if (ms < best_ms) {
            best_ms = ms;
            best_s = cand;
        }
    }
    if (best_s.empty()) return rtg::set_error(RTG_EHIP, "autotune: no candidate ran");
    BatchTune shape = default_tune(p);
    tune_from_string(shape, best_s);
    BatchTune mix = shape;
    for (int b = 0; b < kBuckets; ++b) {
        float bucket_best = 1e30f;
        for (int ci = 0; ci < 4; ++ci) {
            const int32_t s = kBucketCandidates[b][ci];
            if (s < 0) continue;
            BatchTune t = shape;
            t.strat[b] = s;
            float ms = 0.f;
            if (plan_items(ctx, p, t, b) != RTG_OK) continue;
            if (p->bucket_items[b] == 0) break;  // empty bucket: keep the default
            if ((s == kStratBitmap || s == kStratMerge) && !p->all_strategies) continue;  // no owner can use it
            if (time_plan(ctx, p, pe_out, &ms) != RTG_OK) continue;
            if (ms < bucket_best) {
                bucket_best = ms;
                mix.strat[b] = s;
            }
        }
    }
    *best = shape;
    if (std::memcmp(mix.strat, shape.strat, sizeof(mix.strat)) != 0) {
        float ms = 0.f;
        if (plan_items(ctx, p, mix) == RTG_OK && time_plan(ctx, p, pe_out, &ms) == RTG_OK && ms <= best_ms) *best = mix;
    }
    p->autotuned = true;
    return RTG_OK;
}

// Chooses the tune for a batched plan: RTG_TC_TUNE, else the cache entry, else the
// autotuner (when asked for), else the defaults; then builds the work queue for it.
int plan_batched(rtg_ctx *ctx, rtg_tc_plan *p) {
    using Clk = std::chrono::steady_clock;
    const auto ms_since = [](Clk::time_point t) { return std::chrono::duration<double, std::milli>(Clk::now() - t).count(); };
    // Long partner lists are read through buffer resources sized in bytes by a 32-bit int, and
    // lane offsets are int byte offsets: rows are bounded to 2^29 - 1 elements (a degree-ordered
    // CSR's out-degrees stay below sqrt(2m), far under it).
    if (p->g->max_row > (1 << 29) - 1)
        return rtg::set_error(RTG_EUNSUPPORTED, "batched TC: a row longer than 2^29 - 1 elements");
    auto t0 = Clk::now();
    RTG_TRY(plan_graph(ctx, p));
    p->t_plan_graph_ms = ms_since(t0);
    t0 = Clk::now();
    BatchTune tune = default_tune(p);
    const char *env = std::getenv("RTG_TC_TUNE");
    const std::string key = tune_key(ctx, p);
    auto cache = read_tune_cache(p->opts.tune_cache);
    bool have = false;
    if (env) {
        tune_from_string(tune, env);
        have = true;
    } else {
        for (const auto &kv : cache)
            if (kv.first == key) {
                tune_from_string(tune, kv.second);
                have = true;
                p->tune_from_cache = true;
            }
    }
    if (!have && p->opts.autotune) {
        RTG_TRY(autotune(ctx, p, &tune));
        if (p->opts.tune_cache) {
            cache.erase(std::remove_if(cache.begin(), cache.end(), [&](const auto &kv) { return kv.first == key; }), cache.end());
            cache.push_back({key, tune_string(tune)});
            RTG_TRY(write_tune_cache(p->opts.tune_cache, cache));
        }
    }
    p->t_plan_tune_ms = ms_since(t0);
    t0 = Clk::now();
    const int rc = plan_items(ctx, p, tune);
    p->t_plan_items_ms = ms_since(t0);
    return rc;
}

// The batched queue, in up to three kernels: the block bitmap items of a degree-ordered plan
// (items [0, n_bitmap_items)) by kb, the other block items (and the wave items, unless they have
// their own kernel) by the general kernel kg, and the wave items (the queue's last n_wave_items)
// by kw. With both bitmap and wave items (p->side), the wave kernel runs on the context's side
// stream beside the bitmap kernel (its blocks per CU sized by its share of the time, the bitmap
// kernel's grid full): the tail of the heavy wave items no longer follows the bitmap kernel.
// eout: per-edge counts.
int launch_items(rtg_tc_plan *p, hipStream_t st, BatchKernel kb, int32_t grid_b, BatchKernel kg, int32_t grid_g,
                 BatchKernel kw, int32_t grid_w, uint32_t *eout) {
    rtg_ctx *ctx = p->ctx;
    const bool beside = p->side;
    BatchArgs a{p->v_off, p->v_adj, p->d_wofs, p->d_wspan, p->d_small, p->d_items,
                p->n_bitmap_items, p->d_count, p->bitmap_table, 0,
                p->long_min, p->grab, eout, p->v_adj16};
    BatchArgs aw = a;  // the wave items'
    aw.items = p->d_items + (p->n_items - p->n_wave_items);
    aw.n_items = p->n_wave_items;
    aw.max_table = p->max_table;
    aw.wave_table = p->wave_table_wv;
    const size_t lds_w = sizeof(uint32_t) * (size_t)p->wave_table_wv * (p->block_threads / 64);
    if (beside) {
        HIP_TRY(hipEventRecord(ctx->fork, st));
        HIP_TRY(hipStreamWaitEvent(ctx->side, ctx->fork, 0));
        launch_batched(kw, grid_w, p->block_threads, lds_w, ctx->side, aw);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ctx->join, ctx->side));
    }
    if (p->n_bitmap_items > 0) launch_batched(kb, grid_b, p->block_threads, sizeof(uint32_t) * (size_t)p->bitmap_table, st, a);
    const size_t lds = sizeof(uint32_t) * (size_t)p->max_table;
    a.max_table = p->max_table;
    a.wave_table = p->max_table / (p->block_threads / 64);
    const int64_t general_end = p->n_items - p->n_wave_items;
    if (general_end > p->n_bitmap_items) {
        a.items = p->d_items + p->n_bitmap_items;
        a.n_items = general_end - p->n_bitmap_items;
        launch_batched(kg, grid_g, p->block_threads, lds, st, a);
    }
    if (p->n_wave_items > 0 && !beside) launch_batched(kw, grid_w, p->block_threads, lds_w, st, aw);
    HIP_TRY(hipGetLastError());
    if (beside) HIP_TRY(hipStreamWaitEvent(st, ctx->join, 0));
    return RTG_OK;
}

int launch_count(rtg_ctx *ctx, rtg_tc_plan *p) {
    const rtg_graph *g = p->g;
    hipStream_t st = ctx->stream;
    reset_counters<<<1, 64, 0, st>>>(p->d_count);
    HIP_TRY(hipGetLastError());
    const int64_t ne = p->e_end - p->e_begin;
    switch (p->opts.method) {
        case RTG_TC_PLAIN_THREAD:
            if (ne > 0)
                tc_plain_thread<<<grid_cap(ctx, ne, kPlainThreads, 32), kPlainThreads, 0, st>>>(
                    g->d_offsets, g->d_adjs, p->d_row, p->e_begin, p->e_end, p->d_count);
            break;
        case RTG_TC_PLAIN_WAVE:
            if (ne > 0)
                tc_plain_wave<<<grid_cap(ctx, ne * 64, kPlainThreads, 32), kPlainThreads, 0, st>>>(
                    g->d_offsets, g->d_adjs, p->d_row, p->e_begin, p->e_end, p->d_count);
            break;
        case RTG_TC_BATCHED: {
            if (p->n_items == 0) break;
            RTG_TRY(launch_items(p, st, (BatchKernel)p->kernel_bm, p->grid_bm, (BatchKernel)p->kernel, p->grid,
                                 (BatchKernel)p->kernel_wv, p->grid_wv, nullptr));
            break;
        }
        default:
            return rtg::set_error(RTG_EINVAL, "unknown TC method");
    }
    HIP_TRY(hipGetLastError());
    return RTG_OK;
}

// Per-edge counts of the plan's edge range: the per-edge instantiations of the batched kernels
// store each span's count by position (p->d_ecnt; the total is accumulated as in a count), then
// the gather writes the range's counts into eout in edge order (0 for edges without a span).
int launch_per_edge(rtg_ctx *ctx, rtg_tc_plan *p, uint32_t *eout) {
    hipStream_t st = ctx->stream;
    const int64_t ne = p->e_end - p->e_begin;
    reset_counters<<<1, 64, 0, st>>>(p->d_count);
    HIP_TRY(hipGetLastError());
    if (p->n_items > 0)
        RTG_TRY(launch_items(p, st, (BatchKernel)p->kernel_pe_bm, p->grid_pe_bm, (BatchKernel)p->kernel_pe, p->grid_pe,
                             (BatchKernel)p->kernel_pe_wv, p->grid_pe_wv, p->d_ecnt));
    if (ne > 0) {
        gather_edge_counts<<<grid_cap(ctx, ne, 256, 16), 256, 0, st>>>(p->d_epos, ne, p->d_ecnt, eout + p->e_begin);
        HIP_TRY(hipGetLastError());
    }
    return RTG_OK;
}

// Whether a plan counts on the graph's degree-ordered copy: batched plans over the whole edge
// range (an edge range is a range of the graph's own edges); per-edge plans write through the
// copy's edge map. By default graphs of at least 2^20 edges, which repay the copy's one-time
// build (a few ms on 83M edges); RTG_TC_RENUMBER=0 turns it off, =1 also takes smaller graphs.
bool want_renumber(const rtg_tc_plan *p) {
    if (p->opts.method != RTG_TC_BATCHED) return false;
    if (p->e_begin != 0 || p->e_end != p->g->m || p->g->m <= 0 || p->g->n <= 0) return false;
    const char *e = std::getenv("RTG_TC_RENUMBER");
    if (e && *e) return e[0] != '0';
    return p->g->m >= ((int64_t)1 << 20);
}

}  // namespace

int rtg_tc_plan_create_impl(rtg_ctx *ctx, const rtg_graph *g, const rtg_tc_opts *opts, rtg_tc_plan **out) {
    using namespace rtg;
    auto *p = new rtg_tc_plan();
    p->ctx = ctx;
    p->g = g;
    p->opts = *opts;
    p->e_begin = std::max<int64_t>(0, opts->e_begin);
    p->e_end = (opts->e_end <= 0 || opts->e_end > g->m) ? g->m : opts->e_end;
    if (p->e_begin > p->e_end) p->e_begin = p->e_end;
    p->v_off = g->d_offsets;
    p->v_adj = g->d_adjs;
    hipStream_t st = ctx->stream;
    int rc = RTG_OK;
    HIP_TRY(hipEventRecord(ctx->ev[5], st));
    do {
        if (want_renumber(p)) {
            const bool built_copy = g->renum == nullptr;
            const bool built_map = p->opts.per_edge && (built_copy || g->renum->d_emap == nullptr);
            const rtg_renum *r = nullptr;
            if ((rc = graph_renumbered(ctx, g, p->opts.per_edge != 0, &r)) != RTG_OK) break;
            p->v_emap = r->d_emap;
            p->v_s16 = r->d_s16;
            p->v_s256 = r->d_s256;
            p->v_off = r->d_off;
            p->v_adj = r->d_adj;
            p->v_adj16 = r->d_adj16;
            p->v_vinfo = r->d_vinfo;
            p->renumbered = true;
            p->t_renum_ms = (built_copy ? r->t_build_ms : 0.0) + (built_map ? r->t_emap_ms : 0.0);
        }
        if ((rc = dev_alloc(&p->d_count, 4)) != RTG_OK) break;
        // the plain counts read each edge's row; a batched plan only plans with it (in scratch)
        if (opts->method != RTG_TC_BATCHED) {
            if ((rc = dev_alloc(&p->d_row, (size_t)g->m)) != RTG_OK) break;
            if (g->n > 0) {
                edge_rows_kernel<<<grid_cap(ctx, g->n, 256, 16), 256, 0, st>>>(p->v_off, g->n, p->d_row);
                if (hipGetLastError() != hipSuccess) { rc = set_error(RTG_EHIP, "edge_rows launch failed"); break; }
            }
        }
        if (opts->method == RTG_TC_BATCHED) {
            if ((rc = plan_batched(ctx, p)) != RTG_OK) break;
        } else if (opts->method != RTG_TC_PLAIN_THREAD && opts->method != RTG_TC_PLAIN_WAVE) {
            rc = set_error(RTG_EUNSUPPORTED, "rtg_tc_plan_create: method not plannable");
            break;
        }
        if (opts->use_hip_graph) {
            if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) { rc = set_error(RTG_EHIP, "capture begin failed"); break; }
            int lrc = launch_count(ctx, p);
            hipError_t e = hipStreamEndCapture(st, &p->graph);
            if (lrc != RTG_OK) { rc = lrc; break; }
            if (e != hipSuccess) { rc = set_error(RTG_EHIP, "capture end failed"); break; }
            if (hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess) { rc = set_error(RTG_EHIP, "graph instantiate failed"); break; }
        }
    } while (0);
    if (rc == RTG_OK) {
        if (hipEventRecord(ctx->ev[6], st) != hipSuccess || hipEventSynchronize(ctx->ev[6]) != hipSuccess)
            rc = set_error(RTG_EHIP, "plan sync failed");
        else
            p->t_plan_ms = elapsed_ms(ctx->ev[5], ctx->ev[6]);
    }
    if (rc != RTG_OK) {
        rtg_tc_plan_destroy(p);
        return rc;
    }
    *out = p;
    return RTG_OK;
}

static void fill_stats(rtg_tc_plan *p, rtg_tc_stats *s) {
    s->t_plan_ms = p->t_plan_ms;
    s->alg_bytes = p->alg_bytes;
    s->merge_bytes = p->merge_bytes;
    s->staged_bytes = p->opts.method == RTG_TC_BATCHED ? p->staged_bytes : 0;
    s->edges = p->e_end - p->e_begin;
    if (p->opts.method == RTG_TC_BATCHED) {
        // owner buckets: small (wave) owners, block owners up to 1024, larger block owners whose
        // hash table fits, owners too long for a hash table
        s->nbuckets = kBuckets;
        for (int b = 0; b < kBuckets; ++b) {
            s->strategy[b] = p->strat[b];
            s->bucket_items[b] = p->bucket_items[b];
        }
    }
    s->kernel_launches = 2;
}

extern "C" {

RTG_API int rtg_tc_plan_create(rtg_ctx *ctx, const rtg_graph *g, const rtg_tc_opts *opts, rtg_tc_plan **out,
                               rtg_tc_stats *stats) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!g || !opts || !out) return rtg_fail(RTG_EINVAL, "rtg_tc_plan_create: null argument");
    if (!g->oriented) return rtg_fail(RTG_EINVAL, "rtg_tc_plan_create: graph is not degree-oriented");
    *out = nullptr;
    std::lock_guard<std::mutex> lock(ctx->mu);
    RTG_TRY(rtg_tc_plan_create_impl(ctx, g, opts, out));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        fill_stats(*out, stats);
    }
    return RTG_OK;
}

RTG_API int rtg_tc_plan_run(rtg_ctx *ctx, rtg_tc_plan *p, uint64_t *count, rtg_tc_stats *stats) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!p || !count) return rtg_fail(RTG_EINVAL, "rtg_tc_plan_run: null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t st = ctx->stream;
    HIP_TRY(hipEventRecord(ctx->ev[0], st));
    if (p->exec) HIP_TRY(hipGraphLaunch(p->exec, st));
    else RTG_TRY(launch_count(ctx, p));
    HIP_TRY(hipEventRecord(ctx->ev[1], st));
    unsigned long long c = 0;
    HIP_TRY(hipMemcpyAsync(&c, p->d_count, sizeof(c), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *count = c;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        fill_stats(p, stats);
        stats->t_count_ms = rtg::elapsed_ms(ctx->ev[0], ctx->ev[1]);
    }
    return RTG_OK;
}

RTG_API int rtg_tc_plan_run_per_edge(rtg_ctx *ctx, rtg_tc_plan *p, uint32_t *d_out, uint64_t *count,
                                     rtg_tc_stats *stats) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!p || !d_out) return rtg_fail(RTG_EINVAL, "rtg_tc_plan_run_per_edge: null argument");
    if (p->opts.method != RTG_TC_BATCHED || !p->opts.per_edge)
        return rtg_fail(RTG_EINVAL, "rtg_tc_plan_run_per_edge: the plan was not created as a batched per-edge plan");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t st = ctx->stream;
    HIP_TRY(hipEventRecord(ctx->ev[0], st));
    RTG_TRY(launch_per_edge(ctx, p, d_out));
    HIP_TRY(hipEventRecord(ctx->ev[1], st));
    unsigned long long c = 0;
    HIP_TRY(hipMemcpyAsync(&c, p->d_count, sizeof(c), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (count) *count = c;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        fill_stats(p, stats);
        stats->t_count_ms = rtg::elapsed_ms(ctx->ev[0], ctx->ev[1]);
    }
    return RTG_OK;
}

RTG_API int rtg_tc_plan_describe(const rtg_tc_plan *p, char *buf, size_t len) {
    if (!p || !buf || len == 0) return rtg_fail(RTG_EINVAL, "rtg_tc_plan_describe: bad arguments");
    std::string d = "method=" + std::to_string(p->opts.method);
    if (p->opts.method == RTG_TC_BATCHED)
        d += "," + p->tune + ",grid=" + std::to_string(p->grid) + ",table_slots=" + std::to_string(p->max_table) +
             ",wave_owner_max=" + std::to_string(p->wave_owner_max) + ",source=" +
             (p->autotuned ? "autotune" : (p->tune_from_cache ? "cache" : (std::getenv("RTG_TC_TUNE") ? "env" : "default"))) +
             ",search_items=" + std::to_string(p->n_search_items) + ",rev=" + std::to_string(kTuneRevision) +
             ",buckets=" + std::to_string(p->bucket_items[0]) + "/" + std::to_string(p->bucket_items[1]) + "/" +
             std::to_string(p->bucket_items[2]) + "/" + std::to_string(p->bucket_items[3]) + ",bucket_strategy=" +
             kStratName[p->strat[0]] + "/" + kStratName[p->strat[1]] + "/" + kStratName[p->strat[2]] + "/" +
             kStratName[p->strat[3]] + ",all_strategies=" + (p->all_strategies ? "1" : "0") +
             ",plan_graph_ms=" + std::to_string(p->t_plan_graph_ms) + ",plan_tune_ms=" + std::to_string(p->t_plan_tune_ms) +
             ",plan_items_ms=" + std::to_string(p->t_plan_items_ms) + ",renumbered=" + (p->renumbered ? "1" : "0") +
             ",renumber_ms=" + std::to_string(p->t_renum_ms) + ",streamed=" + std::to_string(p->streamed) + ",streamed16=" + std::to_string(p->streamed16) +
             ",bitmap_items=" + std::to_string(p->n_bitmap_items) + ",grid_bitmap=" + std::to_string(p->grid_bm) +
             ",wave_items=" + std::to_string(p->n_wave_items) + ",grid_wave=" + std::to_string(p->grid_wv) +
             ",small_work=" + std::to_string(p->small_work) + ",wave_item_used=" + std::to_string(p->wave_item_used) +
             ",side=" + (p->side ? "1" : "0") +
             ",side_blocks=" + std::to_string(p->side_blocks);
    std::snprintf(buf, len, "%s", d.c_str());
    return RTG_OK;
}

RTG_API int rtg_tc_plan_destroy(rtg_tc_plan *p) {
    if (!p) return RTG_OK;
    if (p->ctx) (void)hipSetDevice(p->ctx->device);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    void *bufs[] = {p->d_row, p->d_wofs, p->d_wspan, p->d_epos, p->d_ecnt, p->d_small, p->d_items, p->d_count, p->d_work};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete p;
    return RTG_OK;
}

RTG_API int rtg_tc_per_edge(rtg_ctx *ctx, const rtg_graph *g, int32_t form, uint32_t *d_out) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!g || (!d_out && g->m > 0)) return rtg_fail(RTG_EINVAL, "rtg_tc_per_edge: null argument");
    if (!g->oriented) return rtg_fail(RTG_EINVAL, "rtg_tc_per_edge: graph is not degree-oriented");
    if (form < 0 || form > 2) return rtg_fail(RTG_EINVAL, "rtg_tc_per_edge: form must be 0, 1 or 2");
    if (g->m == 0) return RTG_OK;
    if (form == 0) {  // a one-off batched per-edge plan (default configuration, no capture)
        rtg_tc_opts o{};
        o.method = RTG_TC_BATCHED;
        o.per_edge = 1;
        rtg_tc_plan *p = nullptr;
        RTG_TRY(rtg_tc_plan_create(ctx, g, &o, &p, nullptr));
        const int rc = rtg_tc_plan_run_per_edge(ctx, p, d_out, nullptr, nullptr);
        rtg_tc_plan_destroy(p);
        return rc;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t st = ctx->stream;
    int32_t *row = nullptr;
    RTG_TRY(dev_alloc(&row, (size_t)g->m));
    DevFree fr;
    fr.p = row;
    edge_rows_kernel<<<grid_cap(ctx, g->n, 256, 16), 256, 0, st>>>(g->d_offsets, g->n, row);
    if (form == 2)
        tc_per_edge_intersect<<<grid_cap(ctx, g->m * 64, kPlainThreads, 32), kPlainThreads, 0, st>>>(
            g->d_offsets, g->d_adjs, row, g->m, d_out);
    else
        tc_per_edge_1a2<<<grid_cap(ctx, g->m, kPlainThreads, 32), kPlainThreads, 0, st>>>(
            g->d_offsets, g->d_adjs, row, g->m, d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return RTG_OK;
}

RTG_API int rtg_tc_count(rtg_ctx *ctx, const rtg_graph *g, const rtg_tc_opts *opts, uint64_t *count,
                         rtg_tc_stats *stats) {
    rtg_tc_opts def{};
    def.method = RTG_TC_BATCHED;
    if (!opts) opts = &def;
    if (opts->method == RTG_TC_2A1) {
        RTG_TRY(rtg::ctx_enter(ctx));
        if (!g || !count) return rtg_fail(RTG_EINVAL, "rtg_tc_count: null argument");
        std::lock_guard<std::mutex> lock(ctx->mu);
        return rtg_tc_count_2a1_impl(ctx, g, count, stats);
    }
    rtg_tc_plan *p = nullptr;
    rtg_tc_stats pst{};
    RTG_TRY(rtg_tc_plan_create(ctx, g, opts, &p, &pst));
    int rc = rtg_tc_plan_run(ctx, p, count, stats);
    if (rc == RTG_OK && stats) stats->t_plan_ms = pst.t_plan_ms;
    rtg_tc_plan_destroy(p);
    return rc;
}

}  // extern "C"

namespace rtg {
// Loads this file's kernels onto the current device (an attribute query of one of them).
int preload_tc() {
    hipFuncAttributes attr;
    return hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&edge_rows_kernel)) == hipSuccess ? RTG_OK : RTG_EHIP;
}
}  // namespace rtg
