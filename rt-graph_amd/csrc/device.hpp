// Device-side building blocks shared by the gfx950 kernels: HIP status checking that reports
// through the C ABI, wave64 reductions and scans, and the context structure.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "rtg_internal.hpp"

#define RTG_WAVE 64

// Evaluates a HIP call; on failure records the error and returns RTG_EHIP from the caller.
#define HIP_TRY(call)                                                                       \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess)                                                               \
            return rtg::set_error(RTG_EHIP, std::string(#call) + ": " + hipGetErrorString(_e) + \
                                                " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

// Same for functions whose status is an int RTG_* code.
#define RTG_TRY(call)                \
    do {                             \
        int _s = (call);             \
        if (_s != RTG_OK) return _s; \
    } while (0)

struct rtg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    int cu_count = 0;
    std::mutex mu;  // serialises entry points that share this context's stream and scratch
    // Grow-only scratch owned by the context (released by rtg_destroy).
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    hipEvent_t ev[8] = {};
    // A second stream (default priority, like the first) for kernels that run beside the main
    // stream's, forked and joined with these events (also under stream capture).
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // Long-pair strategy the set-intersection autotuner chose, per argument set (most recent
    // last). Every strategy returns the same counts; a stale entry only costs speed.
    struct SiTuned {
        const void *off, *el, *pairs;
        int64_t npairs;
        uint32_t num_a, num_b;
        int kind;
    };
    std::vector<SiTuned> si_tuned;
};

namespace rtg {

// Makes the context's device current on this thread.
int ctx_enter(rtg_ctx *ctx);

// One per HIP source file: loads that file's kernels onto the current device (HIP loads a code
// object on the first launch of any of its kernels, a few milliseconds for the largest), so
// that rtg_create pays it once instead of the first build, plan or query.
int preload_csr_build();
int preload_gen_device();
int preload_lookup();
int preload_radix_sort();
int preload_renumber();
int preload_scan();
int preload_serve();
int preload_tc();
int preload_tc_2a1();

// Returns at least `bytes` of device scratch owned by the context (contents undefined).
int ctx_scratch(rtg_ctx *ctx, size_t bytes, void **out);

// Looks up, once per device, the device address of the set-intersection scratch array used by
// rtg_si_pairs_noalloc (si.hip). Called by rtg_create with `device` current; a failure is
// recorded and reported by that entry point.
int si_noalloc_init(int device);

// Carves aligned sub-buffers out of one allocation.
struct Arena {
    char *base = nullptr;
    size_t used = 0;
    size_t cap = 0;
    template <class T>
    T *take(size_t count) {
        size_t off = (used + 255) & ~(size_t)255;
        used = off + count * sizeof(T);
        return reinterpret_cast<T *>(base + off);
    }
    static size_t round(size_t bytes) { return (bytes + 255) & ~(size_t)255; }
};

// Milliseconds between two recorded events.
inline float elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) ms = 0.f;
    return ms;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- device-wide primitives (scan.hip, radix_sort.hip) ----
// Exclusive scan of n values into out (may alias in); the total goes to *d_total (device
// pointer) when non-null. `work` holds scan_work_bytes(n) bytes of device scratch.
size_t scan_work_bytes(int64_t n);
int scan_exclusive_u32_w(rtg_ctx *ctx, const uint32_t *in, uint32_t *out, int64_t n,
                         uint32_t *d_total, void *work);
int scan_exclusive_u64_w(rtg_ctx *ctx, const uint64_t *in, uint64_t *out, int64_t n,
                         uint64_t *d_total, void *work);
// Stable LSD radix sort of n uint64 keys on bits [0, bits). Input in keys[0], the result ends
// in keys[*which] (0 or 1); keys[1] is a same-size buffer. `work` holds
// radix_sort_work_bytes(n) bytes.
size_t radix_sort_work_bytes(int64_t n);
int radix_sort_u64(rtg_ctx *ctx, uint64_t *keys[2], int64_t n, int bits, int *which,
                   int *passes, void *work);
// The same for 32-bit keys, all 32 bits; is_signed orders them as int32 (the bit patterns are
// moved, not converted). Same work size.
int radix_sort_32(rtg_ctx *ctx, uint32_t *keys[2], int64_t n, bool is_signed, int *which,
                  int *passes, void *work);

}  // namespace rtg

// ---- wave64 helpers (device) ----
#if defined(__HIPCC__)
namespace rtgd {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The same sum through DPP lane moves (VALU only, no LDS permutes), returned uniform: row_shr 1,
// 2, 4, 8 leave each row of 16 lanes its sum in its lane 15 (lanes shifted in from outside the
// row read 0), row_bcast 15 adds lane 15's value to row 1 and lane 47's to row 3, row_bcast 31
// lane 31's to rows 2 and 3, so lane 63 ends with the wave's sum.
__device__ __forceinline__ uint32_t wave_sum_dpp(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Inclusive scan across the 64 lanes.
template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
    const int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

// Wave-uniform copies (scalar registers) of a value every lane holds identically. Work-queue
// loops take their exit decision on these, so the compiler sees a uniform loop: a loop whose
// exit it believes divergent can be restructured so that a lane-0-only queue grab at the top
// lands outside the loop the other lanes keep iterating, which then never sees a new item.
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Orders a wave's LDS writes before its later LDS reads (wave-private LDS regions).
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace rtgd
#endif
