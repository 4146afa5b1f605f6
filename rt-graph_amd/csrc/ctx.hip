// Context, errors and device-memory entry points of the C ABI.
#include "device.hpp"

#include <cstdlib>
#include <cstring>

namespace {
thread_local std::string g_last_error;
}

namespace rtg {

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int host_threads() {
    static int n = [] {
        const char *s = std::getenv("RTG_HOST_THREADS");
        int v = s ? std::atoi(s) : 0;
        if (v <= 0) v = (int)std::thread::hardware_concurrency();
        return v > 0 ? v : 1;
    }();
    return n;
}

int ctx_enter(rtg_ctx *ctx) {
    if (!ctx) return set_error(RTG_EINVAL, "null rtg_ctx");
    HIP_TRY(hipSetDevice(ctx->device));
    return RTG_OK;
}

int ctx_scratch(rtg_ctx *ctx, size_t bytes, void **out) {
    if (bytes == 0) bytes = 256;
    if (ctx->scratch_bytes < bytes) {
        if (ctx->scratch) {
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            HIP_TRY(hipFree(ctx->scratch));
            ctx->scratch = nullptr;
            ctx->scratch_bytes = 0;
        }
        size_t want = bytes + bytes / 8;
        if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
            (void)hipGetLastError();
            ctx->scratch = nullptr;
            return set_error(RTG_ENOMEM, "scratch allocation of " + std::to_string(want) + " bytes failed");
        }
        ctx->scratch_bytes = want;
    }
    *out = ctx->scratch;
    return RTG_OK;
}

}  // namespace rtg

extern "C" {

RTG_API const char *rtg_last_error(void) { return g_last_error.c_str(); }

RTG_API int rtg_version(void) { return 1; }

RTG_API int rtg_create(int device, rtg_ctx **out) {
    if (!out) return rtg_fail(RTG_EINVAL, "rtg_create: null out");
    *out = nullptr;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return rtg_fail(RTG_EINVAL, "rtg_create: device " + std::to_string(device) + " out of range (" +
                                        std::to_string(count) + " visible)");
    auto *ctx = new rtg_ctx();
    ctx->device = device;
    int st = RTG_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { st = rtg_fail(RTG_EHIP, "hipSetDevice failed"); break; }
        if (hipGetDeviceProperties(&ctx->prop, device) != hipSuccess) { st = rtg_fail(RTG_EHIP, "hipGetDeviceProperties failed"); break; }
        ctx->cu_count = ctx->prop.multiProcessorCount;
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { st = rtg_fail(RTG_EHIP, "hipStreamCreate failed"); break; }
        if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) { st = rtg_fail(RTG_EHIP, "hipStreamCreate failed"); break; }
        if (hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming) != hipSuccess) { st = rtg_fail(RTG_EHIP, "hipEventCreate failed"); break; }
        for (auto &e : ctx->ev)
            if (hipEventCreate(&e) != hipSuccess) { st = rtg_fail(RTG_EHIP, "hipEventCreate failed"); break; }
        if (st != RTG_OK) break;
        (void)rtg::si_noalloc_init(device);  // status kept per device; rtg_si_pairs_noalloc reports it
        // every other source file's kernels (si.hip's are loaded by the symbol query above)
        for (int (*load)() : {rtg::preload_csr_build, rtg::preload_gen_device, rtg::preload_lookup, rtg::preload_radix_sort,
                              rtg::preload_renumber, rtg::preload_scan, rtg::preload_serve, rtg::preload_tc,
                              rtg::preload_tc_2a1})
            if (load() != RTG_OK) {
                st = rtg_fail(RTG_EHIP, "rtg_create: loading the kernels failed");
                break;
            }
    } while (0);
    if (st != RTG_OK) {
        rtg_destroy(ctx);
        return st;
    }
    *out = ctx;
    return RTG_OK;
}

RTG_API int rtg_destroy(rtg_ctx *ctx) {
    if (!ctx) return RTG_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->join) (void)hipEventDestroy(ctx->join);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RTG_OK;
}

RTG_API int rtg_device_query(rtg_ctx *ctx, rtg_device_info *out) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!out) return rtg_fail(RTG_EINVAL, "rtg_device_query: null out");
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, ctx->prop.name, sizeof(out->name) - 1);
    std::strncpy(out->arch, ctx->prop.gcnArchName, sizeof(out->arch) - 1);
    out->cu_count = ctx->cu_count;
    out->xcd_count = 8;
    out->lds_per_block = (int64_t)ctx->prop.sharedMemPerBlock;
    out->hbm_bytes = (int64_t)ctx->prop.totalGlobalMem;
    out->clock_khz = ctx->prop.clockRate;
    out->mem_clock_khz = ctx->prop.memoryClockRate;
    return RTG_OK;
}

RTG_API int rtg_synchronize(rtg_ctx *ctx) {
    RTG_TRY(rtg::ctx_enter(ctx));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTG_OK;
}

RTG_API int rtg_malloc(rtg_ctx *ctx, size_t bytes, void **d_ptr) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!d_ptr) return rtg_fail(RTG_EINVAL, "rtg_malloc: null out");
    *d_ptr = nullptr;
    if (hipMalloc(d_ptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        *d_ptr = nullptr;
        return rtg_fail(RTG_ENOMEM, "rtg_malloc: " + std::to_string(bytes) + " bytes failed");
    }
    return RTG_OK;
}

RTG_API int rtg_free(rtg_ctx *ctx, void *d_ptr) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!d_ptr) return RTG_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipFree(d_ptr));
    return RTG_OK;
}

RTG_API int rtg_memcpy_h2d(rtg_ctx *ctx, void *d_dst, const void *src, size_t bytes) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!bytes) return RTG_OK;
    HIP_TRY(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTG_OK;
}

RTG_API int rtg_memcpy_d2h(rtg_ctx *ctx, void *dst, const void *d_src, size_t bytes) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!bytes) return RTG_OK;
    HIP_TRY(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTG_OK;
}

RTG_API int rtg_memcpy_d2d(rtg_ctx *ctx, void *d_dst, const void *d_src, size_t bytes) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!bytes) return RTG_OK;
    HIP_TRY(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTG_OK;
}

RTG_API int rtg_memset(rtg_ctx *ctx, void *d_dst, int value, size_t bytes) {
    RTG_TRY(rtg::ctx_enter(ctx));
    if (!bytes) return RTG_OK;
    HIP_TRY(hipMemsetAsync(d_dst, value, bytes, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RTG_OK;
}

}  // extern "C"
