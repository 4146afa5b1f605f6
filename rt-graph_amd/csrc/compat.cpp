// The reference's entry points with their own signatures (declared in include/rtg_compat.h),
// implemented on the C ABI of rtg.h.
#include "../../include/rtg_compat.h"

#include "rtg_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <vector>

namespace {

[[noreturn]] void die(const char *where) {
    std::printf("RTG error: %s: %s\n", where, rtg_last_error());
    std::exit(2);
}

struct Ctx {
    rtg_ctx *c = nullptr;
    ~Ctx() {
        if (c) rtg_destroy(c);
    }
};

struct Dev {
    rtg_ctx *c;
    void *p = nullptr;
    explicit Dev(rtg_ctx *ctx) : c(ctx) {}
    ~Dev() {
        if (p) rtg_free(c, p);
    }
};

}  // namespace

extern "C" {

RTG_API int rtg_count_on_gpu(const int *adjs, const int *offsets, int nodes, int64_t edges, int device_id,
                             int method, uint64_t *count, double *ms) {
    if (!offsets || (edges > 0 && !adjs) || nodes < 0 || !count) return rtg_fail(RTG_EINVAL, "rtg_count_on_gpu: bad arguments");
    Ctx ctx;
    int rc = rtg_create(device_id, &ctx.c);
    if (rc != RTG_OK) return rc;
    rtg_graph *g = nullptr;
    rc = rtg_graph_from_csr(ctx.c, offsets, adjs, nodes, edges, &g);
    if (rc != RTG_OK) return rc;
    rtg_tc_opts opts{};
    opts.method = method;
    rtg_tc_stats st{};
    rc = rtg_tc_count(ctx.c, g, &opts, count, &st);
    rtg_graph_destroy(g);
    if (ms) *ms = st.t_count_ms;
    return rc;
}

RTG_API void CountOnGPU(int *adjs, int *offsets, int nodes, int edges, int device_id) {
    uint64_t count = 0;
    double ms = 0.0;
    if (rtg_count_on_gpu(adjs, offsets, nodes, edges, device_id, RTG_TC_BATCHED, &count, &ms) != RTG_OK)
        die("CountOnGPU");
    std::printf("BSTC: counting time = %f ms, counting result = %u\n", ms, (unsigned)(count & 0xFFFFFFFFu));
    if (count >> 32) std::printf("BSTC: exact 64-bit count = %llu\n", (unsigned long long)count);
}

RTG_API void BSAccessCountFunction(int *data, int data_size, int *query, int query_size, int *bs_results) {
    if (data_size < 0 || query_size < 0 || (data_size > 0 && !data) || (query_size > 0 && (!query || !bs_results))) {
        rtg_fail(RTG_EINVAL, "bad arguments");
        die("BSAccessCountFunction");
    }
    Ctx ctx;
    if (rtg_create(0, &ctx.c) != RTG_OK) die("BSAccessCountFunction");
    Dev dd(ctx.c), dq(ctx.c), dr(ctx.c);
    const size_t nb = sizeof(int) * (size_t)std::max(data_size, 1), qb = sizeof(int) * (size_t)std::max(query_size, 1);
    // the unsorted array goes to the device first; the host copy is sorted separately (:85-95)
    if (rtg_malloc(ctx.c, nb, &dd.p) || rtg_malloc(ctx.c, qb, &dq.p) || rtg_malloc(ctx.c, qb, &dr.p) ||
        rtg_memcpy_h2d(ctx.c, dd.p, data, sizeof(int) * (size_t)data_size) ||
        rtg_memcpy_h2d(ctx.c, dq.p, query, sizeof(int) * (size_t)query_size))
        die("BSAccessCountFunction");
    std::printf("data size = %d\n", data_size);
    const std::clock_t begin = std::clock();
    std::sort(data, data + data_size);
    const std::clock_t end = std::clock();
    std::printf("cpu sort time = %.3f ms\n", double(end - begin) / CLOCKS_PER_SEC * 1000.0);
    double sort_ms = 0.0, bs_ms = 0.0;
    if (rtg_sort_i32(ctx.c, (int32_t *)dd.p, data_size, &sort_ms) ||
        rtg_lookup_i32(ctx.c, (const int32_t *)dd.p, data_size, (const int32_t *)dq.p, query_size, (int32_t *)dr.p, &bs_ms) ||
        rtg_memcpy_d2h(ctx.c, bs_results, dr.p, sizeof(int) * (size_t)query_size))
        die("BSAccessCountFunction");
    std::printf("Sorting time = %f ms, Binary search time = %f ms, Total time = %f ms\n", sort_ms, bs_ms, sort_ms + bs_ms);
}

RTG_API int rtg_set_intersection_tile_host(const uint32_t *set_offsets, const uint32_t *elements, uint32_t num_a,
                                           uint32_t num_b, int32_t method, int device_id, uint32_t *results,
                                           double *ms) {
    if (!set_offsets || !results) return rtg_fail(RTG_EINVAL, "rtg_set_intersection_tile_host: null argument");
    Ctx ctx;
    int rc = rtg_create(device_id, &ctx.c);
    if (rc != RTG_OK) return rc;
    const uint64_t nsets = (uint64_t)num_a + num_b;
    const uint64_t nel = set_offsets[nsets];
    Dev doff(ctx.c), del(ctx.c), dout(ctx.c);
    const size_t ob = sizeof(uint32_t) * (nsets + 1), eb = sizeof(uint32_t) * std::max<uint64_t>(nel, 1),
                 rb = sizeof(uint32_t) * std::max<uint64_t>((uint64_t)num_a * num_b, 1);
    if ((rc = rtg_malloc(ctx.c, ob, &doff.p)) || (rc = rtg_malloc(ctx.c, eb, &del.p)) || (rc = rtg_malloc(ctx.c, rb, &dout.p)) ||
        (rc = rtg_memcpy_h2d(ctx.c, doff.p, set_offsets, ob)) ||
        (nel && (rc = rtg_memcpy_h2d(ctx.c, del.p, elements, sizeof(uint32_t) * nel))))
        return rc;
    rtg_si_stats st{};
    if ((rc = rtg_si_tile(ctx.c, (const uint32_t *)doff.p, (const uint32_t *)del.p, num_a, num_b, method,
                          (uint32_t *)dout.p, &st)))
        return rc;
    if ((uint64_t)num_a * num_b && (rc = rtg_memcpy_d2h(ctx.c, results, dout.p, sizeof(uint32_t) * (uint64_t)num_a * num_b)))
        return rc;
    if (ms) *ms = st.t_kernel_ms;
    return RTG_OK;
}

RTG_API int rtg_binary_search_based_intersection(const uint32_t *set_offsets, const uint32_t *elements, uint32_t num_a,
                                                 uint32_t num_b, uint32_t block_size, int device_id, uint32_t *results,
                                                 double *ms) {
    (void)block_size;
    return rtg_set_intersection_tile_host(set_offsets, elements, num_a, num_b, RTG_SI_BATCHED, device_id, results, ms);
}

}  // extern "C"
