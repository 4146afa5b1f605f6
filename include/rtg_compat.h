/*
 * rtg_compat.h — drop-in replacements, with the reference's own signatures, for the SIMT
 * entry points of the path (the count, the lookup and the set-intersection baselines). A reference driver links librtg instead of compiling the
 * CUDA files and includes this header in place of the reference header named on each line.
 *
 * Error behaviour matches the reference: a device or runtime failure prints a message and
 * calls exit(2) (the reference's CUDA_CHECK, tc/utils/common.h:13-21). Callers that need a
 * status code instead use the rtg_* functions of rtg.h.
 */
#ifndef RTG_COMPAT_H
#define RTG_COMPAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces tc/bs_tc/bs_tc.h:5 (defined tc/bs_tc/bs_tc.cu:232-354): copies the host CSR to
   device `device_id`, counts triangles with the batched kernel and prints
   "BSTC: counting time = %f ms, counting result = %u" (bs_tc.cu:353). The reference's counter
   is a 32-bit uint; %u prints the low 32 bits, and a second line gives the exact 64-bit count
   when it does not fit. */
#ifdef __cplusplus
void CountOnGPU(int *adjs, int *offsets, int nodes, int edges, int device_id = 0);
#else
void CountOnGPU(int *adjs, int *offsets, int nodes, int edges, int device_id);
#endif

/* Same count with a status code and no printing: *count is the exact uint64 triangle count
   and *ms the device time of the counting kernels. */
int rtg_count_on_gpu(const int *adjs, const int *offsets, int nodes, int64_t edges, int device_id,
                     int method, uint64_t *count, double *ms);

/* Replaces rt_search/simple_bs.h:5 (rt_search/simple_bs.cu:80-142): copies `data` and `query`
   to the device, sorts `data` on the host in place (as the reference does at :92) and on the
   device (:99-108), looks every query up on the device and writes the index of the first
   equal element, or -1, to bs_results. Prints the reference's lines "data size = %d",
   "cpu sort time = %.3f ms" and "Sorting time = %f ms, Binary search time = %f ms, Total
   time = %f ms" (:88, :95, :137-139); the times are device times of the sort and the lookup. */
void BSAccessCountFunction(int *data, int data_size, int *query, int query_size, int *bs_results);

/* C form of BinarySearchBasedIntersection (set_intersection/baseline/bs.cu:91-135): sets
   [0, num_a) against [num_a, num_a + num_b) of a set_intersection dataset (util.h:19-28);
   results[a * num_b + b] = |A_a & B_b|. */
int rtg_binary_search_based_intersection(const uint32_t *set_offsets, const uint32_t *elements,
                                         uint32_t num_a, uint32_t num_b, uint32_t block_size,
                                         int device_id, uint32_t *results, double *ms);

/* The same tile with any set-intersection method of rtg.h (RTG_SI_*): the C form behind the
   other baselines' replacements below (hi.cu, ip.cu, bi_naive.cu, bi_dynamic.cu). */
int rtg_set_intersection_tile_host(const uint32_t *set_offsets, const uint32_t *elements,
                                   uint32_t num_a, uint32_t num_b, int32_t method, int device_id,
                                   uint32_t *results, double *ms);

#ifdef __cplusplus
}

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rtg.h"

/* BinarySearchBasedIntersection(const Dataset&, uint blocks, uint block_size,
   std::vector<uint>& results) (bs.cu:21-22), for the reference's Dataset type or any type
   with its field names. */
template <class DatasetT>
inline void BinarySearchBasedIntersection(const DatasetT &dataset, unsigned blocks, unsigned block_size,
                                          std::vector<unsigned> &results) {
    (void)blocks;
    results.resize((size_t)dataset.num_of_sets_a * dataset.num_of_sets_b);
    double ms = 0.0;
    if (rtg_binary_search_based_intersection(dataset.set_offsets.data(), dataset.elements.data(),
                                             dataset.num_of_sets_a, dataset.num_of_sets_b, block_size, 0,
                                             results.data(), &ms) != 0) {
        std::fprintf(stderr, "BinarySearchBasedIntersection failed\n");
        std::exit(2);
    }
    std::printf("Binary Search based Set Intertsection: kernel time = %f ms\n", ms);
}

namespace rtg_compat_detail {
template <class DatasetT>
inline double tile(const DatasetT &dataset, int method, std::vector<unsigned> &results, const char *name) {
    results.resize((size_t)dataset.num_of_sets_a * dataset.num_of_sets_b);
    double ms = 0.0;
    if (rtg_set_intersection_tile_host(dataset.set_offsets.data(), dataset.elements.data(), dataset.num_of_sets_a,
                                       dataset.num_of_sets_b, method, 0, results.data(), &ms) != 0) {
        std::fprintf(stderr, "%s failed\n", name);
        std::exit(2);
    }
    return ms;
}
}  // namespace rtg_compat_detail

/* The other set-intersection baselines with their signatures and report lines; each runs the
   corresponding method of rtg.h (the launch-shape and bucket arguments are accepted and the
   shapes chosen by the library). */
/* HashBasedIntersection, set_intersection/baseline/hi.cu:22-23 -> RTG_SI_HASH */
template <class DatasetT>
inline void HashBasedIntersection(const DatasetT &dataset, unsigned blocks, unsigned block_size, unsigned buckets,
                                  unsigned max_bucket_size, std::vector<unsigned> &results) {
    (void)blocks, (void)block_size, (void)buckets, (void)max_bucket_size;
    const double ms = rtg_compat_detail::tile(dataset, RTG_SI_HASH, results, "HashBasedIntersection");
    std::printf("Hash based Set Intertsection: kernel time = %f ms\n", ms);
}
/* IntersectPathBasedIntersection, ip.cu:32-33 -> RTG_SI_MERGE */
template <class DatasetT>
inline void IntersectPathBasedIntersection(const DatasetT &dataset, unsigned blocks, unsigned block_size,
                                           std::vector<unsigned> &results) {
    (void)blocks, (void)block_size;
    const double ms = rtg_compat_detail::tile(dataset, RTG_SI_MERGE, results, "IntersectPathBasedIntersection");
    std::printf("Intersect Path based Set Intertsection: kernel time = %f ms\n", ms);
}
/* BitmapNaiveBasedIntersection, bi_naive.cu:26-27 -> RTG_SI_BITMAP (its bitmaps are built
   inside the counting launch: the construction time is reported as 0) */
template <class DatasetT>
inline void BitmapNaiveBasedIntersection(const DatasetT &dataset, unsigned blocks, unsigned block_size,
                                         std::vector<unsigned> &results) {
    (void)blocks, (void)block_size;
    const double ms = rtg_compat_detail::tile(dataset, RTG_SI_BITMAP, results, "BitmapNaiveBasedIntersection");
    std::printf("Bitmap(naive) based Set Intertsection:\nconstruction time = %f ms, counting time = %f ms , kernel "
                "time = %f ms\n", 0.0, ms, ms);
}
/* BitmapDynamicBasedIntersection, bi_dynamic.cu:24-25 -> RTG_SI_BITMAP */
template <class DatasetT>
inline void BitmapDynamicBasedIntersection(const DatasetT &dataset, unsigned blocks, unsigned block_size,
                                           std::vector<unsigned> &results) {
    (void)blocks, (void)block_size;
    const double ms = rtg_compat_detail::tile(dataset, RTG_SI_BITMAP, results, "BitmapDynamicBasedIntersection");
    std::printf("Bitmap(dynamic) based Set Intertsection: kernel time = %f ms\n", ms);
}
/* ThreadBinarySearch, bs.cu:25-26 (a thread per element of the A sets) -> RTG_SI_BATCHED */
template <class DatasetT>
inline void ThreadBinarySearch(const DatasetT &dataset, unsigned blocks, unsigned block_size,
                               std::vector<unsigned> &results) {
    (void)blocks, (void)block_size;
    std::printf("-----> thread binary search\n");
    const double ms = rtg_compat_detail::tile(dataset, RTG_SI_BATCHED, results, "ThreadBinarySearch");
    std::printf("Binary Search based Set Intertsection: kernel time = %f ms\n", ms);
}
#endif

#endif /* RTG_COMPAT_H */
