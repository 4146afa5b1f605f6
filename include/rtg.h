/*
 * rtg.h — C ABI of librtg, the MI355X (gfx950) triangle-counting, sorted-set-intersection and
 * sorted-lookup engine.
 *
 * Conventions: every entry point returns an int status (RTG_OK = 0) and never throws; on error
 * rtg_last_error() returns a message for the calling thread. Pointers named d_* are device
 * (HBM) pointers, obtained from rtg_malloc or from another HIP allocator on the context's
 * device; all other pointers are host pointers. Compute entry points run on the context's
 * stream and return after the result is complete.
 *
 * The reference exposes no C ABI; its boundary for this path is the C++ function
 *   void CountOnGPU(int *adjs, int *offsets, int nodes, int edges, int device_id = 0)
 * (tc/bs_tc/bs_tc.h:5, defined tc/bs_tc/bs_tc.cu:232-354) with its siblings
 * BinarySearchBasedIntersection (set_intersection/baseline/bs.cu:21-22) and
 * BSAccessCountFunction (rt_search/simple_bs.h:5). Drop-in replacements with exactly those
 * signatures are declared in rtg_compat.h; this header is the engine underneath them.
 */
#ifndef RTG_H
#define RTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------------------- */
#define RTG_OK 0
#define RTG_EINVAL 1      /* bad argument or precondition (replaces the reference's exit())  */
#define RTG_EHIP 2        /* HIP runtime error (the reference's CUDA_CHECK exit(2) case,      */
                          /* tc/utils/common.h:13-21)                                         */
#define RTG_ENOMEM 3      /* device allocation failed or request exceeds a memory budget      */
#define RTG_EIO 4         /* file could not be opened or parsed (tc/graph/graph.cpp:33-36)    */
#define RTG_EUNSUPPORTED 5
#define RTG_EAGAIN 6      /* server: no free slot / not finished yet                          */
#define RTG_ECANCELED 7   /* server: request cancelled or timed out                           */

const char *rtg_last_error(void);
int rtg_version(void);

/* ---- context: device, stream, scratch -------------------------------------------------- */
typedef struct rtg_ctx rtg_ctx;

typedef struct rtg_device_info {
    char name[64];
    char arch[32];      /* gcnArchName, e.g. "gfx950:sramecc+:xnack-"                        */
    int32_t cu_count;
    int32_t xcd_count;  /* 8 on MI355X                                                       */
    int64_t lds_per_block;
    int64_t hbm_bytes;
    int32_t clock_khz;
    int32_t mem_clock_khz;
} rtg_device_info;

/* Replaces the reference's cudaSetDevice(device_id) inside CountOnGPU (bs_tc.cu:233). */
int rtg_create(int device, rtg_ctx **out);
int rtg_destroy(rtg_ctx *ctx);
int rtg_device_query(rtg_ctx *ctx, rtg_device_info *out);
int rtg_synchronize(rtg_ctx *ctx);

/* Device memory helpers, so hosts without a HIP toolchain (ctypes, cgo) can stage data. */
int rtg_malloc(rtg_ctx *ctx, size_t bytes, void **d_ptr);
int rtg_free(rtg_ctx *ctx, void *d_ptr);
int rtg_memcpy_h2d(rtg_ctx *ctx, void *d_dst, const void *src, size_t bytes);
int rtg_memcpy_d2h(rtg_ctx *ctx, void *dst, const void *d_src, size_t bytes);
int rtg_memcpy_d2d(rtg_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
int rtg_memset(rtg_ctx *ctx, void *d_dst, int value, size_t bytes);

/* ---- graphs: device-resident CSR --------------------------------------------------------- */
typedef struct rtg_graph rtg_graph;

/* Flags for rtg_graph_from_pairs. */
#define RTG_PAIRS_RELABEL 0x1u  /* dense order-preserving relabel of the ids present, with 0- or
                                   1-based detection (Graph::ReadBinaryEdges,
                                   tc/graph/graph.cpp:40-62, intended semantics)            */
#define RTG_PAIRS_ORIENT 0x2u   /* Graph::TCPreprocessing + ConvertToCSR
                                   (tc/graph/graph.cpp:104-177, 71-102): raw-degree
                                   orientation, degree<=1 removal, sort, dedup, CSR.
                                   Without it: symmetric simple CSR (both directions,
                                   sorted, deduplicated, self-loops dropped), for queries.  */
#define RTG_PAIRS_DEVICE 0x4u   /* `pairs` is a device pointer (else host)                  */

typedef struct rtg_build_stats {
    double t_total_ms;      /* device time of the whole build                                */
    double t_relabel_ms;
    double t_orient_ms;     /* degree, orientation, removal                                  */
    double t_sort_ms;       /* radix sort of the packed (src,dst) keys                       */
    double t_csr_ms;        /* dedup, self-loop removal, offsets                             */
    int64_t raw_pairs;      /* input pairs                                                   */
    int64_t valid_pairs;    /* after degree<=1 and self-loop removal (before dedup)          */
    int32_t nodes_in;       /* after relabel                                                 */
    int32_t sort_passes;
} rtg_build_stats;

int rtg_graph_from_pairs(rtg_ctx *ctx, const int32_t *pairs, int64_t npairs, uint32_t flags,
                         rtg_graph **out, rtg_build_stats *stats /* may be NULL */);
/* Copies a host CSR (offsets[n+1], adjs[m]); the caller keeps ownership. Rows must be sorted,
   unique and free of self-loops (the CountOnGPU precondition). Validated on the device.   */
int rtg_graph_from_csr(rtg_ctx *ctx, const int32_t *offsets, const int32_t *adjs, int32_t n,
                       int64_t m, rtg_graph **out);
int rtg_graph_info(const rtg_graph *g, int32_t *n, int64_t *m, int32_t *max_row);
int rtg_graph_download(const rtg_graph *g, int32_t *offsets /* n+1 */, int32_t *adjs /* m */);
/* The degree-ordered copy that batched counting plans over the whole graph count on (no
   counterpart in the reference, which counts on the CSR itself): vertices renamed in
   descending order of total degree (out + in, ties by id), rows renamed and sorted;
   inv[i] = the graph's id of copy vertex i. Built on first use and kept with the graph.
   Any output may be NULL. Non-empty oriented graphs only.                                 */
int rtg_graph_ordered_csr(rtg_ctx *ctx, const rtg_graph *g, int32_t *offsets /* n+1 */, int32_t *adjs /* m */,
                          int32_t *inv /* n */);
int rtg_graph_device_ptrs(const rtg_graph *g, const int32_t **d_offsets, const int32_t **d_adjs);
int rtg_graph_destroy(rtg_graph *g);

/* Preprocessed-CSR cache (checkpoint / resume of ingest + preprocessing; the reference keeps a
   text -> binary edge cache, tc/graph/graph.cpp:179-214). A file holds a header (magic,
   version, oriented flag, n, m, a caller-defined key, a checksum) and the CSR arrays; a read
   with another key, a bad checksum or a wrong size fails with RTG_EIO. `key` identifies the
   input, e.g. rtg_pairs_fingerprint(pairs, npairs, flags). Writes go to "<path>.part" and are
   renamed into place. rtg_csr_file_read uses the two-call convention (offsets == NULL: sizes
   only); both file calls are host-only.                                                      */
int rtg_pairs_fingerprint(const int32_t *pairs, int64_t npairs, int32_t flags, uint64_t *out);
int rtg_csr_file_write(const char *path, uint64_t key, int32_t oriented, const int32_t *offsets,
                       const int32_t *adjs, int32_t n, int64_t m);
int rtg_csr_file_read(const char *path, uint64_t key, int32_t *offsets, int32_t *adjs, int32_t *n,
                      int64_t *m, int32_t *oriented);
int rtg_graph_save(const rtg_graph *g, const char *path, uint64_t key);
int rtg_graph_load(rtg_ctx *ctx, const char *path, uint64_t key, rtg_graph **out);

/* Device-side synthetic inputs (same values as the host generators below). */
int rtg_gen_rmat_device(rtg_ctx *ctx, uint32_t scale, int64_t nsamples, double a, double b,
                        double c, uint64_t seed, int32_t *d_pairs /* 2*nsamples */);
int rtg_gen_uniform_u32_device(rtg_ctx *ctx, int64_t n, uint64_t seed, uint64_t stream,
                               uint32_t *d_out);
int rtg_gen_sorted_array_device(rtg_ctx *ctx, int64_t n, uint64_t seed, uint32_t *d_out);
/* deg is a host array (rtg_gen_powerlaw_degrees); d_pairs holds 2*(total/2 + total) ints. */
int rtg_gen_config_model_device(rtg_ctx *ctx, int32_t n, const int32_t *deg, int64_t total,
                                double p_close, uint64_t seed, int32_t *d_pairs, int64_t *npairs);

/* ---- triangle counting ----------------------------------------------------------------- */
#define RTG_TC_PLAIN_THREAD 0  /* thread per oriented edge, 1A2 form of ListIntersectionKernel
                                  (tc/bs_tc/bs_tc.cu:120-148): for w in N+(u), v in N+(w)?  */
#define RTG_TC_PLAIN_WAVE 1    /* wave64 per oriented edge: |N+(u) & N+(v)| by binary search
                                  of the shorter list in the longer                        */
#define RTG_TC_BATCHED 2       /* degree-bucketed persistent launch, per-bucket strategy      */
#define RTG_TC_2A1 3           /* two-hop multiset + membership (bs_tc.cu:153-229), memory
                                  bounded; for graphs whose wedge count fits the budget     */

typedef struct rtg_tc_opts {
    int32_t method;          /* RTG_TC_*                                                      */
    int32_t use_hip_graph;   /* batched: replay the captured launch sequence                  */
    int64_t e_begin;         /* oriented-edge range [e_begin, e_end) (e_end <= 0: all)         */
    int64_t e_end;
    const char *tune_cache;  /* batched: JSON tuning cache (key: device + graph fingerprint);
                                read if it has the key, written after autotuning             */
    int32_t autotune;        /* batched: time the candidate configurations when the cache has
                                no entry (RTG_TC_TUNE in the environment overrides both)     */
    int32_t per_edge;        /* batched: also record each span's edge, so that
                                rtg_tc_plan_run_per_edge can run (8 more bytes per edge)      */
} rtg_tc_opts;

#define RTG_TC_MAX_BUCKETS 8
typedef struct rtg_tc_stats {
    double t_count_ms;       /* device time of the counting kernels only (bs_tc.cu:313-322)   */
    double t_plan_ms;        /* planning / bucketing (device), reported separately           */
    uint64_t alg_bytes;      /* sum over edges of 24 + list_bytes(min,max) (algorithmic bytes)*/
    uint64_t merge_bytes;    /* 24*m + 4*sum over edges of (d+(u) + d+(v)) (merge-only model)  */
    int64_t edges;           /* oriented edges counted                                        */
    int32_t nbuckets;
    int32_t strategy[RTG_TC_MAX_BUCKETS];
    int64_t bucket_items[RTG_TC_MAX_BUCKETS];
    double bucket_ms[RTG_TC_MAX_BUCKETS];
    int32_t kernel_launches;
    int32_t reserved;
    uint64_t staged_bytes;   /* batched: bytes the owner-staged launch must read at least — each
                                partner list once per edge (4*min(d+(u), d+(v))), each owner's
                                list once per work item, the partner spans (8 B per edge), the
                                queue and the small-owner offsets; 0 for the plain methods     */
} rtg_tc_stats;

int rtg_tc_count(rtg_ctx *ctx, const rtg_graph *g, const rtg_tc_opts *opts, uint64_t *count,
                 rtg_tc_stats *stats /* may be NULL */);

/* A reusable plan: buckets and launch parameters computed once per graph. */
typedef struct rtg_tc_plan rtg_tc_plan;
int rtg_tc_plan_create(rtg_ctx *ctx, const rtg_graph *g, const rtg_tc_opts *opts,
                       rtg_tc_plan **out, rtg_tc_stats *stats);
int rtg_tc_plan_run(rtg_ctx *ctx, rtg_tc_plan *plan, uint64_t *count, rtg_tc_stats *stats);
/* Per-edge counts through a batched plan created with per_edge = 1: for every oriented edge
   e = (u -> v) of the plan's range, d_out[e] = |N+(u) & N+(v)| (triangle a<b<c attributed to
   edge a->b; per-edge outputs as the REDUCE==1 per-ray counts of rt_tc.cpp:64-72); entries
   outside the range are left alone. count (optional) receives the range's total; stats the
   kernel time.                                                                              */
int rtg_tc_plan_run_per_edge(rtg_ctx *ctx, rtg_tc_plan *plan, uint32_t *d_out, uint64_t *count,
                             rtg_tc_stats *stats);
int rtg_tc_plan_destroy(rtg_tc_plan *plan);
/* The plan's configuration as "key=value,..." text (method, tune, grid, table size, and
   whether the tune came from the autotuner, the cache, RTG_TC_TUNE or the defaults). */
int rtg_tc_plan_describe(const rtg_tc_plan *plan, char *buf, size_t len);

/* Per-edge counts for every oriented edge, written to d_out[m]. form 0: |N+(u) & N+(v)|
   (triangle a<b<c attributed to edge a->b) through a one-off batched per-edge plan; form 2:
   the same counts by one wave per edge (no planning: for small graphs); form 1: the 1A2
   attribution (edge a->c, ListIntersectionKernel's order, bs_tc.cu:120-148).               */
int rtg_tc_per_edge(rtg_ctx *ctx, const rtg_graph *g, int32_t form, uint32_t *d_out);

/* ---- sorted-set intersection ---------------------------------------------------------- */
#define RTG_SI_PLAIN 0     /* one wave per pair, binary search of the shorter in the longer    */
#define RTG_SI_BATCHED 1   /* one launch: pairs with a shorter list of <= 64 elements in flat
                              groups, longer ones as 64-element search items                  */
#define RTG_SI_REFERENCE 2 /* block per pair with an LDS splitter sample, the BlockBasedOBS
                              scheme (set_intersection/baseline/bs.cu:138-209)                */
/* Batched variants whose "balanced" long pairs (s + l <= s * (1 + ceil(log2(l + 1))) for
   shorter/longer lengths s, l) take another strategy; skewed ones keep the search items.     */
#define RTG_SI_MERGE 3     /* merge-path segments staged in LDS (baseline/ip.cu:138-287)       */
#define RTG_SI_HASH 4      /* LDS hash table per 1024-element chunk (baseline/hi.cu:134-195)  */
#define RTG_SI_BITMAP 5    /* LDS bitmap per 65,536-value window for dense lists, merge-path
                              segments for the sparser ones (baseline/bi_dynamic.cu:134-169)  */
#define RTG_SI_AUTO 6      /* times BATCHED, MERGE, HASH, BITMAP and REFERENCE once per
                              argument set (offsets, elements, pairs pointers and sizes) and
                              keeps the fastest; t_plan_ms holds the tuning time when it ran */

typedef struct rtg_si_stats {
    double t_kernel_ms;
    double t_plan_ms;
    uint64_t alg_bytes;  /* sum over pairs of 20 + list_bytes(|A|,|B|)                         */
    int32_t nbuckets;
    int32_t strategy;    /* long-pair strategy run: 0 search, 1 merge, 2 hash, 3 bitmap,
                            4 the REFERENCE form chosen by AUTO; -1 for PLAIN and REFERENCE   */
    int64_t bucket_items[RTG_TC_MAX_BUCKETS]; /* batched: units of each strategy (index as
                            strategy): 64-element search items, merge segments of 1024 steps,
                            hash chunks, bitmap windows                                        */
} rtg_si_stats;

/* Lists are sorted, duplicate-free uint32 runs elements[offsets[i] .. offsets[i+1]),
   i < nlists (at most 2^32 - 1). For pair k: d_out[k] = |list(pairs[2k]) & list(pairs[2k+1])|.
   A pair naming an id >= nlists is counted as 0 and the call returns RTG_EINVAL after it ran
   (the device never reads past offsets[nlists]).                                            */
int rtg_si_pairs(rtg_ctx *ctx, const uint32_t *d_offsets, const uint32_t *d_elements,
                 int64_t nlists, const uint32_t *d_pairs, int64_t npairs, int32_t method,
                 uint32_t *d_out, rtg_si_stats *stats);
/* Allocation-free form of rtg_si_pairs, for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It has no
   counterpart in the reference. Same inputs, outputs and results as rtg_si_pairs for the
   methods PLAIN, BATCHED, REFERENCE, MERGE, HASH and BITMAP (AUTO is refused: its tuning keeps
   host-side state), at most RTG_SI_NOALLOC_MAX_PAIRS pairs (more: RTG_EINVAL), no statistics
   beyond the kernel time. Its scratch (per-pair unit offsets, scan sums, flags) is one
   file-scope device array of about 16.1 MiB per GPU, held by the library until the process
   ends and shared by every call on that GPU: calls that overlap on one GPU (two contexts or
   threads on the same device) race and must be serialised by the caller. The call returns
   after the result is complete.                                                             */
#define RTG_SI_NOALLOC_MAX_PAIRS (1 << 21)
int rtg_si_pairs_noalloc(rtg_ctx *ctx, const uint32_t *d_offsets, const uint32_t *d_elements,
                         int64_t nlists, const uint32_t *d_pairs, int64_t npairs, int32_t method,
                         uint32_t *d_out, double *t_kernel_ms /* may be NULL */);
/* The reference's A x B tile mode: sets [0,num_a) against [num_a, num_a+num_b);
   d_counts[a*num_b + b] = |A_a & B_b| (BinarySearchBasedIntersection, bs.cu:91-135).        */
int rtg_si_tile(rtg_ctx *ctx, const uint32_t *d_offsets, const uint32_t *d_elements,
                uint32_t num_a, uint32_t num_b, int32_t method, uint32_t *d_counts,
                rtg_si_stats *stats);

/* ---- sorted-array lookup --------------------------------------------------------------- */
/* out[i] = index of the first element equal to keys[i], or -1 (BSAccessCount,
   rt_search/simple_bs.cu:35-53). data is ascending, duplicates allowed.                     */
int rtg_lookup_u32(rtg_ctx *ctx, const uint32_t *d_data, int64_t n, const uint32_t *d_keys,
                   int64_t q, int32_t *d_out, double *t_kernel_ms);
int rtg_lookup_i32(rtg_ctx *ctx, const int32_t *d_data, int64_t n, const int32_t *d_keys,
                   int64_t q, int32_t *d_out, double *t_kernel_ms);
/* Ascending device sort of n 32-bit keys in place (the thrust::sort of the lookup array in
   BSAccessCountFunction, rt_search/simple_bs.cu:99-108): LSD radix sort, 8-bit digits, with a
   temporary of n keys. t_ms (may be NULL) receives the device time of the sort.             */
int rtg_sort_u32(rtg_ctx *ctx, uint32_t *d_data, int64_t n, double *t_ms);
int rtg_sort_i32(rtg_ctx *ctx, int32_t *d_data, int64_t n, double *t_ms);
/* The same lookup through a search index built once per array, for callers that look up in
   one array many times (the reference re-searches its array per call, simple_bs.cu:80-142;
   this is the analogue of its build-once search structure). The index is a 32-ary tree of
   128-byte nodes over d_data (about n/8 bytes of device memory); d_data is referenced, not
   copied, and must stay allocated and unchanged while the index lives. Keys passed to
   rtg_lookup_index_run have the array's type. levels = tree levels above the data (0: the
   array is searched in LDS alone).                                                          */
typedef struct rtg_lookup_index rtg_lookup_index;
int rtg_lookup_index_create_u32(rtg_ctx *ctx, const uint32_t *d_data, int64_t n,
                                rtg_lookup_index **out);
int rtg_lookup_index_create_i32(rtg_ctx *ctx, const int32_t *d_data, int64_t n,
                                rtg_lookup_index **out);
int rtg_lookup_index_info(const rtg_lookup_index *ix, int32_t *levels, int64_t *bytes);
int rtg_lookup_index_run(rtg_ctx *ctx, const rtg_lookup_index *ix, const void *d_keys,
                         int64_t q, int32_t *d_out, double *t_kernel_ms);
int rtg_lookup_index_destroy(rtg_lookup_index *ix);

/* ---- query server: continuous batching over a device-resident graph and sorted array ---- */
#define RTG_REQ_PAIR 0     /* |N(u) & N(v)| on the symmetric CSR                                  */
#define RTG_REQ_LOOKUP 1   /* first-occurrence lookup of a batch of keys in the sorted array      */
/* request status */
#define RTG_REQ_PENDING 0
#define RTG_REQ_RUNNING 1
#define RTG_REQ_DONE 2
#define RTG_REQ_CANCELLED 3
#define RTG_REQ_TIMEOUT 4
#define RTG_REQ_FAILED 5

typedef struct rtg_server rtg_server;

typedef struct rtg_server_opts {
    int32_t slots;              /* concurrent batches (streams); default 2                      */
    int32_t max_batch_requests; /* requests per batch; default 8192                             */
    int64_t max_batch_pairs;    /* pair queries per batch; default 8192                         */
    int64_t max_batch_keys;     /* lookup keys per batch; default 2^21                          */
    int64_t mem_budget_bytes;   /* admission: bytes of admitted, unfinished requests; 1 GiB     */
    int64_t timeout_us;         /* queued longer than this: dropped with RTG_REQ_TIMEOUT (0=off)*/
    int32_t poll_us;            /* dispatcher wait when idle (submits wake it); default 1000    */
    int32_t null_executor;      /* no device: batches complete after null_latency_us, no results */
    int32_t null_latency_us;
    int32_t copy_staging;       /* 1: stage requests/results through device buffers with copies;
                                   0 (default): kernels read and write pinned host memory     */
} rtg_server_opts;

typedef struct rtg_server_stats {
    int64_t admitted, rejected, completed, cancelled, failed, batches;
    int64_t max_batch_requests, max_busy_slots;
    double mean_batch_requests;
    double mean_us, p50_us, p99_us, p999_us, max_us; /* arrival -> result visible to the host  */
    double throughput_qps, offered_qps;
    double mean_batch_device_us;  /* device time of a batch (copies and kernels)               */
    double mean_batch_pack_us;    /* host time packing a batch into staging buffers            */
    double mean_batch_unpack_us;  /* host time delivering a batch's results                    */
    double mean_batch_slot_us;    /* launch -> results delivered                               */
    double max_batch_device_us;   /* longest batch on the device                               */
    double max_batch_slot_us;     /* longest batch from launch to results delivered            */
    char last_error[128];
} rtg_server_stats;

/* g must be symmetric (rtg_graph_from_pairs without RTG_PAIRS_ORIENT); d_array ascending and
   16-byte aligned (the server builds an rtg_lookup_index over it).
   ctx, g and d_array may be NULL with opts->null_executor. */
int rtg_server_create(rtg_ctx *ctx, const rtg_graph *g, const uint32_t *d_array, int64_t n_array,
                      const rtg_server_opts *opts, rtg_server **out);
/* Submit returns RTG_EAGAIN (request rejected, nothing queued) when the memory budget is full.
   Lookup keys are read and results written (host memory) until the request is terminal. */
int rtg_server_submit_pair(rtg_server *s, int32_t u, int32_t v, uint64_t *req_id);
int rtg_server_submit_lookup(rtg_server *s, const uint32_t *keys, int32_t nkeys, int32_t *out,
                             uint64_t *req_id);
/* Status of a request; a terminal status (DONE, CANCELLED, TIMEOUT, FAILED) is reported once,
   after which the id is forgotten. */
int rtg_server_poll(rtg_server *s, uint64_t req_id, int32_t *status, uint32_t *pair_result,
                    int64_t *latency_ns);
int rtg_server_cancel(rtg_server *s, uint64_t req_id); /* only while pending (else EAGAIN)   */
int rtg_server_get_stats(rtg_server *s, rtg_server_stats *out);
int rtg_server_reset_stats(rtg_server *s);
int rtg_server_destroy(rtg_server *s);                 /* waits for in-flight batches        */

/* Synthetic Poisson stream (the serving workload): rate, count, warm-up (not measured),
   fraction of pair queries (endpoints drawn in proportion to degree), lookup batch size
   U{1..max_keys} of uniform keys. */
typedef struct rtg_stream_spec {
    double rate_qps;
    int64_t nqueries;
    int64_t warmup;
    double pair_frac;
    int32_t max_keys;
    int32_t reserved;
    uint64_t seed;
} rtg_stream_spec;
int rtg_server_replay(rtg_server *s, const rtg_stream_spec *spec, uint32_t *pair_out /* nullable */,
                      uint64_t *lookup_sum /* nullable: per lookup, sum of (result + 2) */,
                      rtg_server_stats *stats);
/* The stream's queries, for checking: pool == NULL returns the key-pool size in *pool_n. */
int rtg_stream_queries(const rtg_graph *g, const rtg_stream_spec *spec, int32_t *kinds, int32_t *a,
                       int32_t *b, int64_t *key_off, int32_t *nkeys, uint32_t *pool,
                       int64_t *pool_n);

/* ---- host-side generators (seeded, counter-based; no GPU needed) ------------------------ */
int rtg_gen_ba(int32_t n, int32_t m, uint64_t seed, int32_t *pairs /* 2*m*(n-m) */,
               int64_t *npairs);
int rtg_gen_rmat(uint32_t scale, int64_t nsamples, double a, double b, double c, uint64_t seed,
                 int32_t *pairs /* 2*nsamples */);
int rtg_gen_powerlaw_degrees(int32_t n, double gamma, double kmax, double mean, uint64_t seed,
                             int32_t *deg, int64_t *total);
int rtg_gen_config_model(int32_t n, const int32_t *deg, int64_t total, double p_close,
                         uint64_t seed, int32_t *pairs /* 2*(total/2 + total) */,
                         int64_t *npairs);
int rtg_gen_zipf_lengths(int64_t nlists, double s, uint32_t kmax, uint64_t seed, uint32_t *len);
int rtg_gen_sorted_lists(int64_t nlists, const uint32_t *offsets, uint32_t range, uint64_t seed,
                         uint32_t *elements);
int rtg_gen_sorted_array(int64_t n, uint64_t seed, uint32_t *out);
int rtg_gen_uniform_u32(int64_t n, uint64_t seed, uint64_t stream, uint32_t *out);

/* ---- host-side file ingest (the reference's formats) ----------------------------------- */
/* Binary little-endian int32 (src,dst) pairs, npairs = file_size/8 (Graph::ReadBinaryEdges,
   tc/graph/graph.cpp:23-39). Two calls: pairs == NULL returns the count only.               */
int rtg_read_binary_pairs(const char *path, int32_t *pairs, int64_t *npairs);
/* Text edge list, one "src dst" per line (Graph::Txt2Bin, tc/graph/graph.cpp:179-214). */
int rtg_read_text_pairs(const char *path, int32_t *pairs, int64_t *npairs);
/* Matrix Market coordinate file, 1-based -> 0-based, symmetric files expanded to both
   directions (Graph::ReadMtx, bfs/graph/graph.cpp:17-122). */
int rtg_read_mtx_pairs(const char *path, int32_t *pairs, int64_t *npairs);
int rtg_write_binary_pairs(const char *path, const int32_t *pairs, int64_t npairs);

#ifdef __cplusplus
}
#endif

#endif /* RTG_H */
