// Triangle-counting plan state shared by the TC sources.
#pragma once

#include "graph.hpp"

#include <string>
#include <vector>

// A work item of the batched kernel: `kind` names the strategy.
//   kItemHash    block item: owner vertex `a`'s out-list in the block's LDS cuckoo hash table,
//                partners wspan[b .. c) streamed against it;
//   kItemWave    wave item: small owners small_list[b .. c), one at a time per wave, each in a
//                wave-private table whose strategy is `a` (0 hash, 2 bitmap, 3 sorted list);
//   kItemSearch  block item whose partners' elements are binary-searched in the owner's list in
//                HBM (no LDS table);
//   kItemBitmap  block item with the owner's list as an LDS bitmap over its value range;
//   kItemSorted  block item with the owner's sorted list in LDS, merged with each partner list.
struct TcItem {
    int32_t kind;
    int32_t a;
    int32_t b;
    int32_t c;
};
constexpr uint32_t kItemHash = 0, kItemWave = 1, kItemSearch = 2, kItemBitmap = 3, kItemSorted = 4;

struct rtg_tc_plan {
    rtg_ctx *ctx = nullptr;
    const rtg_graph *g = nullptr;
    rtg_tc_opts opts{};
    int64_t e_begin = 0, e_end = 0;
    // device arrays
    int32_t *d_row = nullptr;        // edge -> source row (plain methods)
    int32_t *d_wofs = nullptr;       // batched: partner-list offsets per owner (n+1)
    int2 *d_wspan = nullptr;         // batched: partner out-list spans {start, length} by owner
    uint32_t *d_epos = nullptr;      // batched, per-edge plans: each edge's span position (or kNoSpan)
    uint32_t *d_ecnt = nullptr;      // batched, per-edge plans: each span's count, by position
    int32_t *d_small = nullptr;      // batched: small owners, heaviest first
    TcItem *d_items = nullptr;       // batched: work queue, heaviest first
    unsigned long long *d_work = nullptr;  // batched: streamed partner elements per owner (n+1)
    unsigned long long *d_count = nullptr;  // [0] = triangle count, [1] = queue head
    int64_t n_items = 0;
    int64_t n_small = 0;
    int64_t n_work_edges = 0;        // edges with a non-empty shorter list
    // launch parameters
    int32_t max_table = 0;           // LDS hash-table slots per block
    int32_t block_threads = 256;
    int32_t wave_owner_max = 0;      // owners with d+ <= this run as wave items
    int32_t fit_block = 0;           // largest out-degree whose hash table fits a block
    int32_t long_min = 48;           // partner lists at least this long are streamed directly
    int32_t grab = 64;               // partners per wave grab inside a block item
    int32_t grid = 0;
    int32_t unroll = 4;
    void *kernel = nullptr;          // batched: selected kernel instantiation
    void *kernel_pe = nullptr;       // batched, per-edge plans: the per-edge instantiation
    int32_t grid_pe = 0;
    // degree-ordered plans: the block bitmap items (the first n_bitmap_items of the queue) run in
    // their own kernel, counting and per-edge, before the general kernel takes the rest
    int64_t n_bitmap_items = 0;
    void *kernel_bm = nullptr, *kernel_pe_bm = nullptr;
    int32_t grid_bm = 0, grid_pe_bm = 0;
    int32_t bitmap_table = 0;        // the bitmap kernel's table words: the largest of its bitmaps
    // and, when their small owners are hashed, the wave items (the queue's last n_wave_items)
    int64_t n_wave_items = 0;
    void *kernel_wv = nullptr, *kernel_pe_wv = nullptr;
    int32_t grid_wv = 0, grid_pe_wv = 0;
    int32_t wave_table_wv = 0;       // the wave-items kernel's slots per wave
    int64_t small_work = 0;          // streamed elements of the small owners
    bool side = false;               // the wave kernel runs beside the bitmap kernel
    double wave_share = 0.0;         // the wave items' estimated share of the count's time
    int64_t wave_item_used = 0;      // the wave item size planned (the tune's, capped on small graphs)
    int32_t side_blocks = 0;         // the wave kernel's blocks per CU there (the last grid sized)
    // byte models: per-edge (no reuse), merge-only, and the owner-staged launch's reads
    uint64_t alg_bytes = 0, merge_bytes = 0, staged_bytes = 0;
    // captured launch sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double t_plan_ms = 0.0;
    double t_plan_graph_ms = 0.0, t_plan_tune_ms = 0.0, t_plan_items_ms = 0.0;  // host wall time of the stages
    // batched planning state
    uint64_t offsets_fingerprint = 0;  // of the row offsets (tuning-cache key)
    int64_t n_block_items = 0;
    int64_t n_search_items = 0;        // block items searched in HBM
    int64_t bucket_items[4] = {};      // per owner bucket: block items (small owners for bucket 0)
    int32_t strat[4] = {};             // per owner bucket: strategy (0 hash, 1 search, 2 bitmap, 3 merge)
    bool all_strategies = false;       // the queue holds bitmap or sorted-list items
    // The CSR the batched plan counts on: the graph's own, or its degree-ordered copy (full-range
    // counting plans; same count). With the copy, partner spans whose row fits 16 bits carry
    // bit 31 in their start and are streamed from v_adj16.
    const int32_t *v_off = nullptr;
    const int32_t *v_adj = nullptr;
    const uint16_t *v_adj16 = nullptr;
    const int32_t *v_vinfo = nullptr;  // the copy's rows: {start, end, largest entry, 0} each
    const int32_t *v_s16 = nullptr;    // the copy's every 16th entry
    const int32_t *v_s256 = nullptr;   // the copy's every 256th entry
    const int32_t *v_emap = nullptr;   // per-edge plans on the copy: its edge -> the graph's edge
    bool renumbered = false;
    int64_t streamed = 0;              // streamed partner elements (after the cuts)
    int64_t streamed16 = 0;            // of them, in 16-bit spans
    double t_renum_ms = 0.0;           // building the degree-ordered copy (first plan of a graph)
    std::string tune;                       // chosen configuration
    bool autotuned = false, tune_from_cache = false;
};

// Two-hop (2A1) count, tc_2a1.hip. Caller holds ctx->mu.
int rtg_tc_count_2a1_impl(rtg_ctx *ctx, const rtg_graph *g, uint64_t *count, rtg_tc_stats *stats);
