// Sorted-array lookup launchers shared with the query server.
#pragma once

#include "device.hpp"

// A 32-ary implicit search index over an ascending device array (the array itself is level 0
// and stays the caller's). Level L+1 holds, for every run of 32 entries of level L, the last
// entry of the run, i.e. the largest value of its subtree; each level is padded to a multiple
// of 32 entries (one 128-byte line per node) with the type's maximum. Levels are built until
// one has at most kIndexTop entries: that top level is staged in LDS by every block. The
// data must be 16-byte aligned; its last run of 32, when partial, is also kept as a padded
// copy so that every node read is a whole aligned line.
struct rtg_lookup_index {
    static constexpr int kMaxLevels = 6;
    rtg_ctx *ctx = nullptr;
    int32_t device = 0;
    int32_t is_signed = 0;
    const void *d_data = nullptr;
    int64_t n = 0;
    int32_t top = 0;                     // index of the top level (0: the data fits in LDS)
    int64_t len[kMaxLevels + 1] = {};    // real entries per level, len[0] = n
    int64_t off[kMaxLevels + 1] = {};    // element offset of level L >= 1 inside d_levels
    int64_t off_tail = 0;                // 32 entries: the data's last run, padded
    void *d_levels = nullptr;
    int64_t bytes = 0;
};

namespace rtg {
constexpr int kIndexFan = 32;
constexpr int kIndexTop = 8192;

// Enqueues the lookup of q keys on `st` (no synchronisation).
template <class T>
int lookup_launch(rtg_ctx *ctx, hipStream_t st, const T *d_data, int64_t n, const T *d_keys, int64_t q,
                  int32_t *d_out);
// Same results through the index.
template <class T>
int lookup_index_launch(const rtg_lookup_index *ix, hipStream_t st, const T *d_keys, int64_t q, int32_t *d_out);
// Builds the index (allocates, enqueues on the context stream and synchronises).
int lookup_index_build(rtg_ctx *ctx, const void *d_data, int64_t n, int is_signed, rtg_lookup_index **out);
// Enqueues the build of a one-call index whose levels live in the context's scratch (valid until
// the next call that uses the scratch; caller holds ctx->mu).
int lookup_index_transient(rtg_ctx *ctx, const void *d_data, int64_t n, int is_signed, rtg_lookup_index *ix);
void lookup_index_free(rtg_lookup_index *ix);
}  // namespace rtg
