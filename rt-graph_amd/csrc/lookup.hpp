// Sorted-array lookup launcher shared with the query server.
#pragma once

#include "device.hpp"

namespace rtg {
// Enqueues the lookup of q keys on `st` (no synchronisation).
template <class T>
int lookup_launch(rtg_ctx *ctx, hipStream_t st, const T *d_data, int64_t n, const T *d_keys, int64_t q,
                  int32_t *d_out);
}
