// Set-intersection launcher shared with the query server.
#pragma once

#include "device.hpp"

namespace rtg {
// Enqueues the latency-oriented pairwise intersection on `st`: d_items holds nitems pairs
// {k, first} (uint32 each), one per 64 consecutive elements of pair k's shorter list; d_out[k]
// must be zero beforehand and receives |list(pairs[2k]) & list(pairs[2k+1])| for lists
// elements[off[i] .. off[i+1]). Pairs with an empty list have no items.
int si_launch_items(rtg_ctx *ctx, hipStream_t st, const uint32_t *d_off, const uint32_t *d_el, const uint32_t *d_pairs,
                    const uint32_t *d_items, int64_t nitems, uint32_t *d_out);
}  // namespace rtg
