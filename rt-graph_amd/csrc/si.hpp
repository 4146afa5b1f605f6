// Set-intersection launcher shared with the query server.
#pragma once

#include "device.hpp"

namespace rtg {
// Device scratch for si_launch_batched over npairs pairs.
size_t si_batched_work_bytes(int64_t npairs);
// Enqueues the bucketed pairwise intersection (classify + persistent drain) on `st`:
// d_out[k] = |list(pairs[2k]) & list(pairs[2k+1])| for lists elements[off[i] .. off[i+1]).
int si_launch_batched(rtg_ctx *ctx, hipStream_t st, const uint32_t *d_off, const uint32_t *d_el,
                      const uint32_t *d_pairs, int64_t npairs, void *work, uint32_t *d_out);
// Enqueues the latency-oriented form: d_items holds nitems pairs {k, first} (uint32 each), one
// per 64 consecutive elements of pair k's shorter list; d_out[k] must be zero beforehand and
// receives the same value as above. Pairs with an empty list have no items.
int si_launch_items(rtg_ctx *ctx, hipStream_t st, const uint32_t *d_off, const uint32_t *d_el, const uint32_t *d_pairs,
                    const uint32_t *d_items, int64_t nitems, uint32_t *d_out);
}  // namespace rtg
