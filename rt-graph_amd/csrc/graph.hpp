// Device-resident graph (CSR) shared by the counting, intersection and serving code.
#pragma once

#include "device.hpp"

struct rtg_renum;

struct rtg_graph {
    rtg_ctx *ctx = nullptr;
    int32_t n = 0;
    int64_t m = 0;
    int32_t max_row = 0;
    bool oriented = true;
    int32_t *d_offsets = nullptr;  // n + 1
    int32_t *d_adjs = nullptr;     // m
    // Degree-ordered copy for batched counting plans, built by the first plan that asks for it
    // (under the context lock) and kept until the graph is destroyed.
    rtg_renum *renum = nullptr;
};

// An oriented CSR with its vertices renumbered in descending order of total degree (out + in,
// ties by id): a triangle count does not depend on vertex ids, and in this order every
// out-neighbour (an endpoint of higher degree) gets a small id, so an out-list holds only ids
// below its own vertex's id in most rows. Rows are sorted. adj16 holds the low 16 bits of every
// entry, exact for a row whose largest entry (vinfo[v].z) is below 2^16. s16 and s256 sample every
// 16th and every 256th entry, for searches inside a row: s256 is small enough to stay in L2, and
// the 16 entries of s16 between two s256 samples fill one aligned 64-byte line.
struct rtg_renum {
    int32_t n = 0;
    int64_t m = 0;
    int32_t *d_off = nullptr;     // n + 1
    int32_t *d_adj = nullptr;     // m
    uint16_t *d_adj16 = nullptr;  // m
    int32_t *d_vinfo = nullptr;   // 4 n: row v's {start, end, largest entry (-1 when empty), 0}
    int32_t *d_inv = nullptr;     // n: new id -> the graph's id
    int32_t *d_s16 = nullptr;     // 16 * ceil(m / 256) + 16: d_adj[16 i], INT32_MAX past m
    int32_t *d_s256 = nullptr;    // ceil(m / 256): d_adj[256 j]
    int32_t *d_emap = nullptr;    // m: edge of the copy -> the graph's edge (built for per-edge plans)
    int64_t rows16 = 0;           // rows that fit 16 bits
    int64_t entries16 = 0;        // their entries
    float t_build_ms = 0.f;
    float t_emap_ms = 0.f;
};

namespace rtg {
// Largest row length of a device CSR (device reduction, synchronous).
int csr_max_row(rtg_ctx *ctx, const int32_t *d_offsets, int32_t n, int32_t *out);
// Occurrences of each id in ids[0 .. n) (0 <= id < domain) into deg[0 .. domain), which the
// caller has zeroed: a partitioned LDS histogram (ids grouped by id >> 14, each group counted
// in LDS), or device-scope atomics for domains above 2^24 ids. Allocates its own temporaries;
// enqueued on the context stream.
int id_histogram(rtg_ctx *ctx, const int32_t *ids, int64_t n, int64_t domain, uint32_t *deg);
// The graph's degree-ordered copy, built on first use (caller holds ctx->mu); with_emap also
// builds its edge map (once).
int graph_renumbered(rtg_ctx *ctx, const rtg_graph *g, bool with_emap, const rtg_renum **out);
void renum_free(rtg_renum *r);
}
