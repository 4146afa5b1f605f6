// Device-resident graph (CSR) shared by the counting, intersection and serving code.
#pragma once

#include "device.hpp"

struct rtg_graph {
    rtg_ctx *ctx = nullptr;
    int32_t n = 0;
    int64_t m = 0;
    int32_t max_row = 0;
    bool oriented = true;
    int32_t *d_offsets = nullptr;  // n + 1
    int32_t *d_adjs = nullptr;     // m
};

namespace rtg {
// Largest row length of a device CSR (device reduction, synchronous).
int csr_max_row(rtg_ctx *ctx, const int32_t *d_offsets, int32_t n, int32_t *out);
}
