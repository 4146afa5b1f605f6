// Internal helpers shared by the host-side sources of librtg: status codes, the per-thread
// last-error string, and a small std::thread parallel-for for host-side generators.
#pragma once

#include <stdint.h>
#include <stddef.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rtg.h"

#define RTG_API __attribute__((visibility("default")))

namespace rtg {

// Records `msg` as this thread's last error and returns `code`.
int set_error(int code, const std::string &msg);

// Number of host threads for parallel_for: RTG_HOST_THREADS, else hardware_concurrency.
int host_threads();

// Splits [0, n) into contiguous chunks, one per thread, and calls f(lo, hi) on each.
template <class F>
void parallel_for(int64_t n, F f, int64_t min_chunk = 4096) {
    if (n <= 0) return;
    int64_t nt = std::min<int64_t>(host_threads(), (n + min_chunk - 1) / min_chunk);
    if (nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> ts;
    ts.reserve(nt);
    for (int64_t t = 0; t < nt; ++t) {
        int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        ts.emplace_back([=, &f] { f(lo, hi); });
    }
    for (auto &t : ts) t.join();
}

// Whether a device graph is degree-oriented, and setting it (csr_build.hip; for graphs stored
// to and loaded from a file).
bool graph_oriented(const rtg_graph *g);
void graph_set_oriented(rtg_graph *g, bool oriented);

}  // namespace rtg

inline int rtg_fail(int code, const std::string &msg) { return rtg::set_error(code, msg); }
