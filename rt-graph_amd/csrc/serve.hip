// Query server: a device-resident symmetric CSR and a sorted uint32 array answer a stream of
// pair intersections |N(u) & N(v)| and batched first-occurrence lookups, with continuous
// batching.
//
// Requests are admitted against a byte budget (request footprint: keys + results + pair
// slots) and queued. A dispatcher thread owns `slots` execution slots, each with its own HIP
// stream, event, pinned staging buffers and device buffers. Whenever a slot is free and
// requests are waiting, the dispatcher packs as many as the slot holds into one batch and
// launches it at once — the probe-item intersection kernel and the indexed lookup kernel,
// which read the packed requests from the slot's pinned host buffers and write the results
// back there (zero-copy; `copy_staging` stages through device buffers with copies instead) —
// so that requests arriving while earlier batches run join the next free slot and their work
// overlaps on the device. Completed batches are detected by event
// queries; each request records arrival and completion times on the host steady clock.
//
// Results do not depend on how requests are batched: each pair and each key is computed
// independently: a pair's count is a sum of per-element membership tests of its shorter
// list in its longer one, and a key's result depends on the key alone.
//
// With null_executor set, the server runs without a device: a batch "completes" after
// null_latency_us and writes no results (pair results 0, lookup outputs untouched), so that
// admission, batching, cancellation and timeouts can be exercised on any host.
#include "graph.hpp"
#include "lookup.hpp"
#include "rng.hpp"
#include "si.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <memory>
#include <cstring>
#include <deque>
#include <thread>

namespace {

using Clock = std::chrono::steady_clock;

inline int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
}

struct Req {
    uint64_t id = 0;
    int32_t kind = 0;  // RTG_REQ_PAIR / RTG_REQ_LOOKUP
    int32_t u = 0, v = 0;
    const uint32_t *keys = nullptr;
    int32_t nkeys = 0;
    int32_t items = 0;  // pair: 64-element probe items of the shorter list
    int32_t *out = nullptr;
    uint32_t pair_result = 0;
    int64_t t_arrive = 0, t_done = 0;
    int64_t bytes = 0;
    int64_t pos = 0;  // slot-local position of the pair or of the first key
    std::atomic<int32_t> status{RTG_REQ_PENDING};
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, done = nullptr;
    int64_t t_launch = 0, pack_ns = 0;
    uint32_t *h_pairs = nullptr, *h_pair_out = nullptr, *h_keys = nullptr, *h_items = nullptr;
    int32_t *h_lookup_out = nullptr;
    uint32_t *d_pairs = nullptr, *d_pair_out = nullptr, *d_keys = nullptr, *d_items = nullptr;
    int32_t *d_lookup_out = nullptr;
    // device addresses of the pinned host buffers (zero-copy mode)
    const uint32_t *z_pairs = nullptr, *z_items = nullptr, *z_keys = nullptr;
    uint32_t *z_pair_out = nullptr;
    int32_t *z_lookup_out = nullptr;
    std::vector<Req *> reqs;
    bool busy = false;  // dispatcher thread only; other threads read rtg_server::busy_slots
    int64_t fake_deadline = 0;
};

// Requests by id. Ids are consecutive from 1, so the table is a list of fixed chunks of
// pointers: inserting never moves existing entries (no rehash or reallocation while the
// server lock is held), and a chunk is released once every id in it has been collected.
class LiveTable {
   public:
    static constexpr int kBits = 16;
    static constexpr uint64_t kChunk = 1ull << kBits;

    void put(uint64_t id, Req *r) {
        const uint64_t i = id - 1, c = i >> kBits;
        if (c >= chunks_.size()) {
            chunks_.resize(c + 1);
            used_.resize(c + 1, 0);
        }
        if (!chunks_[c]) chunks_[c].reset(new Req *[kChunk]());
        chunks_[c][i & (kChunk - 1)] = r;
        ++used_[c];
        ++size_;
    }
    Req *get(uint64_t id) const {
        if (id == 0) return nullptr;
        const uint64_t i = id - 1, c = i >> kBits;
        return c < chunks_.size() && chunks_[c] ? chunks_[c][i & (kChunk - 1)] : nullptr;
    }
    // Removes a present id; frees its chunk when no id of it can be added or is still present.
    void erase(uint64_t id, uint64_t next_id) {
        const uint64_t i = id - 1, c = i >> kBits;
        chunks_[c][i & (kChunk - 1)] = nullptr;
        --size_;
        if (--used_[c] == 0 && next_id > (c + 1) * kChunk) chunks_[c].reset();
    }
    template <class F>
    void for_each(F f) const {
        for (const auto &ch : chunks_)
            if (ch)
                for (uint64_t k = 0; k < kChunk; ++k)
                    if (ch[k]) f(ch[k]);
    }
    int64_t size() const { return size_; }

   private:
    std::vector<std::unique_ptr<Req *[]>> chunks_;
    std::vector<int64_t> used_;
    int64_t size_ = 0;
};

}  // namespace

struct rtg_server {
    rtg_ctx *ctx = nullptr;
    const rtg_graph *g = nullptr;
    const uint32_t *d_array = nullptr;
    int64_t n_array = 0;
    rtg_lookup_index *index = nullptr;  // built over d_array at creation
    std::vector<int32_t> hoff;          // host copy of the CSR offsets (probe items per pair)
    int64_t max_items = 0;              // probe items per batch
    rtg_server_opts opts{};
    std::vector<Slot> slots;
    std::mutex mu;  // pending, live, stats
    std::deque<Req *> pending;
    LiveTable live;
    std::condition_variable wake;  // dispatcher idle wait; notified by submit
    std::atomic<uint64_t> next_id{1};
    std::atomic<int64_t> inflight_bytes{0};
    std::atomic<bool> stop{false};
    std::thread worker;
    // statistics (under mu)
    int64_t admitted = 0, rejected = 0, completed = 0, cancelled = 0, failed = 0, batches = 0;
    int64_t max_batch = 0, max_busy = 0, batched_requests = 0;
    int64_t busy_slots = 0;  // slots holding a batch (taken and not yet completed)
    int64_t t_first_arrival = 0, t_last_done = 0;
    double device_ms = 0, pack_ms = 0, unpack_ms = 0, slot_ms = 0;  // summed over batches
    double max_device_ms = 0, max_slot_ms = 0;
    std::deque<int64_t> latencies;  // ns, completed requests (chunked: no reallocation under the lock)
    std::string error;
};

namespace {

int64_t footprint(const Req &r) {
    return r.kind == RTG_REQ_PAIR ? 16 : 16 + 8 * (int64_t)r.nkeys;
}

void release(rtg_server *s, Req *r, int32_t status) {
    s->inflight_bytes -= r->bytes;
    r->t_done = now_ns();
    r->status.store(status, std::memory_order_release);
}

// Takes requests from the queue into `slot` while it has room; expired requests are
// cancelled instead. Caller holds s->mu. Each request taken is claimed here, PENDING ->
// RUNNING, under the lock that rtg_server_cancel also holds: a cancel either happened before
// (the request is CANCELLED and dropped) or fails afterwards, so a request is never both
// cancelled and launched.
void take_batch(rtg_server *s, Slot &slot, int64_t now) {
    int64_t np = 0, nk = 0, ni = 0;
    slot.reqs.clear();
    while (!s->pending.empty() && (int64_t)slot.reqs.size() < s->opts.max_batch_requests) {
        Req *r = s->pending.front();
        if (r->status.load(std::memory_order_acquire) != RTG_REQ_PENDING) {  // cancelled while queued
            s->pending.pop_front();
            continue;
        }
        if (s->opts.timeout_us > 0 && now - r->t_arrive > s->opts.timeout_us * 1000) {
            s->pending.pop_front();
            release(s, r, RTG_REQ_TIMEOUT);
            ++s->cancelled;
            continue;
        }
        if (r->kind == RTG_REQ_PAIR) {
            if (np + 1 > s->opts.max_batch_pairs || ni + r->items > s->max_items) break;
        } else {
            if (nk + r->nkeys > s->opts.max_batch_keys) break;
        }
        int32_t expected = RTG_REQ_PENDING;
        if (!r->status.compare_exchange_strong(expected, RTG_REQ_RUNNING, std::memory_order_acq_rel)) {
            s->pending.pop_front();
            continue;
        }
        s->pending.pop_front();
        if (r->kind == RTG_REQ_PAIR) {
            r->pos = np++;
            ni += r->items;
        } else {
            r->pos = nk;
            nk += r->nkeys;
        }
        slot.reqs.push_back(r);
    }
}

__global__ void fill_zero_u32(uint32_t *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void copy_u32(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

unsigned small_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(rtg::ceil_div(n, 256), 64)); }

int launch_slot(rtg_server *s, Slot &slot) {
    int64_t np = 0, nk = 0, ni = 0;
    const int64_t t_pack = now_ns();
    for (Req *r : slot.reqs) {
        if (r->kind == RTG_REQ_PAIR) {
            slot.h_pairs[2 * r->pos] = (uint32_t)r->u;
            slot.h_pairs[2 * r->pos + 1] = (uint32_t)r->v;
            for (int32_t c = 0; c < r->items; ++c) {
                slot.h_items[2 * ni] = (uint32_t)r->pos;
                slot.h_items[2 * ni + 1] = (uint32_t)(64 * c);
                ++ni;
            }
            ++np;
        } else {
            std::memcpy(slot.h_keys + r->pos, r->keys, sizeof(uint32_t) * (size_t)r->nkeys);
            nk += r->nkeys;
        }
    }
    slot.t_launch = now_ns();
    slot.pack_ns = slot.t_launch - t_pack;
    if (s->opts.null_executor) {
        slot.fake_deadline = now_ns() + (int64_t)s->opts.null_latency_us * 1000;
        return RTG_OK;
    }
    hipStream_t st = slot.stream;
    HIP_TRY(hipEventRecord(slot.start, st));
    if (!s->opts.copy_staging) {
        // Zero-copy: the kernels read the packed requests from pinned host memory and write the
        // lookup results there, so a batch is kernels only on one queue (copy engines and the
        // waits between them and the compute queue are not involved). Pair counts accumulate
        // in device memory (atomics) and a small kernel writes them out.
        if (np) {
            fill_zero_u32<<<small_grid(np), 256, 0, st>>>(slot.d_pair_out, np);
            HIP_TRY(hipGetLastError());
            if (ni)
                RTG_TRY(rtg::si_launch_items(s->ctx, st, (const uint32_t *)s->g->d_offsets, (const uint32_t *)s->g->d_adjs,
                                             slot.z_pairs, slot.z_items, ni, slot.d_pair_out));
            copy_u32<<<small_grid(np), 256, 0, st>>>(slot.d_pair_out, slot.z_pair_out, np);
            HIP_TRY(hipGetLastError());
        }
        if (nk) RTG_TRY(rtg::lookup_index_launch<uint32_t>(s->index, st, slot.z_keys, nk, slot.z_lookup_out));
        HIP_TRY(hipEventRecord(slot.done, st));
        return RTG_OK;
    }
    if (np) {
        HIP_TRY(hipMemcpyAsync(slot.d_pairs, slot.h_pairs, 8 * (size_t)np, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemsetAsync(slot.d_pair_out, 0, 4 * (size_t)np, st));
        if (ni) {
            HIP_TRY(hipMemcpyAsync(slot.d_items, slot.h_items, 8 * (size_t)ni, hipMemcpyHostToDevice, st));
            RTG_TRY(rtg::si_launch_items(s->ctx, st, (const uint32_t *)s->g->d_offsets, (const uint32_t *)s->g->d_adjs,
                                         slot.d_pairs, slot.d_items, ni, slot.d_pair_out));
        }
        HIP_TRY(hipMemcpyAsync(slot.h_pair_out, slot.d_pair_out, 4 * (size_t)np, hipMemcpyDeviceToHost, st));
    }
    if (nk) {
        HIP_TRY(hipMemcpyAsync(slot.d_keys, slot.h_keys, 4 * (size_t)nk, hipMemcpyHostToDevice, st));
        RTG_TRY(rtg::lookup_index_launch<uint32_t>(s->index, st, slot.d_keys, nk, slot.d_lookup_out));
        HIP_TRY(hipMemcpyAsync(slot.h_lookup_out, slot.d_lookup_out, 4 * (size_t)nk, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipEventRecord(slot.done, st));
    return RTG_OK;
}

void complete_slot(rtg_server *s, Slot &slot, bool ok) {
    const bool null_exec = s->opts.null_executor != 0;
    const int64_t t_unpack = now_ns();
    float dev_ms = 0.f;
    if (ok && !null_exec && hipEventElapsedTime(&dev_ms, slot.start, slot.done) != hipSuccess) dev_ms = 0.f;
    for (Req *r : slot.reqs) {
        if (ok && !null_exec) {
            if (r->kind == RTG_REQ_PAIR) r->pair_result = slot.h_pair_out[r->pos];
            else std::memcpy(r->out, slot.h_lookup_out + r->pos, sizeof(int32_t) * (size_t)r->nkeys);
        }
    }
    const int64_t t_end = now_ns();
    std::lock_guard<std::mutex> lock(s->mu);
    s->device_ms += dev_ms;
    s->max_device_ms = std::max(s->max_device_ms, (double)dev_ms);
    s->max_slot_ms = std::max(s->max_slot_ms, (double)(t_end - slot.t_launch) / 1e6);
    s->pack_ms += (double)slot.pack_ns / 1e6;
    s->unpack_ms += (double)(t_end - t_unpack) / 1e6;
    s->slot_ms += (double)(t_end - slot.t_launch) / 1e6;
    for (Req *r : slot.reqs) {
        release(s, r, ok ? RTG_REQ_DONE : RTG_REQ_FAILED);
        if (ok) {
            ++s->completed;
            s->latencies.push_back(r->t_done - r->t_arrive);
            s->t_last_done = std::max(s->t_last_done, r->t_done);
        } else {
            ++s->failed;
        }
    }
    slot.reqs.clear();
    --s->busy_slots;
    slot.busy = false;
}

void dispatcher(rtg_server *s) {
    if (!s->opts.null_executor) (void)hipSetDevice(s->ctx->device);
    while (!s->stop.load(std::memory_order_acquire)) {
        bool progress = false;
        int busy = 0;
        for (Slot &slot : s->slots) {
            if (!slot.busy) continue;
            bool finished = false, ok = true;
            if (s->opts.null_executor) {
                finished = now_ns() >= slot.fake_deadline;
            } else {
                const hipError_t q = hipEventQuery(slot.done);
                if (q == hipSuccess) finished = true;
                else if (q != hipErrorNotReady) {
                    finished = true;
                    ok = false;
                    std::lock_guard<std::mutex> lock(s->mu);
                    s->error = std::string("batch failed: ") + hipGetErrorString(q);
                }
            }
            if (finished) {
                complete_slot(s, slot, ok);
                progress = true;
            } else {
                ++busy;
            }
        }
        for (Slot &slot : s->slots) {
            if (slot.busy) continue;
            {
                std::lock_guard<std::mutex> lock(s->mu);
                if (s->pending.empty()) break;
                take_batch(s, slot, now_ns());
                if (slot.reqs.empty()) continue;
                ++s->batches;
                ++s->busy_slots;
                s->batched_requests += (int64_t)slot.reqs.size();
                s->max_batch = std::max<int64_t>(s->max_batch, (int64_t)slot.reqs.size());
            }
            slot.busy = true;
            ++busy;
            progress = true;
            if (launch_slot(s, slot) != RTG_OK) {
                {
                    std::lock_guard<std::mutex> lock(s->mu);
                    s->error = rtg_last_error();
                }
                // whatever was enqueued before the failure still reads the slot's buffers
                if (!s->opts.null_executor) (void)hipStreamSynchronize(slot.stream);
                complete_slot(s, slot, false);
                --busy;
            }
        }
        {
            std::lock_guard<std::mutex> lock(s->mu);
            s->max_busy = std::max<int64_t>(s->max_busy, busy);
        }
        if (progress) continue;
        if (busy > 0) {
            std::this_thread::yield();  // batches in flight: keep polling their events
        } else {
            // idle: sleep until a request arrives (submit notifies) or poll_us passes
            std::unique_lock<std::mutex> lock(s->mu);
            if (s->pending.empty() && !s->stop.load(std::memory_order_acquire))
                s->wake.wait_for(lock, std::chrono::microseconds(s->opts.poll_us));
        }
    }
}

void free_slot(Slot &slot, bool device) {
    if (!device) {
        std::free(slot.h_pairs);
        std::free(slot.h_pair_out);
        std::free(slot.h_keys);
        std::free(slot.h_lookup_out);
        return;
    }
    if (slot.stream) (void)hipStreamSynchronize(slot.stream);
    void *host[] = {slot.h_pairs, slot.h_pair_out, slot.h_keys, slot.h_lookup_out, slot.h_items};
    for (void *p : host)
        if (p) (void)hipHostFree(p);
    void *dev[] = {slot.d_pairs, slot.d_pair_out, slot.d_keys, slot.d_lookup_out, slot.d_items};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (slot.done) (void)hipEventDestroy(slot.done);
    if (slot.start) (void)hipEventDestroy(slot.start);
    if (slot.stream) (void)hipStreamDestroy(slot.stream);
}

int alloc_slot(rtg_server *s, Slot &slot) {
    const int64_t np = s->opts.max_batch_pairs, nk = s->opts.max_batch_keys;
    if (s->opts.null_executor) {
        slot.h_pairs = (uint32_t *)std::malloc(8 * (size_t)np);
        slot.h_pair_out = (uint32_t *)std::malloc(4 * (size_t)np);
        slot.h_keys = (uint32_t *)std::malloc(4 * (size_t)nk);
        slot.h_lookup_out = (int32_t *)std::malloc(4 * (size_t)nk);
        if (!slot.h_pairs || !slot.h_pair_out || !slot.h_keys || !slot.h_lookup_out)
            return rtg::set_error(RTG_ENOMEM, "server: host staging allocation failed");
        return RTG_OK;
    }
    HIP_TRY(hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&slot.start));
    HIP_TRY(hipEventCreate(&slot.done));
    HIP_TRY(hipHostMalloc((void **)&slot.h_pairs, 8 * (size_t)np, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&slot.h_pair_out, 4 * (size_t)np, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&slot.h_keys, 4 * (size_t)nk, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&slot.h_lookup_out, 4 * (size_t)nk, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&slot.h_items, 8 * (size_t)s->max_items, hipHostMallocDefault));
    {
        void *z = nullptr;
        HIP_TRY(hipHostGetDevicePointer(&z, slot.h_pairs, 0));
        slot.z_pairs = (const uint32_t *)z;
        HIP_TRY(hipHostGetDevicePointer(&z, slot.h_items, 0));
        slot.z_items = (const uint32_t *)z;
        HIP_TRY(hipHostGetDevicePointer(&z, slot.h_keys, 0));
        slot.z_keys = (const uint32_t *)z;
        HIP_TRY(hipHostGetDevicePointer(&z, slot.h_pair_out, 0));
        slot.z_pair_out = (uint32_t *)z;
        HIP_TRY(hipHostGetDevicePointer(&z, slot.h_lookup_out, 0));
        slot.z_lookup_out = (int32_t *)z;
    }
    HIP_TRY(hipMalloc((void **)&slot.d_pairs, 8 * (size_t)np));
    HIP_TRY(hipMalloc((void **)&slot.d_pair_out, 4 * (size_t)np));
    HIP_TRY(hipMalloc((void **)&slot.d_keys, 4 * (size_t)nk));
    HIP_TRY(hipMalloc((void **)&slot.d_lookup_out, 4 * (size_t)nk));
    HIP_TRY(hipMalloc((void **)&slot.d_items, 8 * (size_t)s->max_items));
    return RTG_OK;
}

void defaults(rtg_server_opts &o) {
    // Two slots overlap one batch's host work with the other's kernels. (With copy staging,
    // batches whose copies and kernels alternate between the copy engines and the compute queue
    // stalled for up to 7-11 ms now and then, more often with more slots; zero-copy batches
    // never exceeded 0.4 ms of device time at 2e6 queries/s.)
    if (o.slots <= 0) o.slots = 2;
    if (o.max_batch_requests <= 0) o.max_batch_requests = 8192;
    if (o.max_batch_pairs <= 0) o.max_batch_pairs = 8192;
    if (o.max_batch_keys <= 0) o.max_batch_keys = 1 << 21;
    if (o.mem_budget_bytes <= 0) o.mem_budget_bytes = (int64_t)1 << 30;
    if (o.poll_us <= 0) o.poll_us = 1000;
    if (o.null_latency_us <= 0) o.null_latency_us = 50;
}

int submit(rtg_server *s, Req *r, uint64_t *req_id) {
    r->bytes = footprint(*r);
    r->t_arrive = now_ns();
    // Admission by memory: the request's bytes must fit the budget next to everything in flight.
    int64_t cur = s->inflight_bytes.load();
    while (true) {
        if (cur + r->bytes > s->opts.mem_budget_bytes) {
            std::lock_guard<std::mutex> lock(s->mu);
            ++s->rejected;
            delete r;
            return rtg_fail(RTG_EAGAIN, "server: memory budget exhausted, request rejected");
        }
        if (s->inflight_bytes.compare_exchange_weak(cur, cur + r->bytes)) break;
    }
    r->id = s->next_id.fetch_add(1);
    {
        std::lock_guard<std::mutex> lock(s->mu);
        if (s->admitted == 0) s->t_first_arrival = r->t_arrive;
        ++s->admitted;
        s->live.put(r->id, r);
        s->pending.push_back(r);
    }
    s->wake.notify_one();
    *req_id = r->id;
    return RTG_OK;
}

double percentile(std::vector<int64_t> &v, double q) {
    if (v.empty()) return 0.0;
    const size_t k = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + k, v.end());
    return (double)v[k] / 1000.0;
}

}  // namespace

extern "C" {

RTG_API int rtg_server_create(rtg_ctx *ctx, const rtg_graph *g, const uint32_t *d_array, int64_t n_array,
                              const rtg_server_opts *opts, rtg_server **out) {
    if (!out) return rtg_fail(RTG_EINVAL, "rtg_server_create: null out");
    *out = nullptr;
    rtg_server_opts o{};
    if (opts) o = *opts;
    defaults(o);
    if (!o.null_executor) {
        RTG_TRY(rtg::ctx_enter(ctx));
        if (!g || g->oriented) return rtg_fail(RTG_EINVAL, "rtg_server_create: needs a symmetric graph (built without RTG_PAIRS_ORIENT)");
        if (n_array < 0 || (n_array > 0 && !d_array) || n_array > INT32_MAX || ((uintptr_t)d_array & 15))
            return rtg_fail(RTG_EINVAL, "rtg_server_create: bad sorted array (size, or not 16-byte aligned)");
    }
    auto *s = new rtg_server();
    s->ctx = ctx;
    s->g = g;
    s->d_array = d_array;
    s->n_array = n_array;
    s->opts = o;
    if (!o.null_executor) {
        // Probe items per batch: 16 per pair on average, and always room for the longest pair.
        s->max_items = 16 * o.max_batch_pairs + rtg::ceil_div(std::max<int64_t>(g->max_row, 1), 64);
        s->hoff.resize((size_t)g->n + 1);
        if (hipMemcpy(s->hoff.data(), g->d_offsets, sizeof(int32_t) * ((size_t)g->n + 1), hipMemcpyDeviceToHost) != hipSuccess) {
            (void)hipGetLastError();
            delete s;
            return rtg_fail(RTG_EHIP, "rtg_server_create: reading the graph offsets failed");
        }
        std::lock_guard<std::mutex> lock(ctx->mu);
        const int rc = rtg::lookup_index_build(ctx, d_array, n_array, 0, &s->index);
        if (rc != RTG_OK) {
            delete s;
            return rc;
        }
    }
    s->slots.resize((size_t)o.slots);
    for (Slot &slot : s->slots) {
        const int rc = alloc_slot(s, slot);
        if (rc != RTG_OK) {
            for (Slot &sl : s->slots) free_slot(sl, !o.null_executor);
            rtg::lookup_index_free(s->index);
            delete s;
            return rc;
        }
    }
    s->worker = std::thread(dispatcher, s);
    *out = s;
    return RTG_OK;
}

RTG_API int rtg_server_submit_pair(rtg_server *s, int32_t u, int32_t v, uint64_t *req_id) {
    if (!s || !req_id) return rtg_fail(RTG_EINVAL, "rtg_server_submit_pair: null argument");
    if (!s->opts.null_executor && (u < 0 || v < 0 || u >= s->g->n || v >= s->g->n))
        return rtg_fail(RTG_EINVAL, "rtg_server_submit_pair: vertex out of range");
    auto *r = new Req();
    r->kind = RTG_REQ_PAIR;
    r->u = u;
    r->v = v;
    if (!s->opts.null_executor) {
        const int32_t du = s->hoff[u + 1] - s->hoff[u], dv = s->hoff[v + 1] - s->hoff[v];
        r->items = (std::min(du, dv) + 63) / 64;
    }
    return submit(s, r, req_id);
}

RTG_API int rtg_server_submit_lookup(rtg_server *s, const uint32_t *keys, int32_t nkeys, int32_t *out,
                                     uint64_t *req_id) {
    if (!s || !req_id || nkeys <= 0 || !keys || !out) return rtg_fail(RTG_EINVAL, "rtg_server_submit_lookup: bad arguments");
    if (nkeys > s->opts.max_batch_keys) return rtg_fail(RTG_EINVAL, "rtg_server_submit_lookup: more keys than a batch holds");
    auto *r = new Req();
    r->kind = RTG_REQ_LOOKUP;
    r->keys = keys;
    r->nkeys = nkeys;
    r->out = out;
    return submit(s, r, req_id);
}

RTG_API int rtg_server_poll(rtg_server *s, uint64_t req_id, int32_t *status, uint32_t *pair_result,
                            int64_t *latency_ns) {
    if (!s || !status) return rtg_fail(RTG_EINVAL, "rtg_server_poll: null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    Req *r = s->live.get(req_id);
    if (!r) return rtg_fail(RTG_EINVAL, "rtg_server_poll: unknown or already collected request");
    const int32_t st = r->status.load(std::memory_order_acquire);
    *status = st;
    if (st == RTG_REQ_PENDING || st == RTG_REQ_RUNNING) return RTG_OK;
    if (pair_result) *pair_result = r->pair_result;
    if (latency_ns) *latency_ns = r->t_done - r->t_arrive;
    s->live.erase(req_id, s->next_id.load());
    delete r;  // terminal: collected
    return RTG_OK;
}

RTG_API int rtg_server_cancel(rtg_server *s, uint64_t req_id) {
    if (!s) return rtg_fail(RTG_EINVAL, "rtg_server_cancel: null server");
    std::lock_guard<std::mutex> lock(s->mu);
    Req *r = s->live.get(req_id);
    if (!r) return rtg_fail(RTG_EINVAL, "rtg_server_cancel: unknown request");
    int32_t expected = RTG_REQ_PENDING;
    if (!r->status.compare_exchange_strong(expected, RTG_REQ_CANCELLED))
        return rtg_fail(RTG_EAGAIN, "rtg_server_cancel: request already running or finished");
    s->inflight_bytes -= r->bytes;
    r->t_done = now_ns();
    ++s->cancelled;
    auto pit = std::find(s->pending.begin(), s->pending.end(), r);
    if (pit != s->pending.end()) s->pending.erase(pit);
    return RTG_OK;
}

RTG_API int rtg_server_get_stats(rtg_server *s, rtg_server_stats *out) {
    if (!s || !out) return rtg_fail(RTG_EINVAL, "rtg_server_get_stats: null argument");
    std::vector<int64_t> lat;
    std::lock_guard<std::mutex> lock(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->admitted = s->admitted;
    out->rejected = s->rejected;
    out->completed = s->completed;
    out->cancelled = s->cancelled;
    out->failed = s->failed;
    out->batches = s->batches;
    out->max_batch_requests = s->max_batch;
    out->max_busy_slots = s->max_busy;
    out->mean_batch_requests = s->batches ? (double)s->batched_requests / (double)s->batches : 0.0;
    if (s->batches) {
        out->mean_batch_device_us = 1e3 * s->device_ms / (double)s->batches;
        out->mean_batch_pack_us = 1e3 * s->pack_ms / (double)s->batches;
        out->mean_batch_unpack_us = 1e3 * s->unpack_ms / (double)s->batches;
        out->mean_batch_slot_us = 1e3 * s->slot_ms / (double)s->batches;
        out->max_batch_device_us = 1e3 * s->max_device_ms;
        out->max_batch_slot_us = 1e3 * s->max_slot_ms;
    }
    lat.assign(s->latencies.begin(), s->latencies.end());
    if (!lat.empty()) {
        double sum = 0;
        int64_t mx = 0;
        for (int64_t x : lat) {
            sum += (double)x;
            mx = std::max(mx, x);
        }
        out->mean_us = sum / (double)lat.size() / 1000.0;
        out->max_us = (double)mx / 1000.0;
        out->p50_us = percentile(lat, 0.50);
        out->p99_us = percentile(lat, 0.99);
        out->p999_us = percentile(lat, 0.999);
        const double span = (double)(s->t_last_done - s->t_first_arrival) / 1e9;
        out->throughput_qps = span > 0 ? (double)s->completed / span : 0.0;
    }
    if (!s->error.empty()) std::snprintf(out->last_error, sizeof(out->last_error), "%s", s->error.c_str());
    return RTG_OK;
}

RTG_API int rtg_server_reset_stats(rtg_server *s) {
    if (!s) return rtg_fail(RTG_EINVAL, "rtg_server_reset_stats: null server");
    std::lock_guard<std::mutex> lock(s->mu);
    s->admitted = s->rejected = s->completed = s->cancelled = s->failed = s->batches = 0;
    s->max_batch = s->max_busy = s->batched_requests = 0;
    s->t_first_arrival = s->t_last_done = 0;
    s->device_ms = s->pack_ms = s->unpack_ms = s->slot_ms = 0;
    s->max_device_ms = s->max_slot_ms = 0;
    s->latencies.clear();
    return RTG_OK;
}

RTG_API int rtg_server_destroy(rtg_server *s) {
    if (!s) return RTG_OK;
    // Let in-flight batches finish, then stop the dispatcher.
    for (int spin = 0; spin < 200000; ++spin) {
        bool idle;
        {
            std::lock_guard<std::mutex> lock(s->mu);
            idle = s->pending.empty() && s->busy_slots == 0;
        }
        if (idle) break;
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    s->stop.store(true);
    s->wake.notify_all();
    if (s->worker.joinable()) s->worker.join();
    if (!s->opts.null_executor && s->ctx) (void)hipSetDevice(s->ctx->device);
    for (Slot &slot : s->slots) free_slot(slot, !s->opts.null_executor);
    rtg::lookup_index_free(s->index);
    s->pending.clear();  // every queued request is also in `live`
    s->live.for_each([](Req *r) { delete r; });
    delete s;
    return RTG_OK;
}

// Replays a synthetic Poisson stream against the server from a dedicated submitting thread
// and waits for every admitted request. Query i arrives at the i-th point of a Poisson process
// of rate spec->rate_qps (exponential gaps from crand(seed, QUERY, 3i)); with probability
// pair_frac it is a pair (u, v), each endpoint the source of a uniformly drawn CSR entry (so
// drawn in proportion to degree), otherwise a lookup of U{1..max_keys} consecutive keys from a
// pool of uniform uint32 keys. The first `warmup` queries are excluded from the statistics.
// pair_out (optional, nqueries entries) receives pair results; lookup_sum (optional) receives,
// per lookup query, the sum of (result + 2) over its keys.
RTG_API int rtg_server_replay(rtg_server *s, const rtg_stream_spec *spec, uint32_t *pair_out, uint64_t *lookup_sum,
                              rtg_server_stats *stats) {
    if (!s || !spec || spec->nqueries <= 0 || spec->rate_qps <= 0) return rtg_fail(RTG_EINVAL, "rtg_server_replay: bad spec");
    const int64_t nq = spec->nqueries;
    const int32_t max_keys = std::max<int32_t>(1, std::min<int32_t>(spec->max_keys, (int32_t)s->opts.max_batch_keys));
    const uint64_t seed = spec->seed;
    const int64_t pool_n = std::max<int64_t>(1 << 22, 4 * (int64_t)max_keys);
    std::vector<uint32_t> pool((size_t)pool_n);
    for (int64_t i = 0; i < pool_n; ++i) pool[i] = (uint32_t)(rtg::crand(seed, rtg::kStreamQuery ^ 0x706F6F6Cull, (uint64_t)i) >> 32);
    const int64_t m = s->g ? s->g->m : 0;
    std::vector<int32_t> hoff, hadj;
    if (!s->opts.null_executor && m > 0) {
        hoff.resize((size_t)s->g->n + 1);
        hadj.resize((size_t)m);
        RTG_TRY(rtg_graph_download(s->g, hoff.data(), hadj.data()));
    }
    auto src_of = [&](int64_t e) -> int32_t {
        return (int32_t)(std::upper_bound(hoff.begin(), hoff.end(), (int32_t)e) - hoff.begin()) - 1;
    };
    struct Q {
        int64_t t;
        int32_t kind, a, b;
        int64_t key_off;
        int32_t nkeys;
    };
    std::vector<Q> qs((size_t)nq);
    double t = 0.0;
    const uint64_t tp = rtg::prob53(spec->pair_frac);
    int64_t total_keys = 0;
    for (int64_t i = 0; i < nq; ++i) {
        const double u = rtg::u01(rtg::crand(seed, rtg::kStreamQuery, 3 * (uint64_t)i));
        t += -std::log(1.0 - u) / spec->rate_qps;
        Q q{};
        q.t = (int64_t)(t * 1e9);
        const uint64_t r1 = rtg::crand(seed, rtg::kStreamQuery, 3 * (uint64_t)i + 1);
        const uint64_t r2 = rtg::crand(seed, rtg::kStreamQuery, 3 * (uint64_t)i + 2);
        if (rtg::top53(r1) < tp) {
            q.kind = RTG_REQ_PAIR;
            if (m > 0) {
                q.a = src_of((int64_t)((r2 >> 32) % (uint64_t)m));
                q.b = src_of((int64_t)((r2 & 0xFFFFFFFFull) % (uint64_t)m));
            }
        } else {
            q.kind = RTG_REQ_LOOKUP;
            q.nkeys = 1 + (int32_t)rtg::bounded32(r2, (uint32_t)max_keys);
            q.key_off = (int64_t)((r2 & 0xFFFFFFFFull) % (uint64_t)(pool_n - q.nkeys + 1));
            total_keys += q.nkeys;
        }
        qs[i] = q;
    }
    // Lookup outputs: one region per query when the caller wants the sums, else one shared
    // discard region (only the dispatcher thread writes it).
    const bool keep = lookup_sum != nullptr;
    std::vector<int32_t> outs((size_t)std::max<int64_t>(keep ? total_keys : max_keys, 1));
    std::vector<int64_t> out_off((size_t)nq, 0);
    int64_t acc = 0;
    for (int64_t i = 0; i < nq; ++i)
        if (keep && qs[i].kind == RTG_REQ_LOOKUP) {
            out_off[i] = acc;
            acc += qs[i].nkeys;
        }
    std::vector<uint64_t> ids((size_t)nq, 0);
    std::vector<char> admitted((size_t)nq, 0);
    const int64_t warm = std::min<int64_t>(spec->warmup, nq);
    int64_t t0 = now_ns(), tbase = 0;
    int64_t ci = 0;  // oldest request not yet collected
    // Collects request j if it has finished (done = true), recording its pair result.
    auto collect = [&](int64_t j, bool &done) -> int {
        int32_t st = RTG_REQ_PENDING;
        uint32_t pr = 0;
        RTG_TRY(rtg_server_poll(s, ids[j], &st, &pr, nullptr));
        done = st != RTG_REQ_PENDING && st != RTG_REQ_RUNNING;
        if (done) {
            if (pair_out && qs[j].kind == RTG_REQ_PAIR) pair_out[j] = pr;
            admitted[j] = 2;
        }
        return RTG_OK;
    };
    for (int64_t i = 0; i < nq; ++i) {
        if (i == warm) {
            // wait for the warm-up queries, then measure from a clean slate
            for (int64_t j = 0; j < warm; ++j) {
                if (admitted[j] != 1) continue;
                bool done = false;
                while (true) {
                    RTG_TRY(collect(j, done));
                    if (done) break;
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                }
            }
            RTG_TRY(rtg_server_reset_stats(s));
            t0 = now_ns();  // the measured stream starts now, at its own first arrival time
            tbase = qs[i].t;
        }
        const int64_t due = t0 + qs[i].t - tbase;
        // Until the next arrival, collect finished requests in order, as a client would
        // (at most one poll every 2 us, none within 5 us of the arrival).
        for (int64_t now = now_ns(), next_poll = now; now < due; now = now_ns()) {
            if (now < next_poll || due - now < 5000) continue;
            next_poll = now + 2000;
            while (ci < i && admitted[ci] != 1) ++ci;
            if (ci < i) {
                bool done = false;
                RTG_TRY(collect(ci, done));
                if (done) ++ci;
            }
        }
        int rc;
        if (qs[i].kind == RTG_REQ_PAIR) rc = rtg_server_submit_pair(s, qs[i].a, qs[i].b, &ids[i]);
        else rc = rtg_server_submit_lookup(s, pool.data() + qs[i].key_off, qs[i].nkeys, outs.data() + out_off[i], &ids[i]);
        if (rc == RTG_OK) admitted[i] = 1;
        else if (rc != RTG_EAGAIN) return rc;
    }
    // Collect the rest.
    for (int64_t i = ci; i < nq; ++i) {
        if (admitted[i] != 1) continue;
        bool done = false;
        while (true) {
            RTG_TRY(collect(i, done));
            if (done) break;
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
    if (lookup_sum)
        for (int64_t i = 0; i < nq; ++i) {
            lookup_sum[i] = 0;
            if (qs[i].kind != RTG_REQ_LOOKUP || !admitted[i]) continue;
            uint64_t sum = 0;
            for (int32_t j = 0; j < qs[i].nkeys; ++j) sum += (uint64_t)(int64_t)(outs[out_off[i] + j] + 2);
            lookup_sum[i] = sum;
        }
    if (stats) RTG_TRY(rtg_server_get_stats(s, stats));
    if (stats) stats->offered_qps = spec->rate_qps;
    return RTG_OK;
}

// The replay's query list without a server, so a checker can recompute every answer:
// kinds[i], a[i], b[i] (pair endpoints) or key_off[i] / nkeys[i] into the key pool, and the
// pool itself (pool_n entries, pool_n as computed by the replay for this spec).
RTG_API int rtg_stream_queries(const rtg_graph *g, const rtg_stream_spec *spec, int32_t *kinds, int32_t *a, int32_t *b,
                               int64_t *key_off, int32_t *nkeys, uint32_t *pool, int64_t *pool_n) {
    if (!spec || spec->nqueries <= 0 || !pool_n) return rtg_fail(RTG_EINVAL, "rtg_stream_queries: bad spec");
    const int32_t max_keys = std::max<int32_t>(1, spec->max_keys);
    const int64_t pn = std::max<int64_t>(1 << 22, 4 * (int64_t)max_keys);
    if (!pool) {
        *pool_n = pn;
        return RTG_OK;
    }
    for (int64_t i = 0; i < pn; ++i) pool[i] = (uint32_t)(rtg::crand(spec->seed, rtg::kStreamQuery ^ 0x706F6F6Cull, (uint64_t)i) >> 32);
    std::vector<int32_t> hoff;
    int64_t m = 0;
    if (g && g->m > 0) {
        m = g->m;
        hoff.resize((size_t)g->n + 1);
        RTG_TRY(rtg_graph_download(g, hoff.data(), nullptr));
    }
    const uint64_t tp = rtg::prob53(spec->pair_frac);
    for (int64_t i = 0; i < spec->nqueries; ++i) {
        const uint64_t r1 = rtg::crand(spec->seed, rtg::kStreamQuery, 3 * (uint64_t)i + 1);
        const uint64_t r2 = rtg::crand(spec->seed, rtg::kStreamQuery, 3 * (uint64_t)i + 2);
        if (rtg::top53(r1) < tp) {
            kinds[i] = RTG_REQ_PAIR;
            a[i] = b[i] = 0;
            if (m > 0) {
                a[i] = (int32_t)(std::upper_bound(hoff.begin(), hoff.end(), (int32_t)((r2 >> 32) % (uint64_t)m)) - hoff.begin()) - 1;
                b[i] = (int32_t)(std::upper_bound(hoff.begin(), hoff.end(), (int32_t)((r2 & 0xFFFFFFFFull) % (uint64_t)m)) - hoff.begin()) - 1;
            }
            key_off[i] = 0;
            nkeys[i] = 0;
        } else {
            kinds[i] = RTG_REQ_LOOKUP;
            nkeys[i] = 1 + (int32_t)rtg::bounded32(r2, (uint32_t)max_keys);
            key_off[i] = (int64_t)((r2 & 0xFFFFFFFFull) % (uint64_t)(pn - nkeys[i] + 1));
            a[i] = b[i] = 0;
        }
    }
    *pool_n = pn;
    return RTG_OK;
}

}  // extern "C"

namespace rtg {
// Loads this file's kernels onto the current device (an attribute query of one of them).
int preload_serve() {
    hipFuncAttributes attr;
    return hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&fill_zero_u32)) == hipSuccess ? RTG_OK : RTG_EHIP;
}
}  // namespace rtg
