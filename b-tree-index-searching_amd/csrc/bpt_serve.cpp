// Serving loop: a resident tree answers a stream of point / range requests
// with continuous batching (no reference counterpart; the reference runs one
// synchronous batch per process).
//
//  * `slots` HIP streams (default priority), each driven by its own host
//    thread. The threads take work from one FIFO queue, so the host work of
//    different launches (staging copies, HIP calls, waiting) runs in
//    parallel.
//  * A thread that takes a request at the head of the queue also takes the
//    requests of the same kind queued behind it, in order, passing over those
//    of the other kind (up to kFuseQueriesHost/Device queries and kFuseRequests
//    requests), and answers them with one staged copy-in, one search and one
//    copy-out; small requests then share the per-launch cost.
//  * Device memory for requests comes from one pool of `mem_budget` bytes
//    (first-fit, 256-byte granules). A request is admitted when a slot is
//    free AND its bytes fit in the pool (a fused group that does not fit
//    falls back to the head request alone); otherwise it waits in FIFO order,
//    so a small budget makes requests queue, never fail (a request larger
//    than the whole budget is rejected at submit).
//  * Admitted work overlaps on the device: copy-in, search kernels and
//    copy-out of different requests run on different streams.
//  * Pinned staging buffers (hipHostMalloc) per slot carry the host copies.
//  * Device requests (BPT_REQ_DEVICE) skip the host: one alone is searched
//    in place; a fused run is gathered into the pool by one descriptor copy
//    kernel, searched as one batch and scattered back by another. Host and
//    device requests are never fused together.
//  * Latency = completion time - submit time; stats give p50/p99.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bpt_internal.hpp"

namespace {

using clk = std::chrono::steady_clock;

// Queries of one fused launch. Host requests are bound by their staging
// copies, which each slot's thread makes alone: smaller launches spread them
// over the slots (3000 mixed requests, 17.2 M lookups: 2^18 serves 1.6x as
// many requests per second as 2^20). Device requests are bound by the
// per-launch host work, so they fuse more.
constexpr uint64_t kFuseQueriesHost = 1 << 18;
constexpr uint64_t kFuseQueriesDevice = 1 << 20;
constexpr size_t kFuseRequests = 512;        // requests of one fused point launch

struct Request {
    uint64_t ticket = 0;
    bpt_request req{};
    uint64_t bytes = 0;  // device bytes of the request alone
    clk::time_point t_submit, t_done;
    bool done = false;
    bpt_status status = BPT_OK;
};

struct Slot {
    hipStream_t stream = nullptr;
    uint8_t *pin = nullptr;  // pinned staging
    uint64_t pin_bytes = 0;
    std::vector<Request *> group;  // the requests of the launch in flight
    uint64_t off = 0, bytes = 0;   // their pool allocation
    std::thread worker;
};

// First-fit allocator over [0, size) in 256-byte granules.
struct Pool {
    uint64_t size = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    void init(uint64_t s) {
        size = s;
        free_.clear();
        free_[0] = s;
    }
    bool alloc(uint64_t n, uint64_t *off) {
        n = (n + 255) / 256 * 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, o = it->first;
                free_.erase(it);
                if (rest) free_[o + n] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = (n + 255) / 256 * 256;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

}  // namespace

struct bpt_server {
    bpt_tree *tree = nullptr;
    int mode = BPT_POINT_PLAIN;
    uint64_t budget = 0, used = 0, peak = 0;
    uint8_t *d_pool = nullptr;
    Pool pool;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::vector<size_t> pick;  // take_work's scratch: queue positions of the next launch
    std::unordered_map<uint64_t, Request *> reqs;
    uint64_t next_ticket = 1;
    bool stop = false;
    uint32_t inflight = 0, max_inflight = 0;
    std::vector<double> lat_ms;
    uint64_t lookups = 0, completed = 0;
    clk::time_point t_first, t_last;
    bool any = false;
};

namespace {

uint64_t al256(uint64_t b) { return (b + 255) / 256 * 256; }

// Copy pieces of a fused run of k device requests: points gather the
// queries and scatter the values (2k); ranges gather both bounds and scatter
// the counts and the lower bounds asked for (at most 4k).
uint64_t desc_count(int32_t kind, uint64_t k) { return kind == BPT_REQ_POINT ? 2 * k : 4 * k; }

// Pool bytes of one launch: k requests of one kind and one buffer kind, n
// queries in all. Host requests and fused device runs: the queries and
// results in the pool, then the point search's workspace, then (device) the
// copy descriptors. A device request alone is searched in place and needs
// the workspace only.
uint64_t group_bytes(const bpt_tree *t, int mode, int32_t kind, int32_t flags, uint64_t n, uint64_t k) {
    const uint64_t kb = t->key_bytes, vb = t->val_bytes;
    const uint64_t ws = kind == BPT_REQ_POINT ? al256(bpt_point_search_ws_size(t, n, mode)) : 0;
    const bool device = (flags & BPT_REQ_DEVICE) != 0;
    if (device && k == 1) return std::max<uint64_t>(ws, 256);
    uint64_t b = kind == BPT_REQ_POINT ? al256(n * kb) + al256(n * vb) + ws : 2 * al256(n * kb) + 2 * al256(n * 8);
    if (device) b += al256(desc_count(kind, k) * sizeof(bpt::CopyDesc));
    return b;
}

uint64_t request_bytes(const bpt_tree *t, int mode, const bpt_request &r) {
    return group_bytes(t, mode, r.kind, r.flags, r.n, 1);
}

uint64_t staging_bytes(const bpt_tree *t, const bpt_request &r) {
    const uint64_t kb = t->key_bytes, vb = t->val_bytes;
    return r.kind == BPT_REQ_POINT ? r.n * (kb + vb) : r.n * (2 * kb + 16);
}

// Grows the slot's pinned staging buffer to at least `need` bytes.
bpt_status ensure_pinned(Slot &sl, uint64_t need) {
    if (sl.pin_bytes >= need) return BPT_OK;
    if (sl.pin) (void)hipHostFree(sl.pin);
    sl.pin = nullptr;
    sl.pin_bytes = 0;
    const uint64_t bytes = std::max<uint64_t>(need, 1 << 20);
    BPT_HIP(hipHostMalloc(&sl.pin, bytes, hipHostMallocDefault));
    sl.pin_bytes = bytes;
    return BPT_OK;
}

// Queries of the slot's group (all of one kind).
uint64_t group_queries(const Slot &sl) {
    uint64_t n = 0;
    for (const Request *q : sl.group) n += q->req.n;
    return n;
}

// A run of device requests. Alone: searched in place. Fused: the pool holds
// [queries][results][search workspace][descriptors] (points) or
// [lo][hi][first][count][descriptors] (ranges); the descriptors go up
// through the slot's pinned buffer, one copy kernel gathers the requests'
// inputs, the search runs over the whole run, and another copy kernel
// scatters the results to the requests' own buffers.
bpt_status launch_device(bpt_server *S, Slot &sl) {
    using bpt::CopyDesc;
    const bpt_tree *t = S->tree;
    const uint64_t kb = t->key_bytes, vb = t->val_bytes;
    const bool point = sl.group[0]->req.kind == BPT_REQ_POINT;
    const uint64_t n = group_queries(sl), k = sl.group.size();
    uint8_t *dev = S->d_pool + sl.off;
    hipStream_t s = sl.stream;
    const uint64_t wsb = point ? bpt_point_search_ws_size(t, n, S->mode) : 0;
    if (k == 1) {
        const bpt_request &r = sl.group[0]->req;
        if (point) return bpt_point_search_ws(t, r.q, n, r.out, S->mode, wsb ? dev : nullptr, wsb, s);
        return bpt_range_count(t, r.lo, r.hi, n, r.first, r.count, reinterpret_cast<bpt_stream>(s));
    }
    BPT_TRY(ensure_pinned(sl, desc_count(sl.group[0]->req.kind, k) * sizeof(CopyDesc)));
    CopyDesc *h = reinterpret_cast<CopyDesc *>(sl.pin);
    uint64_t ng = 0, ns = 0, max_g = 0, max_s = 0;
    auto gather = [&](const void *src, uint8_t *dst, uint64_t bytes) {
        h[ng++] = CopyDesc{src, dst, bytes};
        max_g = std::max(max_g, bytes);
    };
    std::vector<CopyDesc> sc;  // scatter pieces, placed after the gather pieces
    sc.reserve(2 * k);
    auto scatter = [&](const uint8_t *src, void *dst, uint64_t bytes) {
        sc.push_back(CopyDesc{src, dst, bytes});
        max_s = std::max(max_s, bytes);
    };
    uint64_t at = 0;
    if (point) {
        uint8_t *dq = dev, *dout = dev + al256(n * kb), *ws = dout + al256(n * vb);
        CopyDesc *dd = reinterpret_cast<CopyDesc *>(ws + al256(wsb));
        for (const Request *q : sl.group) {
            gather(q->req.q, dq + at * kb, q->req.n * kb);
            scatter(dout + at * vb, q->req.out, q->req.n * vb);
            at += q->req.n;
        }
        for (const CopyDesc &c : sc) h[ng + ns++] = c;
        BPT_HIP(hipMemcpyAsync(dd, h, (ng + ns) * sizeof(CopyDesc), hipMemcpyHostToDevice, s));
        BPT_TRY(bpt::launch_desc_copy(dd, (uint32_t)ng, max_g, s));
        BPT_TRY(bpt_point_search_ws(t, dq, n, dout, S->mode, wsb ? ws : nullptr, wsb, s));
        BPT_TRY(bpt::launch_desc_copy(dd + ng, (uint32_t)ns, max_s, s));
    } else {
        uint8_t *dlo = dev, *dhi = dev + al256(n * kb), *df = dhi + al256(n * kb), *dc = df + al256(n * 8);
        CopyDesc *dd = reinterpret_cast<CopyDesc *>(dc + al256(n * 8));
        for (const Request *q : sl.group) {
            const bpt_request &r = q->req;
            gather(r.lo, dlo + at * kb, r.n * kb);
            gather(r.hi, dhi + at * kb, r.n * kb);
            if (r.first) scatter(df + at * 8, r.first, r.n * 8);
            scatter(dc + at * 8, r.count, r.n * 8);
            at += r.n;
        }
        for (const CopyDesc &c : sc) h[ng + ns++] = c;
        BPT_HIP(hipMemcpyAsync(dd, h, (ng + ns) * sizeof(CopyDesc), hipMemcpyHostToDevice, s));
        BPT_TRY(bpt::launch_desc_copy(dd, (uint32_t)ng, max_g, s));
        BPT_TRY(bpt_range_count(t, dlo, dhi, n, reinterpret_cast<uint64_t *>(df), reinterpret_cast<uint64_t *>(dc),
                                reinterpret_cast<bpt_stream>(s)));
        BPT_TRY(bpt::launch_desc_copy(dd + ng, (uint32_t)ns, max_s, s));
    }
    return BPT_OK;
}

// Stages, launches and queues the copy-out of the slot's group, in group
// order. Points: staging = [queries][results], device = [queries][results]
// [search workspace]. Ranges: staging = [lo][hi][first][count], device the
// same. Device requests go to launch_device.
bpt_status launch(bpt_server *S, Slot &sl) {
    const bpt_tree *t = S->tree;
    const uint64_t kb = t->key_bytes, vb = t->val_bytes;
    const bpt_request &r0 = sl.group[0]->req;
    if (r0.flags & BPT_REQ_DEVICE) return launch_device(S, sl);
    const bool point = r0.kind == BPT_REQ_POINT;
    const uint64_t n = group_queries(sl);
    bpt_request whole = r0;
    whole.n = n;
    BPT_TRY(ensure_pinned(sl, staging_bytes(t, whole)));
    uint8_t *dev = S->d_pool + sl.off;
    auto al = al256;
    hipStream_t s = sl.stream;
    if (point) {
        uint64_t at = 0;
        for (const Request *q : sl.group) {
            memcpy(sl.pin + at * kb, q->req.q, q->req.n * kb);
            at += q->req.n;
        }
        uint8_t *dq = dev, *dout = dev + al(n * kb), *ws = dout + al(n * vb);
        BPT_HIP(hipMemcpyAsync(dq, sl.pin, n * kb, hipMemcpyHostToDevice, s));
        const uint64_t wsb = bpt_point_search_ws_size(t, n, S->mode);
        BPT_TRY(bpt_point_search_ws(t, dq, n, dout, S->mode, wsb ? ws : nullptr, wsb, s));
        BPT_HIP(hipMemcpyAsync(sl.pin + n * kb, dout, n * vb, hipMemcpyDeviceToHost, s));
    } else {
        uint64_t at = 0;
        for (const Request *q : sl.group) {
            memcpy(sl.pin + at * kb, q->req.lo, q->req.n * kb);
            memcpy(sl.pin + (n + at) * kb, q->req.hi, q->req.n * kb);
            at += q->req.n;
        }
        uint8_t *dlo = dev, *dhi = dev + al(n * kb), *df = dhi + al(n * kb), *dc = df + al(n * 8);
        BPT_HIP(hipMemcpyAsync(dlo, sl.pin, n * kb, hipMemcpyHostToDevice, s));
        BPT_HIP(hipMemcpyAsync(dhi, sl.pin + n * kb, n * kb, hipMemcpyHostToDevice, s));
        BPT_TRY(bpt_range_count(t, dlo, dhi, n, reinterpret_cast<uint64_t *>(df), reinterpret_cast<uint64_t *>(dc),
                                reinterpret_cast<bpt_stream>(s)));
        BPT_HIP(hipMemcpyAsync(sl.pin + 2 * n * kb, df, n * 8, hipMemcpyDeviceToHost, s));
        BPT_HIP(hipMemcpyAsync(sl.pin + 2 * n * kb + n * 8, dc, n * 8, hipMemcpyDeviceToHost, s));
    }
    return BPT_OK;
}

// Copies a finished group's results to the callers' buffers (no lock held).
void copy_out(const bpt_server *S, const Slot &sl, bpt_status st) {
    if (st != BPT_OK || (sl.group[0]->req.flags & BPT_REQ_DEVICE)) return;
    const uint64_t kb = S->tree->key_bytes, vb = S->tree->val_bytes;
    const bool point = sl.group[0]->req.kind == BPT_REQ_POINT;
    const uint64_t n = group_queries(sl);
    uint64_t at = 0;
    for (const Request *q : sl.group) {
        const bpt_request &r = q->req;
        if (point) {
            memcpy(r.out, sl.pin + n * kb + at * vb, r.n * vb);
        } else {
            if (r.first) memcpy(r.first, sl.pin + 2 * n * kb + at * 8, r.n * 8);
            memcpy(r.count, sl.pin + 2 * n * kb + (n + at) * 8, r.n * 8);
        }
        at += r.n;
    }
}

// Takes the next launch off the queue (S->mu held): the head request and
// the requests of its kind and buffer kind queued behind it, in queue order
// (others in between are passed over, not reordered), as far as the fusion
// limits allow; the group's device bytes must fit the pool, else the head
// alone is tried. Returns false when nothing can be admitted now.
bool take_work(bpt_server *S, Slot &sl) {
    if (S->queue.empty()) return false;
    const Request *h = S->queue.front();
    std::vector<size_t> &pick = S->pick;
    pick.assign(1, 0);
    const uint64_t cap = (h->req.flags & BPT_REQ_DEVICE) ? kFuseQueriesDevice : kFuseQueriesHost;
    uint64_t n = h->req.n;
    for (size_t i = 1; i < S->queue.size() && pick.size() < kFuseRequests; ++i) {
        const Request *r = S->queue[i];
        if (r->req.kind != h->req.kind || r->req.flags != h->req.flags) continue;
        if (n + r->req.n > cap) break;  // the kind's order is kept: no later one jumps this one
        n += r->req.n;
        pick.push_back(i);
    }
    uint64_t off = 0, bytes = 0;
    if (pick.size() > 1) {
        bytes = group_bytes(S->tree, S->mode, h->req.kind, h->req.flags, n, pick.size());
        if (!S->pool.alloc(bytes, &off)) pick.resize(1);
    }
    if (pick.size() == 1) {
        bytes = h->bytes;
        if (!S->pool.alloc(bytes, &off)) return false;
    }
    sl.group.clear();
    for (size_t i : pick) sl.group.push_back(S->queue[i]);
    if (pick.size() == 1) {
        S->queue.pop_front();
    } else {  // the rest of the queue, in order
        size_t j = 0, w = 0;
        for (size_t i = 0; i < S->queue.size(); ++i) {
            if (j < pick.size() && pick[j] == i) ++j;
            else S->queue[w++] = S->queue[i];
        }
        S->queue.resize(w);
    }
    sl.off = off;
    sl.bytes = bytes;
    return true;
}

// Bookkeeping of a finished group (S->mu held).
void finish(bpt_server *S, Slot &sl, bpt_status st) {
    const clk::time_point now = clk::now();
    S->pool.release(sl.off, sl.bytes);
    S->used -= (sl.bytes + 255) / 256 * 256;
    for (Request *q : sl.group) {
        q->t_done = now;
        q->status = st;
        q->done = true;
        S->inflight--;
        S->lat_ms.push_back(std::chrono::duration<double, std::milli>(q->t_done - q->t_submit).count());
        S->lookups += q->req.n;
        S->completed++;
    }
    if (!S->any || now > S->t_last) S->t_last = now;
    sl.group.clear();
}

// One thread per slot: admission is FIFO -- a thread takes the head of the
// queue (and the point requests fused with it) only when the bytes fit the
// pool, otherwise every thread waits for memory to be released -- then
// stages, launches and waits on its own stream.
void slot_worker(bpt_server *S, Slot *sl) {
    (void)hipSetDevice(S->tree->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        S->cv_work.wait(lk, [&] { return (S->stop && S->queue.empty()) || take_work(S, *sl); });
        if (sl->group.empty()) break;  // stopping
        S->used += (sl->bytes + 255) / 256 * 256;
        S->peak = std::max(S->peak, S->used);
        S->inflight += (uint32_t)sl->group.size();
        S->max_inflight = std::max(S->max_inflight, S->inflight);
        if (!S->queue.empty()) S->cv_work.notify_one();  // the next head may fit too
        lk.unlock();
        bpt_status st = launch(S, *sl);
        const hipError_t e = hipStreamSynchronize(sl->stream);
        if (st == BPT_OK && e != hipSuccess) st = bpt::hip_fail(e, "request", __FILE__, __LINE__);
        copy_out(S, *sl, st);
        lk.lock();
        finish(S, *sl, st);
        S->cv_done.notify_all();
        S->cv_work.notify_all();  // memory was released
    }
}

// Validates r and makes its Request (nullptr, with the error set, when r is
// rejected).
Request *make_request(bpt_server *S, const bpt_request *r, bpt_status *st) {
    *st = BPT_ERR_INVALID;
    if (!r) {
        bpt::set_error("request is NULL");
        return nullptr;
    }
    if (r->kind != BPT_REQ_POINT && r->kind != BPT_REQ_RANGE) {
        bpt::set_error("request kind %d", r->kind);
        return nullptr;
    }
    if (r->n < 1) {
        bpt::set_error("empty request");
        return nullptr;
    }
    if (r->flags != 0 && r->flags != BPT_REQ_DEVICE) {
        bpt::set_error("request flags %d", r->flags);
        return nullptr;
    }
    if (r->kind == BPT_REQ_POINT ? !(r->q && r->out) : !(r->lo && r->hi && r->count)) {
        bpt::set_error("request buffers are NULL");
        return nullptr;
    }
    if (r->flags & BPT_REQ_DEVICE) {
        const uint64_t kb = S->tree->key_bytes, vb = S->tree->val_bytes;
        auto ok = [](const void *p, uint64_t a) { return reinterpret_cast<uintptr_t>(p) % a == 0; };
        const bool aligned = r->kind == BPT_REQ_POINT
                                 ? ok(r->q, kb) && ok(r->out, vb)
                                 : ok(r->lo, kb) && ok(r->hi, kb) && ok(r->first, 8) && ok(r->count, 8);
        if (!aligned) {
            bpt::set_error("device request buffers not aligned to their element widths");
            return nullptr;
        }
    }
    const uint64_t bytes = request_bytes(S->tree, S->mode, *r);
    if (bytes > S->budget) {
        bpt::set_error("request needs %llu bytes, more than the whole budget", (unsigned long long)bytes);
        *st = BPT_ERR_CAPACITY;
        return nullptr;
    }
    auto *q = new Request;
    q->req = *r;
    q->bytes = bytes;
    *st = BPT_OK;
    return q;
}

// Queues made requests under one lock, in order, and wakes the workers. A
// request's submit time is now unless it already has one (a replayed
// request's scheduled arrival).
void enqueue(bpt_server *S, Request *const *qs, size_t n) {
    if (n == 0) return;
    const clk::time_point now = clk::now();
    {
        std::lock_guard<std::mutex> g(S->mu);
        for (size_t i = 0; i < n; ++i) {
            Request *q = qs[i];
            if (q->t_submit == clk::time_point{}) q->t_submit = now;
            q->ticket = S->next_ticket++;
            if (!S->any || q->t_submit < S->t_first) {
                S->t_first = q->t_submit;
                S->any = true;
            }
            S->reqs[q->ticket] = q;
            S->queue.push_back(q);
        }
    }
    if (n == 1) S->cv_work.notify_one();
    else S->cv_work.notify_all();
}

}  // namespace

using namespace bpt;

extern "C" {

bpt_status bpt_server_create(bpt_tree *t, uint32_t slots, uint64_t mem_budget, int mode, bpt_server **out) {
    BPT_CHECK(t && out, BPT_ERR_INVALID, "NULL argument");
    BPT_CHECK(slots >= 1 && slots <= 64, BPT_ERR_INVALID, "slots %u not in [1, 64]", slots);
    BPT_CHECK(mem_budget >= 4096, BPT_ERR_INVALID, "memory budget too small");
    BPT_CHECK(mode == BPT_POINT_PLAIN || mode == BPT_POINT_GROUPED || mode == BPT_POINT_LEARNED, BPT_ERR_INVALID,
              "mode %d", mode);
    BPT_CHECK(mode != BPT_POINT_LEARNED || t->learned, BPT_ERR_UNSUPPORTED, "tree has no learned model");
    auto *S = new bpt_server;
    S->tree = t;
    S->mode = mode;
    S->budget = (mem_budget + 255) / 256 * 256;
    hipError_t e = hipMalloc(&S->d_pool, S->budget);
    if (e != hipSuccess) {
        delete S;
        return hip_fail(e, "hipMalloc(server pool)", __FILE__, __LINE__);
    }
    S->pool.init(S->budget);
    S->slots.resize(slots);
    for (Slot &sl : S->slots) {
        e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            bpt_server_destroy(S);
            return hip_fail(e, "server stream", __FILE__, __LINE__);
        }
    }
    for (Slot &sl : S->slots) sl.worker = std::thread(slot_worker, S, &sl);
    *out = S;
    return BPT_OK;
}

bpt_status bpt_server_submit(bpt_server *S, const bpt_request *r, uint64_t *ticket) {
    BPT_CHECK(S && r && ticket, BPT_ERR_INVALID, "NULL argument");
    bpt_status st;
    Request *q = make_request(S, r, &st);
    if (!q) return st;
    enqueue(S, &q, 1);
    *ticket = q->ticket;
    return BPT_OK;
}

bpt_status bpt_server_wait(bpt_server *S, uint64_t ticket, double *latency_ms) {
    BPT_CHECK(S, BPT_ERR_INVALID, "server is NULL");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    BPT_CHECK(it != S->reqs.end(), BPT_ERR_NOT_FOUND, "unknown ticket %llu", (unsigned long long)ticket);
    Request *q = it->second;
    S->cv_done.wait(lk, [&] { return q->done; });
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(q->t_done - q->t_submit).count();
    const bpt_status st = q->status;
    S->reqs.erase(it);
    delete q;
    return st;
}

bpt_status bpt_server_get_stats(bpt_server *S, bpt_server_stats *st) {
    BPT_CHECK(S && st, BPT_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> g(S->mu);
    memset(st, 0, sizeof *st);
    st->completed = S->completed;
    st->lookups = S->lookups;
    st->max_inflight = S->max_inflight;
    st->peak_bytes = S->peak;
    st->budget_bytes = S->budget;
    if (!S->lat_ms.empty()) {
        std::vector<double> v = S->lat_ms;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) { return v[std::min<size_t>(v.size() - 1, (size_t)(p * (v.size() - 1) + 0.5))]; };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        st->max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        st->mean_ms = sum / v.size();
    }
    if (S->any) st->elapsed_s = std::chrono::duration<double>(S->t_last - S->t_first).count();
    return BPT_OK;
}

bpt_status bpt_server_reset_stats(bpt_server *S) {
    BPT_CHECK(S, BPT_ERR_INVALID, "server is NULL");
    std::lock_guard<std::mutex> g(S->mu);
    S->lat_ms.clear();
    S->lookups = S->completed = 0;
    S->max_inflight = S->inflight;
    S->peak = S->used;
    S->any = false;
    return BPT_OK;
}

bpt_status bpt_server_destroy(bpt_server *S) {
    if (!S) return BPT_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (Slot &sl : S->slots)
        if (sl.worker.joinable()) sl.worker.join();
    for (auto &kv : S->reqs) delete kv.second;
    for (Slot &sl : S->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.pin) (void)hipHostFree(sl.pin);
    }
    (void)hipFree(S->d_pool);
    delete S;
    return BPT_OK;
}

// Open-loop replay: request i is submitted at arrival_s[i] seconds after the
// start (busy-waiting for its time, independent of completions; the requests
// already due are queued together), then every request is waited for;
// latency_ms[i] = completion - submit.
bpt_status bpt_server_replay(bpt_server *S, const bpt_request *reqs, uint32_t nreq, const double *arrival_s,
                             double *latency_ms) {
    BPT_CHECK(S && (nreq == 0 || (reqs && arrival_s)), BPT_ERR_INVALID, "NULL argument");
    std::vector<uint64_t> tickets(nreq, 0);
    const clk::time_point t0 = clk::now();
    bpt_status st = BPT_OK;
    auto due = [&](uint32_t i) {
        return t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(arrival_s[i]));
    };
    std::vector<Request *> batch;
    std::vector<uint32_t> index;  // request number of each batch entry
    for (uint32_t i = 0; i < nreq;) {
        while (clk::now() < due(i)) std::this_thread::yield();
        // every request already due is queued at once, under one lock
        const clk::time_point now = clk::now();
        const uint32_t first = i;
        batch.clear();
        index.clear();
        for (; i < nreq && (i == first || due(i) <= now); ++i) {
            bpt_status s;
            Request *q = make_request(S, &reqs[i], &s);
            if (q) {
                q->t_submit = due(i);  // latency counts from the scheduled arrival
                batch.push_back(q);
                index.push_back(i);
            } else if (st == BPT_OK) {
                st = s;
            }
        }
        enqueue(S, batch.data(), batch.size());
        for (size_t b = 0; b < batch.size(); ++b) tickets[index[b]] = batch[b]->ticket;
    }
    for (uint32_t i = 0; i < nreq; ++i) {
        if (!tickets[i]) continue;
        double ms = 0;
        bpt_status s = bpt_server_wait(S, tickets[i], &ms);
        if (latency_ms) latency_ms[i] = ms;
        if (s != BPT_OK && st == BPT_OK) st = s;
    }
    return st;
}

}  // extern "C"
