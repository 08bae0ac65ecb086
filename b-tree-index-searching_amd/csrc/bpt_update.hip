// Batched updates of a built tree: upsert (key, value) pairs, erase keys.
//
// No reference counterpart: the reference's pages carry previous/next links
// and a bump allocator (BPTree.h:14-23, BPTree.cu:164) that suggest updates,
// but it never changes a built tree. Here a batch is merged with the tree's
// sorted slots on the device and the flat tree is rebuilt from the merge
// (one level-parallel build, a few milliseconds at 64M keys), so every search
// form keeps its layout and its speed:
//   1. the tree's slots are read out of the leaves in key order;
//   2. the batch is copied and stably sorted by key; for upserts only the last
//      pair of each run of equal keys is kept (the last in the caller's order);
//   3. each batch key finds its place among the slots by binary search; the
//      slots holding it are removed, and an upserted key is counted at its
//      place. A surviving slot's place in the result is its rank among the
//      surviving slots plus the number of upserted keys below it (a scan of
//      those counts), and an upserted pair's place is its rank in the batch
//      plus the number of surviving slots below its place (a merge with no
//      ties left);
//   4. the tree is rebuilt from the result with its own options.
#include <algorithm>

#include "bpt_internal.hpp"

namespace bpt {

template <class K, class V, int LF>
__global__ __launch_bounds__(256) void k_extract_slots(TreeView t, K *__restrict__ keys, V *__restrict__ vals) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += step) {
        keys[i] = slot_key<K, V, LF>(t, i);
        vals[i] = slot_val<K, V, LF>(t, i);
    }
}

// flag[j] = 1 for the last pair of every run of equal keys in sorted b.
template <class K>
__global__ __launch_bounds__(256) void k_run_last(const K *__restrict__ b, uint64_t m, uint32_t *__restrict__ flag) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += step)
        flag[j] = (j + 1 == m || b[j] != b[j + 1]) ? 1u : 0u;
}

template <class K, class V>
__global__ __launch_bounds__(256) void k_compact(const K *__restrict__ k, const V *__restrict__ v,
                                                 const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
                                                 uint64_t m, K *__restrict__ ko, V *__restrict__ vo) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += step)
        if (flag[j]) {
            ko[pos[j]] = k[j];
            if (vo) vo[pos[j]] = v[j];
        }
}

template <class K>
__device__ __forceinline__ uint64_t lower_bound_g(const K *__restrict__ a, uint64_t n, K q) {
    uint64_t lo = 0, len = n;
    while (len > 0) {
        const uint64_t half = len >> 1;
        if (a[lo + half] < q) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *__restrict__ a, uint64_t n, uint32_t v) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) a[i] = v;
}

// Batch key j finds its place among the tree's sorted slots a: q[j] =
// lower_bound(a, b[j]) (kept when q is not NULL). The slots holding b[j] are
// removed (keep = 0).
template <class K>
__global__ __launch_bounds__(256) void k_batch_marks(const K *__restrict__ a, uint64_t n, const K *__restrict__ b,
                                                     uint64_t m, uint32_t *__restrict__ keep, uint32_t *__restrict__ q) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += step) {
        const K key = b[j];
        const uint64_t p = lower_bound_g<K>(a, n, key);
        for (uint64_t r = p; r < n && a[r] == key; ++r) keep[r] = 0;
        if (q) q[j] = (uint32_t)p;
    }
}

// below[p] = number of upserted keys whose place is p. The batch is sorted, so
// q is non-decreasing: the last pair of each run of equal places writes the
// run's length (no atomics, whatever the number of keys at one place).
__global__ __launch_bounds__(256) void k_batch_runs(const uint32_t *__restrict__ q, uint64_t m,
                                                    uint32_t *__restrict__ below) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += step) {
        const uint32_t v = q[j];
        if (j + 1 < m && q[j + 1] == v) continue;
        below[v] = (uint32_t)(j + 1 - lower_bound_g<uint32_t>(q, j + 1, v));
    }
}

// A surviving slot goes after the surviving slots and the upserted keys
// below it: rank[i] + (excl[i] + below[i]), excl = exclusive scan of below.
template <class K, class V>
__global__ __launch_bounds__(256) void k_place_slots(const K *__restrict__ a, const V *__restrict__ av, uint64_t n,
                                                     const uint32_t *__restrict__ keep, const uint32_t *__restrict__ rank,
                                                     const uint32_t *__restrict__ below,
                                                     const uint32_t *__restrict__ excl, K *__restrict__ ko,
                                                     V *__restrict__ vo) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step)
        if (keep[i]) {
            const uint64_t p = (uint64_t)rank[i] + (below ? (uint64_t)excl[i] + below[i] : 0);
            ko[p] = a[i];
            vo[p] = av[i];
        }
}

// Upserted pair j goes after every surviving slot with a smaller key: the
// surviving slots before its place q[j].
template <class K, class V>
__global__ __launch_bounds__(256) void k_place_batch(const K *__restrict__ b, const V *__restrict__ bv, uint64_t m,
                                                     const uint32_t *__restrict__ q, uint64_t n,
                                                     const uint32_t *__restrict__ rank, uint64_t survivors,
                                                     K *__restrict__ ko, V *__restrict__ vo) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += step) {
        const uint64_t qj = q[j];
        const uint64_t p = j + (qj < n ? (uint64_t)rank[qj] : survivors);
        ko[p] = b[j];
        vo[p] = bv[j];
    }
}

namespace {

// Bump allocator over one device buffer (the tree's workspace), 256-byte
// aligned pieces.
struct Arena {
    uint8_t *base = nullptr;
    uint64_t cap = 0, used = 0;
    static uint64_t bytes(uint64_t n) { return (std::max<uint64_t>(n, 1) + 255) & ~255ull; }
    template <class T>
    bpt_status get(uint64_t count, T **out) {
        const uint64_t b = bytes(count * sizeof(T));
        BPT_CHECK(used + b <= cap, BPT_ERR_INVALID, "update workspace too small");
        *out = reinterpret_cast<T *>(base + used);
        used += b;
        return BPT_OK;
    }
};

uint32_t grid_for(uint64_t n) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cu_count() * 16));
}

// Exclusive scan of n u32 flags into rank; returns the total in *total.
bpt_status scan_flags(const uint32_t *flag, uint64_t n, uint32_t *rank, uint64_t *total, Arena &ar, hipStream_t s) {
    *total = 0;
    if (n == 0) return BPT_OK;
    const uint64_t tb = scan_tmp_bytes(n, 4);
    uint8_t *tmp = nullptr;
    BPT_TRY(ar.get<uint8_t>(tb, &tmp));
    BPT_TRY(exclusive_scan<uint32_t>(flag, n, rank, tmp, tb, s));
    uint32_t last_rank = 0, last_flag = 0;
    BPT_HIP(hipMemcpyAsync(&last_rank, rank + n - 1, 4, hipMemcpyDeviceToHost, s));
    BPT_HIP(hipMemcpyAsync(&last_flag, flag + n - 1, 4, hipMemcpyDeviceToHost, s));
    BPT_HIP(hipStreamSynchronize(s));
    *total = (uint64_t)last_rank + last_flag;
    return BPT_OK;
}

// Workspace bytes of one update of an n-slot tree by an m-pair batch (an
// upper bound of what update_impl takes from its arena).
template <class K, class V>
uint64_t update_ws_bytes(uint64_t n, uint64_t m) {
    const auto B = Arena::bytes;
    return B(n * sizeof(K)) + B(n * sizeof(V))                              // the slots
           + B(sort_ws_bytes(m, sizeof(K))) + B(m * sizeof(V))              // sorted batch
           + 2 * B(m * 4) + B(scan_tmp_bytes(m, 4)) + B(m * sizeof(K)) + B(m * sizeof(V))  // dedupe
           + 2 * B(n * 4) + 2 * B((n + 1) * 4) + B(m * 4)                   // marks, ranks, places
           + B(scan_tmp_bytes(n, 4)) + B(scan_tmp_bytes(n + 1, 4))
           + B((n + m) * sizeof(K)) + B((n + m) * sizeof(V));               // the merge
}

template <class K, class V, int LF>
bpt_status update_impl(bpt_tree *t, const void *d_keys, const void *d_vals, uint64_t m, bool upsert, hipStream_t s) {
    const uint64_t n = t->n;
    const uint64_t need = update_ws_bytes<K, V>(n, m);
    BPT_TRY(t->ws.reserve(need));
    Arena ar{reinterpret_cast<uint8_t *>(t->ws.ptr), need, 0};
    // 1. the tree's slots in key order
    K *a = nullptr;
    V *av = nullptr;
    BPT_TRY(ar.get<K>(n, &a));
    BPT_TRY(ar.get<V>(n, &av));
    if (n) {
        k_extract_slots<K, V, LF><<<grid_for(n), 256, 0, s>>>(t->view, a, av);
        BPT_HIP(hipGetLastError());
    }
    // 2. the batch, sorted (stable LSD radix sort), deduplicated for upserts
    K *bk = nullptr;
    V *bv = nullptr;  // erase: keys only
    {
        const uint64_t sb = sort_ws_bytes(m, sizeof(K));
        uint8_t *sws = nullptr;
        BPT_TRY(ar.get<uint8_t>(sb, &sws));
        uint32_t *perm = nullptr;
        BPT_TRY(sort_pairs<K>((const K *)d_keys, nullptr, m, 0, 8 * sizeof(K), sws, sb, &bk, &perm, s));
        if (upsert) {
            BPT_TRY(ar.get<V>(m, &bv));
            BPT_TRY(bpt_gather(d_vals, sizeof(V), perm, m, bv, reinterpret_cast<bpt_stream>(s)));
        }
    }
    uint64_t mb = m;
    if (upsert && m > 1) {
        uint32_t *flag = nullptr, *pos = nullptr;
        BPT_TRY(ar.get<uint32_t>(m, &flag));
        BPT_TRY(ar.get<uint32_t>(m, &pos));
        k_run_last<K><<<grid_for(m), 256, 0, s>>>(bk, m, flag);
        BPT_HIP(hipGetLastError());
        BPT_TRY(scan_flags(flag, m, pos, &mb, ar, s));
        K *ck = nullptr;
        V *cv = nullptr;
        BPT_TRY(ar.get<K>(mb, &ck));
        BPT_TRY(ar.get<V>(mb, &cv));
        k_compact<K, V><<<grid_for(m), 256, 0, s>>>(bk, bv, flag, pos, m, ck, cv);
        BPT_HIP(hipGetLastError());
        bk = ck;
        bv = cv;
    }
    // 3. survivors and the merge
    uint32_t *keep = nullptr, *rank = nullptr, *below = nullptr, *excl = nullptr, *q = nullptr;
    BPT_TRY(ar.get<uint32_t>(n, &keep));
    BPT_TRY(ar.get<uint32_t>(n, &rank));
    const bool counts = upsert && n && mb;
    if (counts) {
        BPT_TRY(ar.get<uint32_t>(n + 1, &below));
        BPT_TRY(ar.get<uint32_t>(n + 1, &excl));
        BPT_TRY(ar.get<uint32_t>(mb, &q));
    }
    uint64_t survivors = 0;
    if (n) {
        k_fill_u32<<<grid_for(n), 256, 0, s>>>(keep, n, 1u);
        if (counts) k_fill_u32<<<grid_for(n + 1), 256, 0, s>>>(below, n + 1, 0u);
        k_batch_marks<K><<<grid_for(mb), 256, 0, s>>>(a, n, bk, mb, keep, q);
        if (counts) k_batch_runs<<<grid_for(mb), 256, 0, s>>>(q, mb, below);
        BPT_HIP(hipGetLastError());
        BPT_TRY(scan_flags(keep, n, rank, &survivors, ar, s));
        if (counts) {
            const uint64_t tb = scan_tmp_bytes(n + 1, 4);
            uint8_t *tmp = nullptr;
            BPT_TRY(ar.get<uint8_t>(tb, &tmp));
            BPT_TRY(exclusive_scan<uint32_t>(below, n + 1, excl, tmp, tb, s));
        }
    }
    const uint64_t total = survivors + (upsert ? mb : 0);
    K *ok = nullptr;
    V *ov = nullptr;
    BPT_TRY(ar.get<K>(total, &ok));
    BPT_TRY(ar.get<V>(total, &ov));
    if (n) {
        // erased keys take no places: only upserted keys count below a slot
        k_place_slots<K, V><<<grid_for(n), 256, 0, s>>>(a, av, n, keep, rank, counts ? below : nullptr, excl, ok,
                                                        ov);
        BPT_HIP(hipGetLastError());
    }
    if (upsert && mb) {
        if (n) {
            k_place_batch<K, V><<<grid_for(mb), 256, 0, s>>>(bk, bv, mb, q, n, rank, survivors, ok, ov);
            BPT_HIP(hipGetLastError());
        } else {
            BPT_HIP(hipMemcpyAsync(ok, bk, mb * sizeof(K), hipMemcpyDeviceToDevice, s));
            BPT_HIP(hipMemcpyAsync(ov, bv, mb * sizeof(V), hipMemcpyDeviceToDevice, s));
        }
    }
    // 4. rebuild with the tree's options (the old leaves are no longer read)
    return tree_rebuild(t, ok, ov, total, s);
}

bpt_status update(bpt_tree *t, const void *d_keys, const void *d_vals, uint64_t m, bool upsert, bpt_stream stream) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    BPT_CHECK(m == 0 || d_keys, BPT_ERR_INVALID, "keys are NULL");
    BPT_CHECK(!upsert || m == 0 || d_vals, BPT_ERR_INVALID, "values are NULL");
    BPT_CHECK(t->n + m < (1ull << 32), BPT_ERR_CAPACITY, "updated tree would exceed 2^32 - 1 keys");
    if (m == 0) return BPT_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_tree(t, [&](auto kt, auto vt, auto, auto lc) -> bpt_status {
        using K = typename decltype(kt)::type;
        using V = typename decltype(vt)::type;
        return update_impl<K, V, decltype(lc)::value>(t, d_keys, d_vals, m, upsert, s);
    });
}

}  // namespace
}  // namespace bpt

using namespace bpt;

extern "C" {

bpt_status bpt_tree_upsert(bpt_tree *t, const void *d_keys, const void *d_vals, uint64_t n, bpt_stream s) {
    return update(t, d_keys, d_vals, n, true, s);
}

bpt_status bpt_tree_erase(bpt_tree *t, const void *d_keys, uint64_t n, bpt_stream s) {
    return update(t, d_keys, nullptr, n, false, s);
}

}  // extern "C"
