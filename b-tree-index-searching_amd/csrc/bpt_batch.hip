// Multi-tree batches: thousands of independent trees (8-byte keys and
// values) in one device arena, searched by ONE persistent launch.
//
// No reference counterpart (the reference drives a single tree, one
// synchronous launch per step). Design:
//  * forest build: every tree's flat layout (leaves of 8 pairs = one 128 B
//    line, internal separator levels root first) is placed in two arenas; two
//    batched kernels write all leaves and all separators of all trees, each
//    thread finding its tree by a binary search over per-tree offsets;
//  * work list: (tree, kind, [begin, end)) items of at most `chunk` queries,
//    point items and range items, built once per batch;
//  * persistent kernel: each block takes the next item from an atomic
//    counter, stages that tree's descriptor and LDS-resident top levels when
//    the tree changes, and walks the item's queries with the cooperative
//    8-lane descent; ranges take two lower bounds (lo and hi + 1), walked in
//    the same load instructions;
//  * the counter reset and the launch are captured into a hipGraph and
//    replayed (the reset is a kernel, not a memset node);
//  * an autotuner times fanout x LDS budget x chunk and keeps the fastest.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "bpt_internal.hpp"

namespace bpt {

struct ForestTreeDesc {  // per tree, host side
    uint64_t n, key_off, leaf_off, inner_off, leaf_slot_off, inner_slot_off;
};

struct WorkItem {
    uint32_t tree;
    uint32_t kind;  // 0 point, 1 range
    uint64_t begin, end;
};

}  // namespace bpt

struct bpt_forest {
    uint32_t ntrees = 0, fanout = 16, leaf_cap = 8, lds_bytes = 32768;
    uint32_t lanes = 8;  // lanes per query of the cooperative descent (8 or 4)
    uint8_t *d_leaves = nullptr, *d_inner = nullptr;
    bpt::TreeView *d_views = nullptr;
    std::vector<bpt::TreeView> views;
    std::vector<bpt::ForestTreeDesc> desc;
    uint64_t leaf_bytes = 0, inner_bytes = 0, total_keys = 0;
};

struct bpt_batch {
    bpt_forest *forest = nullptr;
    bpt::WorkItem *d_items = nullptr;
    uint32_t nitems = 0, chunk = 8192, blocks_per_cu = 8;
    uint32_t *d_counter = nullptr;  // kBatchXcds cursors, then kBatchXcds + 1 range bounds
    const uint64_t *q = nullptr, *lo = nullptr, *hi = nullptr;
    uint64_t *out = nullptr, *first = nullptr, *count = nullptr;
    uint64_t npoints = 0, nranges = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace bpt {

__device__ __forceinline__ uint32_t find_tree(const uint64_t *__restrict__ off, uint32_t ntrees, uint64_t x) {
    uint32_t lo = 0, hi = ntrees;  // largest t with off[t] <= x
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// All leaves of all trees: slot s of the arena -> (tree, local slot).
__global__ __launch_bounds__(256) void k_forest_leaves(const uint64_t *__restrict__ keys,
                                                       const uint64_t *__restrict__ vals,
                                                       const uint64_t *__restrict__ slot_off,
                                                       const uint64_t *__restrict__ key_off,
                                                       uint32_t ntrees, uint64_t nslots, uint8_t *__restrict__ leaves) {
    constexpr int LF = 8;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslots; s += step) {
        const uint32_t t = find_tree(slot_off, ntrees, s);
        const uint64_t local = s - slot_off[t];
        const uint64_t n = key_off[t + 1] - key_off[t];
        const uint64_t k = local < n ? keys[key_off[t] + local] : ~0ull;
        const uint64_t v = local < n ? vals[key_off[t] + local] : 0ull;
        uint8_t *lp = leaves + (s / LF) * 128;  // a leaf is one 128-byte line
        using Fmt = LeafFmt<uint64_t, uint64_t, LF>;
        *reinterpret_cast<uint64_t *>(lp + Fmt::key_off((uint32_t)(s % LF))) = k;
        *reinterpret_cast<uint64_t *>(lp + Fmt::val_off((uint32_t)(s % LF))) = v;
    }
}

// All separators of all trees. A tree's internal levels are described by its
// TreeView (counts, byte offsets inside its inner block).
template <int F>
__global__ __launch_bounds__(256) void k_forest_inner(const uint64_t *__restrict__ keys,
                                                      const uint64_t *__restrict__ sep_off,
                                                      const uint64_t *__restrict__ key_off,
                                                      const TreeView *__restrict__ views, uint32_t ntrees,
                                                      uint64_t nseps, uint64_t *__restrict__ inner_arena) {
    constexpr uint64_t LF = 8;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseps; s += step) {
        const uint32_t t = find_tree(sep_off, ntrees, s);
        const TreeView &v = views[t];
        const uint64_t local = s - sep_off[t];  // separator index inside the tree's inner block
        int k = v.height;                       // levels root first
        while (k > 1 && local >= v.off[k - 1] / 8) --k;
        const uint64_t child = local - v.off[k] / 8;
        uint64_t span = LF;
        for (int j = 1; j < k; ++j) span *= F;  // keys under one node of level k-1
        uint64_t sep = ~0ull;
        if (child < v.count[k - 1]) {
            const uint64_t n = v.n;
            const uint64_t end = (child + 1) * span;
            sep = keys[key_off[t] + (end < n ? end : n) - 1];
        }
        inner_arena[s] = sep;
    }
}

// The items (ordered by tree) are cut into one contiguous range per XCD,
// balanced by lookups, so that an XCD works through its own trees and their
// lines stay in its L2. Blocks are dispatched to the XCDs round-robin (block
// i on XCD i % 8): a block takes items from its XCD's range first, then from
// the others' (in turn from the next XCD), so no XCD idles while work remains.
constexpr uint32_t kBatchXcds = 8;

__global__ void k_reset_counter(uint32_t *c) {
    if (threadIdx.x < kBatchXcds) c[threadIdx.x] = 0;
}

// Next item for a block whose XCD is x, or nitems when every range is done.
__device__ __forceinline__ uint32_t next_item(uint32_t *cursor, const uint32_t *bound, uint32_t x) {
    for (uint32_t k = 0; k < kBatchXcds; ++k) {
        const uint32_t r = (x + k) % kBatchXcds;
        const uint32_t span = bound[r + 1] - bound[r];
        if (cursor[r] >= span) continue;  // a stale read only costs one wasted atomic below
        const uint32_t i = atomicAdd(&cursor[r], 1u);
        if (i < span) return bound[r] + i;
    }
    return 0xFFFFFFFFu;
}

template <int F, int G>
__global__ __launch_bounds__(256) void k_forest_batch(const TreeView *__restrict__ views,
                                                      const WorkItem *__restrict__ items, uint32_t nitems,
                                                      uint32_t *__restrict__ counter, const uint64_t *__restrict__ q,
                                                      uint64_t *__restrict__ out, const uint64_t *__restrict__ lo,
                                                      const uint64_t *__restrict__ hi, uint64_t *__restrict__ first,
                                                      uint64_t *__restrict__ count) {
    constexpr int LF = 8;
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    __shared__ TreeView tv;
    __shared__ uint32_t item_sh;
    uint32_t cur = 0xFFFFFFFFu;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t *bound = counter + kBatchXcds;
    for (;;) {
        if (threadIdx.x == 0) item_sh = next_item(counter, bound, blockIdx.x % kBatchXcds);
        __syncthreads();
        const uint32_t it = item_sh;
        if (it >= nitems) break;  // the same value in every thread of the block
        const WorkItem wi = items[it];
        if (wi.tree != cur) {
            const uint32_t *src = reinterpret_cast<const uint32_t *>(&views[wi.tree]);
            uint32_t *dst = reinterpret_cast<uint32_t *>(&tv);
            for (uint32_t i = threadIdx.x; i < sizeof(TreeView) / 4; i += blockDim.x) dst[i] = src[i];
            __syncthreads();
            const uint4 *s4 = reinterpret_cast<const uint4 *>(tv.inner);
            uint4 *d4 = reinterpret_cast<uint4 *>(lds);
            for (uint32_t i = threadIdx.x; i < tv.lds_bytes / 16; i += blockDim.x) d4[i] = s4[i];
            cur = wi.tree;
        }
        __syncthreads();
        if (wi.kind == 0) {
            for (uint64_t base = wi.begin + (uint64_t)w * 64; base < wi.end; base += (uint64_t)nw * 64) {
                const uint64_t my = base + lane;
                const uint64_t key = my < wi.end ? __builtin_nontemporal_load(&q[my]) : 0ull;
                __builtin_amdgcn_s_setprio(2);  // until the leaf's loads are issued (coop_leaf)
                const uint64_t r = coop_lookup64<uint64_t, uint64_t, F, LF, true, kOpValue, G>(tv, lds, key, ~0ull);
                if (my < wi.end) __builtin_nontemporal_store(r, &out[my]);
            }
        } else {
            // 32 ranges per wave step: lanes 0-31 descend for the lower
            // bounds, lanes 32-63 for the upper ones, so a range's two
            // descents share their load instructions and a node both reach
            // is one line request
            const uint32_t r = lane & 31;
            const bool upper = lane >= 32;
            for (uint64_t base = wi.begin + (uint64_t)w * 32; base < wi.end; base += (uint64_t)nw * 32) {
                const uint64_t my = base + r;
                const bool ok = my < wi.end;
                const uint64_t a = ok ? lo[my] : 0ull, b = ok ? hi[my] : 0ull;
                const bool top = b == ~0ull;
                const uint64_t key = upper ? (top ? b : b + 1) : a;
                const uint64_t p = coop_lookup64<uint64_t, uint64_t, F, LF, false, kOpLowerBound, G>(tv, lds, key, 0ull);
                const uint64_t pa = __shfl(p, (int)r, 64);
                const uint64_t pb0 = __shfl(p, (int)(r + 32), 64);
                const uint64_t pb = top ? tv.n : pb0;
                if (ok && !upper) {
                    first[my] = pa;
                    count[my] = (a > b || pb < pa) ? 0 : pb - pa;
                }
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ synthetic workload
// Counter-based generator (splitmix64, Steele et al. 2014) so that any tree's
// data can be regenerated independently: h(seed, stream, tree, i).
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t hmix(uint64_t seed, uint64_t stream, uint64_t tree, uint64_t i) {
    return splitmix64(splitmix64(splitmix64(seed ^ splitmix64(stream)) ^ tree) ^ i);
}
// Keys of tree t: n draws uniform over [0, 2^64), h(seed, 1 + 16 * attempt,
// t, i), sorted; attempt 0 unless a tree's draws collided (a redraw).
__global__ void k_synth_raw(uint64_t seed, const uint64_t *__restrict__ key_off, uint32_t ntrees, uint64_t total,
                            uint64_t *__restrict__ keys) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step) {
        const uint32_t t = find_tree(key_off, ntrees, g);
        keys[g] = hmix(seed, 1, t, g - key_off[t]);
    }
}

// tree[j] = tree of the element that the key sort put at j.
__global__ void k_tree_of(const uint32_t *__restrict__ perm, const uint64_t *__restrict__ key_off, uint32_t ntrees,
                          uint64_t total, uint32_t *__restrict__ tree) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step)
        tree[g] = find_tree(key_off, ntrees, perm[g]);
}

__global__ void k_gather_u64(const uint64_t *__restrict__ src, const uint32_t *__restrict__ idx, uint64_t n,
                             uint64_t *__restrict__ dst) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += step) dst[g] = src[idx[g]];
}

// Values = rank; dup[t] |= 1 when two keys of tree t are equal.
__global__ void k_rank_and_dups(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ key_off,
                                uint32_t ntrees, uint64_t total, uint64_t *__restrict__ vals,
                                uint32_t *__restrict__ dup) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step) {
        const uint32_t t = find_tree(key_off, ntrees, g);
        vals[g] = g - key_off[t];
        if (g > key_off[t] && keys[g] == keys[g - 1]) atomicOr(&dup[t], 1u);
    }
}

// A redraw of tree t's n keys (before sorting).
__global__ void k_synth_redraw(uint64_t seed, uint32_t t, uint32_t attempt, uint64_t n, uint64_t *__restrict__ keys) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step)
        keys[i] = hmix(seed, 1 + 16ull * attempt, t, i);
}

// Point queries: qpk per key of every tree; a fraction hit_pct/100 are keys
// of the tree (a uniform rank), the rest uniform 64-bit values.
__global__ void k_synth_points(uint64_t seed, const uint64_t *__restrict__ keys, const uint64_t *__restrict__ key_off,
                               uint32_t ntrees, uint32_t qpk, uint32_t hit_pct, uint64_t total,
                               uint64_t *__restrict__ q) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step) {
        const uint32_t t = find_tree(key_off, ntrees, g / qpk);
        const uint64_t n = key_off[t + 1] - key_off[t], j = g - key_off[t] * qpk;
        const uint64_t r = hmix(seed, 2, t, j);
        q[g] = ((r >> 32) % 100 < hit_pct) ? keys[key_off[t] + (r & 0xFFFFFFFFull) % n] : hmix(seed, 3, t, j);
    }
}

// Range queries: start rank uniform, width floor(2^u) ranks with u uniform
// in [lo_exp, hi_exp] (log-uniform), clipped to the tree; [key_s, key_e].
__global__ void k_synth_ranges(uint64_t seed, const uint64_t *__restrict__ keys, const uint64_t *__restrict__ key_off,
                               const uint64_t *__restrict__ range_off, uint32_t ntrees, uint32_t lo_exp,
                               uint32_t hi_exp, uint64_t total, uint64_t *__restrict__ lo, uint64_t *__restrict__ hi) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step) {
        const uint32_t t = find_tree(range_off, ntrees, g);
        const uint64_t n = key_off[t + 1] - key_off[t], j = g - range_off[t];
        const uint64_t r1 = hmix(seed, 4, t, j), r2 = hmix(seed, 5, t, j);
        const uint64_t s = r1 % n;
        const double u = (double)lo_exp + (double)(hi_exp - lo_exp) * ((double)(r2 >> 11) * 0x1.0p-53);
        uint64_t wdt = (uint64_t)floor(exp2(u));
        wdt = wdt < 1 ? 1 : wdt;
        const uint64_t last = (s + wdt < n ? s + wdt : n) - 1;
        lo[g] = keys[key_off[t] + s];
        hi[g] = keys[key_off[t] + last];
    }
}

static uint32_t grid_for(uint64_t n) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cu_count() * 16));
}

template <class T>
static bpt_status upload(const std::vector<T> &h, T **d) {
    BPT_HIP(hipMalloc(d, std::max<size_t>(1, h.size()) * sizeof(T)));
    if (!h.empty()) BPT_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return BPT_OK;
}

}  // namespace bpt

using namespace bpt;

extern "C" {

bpt_status bpt_forest_build(const bpt_batch_config *cfg, uint32_t ntrees, const uint64_t *d_keys,
                            const uint64_t *d_vals, const uint64_t *h_key_off, bpt_forest **out, bpt_stream stream) {
    BPT_CHECK(cfg && out && h_key_off, BPT_ERR_INVALID, "NULL argument");
    BPT_CHECK(ntrees >= 1, BPT_ERR_INVALID, "no trees");
    const uint32_t F = cfg->fanout ? cfg->fanout : 16;
    BPT_CHECK(F == 16 || F == 32 || F == 64, BPT_ERR_INVALID, "fanout %u not in {16,32,64}", F);
    BPT_CHECK(cfg->lanes == 0 || cfg->lanes == 4 || cfg->lanes == 8, BPT_ERR_INVALID, "lanes %u not in {4,8}",
              cfg->lanes);
    hipStream_t s = as_stream(stream);
    auto *f = new bpt_forest;
    f->ntrees = ntrees;
    f->fanout = F;
    f->lds_bytes = cfg->lds_bytes ? std::min<uint32_t>(cfg->lds_bytes, 65536) : 32768;
    f->lanes = cfg->lanes ? cfg->lanes : 8;
    f->views.resize(ntrees);
    f->desc.resize(ntrees);
    const uint64_t LF = 8, kb = 8;
    uint64_t leaf_bytes = 0, inner_bytes = 0;
    std::vector<uint64_t> slot_off(ntrees + 1, 0), sep_off(ntrees + 1, 0), key_off(ntrees + 1, 0);
    for (uint32_t t = 0; t < ntrees; ++t) {
        const uint64_t n = h_key_off[t + 1] - h_key_off[t];
        key_off[t] = h_key_off[t];
        TreeView &v = f->views[t];
        memset(&v, 0, sizeof v);
        v.n = n;
        v.leaf_stride = 128;
        uint64_t cnt = std::max<uint64_t>(1, (n + LF - 1) / LF);
        v.count[0] = cnt;
        int h = 0;
        while (cnt > 1) {
            cnt = (cnt + F - 1) / F;
            v.count[++h] = cnt;
        }
        v.height = h;
        uint64_t off = 0;
        for (int k = h; k >= 1; --k) { v.off[k] = off; off += v.count[k] * F * kb; }
        uint64_t lds = 0;
        int m = 0;
        for (int k = h; k >= 1; --k) {
            const uint64_t b = v.count[k] * F * kb;
            if (lds + b > f->lds_bytes) break;
            lds += b;
            ++m;
        }
        v.lds_levels = m;
        v.lds_bytes = (uint32_t)lds;
        f->desc[t] = {n, h_key_off[t], leaf_bytes, inner_bytes, leaf_bytes / 16, inner_bytes / 8};
        slot_off[t] = leaf_bytes / 16;  // 16 bytes per slot (key + value)
        sep_off[t] = inner_bytes / 8;
        leaf_bytes += v.count[0] * 128;
        inner_bytes += off;
    }
    slot_off[ntrees] = leaf_bytes / 16;
    sep_off[ntrees] = inner_bytes / 8;
    key_off[ntrees] = h_key_off[ntrees];
    f->leaf_bytes = leaf_bytes;
    f->inner_bytes = inner_bytes;
    f->total_keys = h_key_off[ntrees] - h_key_off[0];

    bpt_status st = BPT_OK;
    uint64_t *d_slot = nullptr, *d_sep = nullptr, *d_koff = nullptr;
    auto fail = [&](bpt_status e) { st = e; };
    hipError_t e = hipMalloc(&f->d_leaves, std::max<uint64_t>(leaf_bytes, 16));
    if (e == hipSuccess) e = hipMalloc(&f->d_inner, std::max<uint64_t>(inner_bytes, 16));
    if (e != hipSuccess) fail(hip_fail(e, "hipMalloc(forest)", __FILE__, __LINE__));
    for (uint32_t t = 0; t < ntrees && st == BPT_OK; ++t) {
        f->views[t].leaves = f->d_leaves + f->desc[t].leaf_off;
        f->views[t].inner = f->d_inner + f->desc[t].inner_off;
    }
    if (st == BPT_OK) st = upload(f->views, &f->d_views);
    if (st == BPT_OK) st = upload(slot_off, &d_slot);
    if (st == BPT_OK) st = upload(sep_off, &d_sep);
    if (st == BPT_OK) st = upload(key_off, &d_koff);
    if (st == BPT_OK) {
        k_forest_leaves<<<grid_for(leaf_bytes / 16), 256, 0, s>>>(d_keys, d_vals, d_slot, d_koff, ntrees,
                                                                  leaf_bytes / 16, f->d_leaves);
        const uint64_t nseps = inner_bytes / 8;
        if (nseps) {
            if (F == 16) k_forest_inner<16><<<grid_for(nseps), 256, 0, s>>>(d_keys, d_sep, d_koff, f->d_views, ntrees, nseps, (uint64_t *)f->d_inner);
            if (F == 32) k_forest_inner<32><<<grid_for(nseps), 256, 0, s>>>(d_keys, d_sep, d_koff, f->d_views, ntrees, nseps, (uint64_t *)f->d_inner);
            if (F == 64) k_forest_inner<64><<<grid_for(nseps), 256, 0, s>>>(d_keys, d_sep, d_koff, f->d_views, ntrees, nseps, (uint64_t *)f->d_inner);
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_fail(e, "forest build", __FILE__, __LINE__);
    }
    (void)hipFree(d_slot);
    (void)hipFree(d_sep);
    (void)hipFree(d_koff);
    if (st != BPT_OK) {
        bpt_forest_free(f);
        return st;
    }
    *out = f;
    return BPT_OK;
}

bpt_status bpt_forest_free(bpt_forest *f) {
    if (!f) return BPT_OK;
    (void)hipFree(f->d_leaves);
    (void)hipFree(f->d_inner);
    (void)hipFree(f->d_views);
    delete f;
    return BPT_OK;
}

bpt_status bpt_forest_info(const bpt_forest *f, uint32_t *ntrees, uint32_t *fanout, uint64_t *leaf_bytes,
                           uint64_t *inner_bytes) {
    BPT_CHECK(f, BPT_ERR_INVALID, "forest is NULL");
    if (ntrees) *ntrees = f->ntrees;
    if (fanout) *fanout = f->fanout;
    if (leaf_bytes) *leaf_bytes = f->leaf_bytes;
    if (inner_bytes) *inner_bytes = f->inner_bytes;
    return BPT_OK;
}

bpt_status bpt_batch_prepare(bpt_forest *f, const uint32_t chunk, const uint32_t blocks_per_cu, const uint64_t *d_q,
                             const uint64_t *h_q_off, uint64_t *d_out, const uint64_t *d_lo, const uint64_t *d_hi,
                             const uint64_t *h_r_off, uint64_t *d_first, uint64_t *d_count, bpt_batch **out) {
    BPT_CHECK(f && out, BPT_ERR_INVALID, "NULL argument");
    BPT_CHECK(h_q_off || h_r_off, BPT_ERR_INVALID, "no queries");
    auto *b = new bpt_batch;
    b->forest = f;
    b->chunk = chunk ? chunk : 8192;
    b->blocks_per_cu = blocks_per_cu ? blocks_per_cu : 8;
    b->q = d_q;
    b->out = d_out;
    b->lo = d_lo;
    b->hi = d_hi;
    b->first = d_first;
    b->count = d_count;
    std::vector<WorkItem> items;
    for (uint32_t t = 0; t < f->ntrees; ++t) {
        if (h_q_off) {
            for (uint64_t a = h_q_off[t]; a < h_q_off[t + 1]; a += b->chunk)
                items.push_back({t, 0u, a, std::min<uint64_t>(a + b->chunk, h_q_off[t + 1])});
        }
        if (h_r_off) {
            const uint64_t rc = std::max<uint32_t>(b->chunk / 4, 64);
            for (uint64_t a = h_r_off[t]; a < h_r_off[t + 1]; a += rc)
                items.push_back({t, 1u, a, std::min<uint64_t>(a + rc, h_r_off[t + 1])});
        }
    }
    b->npoints = h_q_off ? h_q_off[f->ntrees] - h_q_off[0] : 0;
    b->nranges = h_r_off ? h_r_off[f->ntrees] - h_r_off[0] : 0;
    b->nitems = (uint32_t)items.size();
    // per-XCD item ranges balanced by lookups (a range query costs two descents)
    std::vector<uint32_t> cnt(2 * kBatchXcds + 1, 0);
    {
        uint64_t total = 0;
        for (const WorkItem &w : items) total += (w.end - w.begin) * (w.kind ? 2 : 1);
        uint64_t acc = 0;
        uint32_t x = 1;
        for (uint32_t i = 0; i < b->nitems && x < kBatchXcds; ++i) {
            acc += (items[i].end - items[i].begin) * (items[i].kind ? 2 : 1);
            while (x < kBatchXcds && acc * kBatchXcds >= total * x) cnt[kBatchXcds + x++] = i + 1;
        }
        while (x <= kBatchXcds) cnt[kBatchXcds + x++] = b->nitems;
    }
    bpt_status st = upload(items, &b->d_items);
    if (st == BPT_OK) st = upload(cnt, &b->d_counter);
    if (st != BPT_OK) {
        bpt_batch_free(b);
        return st;
    }
    *out = b;
    return BPT_OK;
}

static bpt_status batch_enqueue(bpt_batch *b, hipStream_t s) {
    bpt_forest *f = b->forest;
    k_reset_counter<<<1, 64, 0, s>>>(b->d_counter);
    BPT_HIP(hipGetLastError());
    const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)device_cu_count() * b->blocks_per_cu,
                                                       std::max<uint32_t>(b->nitems, 1));
    const size_t lds = std::max<uint32_t>(f->lds_bytes, 16);
    auto go = [&](auto fc, auto gc) {
        k_forest_batch<decltype(fc)::value, decltype(gc)::value><<<grid, 256, lds, s>>>(
            f->d_views, b->d_items, b->nitems, b->d_counter, b->q, b->out, b->lo, b->hi, b->first, b->count);
    };
    auto with_lanes = [&](auto fc) {
        if (f->lanes == 4) go(fc, IC<4>{});
        else go(fc, IC<8>{});
    };
    switch (f->fanout) {
        case 16: with_lanes(IC<16>{}); break;
        case 32: with_lanes(IC<32>{}); break;
        case 64: with_lanes(IC<64>{}); break;
    }
    BPT_HIP(hipGetLastError());
    return BPT_OK;
}

bpt_status bpt_batch_run(bpt_batch *b, int use_graph, bpt_stream stream) {
    BPT_CHECK(b, BPT_ERR_INVALID, "batch is NULL");
    hipStream_t s = as_stream(stream);
    if (!use_graph) return batch_enqueue(b, s);
    if (!b->exec) {
        // capture on a private stream (the legacy default stream cannot capture)
        hipStream_t cs;
        BPT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        bpt_status st = e == hipSuccess ? batch_enqueue(b, cs) : hip_fail(e, "begin capture", __FILE__, __LINE__);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(cs, &g);
        if (st == BPT_OK && e != hipSuccess) st = hip_fail(e, "end capture", __FILE__, __LINE__);
        if (st == BPT_OK) {
            e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) st = hip_fail(e, "graph instantiate", __FILE__, __LINE__);
        }
        b->graph = g;
        (void)hipStreamDestroy(cs);
        if (st != BPT_OK) return st;
    }
    BPT_HIP(hipGraphLaunch(b->exec, s));
    return BPT_OK;
}

bpt_status bpt_batch_free(bpt_batch *b) {
    if (!b) return BPT_OK;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    (void)hipFree(b->d_items);
    (void)hipFree(b->d_counter);
    delete b;
    return BPT_OK;
}

bpt_status bpt_batch_info(const bpt_batch *b, uint32_t *nitems, uint64_t *npoints, uint64_t *nranges) {
    BPT_CHECK(b, BPT_ERR_INVALID, "batch is NULL");
    if (nitems) *nitems = b->nitems;
    if (npoints) *npoints = b->npoints;
    if (nranges) *nranges = b->nranges;
    return BPT_OK;
}

bpt_status bpt_forest_autotune(uint32_t ntrees, const uint64_t *d_keys, const uint64_t *d_vals,
                               const uint64_t *h_key_off, const uint64_t *d_q, const uint64_t *h_q_off,
                               uint64_t *d_out, const uint64_t *d_lo, const uint64_t *d_hi, const uint64_t *h_r_off,
                               uint64_t *d_first, uint64_t *d_count, int reps, bpt_batch_config *best,
                               char *report, uint64_t report_len) {
    BPT_CHECK(best, BPT_ERR_INVALID, "best is NULL");
    const uint32_t fanouts[3] = {16, 32, 64};
    const uint32_t lds_opts[3] = {16384, 32768, 65536};
    const uint32_t chunks[3] = {2048, 8192, 32768};
    const uint32_t lanes_opts[2] = {8, 4};  // 8 or 16 queries per wavefront
    std::string js = "[";
    double best_ms = 1e30;
    hipEvent_t a, z;
    BPT_HIP(hipEventCreate(&a));
    BPT_HIP(hipEventCreate(&z));
    bpt_status st = BPT_OK;
    for (uint32_t F : fanouts) {
        for (uint32_t L : lds_opts) {
            bpt_batch_config c{F, 8, L, 0, 0, 8};
            bpt_forest *f = nullptr;
            st = bpt_forest_build(&c, ntrees, d_keys, d_vals, h_key_off, &f, nullptr);
            if (st != BPT_OK) break;
            for (uint32_t G : lanes_opts) {
                f->lanes = G;
                for (uint32_t ch : chunks) {
                    bpt_batch *b = nullptr;
                    st = bpt_batch_prepare(f, ch, 8, d_q, h_q_off, d_out, d_lo, d_hi, h_r_off, d_first, d_count, &b);
                    if (st != BPT_OK) break;
                    std::vector<float> ms;
                    for (int r = 0; r <= reps && st == BPT_OK; ++r) {
                        (void)hipEventRecord(a, nullptr);
                        st = bpt_batch_run(b, 1, nullptr);
                        (void)hipEventRecord(z, nullptr);
                        (void)hipEventSynchronize(z);
                        float x = 0;
                        (void)hipEventElapsedTime(&x, a, z);
                        if (r > 0) ms.push_back(x);  // the first run is a warm-up
                    }
                    bpt_batch_free(b);
                    if (st != BPT_OK) break;
                    std::sort(ms.begin(), ms.end());
                    const double med = ms.empty() ? 0 : ms[ms.size() / 2];
                    char buf[160];
                    snprintf(buf, sizeof buf,
                             "%s{\"fanout\":%u,\"lds_bytes\":%u,\"lanes\":%u,\"queries_per_wave\":%u,\"chunk\":%u,"
                             "\"ms\":%.4f}",
                             js.size() > 1 ? "," : "", F, L, G, 64 / G, ch, med);
                    js += buf;
                    if (med < best_ms) {
                        best_ms = med;
                        *best = bpt_batch_config{F, 8, L, ch, 8, G};
                    }
                }
                if (st != BPT_OK) break;
            }
            bpt_forest_free(f);
            if (st != BPT_OK) break;
        }
        if (st != BPT_OK) break;
    }
    js += "]";
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(z);
    if (report && report_len) {
        strncpy(report, js.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return st;
}

bpt_status bpt_synth_forest(uint64_t seed, uint32_t ntrees, const uint64_t *h_key_off, uint32_t qpk,
                            uint32_t hit_pct, const uint64_t *h_r_off, uint32_t w_lo_exp, uint32_t w_hi_exp,
                            uint64_t *d_keys, uint64_t *d_vals, uint64_t *d_q, uint64_t *d_lo, uint64_t *d_hi,
                            bpt_stream stream) {
    BPT_CHECK(h_key_off && ntrees && d_keys && d_vals, BPT_ERR_INVALID, "NULL argument");
    BPT_CHECK(w_lo_exp <= w_hi_exp && w_hi_exp < 63, BPT_ERR_INVALID, "width exponents");
    hipStream_t s = as_stream(stream);
    std::vector<uint64_t> koff(h_key_off, h_key_off + ntrees + 1);
    BPT_CHECK(koff[0] == 0, BPT_ERR_INVALID, "h_key_off[0] must be 0");
    for (uint32_t t = 0; t < ntrees; ++t) BPT_CHECK(koff[t + 1] > koff[t], BPT_ERR_INVALID, "empty tree %u", t);
    const uint64_t total = koff[ntrees];
    BPT_CHECK(total < (1ull << 32), BPT_ERR_CAPACITY, "forest of more than 2^32 keys");
    uint32_t tbits = 1;
    while ((1ull << tbits) < ntrees) ++tbits;
    // scratch: the key sort's workspace, later the tree sort's
    const uint64_t ws_bytes = std::max(sort_ws_bytes(total, 8), sort_ws_bytes(total, 4));
    uint64_t *dk = nullptr, *dr = nullptr;
    void *ws = nullptr;
    uint32_t *dup = nullptr;
    BPT_TRY(upload(koff, &dk));
    bpt_status st = BPT_OK;
    hipError_t e = hipMalloc(&ws, ws_bytes);
    if (e == hipSuccess) e = hipMalloc(&dup, ntrees * 4ull);
    if (e != hipSuccess) st = hip_fail(e, "hipMalloc(synthetic forest scratch)", __FILE__, __LINE__);
    // 1. uniform draws, 2. sorted by key (LSD, stable), 3. stably by tree:
    // each tree's keys end sorted in its own range
    if (st == BPT_OK) {
        k_synth_raw<<<grid_for(total), 256, 0, s>>>(seed, dk, ntrees, total, d_keys);
        uint64_t *ks = nullptr;
        uint32_t *perm = nullptr;
        st = sort_pairs<uint64_t>(d_keys, nullptr, total, 0, 64, ws, ws_bytes, &ks, &perm, s);
        if (st == BPT_OK) {
            e = hipMemcpyAsync(d_keys, ks, total * 8, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) st = hip_fail(e, "synthetic keys copy", __FILE__, __LINE__);
        }
        // perm lives in the first workspace: tree ids go to d_vals (free so far)
        uint32_t *tree = reinterpret_cast<uint32_t *>(d_vals);
        if (st == BPT_OK) k_tree_of<<<grid_for(total), 256, 0, s>>>(perm, dk, ntrees, total, tree);
        uint32_t *ts = nullptr, *perm2 = nullptr;
        if (st == BPT_OK) st = sort_pairs<uint32_t>(tree, nullptr, total, 0, tbits, ws, ws_bytes, &ts, &perm2, s);
        if (st == BPT_OK) {
            k_gather_u64<<<grid_for(total), 256, 0, s>>>(d_keys, perm2, total, d_vals);
            e = hipMemcpyAsync(d_keys, d_vals, total * 8, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) st = hip_fail(e, "synthetic keys copy", __FILE__, __LINE__);
        }
    }
    // values = rank; trees whose draws collided are redrawn (a collision
    // among 2^20 uniform 64-bit draws has probability about 2^-25)
    std::vector<uint32_t> hdup(ntrees);
    for (uint32_t attempt = 1; st == BPT_OK; ++attempt) {
        e = hipMemsetAsync(dup, 0, ntrees * 4ull, s);
        if (e == hipSuccess) {
            k_rank_and_dups<<<grid_for(total), 256, 0, s>>>(d_keys, dk, ntrees, total, d_vals, dup);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(hdup.data(), dup, ntrees * 4ull, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { st = hip_fail(e, "synthetic forest", __FILE__, __LINE__); break; }
        bool any = false;
        for (uint32_t t = 0; t < ntrees && st == BPT_OK; ++t) {
            if (!hdup[t]) continue;
            any = true;
            if (attempt > 8) { set_error("tree %u: keys collide after 8 redraws", t); st = BPT_ERR_INVALID; break; }
            const uint64_t n = koff[t + 1] - koff[t];
            uint64_t *seg = d_keys + koff[t], *ks = nullptr;
            uint32_t *perm = nullptr;
            k_synth_redraw<<<grid_for(n), 256, 0, s>>>(seed, t, attempt, n, seg);
            st = sort_pairs<uint64_t>(seg, nullptr, n, 0, 64, ws, ws_bytes, &ks, &perm, s);
            if (st == BPT_OK) {
                e = hipMemcpyAsync(seg, ks, n * 8, hipMemcpyDeviceToDevice, s);
                if (e != hipSuccess) st = hip_fail(e, "redraw copy", __FILE__, __LINE__);
            }
        }
        if (!any) break;
    }
    if (st == BPT_OK && d_q && qpk)
        k_synth_points<<<grid_for(total * qpk), 256, 0, s>>>(seed, d_keys, dk, ntrees, qpk, hit_pct, total * qpk, d_q);
    if (st == BPT_OK && h_r_off && d_lo && d_hi) {
        std::vector<uint64_t> roff(h_r_off, h_r_off + ntrees + 1);
        st = upload(roff, &dr);
        const uint64_t nr = roff[ntrees] - roff[0];
        if (st == BPT_OK && nr)
            k_synth_ranges<<<grid_for(nr), 256, 0, s>>>(seed, d_keys, dk, dr, ntrees, w_lo_exp, w_hi_exp, nr, d_lo,
                                                        d_hi);
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (st == BPT_OK && e != hipSuccess) st = hip_fail(e, "synthetic forest", __FILE__, __LINE__);
    (void)hipFree(dk);
    (void)hipFree(dr);
    (void)hipFree(ws);
    (void)hipFree(dup);
    return st;
}

}  // extern "C"
