// Bulk build: sorted (key, value) arrays -> flat, level-ordered B+tree.
//
// Replaces genBPlusTree (reference BPTree.cu:127-306), which copies PAGESIZE
// chunks into 80 KB pages (assignDataToLeaf, BPTree.cu:73-125) and then runs
// one kernel per internal level whose separators are the child maxima
// (kernelMaxIntermidatePage, BPTree.cu:28-70). Here:
//  * leaves are fixed-size lines of `leaf_cap` keys followed by their values,
//    so one leaf is one 128-byte HBM line for 4/4- and 8/8-byte pairs;
//  * internal levels hold only separators (child maximum, as the reference),
//    children are implicit (child c of node i is node i*F + c one level
//    down), levels are stored root first so the top of the tree is one
//    contiguous prefix that search kernels stage in LDS;
//  * every separator is a closed form of the input: the maximum of a node at
//    level k covers keys [i*LF*F^k, (i+1)*LF*F^k), so all internal levels are
//    written by ONE level-parallel launch that reads the sorted keys directly.
#include <algorithm>
#include <cstring>

#include "bpt_internal.hpp"

namespace bpt {

template <class K, class V, int LF>
__global__ __launch_bounds__(256) void k_build_leaves(const K *__restrict__ keys,
                                                      const V *__restrict__ vals, uint64_t n,
                                                      uint64_t nslots, uint8_t *__restrict__ leaves,
                                                      uint64_t stride) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslots; s += step) {
        const uint64_t leaf = s / LF;
        const uint32_t j = (uint32_t)(s % LF);
        K k = s < n ? keys[s] : key_max<K>();
        V v = s < n ? vals[s] : V(0);
        uint8_t *lp = leaves + leaf * stride;
        *reinterpret_cast<K *>(lp + LeafFmt<K, V, LF>::key_off(j)) = k;
        *reinterpret_cast<V *>(lp + LeafFmt<K, V, LF>::val_off(j)) = v;
    }
}

struct InnerDesc {
    int32_t height;
    uint64_t count[kMaxLevels + 1];
    uint64_t first_slot[kMaxLevels + 1];  // first separator slot of level k in `inner`
    uint64_t span[kMaxLevels + 1];        // keys covered by one node of level k
};

template <class K, int F>
__global__ __launch_bounds__(256) void k_build_inner(const K *__restrict__ keys, uint64_t n,
                                                     InnerDesc d, uint64_t total_slots,
                                                     K *__restrict__ inner) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total_slots; s += step) {
        int k = d.height;  // levels are stored root first
        while (k > 1 && s >= d.first_slot[k - 1]) --k;
        const uint64_t local = s - d.first_slot[k];
        const uint64_t child = local;  // == node * F + c
        K sep = key_max<K>();
        if (child < d.count[k - 1]) {
            uint64_t end = (child + 1) * d.span[k - 1];
            sep = keys[(end < n ? end : n) - 1];
        }
        inner[s] = sep;
    }
}

// Sub-leaf maxima: entry s is the last key of slots [s*sub, (s+1)*sub), the
// largest key there (the key maximum when the run holds no key).
template <class K>
__global__ __launch_bounds__(256) void k_build_sub(const K *__restrict__ keys, uint64_t n, uint64_t nsub,
                                                   uint32_t sub, K *__restrict__ out) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nsub; s += step) {
        const uint64_t first = s * sub, last = first + sub - 1;
        out[s] = first < n ? keys[last < n ? last : n - 1] : key_max<K>();
    }
}

template <class K>
__global__ __launch_bounds__(256) void k_check_sorted(const K *__restrict__ keys, uint64_t n,
                                                      uint32_t *bad) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    uint32_t b = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i + 1 < n; i += step)
        b |= keys[i] > keys[i + 1];
    if (__any(b) && (threadIdx.x & 63) == 0) atomicOr(bad, 1u);
}

uint32_t search_grid(const bpt_tree *t, uint64_t work_items) {
    uint64_t blocks = (work_items + t->block - 1) / t->block;
    uint64_t cap = (uint64_t)device_cu_count() * t->blocks_per_cu;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    return (uint32_t)blocks;
}

static bpt_status compute_layout(bpt_tree *t) {
    const uint64_t LF = t->leaf_cap, F = t->fanout, kb = t->key_bytes;
    TreeView &v = t->view;
    memset(&v, 0, sizeof v);
    v.n = t->n;
    v.leaf_stride = t->leaf_stride;
    uint64_t cnt = (t->n + LF - 1) / LF;
    v.count[0] = cnt;
    int h = 0;
    while (cnt > 1) {
        BPT_CHECK(h < kMaxLevels, BPT_ERR_CAPACITY, "tree taller than %d levels", kMaxLevels);
        cnt = (cnt + F - 1) / F;
        v.count[++h] = cnt;
    }
    v.height = h;
    uint64_t off = 0;
    for (int k = h; k >= 1; --k) {
        v.off[k] = off;
        off += v.count[k] * F * kb;
    }
    // Sub-leaf maxima, after the levels: the largest key of every run of
    // `sub` slots of a leaf (one 64-byte run of (key, value) pairs in the
    // interleaved format), for leaves that hold several such runs. The
    // grouped search stages a bucket's slice of them in LDS and then reads
    // 64 bytes of a leaf per query instead of the whole leaf.
    const uint64_t sub = 64 / (2 * kb);
    if (t->key_bytes == t->val_bytes && h >= 1 && LF > sub && LF % sub == 0) {
        v.sub_slots = (uint32_t)sub;
        v.sub_off = (off + 255) / 256 * 256;
        off = v.sub_off + v.count[0] * (LF / sub) * kb;
    }
    t->inner_bytes = off;
    t->leaf_bytes = v.count[0] * t->leaf_stride;
    // top levels that fit the LDS budget
    uint64_t lds = 0;
    int m = 0;
    for (int k = h; k >= 1; --k) {
        uint64_t b = v.count[k] * F * kb;
        if (lds + b > t->lds_budget) break;
        lds += b;
        ++m;
    }
    v.lds_levels = m;
    v.lds_bytes = (uint32_t)lds;
    return BPT_OK;
}

template <class K, class V, int F, int LF>
static bpt_status build_impl(bpt_tree *t, const void *d_keys, const void *d_vals, hipStream_t s) {
    const TreeView &v = t->view;
    const uint64_t n = t->n;
    if (n == 0) return BPT_OK;
    const uint64_t nslots = v.count[0] * LF;
    {
        uint64_t blocks = std::min<uint64_t>((nslots + 255) / 256, (uint64_t)device_cu_count() * 16);
        k_build_leaves<K, V, LF><<<(uint32_t)blocks, 256, 0, s>>>(
            (const K *)d_keys, (const V *)d_vals, n, nslots, t->d_leaves, t->leaf_stride);
        BPT_HIP(hipGetLastError());
    }
    if (v.height >= 1) {
        InnerDesc d;
        memset(&d, 0, sizeof d);
        d.height = v.height;
        uint64_t span = LF;
        for (int k = 0; k <= v.height; ++k) {
            d.count[k] = v.count[k];
            d.span[k] = span;
            span = (span > (~0ull) / F) ? ~0ull : span * F;
            if (k >= 1) d.first_slot[k] = v.off[k] / sizeof(K);
        }
        // separator slots of every level (level 1, the leaf maxima, is stored last)
        const uint64_t total = (v.off[1] + v.count[1] * F * sizeof(K)) / sizeof(K);
        uint64_t blocks = std::min<uint64_t>((total + 255) / 256, (uint64_t)device_cu_count() * 16);
        k_build_inner<K, F><<<(uint32_t)blocks, 256, 0, s>>>((const K *)d_keys, n, d, total,
                                                             (K *)t->d_inner);
        BPT_HIP(hipGetLastError());
    }
    if (v.sub_slots) {
        const uint64_t nsub = v.count[0] * (LF / v.sub_slots);
        uint64_t blocks = std::min<uint64_t>((nsub + 255) / 256, (uint64_t)device_cu_count() * 16);
        k_build_sub<K><<<(uint32_t)blocks, 256, 0, s>>>((const K *)d_keys, n, nsub, v.sub_slots,
                                                        (K *)(t->d_inner + v.sub_off));
        BPT_HIP(hipGetLastError());
    }
    return BPT_OK;
}

template <class K>
static bpt_status check_sorted(const void *d_keys, uint64_t n, hipStream_t s, bool *sorted) {
    *sorted = true;
    if (n < 2) return BPT_OK;
    uint32_t *d_bad = nullptr;
    BPT_HIP(hipMalloc(&d_bad, sizeof(uint32_t)));
    uint32_t h_bad = 0;
    hipError_t e = hipMemsetAsync(d_bad, 0, sizeof(uint32_t), s);
    if (e == hipSuccess) {
        uint64_t blocks = std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cu_count() * 16);
        k_check_sorted<K><<<(uint32_t)blocks, 256, 0, s>>>((const K *)d_keys, n, d_bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h_bad, d_bad, sizeof h_bad, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_bad);
    if (e != hipSuccess) return hip_fail(e, "sorted check", __FILE__, __LINE__);
    *sorted = (h_bad == 0);
    return BPT_OK;
}

// Device buffers for the current layout. Buffers that hold the new size and
// are at most twice it are kept; otherwise they are reallocated, with 1/16
// slack when `slack` is set so that a run of growing updates rarely
// reallocates. New buffers are allocated before any old one is freed, so a
// failed allocation leaves the tree's buffers and view as they were.
static bpt_status ensure_storage(bpt_tree *t, bool slack) {
    struct Fit {
        uint8_t *p;
        uint64_t cap;
        bool fresh;
    };
    auto fit = [&](uint8_t *cur, uint64_t cap, uint64_t need, const char *what, Fit *out) -> bpt_status {
        *out = {cur, cap, false};
        if (cur && need <= cap && need * 2 >= cap) return BPT_OK;
        *out = {nullptr, 0, true};
        if (need == 0) return BPT_OK;
        const uint64_t want = slack ? need + need / 16 : need;
        BPT_CHECK(t->alloc_limit == 0 || want <= t->alloc_limit, BPT_ERR_NOMEM,
                  "%s: %llu bytes above the tree's allocation limit of %llu", what, (unsigned long long)want,
                  (unsigned long long)t->alloc_limit);
        uint8_t *p = nullptr;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) return hip_fail(e, what, __FILE__, __LINE__);
        *out = {p, want, true};
        return BPT_OK;
    };
    Fit leaves, inner;
    BPT_TRY(fit(t->d_leaves, t->leaf_alloc, t->leaf_bytes, "hipMalloc(leaves)", &leaves));
    bpt_status st = fit(t->d_inner, t->inner_alloc, t->inner_bytes, "hipMalloc(inner)", &inner);
    if (st != BPT_OK) {
        if (leaves.fresh) (void)hipFree(leaves.p);
        return st;
    }
    if (leaves.fresh) {
        (void)hipFree(t->d_leaves);
        t->d_leaves = leaves.p;
        t->leaf_alloc = leaves.cap;
    }
    if (inner.fresh) {
        (void)hipFree(t->d_inner);
        t->d_inner = inner.p;
        t->inner_alloc = inner.cap;
    }
    t->view.inner = t->d_inner;
    t->view.leaves = t->d_leaves;
    return BPT_OK;
}

bpt_status tree_rebuild(bpt_tree *t, const void *d_keys, const void *d_vals, uint64_t n, hipStream_t s) {
    // the new layout and buffers first; if either fails, nothing of the tree
    // has been written and it is restored as it was
    const uint64_t old_n = t->n, old_leaf_bytes = t->leaf_bytes, old_inner_bytes = t->inner_bytes;
    const TreeView old_view = t->view;
    t->n = n;
    bpt_status st = compute_layout(t);
    if (st == BPT_OK) st = ensure_storage(t, true);
    if (st != BPT_OK) {
        t->n = old_n;
        t->leaf_bytes = old_leaf_bytes;
        t->inner_bytes = old_inner_bytes;
        t->view = old_view;
        return st;
    }
    if (t->learned) {
        learned_free(t->learned);
        t->learned = nullptr;
    }
    t->bkt_valid = false;
    st = dispatch_tree(t, [&](auto kt, auto vt, auto fc, auto lc) {
        using K = typename decltype(kt)::type;
        using V = typename decltype(vt)::type;
        return build_impl<K, V, decltype(fc)::value, decltype(lc)::value>(t, d_keys, d_vals, s);
    });
    if (st == BPT_OK) st = bucket_tables_refresh(t, s);
    if (st == BPT_OK) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_fail(e, "rebuild sync", __FILE__, __LINE__);
    }
    if (st != BPT_OK) {  // a device error while writing: leave an empty, usable tree behind
        (void)hipFree(t->d_leaves);
        (void)hipFree(t->d_inner);
        t->d_leaves = t->d_inner = nullptr;
        t->leaf_alloc = t->inner_alloc = 0;
        t->n = 0;
        (void)compute_layout(t);
        t->view.inner = t->view.leaves = nullptr;
    }
    return st;
}

}  // namespace bpt

using namespace bpt;

extern "C" {

bpt_status bpt_tree_build(const bpt_build_desc *desc, const void *d_keys, const void *d_vals,
                          uint64_t n, bpt_tree **out, bpt_stream stream) {
    BPT_CHECK(desc && out, BPT_ERR_INVALID, "desc/out is NULL");
    *out = nullptr;
    const uint32_t kb = desc->key_bytes, vb = desc->val_bytes;
    BPT_CHECK((kb == 4 && vb == 4) || (kb == 8 && vb == 8) || (kb == 8 && vb == 4), BPT_ERR_INVALID,
              "key/value widths %u/%u not supported (4/4, 8/8, 8/4)", kb, vb);
    uint32_t F = desc->fanout ? desc->fanout : 128 / kb;
    uint32_t LF = desc->leaf_cap ? desc->leaf_cap : (kb + vb == 12 ? 8 : 128 / (kb + vb));
    BPT_CHECK(F == 16 || F == 32 || F == 64, BPT_ERR_INVALID, "fanout %u not in {16,32,64}", F);
    BPT_CHECK(LF == 8 || LF == 16 || LF == 32, BPT_ERR_INVALID, "leaf_cap %u not in {8,16,32}", LF);
    BPT_CHECK(n == 0 || (d_keys && d_vals), BPT_ERR_INVALID, "keys/values are NULL");
    hipStream_t s = as_stream(stream);

    if (desc->check_sorted && n > 1) {
        bool sorted = true;
        BPT_TRY(kb == 4 ? check_sorted<uint32_t>(d_keys, n, s, &sorted)
                        : check_sorted<uint64_t>(d_keys, n, s, &sorted));
        BPT_CHECK(sorted, BPT_ERR_UNSORTED, "build keys are not sorted non-decreasing");
    }

    bpt_tree *t = new bpt_tree;
    t->key_bytes = kb;
    t->val_bytes = vb;
    t->fanout = F;
    t->leaf_cap = LF;
    t->n = n;
    t->leaf_stride = ((uint64_t)LF * (kb + vb) + 63) / 64 * 64;
    t->lds_budget = desc->lds_bytes ? desc->lds_bytes : 32768;
    if (t->lds_budget > 65536) t->lds_budget = 65536;
    (void)hipGetDevice(&t->device);

    bpt_status st = compute_layout(t);
    if (st == BPT_OK) st = ensure_storage(t, false);
    if (st == BPT_OK) {
        st = dispatch_tree(t, [&](auto kt, auto vt, auto fc, auto lc) {
            using K = typename decltype(kt)::type;
            using V = typename decltype(vt)::type;
            return build_impl<K, V, decltype(fc)::value, decltype(lc)::value>(t, d_keys, d_vals, s);
        });
    }
    if (st == BPT_OK) st = bucket_tables_refresh(t, s);
    if (st == BPT_OK) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_fail(e, "build sync", __FILE__, __LINE__);
    }
    if (st != BPT_OK) {
        bpt_tree_free(t);
        return st;
    }
    *out = t;
    return BPT_OK;
}

bpt_status bpt_tree_free(bpt_tree *t) {
    if (!t) return BPT_OK;
    if (t->learned) learned_free(t->learned);
    if (t->d_leaves) (void)hipFree(t->d_leaves);
    if (t->d_inner) (void)hipFree(t->d_inner);
    if (t->d_bkt) (void)hipFree(t->d_bkt);
    delete t;
    return BPT_OK;
}

bpt_status bpt_tree_get_info(const bpt_tree *t, bpt_tree_info *info) {
    BPT_CHECK(t && info, BPT_ERR_INVALID, "NULL argument");
    memset(info, 0, sizeof *info);
    info->n = t->n;
    info->key_bytes = t->key_bytes;
    info->val_bytes = t->val_bytes;
    info->fanout = t->fanout;
    info->leaf_cap = t->leaf_cap;
    info->height = (uint32_t)t->view.height;
    info->lds_levels = (uint32_t)t->view.lds_levels;
    info->lds_bytes = t->view.lds_bytes;
    info->leaf_stride = t->leaf_stride;
    info->leaf_bytes = t->leaf_bytes;
    info->inner_bytes = t->inner_bytes;
    for (int k = 0; k <= t->view.height && k < BPT_MAX_LEVELS; ++k) info->level_nodes[k] = t->view.count[k];
    return BPT_OK;
}

bpt_status bpt_tree_get_layout(const bpt_tree *t, const void **d_leaves, const void **d_inner) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    if (d_leaves) *d_leaves = t->d_leaves;
    if (d_inner) *d_inner = t->d_inner;
    return BPT_OK;
}

bpt_status bpt_tree_set_launch(bpt_tree *t, uint32_t block, uint32_t blocks_per_cu) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    if (block) {
        BPT_CHECK(block == 64 || block == 128 || block == 256, BPT_ERR_INVALID,
                  "block %u not 64, 128 or 256 (the search kernels' launch bound)", block);
        t->block = block;
    }
    if (blocks_per_cu) t->blocks_per_cu = blocks_per_cu;
    return BPT_OK;
}

bpt_status bpt_tree_set_option(bpt_tree *t, const char *key, int64_t value) {
    BPT_CHECK(t && key, BPT_ERR_INVALID, "NULL argument");
    if (strcmp(key, "coop") == 0) {
        t->coop = value ? 1 : 0;
        return BPT_OK;
    }
    if (strcmp(key, "leaf_nt") == 0) {
        t->leaf_nt = value ? 1 : 0;
        return BPT_OK;
    }
    if (strcmp(key, "addr32") == 0) {
        t->addr32 = value ? 1 : 0;
        return BPT_OK;
    }
    if (strcmp(key, "group_chunk") == 0) {
        BPT_CHECK(value == 0 || (value >= 256 && value <= (1 << 20) && value % 64 == 0), BPT_ERR_INVALID,
                  "group_chunk must be 0 (by batch size) or a multiple of 64 in [256, 2^20]");
        t->group_chunk = (uint32_t)value;
        return BPT_OK;
    }
    if (strcmp(key, "group_lanes") == 0) {
        BPT_CHECK(value == 4 || value == 8, BPT_ERR_INVALID, "group_lanes must be 4 or 8");
        t->group_lanes = (uint32_t)value;
        return BPT_OK;
    }
    if (strcmp(key, "group_blocks") == 0) {
        BPT_CHECK(value >= 1 && value <= 8, BPT_ERR_INVALID, "group_blocks must be in [1, 8]");
        t->group_blocks = (uint32_t)value;
        return BPT_OK;
    }
    if (strcmp(key, "group_slice") == 0) {
        BPT_CHECK(value == 16384 || value == 32768 || value == 65536 || value == 131072, BPT_ERR_INVALID,
                  "group_slice must be 16384, 32768, 65536 or 131072 bytes");
        t->group_slice = (uint32_t)value;
        return BPT_OK;
    }
    if (strcmp(key, "group_threads") == 0) {
        BPT_CHECK(value == 512 || value == 768 || value == 1024, BPT_ERR_INVALID,
                  "group_threads must be 512, 768 or 1024");
        t->group_threads = (uint32_t)value;
        return BPT_OK;
    }
    if (strcmp(key, "group_sub") == 0) {
        t->group_sub = value ? 1 : 0;
        return BPT_OK;
    }
    if (strcmp(key, "group_min") == 0) {
        BPT_CHECK(value >= 0, BPT_ERR_INVALID, "group_min must be >= 0");
        t->group_min = (uint64_t)value;
        return BPT_OK;
    }
    if (strcmp(key, "group_xcd") == 0) {
        t->group_xcd = value ? 1 : 0;
        return BPT_OK;
    }
    if (strcmp(key, "lanes") == 0) {
        BPT_CHECK(value == 4 || value == 8, BPT_ERR_INVALID, "lanes per query must be 4 or 8");
        t->coop_lanes = (uint32_t)value;
        return BPT_OK;
    }
    if (strcmp(key, "alloc_limit") == 0) {
        BPT_CHECK(value >= 0, BPT_ERR_INVALID, "alloc_limit must be >= 0");
        t->alloc_limit = (uint64_t)value;
        return BPT_OK;
    }
    set_error("unknown tree option '%s'", key);
    return BPT_ERR_NOT_FOUND;
}

bpt_status bpt_tree_set_group_bits(bpt_tree *t, uint32_t bits) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    BPT_CHECK(bits >= 1 && bits <= 16 && bits <= t->key_bytes * 8, BPT_ERR_INVALID,
              "group bits %u not in [1, 16]", bits);
    t->group_bits = bits;
    return BPT_OK;
}

}  // extern "C"
