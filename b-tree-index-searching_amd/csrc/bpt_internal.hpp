// Internal declarations shared by the HIP sources of libbptree_amd.
// gfx950 (CDNA4) only: 64-lane wavefronts, 160 KiB LDS per CU, 128 B L2 lines.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "bptree.h"

namespace bpt {

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
bpt_status hip_fail(hipError_t e, const char *what, const char *file, int line);

#define BPT_HIP(...)                                                         \
    do {                                                                     \
        hipError_t e_ = (__VA_ARGS__);                                       \
        if (e_ != hipSuccess) return ::bpt::hip_fail(e_, #__VA_ARGS__, __FILE__, __LINE__); \
    } while (0)

#define BPT_CHECK(cond, status, ...)                                         \
    do {                                                                     \
        if (!(cond)) { ::bpt::set_error(__VA_ARGS__); return (status); }     \
    } while (0)

#define BPT_TRY(...)                                                         \
    do {                                                                     \
        bpt_status s_ = (__VA_ARGS__);                                       \
        if (s_ != BPT_OK) return s_;                                         \
    } while (0)

inline hipStream_t as_stream(bpt_stream s) { return reinterpret_cast<hipStream_t>(s); }

int device_cu_count();  // CUs of the current device (cached per device)

// One piece of a descriptor copy: `bytes` (a multiple of 4) from src to dst,
// both device memory aligned to 4 bytes (16 for the wide path).
struct CopyDesc {
    const void *src;
    void *dst;
    uint64_t bytes;
};
// Copies the nd pieces of d_desc (a device array) in one launch; max_bytes
// bounds the largest piece (it sizes the grid).
bpt_status launch_desc_copy(const CopyDesc *d_desc, uint32_t nd, uint64_t max_bytes, hipStream_t s);

// ---------------------------------------------------------------- layout
constexpr int kMaxLevels = BPT_MAX_LEVELS;

// What a search kernel needs to walk a tree; passed by value (kernarg).
struct TreeView {
    const uint8_t *inner;   // internal levels, top (root) level first
    const uint8_t *leaves;  // leaf lines
    uint64_t n;             // keys
    uint64_t leaf_stride;   // bytes per leaf
    int32_t height;         // internal levels (0: the single leaf is the root)
    int32_t lds_levels;     // top internal levels copied into LDS
    uint32_t lds_bytes;     // bytes of `inner` copied into LDS (a prefix)
    uint32_t sub_slots;     // slots per sub-leaf of the sub-leaf maxima (0: none)
    uint64_t count[kMaxLevels + 1];  // nodes per level, [0] = leaves
    uint64_t off[kMaxLevels + 1];    // byte offset of level k inside `inner`
    uint64_t sub_off;                // byte offset of the sub-leaf maxima inside `inner`
};

struct Workspace {
    void *ptr = nullptr;
    uint64_t bytes = 0;
    ~Workspace();
    bpt_status reserve(uint64_t need);
};

struct LearnedModel;  // learned position model (bpt_learned.hip)

}  // namespace bpt

struct bpt_tree {
    uint32_t key_bytes = 4, val_bytes = 4, fanout = 32, leaf_cap = 16;
    uint64_t n = 0;
    uint64_t leaf_stride = 128;
    uint8_t *d_leaves = nullptr;
    uint8_t *d_inner = nullptr;
    uint64_t leaf_bytes = 0, inner_bytes = 0;
    uint64_t leaf_alloc = 0, inner_alloc = 0;  // allocated bytes (a rebuild reuses what fits)
    int device = 0;
    uint32_t block = 256, blocks_per_cu = 32;
    uint32_t group_bits = 12;
    uint32_t lds_budget = 32768;
    uint32_t coop = 1;               // cooperative descent when key and value widths match
    uint32_t leaf_nt = 1;            // non-temporal leaf loads
    uint32_t coop_lanes = 8;         // lanes per query of the cooperative descent (4 or 8)
    uint32_t addr32 = 1;             // 32-bit node / leaf offsets when the tree allows them
    uint32_t group_chunk = 0;        // grouped search: queries per work chunk (0: by batch size)
    uint32_t group_xcd = 1;          // grouped search: chunks split per XCD
    uint32_t group_lanes = 4;        // grouped search: lanes per query at the leaf (4 or 8)
    uint32_t group_blocks = 1;       // grouped search: search blocks per CU (at most what LDS holds)
    uint32_t group_threads = 1024;   // grouped search: threads per search block (512, 768 or 1024)
    uint32_t group_slice = 131072;   // grouped search: LDS slice bytes per bucket (16 to 128 KiB)
    uint32_t group_sub = 1;          // grouped search: slices of sub-leaf maxima when the tree has them
    uint64_t group_min = 1ull << 24; // grouped search: smaller batches run the plain descent
    uint64_t alloc_limit = 0;        // largest leaf / inner buffer a rebuild may allocate (0: none)
    // Grouped search: the bucket tables of the default plan (bucket
    // separators, their prefix table, every bucket's slice prefix table),
    // built with the tree and by every update, so that a search need not
    // rebuild them. bkt_sig holds the plan parameters they were built for; a
    // plan with other parameters (other options) builds its own per call.
    uint8_t *d_bkt = nullptr;
    uint64_t bkt_alloc = 0;
    uint64_t bkt_sig[6] = {};
    bool bkt_valid = false;
    bpt::TreeView view{};
    bpt::Workspace ws;             // grouped / learned scratch owned by the tree
    bpt::LearnedModel *learned = nullptr;
};

namespace bpt {

// ---------------------------------------------------------------- dispatch
template <int N> using IC = std::integral_constant<int, N>;
template <class T> struct Tag { using type = T; };

// Calls fn(Tag<K>, Tag<V>, IC<F>, IC<LF>) for the tree's configuration.
template <class Fn>
bpt_status dispatch_tree(const bpt_tree *t, Fn &&fn) {
    auto with_f_lf = [&](auto kt, auto vt) -> bpt_status {
        auto with_lf = [&](auto fc) -> bpt_status {
            switch (t->leaf_cap) {
                case 8: return fn(kt, vt, fc, IC<8>{});
                case 16: return fn(kt, vt, fc, IC<16>{});
                case 32: return fn(kt, vt, fc, IC<32>{});
            }
            set_error("unsupported leaf_cap %u", t->leaf_cap);
            return BPT_ERR_INVALID;
        };
        switch (t->fanout) {
            case 16: return with_lf(IC<16>{});
            case 32: return with_lf(IC<32>{});
            case 64: return with_lf(IC<64>{});
        }
        set_error("unsupported fanout %u", t->fanout);
        return BPT_ERR_INVALID;
    };
    if (t->key_bytes == 4 && t->val_bytes == 4) return with_f_lf(Tag<uint32_t>{}, Tag<uint32_t>{});
    if (t->key_bytes == 8 && t->val_bytes == 8) return with_f_lf(Tag<uint64_t>{}, Tag<uint64_t>{});
    if (t->key_bytes == 8 && t->val_bytes == 4) return with_f_lf(Tag<uint64_t>{}, Tag<uint32_t>{});
    set_error("unsupported key/value widths %u/%u", t->key_bytes, t->val_bytes);
    return BPT_ERR_INVALID;
}

template <class K> __host__ __device__ constexpr K key_max() { return ~K(0); }

// ---------------------------------------------------------------- device helpers
// Number of separators in an F-wide node that are < q. The node is
// F * sizeof(K) bytes, 16-byte aligned; it is read with 128-bit loads, all
// issued before any compare so the loads of one node overlap.
template <class K, int F>
__device__ __forceinline__ uint32_t node_rank(const K *__restrict__ node, K q) {
    constexpr int NV = F * (int)sizeof(K) / 16;
    const uint4 *p = reinterpret_cast<const uint4 *>(node);
    uint4 w[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) w[v] = p[v];
    uint32_t c = 0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        if constexpr (sizeof(K) == 4) {
            c += (w[v].x < q) + (w[v].y < q) + (w[v].z < q) + (w[v].w < q);
        } else {
            uint64_t a = (uint64_t)w[v].x | ((uint64_t)w[v].y << 32);
            uint64_t b = (uint64_t)w[v].z | ((uint64_t)w[v].w << 32);
            c += (a < q) + (b < q);
        }
    }
    return c;
}

// Leaf node index of q: walks the LDS-resident top levels, then the
// global levels. `lds` holds the first t.lds_bytes of t.inner.
template <class K, int F>
__device__ __forceinline__ uint64_t descend(const TreeView &t, const uint8_t *lds, K q) {
    uint64_t node = 0;
    int lvl = t.height;
    const int lds_stop = t.height - t.lds_levels;
    for (; lvl > lds_stop; --lvl) {
        const K *nd = reinterpret_cast<const K *>(lds + t.off[lvl]) + node * F;
        uint64_t c = node * F + node_rank<K, F>(nd, q);
        uint64_t last = t.count[lvl - 1] - 1;
        node = c < last ? c : last;
    }
    for (; lvl >= 1; --lvl) {
        const K *nd = reinterpret_cast<const K *>(t.inner + t.off[lvl]) + node * F;
        uint64_t c = node * F + node_rank<K, F>(nd, q);
        uint64_t last = t.count[lvl - 1] - 1;
        node = c < last ? c : last;
    }
    return node;
}

// ---------------------------------------------------------------- leaf format
// A leaf holds LF (key, value) slots. With equal key and value widths the
// slots are interleaved in 16-byte pieces -- [k0 k1 v0 v1] for 4-byte slots,
// [k0 v0] for 8-byte slots -- so every 16-byte piece holds whole slots and a
// lane that loads one piece can answer for its slots alone. Otherwise (8-byte
// keys, 4-byte values) the LF keys come first, then the LF values.
template <class K, class V, int LF>
struct LeafFmt {
    static constexpr bool kInter = sizeof(K) == sizeof(V);
    static constexpr uint32_t kPair = kInter ? 16u / (2u * (uint32_t)sizeof(K)) : 1u;  // slots per piece
    __host__ __device__ static constexpr uint32_t key_off(uint32_t j) {
        return kInter ? (j / kPair) * 16u + (j % kPair) * (uint32_t)sizeof(K) : j * (uint32_t)sizeof(K);
    }
    __host__ __device__ static constexpr uint32_t val_off(uint32_t j) {
        return kInter ? (j / kPair) * 16u + kPair * (uint32_t)sizeof(K) + (j % kPair) * (uint32_t)sizeof(V)
                      : LF * (uint32_t)sizeof(K) + j * (uint32_t)sizeof(V);
    }
};

// Key and value of slot i of a run of whole 16-byte pieces held in 32-bit
// words (interleaved format).
template <class K>
__device__ __forceinline__ K piece_key(const uint32_t *w, uint32_t i) {
    if constexpr (sizeof(K) == 4) return w[(i / 2) * 4 + (i % 2)];
    else return (uint64_t)w[i * 4] | ((uint64_t)w[i * 4 + 1] << 32);
}
template <class V>
__device__ __forceinline__ V piece_val(const uint32_t *w, uint32_t i) {
    if constexpr (sizeof(V) == 4) return w[(i / 2) * 4 + 2 + (i % 2)];
    else return (uint64_t)w[i * 4 + 2] | ((uint64_t)w[i * 4 + 3] << 32);
}

// All keys (and values) of a leaf in registers, one lane.
template <class K, class V, int LF>
__device__ __forceinline__ void load_leaf(const uint8_t *lp, K (&k)[LF], V *v) {
    using Fmt = LeafFmt<K, V, LF>;
    if constexpr (Fmt::kInter) {
        constexpr int NP = LF * 2 * (int)sizeof(K) / 16;
        uint32_t w[NP * 4];
        const uint4 *p4 = reinterpret_cast<const uint4 *>(lp);
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const uint4 x = p4[i];
            w[4 * i] = x.x; w[4 * i + 1] = x.y; w[4 * i + 2] = x.z; w[4 * i + 3] = x.w;
        }
#pragma unroll
        for (int j = 0; j < LF; ++j) {
            k[j] = piece_key<K>(w, j);
            if (v) v[j] = piece_val<V>(w, j);
        }
    } else {
        const uint4 *p4 = reinterpret_cast<const uint4 *>(lp);
#pragma unroll
        for (int i = 0; i < LF * (int)sizeof(K) / 16; ++i) {
            const uint4 x = p4[i];
            if constexpr (sizeof(K) == 4) {
                k[4 * i] = x.x; k[4 * i + 1] = x.y; k[4 * i + 2] = x.z; k[4 * i + 3] = x.w;
            } else {
                k[2 * i] = (uint64_t)x.x | ((uint64_t)x.y << 32);
                k[2 * i + 1] = (uint64_t)x.z | ((uint64_t)x.w << 32);
            }
        }
        if (v) {
            const uint4 *q4 = reinterpret_cast<const uint4 *>(lp + LF * sizeof(K));
#pragma unroll
            for (int i = 0; i < LF * (int)sizeof(V) / 16; ++i) {
                const uint4 x = q4[i];
                if constexpr (sizeof(V) == 4) {
                    v[4 * i] = x.x; v[4 * i + 1] = x.y; v[4 * i + 2] = x.z; v[4 * i + 3] = x.w;
                } else {
                    v[2 * i] = (uint64_t)x.x | ((uint64_t)x.y << 32);
                    v[2 * i + 1] = (uint64_t)x.z | ((uint64_t)x.w << 32);
                }
            }
        }
    }
}

// Exact lookup of q in leaf `leaf`: value of the first slot equal to q
// (slots past n never match), else `sentinel`.
template <class K, class V, int LF>
__device__ __forceinline__ V leaf_lookup(const TreeView &t, uint64_t leaf, K q, V sentinel) {
    K k[LF];
    V v[LF];
    load_leaf<K, V, LF>(t.leaves + leaf * t.leaf_stride, k, v);
    const uint64_t base = leaf * LF;
    V res = sentinel;
#pragma unroll
    for (int j = LF - 1; j >= 0; --j)
        if (k[j] == q && base + j < t.n) res = v[j];
    return res;
}

// One lane per query from node `node` of level `top` (top = 0: node is the
// leaf) down through the global levels to the value.
template <class K, class V, int F, int LF>
__device__ __forceinline__ V thread_finish(const TreeView &t, K q, uint64_t node, int top, V sentinel) {
    for (int lvl = top; lvl >= 1; --lvl) {
        const K *nd = reinterpret_cast<const K *>(t.inner + t.off[lvl]) + node * F;
        const uint64_t c = node * F + node_rank<K, F>(nd, q);
        const uint64_t last = t.count[lvl - 1] - 1;
        node = c < last ? c : last;
    }
    return leaf_lookup<K, V, LF>(t, node, q, sentinel);
}

// Global lower_bound position of q inside leaf `leaf`.
template <class K, class V, int LF>
__device__ __forceinline__ uint64_t leaf_lower_bound(const TreeView &t, uint64_t leaf, K q) {
    K k[LF];
    load_leaf<K, V, LF>(t.leaves + leaf * t.leaf_stride, k, (V *)nullptr);
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < LF; ++j) c += (k[j] < q);
    uint64_t p = leaf * LF + c;
    return p < t.n ? p : t.n;
}

// Slot i of the sorted key / value sequence, read from the leaf layout.
template <class K, class V, int LF>
__device__ __forceinline__ K slot_key(const TreeView &t, uint64_t i) {
    return *reinterpret_cast<const K *>(t.leaves + (i / LF) * t.leaf_stride +
                                        LeafFmt<K, V, LF>::key_off((uint32_t)(i % LF)));
}
template <class K, class V, int LF>
__device__ __forceinline__ V slot_val(const TreeView &t, uint64_t i) {
    return *reinterpret_cast<const V *>(t.leaves + (i / LF) * t.leaf_stride +
                                        LeafFmt<K, V, LF>::val_off((uint32_t)(i % LF)));
}

// Copies the LDS prefix of the internal levels into shared memory.
__device__ __forceinline__ void stage_top_levels(const TreeView &t, uint8_t *lds) {
    const uint4 *src = reinterpret_cast<const uint4 *>(t.inner);
    uint4 *dst = reinterpret_cast<uint4 *>(lds);
    for (uint32_t i = threadIdx.x; i < t.lds_bytes / 16; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}

// ---------------------------------------------------------------- cooperative descent
// Eight lanes share one node: lane s of a group holds bytes [s*NB, (s+1)*NB)
// of the node, so one load instruction of a wave reads eight whole nodes
// (eight fully used 128-byte lines) instead of one 16-byte piece of 64 lines.
template <int BYTES>
struct Chunk {
    static constexpr int W = BYTES / 4;
    uint32_t w[W];
};

template <int BYTES>
__device__ __forceinline__ void load_chunk(const uint8_t *p, Chunk<BYTES> &c) {
    if constexpr (BYTES == 8) {
        const uint2 v = *reinterpret_cast<const uint2 *>(p);
        c.w[0] = v.x; c.w[1] = v.y;
    } else {
#pragma unroll
        for (int i = 0; i < BYTES / 16; ++i) {
            const uint4 v = reinterpret_cast<const uint4 *>(p)[i];
            c.w[4 * i] = v.x; c.w[4 * i + 1] = v.y; c.w[4 * i + 2] = v.z; c.w[4 * i + 3] = v.w;
        }
    }
}

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

// Same as load_chunk with the non-temporal hint: lines read once (leaves)
// are marked streaming so they do not displace the reused upper levels.
template <int BYTES>
__device__ __forceinline__ void load_chunk_nt(const uint8_t *p, Chunk<BYTES> &c) {
    if constexpr (BYTES == 8) {
        const u32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x2_t *>(p));
        c.w[0] = v.x; c.w[1] = v.y;
    } else {
#pragma unroll
        for (int i = 0; i < BYTES / 16; ++i) {
            const u32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t *>(p) + i);
            c.w[4 * i] = v.x; c.w[4 * i + 1] = v.y; c.w[4 * i + 2] = v.z; c.w[4 * i + 3] = v.w;
        }
    }
}

// Same loads through a buffer resource: `base` is uniform and `off` a 32-bit
// per-lane byte offset (no 64-bit address arithmetic per lane).
template <int BYTES, bool NT>
__device__ __forceinline__ void load_chunk_buf(__amdgpu_buffer_rsrc_t r, uint32_t off, Chunk<BYTES> &c) {
    constexpr int aux = NT ? 2 : 0;  // cache policy: non-temporal
    if constexpr (BYTES == 8) {
        const u32x2_t v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, aux);
        c.w[0] = v.x; c.w[1] = v.y;
    } else {
#pragma unroll
        for (int i = 0; i < BYTES / 16; ++i) {
            const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16u * i, 0, aux);
            c.w[4 * i] = v.x; c.w[4 * i + 1] = v.y; c.w[4 * i + 2] = v.z; c.w[4 * i + 3] = v.w;
        }
    }
}

// Raw buffer resource over [base, base + 4 GiB): stride 0, 32-bit data format
// word of the gfx9 descriptor.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)0xFFFFFFFF, 0x00020000);
}

// Element i mod N of a chunk of N keys, as a chain of compare-selects
// (kept as selects: an indexed register array would be lowered to scratch).
template <class K, int BYTES>
__device__ __forceinline__ K chunk_elem(const Chunk<BYTES> &c, uint32_t i) {
    constexpr int N = BYTES / (int)sizeof(K);
    static_assert((N & (N - 1)) == 0, "power-of-two chunk");
    const uint32_t j = i & (uint32_t)(N - 1);
    K r;
    if constexpr (sizeof(K) == 4) r = c.w[0];
    else r = (uint64_t)c.w[0] | ((uint64_t)c.w[1] << 32);
#pragma unroll
    for (int e = 1; e < N; ++e) {
        K v;
        if constexpr (sizeof(K) == 4) v = c.w[e];
        else v = (uint64_t)c.w[2 * e] | ((uint64_t)c.w[2 * e + 1] << 32);
        r = j == (uint32_t)e ? v : r;
    }
    return r;
}

template <class K, int BYTES>
__device__ __forceinline__ uint32_t chunk_count_lt(const Chunk<BYTES> &c, K q) {
    constexpr int N = BYTES / (int)sizeof(K);
    uint32_t n = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        K v;
        if constexpr (sizeof(K) == 4) v = c.w[j];
        else v = (uint64_t)c.w[2 * j] | ((uint64_t)c.w[2 * j + 1] << 32);
        n += (v < q);
    }
    return n;
}

template <class K, int BYTES>
__device__ __forceinline__ uint32_t chunk_count_eq(const Chunk<BYTES> &c, K q) {
    constexpr int N = BYTES / (int)sizeof(K);
    uint32_t n = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        K v;
        if constexpr (sizeof(K) == 4) v = c.w[j];
        else v = (uint64_t)c.w[2 * j] | ((uint64_t)c.w[2 * j + 1] << 32);
        n += (v == q);
    }
    return n;
}

// Sum over the 8 lanes of a group with DPP lane moves (no LDS traffic):
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], then row_half_mirror, which
// pairs lane i with lane 7-i of the same 8-lane half row (the other quad).
// Every lane has a valid source under these patterns, so bound_ctrl (which
// only matters for invalid sources) is set and no old value is needed.
__device__ __forceinline__ uint32_t group8_sum(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, true);
    return x;
}

// Sum over the G lanes of an aligned group (G = 4: one quad; G = 8).
template <int G>
__device__ __forceinline__ uint32_t group_sum(uint32_t x) {
    static_assert(G == 4 || G == 8, "group of 4 or 8 lanes");
    if constexpr (G == 8) return group8_sum(x);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, true);
    return x;
}

// Branch-free lower bound over n sorted keys (n >= 1): the loop trip count
// depends only on n, so the lanes of a wave stay converged.
template <class K>
__device__ __forceinline__ uint32_t lower_bound_flat(const K *__restrict__ a, uint32_t n, K q) {
    uint32_t base = 0;
    while (n > 1) {
        const uint32_t half = n >> 1;
        base = (a[base + half] < q) ? base + half : base;
        n -= half;
    }
    return base + (a[base] < q);
}

// ---------------------------------------------------------------- cooperative lookup
// One wave looks up 64 queries (one per lane, `myq`) in tree t.
// Phase A (one lane per query): the LDS-resident top levels are walked as a
// single branch-free lower bound over the flattened separators of the lowest
// LDS level -- the separators of one level, concatenated in node order, are
// sorted (pads only trail the last node), and the first separator >= q there
// is exactly the child the root-to-that-level descent reaches.
// Phase B (G lanes per query, G = 8 or 4): group g (lanes Gg..Gg+G-1) walks
// queries Gg..Gg+G-1 through the global levels and the leaf, all G in flight
// at each level; lane s of the group reads bytes [s*NB, (s+1)*NB) of a node,
// so one load instruction fetches 64/G node pieces of whole lines; the leaf
// step is coop_leaf below. Every step is branch-free so the G queries
// interleave.
// OFF32: node and leaf byte offsets fit 32 bits (levels and leaves < 4 GiB),
// so addresses are a uniform base plus a 32-bit lane offset.
// OP kOpValue returns the value of the first key equal to the lane's query
// (else `sentinel`); kOpLowerBound returns the global lower-bound position.
constexpr int kOpValue = 0;
constexpr int kOpLowerBound = 1;

// Leaf step of the cooperative lookup. Group g (lanes Gg..Gg+G-1) holds the
// queries qk[0..G) of its lanes and their leaves node[0..G). Lane s of the
// group loads bytes [s*LB, (s+1)*LB) of each leaf -- whole (key, value) slots
// in the interleaved format -- so one load instruction fetches 64/G leaf
// pieces of whole lines, and the G loads of a query's leaf cover it once.
// The rank r of the query in its leaf is the DPP sum of the lanes' counts; the
// lane whose slots contain r checks the key there and hands its value (or the
// sentinel) to the query's lane with one ds_bpermute. Returns the result of
// the lane's own query.
template <class K, class V, int LF, bool NT, int OP, int G, bool OFF32>
__device__ __forceinline__ typename std::conditional<OP == kOpValue, V, uint64_t>::type
coop_leaf(const TreeView &t, const K (&qk)[G], const uint32_t (&node)[G], V sentinel) {
    static_assert(LeafFmt<K, V, LF>::kInter, "cooperative leaf needs the interleaved format");
    constexpr uint32_t STRIDE = 2 * LF * sizeof(K);
    constexpr int GL = (int)(STRIDE / 16) < G ? (int)(STRIDE / 16) : G;  // lanes holding leaf pieces
    constexpr int LB = (int)STRIDE / GL;               // leaf bytes per such lane
    constexpr uint32_t PPL = LB / (2 * sizeof(K));     // slots per such lane
    static_assert(LB >= 16 && LB % 16 == 0, "whole 16-byte pieces per lane");
    using R = typename std::conditional<OP == kOpValue, V, uint64_t>::type;
    const uint32_t lane = threadIdx.x & 63, g = lane / G, sub = lane % G;
    const bool holds = sub < (uint32_t)GL;
    Chunk<LB> c[G];
#pragma unroll
    for (int k = 0; k < G; ++k)
#pragma unroll
        for (int i = 0; i < LB / 4; ++i) c[k].w[i] = ~0u;  // lanes past GL hold no slots
    if (holds) {
        if constexpr (OFF32) {
            const __amdgpu_buffer_rsrc_t rs = raw_rsrc(t.leaves);
#pragma unroll
            for (int k = 0; k < G; ++k) load_chunk_buf<LB, NT>(rs, node[k] * STRIDE + sub * LB, c[k]);
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const uint8_t *lp = t.leaves + (uint64_t)node[k] * STRIDE + sub * LB;
                if constexpr (NT) load_chunk_nt<LB>(lp, c[k]);
                else load_chunk<LB>(lp, c[k]);
            }
        }
    }
    // The leaf's loads are out; ranking it and storing the result are off the
    // critical path: back to the base priority (the search kernels raise it
    // while a batch walks to its leaf, so that the SIMD issues first for the
    // waves whose next step is a memory access).
    __builtin_amdgcn_s_setprio(0);
    // slot leaf*LF + r holds a key iff leaf < last_leaf or r < last_cnt
    const uint32_t last_leaf = (uint32_t)(t.count[0] - 1);
    const uint32_t last_cnt = (uint32_t)(t.n - (uint64_t)last_leaf * LF);
    R res = OP == kOpValue ? (R)sentinel : (R)0;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        uint32_t cnt = 0;
#pragma unroll
        for (uint32_t i = 0; i < PPL; ++i) cnt += (piece_key<K>(c[k].w, i) < qk[k]);
        const uint32_t r = group_sum<G>(holds ? cnt : 0u);  // lower bound inside the leaf, 0..LF
        if constexpr (OP == kOpValue) {
            const uint32_t local = r - sub * PPL;  // r's slot in this lane, if < PPL
            K key = piece_key<K>(c[k].w, 0);
            V val = piece_val<V>(c[k].w, 0);
#pragma unroll
            for (uint32_t i = 1; i < PPL; ++i) {
                key = local == i ? piece_key<K>(c[k].w, i) : key;
                val = local == i ? piece_val<V>(c[k].w, i) : val;
            }
            const bool hit = holds && local < PPL && key == qk[k] && (node[k] < last_leaf || r < last_cnt);
            const V cand = hit ? val : sentinel;
            const uint32_t owner = r / PPL < (uint32_t)GL ? r / PPL : (uint32_t)GL - 1;  // r = LF: sentinel
            const V got = __shfl(cand, (int)(g * G + owner), 64);
            res = sub == (uint32_t)k ? got : res;
        } else {
            const uint64_t p = (uint64_t)node[k] * LF + r;
            res = sub == (uint32_t)k ? (p < t.n ? p : t.n) : res;
        }
    }
    return res;
}

// Inner levels of phase B: group g (lanes Gg..Gg+G-1) walks its G queries
// qk[] from nodes node[] of level `top` down to their leaves (node[] on
// return), one load instruction per level fetching 64/G whole nodes.
template <class K, int F, int G, bool OFF32>
__device__ __forceinline__ void coop_inner(const TreeView &t, const K (&qk)[G], uint32_t (&node)[G], int top) {
    static_assert(G == 4 || G == 8, "4 or 8 lanes per query");
    constexpr int NB = F * (int)sizeof(K) / G;  // node bytes per lane
    constexpr uint32_t NODE = F * sizeof(K);
    static_assert(NB >= 8, "chunks of at least 8 bytes");
    const uint32_t sub = (threadIdx.x & 63) % G;
    for (int lvl = top; lvl >= 1; --lvl) {
        const uint8_t *lv = t.inner + t.off[lvl];
        const uint32_t last = (uint32_t)t.count[lvl - 1] - 1;
        Chunk<NB> c[G];
        if constexpr (OFF32) {
            const __amdgpu_buffer_rsrc_t rs = raw_rsrc(lv);
#pragma unroll
            for (int k = 0; k < G; ++k) load_chunk_buf<NB, false>(rs, node[k] * NODE + sub * NB, c[k]);
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k) load_chunk<NB>(lv + (uint64_t)node[k] * NODE + sub * NB, c[k]);
        }
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const uint32_t ch = node[k] * F + group_sum<G>(chunk_count_lt<K, NB>(c[k], qk[k]));
            node[k] = ch < last ? ch : last;
        }
    }
}

// Phase B alone: the lane's query `myq` sits in node `mynode` of level `top`
// (top = 0: mynode is its leaf); walks levels top..1 and the leaf.
template <class K, class V, int F, int LF, bool NT, int OP, int G = 8, bool OFF32 = false>
__device__ __forceinline__ typename std::conditional<OP == kOpValue, V, uint64_t>::type
coop_finish64(const TreeView &t, K myq, uint32_t mynode, int top, V sentinel) {
    static_assert(sizeof(K) == sizeof(V), "cooperative lookup needs equal key/value widths");
    const uint32_t lane = threadIdx.x & 63, g = lane / G;
    K qk[G];
    uint32_t node[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        qk[k] = __shfl(myq, (int)(g * G + k), 64);
        node[k] = (uint32_t)__shfl((int)mynode, (int)(g * G + k), 64);
    }
    coop_inner<K, F, G, OFF32>(t, qk, node, top);
    return coop_leaf<K, V, LF, NT, OP, G, OFF32>(t, qk, node, sentinel);
}

// Phase A: the node of the lane's query at the first level walked
// cooperatively (*top; 0 = a leaf), from the LDS-resident levels.
template <class K>
__device__ __forceinline__ uint32_t coop_phase_a(const TreeView &t, const uint8_t *lds, K myq, int *top) {
    const int flat_lvl = t.height - t.lds_levels + 1;
    const bool has_flat = t.lds_levels > 0;
    uint32_t mynode = 0;
    if (has_flat) {
        const K *flat = reinterpret_cast<const K *>(lds + t.off[flat_lvl]);
        const uint32_t flat_n = (uint32_t)t.count[flat_lvl - 1];
        const uint32_t c = lower_bound_flat<K>(flat, flat_n, myq);
        mynode = c < flat_n - 1 ? c : flat_n - 1;
    }
    *top = has_flat ? flat_lvl - 1 : t.height;
    return mynode;
}

// Phase A (LDS levels, one lane per query) then phase B.
template <class K, class V, int F, int LF, bool NT, int OP, int G = 8, bool OFF32 = false>
__device__ __forceinline__ typename std::conditional<OP == kOpValue, V, uint64_t>::type
coop_lookup64(const TreeView &t, const uint8_t *lds, K myq, V sentinel) {
    int top;
    const uint32_t mynode = coop_phase_a<K>(t, lds, myq, &top);
    return coop_finish64<K, V, F, LF, NT, OP, G, OFF32>(t, myq, mynode, top, sentinel);
}

// The leaf of the lane's query: phase A, then the inner levels of phase B
// with G lanes per node (any leaf format).
template <class K, int F, int G, bool OFF32>
__device__ __forceinline__ uint32_t coop_leaf_of(const TreeView &t, const uint8_t *lds, K myq) {
    int top;
    const uint32_t mynode = coop_phase_a<K>(t, lds, myq, &top);
    const uint32_t lane = threadIdx.x & 63, g = lane / G, sub = lane % G;
    K qk[G];
    uint32_t node[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        qk[k] = __shfl(myq, (int)(g * G + k), 64);
        node[k] = (uint32_t)__shfl((int)mynode, (int)(g * G + k), 64);
    }
    coop_inner<K, F, G, OFF32>(t, qk, node, top);
    uint32_t mine = node[0];
#pragma unroll
    for (int k = 1; k < G; ++k) mine = sub == (uint32_t)k ? node[k] : mine;
    return mine;
}

// ---------------------------------------------------------------- host helpers
uint32_t search_grid(const bpt_tree *t, uint64_t work_items);
bpt_status launch_plain(const bpt_tree *t, const void *d_q, uint64_t nq, void *d_out,
                        hipStream_t s);
template <class T>
bpt_status exclusive_scan(const T *d_in, uint64_t n, T *d_out, void *d_tmp, uint64_t tmp_bytes,
                          hipStream_t s);
uint64_t scan_tmp_bytes(uint64_t n, uint64_t elem_bytes);
// LSD radix sort (8-bit stable passes over bits [lo_bit, hi_bit)) of keys with
// u32 values (original indices when vals_in is NULL); results point into ws.
uint64_t sort_ws_bytes(uint64_t n, uint32_t key_bytes);
template <class KT>
bpt_status sort_pairs(const KT *keys_in, const uint32_t *vals_in, uint64_t n, uint32_t lo_bit, uint32_t hi_bit,
                      void *ws, uint64_t ws_bytes, KT **keys_sorted, uint32_t **vals_sorted, hipStream_t s);
// Maxima of two arrays (pass a twice for one); synchronizes s.
template <class KT>
bpt_status pair_max(const KT *a, const KT *b, uint64_t n, uint64_t *ma, uint64_t *mb, hipStream_t s);
// The reference's grouping pass loop over device arrays of n uint64
// (bpt_msd_group.hip): kind 0 buckets arrays[0] by bytes (grouping.cu),
// kind 1 buckets (arrays[0], arrays[1]) by nibble products (grouping2d.cu).
// Every array is permuted with the keys; list / prefix receive the groups.
bpt_status msd_group(uint64_t *const *arrays, int narrays, int kind, uint64_t n, uint64_t cap, int start_pass,
                     std::vector<uint64_t> &list, std::vector<uint64_t> &prefix, hipStream_t s);
int group_start_pass_1d(uint64_t maxval);
int group_start_pass_2d(uint64_t maxa, uint64_t maxb);
bpt_status point_grouped(const bpt_tree *t, const void *d_q, uint64_t nq, void *d_out,
                         void *d_ws, uint64_t ws_bytes, hipStream_t s);
uint64_t point_grouped_ws(const bpt_tree *t, uint64_t nq);
// (Re)builds t's bucket tables on stream s (the caller synchronizes s before
// the tree is searched). Without memory for them the tree keeps none and
// grouped searches build them per call: not an error.
bpt_status bucket_tables_refresh(bpt_tree *t, hipStream_t s);
bpt_status point_learned(const bpt_tree *t, const void *d_q, uint64_t nq, void *d_out,
                         void *d_ws, uint64_t ws_bytes, hipStream_t s);
uint64_t point_learned_ws(const bpt_tree *t, uint64_t nq);
void learned_free(LearnedModel *m);
// Replaces the tree's contents with n sorted (key, value) pairs on the device,
// keeping its widths, shape and options; drops a learned model. Synchronizes s.
bpt_status tree_rebuild(bpt_tree *t, const void *d_keys, const void *d_vals, uint64_t n, hipStream_t s);

}  // namespace bpt
