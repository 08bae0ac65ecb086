// Plain point search, lower bound, 1-D range count / ids, exclusive scan.
//
// Plain search replaces kernelSearchingNaiveOneStep (reference
// BPTreeExactSearch.cu:25-105): one query per lane, root-to-leaf descent over
// child maxima, exact match in the leaf, sentinel on a miss. The reference
// binary-searches 5000-key pages with ~40 dependent 8-byte probes per query;
// here a node is one 128-byte line read with eight independent 16-byte loads
// and ranked with branch-free compares, the top levels come from LDS, and a
// lookup costs one dependent line per global level plus one leaf line.
#include <algorithm>

#include "bpt_internal.hpp"

namespace bpt {

template <class K, class V, int F, int LF>
__global__ __launch_bounds__(256) void k_point_plain(TreeView t, const K *__restrict__ q,
                                                     V *__restrict__ out, uint64_t nq, V sentinel) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += step) {
        const K key = __builtin_nontemporal_load(&q[i]);
        const uint64_t leaf = descend<K, F>(t, lds, key);
        const V r = leaf_lookup<K, V, LF>(t, leaf, key, sentinel);
        __builtin_nontemporal_store(r, &out[i]);
    }
}

// Cooperative single-tree kernel: every wave walks 64 consecutive queries
// at a time with coop_lookup64 (bpt_internal.hpp), G lanes per query.
template <class K, class V, int F, int LF, bool NT, int G, bool OFF32>
__global__ __launch_bounds__(256) void k_point_coop(TreeView t, const K *__restrict__ q,
                                                    V *__restrict__ out, uint64_t nq, V sentinel) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    // the next batch's queries are loaded while this batch is searched
    const uint64_t stride = nwaves * 64;
    uint64_t base = wave * 64;
    K next = base + lane < nq ? __builtin_nontemporal_load(&q[base + lane]) : K(0);
    for (; base < nq; base += stride) {
        const uint64_t my = base + lane;
        const K myq = next;
        next = my + stride < nq ? __builtin_nontemporal_load(&q[my + stride]) : K(0);
        __builtin_amdgcn_s_setprio(2);  // until the leaf's loads are issued (coop_leaf)
        const V res = coop_lookup64<K, V, F, LF, NT, kOpValue, G, OFF32>(t, lds, myq, sentinel);
        if (my < nq) __builtin_nontemporal_store(res, &out[my]);
    }
}

// Trees whose key and value widths differ (a leaf holds its keys, then its
// values): the inner levels cooperatively, G lanes per node, then the leaf
// by the query's own lane.
template <class K, class V, int F, int LF, int G, bool OFF32>
__global__ __launch_bounds__(256) void k_point_coop_inner(TreeView t, const K *__restrict__ q,
                                                          V *__restrict__ out, uint64_t nq, V sentinel) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t stride = ((uint64_t)gridDim.x * blockDim.x) & ~63ull;
    for (uint64_t base = wave * 64; base < nq; base += stride) {
        const uint64_t j = base + lane;
        const K key = j < nq ? __builtin_nontemporal_load(&q[j]) : K(0);
        const uint32_t leaf = coop_leaf_of<K, F, G, OFF32>(t, lds, key);
        const V r = leaf_lookup<K, V, LF>(t, leaf, key, sentinel);
        if (j < nq) __builtin_nontemporal_store(r, &out[j]);
    }
}

template <class K, class V, int F, int LF, int G, bool OFF32>
__global__ __launch_bounds__(256) void k_range_count_coop_inner(TreeView t, const K *__restrict__ lo,
                                                                const K *__restrict__ hi,
                                                                const uint32_t *__restrict__ order,
                                                                uint64_t *__restrict__ first,
                                                                uint64_t *__restrict__ count, uint64_t nq) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t stride = ((uint64_t)gridDim.x * blockDim.x) & ~63ull;
    for (uint64_t base = wave * 64; base < nq; base += stride) {
        const uint64_t j = base + lane;
        const bool valid = j < nq;
        const uint64_t i = valid ? (order ? (uint64_t)order[j] : j) : 0;
        const K a = valid ? lo[i] : K(0), b = valid ? hi[i] : K(0);
        const bool top = b == key_max<K>();
        const K b1 = top ? b : K(b + 1);
        const uint64_t pa = leaf_lower_bound<K, V, LF>(t, coop_leaf_of<K, F, G, OFF32>(t, lds, a), a);
        const uint64_t pb1 = leaf_lower_bound<K, V, LF>(t, coop_leaf_of<K, F, G, OFF32>(t, lds, b1), b1);
        const uint64_t pb = top ? t.n : pb1;
        if (valid) {
            if (first) first[i] = pa;
            count[i] = (a > b || pb < pa) ? 0 : pb - pa;
        }
    }
}

template <class K, class V, int F, int LF, int G, bool OFF32>
__global__ __launch_bounds__(256) void k_lower_bound_coop_inner(TreeView t, const K *__restrict__ q,
                                                                uint64_t *__restrict__ pos, uint64_t nq) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t stride = ((uint64_t)gridDim.x * blockDim.x) & ~63ull;
    for (uint64_t base = wave * 64; base < nq; base += stride) {
        const uint64_t j = base + lane;
        const K key = j < nq ? q[j] : K(0);
        const uint64_t p = leaf_lower_bound<K, V, LF>(t, coop_leaf_of<K, F, G, OFF32>(t, lds, key), key);
        if (j < nq) pos[j] = p;
    }
}

// Calls fn(IC<G>, bool_constant<OFF32>) for the tree's cooperative shape.
template <class Fn>
static void with_coop_shape(const bpt_tree *t, bool off32, Fn &&fn) {
    if (t->coop_lanes == 4) {
        if (off32) fn(IC<4>{}, std::true_type{});
        else fn(IC<4>{}, std::false_type{});
    } else {
        if (off32) fn(IC<8>{}, std::true_type{});
        else fn(IC<8>{}, std::false_type{});
    }
}

template <class K, class V, int F, int LF>
__global__ __launch_bounds__(256) void k_lower_bound(TreeView t, const K *__restrict__ q,
                                                     uint64_t *__restrict__ pos, uint64_t nq) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += step) {
        const K key = q[i];
        const uint64_t leaf = descend<K, F>(t, lds, key);
        pos[i] = leaf_lower_bound<K, V, LF>(t, leaf, key);
    }
}

// Lower bounds through the cooperative descent (equal key and value widths):
// 64 queries per wave at a time, G lanes per node.
template <class K, class V, int F, int LF, int G, bool OFF32>
__global__ __launch_bounds__(256) void k_lower_bound_coop(TreeView t, const K *__restrict__ q,
                                                          uint64_t *__restrict__ pos, uint64_t nq) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t stride = ((uint64_t)gridDim.x * blockDim.x) & ~63ull;
    for (uint64_t base = wave * 64; base < nq; base += stride) {
        const uint64_t j = base + lane;
        const K key = j < nq ? q[j] : K(0);
        const uint64_t p = coop_lookup64<K, V, F, LF, false, kOpLowerBound, G, OFF32>(t, lds, key, V(0));
        if (j < nq) pos[j] = p;
    }
}

// count = upper_bound(hi) - lower_bound(lo); upper_bound(hi) is
// lower_bound(hi + 1), or n when hi is the largest key value.
template <class K, class V, int F, int LF>
__global__ __launch_bounds__(256) void k_range_count(TreeView t, const K *__restrict__ lo,
                                                     const K *__restrict__ hi,
                                                     const uint32_t *__restrict__ order,
                                                     uint64_t *__restrict__ first,
                                                     uint64_t *__restrict__ count, uint64_t nq) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nq; j += step) {
        const uint64_t i = order ? order[j] : j;
        const K a = lo[i], b = hi[i];
        const uint64_t pa = leaf_lower_bound<K, V, LF>(t, descend<K, F>(t, lds, a), a);
        uint64_t pb = t.n;
        if (b != key_max<K>()) {
            const K b1 = b + 1;
            pb = leaf_lower_bound<K, V, LF>(t, descend<K, F>(t, lds, b1), b1);
        }
        if (first) first[i] = pa;
        count[i] = (a > b || pb < pa) ? 0 : pb - pa;
    }
}

// The same counts through the cooperative descent (equal key and value
// widths): each wave takes 64 queries at a time and runs both lower bounds
// with G lanes per node, so the two descents of a group are in flight
// together.
template <class K, class V, int F, int LF, int G, bool OFF32>
__global__ __launch_bounds__(256) void k_range_count_coop(TreeView t, const K *__restrict__ lo,
                                                          const K *__restrict__ hi,
                                                          const uint32_t *__restrict__ order,
                                                          uint64_t *__restrict__ first,
                                                          uint64_t *__restrict__ count, uint64_t nq) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    // 32 ranges per wave step: lane r < 32 descends for lower_bound(lo) of
    // range r and lane r + 32 for lower_bound(hi + 1), so the G-lane groups
    // of the two halves walk a range's two bounds in the same load
    // instructions, and a node or leaf that both bounds reach (level 2 nearly
    // always, often level 1 and the leaf) is one line request, not two.
    const uint32_t lane = threadIdx.x & 63, r = lane & 31;
    const bool upper = lane >= 32;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t stride = (((uint64_t)gridDim.x * blockDim.x) >> 6) * 32;
    for (uint64_t base = wave * 32; base < nq; base += stride) {
        const uint64_t j = base + r;
        const bool valid = j < nq;
        const uint64_t i = valid ? (order ? (uint64_t)order[j] : j) : 0;
        const K a = valid ? lo[i] : K(0), b = valid ? hi[i] : K(0);
        const bool top = b == key_max<K>();
        const K key = upper ? (top ? b : K(b + 1)) : a;
        const uint64_t p = coop_lookup64<K, V, F, LF, false, kOpLowerBound, G, OFF32>(t, lds, key, V(0));
        const uint64_t pa = __shfl(p, (int)r, 64);
        const uint64_t pb1 = __shfl(p, (int)(r + 32), 64);
        const uint64_t pb = top ? t.n : pb1;
        if (valid && !upper) {
            if (first) first[i] = pa;
            count[i] = (a > b || pb < pa) ? 0 : pb - pa;
        }
    }
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint64_t y = __shfl_xor(x, d, 64);
        x = y > x ? y : x;
    }
    return x;
}

constexpr int kRangeIdsUnroll = 8;  // output chunks of 64 ids in flight per wave

// Ids of a group of ranges whose output runs exceed 32-bit relative
// positions: one range at a time, lanes copying consecutive value slots.
template <class K, class V, int LF>
__device__ __forceinline__ void range_ids_wide(const TreeView &t, const uint64_t *__restrict__ first,
                                               const uint64_t *__restrict__ count,
                                               const uint64_t *__restrict__ offsets, uint64_t g0,
                                               uint64_t g1, V *__restrict__ ids) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t i = g0; i < g1; ++i) {
        const uint64_t f = first[i], c = count[i], o = offsets[i];
        for (uint64_t k = lane; k < c; k += 64) ids[o + k] = slot_val<K, V, LF>(t, f + k);
    }
}

// One wave per group of 64 consecutive ranges, range r of the group held by
// lane r. With offsets the exclusive scan of the counts, the group's ids are
// one run of output positions [o0, o0 + span); the wave walks it 64
// positions at a time, U chunks per round so that U loads per lane are in
// flight. A lane finds the range of its position p (the last range starting
// at or before p) between the chunk's first and last range -- two ballots --
// by a binary search over the lanes' starts. Positions outside every
// [offset, offset + count) are skipped, so any non-decreasing offsets are
// accepted; every write stays inside a caller's [offset, offset + count).
// S32: slot numbers fit 32 bits (n <= 2^32).
template <class K, class V, int LF, int U, bool S32>
__global__ __launch_bounds__(256) void k_range_ids(TreeView t, const uint64_t *__restrict__ first,
                                                   const uint64_t *__restrict__ count,
                                                   const uint64_t *__restrict__ offsets, uint64_t nq,
                                                   V *__restrict__ ids) {
    using S = typename std::conditional<S32, uint32_t, uint64_t>::type;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t g0 = wave * 64; g0 < nq; g0 += waves * 64) {
        const uint64_t i = g0 + lane;
        const bool valid = i < nq;
        const uint64_t f = valid ? first[i] : 0, c = valid ? count[i] : 0, o = valid ? offsets[i] : 0;
        const uint64_t o0 = __shfl(o, 0, 64);  // lane 0 always holds a range
        const uint64_t span = wave_max_u64(valid && o >= o0 ? o + c : o0) - o0;
        if (span > 0xFFFFFFFFull - 64 * U) {
            range_ids_wide<K, V, LF>(t, first, count, offsets, g0, g0 + 64 < nq ? g0 + 64 : nq, ids);
            continue;
        }
        // relative start; a lane without a range (or starting before o0) never matches
        const uint32_t rel = valid && o >= o0 ? (uint32_t)(o - o0) : 0xFFFFFFFFu;
        const uint32_t cnt = rel != 0xFFFFFFFFu ? (uint32_t)c : 0u;
        const S sb = (S)f - (S)rel;  // slot of position p in this range: sb + p
        const uint32_t nxt = __shfl(rel, (int)((lane + 1) & 63), 64);
        // dense: every range starts at or after o0 and ends where the next
        // one starts (offsets are the exclusive scan of the counts)
        const bool has_next = lane < 63 && i + 1 < nq;
        const bool dense = __all(!valid || (rel != 0xFFFFFFFFu && (!has_next || nxt == rel + cnt)));
        const uint32_t end = (uint32_t)span;
        for (uint32_t base = 0; base < end; base += 64 * U) {
            V v[U];
            uint32_t pos[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t P = base + 64u * u, p = P + lane;
                const uint32_t ra = (uint32_t)__popcll(__ballot(rel <= P)) - 1;
                const uint32_t rb = (uint32_t)__popcll(__ballot(rel <= P + 63)) - 1;
                uint32_t r = ra;
                if (rb > ra) {
                    for (uint32_t st = 1u << (31 - __clz(rb - ra)); st; st >>= 1) {
                        const uint32_t cand = r + st < rb ? r + st : rb;
                        r = __shfl(rel, (int)cand, 64) <= p ? cand : r;
                    }
                }
                bool in = p < end;
                if (!dense) {
                    const uint32_t rr = __shfl(rel, (int)r, 64), cr = __shfl(cnt, (int)r, 64);
                    in = in && p >= rr && p - rr < cr;
                }
                const S slot = __shfl(sb, (int)r, 64) + (S)p;
                v[u] = in ? slot_val<K, V, LF>(t, (uint64_t)slot) : V(0);
                pos[u] = p;
                ok[u] = in;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (ok[u]) ids[o0 + pos[u]] = v[u];
        }
    }
}

template <class V>
__global__ void k_fill(V *out, uint64_t n, V value) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) out[i] = value;
}

// ------------------------------------------------------------------ scan
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr uint64_t kScanTile = kScanBlock * kScanItems;

template <class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *sh_warp, T *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) sh_warp[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T acc = 0;
        for (int i = 0; i < kScanBlock / 64; ++i) {
            T s = sh_warp[i];
            sh_warp[i] = acc;
            acc += s;
        }
        sh_warp[kScanBlock / 64] = acc;
    }
    __syncthreads();
    T excl = x - v + sh_warp[w];
    *total = sh_warp[kScanBlock / 64];
    __syncthreads();
    return excl;
}

// A thread's kScanItems consecutive items: 16-byte loads when the array is
// 16-byte aligned and all the items exist (the items of a thread start on a
// 32-byte boundary), element loads otherwise.
template <class T>
__device__ __forceinline__ void scan_load(const T *__restrict__ in, uint64_t base, uint64_t n, T (&v)[kScanItems]) {
    if ((reinterpret_cast<uintptr_t>(in) & 15) == 0 && base + kScanItems <= n) {
        const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
        constexpr int NV = kScanItems * (int)sizeof(T) / 16;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const uint4 x = p[i];
            if constexpr (sizeof(T) == 4) {
                v[4 * i] = x.x; v[4 * i + 1] = x.y; v[4 * i + 2] = x.z; v[4 * i + 3] = x.w;
            } else {
                v[2 * i] = (uint64_t)x.x | ((uint64_t)x.y << 32);
                v[2 * i + 1] = (uint64_t)x.z | ((uint64_t)x.w << 32);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) v[j] = (base + j < n) ? in[base + j] : T(0);
    }
}

template <class T>
__device__ __forceinline__ void scan_store(T *__restrict__ out, uint64_t base, uint64_t n, const T (&v)[kScanItems]) {
    if ((reinterpret_cast<uintptr_t>(out) & 15) == 0 && base + kScanItems <= n) {
        uint4 *p = reinterpret_cast<uint4 *>(out + base);
        constexpr int NV = kScanItems * (int)sizeof(T) / 16;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            uint4 x;
            if constexpr (sizeof(T) == 4) {
                x.x = v[4 * i]; x.y = v[4 * i + 1]; x.z = v[4 * i + 2]; x.w = v[4 * i + 3];
            } else {
                x.x = (uint32_t)v[2 * i]; x.y = (uint32_t)(v[2 * i] >> 32);
                x.z = (uint32_t)v[2 * i + 1]; x.w = (uint32_t)(v[2 * i + 1] >> 32);
            }
            p[i] = x;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j)
            if (base + j < n) out[base + j] = v[j];
    }
}

template <class T>
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const T *__restrict__ in, uint64_t n,
                                                            T *__restrict__ block_sums) {
    __shared__ T sh[kScanBlock / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    scan_load<T>(in, base, n, v);
    T s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) s += v[j];
    T total;
    (void)block_exclusive_scan<T>(s, sh, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

template <class T>
__global__ __launch_bounds__(kScanBlock) void k_scan_down(const T *__restrict__ in, uint64_t n,
                                                          const T *__restrict__ block_base,
                                                          T *__restrict__ out) {
    __shared__ T sh[kScanBlock / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    scan_load<T>(in, base, n, v);
    T s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) s += v[j];
    T total;
    T run = block_exclusive_scan<T>(s, sh, &total) + (block_base ? block_base[blockIdx.x] : T(0));
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const T x = v[j];
        v[j] = run;
        run += x;
    }
    scan_store<T>(out, base, n, v);
}

uint64_t scan_tmp_bytes(uint64_t n, uint64_t elem_bytes) {
    uint64_t bytes = 0;
    while (n > kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        bytes += 2 * ((n * elem_bytes + 255) / 256 * 256);
    }
    return bytes + 256;
}

// Exclusive scan, in-place allowed (d_in == d_out). Reduce-then-scan,
// recursing on the per-tile sums.
template <class T>
bpt_status exclusive_scan(const T *d_in, uint64_t n, T *d_out, void *d_tmp, uint64_t tmp_bytes,
                          hipStream_t s) {
    if (n == 0) return BPT_OK;
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        k_scan_down<T><<<1, kScanBlock, 0, s>>>(d_in, n, nullptr, d_out);
        BPT_HIP(hipGetLastError());
        return BPT_OK;
    }
    const uint64_t chunk = (tiles * sizeof(T) + 255) / 256 * 256;
    BPT_CHECK(tmp_bytes >= 2 * chunk, BPT_ERR_INVALID, "scan workspace too small");
    T *sums = reinterpret_cast<T *>(d_tmp);
    T *bases = reinterpret_cast<T *>(reinterpret_cast<uint8_t *>(d_tmp) + chunk);
    BPT_CHECK(tiles <= 0x7fffffffull, BPT_ERR_CAPACITY, "scan too large");
    k_scan_reduce<T><<<(uint32_t)tiles, kScanBlock, 0, s>>>(d_in, n, sums);
    BPT_HIP(hipGetLastError());
    BPT_TRY(exclusive_scan<T>(sums, tiles, bases, reinterpret_cast<uint8_t *>(d_tmp) + 2 * chunk,
                              tmp_bytes - 2 * chunk, s));
    k_scan_down<T><<<(uint32_t)tiles, kScanBlock, 0, s>>>(d_in, n, bases, d_out);
    BPT_HIP(hipGetLastError());
    return BPT_OK;
}

template bpt_status exclusive_scan<uint32_t>(const uint32_t *, uint64_t, uint32_t *, void *, uint64_t, hipStream_t);
template bpt_status exclusive_scan<uint64_t>(const uint64_t *, uint64_t, uint64_t *, void *, uint64_t, hipStream_t);

// ------------------------------------------------------------------ launches
static size_t lds_of(const bpt_tree *t) { return t->view.lds_bytes ? t->view.lds_bytes : 16; }

bpt_status launch_plain(const bpt_tree *t, const void *d_q, uint64_t nq, void *d_out, hipStream_t s) {
    if (nq == 0) return BPT_OK;
    if (t->n == 0) {
        uint32_t g = search_grid(t, nq);
        if (t->val_bytes == 4) k_fill<uint32_t><<<g, 256, 0, s>>>((uint32_t *)d_out, nq, BPT_SENTINEL32);
        else k_fill<uint64_t><<<g, 256, 0, s>>>((uint64_t *)d_out, nq, BPT_SENTINEL64);
        BPT_HIP(hipGetLastError());
        return BPT_OK;
    }
    return dispatch_tree(t, [&](auto kt, auto vt, auto fc, auto lc) -> bpt_status {
        using K = typename decltype(kt)::type;
        using V = typename decltype(vt)::type;
        constexpr int F = decltype(fc)::value, LF = decltype(lc)::value;
        const V sentinel = ~V(0);
        if constexpr (sizeof(K) == sizeof(V)) {
            if (t->coop && t->view.count[0] < (1ull << 32)) {
                const uint32_t g = search_grid(t, nq);
                const size_t l = lds_of(t);
                const bool off32 = t->addr32 && t->leaf_bytes <= (1ull << 32) && t->inner_bytes <= (1ull << 32);
                auto go = [&](auto nt, auto lanes, auto o32) {
                    k_point_coop<K, V, F, LF, decltype(nt)::value, decltype(lanes)::value, decltype(o32)::value>
                        <<<g, t->block, l, s>>>(t->view, (const K *)d_q, (V *)d_out, nq, sentinel);
                };
                auto with_lanes = [&](auto nt) {
                    if (t->coop_lanes == 4) {
                        if (off32) go(nt, IC<4>{}, std::true_type{});
                        else go(nt, IC<4>{}, std::false_type{});
                    } else {
                        if (off32) go(nt, IC<8>{}, std::true_type{});
                        else go(nt, IC<8>{}, std::false_type{});
                    }
                };
                if (t->leaf_nt) with_lanes(std::true_type{});
                else with_lanes(std::false_type{});
                BPT_HIP(hipGetLastError());
                return BPT_OK;
            }
        }
        if (t->coop && t->view.count[0] < (1ull << 32)) {  // key and value widths differ
            with_coop_shape(t, t->addr32 && t->inner_bytes <= (1ull << 32), [&](auto lanes, auto o32) {
                k_point_coop_inner<K, V, F, LF, decltype(lanes)::value, decltype(o32)::value>
                    <<<search_grid(t, nq), t->block, lds_of(t), s>>>(t->view, (const K *)d_q, (V *)d_out, nq,
                                                                     sentinel);
            });
            BPT_HIP(hipGetLastError());
            return BPT_OK;
        }
        k_point_plain<K, V, F, LF><<<search_grid(t, nq), t->block, lds_of(t), s>>>(
            t->view, (const K *)d_q, (V *)d_out, nq, sentinel);
        BPT_HIP(hipGetLastError());
        return BPT_OK;
    });
}

}  // namespace bpt

using namespace bpt;

extern "C" {

uint64_t bpt_point_search_ws_size(const bpt_tree *t, uint64_t nq, int mode) {
    if (!t) return 0;
    if (mode == BPT_POINT_GROUPED) return point_grouped_ws(t, nq);
    if (mode == BPT_POINT_LEARNED) return point_learned_ws(t, nq);
    return 0;
}

bpt_status bpt_point_search_ws(const bpt_tree *t, const void *d_q, uint64_t nq, void *d_out, int mode,
                               void *d_ws, uint64_t ws_bytes, bpt_stream stream) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    BPT_CHECK(nq == 0 || (d_q && d_out), BPT_ERR_INVALID, "queries/out are NULL");
    hipStream_t s = as_stream(stream);
    switch (mode) {
        case BPT_POINT_PLAIN: return launch_plain(t, d_q, nq, d_out, s);
        case BPT_POINT_GROUPED: return point_grouped(t, d_q, nq, d_out, d_ws, ws_bytes, s);
        case BPT_POINT_LEARNED: return point_learned(t, d_q, nq, d_out, d_ws, ws_bytes, s);
    }
    set_error("unknown point mode %d", mode);
    return BPT_ERR_INVALID;
}

bpt_status bpt_point_search(bpt_tree *t, const void *d_q, uint64_t nq, void *d_out, int mode,
                            bpt_stream stream) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    if (mode == BPT_POINT_PLAIN) return bpt_point_search_ws(t, d_q, nq, d_out, mode, nullptr, 0, stream);
    const uint64_t need = bpt_point_search_ws_size(t, nq, mode);
    BPT_TRY(t->ws.reserve(need));
    return bpt_point_search_ws(t, d_q, nq, d_out, mode, t->ws.ptr, t->ws.bytes, stream);
}

bpt_status bpt_lower_bound(const bpt_tree *t, const void *d_q, uint64_t nq, uint64_t *d_pos,
                           bpt_stream stream) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    BPT_CHECK(nq == 0 || (d_q && d_pos), BPT_ERR_INVALID, "queries/pos are NULL");
    if (nq == 0) return BPT_OK;
    hipStream_t s = as_stream(stream);
    if (t->n == 0) {
        k_fill<uint64_t><<<search_grid(t, nq), 256, 0, s>>>(d_pos, nq, 0);
        BPT_HIP(hipGetLastError());
        return BPT_OK;
    }
    return dispatch_tree(t, [&](auto kt, auto vt, auto fc, auto lc) -> bpt_status {
        using K = typename decltype(kt)::type;
        using V = typename decltype(vt)::type;
        constexpr int F = decltype(fc)::value, LF = decltype(lc)::value;
        const uint32_t g = search_grid(t, nq);
        const size_t l = lds_of(t);
        if constexpr (sizeof(K) == sizeof(V)) {
            if (t->coop && t->view.count[0] < (1ull << 32)) {
                const bool off32 = t->addr32 && t->leaf_bytes <= (1ull << 32) && t->inner_bytes <= (1ull << 32);
                auto go = [&](auto lanes, auto o32) {
                    k_lower_bound_coop<K, V, F, LF, decltype(lanes)::value, decltype(o32)::value>
                        <<<g, t->block, l, s>>>(t->view, (const K *)d_q, d_pos, nq);
                };
                if (t->coop_lanes == 4) {
                    if (off32) go(IC<4>{}, std::true_type{});
                    else go(IC<4>{}, std::false_type{});
                } else {
                    if (off32) go(IC<8>{}, std::true_type{});
                    else go(IC<8>{}, std::false_type{});
                }
                BPT_HIP(hipGetLastError());
                return BPT_OK;
            }
        }
        if (t->coop && t->view.count[0] < (1ull << 32)) {  // key and value widths differ
            with_coop_shape(t, t->addr32 && t->inner_bytes <= (1ull << 32), [&](auto lanes, auto o32) {
                k_lower_bound_coop_inner<K, V, F, LF, decltype(lanes)::value, decltype(o32)::value>
                    <<<g, t->block, l, s>>>(t->view, (const K *)d_q, d_pos, nq);
            });
            BPT_HIP(hipGetLastError());
            return BPT_OK;
        }
        k_lower_bound<K, V, F, LF><<<g, t->block, l, s>>>(t->view, (const K *)d_q, d_pos, nq);
        BPT_HIP(hipGetLastError());
        return BPT_OK;
    });
}

}  // extern "C"

namespace bpt {
bpt_status range_count_ordered(const bpt_tree *t, const void *d_lo, const void *d_hi, uint64_t nq,
                               const uint32_t *d_order, uint64_t *d_first, uint64_t *d_count,
                               hipStream_t s) {
    if (nq == 0) return BPT_OK;
    if (t->n == 0) {
        k_fill<uint64_t><<<search_grid(t, nq), 256, 0, s>>>(d_count, nq, 0);
        if (d_first) k_fill<uint64_t><<<search_grid(t, nq), 256, 0, s>>>(d_first, nq, 0);
        BPT_HIP(hipGetLastError());
        return BPT_OK;
    }
    return dispatch_tree(t, [&](auto kt, auto vt, auto fc, auto lc) -> bpt_status {
        using K = typename decltype(kt)::type;
        using V = typename decltype(vt)::type;
        constexpr int F = decltype(fc)::value, LF = decltype(lc)::value;
        const uint32_t g = search_grid(t, nq);
        const size_t l = lds_of(t);
        if constexpr (sizeof(K) == sizeof(V)) {
            if (t->coop && t->view.count[0] < (1ull << 32)) {
                const bool off32 = t->addr32 && t->leaf_bytes <= (1ull << 32) && t->inner_bytes <= (1ull << 32);
                auto go = [&](auto lanes, auto o32) {
                    k_range_count_coop<K, V, F, LF, decltype(lanes)::value, decltype(o32)::value>
                        <<<g, t->block, l, s>>>(t->view, (const K *)d_lo, (const K *)d_hi, d_order, d_first,
                                                d_count, nq);
                };
                if (t->coop_lanes == 4) {
                    if (off32) go(IC<4>{}, std::true_type{});
                    else go(IC<4>{}, std::false_type{});
                } else {
                    if (off32) go(IC<8>{}, std::true_type{});
                    else go(IC<8>{}, std::false_type{});
                }
                BPT_HIP(hipGetLastError());
                return BPT_OK;
            }
        }
        if (t->coop && t->view.count[0] < (1ull << 32)) {  // key and value widths differ
            with_coop_shape(t, t->addr32 && t->inner_bytes <= (1ull << 32), [&](auto lanes, auto o32) {
                k_range_count_coop_inner<K, V, F, LF, decltype(lanes)::value, decltype(o32)::value>
                    <<<g, t->block, l, s>>>(t->view, (const K *)d_lo, (const K *)d_hi, d_order, d_first, d_count,
                                            nq);
            });
            BPT_HIP(hipGetLastError());
            return BPT_OK;
        }
        k_range_count<K, V, F, LF><<<g, t->block, l, s>>>(t->view, (const K *)d_lo, (const K *)d_hi, d_order,
                                                          d_first, d_count, nq);
        BPT_HIP(hipGetLastError());
        return BPT_OK;
    });
}
}  // namespace bpt

extern "C" {

bpt_status bpt_range_count(const bpt_tree *t, const void *d_lo, const void *d_hi, uint64_t nq,
                           uint64_t *d_first, uint64_t *d_count, bpt_stream stream) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    BPT_CHECK(nq == 0 || (d_lo && d_hi && d_count), BPT_ERR_INVALID, "NULL argument");
    return range_count_ordered(t, d_lo, d_hi, nq, nullptr, d_first, d_count, as_stream(stream));
}

bpt_status bpt_range_ids(const bpt_tree *t, const uint64_t *d_first, const uint64_t *d_count,
                         const uint64_t *d_offsets, uint64_t nq, void *d_ids, bpt_stream stream) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    BPT_CHECK(nq == 0 || (d_first && d_count && d_offsets), BPT_ERR_INVALID, "NULL argument");
    if (nq == 0 || t->n == 0) return BPT_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_tree(t, [&](auto kt, auto vt, auto, auto lc) -> bpt_status {
        using K = typename decltype(kt)::type;
        using V = typename decltype(vt)::type;
        constexpr int LF = decltype(lc)::value;
        const uint64_t groups = (nq + 63) / 64;  // one wave per 64 ranges
        const uint32_t blocks = (uint32_t)std::min<uint64_t>((groups + 3) / 4, (uint64_t)device_cu_count() * 8);
        if (t->n <= (1ull << 32))
            k_range_ids<K, V, LF, kRangeIdsUnroll, true>
                <<<blocks, 256, 0, s>>>(t->view, d_first, d_count, d_offsets, nq, (V *)d_ids);
        else
            k_range_ids<K, V, LF, kRangeIdsUnroll, false>
                <<<blocks, 256, 0, s>>>(t->view, d_first, d_count, d_offsets, nq, (V *)d_ids);
        BPT_HIP(hipGetLastError());
        return BPT_OK;
    });
}

bpt_status bpt_exclusive_scan_u64(const uint64_t *d_in, uint64_t n, uint64_t *d_out, uint64_t *h_total,
                                  bpt_stream stream) {
    BPT_CHECK(n == 0 || (d_in && d_out), BPT_ERR_INVALID, "NULL argument");
    hipStream_t s = as_stream(stream);
    if (n == 0) {
        if (h_total) *h_total = 0;
        return BPT_OK;
    }
    uint64_t tmp_bytes = scan_tmp_bytes(n, 8);
    void *tmp = nullptr;
    BPT_HIP(hipMalloc(&tmp, tmp_bytes));
    uint64_t last_in = 0;
    bpt_status st = BPT_OK;
    if (h_total) {
        hipError_t e = hipMemcpyAsync(&last_in, d_in + n - 1, 8, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) st = hip_fail(e, "scan tail copy", __FILE__, __LINE__);
    }
    if (st == BPT_OK) st = exclusive_scan<uint64_t>(d_in, n, d_out, tmp, tmp_bytes, s);
    uint64_t last_out = 0;
    if (st == BPT_OK && h_total) {
        hipError_t e = hipMemcpyAsync(&last_out, d_out + n - 1, 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_fail(e, "scan total", __FILE__, __LINE__);
        *h_total = last_out + last_in;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(tmp);
    return st;
}

}  // extern "C"
