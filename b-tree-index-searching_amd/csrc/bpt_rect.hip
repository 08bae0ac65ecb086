// 2-D rectangle search over points keyed by their 64-bit Morton code.
//
// No reference counterpart: the reference's "2D" forms (BtreeRangeSearch.cu
// :377-783) only reorder 1-D [start, end] queries by a 2-D bucket or a Z-order
// key of the pair, and never return ids (results are freed on the device,
// BtreeRangeSearch.cu:563-564). Here the tree holds key = Morton(x, y) with x
// in the even bits (the bit order of zOrder2d.cu:34-51) and value = point id,
// and each rectangle [x0, x1] x [y0, y1] is answered exactly:
//   * the codes of points inside lie in [Z(x0, y0), Z(x1, y1)];
//   * the scan of that interval skips runs outside the rectangle with
//     Tropf & Herzog's BIGMIN (the next code inside the rectangle), staying in
//     the current leaf when the jump target is near and descending the tree
//     again otherwise;
//   * pass 1 counts, with the first descent (to Zmin) run cooperatively;
//     walks longer than a budget are deferred to a pass that splits each
//     such rectangle over a whole wave;
//   * the caller scans counts into offsets; pass 2 lists the rectangles with
//     hits, walks them again, writes their ids and sorts each run ascending.
// The 2D-grouping and Z-order modes reorder the rectangles of the count pass
// by the keys of grouping2d.cu:24-31 / zOrder2d.cu:27-54 applied to
// (Zmin, Zmax), so that neighbouring lanes walk neighbouring parts of the
// tree; results are identical in every mode.
#include <algorithm>

#include "bpt_internal.hpp"

namespace bpt {

template <class KT>
bpt_status query_order(const KT *lo, const KT *hi, uint64_t n, int mode, void *ws, uint64_t ws_bytes,
                       uint32_t **order, hipStream_t s);
uint64_t query_order_ws_any(uint64_t n);

__device__ __forceinline__ uint64_t spread(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

__device__ __forceinline__ uint32_t compact(uint64_t x) {
    x &= 0x5555555555555555ull;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
    x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
    return (uint32_t)x;
}

// BIGMIN (Tropf & Herzog, 1981): the smallest Morton code above z inside the
// box [x0, x1] x [y0, y1], for z = Morton(vx, vy) outside the box; 0 (<= z)
// when there is none. Their algorithm walks the 64 code bits one at a time;
// this closed form gets the same code from a few bit scans:
//   * j, the highest code bit where z's prefix leaves the box: x bit b sits at
//     2b and y bit b at 2b + 1, and a coordinate below (above) its range
//     leaves at the top bit where it differs from its lower (upper) bound;
//   * every code above z first differs from z at a bit i where it has 1 and z
//     has 0; with i >= j the prefix above i stays inside the box, and setting
//     bit i keeps it there iff that coordinate is below its upper bound and
//     the bit is no higher than the top bit where the two differ;
//   * the lowest such i gives the smallest code: keep the prefix above i, set
//     bit i, and complete each coordinate with the least value in its range,
//     max(lower bound, prefix with zeros below).
__device__ __forceinline__ uint64_t bigmin(uint32_t vx, uint32_t vy, uint32_t x0, uint32_t y0, uint32_t x1,
                                           uint32_t y1) {
    int j = 0;
    if (vx < x0) j = 2 * (31 - __clz(vx ^ x0));
    else if (vx > x1) j = 2 * (31 - __clz(vx ^ x1));
    if (vy < y0) j = max(j, 2 * (31 - __clz(vy ^ y0)) + 1);
    else if (vy > y1) j = max(j, 2 * (31 - __clz(vy ^ y1)) + 1);
    const uint32_t cx = vx < x1 ? ~vx & (uint32_t)((2ull << (31 - __clz(vx ^ x1))) - 1) : 0u;
    const uint32_t cy = vy < y1 ? ~vy & (uint32_t)((2ull << (31 - __clz(vy ^ y1))) - 1) : 0u;
    const uint64_t c = (spread(cx) | (spread(cy) << 1)) & (~0ull << j);
    if (!c) return 0;
    const int i = __ffsll((unsigned long long)c) - 1;
    const int kx = (i + 1) >> 1, ky = i >> 1;  // bits of x and of y below code bit i
    uint64_t px = vx, py = vy;
    if (i & 1) py |= 1ull << ky;
    else px |= 1ull << kx;
    px = (px >> kx) << kx;
    py = (py >> ky) << ky;
    const uint32_t mx = px > x0 ? (uint32_t)px : x0, my = py > y0 ? (uint32_t)py : y0;
    return spread(mx) | (spread(my) << 1);
}

template <int F, int LF>
__device__ __forceinline__ uint64_t lb_u64(const TreeView &t, const uint8_t *lds, uint64_t z) {
    return leaf_lower_bound<uint64_t, uint32_t, LF>(t, descend<uint64_t, F>(t, lds, z), z);
}

// lower_bound(z) when every key of `leaf` is below z: the leaf's level-1
// node first (its children's maxima; padding entries are the largest key),
// which holds the answer unless all its children end below z, then a
// descent from the top.
template <int F, int LF>
__device__ __forceinline__ uint64_t lb_after_leaf(const TreeView &t, const uint8_t *lds, uint64_t leaf, uint64_t z) {
    if (t.height >= 1) {
        const uint64_t node = leaf / F;
        const uint64_t *sep = reinterpret_cast<const uint64_t *>(t.inner + t.off[1]) + node * F;
        const uint64_t child = node * F + node_rank<uint64_t, F>(sep, z);
        if (child < (node + 1) * F && child < t.count[0])
            return leaf_lower_bound<uint64_t, uint32_t, LF>(t, child, z);
    }
    return lb_u64<F, LF>(t, lds, z);
}

// Walks rectangle r from p = lower_bound(zmin); calls emit(slot) for every
// point inside, in code order. Gives up (returns false) before its
// budget+1-th key read; the points emitted so far are then a prefix.
template <int F, int LF, class Emit>
__device__ __forceinline__ bool rect_walk_budget(const TreeView &t, const uint8_t *lds, uint4 r, uint64_t p,
                                                 uint32_t budget, Emit &&emit) {
    const uint32_t x0 = r.x, y0 = r.y, x1 = r.z, y1 = r.w;
    if (x0 > x1 || y0 > y1) return true;
    const uint64_t zmax = spread(x1) | (spread(y1) << 1);
    uint32_t reads = 0;
    while (p < t.n) {
        if (reads++ == budget) return false;
        const uint64_t k = slot_key<uint64_t, uint32_t, LF>(t, p);
        if (k > zmax) break;
        const uint32_t x = compact(k), y = compact(k >> 1);
        if (x >= x0 && x <= x1 && y >= y0 && y <= y1) {
            emit(p);
            ++p;
            continue;
        }
        const uint64_t nz = bigmin(x, y, x0, y0, x1, y1);
        if (nz <= k) break;  // no code of the rectangle above k
        // stay inside the current leaf while its keys are below the target
        const uint64_t leaf_end = (p / LF + 1) * LF;
        uint64_t q = p + 1;
        while (q < leaf_end && q < t.n && slot_key<uint64_t, uint32_t, LF>(t, q) < nz) ++q;
        if (q < leaf_end || q >= t.n) {
            p = q;
        } else {
            p = lb_after_leaf<F, LF>(t, lds, p / LF, nz);
        }
    }
    return true;
}

template <int F, int LF, class Emit>
__device__ __forceinline__ void rect_walk_from(const TreeView &t, const uint8_t *lds, uint4 r, uint64_t p,
                                               Emit &&emit) {
    (void)rect_walk_budget<F, LF>(t, lds, r, p, ~0u, emit);
}

template <int F, int LF, class Emit>
__device__ __forceinline__ void rect_walk(const TreeView &t, const uint8_t *lds, uint4 r, Emit &&emit) {
    if (r.x > r.z || r.y > r.w) return;
    const uint64_t zmin = spread(r.x) | (spread(r.y) << 1);
    rect_walk_from<F, LF>(t, lds, r, lb_u64<F, LF>(t, lds, zmin), emit);
}

template <int F, int LF>
__global__ __launch_bounds__(256) void k_rect_count(TreeView t, const uint4 *__restrict__ rects,
                                                    const uint32_t *__restrict__ order, uint64_t nr,
                                                    uint64_t *__restrict__ counts) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nr; j += step) {
        const uint64_t i = order ? order[j] : j;
        uint64_t c = 0;
        rect_walk<F, LF>(t, lds, rects[i], [&](uint64_t) { ++c; });
        counts[i] = c;
    }
}

// The same counts with the first descent of each rectangle -- to
// lower_bound(Zmin), nearly all of a rectangle's tree reads when few points
// fall inside -- run cooperatively: each wave takes 64 rectangles at a time,
// G lanes per inner node, then one lane per rectangle for its leaf and walk.
// A walk that needs more than kRectBudget key reads would hold its wave up:
// its rectangle is queued instead (per block in LDS, then appended to the
// global list `deferred` with one atomic per block) and counted afterwards
// by a whole wave (k_rect_count_deferred). When a block's queue is full the
// lane finishes the walk itself.
constexpr uint32_t kRectBudget = 8;      // key reads of a walk before it is deferred
constexpr uint32_t kRectQueue = 1024;    // deferred rectangles queued per block
template <int F, int LF, int G, bool OFF32>
__global__ __launch_bounds__(256) void k_rect_count_coop(TreeView t, const uint4 *__restrict__ rects,
                                                         const uint32_t *__restrict__ order, uint64_t nr,
                                                         uint64_t *__restrict__ counts,
                                                         uint32_t *__restrict__ deferred,
                                                         uint64_t *__restrict__ ndeferred) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    __shared__ uint32_t dq[kRectQueue];
    __shared__ uint32_t dn;
    __shared__ uint64_t dbase;
    if (threadIdx.x == 0) dn = 0;
    stage_top_levels(t, lds);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t stride = ((uint64_t)gridDim.x * blockDim.x) & ~63ull;
    for (uint64_t base = wave * 64; base < nr; base += stride) {
        const uint64_t j = base + lane;
        const bool valid = j < nr;
        const uint64_t i = valid ? (order ? (uint64_t)order[j] : j) : 0;
        const uint4 r = valid ? rects[i] : make_uint4(1, 1, 0, 0);  // empty: x0 > x1
        const uint64_t zmin = spread(r.x) | (spread(r.y) << 1);
        const uint32_t leaf = coop_leaf_of<uint64_t, F, G, OFF32>(t, lds, zmin);
        uint64_t c = 0;
        if (valid && r.x <= r.z && r.y <= r.w) {
            const uint64_t p = leaf_lower_bound<uint64_t, uint32_t, LF>(t, leaf, zmin);
            if (!rect_walk_budget<F, LF>(t, lds, r, p, kRectBudget, [&](uint64_t) { ++c; })) {
                const uint32_t at = atomicAdd(&dn, 1u);
                if (at < kRectQueue) {
                    dq[at] = (uint32_t)i;
                    continue;  // counted by k_rect_count_deferred
                }
                c = 0;
                rect_walk_from<F, LF>(t, lds, r, p, [&](uint64_t) { ++c; });
            }
        }
        if (valid) counts[i] = c;
    }
    __syncthreads();
    const uint32_t n = dn < kRectQueue ? dn : kRectQueue;
    if (n == 0) return;  // the same in every thread of the block
    if (threadIdx.x == 0) dbase = atomicAdd(reinterpret_cast<unsigned long long *>(ndeferred), (unsigned long long)n);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) deferred[dbase + k] = dq[k];
}

// Counts of the deferred rectangles: each wave takes one at a time, lane l
// walks sub-rectangle (l % 8, l / 8) of an 8 x 8 split, and the lanes' counts
// are summed (the sub-rectangles cover the rectangle once). (Cooperative
// descents to the sub-rectangles' first codes measured no faster: the pass
// waits on its longest walks.)
template <int F, int LF>
__global__ __launch_bounds__(256) void k_rect_count_deferred(TreeView t, const uint4 *__restrict__ rects,
                                                             const uint32_t *__restrict__ deferred,
                                                             const uint64_t *__restrict__ ndeferred,
                                                             uint64_t *__restrict__ counts) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint32_t lane = threadIdx.x & 63, kx = lane & 7, ky = lane >> 3;
    const uint64_t nd = *ndeferred;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t k = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < nd; k += nwaves) {
        const uint32_t i = deferred[k];
        const uint4 r = rects[i];
        const uint64_t w = (uint64_t)r.z - r.x + 1, h = (uint64_t)r.w - r.y + 1;
        const uint64_t xa = r.x + w * kx / 8, xe = r.x + w * (kx + 1) / 8;
        const uint64_t ya = r.y + h * ky / 8, ye = r.y + h * (ky + 1) / 8;
        uint64_t c = 0;
        if (xa < xe && ya < ye) {
            const uint4 sub = make_uint4((uint32_t)xa, (uint32_t)ya, (uint32_t)(xe - 1), (uint32_t)(ye - 1));
            rect_walk<F, LF>(t, lds, sub, [&](uint64_t) { ++c; });
        }
        for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
        if (lane == 0) counts[i] = c;
    }
}

__global__ void k_rect_zero(uint64_t *p) { *p = 0; }

// counts of an empty tree (a kernel, not a memset: callers may capture the
// search into a graph)
__global__ void k_rect_zero_n(uint64_t *__restrict__ p, uint64_t n) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) p[i] = 0;
}

constexpr uint32_t kRectSortSmall = 8;    // runs sorted by their own lane, in registers
constexpr uint64_t kRectSplit = 4;        // longer runs are walked by a whole wave
constexpr uint32_t kRectSortCap = 1024;   // runs sorted by the whole wave in LDS

// Odd-even transposition network over kRectSortSmall registers.
__device__ __forceinline__ void sort_small(uint32_t *p, uint32_t c) {
    uint32_t v[kRectSortSmall];
#pragma unroll
    for (uint32_t k = 0; k < kRectSortSmall; ++k) v[k] = k < c ? p[k] : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t r = 0; r < kRectSortSmall; ++r)
#pragma unroll
        for (uint32_t k = r & 1; k + 1 < kRectSortSmall; k += 2) {
            const uint32_t a = v[k] < v[k + 1] ? v[k] : v[k + 1];
            const uint32_t b = v[k] < v[k + 1] ? v[k + 1] : v[k];
            v[k] = a;
            v[k + 1] = b;
        }
#pragma unroll
    for (uint32_t k = 0; k < kRectSortSmall; ++k)
        if (k < c) p[k] = v[k];
}

// The whole wave sorts p[0, c), c <= kRectSortCap: a bitonic sort of the run
// padded to a power of two in the wave's LDS buffer.
__device__ __forceinline__ void wave_sort_lds(uint32_t *p, uint32_t c, uint32_t *buf) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t m = 64;
    while (m < c) m <<= 1;
    for (uint32_t k = lane; k < m; k += 64) buf[k] = k < c ? p[k] : 0xFFFFFFFFu;
    for (uint32_t size = 2; size <= m; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            for (uint32_t q = lane; q < m / 2; q += 64) {
                const uint32_t a = 2 * q - (q & (stride - 1)), b = a + stride;
                const uint32_t x = buf[a], y = buf[b];
                if ((x > y) == ((a & size) == 0)) {
                    buf[a] = y;
                    buf[b] = x;
                }
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    for (uint32_t k = lane; k < c; k += 64) p[k] = buf[k];
}

// The whole wave sorts p[0, c) of any length in place, in global memory: a
// bitonic network in its all-ascending form (each merge starts by comparing
// position i with its mirror in the block, i ^ (size - 1)), so that the
// positions past c act as +infinity and are never read or written.
__device__ __forceinline__ void wave_sort_global(uint32_t *p, uint64_t c) {
    const uint32_t lane = threadIdx.x & 63;
    uint64_t m = 1;
    while (m < c) m <<= 1;
    for (uint64_t size = 2; size <= m; size <<= 1) {
        for (uint64_t stride = size >> 1; stride > 0; stride >>= 1) {
            const bool mirror = stride == (size >> 1);
            for (uint64_t q = lane; q < m / 2; q += 64) {
                const uint64_t a = 2 * q - (q & (stride - 1));
                const uint64_t b = mirror ? a ^ (size - 1) : a + stride;
                if (b < c) {
                    const uint32_t x = p[a], y = p[b];
                    if (x > y) {
                        p[a] = y;
                        p[b] = x;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
    }
}

// The hit count of one rectangle by one wave: lane l walks sub-rectangle
// (l % 8, l / 8) of an 8 x 8 split (the split of the ids pass), and the
// lanes' counts are summed. For the last rectangle of an ids pass, which has
// no next offset to give its count; it also clears the pass's two list
// lengths.
template <int F, int LF>
__global__ __launch_bounds__(64) void k_rect_count_split(TreeView t, const uint4 *__restrict__ rect,
                                                         uint64_t *__restrict__ count,
                                                         uint64_t *__restrict__ nlists) {
    if (threadIdx.x < 2) nlists[threadIdx.x] = 0;
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint32_t lane = threadIdx.x & 63, kx = lane & 7, ky = lane >> 3;
    const uint4 r = *rect;
    const uint64_t w = (uint64_t)r.z - r.x + 1, h = (uint64_t)r.w - r.y + 1;
    const uint64_t xa = r.x + w * kx / 8, xe = r.x + w * (kx + 1) / 8;
    const uint64_t ya = r.y + h * ky / 8, ye = r.y + h * (ky + 1) / 8;
    uint64_t c = 0;
    if (r.x <= r.z && r.y <= r.w && xa < xe && ya < ye) {
        const uint4 sub = make_uint4((uint32_t)xa, (uint32_t)ya, (uint32_t)(xe - 1), (uint32_t)(ye - 1));
        rect_walk<F, LF>(t, lds, sub, [&](uint64_t) { ++c; });
    }
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane == 0) *count = c;
}

// The rectangles with hits, listed for the ids pass in two lists: narrow
// (1 to kRectSplit ids) and wide (more). A rectangle's count is the gap to
// the next offset (offsets are the exclusive scan of the counts); the last
// rectangle's, which has no next offset, is *last_count (counted just
// before). Each block scans a contiguous range, collects its entries in LDS
// and appends them with one atomic per list and flush; the lists' order does
// not matter (every rectangle writes its own run).
constexpr uint32_t kHitsPerThread = 16;
constexpr uint32_t kHitsRound = 256 * kHitsPerThread;  // rectangles per block and round
constexpr uint32_t kHitsCap = 2 * kHitsRound;          // LDS entries per list
__global__ __launch_bounds__(256) void k_rect_hits(const uint64_t *__restrict__ offsets, uint64_t nr,
                                                   const uint64_t *__restrict__ last_count,
                                                   uint32_t *__restrict__ narrow, uint32_t *__restrict__ wide,
                                                   uint64_t *__restrict__ nlists) {
    __shared__ uint32_t ln[kHitsCap], lw[kHitsCap];
    __shared__ uint32_t cn, cw;
    __shared__ uint64_t bn, bw;
    const uint64_t per = (nr + gridDim.x - 1) / gridDim.x;
    const uint64_t lo = std::min<uint64_t>(nr, (uint64_t)blockIdx.x * per), hi = std::min<uint64_t>(nr, lo + per);
    const uint64_t lastc = *last_count;
    if (threadIdx.x == 0) cn = cw = 0;
    __syncthreads();
    auto flush = [&]() {  // every thread of the block
        if (threadIdx.x == 0) {
            bn = cn ? atomicAdd(reinterpret_cast<unsigned long long *>(nlists), (unsigned long long)cn) : 0;
            bw = cw ? atomicAdd(reinterpret_cast<unsigned long long *>(nlists + 1), (unsigned long long)cw) : 0;
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < cn; k += blockDim.x) narrow[bn + k] = ln[k];
        for (uint32_t k = threadIdx.x; k < cw; k += blockDim.x) wide[bw + k] = lw[k];
        __syncthreads();
        if (threadIdx.x == 0) cn = cw = 0;
        __syncthreads();
    };
    for (uint64_t base = lo; base < hi; base += kHitsRound) {
        uint64_t a[kHitsPerThread], b[kHitsPerThread];
#pragma unroll
        for (uint32_t e = 0; e < kHitsPerThread; ++e) {
            const uint64_t i = base + e * blockDim.x + threadIdx.x;
            a[e] = i < hi ? offsets[i] : 0;
            b[e] = i < hi && i + 1 < nr ? offsets[i + 1] : 0;
        }
#pragma unroll
        for (uint32_t e = 0; e < kHitsPerThread; ++e) {
            const uint64_t i = base + e * blockDim.x + threadIdx.x;
            if (i < hi) {
                const uint64_t c = i + 1 < nr ? b[e] - a[e] : lastc;
                if (c > kRectSplit) lw[atomicAdd(&cw, 1u)] = (uint32_t)i;
                else if (c > 0) ln[atomicAdd(&cn, 1u)] = (uint32_t)i;
            }
        }
        __syncthreads();
        const bool full = cn > kHitsCap - kHitsRound || cw > kHitsCap - kHitsRound;  // room for the next round?
        __syncthreads();  // every thread has read cn, cw before the next round adds to them
        if (full) flush();
    }
    __syncthreads();
    if (cn || cw) flush();
}

// Ids of the listed rectangles; rectangles without points cost nothing here.
// Wide rectangles go one per wave at a time to the whole wave: lane l takes
// sub-rectangle (l % 8, l / 8) of an 8 x 8 split of it (empty where a side is
// shorter than 8), and the lanes write through a counter in LDS; the
// sub-rectangles cover the rectangle once, so the run holds the same ids.
// Narrow rectangles go 64 to a wave (from the last wave down, so that the
// waves with wide ones take fewer): their lanes walk and write the ids in
// code order, and a walk that needs more than kRectBudget key reads goes to
// the whole wave as a wide one (its ids so far are written again). Runs of up
// to kRectSortSmall ids are then sorted by one lane in registers, longer runs
// by the whole wave: in LDS up to kRectSortCap ids, in place in global memory
// beyond.
template <int F, int LF>
__global__ __launch_bounds__(256) void k_rect_ids(TreeView t, const uint4 *__restrict__ rects,
                                                  const uint32_t *__restrict__ narrow,
                                                  const uint32_t *__restrict__ wide,
                                                  const uint64_t *__restrict__ nlists, uint64_t nr,
                                                  const uint64_t *__restrict__ offsets,
                                                  const uint64_t *__restrict__ last_count,
                                                  uint32_t *__restrict__ ids) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    uint32_t *buf = reinterpret_cast<uint32_t *>(lds + ((t.lds_bytes + 15) & ~15ull)) +
                    (threadIdx.x >> 6) * kRectSortCap;
    const uint32_t lane = threadIdx.x & 63, kx = lane & 7, ky = lane >> 3;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t nn = nlists[0], nw = nlists[1];
    auto count_of = [&](uint64_t i) -> uint64_t { return i + 1 < nr ? offsets[i + 1] - offsets[i] : *last_count; };
    // the whole wave writes the sc ids of rectangle i at ids[so ...]
    auto walk_wide = [&](uint64_t i, uint64_t so, uint64_t sc) {
        uint32_t *ctr = buf;  // the sorts use buf only after the walk
        const uint4 r = rects[i];
        if (lane == 0) *ctr = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        const uint64_t w = (uint64_t)r.z - r.x + 1, h = (uint64_t)r.w - r.y + 1;
        const uint64_t xa = r.x + w * kx / 8, xe = r.x + w * (kx + 1) / 8;
        const uint64_t ya = r.y + h * ky / 8, ye = r.y + h * (ky + 1) / 8;
        if (xa < xe && ya < ye) {
            const uint4 sub = make_uint4((uint32_t)xa, (uint32_t)ya, (uint32_t)(xe - 1), (uint32_t)(ye - 1));
            rect_walk<F, LF>(t, lds, sub, [&](uint64_t slot) {
                const uint32_t at = atomicAdd(ctr, 1u);
                if (at < sc) ids[so + at] = slot_val<uint64_t, uint32_t, LF>(t, slot);
            });
        }
        // the lanes' writes are visible to the whole wave (one CU, one L1)
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    };
    for (uint64_t k = wave; k < nw; k += nwaves) {
        const uint64_t i = wide[k];
        const uint64_t so = offsets[i], sc = count_of(i);
        walk_wide(i, so, sc);
        if (sc <= kRectSortSmall) {
            if (lane == 0) sort_small(ids + so, (uint32_t)sc);
        } else if (sc <= kRectSortCap) {
            wave_sort_lds(ids + so, (uint32_t)sc, buf);
        } else {
            wave_sort_global(ids + so, sc);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    for (uint64_t b = nwaves - 1 - wave; b * 64 < nn; b += nwaves) {
        const uint64_t j = b * 64 + lane;
        uint64_t o = 0, c = 0, i = 0, known = 0;
        bool spread_me = false;
        if (j < nn) {
            i = narrow[j];
            o = offsets[i];
            known = count_of(i);
            uint32_t *out = ids + o;
            const uint4 r = rects[i];
            const uint64_t zmin = spread(r.x) | (spread(r.y) << 1);
            if (!rect_walk_budget<F, LF>(t, lds, r, lb_u64<F, LF>(t, lds, zmin), kRectBudget,
                                         [&](uint64_t slot) { out[c++] = slot_val<uint64_t, uint32_t, LF>(t, slot); }))
                spread_me = true;
        }
        for (uint64_t sp = __ballot(spread_me); sp; sp &= sp - 1) {
            const int src = __ffsll((unsigned long long)sp) - 1;
            walk_wide(__shfl(i, src, 64), __shfl(o, src, 64), __shfl(known, src, 64));
        }
        if (spread_me) c = known;
        if (c > 1) sort_small(ids + o, (uint32_t)c);  // c <= kRectSplit <= kRectSortSmall
    }
}

__global__ void k_rect_zbounds(const uint4 *__restrict__ rects, uint64_t nr, uint64_t *__restrict__ zlo,
                               uint64_t *__restrict__ zhi) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += step) {
        const uint4 r = rects[i];
        zlo[i] = spread(r.x) | (spread(r.y) << 1);
        zhi[i] = spread(r.z) | (spread(r.w) << 1);
    }
}

// Grouped query order for rectangles: keys from (Zmin, Zmax). The scratch
// space is the tree's own workspace (kept between calls, grown on demand).
static uint64_t rect_order_bytes(uint64_t nr, int mode) {
    if (mode == BPT_GROUP_NONE) return 0;
    const uint64_t zb = (nr * 8 + 255) / 256 * 256;
    return (2 * zb + query_order_ws_any(nr) + 255) / 256 * 256;
}

static bpt_status rect_order(bpt_tree *t, const uint32_t *d_rects, uint64_t nr, int mode, uint32_t **order,
                             hipStream_t s) {
    *order = nullptr;
    if (mode == BPT_GROUP_NONE) return BPT_OK;
    BPT_CHECK(mode == BPT_GROUP_GRID2D || mode == BPT_GROUP_ZORDER, BPT_ERR_INVALID, "group mode %d", mode);
    BPT_CHECK(nr < (1ull << 32), BPT_ERR_CAPACITY, "more than 2^32 rectangles");
    const uint64_t zb = (nr * 8 + 255) / 256 * 256;
    const uint64_t bytes = rect_order_bytes(nr, mode);
    BPT_TRY(t->ws.reserve(bytes));
    void *ws = t->ws.ptr;
    uint64_t *zlo = reinterpret_cast<uint64_t *>(ws);
    uint64_t *zhi = reinterpret_cast<uint64_t *>(reinterpret_cast<uint8_t *>(ws) + zb);
    uint32_t g = (uint32_t)std::min<uint64_t>((nr + 255) / 256, (uint64_t)device_cu_count() * 16);
    k_rect_zbounds<<<std::max<uint32_t>(g, 1), 256, 0, s>>>((const uint4 *)d_rects, nr, zlo, zhi);
    BPT_HIP(hipGetLastError());
    return query_order<uint64_t>(zlo, zhi, nr, mode, reinterpret_cast<uint8_t *>(ws) + 2 * zb,
                                 bytes - 2 * zb, order, s);
}

// The search grid cut to the blocks the GPU holds at once. Each block stages
// the top levels into LDS; the rectangle kernels hold few blocks per CU, so a
// larger grid only stages them again.
template <class Kernel>
static bpt_status resident_grid(const bpt_tree *t, Kernel kernel, uint64_t nr, size_t lds, uint32_t *grid) {
    int per_cu = 0;
    BPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(kernel),
                                                         (int)t->block, lds));
    *grid = search_grid(t, nr);
    if (per_cu > 0) *grid = (uint32_t)std::min<uint64_t>(*grid, (uint64_t)device_cu_count() * per_cu);
    return BPT_OK;
}

static bpt_status rect_run(bpt_tree *t, const uint32_t *d_rects, uint64_t nr, int mode,
                           uint64_t *d_counts, const uint64_t *d_offsets, uint32_t *d_ids, hipStream_t s) {
    BPT_CHECK(t->key_bytes == 8 && t->val_bytes == 4, BPT_ERR_UNSUPPORTED,
              "rectangle search needs a tree of 8-byte Morton keys and 4-byte ids");
    if (nr == 0) return BPT_OK;
    if (t->n == 0) {
        if (d_counts) {
            k_rect_zero_n<<<(uint32_t)std::min<uint64_t>((nr + 255) / 256, 4096), 256, 0, s>>>(d_counts, nr);
            BPT_HIP(hipGetLastError());
        }
        return BPT_OK;
    }
    // the count pass walks in the requested order; the ids pass walks only the
    // rectangles with hits, in the caller's order (results are the same)
    // the count pass's scratch: the query order, then the deferred-walk list
    // (reserved together, so that the order's pointers stay valid)
    const bool coop = !d_ids && t->coop && t->view.count[0] < (1ull << 32);
    BPT_CHECK(nr < (1ull << 32), BPT_ERR_CAPACITY, "more than 2^32 rectangles");
    const uint64_t obytes = d_ids ? 0 : rect_order_bytes(nr, mode);
    if (coop) BPT_TRY(t->ws.reserve(obytes + 256 + nr * 4));
    uint32_t *order = nullptr;
    bpt_status st = d_ids ? BPT_OK : rect_order(t, d_rects, nr, mode, &order, s);
    if (st == BPT_OK) {
        st = dispatch_tree(t, [&](auto, auto, auto fc, auto lc) -> bpt_status {
            constexpr int F = decltype(fc)::value, LF = decltype(lc)::value;
            const size_t lds = t->view.lds_bytes ? t->view.lds_bytes : 16;
            if (d_ids) {
                // the last rectangle's count (it has no next offset), so that a
                // large one is walked by a whole wave like the others
                // the last rectangle's count and the two list lengths, then the
                // narrow and wide lists (nr entries of room each)
                const uint64_t lb = (nr * 4 + 255) / 256 * 256;
                BPT_TRY(t->ws.reserve(256 + 2 * lb));
                uint64_t *last = static_cast<uint64_t *>(t->ws.ptr);
                uint64_t *nlists = last + 1;
                uint32_t *narrow = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(t->ws.ptr) + 256);
                uint32_t *wide = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(t->ws.ptr) + 256 + lb);
                k_rect_count_split<F, LF><<<1, 64, lds, s>>>(t->view, (const uint4 *)d_rects + (nr - 1), last,
                                                             nlists);
                BPT_HIP(hipGetLastError());
                const uint32_t hg = (uint32_t)std::min<uint64_t>((nr + kHitsRound - 1) / kHitsRound, 512);
                k_rect_hits<<<hg, 256, 0, s>>>(d_offsets, nr, last, narrow, wide, nlists);
                BPT_HIP(hipGetLastError());
                const size_t ids_lds = ((t->view.lds_bytes + 15) & ~(size_t)15) +
                                       (size_t)(t->block / 64) * kRectSortCap * sizeof(uint32_t);
                uint32_t grid = 0;
                BPT_TRY(resident_grid(t, k_rect_ids<F, LF>, nr, ids_lds, &grid));
                k_rect_ids<F, LF><<<grid, t->block, ids_lds, s>>>(t->view, (const uint4 *)d_rects, narrow, wide,
                                                                nlists, nr, d_offsets, last, d_ids);
            } else if (coop) {
                const bool off32 = t->addr32 && t->inner_bytes <= (1ull << 32);
                uint8_t *dw = static_cast<uint8_t *>(t->ws.ptr) + obytes;
                uint64_t *ndef = reinterpret_cast<uint64_t *>(dw);
                uint32_t *def = reinterpret_cast<uint32_t *>(dw + 256);
                k_rect_zero<<<1, 1, 0, s>>>(ndef);
                BPT_HIP(hipGetLastError());
                auto go = [&](auto lanes, auto o32) -> bpt_status {
                    constexpr auto kernel = k_rect_count_coop<F, LF, decltype(lanes)::value, decltype(o32)::value>;
                    uint32_t grid = 0;
                    BPT_TRY(resident_grid(t, kernel, nr, lds, &grid));
                    kernel<<<grid, t->block, lds, s>>>(t->view, (const uint4 *)d_rects, order, nr, d_counts, def,
                                                       ndef);
                    return BPT_OK;
                };
                if (t->coop_lanes == 4) {
                    BPT_TRY(off32 ? go(IC<4>{}, std::true_type{}) : go(IC<4>{}, std::false_type{}));
                } else {
                    BPT_TRY(off32 ? go(IC<8>{}, std::true_type{}) : go(IC<8>{}, std::false_type{}));
                }
                BPT_HIP(hipGetLastError());
                uint32_t grid = 0;
                BPT_TRY(resident_grid(t, k_rect_count_deferred<F, LF>, nr, lds, &grid));
                k_rect_count_deferred<F, LF><<<grid, t->block, lds, s>>>(t->view, (const uint4 *)d_rects, def, ndef,
                                                                        d_counts);
            } else {
                k_rect_count<F, LF><<<search_grid(t, nr), t->block, lds, s>>>(
                    t->view, (const uint4 *)d_rects, order, nr, d_counts);
            }
            BPT_HIP(hipGetLastError());
            return BPT_OK;
        });
    }
    return st;
}

}  // namespace bpt

using namespace bpt;

extern "C" {

bpt_status bpt_rect_count(bpt_tree *t, const uint32_t *d_rects, uint64_t nr, int group_mode,
                          uint64_t *d_counts, bpt_stream s) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    BPT_CHECK(nr == 0 || (d_rects && d_counts), BPT_ERR_INVALID, "NULL argument");
    return rect_run(t, d_rects, nr, group_mode, d_counts, nullptr, nullptr, as_stream(s));
}

bpt_status bpt_rect_ids(bpt_tree *t, const uint32_t *d_rects, uint64_t nr, int group_mode,
                        const uint64_t *d_offsets, uint32_t *d_ids, bpt_stream s) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    BPT_CHECK(nr == 0 || (d_rects && d_offsets && d_ids), BPT_ERR_INVALID, "NULL argument");
    return rect_run(t, d_rects, nr, group_mode, nullptr, d_offsets, d_ids, as_stream(s));
}

}  // extern "C"
