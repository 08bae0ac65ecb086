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
//   * pass 1 counts, the caller scans counts into offsets, pass 2 writes ids,
//     pass 3 sorts each rectangle's ids ascending.
// The 2D-grouping and Z-order modes reorder the rectangles by the keys of
// grouping2d.cu:24-31 / zOrder2d.cu:27-54 applied to (Zmin, Zmax), so that
// neighbouring lanes walk neighbouring parts of the tree; results are
// identical in every mode.
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

// Tropf & Herzog (1981): smallest Morton code > z inside the box spanned by
// zmin and zmax, for z outside the box and zmin <= z <= zmax.
__device__ __forceinline__ uint64_t bigmin(uint64_t z, uint64_t zmin, uint64_t zmax) {
    uint64_t big = 0;
    for (int i = 63; i >= 0; --i) {
        const uint64_t bit = 1ull << i;
        const uint64_t dm = (i & 1) ? 0xAAAAAAAAAAAAAAAAull : 0x5555555555555555ull;
        const uint64_t low = dm & (bit - 1);
        const int v = (z & bit) != 0, lo = (zmin & bit) != 0, hi = (zmax & bit) != 0;
        if (!v && !lo && hi) {
            big = (zmin | bit) & ~low;
            zmax = (zmax & ~bit) | low;
        } else if (!v && lo && hi) {
            return zmin;
        } else if (v && !lo && !hi) {
            return big;
        } else if (v && !lo && hi) {
            zmin = (zmin | bit) & ~low;
        }
    }
    return big;
}

template <int F, int LF>
__device__ __forceinline__ uint64_t lb_u64(const TreeView &t, const uint8_t *lds, uint64_t z) {
    return leaf_lower_bound<uint64_t, uint32_t, LF>(t, descend<uint64_t, F>(t, lds, z), z);
}

// Walks rectangle r; calls emit(slot) for every point inside, in code order.
template <int F, int LF, class Emit>
__device__ __forceinline__ void rect_walk(const TreeView &t, const uint8_t *lds, uint4 r, Emit &&emit) {
    const uint32_t x0 = r.x, y0 = r.y, x1 = r.z, y1 = r.w;
    if (x0 > x1 || y0 > y1) return;
    const uint64_t zmin = spread(x0) | (spread(y0) << 1);
    const uint64_t zmax = spread(x1) | (spread(y1) << 1);
    uint64_t p = lb_u64<F, LF>(t, lds, zmin);
    while (p < t.n) {
        const uint64_t k = slot_key<uint64_t, uint32_t, LF>(t, p);
        if (k > zmax) break;
        const uint32_t x = compact(k), y = compact(k >> 1);
        if (x >= x0 && x <= x1 && y >= y0 && y <= y1) {
            emit(p);
            ++p;
            continue;
        }
        const uint64_t nz = bigmin(k, zmin, zmax);
        if (nz <= k) break;  // no code of the rectangle above k
        // stay inside the current leaf while its keys are below the target
        const uint64_t leaf_end = (p / LF + 1) * LF;
        uint64_t q = p + 1;
        while (q < leaf_end && q < t.n && slot_key<uint64_t, uint32_t, LF>(t, q) < nz) ++q;
        if (q < leaf_end || q >= t.n) {
            p = q;
        } else {
            p = lb_u64<F, LF>(t, lds, nz);
        }
    }
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

// In-place heapsort of a short id run (one thread).
__device__ void sift(uint32_t *a, uint64_t start, uint64_t end) {
    uint64_t root = start;
    while (2 * root + 1 < end) {
        uint64_t child = 2 * root + 1;
        if (child + 1 < end && a[child] < a[child + 1]) ++child;
        if (a[root] < a[child]) {
            uint32_t tmp = a[root]; a[root] = a[child]; a[child] = tmp;
            root = child;
        } else {
            return;
        }
    }
}

__device__ void sort_run(uint32_t *a, uint64_t n) {
    if (n < 2) return;
    if (n <= 16) {
        for (uint64_t i = 1; i < n; ++i) {
            uint32_t v = a[i];
            uint64_t j = i;
            while (j > 0 && a[j - 1] > v) { a[j] = a[j - 1]; --j; }
            a[j] = v;
        }
        return;
    }
    for (uint64_t s = n / 2; s-- > 0;) sift(a, s, n);
    for (uint64_t e = n - 1; e > 0; --e) {
        uint32_t tmp = a[0]; a[0] = a[e]; a[e] = tmp;
        sift(a, 0, e);
    }
}

template <int F, int LF>
__global__ __launch_bounds__(256) void k_rect_ids(TreeView t, const uint4 *__restrict__ rects,
                                                  const uint32_t *__restrict__ order, uint64_t nr,
                                                  const uint64_t *__restrict__ offsets,
                                                  uint32_t *__restrict__ ids) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    stage_top_levels(t, lds);
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nr; j += step) {
        const uint64_t i = order ? order[j] : j;
        uint32_t *out = ids + offsets[i];
        uint64_t c = 0;
        rect_walk<F, LF>(t, lds, rects[i], [&](uint64_t slot) {
            out[c++] = slot_val<uint64_t, uint32_t, LF>(t, slot);
        });
        sort_run(out, c);
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
static bpt_status rect_order(bpt_tree *t, const uint32_t *d_rects, uint64_t nr, int mode, uint32_t **order,
                             hipStream_t s) {
    *order = nullptr;
    if (mode == BPT_GROUP_NONE) return BPT_OK;
    BPT_CHECK(mode == BPT_GROUP_GRID2D || mode == BPT_GROUP_ZORDER, BPT_ERR_INVALID, "group mode %d", mode);
    BPT_CHECK(nr < (1ull << 32), BPT_ERR_CAPACITY, "more than 2^32 rectangles");
    const uint64_t zb = (nr * 8 + 255) / 256 * 256;
    const uint64_t bytes = 2 * zb + query_order_ws_any(nr);
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

static bpt_status rect_run(bpt_tree *t, const uint32_t *d_rects, uint64_t nr, int mode,
                           uint64_t *d_counts, const uint64_t *d_offsets, uint32_t *d_ids, hipStream_t s) {
    BPT_CHECK(t->key_bytes == 8 && t->val_bytes == 4, BPT_ERR_UNSUPPORTED,
              "rectangle search needs a tree of 8-byte Morton keys and 4-byte ids");
    if (nr == 0) return BPT_OK;
    if (t->n == 0) {
        if (d_counts) BPT_HIP(hipMemsetAsync(d_counts, 0, nr * 8, s));
        return BPT_OK;
    }
    uint32_t *order = nullptr;
    bpt_status st = rect_order(t, d_rects, nr, mode, &order, s);
    if (st == BPT_OK) {
        st = dispatch_tree(t, [&](auto, auto, auto fc, auto lc) -> bpt_status {
            constexpr int F = decltype(fc)::value, LF = decltype(lc)::value;
            const size_t lds = t->view.lds_bytes ? t->view.lds_bytes : 16;
            if (d_ids) {
                k_rect_ids<F, LF><<<search_grid(t, nr), t->block, lds, s>>>(
                    t->view, (const uint4 *)d_rects, order, nr, d_offsets, d_ids);
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
