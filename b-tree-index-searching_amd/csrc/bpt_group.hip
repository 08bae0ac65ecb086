// Query grouping on the device: radix partition of point queries on their top
// key bits, stable LSD radix sort (the grouping front ends of range search),
// Morton / Z-order / 2-D grouping keys.
//
// The reference groups queries with a host-driven MSD loop (genGrouping,
// grouping.cu:1677-1851): 8-bit passes, a device histogram copied to the host
// after every pass, a 256-entry prefix sum on one block, relocation, and a
// host loop merging adjacent buckets below 33000 queries. Results come back
// in grouped order and overwrite the caller's queries. Here grouping is one
// on-device partition with no host round trip, and results are returned in
// the caller's order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "bpt_internal.hpp"

namespace bpt {

bpt_status range_count_ordered(const bpt_tree *t, const void *d_lo, const void *d_hi, uint64_t nq,
                               const uint32_t *d_order, uint64_t *d_first, uint64_t *d_count,
                               hipStream_t s);

static inline uint64_t align256(uint64_t b) { return (b + 255) / 256 * 256; }

// First position of every digit value b in [0, nb] of keys sorted on the
// digit (key >> shift) & mask; position nb is n.
template <class KT>
__global__ void k_digit_starts(const KT *__restrict__ ks, uint64_t n, uint32_t shift, uint32_t mask, uint32_t nb,
                               uint64_t *__restrict__ start) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= nb; b += gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = n;
        if (b < nb) {
            while (lo < hi) {
                const uint64_t mid = lo + (hi - lo) / 2;
                if (((uint32_t)(ks[mid] >> shift) & mask) < b) lo = mid + 1; else hi = mid;
            }
        }
        start[b] = b < nb ? lo : n;
    }
}

// ==================================================================== gather
template <class V>
__global__ __launch_bounds__(256) void k_gather(const V *__restrict__ src, const uint32_t *__restrict__ idx,
                                                uint64_t n, V *__restrict__ dst) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step)
        __builtin_nontemporal_store(src[__builtin_nontemporal_load(&idx[i])], &dst[i]);
}

// ==================================================================== stable LSD radix sort
// Sorts (key, u32 value) pairs by key bits [lo_bit, hi_bit), 8 bits a pass.
// Each pass: per-tile digit histogram (tile-major), exclusive scan in
// digit-major order, stable scatter. Stability inside a tile comes from
// processing 256-element rows in order with per-wave digit counts.
constexpr int kSortBlock = 256;
constexpr int kSortRows = 16;
constexpr uint32_t kSortTile = kSortBlock * kSortRows;  // 4096

template <class KT>
__global__ __launch_bounds__(kSortBlock) void k_sort_hist(const KT *__restrict__ keys, uint64_t n,
                                                          uint32_t shift, uint32_t mask,
                                                          uint32_t ntiles, uint32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRows; ++r) {
        const uint64_t i = base + (uint64_t)r * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(uint32_t)(keys[i] >> shift) & mask], 1u);
    }
    __syncthreads();
    // digit-major: hist[d * ntiles + tile]
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

template <class KT>
__global__ __launch_bounds__(kSortBlock) void k_sort_scatter(const KT *__restrict__ keys_in,
                                                             const uint32_t *__restrict__ vals_in,
                                                             uint64_t n, uint32_t shift, uint32_t mask,
                                                             uint32_t nbits, uint32_t ntiles,
                                                             const uint32_t *__restrict__ offs,
                                                             KT *__restrict__ keys_out,
                                                             uint32_t *__restrict__ vals_out) {
    constexpr int W = kSortBlock / 64;
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[W][256];
    run[threadIdx.x] = offs[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
    const int w = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRows; ++r) {
        for (int k = 0; k < W; ++k) wcnt[k][threadIdx.x] = 0;
        __syncthreads();
        const uint64_t i = base + (uint64_t)r * kSortBlock + threadIdx.x;
        const bool ok = i < n;
        KT key = ok ? keys_in[i] : KT(0);
        uint32_t d = ok ? ((uint32_t)(key >> shift) & mask) : 0xFFFFFFFFu;
        // lanes of this wave with the same digit
        uint64_t peers = __ballot(ok);
        for (uint32_t b = 0; b < nbits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lanemask_lt());
        if (ok && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        // thread t owns digit t: per-wave bases, then advance the running offset
        {
            uint32_t acc = run[threadIdx.x];
            for (int k = 0; k < W; ++k) {
                const uint32_t c = wcnt[k][threadIdx.x];
                wcnt[k][threadIdx.x] = acc;
                acc += c;
            }
            run[threadIdx.x] = acc;
        }
        __syncthreads();
        if (ok) {
            const uint32_t dst = wcnt[w][d] + rank;
            keys_out[dst] = key;
            vals_out[dst] = vals_in ? vals_in[i] : (uint32_t)i;
        }
        __syncthreads();
    }
}

uint64_t sort_ws_bytes(uint64_t n, uint32_t key_bytes) {
    const uint64_t ntiles = (n + kSortTile - 1) / kSortTile;
    return 2 * align256(n * key_bytes) + 2 * align256(n * 4) + align256(ntiles * 256 * 4) +
           scan_tmp_bytes(ntiles * 256, 4);
}

// Sorted values (original indices when vals_in is NULL) end in *vals_sorted,
// keys in *keys_sorted; both point into ws.
template <class KT>
bpt_status sort_pairs(const KT *keys_in, const uint32_t *vals_in, uint64_t n, uint32_t lo_bit,
                      uint32_t hi_bit, void *ws, uint64_t ws_bytes, KT **keys_sorted,
                      uint32_t **vals_sorted, hipStream_t s) {
    BPT_CHECK(n < (1ull << 32), BPT_ERR_CAPACITY, "sort of more than 2^32 elements");
    BPT_CHECK(ws_bytes >= sort_ws_bytes(n, sizeof(KT)), BPT_ERR_INVALID, "sort workspace too small");
    const uint64_t ntiles = (n + kSortTile - 1) / kSortTile;
    uint8_t *p = reinterpret_cast<uint8_t *>(ws);
    KT *kbuf[2];
    uint32_t *vbuf[2];
    kbuf[0] = reinterpret_cast<KT *>(p); p += align256(n * sizeof(KT));
    kbuf[1] = reinterpret_cast<KT *>(p); p += align256(n * sizeof(KT));
    vbuf[0] = reinterpret_cast<uint32_t *>(p); p += align256(n * 4);
    vbuf[1] = reinterpret_cast<uint32_t *>(p); p += align256(n * 4);
    uint32_t *hist = reinterpret_cast<uint32_t *>(p); p += align256(ntiles * 256 * 4);
    void *scan_tmp = p;
    const uint64_t scan_bytes = scan_tmp_bytes(ntiles * 256, 4);
    const KT *kin = keys_in;
    const uint32_t *vin = vals_in;
    int cur = 0;
    if (n == 0) { *keys_sorted = kbuf[0]; *vals_sorted = vbuf[0]; return BPT_OK; }
    for (uint32_t b = lo_bit; b < hi_bit; b += 8) {
        const uint32_t nbits = std::min<uint32_t>(8, hi_bit - b);
        const uint32_t mask = (1u << nbits) - 1;
        k_sort_hist<KT><<<(uint32_t)ntiles, kSortBlock, 0, s>>>(kin, n, b, mask, (uint32_t)ntiles, hist);
        BPT_HIP(hipGetLastError());
        BPT_TRY(exclusive_scan<uint32_t>(hist, ntiles * 256, hist, scan_tmp, scan_bytes, s));
        k_sort_scatter<KT><<<(uint32_t)ntiles, kSortBlock, 0, s>>>(kin, vin, n, b, mask, nbits,
                                                                  (uint32_t)ntiles, hist, kbuf[cur], vbuf[cur]);
        BPT_HIP(hipGetLastError());
        kin = kbuf[cur];
        vin = vbuf[cur];
        cur ^= 1;
    }
    if (vin == nullptr) {
        // zero passes: identity permutation
        BPT_CHECK(false, BPT_ERR_INVALID, "empty bit range");
    }
    *keys_sorted = const_cast<KT *>(kin);
    *vals_sorted = const_cast<uint32_t *>(vin);
    return BPT_OK;
}

template bpt_status sort_pairs<uint32_t>(const uint32_t *, const uint32_t *, uint64_t, uint32_t, uint32_t,
                                         void *, uint64_t, uint32_t **, uint32_t **, hipStream_t);
template bpt_status sort_pairs<uint64_t>(const uint64_t *, const uint32_t *, uint64_t, uint32_t, uint32_t,
                                         void *, uint64_t, uint64_t **, uint32_t **, hipStream_t);

// ==================================================================== keys
__device__ __host__ __forceinline__ uint64_t spread32(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

// Quotient of a Z-order coordinate by its interval; a zero interval (an
// all-zero side, see zorder_interval) gives 0.
__host__ __device__ __forceinline__ uint64_t zorder_div(uint64_t v, uint64_t d) { return d ? v / d : 0; }

__global__ void k_morton(const uint32_t *__restrict__ x, const uint32_t *__restrict__ y, uint64_t n,
                         uint64_t *__restrict__ z) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step)
        z[i] = spread32(x[i]) | (spread32(y[i]) << 1);
}

// Maxima of a and b: wave shuffles, then the block's waves through LDS, then
// one atomic per block (atomics of every wave on two words serialize).
template <class KT>
__global__ __launch_bounds__(256) void k_max2(const KT *__restrict__ a, const KT *__restrict__ b, uint64_t n,
                                              unsigned long long *__restrict__ out) {
    __shared__ unsigned long long wa[4], wb[4];
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long ma = 0, mb = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        ma = std::max<unsigned long long>(ma, (unsigned long long)a[i]);
        mb = std::max<unsigned long long>(mb, (unsigned long long)b[i]);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        ma = std::max<unsigned long long>(ma, __shfl_xor(ma, d, 64));
        mb = std::max<unsigned long long>(mb, __shfl_xor(mb, d, 64));
    }
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wa[w] = ma;
        wb[w] = mb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t k = 1; k < blockDim.x / 64; ++k) {
            ma = std::max<unsigned long long>(ma, wa[k]);
            mb = std::max<unsigned long long>(mb, wb[k]);
        }
        atomicMax(&out[0], ma);
        atomicMax(&out[1], mb);
    }
}

template <class KT>
__global__ void k_zorder_keys(const KT *__restrict__ a, const KT *__restrict__ b, uint64_t n,
                              uint64_t ix, uint64_t iy, uint64_t *__restrict__ out) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const uint32_t xg = (uint32_t)zorder_div((uint64_t)a[i], ix);
        const uint32_t yg = (uint32_t)zorder_div((uint64_t)b[i], iy);
        out[i] = spread32(xg) | (spread32(yg) << 1);
    }
}

__device__ __forceinline__ uint32_t bucket2d(uint64_t a, uint64_t b, int pass) {
    // grouping2d.cu:24-31: BIT_PASS 3 -> 22 passes of a 64-bit key, 4-bit nibbles
    const int shift = 3 * (21 - pass);
    if (pass > 15 || shift < 0 || shift >= 64) return 0;
    return (uint32_t)(((a >> shift) & 0xF) * ((b >> shift) & 0xF));
}

template <class KT>
__global__ void k_group2d_keys(const KT *__restrict__ a, const KT *__restrict__ b, uint64_t n,
                               int pass0, uint32_t *__restrict__ out) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const uint64_t x = a[i], y = b[i];
        out[i] = (bucket2d(x, y, pass0) << 24) | (bucket2d(x, y, pass0 + 1) << 16) |
                 (bucket2d(x, y, pass0 + 2) << 8) | bucket2d(x, y, pass0 + 3);
    }
}

__host__ __device__ static int bitwidth(uint64_t v) { int b = 0; while (v) { v >>= 1; ++b; } return b; }

template <class KT>
bpt_status pair_max(const KT *a, const KT *b, uint64_t n, uint64_t *ma, uint64_t *mb, hipStream_t s) {
    unsigned long long *d = nullptr;
    BPT_HIP(hipMalloc(&d, 16));
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemsetAsync(d, 0, 16, s);
    if (e == hipSuccess) {
        uint32_t g = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cu_count() * 2);
        k_max2<KT><<<std::max<uint32_t>(g, 1), 256, 0, s>>>(a, b, n, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "pair max", __FILE__, __LINE__);
    *ma = h[0];
    *mb = h[1];
    return BPT_OK;
}

// zOrder2d.cu:129-130 computes int interval = ceil((float)max / 2^32); the
// kernel (:57, :27) carries it as ulli then int, and :29 divides the ulli
// input by it, converting the int back to ulli (sign-extended). This is that
// 64-bit divisor. A quotient of 2^31 or more does not fit the int: the
// reference's host is x86, whose float-to-int conversion gives INT_MIN, so
// the divisor is 0xFFFFFFFF80000000 (every quotient 0 or 1). That outcome is
// written out, because the device's own conversion saturates instead.
// max == 0 gives 0, a division by zero in the reference; zorder_div (above)
// takes every quotient as 0 then (every coordinate of that side is 0 anyway).
__host__ __device__ static uint64_t zorder_interval(uint64_t m) {
    const float r = (float)m / 4294967296.0f;
    if (r >= 2147483648.0f) return 0xFFFFFFFF80000000ull;
    return (uint64_t)(int64_t)(int)ceilf(r);
}

// grouping2d.cu:610-635
__host__ __device__ static int group2d_start_pass(uint64_t ma, uint64_t mb) {
    int a = 16 - (int)ceilf((float)bitwidth(ma) / 3.0f);
    int b = 16 - (int)ceilf((float)bitwidth(mb) / 3.0f);
    return a > b ? a : b;
}

// The same keys with their parameters taken from the maxima in device memory
// (mx[0], mx[1], written by k_max2), so no host round trip sits between.
// Only the top 32 bits of the Z key are kept: the order uses its top 16.
template <class KT>
__global__ void k_zorder_keys_dev(const KT *__restrict__ a, const KT *__restrict__ b, uint64_t n,
                                  const unsigned long long *__restrict__ mx, uint32_t *__restrict__ out) {
    const uint64_t ix = zorder_interval(mx[0]), iy = zorder_interval(mx[1]);
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const uint32_t xg = (uint32_t)zorder_div((uint64_t)a[i], ix);
        const uint32_t yg = (uint32_t)zorder_div((uint64_t)b[i], iy);
        out[i] = (uint32_t)((spread32(xg) | (spread32(yg) << 1)) >> 32);
    }
}

template <class KT>
__global__ void k_group2d_keys_dev(const KT *__restrict__ a, const KT *__restrict__ b, uint64_t n,
                                   const unsigned long long *__restrict__ mx, uint32_t *__restrict__ out) {
    int p0 = group2d_start_pass(mx[0], mx[1]);
    if (p0 < 0) p0 = 0;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const uint64_t x = a[i], y = b[i];
        out[i] = (bucket2d(x, y, p0) << 24) | (bucket2d(x, y, p0 + 1) << 16) | (bucket2d(x, y, p0 + 2) << 8) |
                 bucket2d(x, y, p0 + 3);
    }
}

template <class KT>
bpt_status zorder_keys(const KT *a, const KT *b, uint64_t n, uint64_t *out, hipStream_t s) {
    if (n == 0) return BPT_OK;
    uint64_t ma, mb;
    BPT_TRY(pair_max<KT>(a, b, n, &ma, &mb, s));
    uint32_t g = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cu_count() * 16);
    k_zorder_keys<KT><<<g, 256, 0, s>>>(a, b, n, zorder_interval(ma), zorder_interval(mb), out);
    BPT_HIP(hipGetLastError());
    return BPT_OK;
}

template <class KT>
bpt_status group2d_keys(const KT *a, const KT *b, uint64_t n, uint32_t *out, hipStream_t s) {
    if (n == 0) return BPT_OK;
    uint64_t ma, mb;
    BPT_TRY(pair_max<KT>(a, b, n, &ma, &mb, s));
    int p0 = group2d_start_pass(ma, mb);
    if (p0 < 0) p0 = 0;  // passes below 0 shift by 64 or more: every bucket is 0
    uint32_t g = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cu_count() * 16);
    k_group2d_keys<KT><<<g, 256, 0, s>>>(a, b, n, p0, out);
    BPT_HIP(hipGetLastError());
    return BPT_OK;
}

template bpt_status pair_max<uint64_t>(const uint64_t *, const uint64_t *, uint64_t, uint64_t *, uint64_t *,
                                      hipStream_t);
int group_start_pass_2d(uint64_t maxa, uint64_t maxb) { return group2d_start_pass(maxa, maxb); }
// grouping.cu:1736-1744: 8 - ceil(bits of the maximum / 8), in float as there
int group_start_pass_1d(uint64_t maxval) { return 8 - (int)ceilf((float)bitwidth(maxval) / 8.0f); }
template bpt_status zorder_keys<uint32_t>(const uint32_t *, const uint32_t *, uint64_t, uint64_t *, hipStream_t);
template bpt_status zorder_keys<uint64_t>(const uint64_t *, const uint64_t *, uint64_t, uint64_t *, hipStream_t);
template bpt_status group2d_keys<uint32_t>(const uint32_t *, const uint32_t *, uint64_t, uint32_t *, hipStream_t);
template bpt_status group2d_keys<uint64_t>(const uint64_t *, const uint64_t *, uint64_t, uint32_t *, hipStream_t);

// Order of n (lo, hi) queries under a grouping mode: *order points into ws.
// Layout: maxima (256 B) | keys (n * 8) | sort scratch.
uint64_t query_order_ws(uint64_t n) {
    return 256 + align256(n * 8) + sort_ws_bytes(n, 4);
}

// Clears the two maxima k_max2 accumulates into.
__global__ void k_zero_max2(unsigned long long *mx) {
    if (threadIdx.x < 2) mx[threadIdx.x] = 0;
}

template <class KT>
bpt_status query_order(const KT *lo, const KT *hi, uint64_t n, int mode, void *ws, uint64_t ws_bytes,
                       uint32_t **order, hipStream_t s) {
    BPT_CHECK(ws_bytes >= query_order_ws(n), BPT_ERR_INVALID, "order workspace too small");
    BPT_CHECK(mode == BPT_GROUP_GRID2D || mode == BPT_GROUP_ZORDER, BPT_ERR_INVALID, "unknown group mode %d", mode);
    unsigned long long *mx = reinterpret_cast<unsigned long long *>(ws);
    uint8_t *p = reinterpret_cast<uint8_t *>(ws) + 256;
    uint8_t *sort_ws = p + align256(n * 8);
    ws_bytes -= 256;
    uint32_t *ksorted = nullptr;
    *order = nullptr;
    if (n == 0) return BPT_OK;
    const uint32_t g = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cu_count() * 16);
    k_zero_max2<<<1, 64, 0, s>>>(mx);  // a kernel, not a memset: callers may capture the search into a graph
    k_max2<KT><<<(uint32_t)std::min<uint64_t>(g, (uint64_t)device_cu_count() * 2), 256, 0, s>>>(lo, hi, n, mx);
    BPT_HIP(hipGetLastError());
    if (mode == BPT_GROUP_GRID2D) {
        uint32_t *keys = reinterpret_cast<uint32_t *>(p);
        k_group2d_keys_dev<KT><<<g, 256, 0, s>>>(lo, hi, n, mx, keys);
        BPT_HIP(hipGetLastError());
        // the first two passes' buckets (the top 16 bits) order the groups: 65536
        // groups are enough for neighbouring lanes to share subtrees, and two
        // radix passes cost half of four
        return sort_pairs<uint32_t>(keys, nullptr, n, 16, 32, sort_ws, ws_bytes - align256(n * 8), &ksorted,
                                    order, s);
    }
    if (mode == BPT_GROUP_ZORDER) {
        uint32_t *z = reinterpret_cast<uint32_t *>(p);
        k_zorder_keys_dev<KT><<<g, 256, 0, s>>>(lo, hi, n, mx, z);
        BPT_HIP(hipGetLastError());
        // the top 16 bits of the 64-bit Z key order the groups (as above)
        return sort_pairs<uint32_t>(z, nullptr, n, 16, 32, sort_ws, ws_bytes - align256(n * 8), &ksorted, order, s);
    }
    set_error("unknown group mode %d", mode);
    return BPT_ERR_INVALID;
}

template bpt_status query_order<uint32_t>(const uint32_t *, const uint32_t *, uint64_t, int, void *, uint64_t,
                                          uint32_t **, hipStream_t);
template bpt_status query_order<uint64_t>(const uint64_t *, const uint64_t *, uint64_t, int, void *, uint64_t,
                                          uint32_t **, hipStream_t);

uint64_t query_order_ws_any(uint64_t n) {
    return 256 + align256(n * 8) + std::max(sort_ws_bytes(n, 4), sort_ws_bytes(n, 8));
}

}  // namespace bpt

using namespace bpt;

extern "C" {

bpt_status bpt_radix_group(const void *d_keys, uint32_t key_bytes, uint64_t n, uint32_t shift,
                           uint32_t bits, uint32_t *d_perm, uint64_t *d_bucket_start, bpt_stream stream) {
    BPT_CHECK(key_bytes == 4 || key_bytes == 8, BPT_ERR_INVALID, "key_bytes %u", key_bytes);
    BPT_CHECK(bits >= 1 && bits <= 16 && shift + bits <= key_bytes * 8, BPT_ERR_INVALID,
              "bucket bits [%u, %u) outside the key", shift, shift + bits);
    BPT_CHECK(n == 0 || (d_keys && d_perm), BPT_ERR_INVALID, "NULL argument");
    hipStream_t s = as_stream(stream);
    const uint32_t nb = 1u << bits;
    if (n == 0) {
        if (d_bucket_start) BPT_HIP(hipMemsetAsync(d_bucket_start, 0, (nb + 1) * 8ull, s));
        BPT_HIP(hipStreamSynchronize(s));
        return BPT_OK;
    }
    const uint64_t ws_bytes = sort_ws_bytes(n, key_bytes);
    void *ws = nullptr;
    BPT_HIP(hipMalloc(&ws, ws_bytes));
    bpt_status st;
    uint64_t *bucket_keys_u64 = nullptr;
    uint32_t *sorted_u32 = nullptr;
    uint32_t *perm = nullptr;
    if (key_bytes == 4) {
        uint32_t *ks = nullptr;
        st = sort_pairs<uint32_t>((const uint32_t *)d_keys, nullptr, n, shift, shift + bits, ws, ws_bytes, &ks,
                                  &perm, s);
        sorted_u32 = ks;
    } else {
        uint64_t *ks = nullptr;
        st = sort_pairs<uint64_t>((const uint64_t *)d_keys, nullptr, n, shift, shift + bits, ws, ws_bytes, &ks,
                                  &perm, s);
        bucket_keys_u64 = ks;
    }
    if (st == BPT_OK) {
        hipError_t e = hipMemcpyAsync(d_perm, perm, n * 4, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) st = hip_fail(e, "perm copy", __FILE__, __LINE__);
    }
    if (st == BPT_OK && d_bucket_start) {
        // bucket starts = lower bound of every bucket in the sorted digit sequence
        const uint32_t g = (nb + 1 + 255) / 256;
        if (key_bytes == 4) {
            k_digit_starts<uint32_t><<<g, 256, 0, s>>>(sorted_u32, n, shift, nb - 1, nb, d_bucket_start);
        } else {
            k_digit_starts<uint64_t><<<g, 256, 0, s>>>(bucket_keys_u64, n, shift, nb - 1, nb, d_bucket_start);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) st = hip_fail(e, "k_digit_starts", __FILE__, __LINE__);
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(ws);
    return st;
}

bpt_status bpt_sort_perm(const void *d_keys, uint32_t key_bytes, uint64_t n, uint32_t lo_bit, uint32_t hi_bit,
                         uint32_t *d_perm, bpt_stream stream) {
    BPT_CHECK(key_bytes == 4 || key_bytes == 8, BPT_ERR_INVALID, "key_bytes %u", key_bytes);
    BPT_CHECK(lo_bit < hi_bit && hi_bit <= key_bytes * 8, BPT_ERR_INVALID, "bit range [%u, %u)", lo_bit, hi_bit);
    BPT_CHECK(n < (1ull << 32), BPT_ERR_CAPACITY, "sort of more than 2^32 elements");
    BPT_CHECK(n == 0 || (d_keys && d_perm), BPT_ERR_INVALID, "NULL argument");
    if (n == 0) return BPT_OK;
    hipStream_t s = as_stream(stream);
    const uint64_t ws_bytes = sort_ws_bytes(n, key_bytes);
    void *ws = nullptr;
    BPT_HIP(hipMalloc(&ws, ws_bytes));
    uint32_t *perm = nullptr;
    bpt_status st;
    if (key_bytes == 4) {
        uint32_t *ks = nullptr;
        st = sort_pairs<uint32_t>((const uint32_t *)d_keys, nullptr, n, lo_bit, hi_bit, ws, ws_bytes, &ks, &perm, s);
    } else {
        uint64_t *ks = nullptr;
        st = sort_pairs<uint64_t>((const uint64_t *)d_keys, nullptr, n, lo_bit, hi_bit, ws, ws_bytes, &ks, &perm, s);
    }
    if (st == BPT_OK) {
        hipError_t e = hipMemcpyAsync(d_perm, perm, n * 4, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_fail(e, "sort perm copy", __FILE__, __LINE__);
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(ws);
    return st;
}

bpt_status bpt_gather(const void *d_src, uint32_t elem_bytes, const uint32_t *d_idx, uint64_t n, void *d_dst,
                      bpt_stream stream) {
    BPT_CHECK(elem_bytes == 4 || elem_bytes == 8, BPT_ERR_INVALID, "elem_bytes %u", elem_bytes);
    BPT_CHECK(n == 0 || (d_src && d_idx && d_dst), BPT_ERR_INVALID, "NULL argument");
    if (n == 0) return BPT_OK;
    hipStream_t s = as_stream(stream);
    const uint32_t g = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cu_count() * 16);
    if (elem_bytes == 4) k_gather<uint32_t><<<g, 256, 0, s>>>((const uint32_t *)d_src, d_idx, n, (uint32_t *)d_dst);
    else k_gather<uint64_t><<<g, 256, 0, s>>>((const uint64_t *)d_src, d_idx, n, (uint64_t *)d_dst);
    BPT_HIP(hipGetLastError());
    return BPT_OK;
}

bpt_status bpt_sort_pairs(void *d_keys, void *d_vals, uint32_t key_bytes, uint32_t val_bytes, uint64_t n,
                          bpt_stream stream) {
    BPT_CHECK(val_bytes == 4 || val_bytes == 8, BPT_ERR_INVALID, "val_bytes %u", val_bytes);
    if (n < 2) return BPT_OK;
    hipStream_t s = as_stream(stream);
    uint32_t *perm = nullptr;
    void *tmp = nullptr;
    BPT_HIP(hipMalloc(&perm, n * 4));
    hipError_t e = hipMalloc(&tmp, n * std::max(key_bytes, val_bytes));
    if (e != hipSuccess) {
        (void)hipFree(perm);
        return hip_fail(e, "hipMalloc(sort tmp)", __FILE__, __LINE__);
    }
    bpt_status st = bpt_sort_perm(d_keys, key_bytes, n, 0, key_bytes * 8, perm, stream);
    if (st == BPT_OK) st = bpt_gather(d_keys, key_bytes, perm, n, tmp, stream);
    if (st == BPT_OK) {
        e = hipMemcpyAsync(d_keys, tmp, n * key_bytes, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) st = hip_fail(e, "sort keys copy", __FILE__, __LINE__);
    }
    if (st == BPT_OK && d_vals) st = bpt_gather(d_vals, val_bytes, perm, n, tmp, stream);
    if (st == BPT_OK && d_vals) {
        e = hipMemcpyAsync(d_vals, tmp, n * val_bytes, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) st = hip_fail(e, "sort values copy", __FILE__, __LINE__);
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(perm);
    (void)hipFree(tmp);
    return st;
}

bpt_status bpt_morton_encode(const uint32_t *d_x, const uint32_t *d_y, uint64_t n, uint64_t *d_z,
                             bpt_stream stream) {
    BPT_CHECK(n == 0 || (d_x && d_y && d_z), BPT_ERR_INVALID, "NULL argument");
    if (n == 0) return BPT_OK;
    uint32_t g = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cu_count() * 16);
    k_morton<<<g, 256, 0, as_stream(stream)>>>(d_x, d_y, n, d_z);
    BPT_HIP(hipGetLastError());
    return BPT_OK;
}

bpt_status bpt_zorder2d(const uint64_t *d_a, const uint64_t *d_b, uint64_t n, uint64_t *d_out,
                        bpt_stream stream) {
    BPT_CHECK(n == 0 || (d_a && d_b && d_out), BPT_ERR_INVALID, "NULL argument");
    BPT_TRY(zorder_keys<uint64_t>(d_a, d_b, n, d_out, as_stream(stream)));
    BPT_HIP(hipStreamSynchronize(as_stream(stream)));
    return BPT_OK;
}

bpt_status bpt_group2d_keys(const uint64_t *d_a, const uint64_t *d_b, uint64_t n, uint32_t *d_out,
                            bpt_stream stream) {
    BPT_CHECK(n == 0 || (d_a && d_b && d_out), BPT_ERR_INVALID, "NULL argument");
    BPT_TRY(group2d_keys<uint64_t>(d_a, d_b, n, d_out, as_stream(stream)));
    BPT_HIP(hipStreamSynchronize(as_stream(stream)));
    return BPT_OK;
}

bpt_status bpt_range_count_grouped(bpt_tree *t, const void *d_lo, const void *d_hi, uint64_t nq, int group_mode,
                                   uint64_t *d_first, uint64_t *d_count, bpt_stream stream) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    BPT_CHECK(nq == 0 || (d_lo && d_hi && d_count), BPT_ERR_INVALID, "NULL argument");
    hipStream_t s = as_stream(stream);
    if (group_mode == BPT_GROUP_NONE || nq == 0)
        return range_count_ordered(t, d_lo, d_hi, nq, nullptr, d_first, d_count, s);
    BPT_CHECK(group_mode == BPT_GROUP_GRID2D || group_mode == BPT_GROUP_ZORDER, BPT_ERR_INVALID,
              "unknown group mode %d", group_mode);
    BPT_CHECK(nq < (1ull << 32), BPT_ERR_CAPACITY, "grouped range search of more than 2^32 queries");
    // the order's scratch is the tree's workspace (kept between calls)
    const uint64_t ws_bytes = query_order_ws_any(nq);
    BPT_TRY(t->ws.reserve(ws_bytes));
    uint32_t *order = nullptr;
    bpt_status st = t->key_bytes == 4
                        ? query_order<uint32_t>((const uint32_t *)d_lo, (const uint32_t *)d_hi, nq, group_mode,
                                                t->ws.ptr, ws_bytes, &order, s)
                        : query_order<uint64_t>((const uint64_t *)d_lo, (const uint64_t *)d_hi, nq, group_mode,
                                                t->ws.ptr, ws_bytes, &order, s);
    if (st == BPT_OK) st = range_count_ordered(t, d_lo, d_hi, nq, order, d_first, d_count, s);
    return st;
}

}  // extern "C"
