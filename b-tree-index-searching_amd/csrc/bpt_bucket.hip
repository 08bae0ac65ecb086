// Grouped point search: queries are partitioned into leaf-range buckets, so
// that each bucket's slice of an upper level fits in LDS and a lookup reads
// one line (the leaf) from global memory instead of one per level.
//
// The reference groups queries on their key bits (genGrouping,
// grouping.cu:1677-1851) and then searches each group from the root
// (kernelSearchingKeyGroupTwoStep, BPTreeExactSearch.cu:206-294). Here a
// bucket is a range of leaves chosen from the tree itself, so buckets are
// balanced in leaves whatever the key distribution:
//   bucket b = leaves [b*lpb, (b+1)*lpb); its separator = its largest key;
//   slice    = the bucket's sub-leaf maxima (the largest key of every
//              64-byte half or quarter of a leaf, kept by the tree), or its
//              entries of level s (s = 1: the leaf maxima), at most the
//              tree's group_slice bytes (16 to 128 KiB), staged in LDS by the
//              search kernel. With sub-leaf maxima a query reads the 64 bytes
//              of its leaf that hold its key, not the whole leaf.
// Passes over the queries (tiles of T queries):
//   k_bkt_hist       bucket counts of each tile, one row per tile
//   column scan      first slot of every (tile, bucket) run in bucket-major
//                    order: per-group column sums (k_bkt_colsum), a scan
//                    across groups and buckets (k_bkt_colscan), per-row
//                    offsets (k_bkt_colprefix); every row stays contiguous
//   k_bkt_scatter    tile counting-sorted in LDS, written out as runs
//   k_bkt_search     chunks of the bucket-ordered queries, grid-strided so
//                    that the blocks in flight share a window of buckets
//                    (their leaves stay in cache); lower bound in the LDS
//                    slice, then levels s-1..1 and the leaf (or the sub-leaf)
//                    from global memory
//   k_bkt_unpermute  each tile gathers its runs back through LDS into the
//                    caller's order
// Every step is deterministic in its results (the order inside a run is not,
// but each query carries its position in the tile).
#include <algorithm>

#include "bpt_internal.hpp"

namespace bpt {

constexpr int kBBlock = 256;
constexpr uint32_t kMaxSliceBytes = 131072; // LDS slice of the search kernel (at most)
constexpr uint32_t kMaxBuckets = 1024;     // runs per partition tile (one partition thread each)
constexpr size_t kMaxLds = 160 * 1024;     // LDS of a workgroup (gfx950)
constexpr uint64_t kBucketChunk = 1ull << 28;  // queries per partition (u32 slots)

constexpr uint32_t kColGroup = 128;          // tiles per column-sum group
constexpr uint32_t kColStep = 8;             // column loads in flight per thread

struct BucketDev {
    uint32_t nb;       // buckets
    uint32_t level;    // slice level s (1 = leaf maxima)
    uint64_t lpb;      // leaves per bucket
    uint32_t epb;      // slice entries per bucket (entries of level s)
    uint32_t tile;     // queries per partition tile (T)
    uint32_t ntiles;
    uint32_t ngroups;  // column-sum groups of kColGroup tiles
    uint32_t chunk;    // queries per search work chunk
    uint32_t xcd;      // search chunks split per XCD
    uint32_t sbits;    // prefix-table bits of a search slice
    uint32_t spl;      // sub-leaves per leaf when the slice holds sub-leaf maxima (0: level `level`)
};

static inline uint64_t align256(uint64_t b) { return (b + 255) / 256 * 256; }

// Partition tile: T queries, kPBlock threads with T / kPBlock each (all of a
// thread's loads are issued before its first bucket search). Large tiles keep
// the (tile, bucket) runs long, so the runs are written and read back as
// whole lines; a tile's staging uses up to ~144 KiB of the 160 KiB of LDS.
constexpr int kPBlock = 1024;
static_assert(kMaxBuckets <= (uint32_t)kPBlock, "one partition thread per bucket");
template <class K> constexpr uint32_t tile_of() { return sizeof(K) == 4 ? 16384u : 8192u; }
template <class K> constexpr int per_thread() { return (int)(tile_of<K>() / kPBlock); }

// Key-prefix index over a sorted array a[0, n) with a[0] >= base: entry p of
// its table (p < 2^bits) is the lower bound of base + (p << shift) in a, entry
// 2^bits is n. The lower bound of a key with prefix p then lies in
// [tab[p], tab[p + 1]], so a search touches one table entry and a window of
// about n / 2^bits elements instead of all log2(n) levels (fewer random LDS
// reads, which is where the bank conflicts are).
template <class K>
struct PrefixIndex {
    K base;
    uint32_t shift;
};

template <class K>
__device__ __forceinline__ int bit_width(K v) {
    if constexpr (sizeof(K) == 4) return v ? 32 - __clz((int)v) : 0;
    else return v ? 64 - __clzll((long long)v) : 0;
}

template <class K>
__device__ __forceinline__ PrefixIndex<K> prefix_index(K lo, K hi, uint32_t bits) {
    const int bw = bit_width<K>(hi - lo);
    return {lo, bw > (int)bits ? (uint32_t)(bw - (int)bits) : 0u};
}

// Start key of prefix p (saturating at the largest key).
template <class K>
__device__ __forceinline__ K prefix_start(const PrefixIndex<K> &pi, uint32_t p) {
    const K off = (K)p << pi.shift;
    const K s = pi.base + off;
    return s < pi.base ? ~K(0) : s;
}

template <class K>
__device__ __forceinline__ uint32_t prefix_of(const PrefixIndex<K> &pi, K q, uint32_t bits) {
    const K dlt = q > pi.base ? q - pi.base : K(0);
    const K p = dlt >> pi.shift;
    const K top = ((K)1 << bits) - 1;
    return (uint32_t)(p < top ? p : top);
}

// Lower bound of q in a[lo, hi) (hi when every element there is < q).
template <class K>
__device__ __forceinline__ uint32_t lower_bound_window(const K *a, uint32_t lo, uint32_t hi, K q) {
    return hi > lo ? lo + lower_bound_flat<K>(a + lo, hi - lo, q) : lo;
}

// Prefix table of a sorted LDS array a[0, n), n >= 1, built by all threads of
// the block, with pi = prefix_index(a[0], a[n-1], bits): tab[p] = the lower
// bound of prefix_start(p) for p < 2^bits, tab[2^bits] = n. With f(x) =
// (x - base) >> shift, a[e] >= prefix_start(p) exactly when f(a[e]) >= p, so
// entry e takes the prefixes in (f(a[e-1]), f(a[e])] and n takes those past
// f(a[n-1]) (up to half the table, written by all threads): one pass over the
// array, every table entry written once, instead of a binary search per table
// entry. A gap in the keys makes one entry's range long (it is written by that
// entry's thread).
template <class K>
__device__ void build_prefix_table(const K *a, uint32_t n, const PrefixIndex<K> &pi, uint32_t bits, uint16_t *tab) {
    const uint32_t np = 1u << bits;
    auto f = [&](K x) -> uint32_t { return (uint32_t)((K)(x - pi.base) >> pi.shift); };  // < np for x in a
    for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
        const uint32_t lo = e == 0 ? 0u : f(a[e - 1]) + 1;
        const uint32_t hi = f(a[e]) + 1;
        for (uint32_t p = lo; p < hi; ++p) tab[p] = (uint16_t)e;
    }
    for (uint32_t p = f(a[n - 1]) + 1 + threadIdx.x; p <= np; p += blockDim.x) tab[p] = (uint16_t)n;
}

constexpr uint32_t kBucketTabBits = 10;  // prefix table of the bucket separators
constexpr uint32_t kMaxSliceTabBits = 13;  // prefix table of a search slice (at most)

template <class K>
__device__ __forceinline__ uint32_t bucket_of(const K *bsep, uint32_t nb, const uint16_t *btab,
                                              const PrefixIndex<K> &pi, K q) {
    const uint32_t p = prefix_of<K>(pi, q, kBucketTabBits);
    const uint32_t b = lower_bound_window<K>(bsep, btab[p], btab[p + 1], q);
    return b < nb ? b : nb - 1;
}

// Blocks are dispatched to the 8 XCDs round-robin (block i on XCD i % 8).
// Logical index of block i such that every XCD owns a contiguous range of
// [0, n) (n = number of blocks): neighbouring tiles then share an L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t i, uint32_t n) {
    const uint32_t x = i % kXcds, k = i / kXcds, per = n / kXcds, rem = n % kXcds;
    return x * per + (x < rem ? x : rem) + k;
}

// In-place exclusive scan of a[0, n) in LDS, all threads of the block;
// returns the total. wsum: blockDim.x / 64 entries of LDS.
__device__ uint32_t block_exclusive_scan(uint32_t *a, uint32_t n, uint32_t *wsum) {
    const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t local = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < n) local += a[b0 + k];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = local;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t run = x - local, total = 0;
    for (uint32_t k = 0; k < nw; ++k) {
        if (k < w) run += wsum[k];
        total += wsum[k];
    }
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < n) {
            const uint32_t v = a[b0 + k];
            a[b0 + k] = run;
            run += v;
        }
    __syncthreads();
    return total;
}

// Bucket separators: the largest key of every bucket (a leaf maximum).
template <class K>
__global__ void k_bkt_seps(TreeView t, BucketDev d, K *__restrict__ bsep) {
    const K *leafmax = reinterpret_cast<const K *>(t.inner + t.off[1]);
    const uint64_t L = t.count[0];
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < d.nb; b += gridDim.x * blockDim.x)
        bsep[b] = leafmax[std::min<uint64_t>((uint64_t)(b + 1) * d.lpb, L) - 1];
}

// Prefix table of the bucket separators: hdr = {base, shift}, btab[2^10 + 1].
template <class K>
__global__ __launch_bounds__(kBBlock) void k_bkt_btab(TreeView t, BucketDev d, const K *__restrict__ bsep,
                                                      uint64_t *__restrict__ hdr, uint16_t *__restrict__ btab) {
    const K kmin = *reinterpret_cast<const K *>(t.leaves);  // smallest key (slot 0)
    const PrefixIndex<K> pi = prefix_index<K>(kmin, bsep[d.nb - 1], kBucketTabBits);
    const uint32_t nent = 1u << kBucketTabBits;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= nent; p += gridDim.x * blockDim.x)
        btab[p] = (uint16_t)(p == nent ? d.nb : lower_bound_flat<K>(bsep, d.nb, prefix_start<K>(pi, p)));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        hdr[0] = (uint64_t)pi.base;
        hdr[1] = pi.shift;
    }
}

// Entries per bucket in the table of slice prefix tables: 2^sbits + 1,
// rounded up to whole 16-byte pieces (room for one word past the table).
__host__ __device__ inline uint32_t stab_stride(uint32_t sbits) { return ((1u << sbits) + 2 + 7) / 8 * 8; }

// The prefix table of every bucket's search slice, built once per call from
// the slice source in global memory: stabs[b * stride ...] is bucket b's
// table, spi[2b], spi[2b + 1] its base and shift. The search blocks copy a
// bucket's table with its slice instead of each rebuilding it.
template <class K>
__global__ __launch_bounds__(kPBlock) void k_bkt_stabs(BucketDev d, const K *__restrict__ level, uint64_t level_count,
                                                       uint16_t *__restrict__ stabs, uint64_t *__restrict__ spi) {
    const uint32_t b = blockIdx.x;
    const uint64_t first = (uint64_t)b * d.epb;
    const uint32_t eb = (uint32_t)std::min<uint64_t>(d.epb, level_count - first);
    const K *a = level + first;
    const PrefixIndex<K> pi = prefix_index<K>(a[0], a[eb - 1], d.sbits);
    build_prefix_table<K>(a, eb, pi, d.sbits, stabs + (uint64_t)b * stab_stride(d.sbits));
    if (threadIdx.x == 0) {
        spi[2 * b] = (uint64_t)pi.base;
        spi[2 * b + 1] = pi.shift;
    }
}

template <class K>
__global__ __launch_bounds__(kPBlock) void k_bkt_hist(const K *__restrict__ q, uint64_t nq, BucketDev d,
                                                      const K *__restrict__ bsep_g, const uint64_t *__restrict__ hdr,
                                                      const uint16_t *__restrict__ btab_g,
                                                      uint32_t *__restrict__ hist) {
    extern __shared__ uint4 lds_raw[];
    K *bsep = reinterpret_cast<K *>(lds_raw);
    uint32_t *cnt = reinterpret_cast<uint32_t *>(bsep + d.nb);
    uint16_t *btab = reinterpret_cast<uint16_t *>(cnt + d.nb);
    for (uint32_t b = threadIdx.x; b < d.nb; b += blockDim.x) {
        bsep[b] = bsep_g[b];
        cnt[b] = 0;
    }
    for (uint32_t p = threadIdx.x; p <= (1u << kBucketTabBits); p += blockDim.x) btab[p] = btab_g[p];
    const PrefixIndex<K> pi{(K)hdr[0], (uint32_t)hdr[1]};
    const uint32_t tile = xcd_contiguous(blockIdx.x, gridDim.x);
    constexpr int E = per_thread<K>();
    const uint64_t lo = (uint64_t)tile * d.tile;
    const uint32_t cq = (uint32_t)(std::min<uint64_t>(lo + d.tile, nq) - lo);
    K key[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = e * kPBlock + threadIdx.x;
        key[e] = i < cq ? __builtin_nontemporal_load(&q[lo + i]) : K(0);
    }
    __syncthreads();
    uint32_t bk[E];
#pragma unroll
    for (int e = 0; e < E; ++e) bk[e] = bucket_of<K>(bsep, d.nb, btab, pi, key[e]);
#pragma unroll
    for (int e = 0; e < E; ++e)
        if (e * kPBlock + threadIdx.x < cq) atomicAdd(&cnt[bk[e]], 1u);
    __syncthreads();
    uint32_t *row = hist + (uint64_t)tile * d.nb;
    for (uint32_t b = threadIdx.x; b < d.nb; b += blockDim.x) row[b] = cnt[b];
}

// Column sums of groups of kColGroup rows: S[g][b].
__global__ __launch_bounds__(kBBlock) void k_bkt_colsum(const uint32_t *__restrict__ hist, BucketDev d,
                                                        uint32_t *__restrict__ S) {
    const uint32_t g = blockIdx.x;
    const uint32_t t0 = g * kColGroup, t1 = std::min(t0 + kColGroup, d.ntiles);
    for (uint32_t b = threadIdx.x; b < d.nb; b += blockDim.x) {
        uint32_t acc = 0;
        for (uint32_t t = t0; t < t1; t += kColStep) {  // kColStep independent loads at a time
            uint32_t v[kColStep];
#pragma unroll
            for (uint32_t k = 0; k < kColStep; ++k) v[k] = t + k < t1 ? hist[(uint64_t)(t + k) * d.nb + b] : 0u;
#pragma unroll
            for (uint32_t k = 0; k < kColStep; ++k) acc += v[k];
        }
        S[(uint64_t)g * d.nb + b] = acc;
    }
}

// One block of kScanThreads: S[g][b] becomes the count of bucket b in the
// groups before g; bstart[b] = first slot of bucket b, bstart[nb] = total.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_bkt_colscan(uint32_t *__restrict__ S, BucketDev d,
                                                              uint32_t *__restrict__ bstart) {
    __shared__ uint32_t tot[kMaxBuckets];
    __shared__ uint32_t wsum[kScanThreads / 64];
    for (uint32_t b = threadIdx.x; b < d.nb; b += blockDim.x) {
        uint32_t run = 0;
        for (uint32_t g = 0; g < d.ngroups; g += kColStep) {  // loads first: kColStep in flight
            uint32_t v[kColStep];
#pragma unroll
            for (uint32_t k = 0; k < kColStep; ++k) v[k] = g + k < d.ngroups ? S[(uint64_t)(g + k) * d.nb + b] : 0u;
#pragma unroll
            for (uint32_t k = 0; k < kColStep; ++k) {
                if (g + k < d.ngroups) S[(uint64_t)(g + k) * d.nb + b] = run;
                run += v[k];
            }
        }
        tot[b] = run;
    }
    __syncthreads();
    const uint32_t total = block_exclusive_scan(tot, d.nb, wsum);
    for (uint32_t b = threadIdx.x; b < d.nb; b += blockDim.x) bstart[b] = tot[b];
    if (threadIdx.x == 0) bstart[d.nb] = total;
}

// off[t][b] = first slot of tile t's run of bucket b.
__global__ __launch_bounds__(kBBlock) void k_bkt_colprefix(const uint32_t *__restrict__ hist,
                                                           const uint32_t *__restrict__ S,
                                                           const uint32_t *__restrict__ bstart, BucketDev d,
                                                           uint32_t *__restrict__ off) {
    const uint32_t g = blockIdx.x;
    const uint32_t t0 = g * kColGroup, t1 = std::min(t0 + kColGroup, d.ntiles);
    for (uint32_t b = threadIdx.x; b < d.nb; b += blockDim.x) {
        uint32_t run = bstart[b] + S[(uint64_t)g * d.nb + b];
        for (uint32_t t = t0; t < t1; t += kColStep) {  // kColStep independent loads at a time
            uint32_t v[kColStep];
#pragma unroll
            for (uint32_t k = 0; k < kColStep; ++k) v[k] = t + k < t1 ? hist[(uint64_t)(t + k) * d.nb + b] : 0u;
#pragma unroll
            for (uint32_t k = 0; k < kColStep; ++k) {
                if (t + k < t1) off[(uint64_t)(t + k) * d.nb + b] = run;
                run += v[k];
            }
        }
    }
}

// LDS (widest first, so every array is aligned):
//   sk[T] | bsep[nb] | delta[nb] | lcur[nb] | sbl[T] | btab[2^10 + 1]
// sbl[j] = (bucket << 16) | position in the tile of tile-sorted element j;
// delta[b] = first global slot of bucket b's run minus its first tile slot.
template <class K>
__global__ __launch_bounds__(kPBlock) void k_bkt_scatter(const K *__restrict__ q, uint64_t nq, BucketDev d,
                                                         const K *__restrict__ bsep_g,
                                                         const uint64_t *__restrict__ hdr,
                                                         const uint16_t *__restrict__ btab_g,
                                                         const uint32_t *__restrict__ hist,
                                                         const uint32_t *__restrict__ off,
                                                         K *__restrict__ qs, uint16_t *__restrict__ lidx) {
    extern __shared__ uint4 lds_raw[];
    __shared__ uint32_t wsum[kPBlock / 64];
    K *sk = reinterpret_cast<K *>(lds_raw);
    K *bsep = sk + d.tile;
    uint32_t *delta = reinterpret_cast<uint32_t *>(bsep + d.nb);
    uint32_t *lcur = delta + d.nb;
    uint32_t *sbl = lcur + d.nb;
    uint16_t *btab = reinterpret_cast<uint16_t *>(sbl + d.tile);
    constexpr int E = per_thread<K>();
    // persistent, one round of gridDim.x consecutive tiles at a time; in each
    // round an XCD's blocks take consecutive tiles, whose runs of one bucket
    // are adjacent in the output (their writes meet in that XCD's L2)
    const uint32_t t0 = xcd_contiguous(blockIdx.x, gridDim.x), t1 = d.ntiles;
    auto load_tile = [&](uint32_t tile, K (&key)[E]) {
        const uint64_t lo = (uint64_t)tile * d.tile;
        const uint32_t cnt = tile < t1 ? (uint32_t)(std::min<uint64_t>(lo + d.tile, nq) - lo) : 0u;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = e * kPBlock + threadIdx.x;
            key[e] = i < cnt ? __builtin_nontemporal_load(&q[lo + i]) : K(0);
        }
    };
    // thread b < nb (nb <= kMaxBuckets = kPBlock) holds bucket b's entries of
    // a tile's histogram row and offset row, loaded with the tile's queries
    uint32_t hrow = 0, orow = 0;
    auto load_rows = [&](uint32_t tile) {
        if (tile < t1 && threadIdx.x < d.nb) {
            hrow = hist[(uint64_t)tile * d.nb + threadIdx.x];  // the tile's count of bucket b
            orow = off[(uint64_t)tile * d.nb + threadIdx.x];   // first slot of its run
        }
    };
    K key[E];  // issued first: their latency overlaps the staging and the scan
    load_tile(t0, key);
    load_rows(t0);
    for (uint32_t b = threadIdx.x; b < d.nb; b += blockDim.x) bsep[b] = bsep_g[b];
    for (uint32_t p = threadIdx.x; p <= (1u << kBucketTabBits); p += blockDim.x) btab[p] = btab_g[p];
    const PrefixIndex<K> pi{(K)hdr[0], (uint32_t)hdr[1]};
    for (uint32_t tile = t0; tile < t1; tile += gridDim.x) {
        const uint64_t lo = (uint64_t)tile * d.tile;
        const uint32_t cnt = (uint32_t)(std::min<uint64_t>(lo + d.tile, nq) - lo);
        __syncthreads();  // the previous tile's write-out has read sk, sbl and delta
        if (threadIdx.x < d.nb) lcur[threadIdx.x] = hrow;
        const uint32_t orun = orow;
        __syncthreads();
        block_exclusive_scan(lcur, d.nb, wsum);  // first tile slot of every bucket
        if (threadIdx.x < d.nb) delta[threadIdx.x] = orun - lcur[threadIdx.x];
        __syncthreads();
        uint32_t bk[E];
#pragma unroll
        for (int e = 0; e < E; ++e) bk[e] = bucket_of<K>(bsep, d.nb, btab, pi, key[e]);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t i = e * kPBlock + threadIdx.x;
            if (i < cnt) {
                const uint32_t slot = atomicAdd(&lcur[bk[e]], 1u);
                sk[slot] = key[e];
                sbl[slot] = (bk[e] << 16) | i;
            }
        }
        load_tile(tile + gridDim.x, key);  // the next tile's queries and rows arrive during the write-out
        load_rows(tile + gridDim.x);
        __syncthreads();
        // tile-sorted element j goes to its bucket run: consecutive j, consecutive slots
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t j = e * kPBlock + threadIdx.x;
            if (j < cnt) {
                const uint32_t x = sbl[j];
                const uint32_t g = delta[x >> 16] + j;
                qs[g] = sk[j];
                lidx[g] = (uint16_t)(x & 0xFFFFu);
            }
        }
    }
}

// Leaf step on a sub-leaf: the lane's query `myq` lies in sub-leaf `mysub`
// (64 bytes of a leaf: slots [mysub*SUB, (mysub+1)*SUB) of the tree, SUB =
// 64 / (2 * key bytes)). Group g (lanes 4g..4g+3) takes the 4 queries of its
// lanes; lane s of the group loads 16-byte piece s of each query's sub-leaf,
// so one load instruction fetches 16 sub-leaves and a query costs 64 bytes
// instead of the whole leaf. The rank r of the query in its sub-leaf is the
// DPP sum of the lanes' counts; the lane holding slot r checks the key and
// hands the value (or the sentinel) to the query's lane.
// subleaf_issue broadcasts the group's queries and issues the loads;
// subleaf_finish ranks and answers. (Issuing the next batch's loads before
// finishing the current one measured slower: the wave's loads are not what
// bounds this step.)
constexpr int kSubLanes = 4;

template <class K>
struct SubLeafBatch {
    K qk[kSubLanes];
    uint32_t sb[kSubLanes];
    Chunk<16> c[kSubLanes];
};

template <class K, class V, int LF, bool OFF32>
__device__ __forceinline__ void subleaf_issue(const TreeView &t, K myq, uint32_t mysub, SubLeafBatch<K> &b) {
    constexpr int G = kSubLanes;
    constexpr uint32_t SUB = 64 / (2 * sizeof(K));  // slots per sub-leaf
    constexpr uint32_t SPL = LF / SUB;              // sub-leaves per leaf
    constexpr uint32_t STRIDE = 2 * LF * sizeof(K);
    static_assert(sizeof(K) == sizeof(V) && LF % SUB == 0 && SPL >= 2, "interleaved leaves of several sub-leaves");
    const uint32_t lane = threadIdx.x & 63, g = lane / G, sl = lane % G;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        b.qk[k] = __shfl(myq, (int)(g * G + k), 64);
        b.sb[k] = (uint32_t)__shfl((int)mysub, (int)(g * G + k), 64);
    }
    if constexpr (OFF32) {
        const __amdgpu_buffer_rsrc_t rs = raw_rsrc(t.leaves);
#pragma unroll
        for (int k = 0; k < G; ++k)
            load_chunk_buf<16, false>(rs, (b.sb[k] / SPL) * STRIDE + (b.sb[k] % SPL) * 64 + sl * 16, b.c[k]);
    } else {
#pragma unroll
        for (int k = 0; k < G; ++k)
            load_chunk<16>(t.leaves + (uint64_t)(b.sb[k] / SPL) * STRIDE + (b.sb[k] % SPL) * 64 + sl * 16, b.c[k]);
    }
}

template <class K, class V>
__device__ __forceinline__ V subleaf_finish(const SubLeafBatch<K> &b, uint32_t nsb, uint32_t nrem, V sentinel) {
    constexpr int G = kSubLanes;
    constexpr uint32_t SUB = 64 / (2 * sizeof(K));
    constexpr uint32_t PPL = 16 / (2 * sizeof(K));  // slots per lane
    const uint32_t lane = threadIdx.x & 63, g = lane / G, sl = lane % G;
    V res = sentinel;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        uint32_t cnt = 0;
#pragma unroll
        for (uint32_t i = 0; i < PPL; ++i) cnt += (piece_key<K>(b.c[k].w, i) < b.qk[k]);
        const uint32_t r = group_sum<G>(cnt);  // lower bound inside the sub-leaf, 0..SUB
        const uint32_t local = r - sl * PPL;   // r's slot in this lane, if < PPL
        K key = piece_key<K>(b.c[k].w, 0);
        V val = piece_val<V>(b.c[k].w, 0);
#pragma unroll
        for (uint32_t i = 1; i < PPL; ++i) {
            key = local == i ? piece_key<K>(b.c[k].w, i) : key;
            val = local == i ? piece_val<V>(b.c[k].w, i) : val;
        }
        // slot sb*SUB + r of the tree holds a key iff it is below n = nsb*SUB + nrem
        const bool valid = (b.sb[k] < nsb) | ((b.sb[k] == nsb) & (r < nrem));
        const bool hit = (local < PPL) & (key == b.qk[k]) & valid;
        const V cand = hit ? val : sentinel;
        const uint32_t owner = r / PPL < (uint32_t)G ? r / PPL : (uint32_t)G - 1;  // r = SUB: sentinel
        const V got = __shfl(cand, (int)(g * G + owner), 64);
        res = sl == (uint32_t)k ? got : res;
    }
    return res;
}

// LDS: bstart[nb + 1] | slice[epb] | stab[stab_stride] (16-byte aligned pieces)
// Work: the bucket-ordered queries in chunks of d.chunk. With d.xcd the
// chunks are split into 8 contiguous parts, one per XCD, and the blocks of an
// XCD stride through their part together, so an XCD works on about one bucket
// at a time and that bucket's leaves stay in its L2.
constexpr int kMaxSearchBlock = 1024;  // waves of a block share one staged slice
constexpr int kPieceBatches = 8;   // batches of 64 queries a wave loads at once

template <class K, class V, int F, int LF, bool COOP, int G, bool OFF32, bool SUBL>
__global__ __launch_bounds__(kMaxSearchBlock) void k_bkt_search(TreeView t, BucketDev d, const K *__restrict__ qs,
                                                        const uint32_t *__restrict__ bstart_g,
                                                        const uint16_t *__restrict__ stabs,
                                                        const uint64_t *__restrict__ spi, uint64_t nq,
                                                        V *__restrict__ res, V sentinel) {
    extern __shared__ uint4 lds_raw[];
    uint32_t *bstart = reinterpret_cast<uint32_t *>(lds_raw);
    K *slice = reinterpret_cast<K *>(lds_raw + (d.nb + 1 + 3) / 4);
    uint16_t *stab = reinterpret_cast<uint16_t *>(lds_raw + (d.nb + 1 + 3) / 4 + (d.epb * sizeof(K) + 15) / 16);
    for (uint32_t b = threadIdx.x; b <= d.nb; b += blockDim.x) bstart[b] = bstart_g[b];
    __syncthreads();
    // slice source: the sub-leaf maxima, or level s (its entries = nodes of level s-1)
    const K *level = reinterpret_cast<const K *>(t.inner + (SUBL ? t.sub_off : t.off[d.level]));
    // (sub-leaves past the last key hold no key: they are left out of the slices)
    const uint64_t level_count = SUBL ? (t.n + t.sub_slots - 1) / t.sub_slots : t.count[d.level - 1];
    // slot sb*SUB + r holds a key iff sb < nsb, or sb == nsb and r < nrem (32-bit: sub-leaf indices are)
    const uint32_t nsb = SUBL ? (uint32_t)(t.n / t.sub_slots) : 0u;
    const uint32_t nrem = SUBL ? (uint32_t)(t.n % t.sub_slots) : 0u;
    const int top = (int)d.level - 1;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint64_t nchunks = (nq + d.chunk - 1) / d.chunk;
    uint64_t c0 = 0, c1 = nchunks, cstep = gridDim.x, cfirst = blockIdx.x;
    if (d.xcd && gridDim.x % kXcds == 0) {
        const uint32_t x = blockIdx.x % kXcds;
        const uint64_t per = nchunks / kXcds, rem = nchunks % kXcds;
        c0 = x * per + (x < rem ? x : rem);
        c1 = c0 + per + (x < rem ? 1 : 0);
        cstep = gridDim.x / kXcds;
        cfirst = c0 + blockIdx.x / kXcds;
    }
    uint32_t cur = ~0u;  // bucket whose slice is staged
    uint32_t eb = 0;
    uint64_t first = 0;
    PrefixIndex<K> pi{K(0), 0};
    // Pieces are handed out by a block-wide ticket counter, so the waves finish
    // a bucket together before the next slice is staged. Tickets number the
    // pieces of the block's segments in order; seg_base is the first ticket of
    // the current segment (the same in every wave).
    __shared__ uint32_t tickets;
    if (threadIdx.x == 0) tickets = 0;
    __syncthreads();
    auto claim = [&]() -> uint32_t {
        uint32_t tk = 0;
        if (lane == 0) tk = atomicAdd(&tickets, 1u);
        return (uint32_t)__builtin_amdgcn_readfirstlane((int)tk);
    };
    uint32_t my_ticket = claim();
    uint32_t seg_base = 0;
    for (uint64_t chunk = cfirst; chunk < c1; chunk += cstep) {
        uint64_t pos = chunk * d.chunk;
        const uint64_t end = std::min<uint64_t>(pos + d.chunk, nq);
        while (pos < end) {  // the same in every thread of the block
            uint32_t lo = 0, hi = d.nb;  // last bucket b with bstart[b] <= pos
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (bstart[mid] <= pos) lo = mid; else hi = mid;
            }
            const uint32_t b = lo;
            const uint64_t seg_end = std::min<uint64_t>(end, bstart[b + 1]);
            if (b != cur) {
                __syncthreads();  // every wave is done with the previous slice
                first = (uint64_t)b * d.epb;
                eb = (uint32_t)std::min<uint64_t>(d.epb, level_count - first);
                // 16-byte copies when the slice starts on a 16-byte boundary (all but
                // odd-sized single buckets), element copies for the rest
                const uint32_t per16 = 16 / sizeof(K);
                const uint32_t nvec = (first % per16 == 0) ? eb / per16 : 0;
                const uint4 *src4 = reinterpret_cast<const uint4 *>(level + first);
                for (uint32_t e = threadIdx.x; e < nvec; e += blockDim.x) reinterpret_cast<uint4 *>(slice)[e] = src4[e];
                for (uint32_t e = nvec * per16 + threadIdx.x; e < eb; e += blockDim.x) slice[e] = level[first + e];
                // the bucket's prefix table, built once per call by k_bkt_stabs
                const uint32_t tw = stab_stride(d.sbits) / 8;  // 16-byte pieces
                const uint4 *tsrc = reinterpret_cast<const uint4 *>(stabs + (uint64_t)b * stab_stride(d.sbits));
                for (uint32_t e = threadIdx.x; e < tw; e += blockDim.x) reinterpret_cast<uint4 *>(stab)[e] = tsrc[e];
                pi = PrefixIndex<K>{(K)spi[2 * b], (uint32_t)spi[2 * b + 1]};
                __syncthreads();
                cur = b;
            }
            // A wave takes pieces of kPieceBatches batches of 64 consecutive
            // queries and loads a piece's queries with back-to-back loads
            // before its first search. A wave's loads complete in issue order,
            // so a query load issued before a batch's leaf loads is waited for
            // with them: one wait on the query stream per piece, not per batch.
            // Loads past the segment's end re-read its last query (in bounds;
            // their results are dropped).
            const uint64_t piece = (uint64_t)kPieceBatches * 64;
            const uint32_t npieces = (uint32_t)((seg_end - pos + piece - 1) / piece);
            for (; my_ticket < seg_base + npieces; my_ticket = claim()) {
                const uint64_t p0 = pos + (uint64_t)(my_ticket - seg_base) * piece;
                K qv[kPieceBatches];
#pragma unroll
                for (int j = 0; j < kPieceBatches; ++j) {
                    const uint64_t i = p0 + (uint64_t)j * 64 + lane;
                    qv[j] = __builtin_nontemporal_load(&qs[i < seg_end ? i : seg_end - 1]);
                }
                // the node (or sub-leaf) of a query: lower bound in the LDS slice
                auto locate = [&](K key) -> uint32_t {
                    const uint32_t p = prefix_of<K>(pi, key, d.sbits);
                    const uint32_t at = lower_bound_window<K>(slice, stab[p], stab[p + 1], key);
                    return (uint32_t)first + (at < eb ? at : eb - 1);
                };
                V rv[kPieceBatches];
                if constexpr (SUBL) {
                    // Every batch of the piece runs (past the segment's end on its
                    // last query, in bounds); only results inside it are stored.
                    // Leaves are reused by every query of the bucket: regular
                    // (temporal) loads.
#pragma unroll
                    for (int j = 0; j < kPieceBatches; ++j) {
                        SubLeafBatch<K> bt;
                        // raised priority from the slice search to the issued
                        // loads, the batch's critical path; the ranking is not
                        __builtin_amdgcn_s_setprio(2);
                        const uint32_t sub = locate(qv[j]);
                        subleaf_issue<K, V, LF, OFF32>(t, qv[j], sub, bt);
                        __builtin_amdgcn_s_setprio(0);
                        rv[j] = subleaf_finish<K, V>(bt, nsb, nrem, sentinel);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < kPieceBatches; ++j) {
                        rv[j] = sentinel;
                        if (p0 + (uint64_t)j * 64 < seg_end) {  // the same in every lane of the wave
                            const K key = qv[j];
                            const uint32_t node = locate(key);
                            // leaves are reused by every query of the bucket: regular (temporal) loads
                            if constexpr (COOP)
                                rv[j] = coop_finish64<K, V, F, LF, false, kOpValue, G, OFF32>(t, key, node, top,
                                                                                              sentinel);
                            else rv[j] = thread_finish<K, V, F, LF>(t, key, node, top, sentinel);
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < kPieceBatches; ++j) {
                    const uint64_t i = p0 + (uint64_t)j * 64 + lane;
                    if (i < seg_end) __builtin_nontemporal_store(rv[j], &res[i]);
                }
            }
            seg_base += npieces;
            pos = seg_end;
        }
    }
}

// LDS: delta[nb] | lstart[nb + 1] | outt[T] (V), whose first 2T bytes hold
// bid[T] (u16, bucket of every tile-sorted position) before the results
// arrive. Each wave paints whole buckets' runs of bid with its 64 lanes, so the
// cost follows the run lengths, not a search per element.
template <class K, class V>
__global__ __launch_bounds__(kPBlock) void k_bkt_unpermute(const uint32_t *__restrict__ hist,
                                                           const uint32_t *__restrict__ off, BucketDev d,
                                                           const uint16_t *__restrict__ lidx,
                                                           const V *__restrict__ res, uint64_t nq,
                                                           V *__restrict__ out) {
    extern __shared__ uint4 lds_raw[];
    __shared__ uint32_t wsum[kPBlock / 64];
    uint32_t *delta = reinterpret_cast<uint32_t *>(lds_raw);
    uint32_t *lstart = delta + d.nb;
    V *outt = reinterpret_cast<V *>(lds_raw + (2 * d.nb + 1 + 3) / 4);
    uint16_t *bid = reinterpret_cast<uint16_t *>(outt);
    const uint32_t tile = xcd_contiguous(blockIdx.x, gridDim.x);
    // thread b < nb (nb <= kMaxBuckets = kPBlock) loads bucket b's count and run
    // offset of the tile together
    const uint32_t hrow = threadIdx.x < d.nb ? hist[(uint64_t)tile * d.nb + threadIdx.x] : 0u;
    const uint32_t orow = threadIdx.x < d.nb ? off[(uint64_t)tile * d.nb + threadIdx.x] : 0u;
    if (threadIdx.x < d.nb) lstart[threadIdx.x] = hrow;
    __syncthreads();
    const uint32_t total = block_exclusive_scan(lstart, d.nb, wsum);
    if (threadIdx.x < d.nb) delta[threadIdx.x] = orow - lstart[threadIdx.x];
    if (threadIdx.x == 0) lstart[d.nb] = total;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (uint32_t b = w; b < d.nb; b += nw)
        for (uint32_t j = lstart[b] + lane; j < lstart[b + 1]; j += 64) bid[j] = (uint16_t)b;
    __syncthreads();
    constexpr int E = per_thread<K>();
    const uint64_t lo = (uint64_t)tile * d.tile;
    const uint32_t cnt = (uint32_t)(std::min<uint64_t>(lo + d.tile, nq) - lo);
    uint32_t g[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t j = e * kPBlock + threadIdx.x;
        g[e] = j < cnt ? delta[bid[j]] + j : 0u;
    }
    __syncthreads();  // bid is dead: outt reuses its bytes
    uint16_t li[E];
    V rv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const bool ok = e * kPBlock + threadIdx.x < cnt;
        li[e] = ok ? lidx[g[e]] : (uint16_t)0;
        rv[e] = ok ? __builtin_nontemporal_load(&res[g[e]]) : V(0);
    }
#pragma unroll
    for (int e = 0; e < E; ++e)
        if (e * kPBlock + threadIdx.x < cnt) outt[li[e]] = rv[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = e * kPBlock + threadIdx.x;
        if (i < cnt) __builtin_nontemporal_store(outt[i], &out[lo + i]);
    }
}

// ------------------------------------------------------------------ host side
// The bucket tables of a plan (they depend on the tree and the plan's
// nb, level, lpb, epb, sbits and spl, not on the queries): bucket
// separators, their prefix table (hdr, btab) and every bucket's slice prefix
// table (stabs, spi), at these offsets of one block.
struct TableLayout {
    uint64_t bsep, hdr, btab, stabs, spi, total;
};

static TableLayout table_layout(const BucketDev &d, uint32_t key_bytes) {
    TableLayout L{};
    uint64_t off = 0;
    L.bsep = off;   off += align256((uint64_t)d.nb * key_bytes);
    L.hdr = off;    off += align256(16);
    L.btab = off;   off += align256(((1ull << kBucketTabBits) + 1) * 2);
    L.stabs = off;  off += align256((uint64_t)d.nb * stab_stride(d.sbits) * 2);
    L.spi = off;    off += align256((uint64_t)d.nb * 16);
    L.total = off;
    return L;
}

static void table_sig(const BucketDev &d, uint64_t (&sig)[6]) {
    sig[0] = d.nb;
    sig[1] = d.level;
    sig[2] = d.lpb;
    sig[3] = d.epb;
    sig[4] = d.sbits;
    sig[5] = d.spl;
}

struct BucketPlan {
    BucketDev d;
    uint32_t grid;                 // search blocks
    uint64_t hist_entries;
    uint64_t off_qs, off_lidx, off_res, off_hist, off_off, off_cols, off_bstart, off_tables, total;
    size_t lds_hist, lds_scatter, lds_search, lds_unpermute;
};

static BucketPlan bucket_plan(const bpt_tree *t, uint64_t nq) {
    BucketPlan p{};
    const TreeView &v = t->view;
    const uint64_t L = v.count[0], F = t->fanout;
    const uint64_t slice_cap = std::min<uint32_t>(t->group_slice, kMaxSliceBytes) / t->key_bytes;
    const uint64_t nb_cap = std::min<uint64_t>(1ull << t->group_bits, kMaxBuckets);
    BucketDev &d = p.d;
    const uint64_t spl = v.sub_slots ? t->leaf_cap / v.sub_slots : 0;  // sub-leaves per leaf
    if (spl && t->group_sub && t->coop && L * spl < (1ull << 32) &&
        (L + slice_cap / spl - 1) / (slice_cap / spl) <= nb_cap) {
        // slice of sub-leaf maxima: the search reads 64 bytes of a leaf
        d.level = 1;
        d.spl = (uint32_t)spl;
        d.lpb = std::min(slice_cap / spl, L);
        d.nb = (uint32_t)((L + d.lpb - 1) / d.lpb);
        d.epb = (uint32_t)(d.lpb * spl);
    } else {
        uint64_t span = slice_cap;  // leaves per bucket at level s = slice_cap * F^(s-1)
        uint32_t s = 1;
        for (;; ++s) {
            if (span >= L || (L + span - 1) / span <= nb_cap || s == (uint32_t)v.height) break;
            span *= F;
        }
        d.level = s;
        d.spl = 0;
        d.lpb = std::min(span, L);
        d.nb = (uint32_t)((L + d.lpb - 1) / d.lpb);
        uint64_t per = 1;  // leaves per entry of level s
        for (uint32_t k = 1; k < s; ++k) per *= F;
        d.epb = (uint32_t)((d.lpb + per - 1) / per);
    }
    // prefix table of about epb / 4 entries: windows of ~4 slice entries
    d.sbits = 4;
    while (d.sbits < kMaxSliceTabBits && (4ull << d.sbits) < d.epb) ++d.sbits;
    d.tile = t->key_bytes == 4 ? tile_of<uint32_t>() : tile_of<uint64_t>();
    d.ntiles = (uint32_t)std::max<uint64_t>(1, (nq + d.tile - 1) / d.tile);
    d.ngroups = (d.ntiles + kColGroup - 1) / kColGroup;
    // search blocks per CU: as asked, at most as many as the CU's LDS holds
    const size_t slice_lds = ((d.nb + 1 + 3) / 4) * 16 + ((size_t)d.epb * t->key_bytes + 15) / 16 * 16 +
                             (size_t)stab_stride(d.sbits) * 2;
    const uint64_t per_cu = std::max<uint64_t>(1, std::min<uint64_t>(t->group_blocks, kMaxLds / slice_lds));
    const uint64_t cap = (uint64_t)device_cu_count() * per_cu;
    if (t->group_chunk) {
        d.chunk = t->group_chunk;
    } else {
        // about 16 chunks per search block: long enough runs of one bucket per
        // slice staging, enough chunks for every block (32768 at 2^28 queries)
        uint64_t c = 1024;
        while (c < 32768 && c * 2 * cap * 16 <= nq) c *= 2;
        d.chunk = (uint32_t)c;
    }
    const uint64_t nchunks = std::max<uint64_t>(1, (nq + d.chunk - 1) / d.chunk);
    p.grid = (uint32_t)std::min(cap, nchunks);
    // split per XCD only when every XCD gets whole strides of chunks
    d.xcd = (t->group_xcd && p.grid % kXcds == 0 && nchunks >= 2ull * p.grid) ? 1u : 0u;
    p.hist_entries = (uint64_t)d.nb * d.ntiles;
    uint64_t off = 0;
    p.off_qs = off;     off += align256(nq * t->key_bytes);
    p.off_lidx = off;   off += align256(nq * 2);
    p.off_res = off;    off += align256(nq * t->val_bytes);
    p.off_hist = off;   off += align256(p.hist_entries * 4);
    p.off_off = off;    off += align256(p.hist_entries * 4);
    p.off_cols = off;   off += align256((uint64_t)d.ngroups * d.nb * 4);
    p.off_bstart = off; off += align256(((uint64_t)d.nb + 1) * 4);
    p.off_tables = off; off += table_layout(d, t->key_bytes).total;  // used when the tree's are not
    p.total = off;
    const size_t kb = t->key_bytes, vb = t->val_bytes;
    const size_t btab_bytes = ((1u << kBucketTabBits) + 1) * 2;
    p.lds_hist = d.nb * (kb + 4) + btab_bytes;
    p.lds_scatter = d.nb * (kb + 8) + (size_t)d.tile * (kb + 4) + btab_bytes;
    p.lds_search = slice_lds;
    p.lds_unpermute = ((2 * d.nb + 1 + 3) / 4) * 16 + (size_t)d.tile * vb;
    return p;
}

// Batches below t->group_min run the plain descent (same results): the
// partition's fixed cost (seven launches) and the slice staging are repaid
// only by large batches. On MI355X the grouped form overtakes the plain
// descent at about 2^24 queries on 16M-64M-key trees (2^22 on 4M keys, 2^25 on
// trees whose leaves fit the L2).
static bool grouped_runs_plain(const bpt_tree *t, uint64_t nq) {
    return t->n == 0 || t->view.height == 0 || nq < t->group_min;
}

uint64_t point_grouped_ws(const bpt_tree *t, uint64_t nq) {
    if (grouped_runs_plain(t, nq)) return 0;
    return bucket_plan(t, std::min(nq, kBucketChunk)).total;
}

// Builds the bucket tables of plan d for tree t into the block at tb.
template <class K>
static bpt_status fill_tables(const bpt_tree *t, const BucketDev &d, uint8_t *tb, hipStream_t s) {
    const TableLayout L = table_layout(d, t->key_bytes);
    K *bsep = reinterpret_cast<K *>(tb + L.bsep);
    k_bkt_seps<K><<<(d.nb + 255) / 256, 256, 0, s>>>(t->view, d, bsep);
    BPT_HIP(hipGetLastError());
    k_bkt_btab<K><<<((1u << kBucketTabBits) + 256) / 256, kBBlock, 0, s>>>(
        t->view, d, bsep, reinterpret_cast<uint64_t *>(tb + L.hdr), reinterpret_cast<uint16_t *>(tb + L.btab));
    BPT_HIP(hipGetLastError());
    // slice prefix tables: the slice source as the search kernel reads it
    const TreeView &v = t->view;
    const bool subl = d.spl != 0;
    const K *level = reinterpret_cast<const K *>(v.inner + (subl ? v.sub_off : v.off[d.level]));
    const uint64_t level_count = subl ? (v.n + v.sub_slots - 1) / v.sub_slots : v.count[d.level - 1];
    k_bkt_stabs<K><<<d.nb, kPBlock, 0, s>>>(d, level, level_count, reinterpret_cast<uint16_t *>(tb + L.stabs),
                                            reinterpret_cast<uint64_t *>(tb + L.spi));
    BPT_HIP(hipGetLastError());
    return BPT_OK;
}

template <class K, class V, int F, int LF>
static bpt_status bucket_chunk(const bpt_tree *t, const K *q, uint64_t nq, V *out, uint8_t *ws, hipStream_t s) {
    const BucketPlan p = bucket_plan(t, nq);
    const BucketDev &d = p.d;
    K *qs = reinterpret_cast<K *>(ws + p.off_qs);
    uint16_t *lidx = reinterpret_cast<uint16_t *>(ws + p.off_lidx);
    V *res = reinterpret_cast<V *>(ws + p.off_res);
    uint32_t *hist = reinterpret_cast<uint32_t *>(ws + p.off_hist);
    uint32_t *offs = reinterpret_cast<uint32_t *>(ws + p.off_off);
    uint32_t *cols = reinterpret_cast<uint32_t *>(ws + p.off_cols);
    uint32_t *bstart = reinterpret_cast<uint32_t *>(ws + p.off_bstart);
    BPT_CHECK(p.lds_scatter <= kMaxLds && p.lds_search <= kMaxLds && p.lds_unpermute <= kMaxLds, BPT_ERR_INVALID,
              "bucket plan exceeds LDS");
    // the tree's own tables when they were built for this plan, else this call's
    uint64_t sig[6];
    table_sig(d, sig);
    const bool own = t->bkt_valid && std::equal(sig, sig + 6, t->bkt_sig);
    uint8_t *tb = own ? t->d_bkt : ws + p.off_tables;
    if (!own) BPT_TRY(fill_tables<K>(t, d, tb, s));
    const TableLayout L = table_layout(d, t->key_bytes);
    const K *bsep = reinterpret_cast<const K *>(tb + L.bsep);
    const uint64_t *hdr = reinterpret_cast<const uint64_t *>(tb + L.hdr);
    const uint16_t *btab = reinterpret_cast<const uint16_t *>(tb + L.btab);
    const uint16_t *stabs = reinterpret_cast<const uint16_t *>(tb + L.stabs);
    const uint64_t *spi = reinterpret_cast<const uint64_t *>(tb + L.spi);
    k_bkt_hist<K><<<d.ntiles, kPBlock, p.lds_hist, s>>>(q, nq, d, bsep, hdr, btab, hist);
    BPT_HIP(hipGetLastError());
    k_bkt_colsum<<<d.ngroups, kBBlock, 0, s>>>(hist, d, cols);
    BPT_HIP(hipGetLastError());
    k_bkt_colscan<<<1, kScanThreads, 0, s>>>(cols, d, bstart);
    BPT_HIP(hipGetLastError());
    k_bkt_colprefix<<<d.ngroups, kBBlock, 0, s>>>(hist, cols, bstart, d, offs);
    BPT_HIP(hipGetLastError());
    k_bkt_scatter<K><<<std::min<uint32_t>(d.ntiles, (uint32_t)device_cu_count()), kPBlock, p.lds_scatter, s>>>(
        q, nq, d, bsep, hdr, btab, hist, offs, qs, lidx);
    BPT_HIP(hipGetLastError());
    const V sentinel = ~V(0);
    bool coop = false;
    if constexpr (sizeof(K) == sizeof(V)) coop = t->coop && t->view.count[0] < (1ull << 32);
    const bool off32 = t->addr32 && t->leaf_bytes <= (1ull << 32) && t->inner_bytes <= (1ull << 32);
    auto go = [&](auto cp, auto lanes, auto o32, auto subl) {
        k_bkt_search<K, V, F, LF, decltype(cp)::value, decltype(lanes)::value, decltype(o32)::value,
                     decltype(subl)::value>
            <<<p.grid, t->group_threads, p.lds_search, s>>>(t->view, d, qs, bstart, stabs, spi, nq, res, sentinel);
    };
    auto go_leaf = [&](auto cp, auto lanes, auto o32) { go(cp, lanes, o32, std::false_type{}); };
    if constexpr (sizeof(K) == sizeof(V) && LF * sizeof(K) > 32) {  // leaves of several sub-leaves
        if (d.spl) {
            if (off32) go(std::true_type{}, IC<4>{}, std::true_type{}, std::true_type{});
            else go(std::true_type{}, IC<4>{}, std::false_type{}, std::true_type{});
            BPT_HIP(hipGetLastError());
            k_bkt_unpermute<K, V><<<d.ntiles, kPBlock, p.lds_unpermute, s>>>(hist, offs, d, lidx, res, nq, out);
            BPT_HIP(hipGetLastError());
            return BPT_OK;
        }
    }
    BPT_CHECK(d.spl == 0, BPT_ERR_INVALID, "sub-leaf slice planned for a tree without sub-leaves");
    if constexpr (sizeof(K) == sizeof(V)) {
        if (coop) {
            if (t->group_lanes == 4) {
                if (off32) go_leaf(std::true_type{}, IC<4>{}, std::true_type{});
                else go_leaf(std::true_type{}, IC<4>{}, std::false_type{});
            } else {
                if (off32) go_leaf(std::true_type{}, IC<8>{}, std::true_type{});
                else go_leaf(std::true_type{}, IC<8>{}, std::false_type{});
            }
        } else {
            go_leaf(std::false_type{}, IC<8>{}, std::false_type{});
        }
    } else {
        go_leaf(std::false_type{}, IC<8>{}, std::false_type{});
    }
    BPT_HIP(hipGetLastError());
    k_bkt_unpermute<K, V><<<d.ntiles, kPBlock, p.lds_unpermute, s>>>(hist, offs, d, lidx, res, nq, out);
    BPT_HIP(hipGetLastError());
    return BPT_OK;
}

bpt_status bucket_tables_refresh(bpt_tree *t, hipStream_t s) {
    t->bkt_valid = false;
    if (t->n == 0 || t->view.height == 0) return BPT_OK;
    const BucketPlan p = bucket_plan(t, 1);  // the tables do not depend on the batch
    const TableLayout L = table_layout(p.d, t->key_bytes);
    if (L.total > t->bkt_alloc) {  // with a quarter to spare, so that growing updates rarely reallocate
        if (t->d_bkt) (void)hipFree(t->d_bkt);
        t->d_bkt = nullptr;
        t->bkt_alloc = 0;
        const uint64_t want = align256(L.total + L.total / 4);
        if (t->alloc_limit && want > t->alloc_limit) return BPT_OK;
        if (hipMalloc(&t->d_bkt, want) != hipSuccess) {
            (void)hipGetLastError();  // no tables of its own: searches build them per call
            t->d_bkt = nullptr;
            return BPT_OK;
        }
        t->bkt_alloc = want;
    }
    BPT_TRY(dispatch_tree(t, [&](auto kt, auto, auto, auto) -> bpt_status {
        return fill_tables<typename decltype(kt)::type>(t, p.d, t->d_bkt, s);
    }));
    table_sig(p.d, t->bkt_sig);
    t->bkt_valid = true;
    return BPT_OK;
}

bpt_status point_grouped(const bpt_tree *t, const void *d_q, uint64_t nq, void *d_out, void *d_ws,
                         uint64_t ws_bytes, hipStream_t s) {
    if (nq == 0) return BPT_OK;
    if (grouped_runs_plain(t, nq)) return launch_plain(t, d_q, nq, d_out, s);
    BPT_CHECK(d_ws && ws_bytes >= point_grouped_ws(t, nq), BPT_ERR_INVALID,
              "grouped search needs a workspace of %llu bytes",
              (unsigned long long)point_grouped_ws(t, nq));
    return dispatch_tree(t, [&](auto kt, auto vt, auto fc, auto lc) -> bpt_status {
        using K = typename decltype(kt)::type;
        using V = typename decltype(vt)::type;
        constexpr int F = decltype(fc)::value, LF = decltype(lc)::value;
        for (uint64_t b = 0; b < nq; b += kBucketChunk) {
            const uint64_t m = std::min(kBucketChunk, nq - b);
            if (grouped_runs_plain(t, m))  // a short last piece
                BPT_TRY(launch_plain(t, (const K *)d_q + b, m, (V *)d_out + b, s));
            else
                BPT_TRY((bucket_chunk<K, V, F, LF>(t, (const K *)d_q + b, m, (V *)d_out + b, (uint8_t *)d_ws, s)));
        }
        return BPT_OK;
    });
}

}  // namespace bpt
