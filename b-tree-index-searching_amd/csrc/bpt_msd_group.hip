// The reference's MSD grouping pass loop, for the reference-named call
// surface (genGrouping, genGroupingWithOnedata, genGroupingWithTwodata,
// genGrouping2d), whose group lists callers read.
//
// The reference (grouping.cu:776-1363,1365-1851; grouping2d.cu:349-743)
// runs, per pass, a device histogram of every group still to split, copies
// it to the host, relocates each group's elements to their bucket's run, and
// merges adjacent buckets on the host below the group cap. Here, per pass:
//  - one histogram launch covers every group to split (chunks of 16384
//    elements, LDS counts, one global add per bucket and chunk);
//  - the host reads the counts and runs the bookkeeping of the reference's
//    two forms (fewer than 10 groups: grouping.cu:776-884; otherwise the
//    single-kernel form :1288-1363, whose merge may join neighbouring
//    groups), which is the only host work;
//  - the split groups are reordered by bucket with one stable on-device
//    radix sort of (group, bucket) keys, which moves every payload array
//    with its key; groups not split keep their order.
// The order inside a final group is this sort's (stable), where the
// reference's atomics leave it unspecified.
#include <algorithm>
#include <vector>

#include "bpt_internal.hpp"

namespace bpt {

namespace {

constexpr uint64_t kHistChunk = 16384;

struct MsdChunk {
    uint64_t begin, end;
    uint64_t slot;
};

// Bucket of element i at `pass`: kind 0 grouping.cu:25-29 (8-bit MSD digit),
// kind 1 grouping2d.cu:24-31 (product of the 4-bit nibbles of a and b at a
// 3-bit stride). A shift outside [0, 64) gives 0, as the reference's 64-bit
// device shift does.
__device__ __forceinline__ uint32_t msd_bucket(const uint64_t *__restrict__ a, const uint64_t *__restrict__ b,
                                               uint64_t i, int kind, int pass) {
    if (kind == 0) {
        const int shift = 8 * (7 - pass);
        return (shift < 0 || shift >= 64) ? 0u : (uint32_t)((a[i] >> shift) & 0xFF);
    }
    const int shift = 3 * (21 - pass);
    if (shift < 0 || shift >= 64) return 0u;
    return (uint32_t)(((a[i] >> shift) & 0xF) * ((b[i] >> shift) & 0xF));
}

__global__ __launch_bounds__(256) void k_msd_hist(const uint64_t *__restrict__ a, const uint64_t *__restrict__ b,
                                                  int kind, int pass, const MsdChunk *__restrict__ chunks,
                                                  unsigned long long *__restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const MsdChunk c = chunks[blockIdx.x];
    for (uint64_t i = c.begin + threadIdx.x; i < c.end; i += 256) atomicAdd(&h[msd_bucket(a, b, i, kind, pass)], 1u);
    __syncthreads();
    const uint32_t v = h[threadIdx.x];
    if (v) atomicAdd(&hist[c.slot * 256 + threadIdx.x], (unsigned long long)v);
}

// comp[i] = (group of position i) << 8 | (its bucket if that group is split).
// The group is the last one whose start is <= i (empty groups never win:
// the next group shares their start).
__global__ __launch_bounds__(256) void k_msd_composite(const uint64_t *__restrict__ a, const uint64_t *__restrict__ b,
                                                       int kind, int pass, uint64_t n,
                                                       const uint64_t *__restrict__ gstart, uint32_t ng,
                                                       const uint8_t *__restrict__ split, uint64_t *__restrict__ comp) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        uint32_t lo = 0, hi = ng;  // first group starting after i
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (gstart[mid] <= i) lo = mid + 1; else hi = mid;
        }
        const uint32_t g = lo - 1;
        comp[i] = ((uint64_t)g << 8) | (split[g] ? msd_bucket(a, b, i, kind, pass) : 0u);
    }
}

__global__ __launch_bounds__(256) void k_msd_gather(const uint64_t *__restrict__ src, const uint32_t *__restrict__ idx,
                                                    uint64_t n, uint64_t *__restrict__ dst) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) dst[i] = src[idx[i]];
}

// Greedy merge of grouping.cu:855-872 / :1333-1352 over counts[0, nb): from
// each non-empty count, absorb the following ones while the running total
// stays below the cap. Appends (size, index of the first count).
void merge_counts(const uint64_t *counts, uint64_t nb, uint64_t cap, std::vector<uint64_t> &size,
                  std::vector<uint64_t> &first) {
    for (uint64_t j = 0; j < nb; ++j) {
        if (counts[j] == 0) continue;
        uint64_t cur = counts[j], k = j;
        while (cur < cap && k + 1 < nb && cur + counts[k + 1] < cap) cur += counts[++k];
        size.push_back(cur);
        first.push_back(j);
        j = k;
    }
}

// Device scratch freed on scope exit; alloc() replaces what it held.
struct DeviceBuffer {
    void *p = nullptr;
    ~DeviceBuffer() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    bpt_status alloc(uint64_t bytes) {
        release();
        BPT_HIP(hipMalloc(&p, std::max<uint64_t>(bytes, 256)));
        return BPT_OK;
    }
};

}  // namespace

bpt_status msd_group(uint64_t *const *arrays, int narrays, int kind, uint64_t n, uint64_t cap, int start_pass,
                     std::vector<uint64_t> &list, std::vector<uint64_t> &prefix, hipStream_t s) {
    BPT_CHECK(narrays >= 1 + kind && narrays <= 4 && (kind == 0 || kind == 1), BPT_ERR_INVALID,
              "msd_group: %d arrays for kind %d", narrays, kind);
    BPT_CHECK(n < (1ull << 32), BPT_ERR_CAPACITY, "grouping of more than 2^32 elements");
    list.assign(1, n);
    prefix.assign(1, 0);
    const int end_pass = kind == 0 ? 8 : 16;
    if (n == 0 && start_pass >= end_pass) return BPT_OK;
    const uint64_t *a = arrays[0];
    const uint64_t *b = kind == 1 ? arrays[1] : arrays[0];
    const uint64_t sort_bytes = sort_ws_bytes(n, 8);
    DeviceBuffer comp, tmp, sort_ws, hist, chunks, gstart, split;
    BPT_TRY(comp.alloc(n * 8));
    BPT_TRY(tmp.alloc(n * 8));
    BPT_TRY(sort_ws.alloc(sort_bytes));
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>((n + 255) / 256, (uint64_t)device_cu_count() * 16));
    std::vector<uint64_t> counts, nl, np, msize, mfirst;
    std::vector<MsdChunk> ch;
    std::vector<uint8_t> split_h;
    std::vector<uint64_t> cand;  // groups histogrammed this pass
    uint64_t hist_cap = 0, chunk_cap = 0, group_cap = 0;
    for (int pass = start_pass; pass < end_pass; ++pass) {
        const uint64_t ng = list.size();
        const bool few = ng < 10;
        // groups to histogram: below 10 groups every non-empty group not below
        // the cap (grouping.cu:783-790); otherwise every group above the cap
        // (the kernel's test, grouping.cu:703)
        cand.clear();
        for (uint64_t i = 0; i < ng; ++i)
            if (few ? (list[i] != 0 && list[i] >= cap) : list[i] > cap) cand.push_back(i);
        counts.assign(cand.size() * 256, 0);
        if (!cand.empty()) {
            ch.clear();
            for (uint64_t c = 0; c < cand.size(); ++c) {
                const uint64_t g = cand[c];
                for (uint64_t x = prefix[g]; x < prefix[g] + list[g]; x += kHistChunk)
                    ch.push_back({x, std::min(prefix[g] + list[g], x + kHistChunk), c});
            }
            if (cand.size() * 256 > hist_cap) {
                hist_cap = std::max<uint64_t>(cand.size() * 256, 4096);
                BPT_TRY(hist.alloc(hist_cap * 8));
            }
            if (ch.size() > chunk_cap) {
                chunk_cap = std::max<uint64_t>(ch.size(), 1024);
                BPT_TRY(chunks.alloc(chunk_cap * sizeof(MsdChunk)));
            }
            BPT_HIP(hipMemcpyAsync(chunks.p, ch.data(), ch.size() * sizeof(MsdChunk), hipMemcpyHostToDevice, s));
            BPT_HIP(hipMemsetAsync(hist.p, 0, cand.size() * 256 * 8, s));
            BPT_CHECK(ch.size() < (1ull << 31), BPT_ERR_CAPACITY, "too many histogram chunks");
            k_msd_hist<<<(uint32_t)ch.size(), 256, 0, s>>>(a, b, kind, pass, static_cast<const MsdChunk *>(chunks.p),
                                                           static_cast<unsigned long long *>(hist.p));
            BPT_HIP(hipGetLastError());
            BPT_HIP(hipMemcpyAsync(counts.data(), hist.p, cand.size() * 256 * 8, hipMemcpyDeviceToHost, s));
            BPT_HIP(hipStreamSynchronize(s));
        }
        // bookkeeping of the two forms; split_h marks the groups whose
        // elements move to their bucket runs
        split_h.assign(ng, 0);
        nl.clear();
        np.clear();
        if (few) {
            uint64_t c = 0;
            for (uint64_t i = 0; i < ng; ++i) {
                if (list[i] == 0) continue;
                if (list[i] < cap) {
                    nl.push_back(list[i]);
                    np.push_back(prefix[i]);
                    continue;
                }
                const uint64_t *h = &counts[256 * c++];
                int valid = 0;
                for (int j = 0; j < 256; ++j) valid += h[j] != 0;
                if (valid == 1 && ng == 1) {  // grouping.cu:840-844: stays whole, nothing moves
                    nl.push_back(list[i]);
                    np.push_back(prefix[i]);
                    continue;
                }
                split_h[i] = 1;
                uint64_t start[256], acc = 0;
                for (int j = 0; j < 256; ++j) { start[j] = acc; acc += h[j]; }
                msize.clear();
                mfirst.clear();
                merge_counts(h, 256, cap, msize, mfirst);
                for (size_t t = 0; t < msize.size(); ++t) {
                    nl.push_back(msize[t]);
                    np.push_back(prefix[i] + start[mfirst[t]]);
                }
            }
        } else {
            // one count per bucket of a split group, one per other group in
            // its first slot; one merge over all slots, prefixes run on
            std::vector<uint64_t> slots(256 * ng, 0);
            uint64_t c = 0;
            for (uint64_t i = 0; i < ng; ++i) {
                if (list[i] > cap) {
                    std::copy(&counts[256 * c], &counts[256 * c] + 256, &slots[256 * i]);
                    ++c;
                    split_h[i] = 1;
                } else {
                    slots[256 * i] = list[i];
                }
            }
            msize.clear();
            mfirst.clear();
            merge_counts(slots.data(), slots.size(), cap, msize, mfirst);
            uint64_t run = 0;
            for (uint64_t v : msize) {
                nl.push_back(v);
                np.push_back(run);
                run += v;
            }
        }
        // move every split group's elements to their bucket runs
        if (std::find(split_h.begin(), split_h.end(), 1) != split_h.end()) {
            if (ng > group_cap) {
                group_cap = std::max<uint64_t>(ng, 1024);
                BPT_TRY(gstart.alloc(group_cap * 8));
                BPT_TRY(split.alloc(group_cap));
            }
            BPT_HIP(hipMemcpyAsync(gstart.p, prefix.data(), ng * 8, hipMemcpyHostToDevice, s));
            BPT_HIP(hipMemcpyAsync(split.p, split_h.data(), ng, hipMemcpyHostToDevice, s));
            k_msd_composite<<<grid, 256, 0, s>>>(a, b, kind, pass, n, static_cast<const uint64_t *>(gstart.p),
                                                 (uint32_t)ng, static_cast<const uint8_t *>(split.p),
                                                 static_cast<uint64_t *>(comp.p));
            BPT_HIP(hipGetLastError());
            uint32_t gbits = 0;
            while ((1ull << gbits) < ng) ++gbits;
            uint64_t *ks = nullptr;
            uint32_t *perm = nullptr;
            BPT_TRY(sort_pairs<uint64_t>(static_cast<const uint64_t *>(comp.p), nullptr, n, 0, 8 + gbits, sort_ws.p,
                                         sort_bytes, &ks, &perm, s));
            for (int k = 0; k < narrays; ++k) {
                k_msd_gather<<<grid, 256, 0, s>>>(arrays[k], perm, n, static_cast<uint64_t *>(tmp.p));
                BPT_HIP(hipGetLastError());
                BPT_HIP(hipMemcpyAsync(arrays[k], tmp.p, n * 8, hipMemcpyDeviceToDevice, s));
            }
        }
        list.swap(nl);
        prefix.swap(np);
        // an empty list ends the loop (n == 0; the reference reads
        // groupList[0] of the empty list there)
        if (list.empty()) break;
        if (*std::max_element(list.begin(), list.end()) <= cap) break;
    }
    BPT_HIP(hipStreamSynchronize(s));
    return BPT_OK;
}

}  // namespace bpt
