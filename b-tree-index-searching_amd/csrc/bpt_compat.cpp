// The reference's C++ call surface (BPTree.h, grouping.h, grouping2d.h,
// zOrder2d.h of sc25index/B-tree-index-searching) over the C ABI.
//
// Searches and partitions run on the device through bpt_* calls. What stays
// on the host is what the reference itself keeps on the host: copying the
// caller's arrays, and the group-list bookkeeping of its grouping loop (the
// greedy merge of adjacent buckets below the group cap).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bptree.h"
#include "compat/BPTree.h"
#include "compat/grouping.h"
#include "compat/grouping2d.h"
#include "compat/zOrder2d.h"

namespace {

bool ok(bpt_status st, const char *what) {
    if (st == BPT_OK) return true;
    fprintf(stderr, "%s: %s\n", what, bpt_last_error());
    return false;
}

// Device copy of a host array; nullptr on failure.
template <class T>
T *to_device(const T *h, ulli n) {
    void *d = nullptr;
    if (!ok(bpt_device_alloc(n * sizeof(T), &d), "device alloc")) return nullptr;
    if (n && !ok(bpt_copy_h2d(d, h, n * sizeof(T), nullptr), "copy to device")) {
        bpt_device_free(d);
        return nullptr;
    }
    return static_cast<T *>(d);
}

template <class T>
bool to_host(T *h, const T *d, ulli n) {
    return n == 0 || ok(bpt_copy_d2h(h, d, n * sizeof(T), nullptr), "copy to host");
}

ulli *point_search(ulli *input_key, ulli n, BTPage *tree, int mode) {
    if (!tree) return nullptr;
    ulli *res = new ulli[n ? n : 1];
    ulli *dq = to_device(input_key, n);
    void *dout = nullptr;
    bool good = dq && ok(bpt_device_alloc(n * 8, &dout), "device alloc") &&
                ok(bpt_point_search(tree, dq, n, dout, mode, nullptr), "point search") &&
                to_host(res, static_cast<const ulli *>(dout), n);
    bpt_device_free(dq);
    bpt_device_free(dout);
    if (!good) {
        delete[] res;
        return nullptr;
    }
    return res;
}

ulli *range_counts(ulli *lo, ulli *hi, ulli n, BTPage *tree, int mode) {
    if (!tree) return nullptr;
    ulli *res = new ulli[n ? n : 1];
    ulli *dlo = to_device(lo, n), *dhi = to_device(hi, n);
    void *dc = nullptr;
    bool good = dlo && dhi && ok(bpt_device_alloc(n * 8, &dc), "device alloc") &&
                ok(bpt_range_count_grouped(tree, dlo, dhi, n, mode, nullptr, static_cast<uint64_t *>(dc), nullptr),
                   "range count") &&
                to_host(res, static_cast<const ulli *>(dc), n);
    bpt_device_free(dlo);
    bpt_device_free(dhi);
    bpt_device_free(dc);
    if (!good) {
        delete[] res;
        return nullptr;
    }
    return res;
}

int bit_width(ulli v) {
    int b = 0;
    while (v) { v >>= 1; ++b; }
    return b;
}

// The reference's grouping loop over passes [p0, pend) (grouping.cu:1746-1791,
// grouping2d.cu:635-680), run on the bucket counts of the grouped order:
// fewer than 10 groups -> per-group split of groups >= cap with the merge of
// grouping.cu:855-872 (a single group with one non-empty bucket stays whole);
// otherwise -> groups > cap split, small groups passed through as one entry,
// and one merge over all groups' buckets (grouping.cu:1333-1352). Stops once
// no group exceeds cap. bucket(i, pass) is the bucket of grouped element i.
template <class BucketFn>
void reference_group_lists(ulli n, ulli cap, int p0, int pend, BucketFn bucket, std::vector<ulli> &list,
                           std::vector<ulli> &prefix) {
    list.assign(1, n);
    prefix.assign(1, 0);
    for (int pass = p0; pass < pend; ++pass) {
        std::vector<ulli> nl, np;
        if (list.size() < 10) {
            for (size_t i = 0; i < list.size(); ++i) {
                if (list[i] == 0) continue;
                if (list[i] < cap) {
                    nl.push_back(list[i]);
                    np.push_back(prefix[i]);
                    continue;
                }
                ulli hist[256] = {0}, pre[256];
                for (ulli k = prefix[i]; k < prefix[i] + list[i]; ++k) hist[bucket(k, pass)]++;
                int valid = 0;
                for (int j = 0; j < 256; ++j) valid += hist[j] != 0;
                if (valid == 1 && list.size() == 1) {
                    nl.push_back(list[i]);
                    np.push_back(prefix[i]);
                    continue;
                }
                ulli acc = 0;
                for (int j = 0; j < 256; ++j) { pre[j] = acc; acc += hist[j]; }
                for (int j = 0; j < 256; ++j) {
                    if (!hist[j]) continue;
                    int k = 0;
                    ulli cur = hist[j];
                    while (cur < cap && j + k + 1 < 256) {
                        if (cur + hist[j + k + 1] < cap) { cur += hist[j + k + 1]; ++k; }
                        else break;
                    }
                    nl.push_back(cur);
                    np.push_back(prefix[i] + pre[j]);
                    j += k;
                }
            }
        } else {
            std::vector<ulli> out(256 * list.size(), 0);
            for (size_t i = 0; i < list.size(); ++i) {
                if (list[i] > cap) {
                    for (ulli k = prefix[i]; k < prefix[i] + list[i]; ++k) out[256 * i + bucket(k, pass)]++;
                } else {
                    out[256 * i] = list[i];
                }
            }
            ulli pt = 0;
            for (size_t i = 0; i < out.size(); ++i) {
                if (!out[i]) continue;
                size_t j = 0;
                ulli cur = out[i];
                while (cur < cap && i + j + 1 < out.size()) {
                    if (cur + out[i + j + 1] < cap) { cur += out[i + j + 1]; ++j; }
                    else break;
                }
                nl.push_back(cur);
                np.push_back(pt);
                pt += cur;
                i += j;
            }
        }
        list.swap(nl);
        prefix.swap(np);
        ulli mx = 0;
        for (ulli v : list) mx = std::max(mx, v);
        if (mx <= cap) break;
    }
}

// Sorts host keys (with up to two payloads) on the device by key bits
// [lo_bit, 64); results go to new[]'d host arrays.
bool device_group_by_key(const ulli *keys, const ulli *d1, const ulli *d2, ulli n, int lo_bit, ulli *out_k,
                         ulli *out_1, ulli *out_2) {
    ulli *dk = to_device(keys, n);
    void *perm = nullptr, *tmp = nullptr;
    bool good = dk && ok(bpt_device_alloc(n * 4, &perm), "device alloc") &&
                ok(bpt_device_alloc(n * 8, &tmp), "device alloc") &&
                ok(bpt_sort_perm(dk, 8, n, (uint32_t)lo_bit, 64, static_cast<uint32_t *>(perm), nullptr), "sort") &&
                ok(bpt_gather(dk, 8, static_cast<uint32_t *>(perm), n, tmp, nullptr), "gather") &&
                ok(bpt_device_sync(), "sync") && to_host(out_k, static_cast<const ulli *>(tmp), n);
    const ulli *payload[2] = {d1, d2};
    ulli *outp[2] = {out_1, out_2};
    for (int p = 0; p < 2 && good; ++p) {
        if (!payload[p]) continue;
        ulli *dp = to_device(payload[p], n);
        good = dp && ok(bpt_gather(dp, 8, static_cast<uint32_t *>(perm), n, tmp, nullptr), "gather") &&
               ok(bpt_device_sync(), "sync") && to_host(outp[p], static_cast<const ulli *>(tmp), n);
        bpt_device_free(dp);
    }
    bpt_device_free(dk);
    bpt_device_free(perm);
    bpt_device_free(tmp);
    return good;
}

}  // namespace

// ------------------------------------------------------------------ BPTree.h
void *genBPlusTree(ulli *devInputKey, ulli *devInput, ulli inputSize, std::vector<ulli> groupList,
                   std::vector<ulli> groupPrefix, ulli *&dev_btree_out, ulli &treeLevel_out,
                   ulli *&btreeIndex_out, ulli &bTreeSize_out, BTPage *&dev_globalPointer_in,
                   ulli *&dev_currentPage_in) {
    (void)groupList;
    (void)groupPrefix;
    bpt_build_desc d;
    memset(&d, 0, sizeof d);
    d.key_bytes = 8;
    d.val_bytes = 8;
    d.check_sorted = 1;
    bpt_tree *t = nullptr;
    if (!ok(bpt_tree_build(&d, devInputKey, devInput, inputSize, &t, nullptr), "genBPlusTree")) return nullptr;
    // the reference consumes the caller's sorted buffers (BPTree.cu:210-211)
    (void)hipFree(devInput);
    (void)hipFree(devInputKey);
    bpt_tree_info info;
    bpt_tree_get_info(t, &info);
    ulli pages = 0;
    for (uint32_t k = 0; k <= info.height; ++k) pages += info.level_nodes[k];
    const ulli root = 0;
    void *a = nullptr, *b = nullptr;
    if (!ok(bpt_device_alloc(8, &a), "genBPlusTree") || !ok(bpt_device_alloc(8, &b), "genBPlusTree") ||
        !ok(bpt_copy_h2d(a, &root, 8, nullptr), "genBPlusTree") || !ok(bpt_copy_h2d(b, &pages, 8, nullptr), "genBPlusTree")) {
        bpt_device_free(a);
        bpt_device_free(b);
        bpt_tree_free(t);
        return nullptr;
    }
    dev_btree_out = static_cast<ulli *>(a);
    dev_currentPage_in = static_cast<ulli *>(b);
    treeLevel_out = info.height + 1;
    btreeIndex_out = nullptr;
    bTreeSize_out = inputSize;
    dev_globalPointer_in = t;
    return t;
}

void freeBPlusTree(BTPage *dev_globalPointer, ulli *dev_btree, ulli *dev_currentPage) {
    bpt_tree_free(dev_globalPointer);
    bpt_device_free(dev_btree);
    bpt_device_free(dev_currentPage);
}

void *SearchingBTreeGPUOnlyOneState(ulli *input_key, ulli inputSize, ulli *dev_Btree, BTPage *dev_globalPointer) {
    (void)dev_Btree;
    return point_search(input_key, inputSize, dev_globalPointer, BPT_POINT_PLAIN);
}

void *SearchingBTreeGPUCombineStep(ulli *input_key, ulli inputSize, ulli *dev_Btree, BTPage *dev_globalPointer) {
    (void)dev_Btree;
    return point_search(input_key, inputSize, dev_globalPointer, BPT_POINT_GROUPED);
}

void *SearchingBTreeGPURange(ulli *input_key_start, ulli *input_key_end, ulli inputSize, ulli *dev_Btree,
                             BTPage *dev_globalPointer) {
    (void)dev_Btree;
    return range_counts(input_key_start, input_key_end, inputSize, dev_globalPointer, BPT_GROUP_NONE);
}

void *SearchingBTreeGPURangeWithGrouping(ulli *input_key_start, ulli *input_key_end, ulli inputSize,
                                         ulli *dev_Btree, BTPage *dev_globalPointer) {
    (void)dev_Btree;
    return range_counts(input_key_start, input_key_end, inputSize, dev_globalPointer, BPT_GROUP_GRID2D);
}

void *SearchingBTreeGPURangeWithGroupingZorder(ulli *input_key_start, ulli *input_key_end, ulli inputSize,
                                               ulli *dev_Btree, BTPage *dev_globalPointer) {
    (void)dev_Btree;
    return range_counts(input_key_start, input_key_end, inputSize, dev_globalPointer, BPT_GROUP_ZORDER);
}

void *SearchingBTreeGPURangeResults(ulli *input_key_start, ulli *input_key_end, ulli inputSize, ulli *dev_Btree,
                                    BTPage *dev_globalPointer, int groupMode, ulli *&output_offset, ulli *&ids) {
    (void)dev_Btree;
    output_offset = nullptr;
    ids = nullptr;
    if (!dev_globalPointer || groupMode < BPT_GROUP_NONE || groupMode > BPT_GROUP_ZORDER) return nullptr;
    const ulli n = inputSize;
    ulli *counts = new ulli[n ? n : 1];
    ulli *offs = new ulli[n + 1];
    ulli *dlo = to_device(input_key_start, n), *dhi = to_device(input_key_end, n);
    void *dfirst = nullptr, *dcount = nullptr, *doff = nullptr, *dids = nullptr;
    uint64_t total = 0;
    bool good = dlo && dhi && ok(bpt_device_alloc(n * 8, &dfirst), "device alloc") &&
                ok(bpt_device_alloc(n * 8, &dcount), "device alloc") && ok(bpt_device_alloc(n * 8, &doff), "device alloc") &&
                ok(bpt_range_count_grouped(dev_globalPointer, dlo, dhi, n, groupMode, static_cast<uint64_t *>(dfirst),
                                           static_cast<uint64_t *>(dcount), nullptr),
                   "range count") &&
                ok(bpt_exclusive_scan_u64(static_cast<const uint64_t *>(dcount), n, static_cast<uint64_t *>(doff), &total,
                                          nullptr),
                   "count scan");
    ulli *h_ids = good ? new ulli[total ? total : 1] : nullptr;
    good = good && ok(bpt_device_alloc(total ? total * 8 : 8, &dids), "device alloc") &&
           ok(bpt_range_ids(dev_globalPointer, static_cast<const uint64_t *>(dfirst), static_cast<const uint64_t *>(dcount),
                            static_cast<const uint64_t *>(doff), n, dids, nullptr),
              "range ids") &&
           to_host(counts, static_cast<const ulli *>(dcount), n) && to_host(offs, static_cast<const ulli *>(doff), n) &&
           to_host(h_ids, static_cast<const ulli *>(dids), total);
    bpt_device_free(dlo);
    bpt_device_free(dhi);
    bpt_device_free(dfirst);
    bpt_device_free(dcount);
    bpt_device_free(doff);
    bpt_device_free(dids);
    if (!good) {
        delete[] counts;
        delete[] offs;
        delete[] h_ids;
        return nullptr;
    }
    offs[n] = total;
    output_offset = offs;
    ids = h_ids;
    return counts;
}

// ------------------------------------------------------------------ grouping.h
int bucketIdentify(ulli input, int nPass) {
    const int shift = 8 * (7 - nPass);
    return (shift < 0 || shift >= 64) ? 0 : (int)((input >> shift) & 0xFF);
}

static void group1d(ulli *input, ulli *d1, ulli *d2, ulli n, int start_pass, ulli cap, ulli *ok_k, ulli *ok_1,
                    ulli *ok_2, std::vector<ulli> &groupList, std::vector<ulli> &groupPrefix, bool *good) {
    // grouped order = key order over the bytes from the start pass down
    const int lo_bit = 0;
    *good = device_group_by_key(input, d1, d2, n, lo_bit, ok_k, ok_1, ok_2);
    if (!*good) return;
    reference_group_lists(n, cap, start_pass, 8, [&](ulli i, int p) { return bucketIdentify(ok_k[i], p); },
                          groupList, groupPrefix);
}

void *genGrouping(ulli *input, ulli inputSize, std::vector<ulli> &groupList, std::vector<ulli> &groupPrefix,
                  ulli Local_Group_Size) {
    ulli mx = 0;
    for (ulli i = 0; i < inputSize; ++i) mx = std::max(mx, input[i]);
    const int start = 8 - (int)ceilf((float)bit_width(mx) / 8.0f);  // grouping.cu:1736-1744
    std::vector<ulli> keys(inputSize ? inputSize : 1);
    bool good = false;
    group1d(input, nullptr, nullptr, inputSize, start, Local_Group_Size, keys.data(), nullptr, nullptr, groupList,
            groupPrefix, &good);
    if (!good) return nullptr;
    return to_device(keys.data(), inputSize);  // device pointer, caller frees
}

void genGroupingWithOnedata(ulli *input, ulli *input_data, ulli inputSize, ulli *&output, ulli *&output_data,
                            std::vector<ulli> &groupList, std::vector<ulli> &groupPrefix, ulli Local_Group_Size) {
    output = new ulli[inputSize ? inputSize : 1];
    output_data = new ulli[inputSize ? inputSize : 1];
    bool good = false;
    group1d(input, input_data, nullptr, inputSize, 0, Local_Group_Size, output, output_data, nullptr, groupList,
            groupPrefix, &good);
    if (!good) fprintf(stderr, "genGroupingWithOnedata failed\n");
}

void genGroupingWithTwodata(ulli *input, ulli *input_data, ulli *input_data2, ulli inputSize, ulli *&output,
                            ulli *&output_data, ulli *&output_data2, std::vector<ulli> &groupList,
                            std::vector<ulli> &groupPrefix, ulli Local_Group_Size) {
    output = new ulli[inputSize ? inputSize : 1];
    output_data = new ulli[inputSize ? inputSize : 1];
    output_data2 = new ulli[inputSize ? inputSize : 1];
    bool good = false;
    // genGroupingWithTwodata starts at pass 0 (grouping.cu:1365-1529)
    group1d(input, input_data, input_data2, inputSize, 0, Local_Group_Size, output, output_data, output_data2,
            groupList, groupPrefix, &good);
    if (!good) fprintf(stderr, "genGroupingWithTwodata failed\n");
}

// ------------------------------------------------------------------ grouping2d.h
int bucketIdentify2d(ulli input, ulli input2, int nPass) {
    const int shift = 3 * (21 - nPass);
    if (shift < 0 || shift >= 64) return 0;
    return (int)(((input >> shift) & 0xF) * ((input2 >> shift) & 0xF));
}

void genGrouping2d(ulli *input, ulli *input2, ulli inputSize, std::vector<ulli> &groupList, ulli *&output,
                   ulli *&output2, std::vector<ulli> &groupPrefix, ulli Local_Group_Size) {
    const ulli n = inputSize;
    output = new ulli[n ? n : 1];
    output2 = new ulli[n ? n : 1];
    ulli ma = 0, mb = 0;
    for (ulli i = 0; i < n; ++i) { ma = std::max(ma, input[i]); mb = std::max(mb, input2[i]); }
    int sa = 16 - (int)ceilf((float)bit_width(ma) / 3.0f), sb = 16 - (int)ceilf((float)bit_width(mb) / 3.0f);
    const int p0 = std::max(sa, sb);  // grouping2d.cu:615-635
    const int q0 = std::max(p0, 0);   // passes below 0 put everything in bucket 0
    // grouped order: by the 2-D buckets of up to 8 successive passes from q0
    std::vector<ulli> comp(n ? n : 1);
    for (ulli i = 0; i < n; ++i) {
        ulli c = 0;
        for (int k = 0; k < 8; ++k) {
            const int p = q0 + k;
            c = (c << 8) | (p < 16 ? (ulli)bucketIdentify2d(input[i], input2[i], p) : 0);
        }
        comp[i] = c;
    }
    std::vector<ulli> sorted_comp(n ? n : 1);
    if (!device_group_by_key(comp.data(), input, input2, n, 0, sorted_comp.data(), output, output2)) {
        fprintf(stderr, "genGrouping2d failed\n");
        return;
    }
    reference_group_lists(n, Local_Group_Size, p0, 16,
                          [&](ulli i, int p) { return bucketIdentify2d(output[i], output2[i], p); }, groupList,
                          groupPrefix);
}

// ------------------------------------------------------------------ zOrder2d.h
void genGrouping2dZ(ulli *input, ulli *input2, ulli *&output, ulli inputSize) {
    output = new ulli[inputSize ? inputSize : 1];
    ulli *da = to_device(input, inputSize), *db = to_device(input2, inputSize);
    void *dz = nullptr;
    bool good = da && db && ok(bpt_device_alloc(inputSize * 8, &dz), "device alloc") &&
                ok(bpt_zorder2d(reinterpret_cast<const uint64_t *>(da), reinterpret_cast<const uint64_t *>(db), inputSize,
                                static_cast<uint64_t *>(dz), nullptr),
                   "genGrouping2dZ") &&
                to_host(output, static_cast<const ulli *>(dz), inputSize);
    if (!good) fprintf(stderr, "genGrouping2dZ failed\n");
    bpt_device_free(da);
    bpt_device_free(db);
    bpt_device_free(dz);
}
