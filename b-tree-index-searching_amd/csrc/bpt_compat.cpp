// The reference's C++ call surface (BPTree.h, grouping.h, grouping2d.h,
// zOrder2d.h of sc25index/B-tree-index-searching) over the C ABI.
//
// Searches and partitions run on the device through bpt_* calls. What stays
// on the host is what the reference itself keeps on the host: copying the
// caller's arrays, and the group-list bookkeeping of its grouping loop (the
// greedy merge of adjacent buckets below the group cap, over bucket counts
// the device computes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bpt_internal.hpp"
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

// Runs the reference's grouping pass loop (msd_group) on device copies of
// the caller's host arrays. arrays[0] (and arrays[1] for kind 1) are the keys;
// the others move with them. start_pass < -100 takes the start pass from the
// device maxima (grouping.cu:1722-1744, grouping2d.cu:588-635). On success
// the grouped arrays are left on the device in dev[], owned by the caller.
bool device_grouping(const ulli *const *host, int narrays, int kind, ulli n, ulli cap, int start_pass,
                     std::vector<ulli> &groupList, std::vector<ulli> &groupPrefix, ulli **dev) {
    for (int k = 0; k < narrays; ++k) dev[k] = nullptr;
    bool good = true;
    for (int k = 0; k < narrays && good; ++k) {
        void *d = nullptr;
        good = ok(bpt_device_alloc(n ? n * 8 : 8, &d), "device alloc");
        dev[k] = static_cast<ulli *>(d);
        good = good && (n == 0 || ok(bpt_copy_h2d(d, host[k], n * 8, nullptr), "copy to device"));
    }
    if (good && start_pass < -100) {
        uint64_t ma = 0, mb = 0;
        good = ok(n == 0 ? BPT_OK
                         : bpt::pair_max<uint64_t>(reinterpret_cast<uint64_t *>(dev[0]),
                                                   reinterpret_cast<uint64_t *>(dev[kind == 1 ? 1 : 0]), n, &ma,
                                                   &mb, nullptr),
                  "grouping maximum");
        start_pass = kind == 0 ? bpt::group_start_pass_1d(ma) : bpt::group_start_pass_2d(ma, mb);
    }
    std::vector<uint64_t> list, prefix;
    if (good) {
        uint64_t *arr[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int k = 0; k < narrays; ++k) arr[k] = reinterpret_cast<uint64_t *>(dev[k]);
        good = ok(bpt::msd_group(arr, narrays, kind, n, cap, start_pass, list, prefix, nullptr), "grouping");
    }
    if (!good) {
        for (int k = 0; k < narrays; ++k) {
            bpt_device_free(dev[k]);
            dev[k] = nullptr;
        }
        return false;
    }
    groupList.assign(list.begin(), list.end());
    groupPrefix.assign(prefix.begin(), prefix.end());
    return true;
}

// The grouped device arrays to new[]'d host arrays, freeing the device copies.
bool grouped_to_host(ulli **dev, int narrays, ulli n, ulli **host_out) {
    bool good = true;
    for (int k = 0; k < narrays; ++k) {
        good = good && to_host(host_out[k], dev[k], n);
        bpt_device_free(dev[k]);
        dev[k] = nullptr;
    }
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
// The group lists are those of the reference's pass loop; the caller's
// vectors are replaced (the reference appends to them, which every one of its
// call sites passes empty). The reference's small-group copy from offset 0
// (grouping.cu:789) is not reproduced: every element keeps its key and
// payloads.
int bucketIdentify(ulli input, int nPass) {
    const int shift = 8 * (7 - nPass);
    return (shift < 0 || shift >= 64) ? 0 : (int)((input >> shift) & 0xFF);
}

void *genGrouping(ulli *input, ulli inputSize, std::vector<ulli> &groupList, std::vector<ulli> &groupPrefix,
                  ulli Local_Group_Size) {
    const ulli *h[1] = {input};
    ulli *d[1];
    // start pass from the maximum (grouping.cu:1722-1744)
    if (!device_grouping(h, 1, 0, inputSize, Local_Group_Size, -1000, groupList, groupPrefix, d)) return nullptr;
    return d[0];  // device pointer, the caller frees it (BPTreeExactSearch.cu:699)
}

void genGroupingWithOnedata(ulli *input, ulli *input_data, ulli inputSize, ulli *&output, ulli *&output_data,
                            std::vector<ulli> &groupList, std::vector<ulli> &groupPrefix, ulli Local_Group_Size) {
    output = new ulli[inputSize ? inputSize : 1];
    output_data = new ulli[inputSize ? inputSize : 1];
    const ulli *h[2] = {input, input_data};
    ulli *d[2], *out[2] = {output, output_data};
    // passes from 0 (grouping.cu:1567)
    if (!device_grouping(h, 2, 0, inputSize, Local_Group_Size, 0, groupList, groupPrefix, d) ||
        !grouped_to_host(d, 2, inputSize, out))
        fprintf(stderr, "genGroupingWithOnedata failed\n");
}

void genGroupingWithTwodata(ulli *input, ulli *input_data, ulli *input_data2, ulli inputSize, ulli *&output,
                            ulli *&output_data, ulli *&output_data2, std::vector<ulli> &groupList,
                            std::vector<ulli> &groupPrefix, ulli Local_Group_Size) {
    output = new ulli[inputSize ? inputSize : 1];
    output_data = new ulli[inputSize ? inputSize : 1];
    output_data2 = new ulli[inputSize ? inputSize : 1];
    const ulli *h[3] = {input, input_data, input_data2};
    ulli *d[3], *out[3] = {output, output_data, output_data2};
    // passes from 0 (grouping.cu:1413)
    if (!device_grouping(h, 3, 0, inputSize, Local_Group_Size, 0, groupList, groupPrefix, d) ||
        !grouped_to_host(d, 3, inputSize, out))
        fprintf(stderr, "genGroupingWithTwodata failed\n");
}

// ------------------------------------------------------------------ grouping2d.h
int bucketIdentify2d(ulli input, ulli input2, int nPass) {
    const int shift = 3 * (21 - nPass);
    if (shift < 0 || shift >= 64) return 0;
    return (int)(((input >> shift) & 0xF) * ((input2 >> shift) & 0xF));
}

void genGrouping2d(ulli *input, ulli *input2, ulli inputSize, std::vector<ulli> &groupList, ulli *&output,
                   ulli *&output2, std::vector<ulli> &groupPrefix, ulli Local_Group_Size) {
    output = new ulli[inputSize ? inputSize : 1];
    output2 = new ulli[inputSize ? inputSize : 1];
    const ulli *h[2] = {input, input2};
    ulli *d[2], *out[2] = {output, output2};
    // start pass from both maxima (grouping2d.cu:588-635)
    if (!device_grouping(h, 2, 1, inputSize, Local_Group_Size, -1000, groupList, groupPrefix, d) ||
        !grouped_to_host(d, 2, inputSize, out))
        fprintf(stderr, "genGrouping2d failed\n");
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
