// Drop-in driver with the reference's command line (main.cu:19-360):
//   bptree_main <n> <keyfile> <datafile>                      build only
//   bptree_main <n> <keyfile> <datafile> <nq> <qfile>          point search
//   bptree_main <n> <keyfile> <datafile> <nq> <startfile> <endfile>   range search
// Files are raw little-endian uint64 (main.cu:38-56). Keys are sorted with
// their data on the device first (main.cu:106-118), the tree is built with
// genBPlusTree, and the searches run through the reference-named functions.
// Extra (no reference counterpart): `-o <prefix>` writes each search's
// results as raw uint64 to <prefix>.<form>.bin; with `--ids`, range searches
// also write the two-phase ids of SearchingBTreeGPURangeResults
// (<prefix>.range_offsets.bin, inputSize + 1 entries, and <prefix>.range_ids.bin).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "bptree.h"
#include "compat/BPTree.h"

static bool read_u64(const char *path, std::vector<ulli> &v, ulli n) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { fprintf(stderr, "Error: cannot open %s\n", path); return false; }
    v.resize(n);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(ulli)));
    if ((ulli)f.gcount() != n * sizeof(ulli)) { fprintf(stderr, "Error: %s is short\n", path); return false; }
    return true;
}

static void write_u64(const std::string &path, const ulli *v, ulli n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v), (std::streamsize)(n * sizeof(ulli)));
}

int main(int argc, char **argv) {
    std::string out_prefix;
    bool want_ids = false;
    std::vector<char *> args;
    for (int i = 0; i < argc; ++i) {
        if (!strcmp(argv[i], "-o") && i + 1 < argc) { out_prefix = argv[++i]; continue; }
        if (!strcmp(argv[i], "--ids")) { want_ids = true; continue; }
        args.push_back(argv[i]);
    }
    const int na = (int)args.size();
    if (na < 4) {
        fprintf(stderr, "usage: %s inputSize fileNameKey filenameData [nq qfile | nq startfile endfile] "
                        "[-o prefix [--ids]]\n", argv[0]);
        return -1;
    }
    const ulli n = strtoull(args[1], nullptr, 10);
    std::vector<ulli> key, data;
    if (!read_u64(args[2], key, n) || !read_u64(args[3], data, n)) return -1;

    void *dk = nullptr, *dv = nullptr;
    if (bpt_device_alloc(n * 8, &dk) || bpt_device_alloc(n * 8, &dv) || bpt_copy_h2d(dk, key.data(), n * 8, nullptr) ||
        bpt_copy_h2d(dv, data.data(), n * 8, nullptr) || bpt_sort_pairs(dk, dv, 8, 8, n, nullptr)) {
        fprintf(stderr, "input preparation failed: %s\n", bpt_last_error());
        return -1;
    }
    std::vector<ulli> groupList, groupPrefix;
    for (ulli off = 0; off < n; off += PAGESIZE) {
        groupPrefix.push_back(off);
        groupList.push_back(off + PAGESIZE <= n ? PAGESIZE : n - off);
    }
    ulli *dev_btree = nullptr, *dev_currentPage = nullptr, *btreeIndex = nullptr;
    ulli treeLevel = 0, treeSize = 0;
    BTPage *pool = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    if (!genBPlusTree(static_cast<ulli *>(dk), static_cast<ulli *>(dv), n, groupList, groupPrefix, dev_btree, treeLevel,
                      btreeIndex, treeSize, pool, dev_currentPage)) {
        return -1;
    }
    auto t1 = std::chrono::steady_clock::now();
    printf("Building the Tree %g\n", std::chrono::duration<double>(t1 - t0).count());

    int rc = 0;
    if (na == 6) {
        const ulli nq = strtoull(args[4], nullptr, 10);
        std::vector<ulli> q;
        if (!read_u64(args[5], q, nq)) return -1;
        ulli *r0 = static_cast<ulli *>(SearchingBTreeGPUCombineStep(q.data(), nq, dev_btree, pool));
        ulli *r1 = static_cast<ulli *>(SearchingBTreeGPUOnlyOneState(q.data(), nq, dev_btree, pool));
        if (!r0 || !r1) rc = 1;
        else {
            ulli found = 0;
            for (ulli i = 0; i < nq; ++i) found += r1[i] != INVALID_DATA;
            printf("Found: %g%% of key\n", nq ? 100.0 * (double)found / (double)nq : 0.0);
            if (!out_prefix.empty()) {
                write_u64(out_prefix + ".grouped.bin", r0, nq);
                write_u64(out_prefix + ".plain.bin", r1, nq);
            }
        }
        delete[] r0;
        delete[] r1;
    } else if (na == 7) {
        const ulli nq = strtoull(args[4], nullptr, 10);
        std::vector<ulli> s, e;
        if (!read_u64(args[5], s, nq) || !read_u64(args[6], e, nq)) return -1;
        const char *names[3] = {"range", "range2d", "rangez"};
        void *(*fns[3])(ulli *, ulli *, ulli, ulli *, BTPage *) = {
            SearchingBTreeGPURange, SearchingBTreeGPURangeWithGrouping, SearchingBTreeGPURangeWithGroupingZorder};
        for (int f = 0; f < 3; ++f) {
            ulli *r = static_cast<ulli *>(fns[f](s.data(), e.data(), nq, dev_btree, pool));
            if (!r) { rc = 1; continue; }
            if (!out_prefix.empty()) write_u64(out_prefix + "." + names[f] + ".bin", r, nq);
            delete[] r;
        }
        if (!out_prefix.empty() && want_ids) {
            ulli *offsets = nullptr, *ids = nullptr;
            ulli *r = static_cast<ulli *>(
                SearchingBTreeGPURangeResults(s.data(), e.data(), nq, dev_btree, pool, 0, offsets, ids));
            if (!r) rc = 1;
            else {
                write_u64(out_prefix + ".range_offsets.bin", offsets, nq + 1);
                write_u64(out_prefix + ".range_ids.bin", ids, offsets[nq]);
                printf("Range ids: %llu\n", offsets[nq]);
            }
            delete[] r;
            delete[] offsets;
            delete[] ids;
        }
    }
    freeBPlusTree(pool, dev_btree, dev_currentPage);
    return rc;
}
