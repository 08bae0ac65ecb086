// Driver of the reference-named grouping functions (include/compat/grouping.h,
// grouping2d.h, zOrder2d.h), as a caller linked against the reference would
// call them, writing what they return:
//   bptree_groups genGrouping            <n> <cap> <out> <keys>
//   bptree_groups genGroupingWithOnedata <n> <cap> <out> <keys> <data>
//   bptree_groups genGroupingWithTwodata <n> <cap> <out> <keys> <data> <data2>
//   bptree_groups genGrouping2d          <n> <cap> <out> <a> <b>
//   bptree_groups genGrouping2dZ         <n> 0     <out> <a> <b>
//   bptree_groups bucketIdentify         <n> 0     <out> <keys>     (passes 0..7)
//   bptree_groups bucketIdentify2d       <n> 0     <out> <a> <b>    (passes -6..15)
// Inputs and outputs are raw little-endian uint64. Outputs: <out>.list and
// <out>.prefix (the group lists), <out>.0, <out>.1, <out>.2 (the grouped
// arrays, keys first; for genGrouping the keys are read back from the device
// pointer it returns; for genGrouping2dZ <out>.0 holds the Z keys; for the
// bucket functions <out>.0 holds n values per pass, pass-major). The time of
// the call is printed as "<function> ms=<t>".
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "bptree.h"
#include "compat/grouping.h"
#include "compat/grouping2d.h"
#include "compat/utils.h"
#include "compat/zOrder2d.h"

static bool read_u64(const char *path, std::vector<ulli> &v, ulli n) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { fprintf(stderr, "Error: cannot open %s\n", path); return false; }
    v.resize(n ? n : 1);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(ulli)));
    if ((ulli)f.gcount() != n * sizeof(ulli)) { fprintf(stderr, "Error: %s is short\n", path); return false; }
    return true;
}

static void write_u64(const std::string &path, const ulli *v, ulli n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v), (std::streamsize)(n * sizeof(ulli)));
}

int main(int argc, char **argv) {
    if (argc < 6) {
        fprintf(stderr, "usage: %s <function> <n> <cap> <out> <in> [<in2> <in3>]\n", argv[0]);
        return 2;
    }
    const std::string fn = argv[1];
    const ulli n = strtoull(argv[2], nullptr, 10);
    const ulli cap = strtoull(argv[3], nullptr, 10);
    const std::string out = argv[4];
    const int nin = argc - 5;
    std::vector<ulli> in[3];
    for (int k = 0; k < nin && k < 3; ++k)
        if (!read_u64(argv[5 + k], in[k], n)) return 1;
    std::vector<ulli> list, prefix;
    ulli *o[3] = {nullptr, nullptr, nullptr};
    int nout = 0;
    const auto t0 = std::chrono::steady_clock::now();
    if (fn == "genGrouping" && nin == 1) {
        ulli *d = static_cast<ulli *>(genGrouping(in[0].data(), n, list, prefix, cap));
        if (!d) return 1;
        o[0] = new ulli[n ? n : 1];
        if (n && bpt_copy_d2h(o[0], d, n * 8, nullptr) != BPT_OK) return 1;
        (void)hipFree(d);  // the caller frees the returned device buffer
        nout = 1;
    } else if (fn == "genGroupingWithOnedata" && nin == 2) {
        genGroupingWithOnedata(in[0].data(), in[1].data(), n, o[0], o[1], list, prefix, cap);
        nout = 2;
    } else if (fn == "genGroupingWithTwodata" && nin == 3) {
        genGroupingWithTwodata(in[0].data(), in[1].data(), in[2].data(), n, o[0], o[1], o[2], list, prefix, cap);
        nout = 3;
    } else if (fn == "genGrouping2d" && nin == 2) {
        genGrouping2d(in[0].data(), in[1].data(), n, list, o[0], o[1], prefix, cap);
        nout = 2;
    } else if (fn == "genGrouping2dZ" && nin == 2) {
        genGrouping2dZ(in[0].data(), in[1].data(), o[0], n);
        nout = 1;
    } else if (fn == "bucketIdentify" && nin == 1) {
        o[0] = new ulli[8 * n + 1];
        for (int p = 0; p < 8; ++p)
            for (ulli i = 0; i < n; ++i) o[0][p * n + i] = (ulli)bucketIdentify(in[0][i], p);
        write_u64(out + ".0", o[0], 8 * n);
        return 0;
    } else if (fn == "bucketIdentify2d" && nin == 2) {
        o[0] = new ulli[22 * n + 1];
        for (int p = -6; p < 16; ++p)
            for (ulli i = 0; i < n; ++i) o[0][(p + 6) * n + i] = (ulli)bucketIdentify2d(in[0][i], in[1][i], p);
        write_u64(out + ".0", o[0], 22 * n);
        return 0;
    } else {
        fprintf(stderr, "unknown function %s with %d inputs\n", fn.c_str(), nin);
        return 2;
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    printf("%s ms=%.3f groups=%zu\n", fn.c_str(), ms, list.size());
    write_u64(out + ".list", list.data(), list.size());
    write_u64(out + ".prefix", prefix.data(), prefix.size());
    for (int k = 0; k < nout; ++k) {
        write_u64(out + "." + std::to_string(k), o[k], n);
        delete[] o[k];
    }
    return 0;
}
