// Point search that allocates nothing, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops). It has
// no counterpart in the reference.
//
// The grouped mode needs scratch (partitioned queries, their results, tile
// counts); the plain and learned modes need none. bpt_point_search takes it from the
// tree's workspace, which grows with hipMalloc. This entry point uses a
// file-scope device array instead: every GPU the library runs on holds its
// own copy of it until the process ends. Host code never passes the array's
// name to a kernel (in host code that name is only a stand-in, not a device
// address): the address is looked up once per device with
// hipGetSymbolAddress, with that device current, and that pointer is what
// the kernels receive.
//
// Every call on one GPU shares that GPU's scratch, which keeps the last
// call's contents (the kernels write every word they read, so nothing needs
// clearing between calls); calls that overlap on one GPU race.
#include <mutex>

#include "bpt_internal.hpp"

namespace {

__device__ __align__(256) uint8_t scratch[BPT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The scratch of the current device; looked up on its first use there.
bpt_status scratch_of_current_device(void **out) {
    int dev = 0;
    BPT_HIP(hipGetDevice(&dev));
    BPT_CHECK(dev >= 0 && dev < kMaxDevices, BPT_ERR_INVALID, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        BPT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        BPT_CHECK(p != nullptr, BPT_ERR_HIP, "no device address for the scratch on device %d", dev);
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return BPT_OK;
}

}  // namespace

extern "C" {

bpt_status bpt_point_search_noalloc(const bpt_tree *t, const void *d_q, uint64_t nq, void *d_out, int mode,
                                    bpt_stream stream) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    BPT_CHECK(mode == BPT_POINT_PLAIN || mode == BPT_POINT_GROUPED || mode == BPT_POINT_LEARNED, BPT_ERR_INVALID,
              "unknown point mode %d", mode);
    const uint64_t need = bpt_point_search_ws_size(t, nq, mode);
    BPT_CHECK(need <= BPT_NOALLOC_SCRATCH_BYTES, BPT_ERR_CAPACITY,
              "point search of %llu queries needs %llu bytes of scratch; the no-alloc scratch holds %llu "
              "(use bpt_point_search or bpt_point_search_ws, or split the batch)",
              (unsigned long long)nq, (unsigned long long)need, (unsigned long long)BPT_NOALLOC_SCRATCH_BYTES);
    void *ws = nullptr;
    if (need) BPT_TRY(scratch_of_current_device(&ws));
    return bpt_point_search_ws(t, d_q, nq, d_out, mode, ws, need ? BPT_NOALLOC_SCRATCH_BYTES : 0, stream);
}

}  // extern "C"
