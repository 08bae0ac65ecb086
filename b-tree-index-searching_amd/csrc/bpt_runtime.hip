// Error reporting, device plumbing and timers of the C ABI.
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "bpt_internal.hpp"

namespace bpt {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

bpt_status hip_fail(hipError_t e, const char *what, const char *file, int line) {
    set_error("%s failed: %s (%d) at %s:%d", what, hipGetErrorString(e), (int)e, file, line);
    (void)hipGetLastError();  // clear the sticky non-fatal error
    return (e == hipErrorOutOfMemory) ? BPT_ERR_NOMEM : BPT_ERR_HIP;
}

int device_cu_count() {
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> g(mu);
    if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
            cu = 256;
        cache[dev] = cu;
    }
    return cache[dev];
}

Workspace::~Workspace() {
    if (ptr) (void)hipFree(ptr);
}

bpt_status Workspace::reserve(uint64_t need) {
    if (need <= bytes) return BPT_OK;
    if (ptr) {
        BPT_HIP(hipDeviceSynchronize());
        BPT_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
    }
    BPT_HIP(hipMalloc(&ptr, need));
    bytes = need;
    return BPT_OK;
}

}  // namespace bpt

namespace bpt {
// Device-to-device copy with 16-byte loads/stores (also the known-size
// reference kernel for HBM byte counters: it reads and writes `bytes`).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// One non-temporal 16-byte load and store per thread and step; launched with
// 4 blocks of 256 threads per CU. On MI355X a 1 GiB copy moved 6.0 TB/s (read
// + write) this way, against 4.1-5.6 TB/s with 4 loads in flight per thread
// or 8-64 blocks per CU, and 5.2 TB/s for hipMemcpyAsync.
__global__ __launch_bounds__(256) void k_copy_u4(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                 uint64_t n16) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += step)
        __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

// Descriptor copy: blockIdx.y picks the piece, the x blocks stride over it
// with 16-byte accesses when its ends and size allow, else 4-byte ones.
__global__ __launch_bounds__(256) void k_desc_copy(const CopyDesc *__restrict__ d) {
    const CopyDesc c = d[blockIdx.y];
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (((reinterpret_cast<uintptr_t>(c.src) | reinterpret_cast<uintptr_t>(c.dst) | c.bytes) & 15) == 0) {
        const u32x4 *s = reinterpret_cast<const u32x4 *>(c.src);
        u32x4 *o = reinterpret_cast<u32x4 *>(c.dst);
        for (uint64_t i = first; i < c.bytes / 16; i += step) o[i] = s[i];
    } else {
        const uint32_t *s = reinterpret_cast<const uint32_t *>(c.src);
        uint32_t *o = reinterpret_cast<uint32_t *>(c.dst);
        for (uint64_t i = first; i < c.bytes / 4; i += step) o[i] = s[i];
    }
}

bpt_status launch_desc_copy(const CopyDesc *d_desc, uint32_t nd, uint64_t max_bytes, hipStream_t s) {
    if (nd == 0 || max_bytes == 0) return BPT_OK;
    BPT_CHECK(nd <= 65535, BPT_ERR_INVALID, "%u copy pieces (at most 65535)", nd);
    // one pass of a block moves 4 KiB with 16-byte accesses
    const uint64_t bx = std::min<uint64_t>(std::max<uint64_t>((max_bytes + 4095) / 4096, 1), 128);
    k_desc_copy<<<dim3((uint32_t)bx, nd), 256, 0, s>>>(d_desc);
    BPT_HIP(hipGetLastError());
    return BPT_OK;
}
}  // namespace bpt

struct bpt_timer {
    hipEvent_t a = nullptr, b = nullptr;
};

extern "C" {

const char *bpt_last_error(void) { return bpt::g_last_error.c_str(); }
const char *bpt_version(void) { return "bptree_amd 0.1 (gfx950)"; }

bpt_status bpt_device_count(int *count) {
    BPT_CHECK(count, BPT_ERR_INVALID, "count is NULL");
    BPT_HIP(hipGetDeviceCount(count));
    return BPT_OK;
}

bpt_status bpt_device_info(int device, int *cu_count, uint64_t *hbm_bytes, char *name, int name_len) {
    hipDeviceProp_t p;
    BPT_HIP(hipGetDeviceProperties(&p, device));
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    return BPT_OK;
}

bpt_status bpt_device_alloc(uint64_t bytes, void **d_ptr) {
    BPT_CHECK(d_ptr, BPT_ERR_INVALID, "d_ptr is NULL");
    *d_ptr = nullptr;
    if (bytes == 0) return BPT_OK;
    BPT_HIP(hipMalloc(d_ptr, bytes));
    return BPT_OK;
}

bpt_status bpt_device_free(void *d_ptr) {
    if (d_ptr) BPT_HIP(hipFree(d_ptr));
    return BPT_OK;
}

bpt_status bpt_host_alloc(uint64_t bytes, void **h_ptr) {
    BPT_CHECK(h_ptr, BPT_ERR_INVALID, "h_ptr is NULL");
    *h_ptr = nullptr;
    if (bytes == 0) return BPT_OK;
    void *p = nullptr;
    BPT_HIP(hipHostMalloc(&p, bytes, hipHostMallocMapped | hipHostMallocPortable));
    void *d = nullptr;
    const hipError_t e = hipHostGetDevicePointer(&d, p, 0);
    if (e != hipSuccess || d != p) {  // the contract is one address for host and device
        (void)hipHostFree(p);
        if (e != hipSuccess) return bpt::hip_fail(e, "hipHostGetDevicePointer", __FILE__, __LINE__);
        bpt::set_error("pinned memory maps to another device address");
        return BPT_ERR_UNSUPPORTED;
    }
    *h_ptr = p;
    return BPT_OK;
}

bpt_status bpt_host_free(void *h_ptr) {
    if (h_ptr) BPT_HIP(hipHostFree(h_ptr));
    return BPT_OK;
}

bpt_status bpt_memset(void *d_ptr, int value, uint64_t bytes, bpt_stream s) {
    if (bytes == 0) return BPT_OK;
    BPT_HIP(hipMemsetAsync(d_ptr, value, bytes, bpt::as_stream(s)));
    return BPT_OK;
}

bpt_status bpt_copy_h2d(void *d_dst, const void *h_src, uint64_t bytes, bpt_stream s) {
    if (bytes == 0) return BPT_OK;
    BPT_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, bpt::as_stream(s)));
    BPT_HIP(hipStreamSynchronize(bpt::as_stream(s)));
    return BPT_OK;
}

bpt_status bpt_copy_d2h(void *h_dst, const void *d_src, uint64_t bytes, bpt_stream s) {
    if (bytes == 0) return BPT_OK;
    BPT_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, bpt::as_stream(s)));
    BPT_HIP(hipStreamSynchronize(bpt::as_stream(s)));
    return BPT_OK;
}

bpt_status bpt_copy_d2d(void *d_dst, const void *d_src, uint64_t bytes, bpt_stream s) {
    if (bytes == 0) return BPT_OK;
    hipStream_t h = bpt::as_stream(s);
    const bool aligned = ((reinterpret_cast<uintptr_t>(d_dst) | reinterpret_cast<uintptr_t>(d_src) | bytes) & 15) == 0;
    if (aligned) {
        const uint64_t n16 = bytes / 16;
        uint64_t blocks = (n16 + 255) / 256;
        const uint64_t cap = (uint64_t)bpt::device_cu_count() * 4;
        if (blocks > cap) blocks = cap;
        bpt::k_copy_u4<<<(uint32_t)blocks, 256, 0, h>>>((const bpt::u32x4 *)d_src, (bpt::u32x4 *)d_dst, n16);
        BPT_HIP(hipGetLastError());
    } else {
        BPT_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, h));
    }
    BPT_HIP(hipStreamSynchronize(h));
    return BPT_OK;
}

bpt_status bpt_stream_create(bpt_stream *s) {
    BPT_CHECK(s, BPT_ERR_INVALID, "s is NULL");
    hipStream_t h;
    BPT_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = reinterpret_cast<bpt_stream>(h);
    return BPT_OK;
}

bpt_status bpt_stream_destroy(bpt_stream s) {
    if (s) BPT_HIP(hipStreamDestroy(bpt::as_stream(s)));
    return BPT_OK;
}

bpt_status bpt_stream_sync(bpt_stream s) {
    BPT_HIP(hipStreamSynchronize(bpt::as_stream(s)));
    return BPT_OK;
}

bpt_status bpt_device_sync(void) {
    BPT_HIP(hipDeviceSynchronize());
    return BPT_OK;
}

bpt_status bpt_timer_create(bpt_timer **t) {
    BPT_CHECK(t, BPT_ERR_INVALID, "t is NULL");
    auto *x = new bpt_timer;
    hipError_t e = hipEventCreate(&x->a);
    if (e == hipSuccess) e = hipEventCreate(&x->b);
    if (e != hipSuccess) {
        if (x->a) (void)hipEventDestroy(x->a);
        delete x;
        return bpt::hip_fail(e, "hipEventCreate", __FILE__, __LINE__);
    }
    *t = x;
    return BPT_OK;
}

bpt_status bpt_timer_destroy(bpt_timer *t) {
    if (!t) return BPT_OK;
    (void)hipEventDestroy(t->a);
    (void)hipEventDestroy(t->b);
    delete t;
    return BPT_OK;
}

bpt_status bpt_timer_start(bpt_timer *t, bpt_stream s) {
    BPT_CHECK(t, BPT_ERR_INVALID, "timer is NULL");
    BPT_HIP(hipEventRecord(t->a, bpt::as_stream(s)));
    return BPT_OK;
}

bpt_status bpt_timer_stop(bpt_timer *t, bpt_stream s) {
    BPT_CHECK(t, BPT_ERR_INVALID, "timer is NULL");
    BPT_HIP(hipEventRecord(t->b, bpt::as_stream(s)));
    return BPT_OK;
}

bpt_status bpt_timer_ms(bpt_timer *t, float *ms) {
    BPT_CHECK(t && ms, BPT_ERR_INVALID, "NULL argument");
    BPT_HIP(hipEventSynchronize(t->b));
    BPT_HIP(hipEventElapsedTime(ms, t->a, t->b));
    return BPT_OK;
}

}  // extern "C"
