// Learned position model attached to a tree (no reference counterpart).
//
// Root: an FP32 MLP 1-H-H-1 (ReLU, H = 32) over the feature
// x = log2(key + 1) / key_bits, trained on the GPU by the kernels below
// (forward + backward per sample, per-block gradient reduction in LDS, Adam;
// one training step = three launches captured in a hipGraph and replayed) to
// predict the rank fraction of a key. Its output picks one of M linear leaf
// models (j = floor(y * M)). Each leaf model interpolates position between
// the first and last keys routed to it; its error bounds [elo, ehi] are taken
// over every key routed to it with the exact query-time arithmetic, so a key
// present in the tree always lies in its model's window. A query computes
// the window, takes a branch-free lower bound inside it, and compares.
// Models whose window exceeds max_window answer through the plain descent.
//
// Root evaluation: a ReLU network of one scalar input is piecewise linear in
// x. After training, the host lists the breakpoints of the first layer
// (w1 x + b1 = 0) and, on each of those pieces, of the second layer, and
// tabulates y = alpha * x + beta per piece; queries find their piece by a
// branch-free search over the breakpoints in LDS and do one FMA. The
// layer-by-layer forward is kept as a root mode for comparison. Routing is
// whatever function the chosen mode computes; leaf models are fitted with
// that same function, which is what keeps results exact.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "bpt_internal.hpp"

namespace bpt {

constexpr int kH = 32;
constexpr int kNP = 3 * kH + kH * kH + kH + 1;  // 1153 parameters
constexpr int kMaxBreaks = kH + kH * (kH + 1);  // 1088
constexpr int kTile = 64;                       // samples per block tile (one wave)

enum : uint32_t { kModelEmpty = 1, kModelFallback = 2 };

struct LeafModel {  // 48 bytes
    uint64_t k0, k1;
    uint32_t p0;     // position of the first key routed here
    float slope;     // positions per 2^shift key units
    int32_t elo, ehi;
    uint32_t flags, cnt;
    uint32_t shift, pad_;  // keys are compared in units of 2^shift (spans of 2^47 and more)
};

// The lookup's copy of a LeafModel, 16 bytes (65536 models: 1 MB, small
// enough to stay in a CU's L2 beside the tree's traffic):
//   w[0]      bits 0..31 of k0 >> shift
//   w[1]      bits 0..15: bits 32..47 of k0 >> shift; bits 16..20: shift;
//             bits 21..31: the window width ehi - elo rounded up to
//             m * 2^e (7-bit m, 4-bit e), so widths below 128 are exact
//   w[2]      p0 + elo, the window's first slot at offset 0 (int32);
//             INT32_MIN: a fallback model; INT32_MAX: an empty one (a miss)
//   w[3]      slope (float bits)
// shift makes a model's span in key units below 2^47, so for its own keys
// delta = ((key >> shift) - (k0 >> shift)) mod 2^48 is exact.
struct alignas(16) PackedModel {
    uint32_t w[4];
};
static_assert(sizeof(PackedModel) == 16, "16-byte packed leaf model");
constexpr uint64_t kDeltaMask = (1ull << 48) - 1;
constexpr int32_t kPackedFallback = INT32_MIN;
constexpr int32_t kPackedEmpty = INT32_MAX;
constexpr uint32_t kMaxPackedWidth = 127u << 15;

// Width code of w[1] bits 21..31: the least m * 2^e >= w with m < 128.
__host__ __device__ inline uint32_t width_code(uint32_t w) {
    uint32_t e = 0;
    while (((w + (1u << e) - 1) >> e) > 127u) ++e;
    return ((w + (1u << e) - 1) >> e) | (e << 7);
}
__host__ __device__ inline uint32_t width_of(uint32_t code) { return (code & 127u) << (code >> 7); }

struct LearnedView {
    const PackedModel *models;
    const float *params;  // device, layer-by-layer mode
    const float *bx;      // breakpoints (nbreak)
    const float *alpha, *beta;  // nbreak + 1 pieces
    uint32_t nmodels, nbreak, root, pad_;
};

struct LearnedModel {
    LearnedView view{};
    LeafModel *d_models = nullptr;
    PackedModel *d_packed = nullptr;
    float *d_params = nullptr, *d_pwl = nullptr;
    std::vector<float> params;
    bpt_learned_stats stats{};
    uint32_t max_window = 256;
    uint64_t packed_fallbacks = 0;  // models whose queries the fallback pass answers
};

void learned_free(LearnedModel *m) {
    if (!m) return;
    (void)hipFree(m->d_models);
    (void)hipFree(m->d_packed);
    (void)hipFree(m->d_params);
    (void)hipFree(m->d_pwl);
    delete m;
}

template <class K>
__device__ __forceinline__ float key_feature(K key) {
    return log2f((float)key + 1.0f) * (1.0f / (8.0f * sizeof(K)));
}

// ------------------------------------------------------------------ training
// Forward + backward of one 64-sample tile per iteration; per-block sums of
// the gradient (and of the loss) go to partial[block][kNP + 1].
__global__ __launch_bounds__(kTile) void k_mlp_grad(const float *__restrict__ params, const float *__restrict__ xs,
                                                    const float *__restrict__ ts, uint32_t m, float inv_m,
                                                    float *__restrict__ partial) {
    __shared__ float P[kNP];
    __shared__ float s_h1[kTile][kH + 1], s_d2[kTile][kH + 1], s_d1[kTile][kH + 1], s_h2[kTile][kH + 1];
    __shared__ float s_x[kTile], s_dy[kTile], s_loss[kTile];
    const int tid = threadIdx.x;
    for (int i = tid; i < kNP; i += kTile) P[i] = params[i];
    __syncthreads();
    const float *w1 = P, *b1 = P + kH, *W2 = P + 2 * kH, *b2 = W2 + kH * kH, *w3 = b2 + kH, *b3 = w3 + kH;
    constexpr int PER = kH * kH / kTile;  // W2 entries per thread
    static_assert(kH * kH % kTile == 0 && kH % PER == 0, "a thread's W2 entries lie in one row");
    float gW2[PER];  // W2 entries tid*PER .. tid*PER+PER-1
#pragma unroll
    for (int r = 0; r < PER; ++r) gW2[r] = 0.f;
    float gv[4] = {0.f, 0.f, 0.f, 0.f};  // tid < 32: w1, b1, b2, w3 of unit tid
    float gb3 = 0.f, gloss = 0.f;
    for (uint64_t tile = blockIdx.x; tile * kTile < m; tile += gridDim.x) {
        const uint64_t s = tile * kTile + tid;
        const bool valid = s < m;
        const float x = valid ? xs[s] : 0.f, t = valid ? ts[s] : 0.f;
        // ReLU masks are kept as bit words: register arrays are indexed only in
        // fully unrolled loops (a variable index would put them in scratch)
        float h1[kH];
        uint32_t m1 = 0, m2 = 0;
#pragma unroll
        for (int k = 0; k < kH; ++k) {
            const float a = fmaf(w1[k], x, b1[k]);
            h1[k] = a > 0.f ? a : 0.f;
            m1 |= (a > 0.f ? 1u : 0u) << k;
        }
        float y = b3[0];
#pragma unroll 1
        for (int j = 0; j < kH; ++j) {
            float a = b2[j];
#pragma unroll
            for (int k = 0; k < kH; ++k) a = fmaf(W2[j * kH + k], h1[k], a);
            m2 |= (a > 0.f ? 1u : 0u) << j;
            const float h = a > 0.f ? a : 0.f;
            s_h2[tid][j] = h;
            y = fmaf(w3[j], h, y);
        }
        const float e = y - t;
        const float dy = valid ? 2.f * e * inv_m : 0.f;
        s_loss[tid] = valid ? e * e * inv_m : 0.f;
        float d2[kH];
#pragma unroll
        for (int j = 0; j < kH; ++j) {
            d2[j] = ((m2 >> j) & 1u) ? dy * w3[j] : 0.f;
            s_d2[tid][j] = d2[j];
            s_h1[tid][j] = h1[j];
        }
#pragma unroll 1
        for (int k = 0; k < kH; ++k) {
            float d = 0.f;
#pragma unroll
            for (int j = 0; j < kH; ++j) d = fmaf(d2[j], W2[j * kH + k], d);
            s_d1[tid][k] = ((m1 >> k) & 1u) ? d : 0.f;
        }
        s_x[tid] = x;
        s_dy[tid] = dy;
        __syncthreads();
        {
            // this thread's W2 entries share one row j and run over PER
            // consecutive columns from k0
            const int j = (tid * PER) / kH, k0 = (tid * PER) % kH;
#pragma unroll 2
            for (int q = 0; q < kTile; ++q) {
                const float dj = s_d2[q][j];
#pragma unroll
                for (int r = 0; r < PER; ++r) gW2[r] = fmaf(dj, s_h1[q][k0 + r], gW2[r]);
            }
        }
        if (tid < kH) {
            float a = 0.f, b = 0.f, c = 0.f, d = 0.f;
            for (int q = 0; q < kTile; ++q) {
                a = fmaf(s_d1[q][tid], s_x[q], a);
                b += s_d1[q][tid];
                c += s_d2[q][tid];
                d = fmaf(s_dy[q], s_h2[q][tid], d);
            }
            gv[0] += a; gv[1] += b; gv[2] += c; gv[3] += d;
        } else if (tid == kH) {
            float a = 0.f, l = 0.f;
            for (int q = 0; q < kTile; ++q) { a += s_dy[q]; l += s_loss[q]; }
            gb3 += a;
            gloss += l;
        }
        __syncthreads();
    }
    float *out = partial + (uint64_t)blockIdx.x * (kNP + 1);
#pragma unroll
    for (int r = 0; r < PER; ++r) out[2 * kH + tid * PER + r] = gW2[r];
    if (tid < kH) {
        out[tid] = gv[0];                              // w1
        out[kH + tid] = gv[1];                         // b1
        out[2 * kH + kH * kH + tid] = gv[2];           // b2
        out[3 * kH + kH * kH + tid] = gv[3];           // w3
    } else if (tid == kH) {
        out[kNP - 1] = gb3;
        out[kNP] = gloss;
    }
}

__global__ void k_reduce_partials(const float *__restrict__ partial, uint32_t blocks, float *__restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kNP) return;
    float s = 0.f;
    for (uint32_t b = 0; b < blocks; ++b) s += partial[(uint64_t)b * (kNP + 1) + p];
    grad[p] = s;
}

// Adam (Kingma & Ba 2015); one block so the step counter updates once.
__global__ __launch_bounds__(1024) void k_adam(float *__restrict__ params, const float *__restrict__ grad,
                                               float *__restrict__ mom, float *__restrict__ vel,
                                               uint32_t *__restrict__ step, float lr) {
    const uint32_t t = *step + 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    for (int p = threadIdx.x; p < kNP; p += blockDim.x) {
        const float g = grad[p];
        const float mm = b1 * mom[p] + (1.f - b1) * g;
        const float vv = b2 * vel[p] + (1.f - b2) * g * g;
        mom[p] = mm;
        vel[p] = vv;
        params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    }
    __syncthreads();
    if (threadIdx.x == 0) *step = t;
}

// ------------------------------------------------------------------ routing
__device__ __forceinline__ float mlp_forward(const float *P, float x) {
    const float *w1 = P, *b1 = P + kH, *W2 = P + 2 * kH, *b2 = W2 + kH * kH, *w3 = b2 + kH, *b3 = w3 + kH;
    float h1[kH];
#pragma unroll
    for (int k = 0; k < kH; ++k) {
        const float a = fmaf(w1[k], x, b1[k]);
        h1[k] = a > 0.f ? a : 0.f;
    }
    float y = b3[0];
    for (int j = 0; j < kH; ++j) {
        float a = b2[j];
#pragma unroll
        for (int k = 0; k < kH; ++k) a = fmaf(W2[j * kH + k], h1[k], a);
        y = fmaf(w3[j], a > 0.f ? a : 0.f, y);
    }
    return y;
}

// LDS image: root 0: [bx nbreak][alpha nbreak+1][beta nbreak+1];
// root 1: [params kNP]; roots 2 and 3 need none. LAYERWISE = false leaves the
// layer-by-layer network (root 1) out of the code, and with it the registers
// its hidden layer takes; such callers never see root 1.
template <class K, bool LAYERWISE = true>
__device__ __forceinline__ uint32_t route(const LearnedView &v, const float *sh, K key) {
    const float x = key_feature<K>(key);
    float y;
    if (v.root == 0) {
        const uint32_t s = v.nbreak ? lower_bound_flat<float>(sh, v.nbreak, x) : 0u;
        y = fmaf(sh[v.nbreak + s], x, sh[2 * v.nbreak + 1 + s]);
    } else if (LAYERWISE && v.root == 1) {
        y = mlp_forward(sh, x);
    } else if (v.root == 2) {
        y = x;
    } else {
        y = (float)key * (1.0f / (sizeof(K) == 4 ? 4294967296.0f : 18446744073709551616.0f));
    }
    const float f = floorf(y * (float)v.nmodels);
    return f <= 0.f ? 0u : (f >= (float)(v.nmodels - 1) ? v.nmodels - 1 : (uint32_t)f);
}

// Offset of a key from its model's first position: floor(slope * delta),
// delta = (key - k0) mod 2^48, clamped to +-2^40. One FP32 multiply and a
// floor, computed by this one function at fit time and at query time, so the
// error bounds hold exactly for the query-time arithmetic.
__device__ __forceinline__ int64_t model_offset(float slope, uint64_t delta) {
    float x = floorf(slope * (float)delta);
    x = fminf(fmaxf(x, -1099511627776.0f), 1099511627776.0f);
    return (int64_t)x;
}

__device__ __forceinline__ int64_t model_pred(const LeafModel &m, uint64_t key) {
    return (int64_t)m.p0 + model_offset(m.slope, ((key >> m.shift) - (m.k0 >> m.shift)) & kDeltaMask);
}

__device__ __forceinline__ bool packed_fallback(const PackedModel &m) { return (int32_t)m.w[2] == kPackedFallback; }

// The window [lo, hi] in slots of `key` under packed model m (lo > hi: none;
// an empty model's window starts past every tree).
__device__ __forceinline__ void packed_window(const PackedModel &m, uint64_t key, uint64_t n, int64_t *lo,
                                              int64_t *hi) {
    const uint32_t sh = (m.w[1] >> 16) & 31u;
    const uint64_t k0s = (uint64_t)m.w[0] | ((uint64_t)(m.w[1] & 0xFFFFu) << 32);
    const int64_t a =
        (int64_t)(int32_t)m.w[2] + model_offset(__uint_as_float(m.w[3]), ((key >> sh) - k0s) & kDeltaMask);
    const int64_t b = a + (int64_t)width_of(m.w[1] >> 21);
    *lo = a < 0 ? 0 : a;
    *hi = b > (int64_t)n - 1 ? (int64_t)n - 1 : b;
}

__device__ void stage_learned(const LearnedView &v, float *sh) {
    if (v.root == 0) {
        for (uint32_t i = threadIdx.x; i < v.nbreak; i += blockDim.x) sh[i] = v.bx[i];
        for (uint32_t i = threadIdx.x; i <= v.nbreak; i += blockDim.x) {
            sh[v.nbreak + i] = v.alpha[i];
            sh[2 * v.nbreak + 1 + i] = v.beta[i];
        }
    } else if (v.root == 1) {
        for (uint32_t i = threadIdx.x; i < kNP; i += blockDim.x) sh[i] = v.params[i];
    }
    __syncthreads();
}

static size_t learned_lds(const LearnedView &v) {
    if (v.root == 0) return sizeof(float) * (3 * v.nbreak + 2);
    if (v.root == 1) return sizeof(float) * kNP;
    return 16;
}


template <class K, class V, int LF>
__global__ __launch_bounds__(256) void k_route_keys(TreeView t, LearnedView v, uint32_t *__restrict__ routes,
                                                    uint32_t *__restrict__ pmin, uint32_t *__restrict__ pmax,
                                                    uint32_t *__restrict__ cnt) {
    extern __shared__ float sh_l[];
    stage_learned(v, sh_l);
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += step) {
        const uint32_t j = route<K>(v, sh_l, slot_key<K, V, LF>(t, i));
        routes[i] = j;
        atomicMin(&pmin[j], (uint32_t)i);
        atomicMax(&pmax[j], (uint32_t)i);
        atomicAdd(&cnt[j], 1u);
    }
}

template <class K, class V, int LF>
__global__ void k_fit_models(TreeView t, uint32_t nmodels, const uint32_t *__restrict__ pmin,
                             const uint32_t *__restrict__ pmax, const uint32_t *__restrict__ cnt,
                             LeafModel *__restrict__ models) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nmodels) return;
    LeafModel m{};
    m.cnt = cnt[j];
    if (m.cnt == 0) {
        m.flags = kModelEmpty;
    } else {
        const uint64_t a = pmin[j], b = pmax[j];
        m.k0 = (uint64_t)slot_key<K, V, LF>(t, a);
        m.k1 = (uint64_t)slot_key<K, V, LF>(t, b);
        m.p0 = (uint32_t)a;
        const uint64_t span = m.k1 - m.k0;
        m.shift = span >> 47 ? 64u - (uint32_t)__clzll((long long)span) - 47u : 0u;  // span >> shift < 2^47
        const uint64_t ds = (m.k1 >> m.shift) - (m.k0 >> m.shift);
        m.slope = ds ? (float)((double)(b - a) / (double)ds) : 0.0f;
        m.elo = 0x7fffffff;
        m.ehi = -0x7fffffff;
    }
    models[j] = m;
}

template <class K, class V, int LF>
__global__ void k_model_errors(TreeView t, const uint32_t *__restrict__ routes, LeafModel *__restrict__ models) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += step) {
        LeafModel &m = models[routes[i]];
        const int64_t e = (int64_t)i - model_pred(m, (uint64_t)slot_key<K, V, LF>(t, i));
        const int32_t ec = e > 0x3fffffff ? 0x3fffffff : (e < -0x3fffffff ? -0x3fffffff : (int32_t)e);
        atomicMin(&m.elo, ec);
        atomicMax(&m.ehi, ec);
    }
}

__global__ void k_finalize_models(LeafModel *__restrict__ models, uint32_t nmodels, uint32_t max_window,
                                  unsigned long long *__restrict__ acc /* fallback models, fallback keys, sum window*cnt, max window, used */) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nmodels) return;
    LeafModel &m = models[j];
    if (m.flags & kModelEmpty) return;
    const uint64_t w = (uint64_t)((int64_t)m.ehi - (int64_t)m.elo + 1);
    if (w > max_window || w - 1 > kMaxPackedWidth) {  // the packed width is at most 127 * 2^15
        m.flags |= kModelFallback;
        atomicAdd(&acc[0], 1ull);
        atomicAdd(&acc[1], (unsigned long long)m.cnt);
    }
    atomicAdd(&acc[2], (unsigned long long)(w * m.cnt));
    atomicMax(&acc[3], (unsigned long long)w);
    atomicAdd(&acc[4], 1ull);
}

// Also counts the packed models the lookup sends to the fallback pass in
// *fallbacks.
__global__ void k_pack_models(const LeafModel *__restrict__ models, uint32_t nmodels, PackedModel *__restrict__ out,
                              unsigned long long *__restrict__ fallbacks) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nmodels) return;
    const LeafModel m = models[j];
    PackedModel p;
    p.w[0] = p.w[1] = p.w[3] = 0;
    p.w[2] = (uint32_t)kPackedEmpty;
    if (!(m.flags & kModelEmpty)) {
        if (m.flags & kModelFallback) {
            p.w[2] = (uint32_t)kPackedFallback;
            atomicAdd(fallbacks, 1ull);
        } else {
            const uint64_t k0s = m.k0 >> m.shift;
            p.w[0] = (uint32_t)k0s;
            p.w[1] = (uint32_t)((k0s >> 32) & 0xFFFFu) | (m.shift << 16) |
                     (width_code((uint32_t)(m.ehi - m.elo)) << 21);
            p.w[2] = (uint32_t)((int32_t)m.p0 + m.elo);  // elo <= 0: the first key's error is 0
            p.w[3] = __float_as_uint(m.slope);
        }
    }
    out[j] = p;
}

// ------------------------------------------------------------------ query
template <class K, class V, int F, int LF>
__global__ __launch_bounds__(256) void k_point_learned(TreeView t, LearnedView v, const K *__restrict__ q,
                                                       V *__restrict__ out, uint64_t nq, V sentinel,
                                                       uint32_t lds_tree_bytes) {
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    float *sh = reinterpret_cast<float *>(lds + lds_tree_bytes);
    stage_top_levels(t, lds);
    stage_learned(v, sh);
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += step) {
        const K key = __builtin_nontemporal_load(&q[i]);
        const PackedModel m = v.models[route<K>(v, sh, key)];
        V res = sentinel;
        if (packed_fallback(m)) {
            res = leaf_lookup<K, V, LF>(t, descend<K, F>(t, lds, key), key, sentinel);
        } else {
            int64_t lo, hi;
            packed_window(m, (uint64_t)key, t.n, &lo, &hi);
            if (lo <= hi) {
                // branch-free lower bound over slots [lo, hi]
                uint64_t base = (uint64_t)lo, len = (uint64_t)(hi - lo + 1);
                while (len > 1) {
                    const uint64_t half = len >> 1;
                    base = (slot_key<K, V, LF>(t, base + half) < key) ? base + half : base;
                    len -= half;
                }
                const uint64_t p = base + (slot_key<K, V, LF>(t, base) < key);
                if (p <= (uint64_t)hi && slot_key<K, V, LF>(t, p) == key) res = slot_val<K, V, LF>(t, p);
            }
        }
        __builtin_nontemporal_store(res, &out[i]);
    }
}

// Cooperative form (equal key/value widths). Phase A, one lane per query:
// route, model and window [lo, hi] in slots -> leaf range [la, lb] -> level-1
// node range [ja, jb]. A window over several level-1 nodes picks its node by
// counting the maxima of nodes ja..jb-1 (the level-2 entries, contiguous)
// below the query: the lane reads the first two itself, both loads in flight
// together, and searches on only for windows over more than three nodes. A
// window inside one node reads no maxima. Phase B is the plain descent's
// cooperative step from that level-1 node: eight lanes per query, the group's
// eight queries in flight together, through the node and then the leaf.
// Phase A keeps only the lane's own query state, so the kernel holds as many
// waves per CU as the plain descent.
// Queries of fallback models are queued by their wave in its own LDS queue
// (key and position); whenever 64 are queued the wave answers them with the
// plain cooperative descent from the LDS-resident top levels, all 64 lanes
// busy, and the rest at its end. No second pass re-reads them from memory.
// LDS: top levels [lds_bytes] | root tables | per wave: keys[kFbQ], positions[kFbQ]
constexpr uint32_t kFbQ = 128;  // queue entries per wave (at most 63 + 64 at once)

template <class K>
__host__ __device__ constexpr size_t fbq_bytes_per_wave() { return kFbQ * (sizeof(K) + sizeof(uint32_t)); }

// LAYERWISE selects the root-1 form of route (the network evaluated layer by
// layer), whose hidden layer would cost the other roots two waves per SIMD.
template <class K, class V, int F, int LF, bool NT, bool OFF32, bool LAYERWISE>
__global__ __launch_bounds__(256) void k_point_learned_coop(TreeView t, LearnedView v, const K *__restrict__ q,
                                                            V *__restrict__ out, uint64_t nq, V sentinel,
                                                            uint32_t root_off, uint32_t fbq_off) {
    static_assert(sizeof(K) == sizeof(V), "cooperative leaf needs equal key/value widths");
    extern __shared__ uint4 lds_raw[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_raw);
    float *sh = reinterpret_cast<float *>(lds + root_off);
    const uint32_t lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
    K *fbk = reinterpret_cast<K *>(lds + fbq_off + wib * fbq_bytes_per_wave<K>());
    uint32_t *fbp = reinterpret_cast<uint32_t *>(fbk + kFbQ);
    stage_top_levels(t, lds);
    stage_learned(v, sh);
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    // level-2 entries: the maxima of the level-1 nodes (trees of height >= 2)
    const K *nmax = reinterpret_cast<const K *>(t.inner + t.off[t.height >= 2 ? 2 : 1]);
    uint32_t nfb = 0;  // queued fallback queries of this wave (the same in every lane)
    // answers the first cnt (<= 64) queued queries, one per lane
    auto drain = [&](uint32_t cnt) {
        const bool has = lane < cnt;
        const K key = has ? fbk[lane] : K(0);
        const uint32_t pos = has ? fbp[lane] : 0u;
        const V r = coop_lookup64<K, V, F, LF, NT, kOpValue, 8, OFF32>(t, lds, key, sentinel);
        if (has) __builtin_nontemporal_store(r, &out[pos]);
    };
    for (uint64_t base = wave * 64; base < nq; base += nwaves * 64) {
        __builtin_amdgcn_s_setprio(2);  // until the leaf's loads are issued (coop_leaf)
        const uint64_t my = base + lane;
        const K myq = my < nq ? __builtin_nontemporal_load(&q[my]) : K(0);
        const PackedModel m = v.models[route<K, LAYERWISE>(v, sh, myq)];
        const bool fb = my < nq && packed_fallback(m);
        const uint64_t fbm = __ballot(fb);
        if (fbm) {
            if (fb) {
                const uint32_t slot = nfb + (uint32_t)__popcll(fbm & ((1ull << lane) - 1));
                fbk[slot] = myq;
                fbp[slot] = (uint32_t)my;
            }
            nfb += (uint32_t)__popcll(fbm);
        }
        // an empty window (a miss) and a queued query walk node 0 with the group
        bool live = false;
        uint32_t node = 0;
        if (!fb && my < nq) {
            int64_t lo, hi;
            packed_window(m, (uint64_t)myq, t.n, &lo, &hi);
            if (lo <= hi) {
                live = true;
                const uint32_t ja = (uint32_t)(lo / LF) / F, jb = (uint32_t)(hi / LF) / F;
                node = ja;
                if (ja < jb) {
                    const K e0 = nmax[ja];
                    const K e1 = ja + 1 < jb ? nmax[ja + 1] : key_max<K>();
                    node = ja + (uint32_t)(e0 < myq) + (uint32_t)(e1 < myq);
                    if (node == ja + 2 && node < jb) {  // wider windows: lower bound over the rest
                        uint32_t len = jb - node;
                        while (len > 0) {
                            const uint32_t half = len >> 1;
                            if (nmax[node + half] < myq) {
                                node += half + 1;
                                len -= half + 1;
                            } else {
                                len = half;
                            }
                        }
                    }
                }
            }
        }
        const V got = coop_finish64<K, V, F, LF, NT, kOpValue, 8, OFF32>(t, myq, node, 1, sentinel);
        if (my < nq && !fb) __builtin_nontemporal_store(live ? got : sentinel, &out[my]);
        if (nfb >= 64) {
            drain(64);
            const uint32_t rest = nfb - 64;  // < 64: moved to the front of the queue
            const K k2 = lane < rest ? fbk[64 + lane] : K(0);
            const uint32_t p2 = lane < rest ? fbp[64 + lane] : 0u;
            if (lane < rest) {
                fbk[lane] = k2;
                fbp[lane] = p2;
            }
            nfb = rest;
        }
    }
    if (nfb) drain(nfb);
}

// ------------------------------------------------------------------ host side
static void init_params(std::vector<float> &p, uint32_t seed) {
    std::mt19937_64 g(seed);
    std::normal_distribution<double> n1(0.0, 1.0);
    p.assign(kNP, 0.f);
    float *w1 = p.data(), *b1 = w1 + kH, *W2 = b1 + kH, *b2 = W2 + kH * kH, *w3 = b2 + kH, *b3 = w3 + kH;
    // first-layer hinges spread over [0, 1]: w1 = 1, b1 = -k/H (+ jitter)
    for (int k = 0; k < kH; ++k) {
        w1[k] = 1.f;
        b1[k] = -(float)k / kH + 0.01f * (float)n1(g);
    }
    const double s2 = std::sqrt(2.0 / kH);
    for (int i = 0; i < kH * kH; ++i) W2[i] = (float)(s2 * n1(g));
    for (int j = 0; j < kH; ++j) {
        b2[j] = 0.01f;
        w3[j] = (float)(std::sqrt(1.0 / kH) * n1(g));
    }
    b3[0] = 0.f;
}

// Piecewise-linear form of the network on [0, 1]: sorted breakpoints and
// (alpha, beta) per piece, computed in double from the FP32 parameters.
static void compile_pwl(const std::vector<float> &pf, std::vector<float> &bx, std::vector<float> &al,
                        std::vector<float> &be) {
    std::vector<double> p(pf.begin(), pf.end());
    const double *w1 = p.data(), *b1 = w1 + kH, *W2 = b1 + kH, *b2 = W2 + kH * kH, *w3 = b2 + kH, *b3 = w3 + kH;
    std::vector<double> bp;
    for (int k = 0; k < kH; ++k)
        if (w1[k] != 0.0) {
            const double x = -b1[k] / w1[k];
            if (x > 0.0 && x < 1.0) bp.push_back(x);
        }
    std::sort(bp.begin(), bp.end());
    std::vector<double> cuts = {0.0};
    cuts.insert(cuts.end(), bp.begin(), bp.end());
    cuts.push_back(1.0);
    std::vector<double> all = bp;
    for (size_t s = 0; s + 1 < cuts.size(); ++s) {
        const double lo = cuts[s], hi = cuts[s + 1], mid = 0.5 * (lo + hi);
        for (int j = 0; j < kH; ++j) {
            double c = 0, d = b2[j];  // a2_j(x) = c x + d on this piece
            for (int k = 0; k < kH; ++k)
                if (w1[k] * mid + b1[k] > 0) { c += W2[j * kH + k] * w1[k]; d += W2[j * kH + k] * b1[k]; }
            if (c != 0.0) {
                const double x = -d / c;
                if (x > lo && x < hi) all.push_back(x);
            }
        }
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    std::vector<double> edges = {0.0};
    edges.insert(edges.end(), all.begin(), all.end());
    edges.push_back(1.0);
    bx.clear();
    al.clear();
    be.clear();
    for (double x : all) bx.push_back((float)x);
    for (size_t s = 0; s + 1 < edges.size(); ++s) {
        const double mid = 0.5 * (edges[s] + edges[s + 1]);
        double a = 0, b = b3[0];
        for (int j = 0; j < kH; ++j) {
            double c = 0, d = b2[j];
            for (int k = 0; k < kH; ++k)
                if (w1[k] * mid + b1[k] > 0) { c += W2[j * kH + k] * w1[k]; d += W2[j * kH + k] * b1[k]; }
            if (c * mid + d > 0) { a += w3[j] * c; b += w3[j] * d; }
        }
        al.push_back((float)a);
        be.push_back((float)b);
    }
}

// One loss/gradient evaluation on device buffers.
static bpt_status mlp_grad_dev(const float *d_params, const float *d_x, const float *d_t, uint32_t m,
                               float *d_partial, uint32_t blocks, float *d_grad, hipStream_t s) {
    k_mlp_grad<<<blocks, kTile, 0, s>>>(d_params, d_x, d_t, m, 1.0f / (float)m, d_partial);
    BPT_HIP(hipGetLastError());
    k_reduce_partials<<<(kNP + 1 + 255) / 256, 256, 0, s>>>(d_partial, blocks, d_grad);
    BPT_HIP(hipGetLastError());
    return BPT_OK;
}

template <class K, class V, int LF>
__global__ void k_training_samples(TreeView t, uint32_t m, float *__restrict__ xs, float *__restrict__ ts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t pos = (uint64_t)(((double)i + 0.5) * (double)t.n / (double)m);
    const uint64_t p = pos < t.n ? pos : t.n - 1;
    xs[i] = key_feature<K>(slot_key<K, V, LF>(t, p));
    ts[i] = (float)((double)p / (double)t.n);
}

template <class K, class V, int F, int LF>
static bpt_status learned_build_impl(bpt_tree *t, const bpt_learned_config &cfg, LearnedModel *lm, hipStream_t s) {
    const uint32_t M = cfg.leaf_models, m = std::min<uint64_t>(cfg.train_samples, t->n);
    const uint32_t blocks = std::min<uint32_t>(256, std::max<uint32_t>(1, (m + kTile - 1) / kTile));
    float *d_x = nullptr, *d_t = nullptr, *d_partial = nullptr, *d_grad = nullptr, *d_mom = nullptr, *d_vel = nullptr;
    uint32_t *d_step = nullptr, *d_routes = nullptr, *d_pmin = nullptr, *d_pmax = nullptr, *d_cnt = nullptr;
    unsigned long long *d_acc = nullptr;
    bpt_status st = BPT_OK;
    auto H = [&](hipError_t e, const char *w) { if (st == BPT_OK && e != hipSuccess) st = hip_fail(e, w, __FILE__, __LINE__); };
    H(hipMalloc(&d_x, m * 4ull), "hipMalloc");
    H(hipMalloc(&d_t, m * 4ull), "hipMalloc");
    H(hipMalloc(&d_partial, (uint64_t)blocks * (kNP + 1) * 4), "hipMalloc");
    H(hipMalloc(&d_grad, (kNP + 1) * 4), "hipMalloc");
    H(hipMalloc(&d_mom, kNP * 4), "hipMalloc");
    H(hipMalloc(&d_vel, kNP * 4), "hipMalloc");
    H(hipMalloc(&d_step, 4), "hipMalloc");
    H(hipMalloc(&lm->d_params, kNP * 4), "hipMalloc");
    init_params(lm->params, cfg.seed);
    const bool train = cfg.root <= 1;
    auto t0 = std::chrono::steady_clock::now();
    if (st == BPT_OK && train) {
        H(hipMemcpyAsync(lm->d_params, lm->params.data(), kNP * 4, hipMemcpyHostToDevice, s), "params");
        H(hipMemsetAsync(d_mom, 0, kNP * 4, s), "memset");
        H(hipMemsetAsync(d_vel, 0, kNP * 4, s), "memset");
        H(hipMemsetAsync(d_step, 0, 4, s), "memset");
        if (st == BPT_OK) {
            k_training_samples<K, V, LF><<<(m + 255) / 256, 256, 0, s>>>(t->view, m, d_x, d_t);
            H(hipGetLastError(), "samples");
        }
        if (st == BPT_OK) st = mlp_grad_dev(lm->d_params, d_x, d_t, m, d_partial, blocks, d_grad, s);
        float l0 = 0;
        H(hipMemcpyAsync(&l0, d_grad + kNP, 4, hipMemcpyDeviceToHost, s), "loss");
        H(hipStreamSynchronize(s), "sync");
        lm->stats.initial_loss = l0;
        // one Adam step = gradient + reduction + update, captured once and replayed
        hipStream_t cs = nullptr;
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        H(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "stream");
        H(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "capture");
        if (st == BPT_OK) {
            bpt_status s2 = mlp_grad_dev(lm->d_params, d_x, d_t, m, d_partial, blocks, d_grad, cs);
            k_adam<<<1, 1024, 0, cs>>>(lm->d_params, d_grad, d_mom, d_vel, d_step, cfg.lr);
            hipError_t e1 = hipGetLastError();
            hipError_t e2 = hipStreamEndCapture(cs, &g);
            if (s2 != BPT_OK) st = s2;
            H(e1, "adam");
            H(e2, "end capture");
        }
        H(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "instantiate");
        H(hipStreamSynchronize(s), "sync");
        for (uint32_t k = 0; k < cfg.train_steps && st == BPT_OK; ++k) H(hipGraphLaunch(ge, cs), "graph launch");
        H(hipStreamSynchronize(cs), "sync");
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        if (cs) (void)hipStreamDestroy(cs);
        if (st == BPT_OK) st = mlp_grad_dev(lm->d_params, d_x, d_t, m, d_partial, blocks, d_grad, s);
        float l1 = 0;
        H(hipMemcpyAsync(&l1, d_grad + kNP, 4, hipMemcpyDeviceToHost, s), "loss");
        H(hipMemcpyAsync(lm->params.data(), lm->d_params, kNP * 4, hipMemcpyDeviceToHost, s), "params");
        H(hipStreamSynchronize(s), "sync");
        lm->stats.final_loss = l1;
    } else {
        H(hipMemcpy(lm->d_params, lm->params.data(), kNP * 4, hipMemcpyHostToDevice), "params");
    }
    auto t1 = std::chrono::steady_clock::now();
    lm->stats.train_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();

    // root evaluation tables
    std::vector<float> bx, al, be;
    if (cfg.root == 0) compile_pwl(lm->params, bx, al, be);
    lm->stats.breakpoints = (uint32_t)bx.size();
    const uint32_t nb = (uint32_t)bx.size();
    H(hipMalloc(&lm->d_pwl, (3 * (size_t)nb + 2) * 4 + 16), "hipMalloc");
    if (st == BPT_OK && cfg.root == 0) {
        std::vector<float> img(bx);
        img.insert(img.end(), al.begin(), al.end());
        img.insert(img.end(), be.begin(), be.end());
        H(hipMemcpy(lm->d_pwl, img.data(), img.size() * 4, hipMemcpyHostToDevice), "pwl");
    }
    LearnedView &v = lm->view;
    v.params = lm->d_params;
    v.bx = lm->d_pwl;
    v.alpha = lm->d_pwl + nb;
    v.beta = lm->d_pwl + 2 * nb + 1;
    v.nmodels = M;
    v.nbreak = nb;
    v.root = cfg.root;

    // leaf models
    H(hipMalloc(&lm->d_models, (size_t)M * sizeof(LeafModel)), "hipMalloc");
    H(hipMalloc(&d_routes, t->n * 4ull), "hipMalloc");
    H(hipMalloc(&d_pmin, M * 4ull), "hipMalloc");
    H(hipMalloc(&d_pmax, M * 4ull), "hipMalloc");
    H(hipMalloc(&d_cnt, M * 4ull), "hipMalloc");
    H(hipMalloc(&d_acc, 6 * 8), "hipMalloc");
    H(hipMemsetAsync(d_pmin, 0xFF, M * 4ull, s), "memset");
    H(hipMemsetAsync(d_pmax, 0, M * 4ull, s), "memset");
    H(hipMemsetAsync(d_cnt, 0, M * 4ull, s), "memset");
    H(hipMemsetAsync(d_acc, 0, 6 * 8, s), "memset");
    H(hipMalloc(&lm->d_packed, (size_t)M * sizeof(PackedModel)), "hipMalloc");
    v.models = lm->d_packed;
    if (st == BPT_OK) {
        const uint32_t g = (uint32_t)std::min<uint64_t>((t->n + 255) / 256, (uint64_t)device_cu_count() * 16);
        k_route_keys<K, V, LF><<<g, 256, learned_lds(v), s>>>(t->view, v, d_routes, d_pmin, d_pmax, d_cnt);
        k_fit_models<K, V, LF><<<(M + 255) / 256, 256, 0, s>>>(t->view, M, d_pmin, d_pmax, d_cnt, lm->d_models);
        k_model_errors<K, V, LF><<<g, 256, 0, s>>>(t->view, d_routes, lm->d_models);
        k_finalize_models<<<(M + 255) / 256, 256, 0, s>>>(lm->d_models, M, cfg.max_window, d_acc);
        k_pack_models<<<(M + 255) / 256, 256, 0, s>>>(lm->d_models, M, lm->d_packed, d_acc + 5);
        H(hipGetLastError(), "leaf models");
    }
    unsigned long long acc[6] = {0, 0, 0, 0, 0, 0};
    H(hipMemcpyAsync(acc, d_acc, sizeof acc, hipMemcpyDeviceToHost, s), "stats");
    H(hipStreamSynchronize(s), "sync");
    auto t2 = std::chrono::steady_clock::now();
    lm->stats.fit_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    lm->stats.fallback_models = (uint32_t)acc[0];
    lm->stats.fallback_keys = acc[1];
    lm->stats.mean_window = t->n ? (double)acc[2] / (double)t->n : 0.0;
    lm->stats.max_window_seen = (uint32_t)acc[3];
    lm->stats.models_used = (uint32_t)acc[4];
    lm->packed_fallbacks = acc[5];
    lm->max_window = cfg.max_window;
    for (void *p : {(void *)d_x, (void *)d_t, (void *)d_partial, (void *)d_grad, (void *)d_mom, (void *)d_vel,
                    (void *)d_step, (void *)d_routes, (void *)d_pmin, (void *)d_pmax, (void *)d_cnt, (void *)d_acc})
        (void)hipFree(p);
    return st;
}

uint64_t point_learned_ws(const bpt_tree *, uint64_t) { return 0; }  // fallback queues live in LDS

bpt_status point_learned(const bpt_tree *t, const void *d_q, uint64_t nq, void *d_out, void *d_ws, uint64_t ws_bytes,
                         hipStream_t s) {
    BPT_CHECK(t->learned, BPT_ERR_UNSUPPORTED, "no learned model attached to this tree (bpt_learned_build)");
    (void)d_ws;  // no scratch: fallback queues live in LDS
    (void)ws_bytes;
    if (nq == 0) return BPT_OK;
    if (t->n == 0) return launch_plain(t, d_q, nq, d_out, s);
    return dispatch_tree(t, [&](auto kt, auto vt, auto fc, auto lc) -> bpt_status {
        using K = typename decltype(kt)::type;
        using V = typename decltype(vt)::type;
        constexpr int F = decltype(fc)::value, LF = decltype(lc)::value;
        TreeView gv = t->view;  // fallback descents read every level from global memory
        gv.lds_levels = 0;
        gv.lds_bytes = 0;
        const size_t lds = learned_lds(t->learned->view);
        if constexpr (sizeof(K) == sizeof(V)) {
            if (t->coop && t->view.height >= 1 && t->view.count[0] < (1ull << 32) && nq < (1ull << 32)) {
                const uint32_t root_off = (t->view.lds_bytes + 15) / 16 * 16;
                const uint32_t fbq_off = root_off + (uint32_t)((lds + 15) / 16 * 16);
                const size_t lds_all = fbq_off + (size_t)(t->block / 64) * fbq_bytes_per_wave<K>();
                const bool off32 = t->addr32 && t->leaf_bytes <= (1ull << 32) && t->inner_bytes <= (1ull << 32);
                auto go = [&](auto nt, auto o32, auto lw) {
                    k_point_learned_coop<K, V, F, LF, decltype(nt)::value, decltype(o32)::value, decltype(lw)::value>
                        <<<search_grid(t, nq), t->block, lds_all, s>>>(t->view, t->learned->view, (const K *)d_q,
                                                                        (V *)d_out, nq, ~V(0), root_off, fbq_off);
                };
                auto with_o32 = [&](auto nt) {
                    if (t->learned->view.root == 1) go(nt, std::false_type{}, std::true_type{});
                    else if (off32) go(nt, std::true_type{}, std::false_type{});
                    else go(nt, std::false_type{}, std::false_type{});
                };
                if (t->leaf_nt) with_o32(std::true_type{});
                else with_o32(std::false_type{});
                BPT_HIP(hipGetLastError());
                return BPT_OK;
            }
        }
        k_point_learned<K, V, F, LF><<<search_grid(t, nq), t->block, lds, s>>>(
            gv, t->learned->view, (const K *)d_q, (V *)d_out, nq, ~V(0), 0);
        BPT_HIP(hipGetLastError());
        return BPT_OK;
    });
}

}  // namespace bpt

using namespace bpt;

extern "C" {

bpt_status bpt_learned_build(bpt_tree *t, const bpt_learned_config *cfg_in, bpt_learned_stats *stats,
                             bpt_stream stream) {
    BPT_CHECK(t, BPT_ERR_INVALID, "tree is NULL");
    bpt_learned_config cfg{65536, 256, 1u << 20, 2000, 3e-3f, 1, 0};
    if (cfg_in) cfg = *cfg_in;
    if (!cfg.leaf_models) cfg.leaf_models = 65536;
    if (!cfg.max_window) cfg.max_window = 256;
    if (!cfg.train_samples) cfg.train_samples = 1u << 20;
    if (cfg.lr <= 0.f) cfg.lr = 3e-3f;
    BPT_CHECK(cfg.root <= 3, BPT_ERR_INVALID, "root mode %u", cfg.root);
    BPT_CHECK(t->n >= 1 && t->n < (1ull << 31), BPT_ERR_CAPACITY, "learned index needs 1 <= n < 2^31 keys");
    BPT_CHECK(cfg.leaf_models <= (1u << 24), BPT_ERR_CAPACITY, "too many leaf models");
    if (t->learned) {
        learned_free(t->learned);
        t->learned = nullptr;
    }
    auto *lm = new LearnedModel;
    hipStream_t s = as_stream(stream);
    bpt_status st = dispatch_tree(t, [&](auto kt, auto vt, auto fc, auto lc) -> bpt_status {
        using K = typename decltype(kt)::type;
        using V = typename decltype(vt)::type;
        return learned_build_impl<K, V, decltype(fc)::value, decltype(lc)::value>(t, cfg, lm, s);
    });
    if (st != BPT_OK) {
        learned_free(lm);
        return st;
    }
    t->learned = lm;
    if (stats) *stats = lm->stats;
    return BPT_OK;
}

bpt_status bpt_learned_params(const bpt_tree *t, float *h_params, uint32_t count) {
    BPT_CHECK(t && t->learned && h_params, BPT_ERR_INVALID, "no learned model");
    BPT_CHECK(count >= (uint32_t)kNP, BPT_ERR_INVALID, "need %d floats", kNP);
    memcpy(h_params, t->learned->params.data(), kNP * sizeof(float));
    return BPT_OK;
}

bpt_status bpt_mlp_loss_grad(const float *h_params, uint32_t hidden, const float *h_x, const float *h_t, uint64_t m,
                             float *h_grad, float *h_loss) {
    BPT_CHECK(hidden == (uint32_t)kH, BPT_ERR_UNSUPPORTED, "hidden width %u (built for %d)", hidden, kH);
    BPT_CHECK(m >= 1 && m < (1ull << 31) && h_params && h_x && h_t && h_grad, BPT_ERR_INVALID, "bad argument");
    const uint32_t blocks = std::min<uint32_t>(256, (uint32_t)((m + kTile - 1) / kTile));
    float *d_p = nullptr, *d_x = nullptr, *d_t = nullptr, *d_part = nullptr, *d_g = nullptr;
    bpt_status st = BPT_OK;
    auto H = [&](hipError_t e, const char *w) { if (st == BPT_OK && e != hipSuccess) st = hip_fail(e, w, __FILE__, __LINE__); };
    H(hipMalloc(&d_p, kNP * 4), "hipMalloc");
    H(hipMalloc(&d_x, m * 4), "hipMalloc");
    H(hipMalloc(&d_t, m * 4), "hipMalloc");
    H(hipMalloc(&d_part, (uint64_t)blocks * (kNP + 1) * 4), "hipMalloc");
    H(hipMalloc(&d_g, (kNP + 1) * 4), "hipMalloc");
    H(hipMemcpy(d_p, h_params, kNP * 4, hipMemcpyHostToDevice), "copy");
    H(hipMemcpy(d_x, h_x, m * 4, hipMemcpyHostToDevice), "copy");
    H(hipMemcpy(d_t, h_t, m * 4, hipMemcpyHostToDevice), "copy");
    if (st == BPT_OK) st = mlp_grad_dev(d_p, d_x, d_t, (uint32_t)m, d_part, blocks, d_g, nullptr);
    std::vector<float> g(kNP + 1);
    H(hipMemcpy(g.data(), d_g, (kNP + 1) * 4, hipMemcpyDeviceToHost), "copy");
    if (st == BPT_OK) {
        memcpy(h_grad, g.data(), kNP * 4);
        if (h_loss) *h_loss = g[kNP];
    }
    for (void *p : {(void *)d_p, (void *)d_x, (void *)d_t, (void *)d_part, (void *)d_g}) (void)hipFree(p);
    return st;
}

}  // extern "C"
