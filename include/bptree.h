/*
 * bptree.h — C ABI of the MI355X (gfx950) B+tree index library
 * (libbptree_amd.so). Plain C: opaque handles, pointers and sizes.
 *
 * Every pointer named d_* is device memory (hipMalloc / bpt_device_alloc);
 * every call is asynchronous on the given stream unless it says otherwise;
 * NULL means the default stream. Functions return a bpt_status; on failure
 * bpt_last_error() (thread-local) describes the cause. Nothing here calls
 * exit().
 *
 * Reference call surface each entry point replaces (paths relative to the
 * sc25index/B-tree-index-searching tree) is named in its comment. The C++
 * compatibility headers under include/compat/ keep the reference's exact
 * names and forward to these functions.
 */
#ifndef BPTREE_AMD_H
#define BPTREE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bpt_tree bpt_tree;
typedef void *bpt_stream; /* a hipStream_t */

typedef enum bpt_status {
    BPT_OK = 0,
    BPT_ERR_INVALID = 1,     /* bad argument / unsupported combination */
    BPT_ERR_UNSORTED = 2,    /* build input keys not non-decreasing */
    BPT_ERR_NOMEM = 3,       /* device allocation failed */
    BPT_ERR_HIP = 4,         /* HIP runtime error */
    BPT_ERR_CAPACITY = 5,    /* size above a documented maximum */
    BPT_ERR_UNSUPPORTED = 6, /* mode not available for this tree */
    BPT_ERR_NOT_FOUND = 7    /* no such handle / entry */
} bpt_status;

/* Miss sentinel: all ones of the value width (INVALID_DATA, BPTree.h:12). */
#define BPT_SENTINEL32 0xFFFFFFFFu
#define BPT_SENTINEL64 0xFFFFFFFFFFFFFFFFull

const char *bpt_last_error(void);
const char *bpt_version(void);

/* ------------------------------------------------------------------ */
/* Device plumbing (used by hosts that do not bring their own runtime) */
/* ------------------------------------------------------------------ */
bpt_status bpt_device_count(int *count);
bpt_status bpt_device_info(int device, int *cu_count, uint64_t *hbm_bytes, char *name, int name_len);
bpt_status bpt_device_alloc(uint64_t bytes, void **d_ptr);
bpt_status bpt_device_free(void *d_ptr);
/* Pinned host memory mapped into the device's address space at the same
 * address: the host reads and writes it through the pointer, and kernels
 * (the serving loop's BPT_REQ_DEVICE requests, bpt_copy_*) can take the
 * same pointer, reaching it over PCIe. */
bpt_status bpt_host_alloc(uint64_t bytes, void **h_ptr);
bpt_status bpt_host_free(void *h_ptr);
bpt_status bpt_memset(void *d_ptr, int value, uint64_t bytes, bpt_stream s);
/* Synchronous copies (the stream is synchronized before return). */
bpt_status bpt_copy_h2d(void *d_dst, const void *h_src, uint64_t bytes, bpt_stream s);
bpt_status bpt_copy_d2h(void *h_dst, const void *d_src, uint64_t bytes, bpt_stream s);
bpt_status bpt_copy_d2d(void *d_dst, const void *d_src, uint64_t bytes, bpt_stream s);
bpt_status bpt_stream_create(bpt_stream *s);
bpt_status bpt_stream_destroy(bpt_stream s);
bpt_status bpt_stream_sync(bpt_stream s);
bpt_status bpt_device_sync(void);
/* HIP-event timer on a stream: record start, record stop, read ms. */
typedef struct bpt_timer bpt_timer;
bpt_status bpt_timer_create(bpt_timer **t);
bpt_status bpt_timer_destroy(bpt_timer *t);
bpt_status bpt_timer_start(bpt_timer *t, bpt_stream s);
bpt_status bpt_timer_stop(bpt_timer *t, bpt_stream s);
bpt_status bpt_timer_ms(bpt_timer *t, float *ms); /* waits for the stop event */

/* ------------------------------------------------------------------ */
/* Bulk build                                                          */
/* ------------------------------------------------------------------ */
/* Replaces genBPlusTree (BPTree.h:26-29, BPTree.cu:127-306): page pool +
 * leaf copy + one kernel per internal level. Here: a flat, level-ordered
 * layout; leaves of `leaf_cap` (key, value) slots, interleaved in 16-byte
 * pieces when key and value widths are equal ([k0 k1 v0 v1] for 4-byte,
 * [k0 v0] for 8-byte slots), otherwise all keys then all values; internal
 * nodes of `fanout` separators (child maximum, as the reference's
 * kernelMaxIntermidatePage, BPTree.cu:54), children implicit. */
typedef struct bpt_build_desc {
    uint32_t key_bytes;    /* 4 or 8 */
    uint32_t val_bytes;    /* 4 or 8 */
    uint32_t fanout;       /* internal separators per node: 16, 32 or 64 (0: 128 B nodes) */
    uint32_t leaf_cap;     /* slots per leaf: 8, 16 or 32 (0: 128 B leaves) */
    uint32_t check_sorted; /* non-zero: reject keys that decrease (BPT_ERR_UNSORTED) */
    uint32_t lds_bytes;    /* LDS budget for the top levels (0: 32 KiB) */
} bpt_build_desc;

/* Builds from n keys sorted non-decreasing and their values, both on the
 * device; the inputs are not modified and remain the caller's. n == 0 gives
 * an empty tree on which every lookup misses. Synchronizes `s`. The tree also
 * keeps the grouped search's bucket tables for its default options (5 MB
 * for 64M uint32 keys; rebuilt by every update); without device memory for
 * them it keeps none, and grouped searches build them per call. */
bpt_status bpt_tree_build(const bpt_build_desc *desc, const void *d_keys, const void *d_vals,
                          uint64_t n, bpt_tree **out, bpt_stream s);
bpt_status bpt_tree_free(bpt_tree *t);

#define BPT_MAX_LEVELS 16
typedef struct bpt_tree_info {
    uint64_t n;
    uint32_t key_bytes, val_bytes, fanout, leaf_cap;
    uint32_t height;     /* internal levels above the leaves */
    uint32_t lds_levels; /* top internal levels staged in LDS by the search kernels */
    uint64_t lds_bytes;
    uint64_t leaf_stride;               /* bytes per leaf */
    uint64_t leaf_bytes, inner_bytes;   /* device bytes */
    uint64_t level_nodes[BPT_MAX_LEVELS]; /* [0] = leaves, [height] = 1 */
} bpt_tree_info;
bpt_status bpt_tree_get_info(const bpt_tree *t, bpt_tree_info *info);
/* Device pointers of the flat layout (for tests and the learned index). */
bpt_status bpt_tree_get_layout(const bpt_tree *t, const void **d_leaves, const void **d_inner);
/* Launch shape of the search kernels: block of 64, 128 or 256 threads,
 * blocks per CU of the persistent grid (0 keeps the default). */
bpt_status bpt_tree_set_launch(bpt_tree *t, uint32_t block, uint32_t blocks_per_cu);

/* Batched updates (no reference counterpart: its pages carry next/previous
 * links and a bump allocator, BPTree.h:14-23 and BPTree.cu:164, but a built
 * tree is never changed). Device arrays of the tree's key and value widths.
 * bpt_tree_upsert: each key takes its new value, or is added; of equal keys
 * in the batch the last one (in array order) wins. bpt_tree_erase: every slot
 * whose key is one of d_keys is removed (absent keys are ignored). The batch
 * is merged with the tree's slots on the device and the tree is rebuilt with
 * its own options (O(tree + batch) per call); a trained learned model is
 * dropped. The merge's scratch is the tree's workspace (about twice the
 * tree's slot bytes plus 16 B per slot), kept for later calls; the leaf and
 * inner buffers are reused when the new tree fits them. Synchronize s.
 * When a step fails before the new tree is written (argument checks, the
 * workspace, the merge, allocating new leaf or inner buffers), the tree is
 * left unchanged, learned model included; only a device error while the new
 * tree is written leaves it empty. */
bpt_status bpt_tree_upsert(bpt_tree *t, const void *d_keys, const void *d_vals, uint64_t n, bpt_stream s);
bpt_status bpt_tree_erase(bpt_tree *t, const void *d_keys, uint64_t n, bpt_stream s);

/* Search-kernel options: "coop" (1 = several lanes share each node, the
 * default: the whole descent when key and value widths match, the inner
 * levels otherwise, with the leaf read by the query's lane; 0 = one lane per
 * query),
 * "lanes" (lanes per query of the cooperative descent: 8, the default, or 4),
 * "leaf_nt" (1 = non-temporal leaf loads, the default),
 * "addr32" (1 = 32-bit byte offsets into levels and leaves below 4 GiB, the
 * default; 0 = 64-bit addresses throughout), and for the grouped mode
 * "group_chunk" (queries per search work chunk; 0, the default, picks about
 * 16 chunks per search block, a power of two in [1024, 32768]),
 * "group_lanes" (lanes per query at the leaf: 4, the default, or 8),
 * "group_blocks" (search blocks per CU, default 1; at most as many as the
 * CU's LDS holds),
 * "group_threads" (threads per search block: 512, 768 or 1024, the default),
 * "group_slice" (LDS bytes of one bucket's slice of the tree: 16384, 32768,
 * 65536 or 131072, the default; a larger slice means fewer, wider buckets
 * and fewer search blocks per CU),
 * "group_sub" (1 = slices of sub-leaf maxima, so that a query reads 64
 * bytes of its leaf, the default where the tree's leaves span several
 * 64-byte runs and its key and value widths match; 0 = leaf maxima and
 * whole leaves),
 * "group_xcd" (1 = each XCD searches its own contiguous part of the
 * bucket-ordered queries, the default) and
 * "group_min" (queries; a grouped batch below it runs the plain descent,
 * with the same results and no workspace; default 2^24, about where the
 * grouped kernels overtake the plain descent on MI355X for trees of 16M-64M
 * keys; 0 runs the grouped kernels at every size). "alloc_limit" (bytes; 0, the
 * default, means none) caps each leaf or inner buffer that an update's
 * rebuild allocates: a larger one fails with BPT_ERR_NOMEM, as an exhausted
 * device would, and the tree stays unchanged. */
bpt_status bpt_tree_set_option(bpt_tree *t, const char *key, int64_t value);

/* ------------------------------------------------------------------ */
/* Point search                                                        */
/* ------------------------------------------------------------------ */
typedef enum bpt_point_mode {
    BPT_POINT_PLAIN = 0,   /* SearchingBTreeGPUOnlyOneState: root-to-leaf descent */
    BPT_POINT_GROUPED = 1, /* SearchingBTreeGPUCombineStep: on-device radix grouping first */
    BPT_POINT_LEARNED = 2  /* learned position model + window search (bpt_learned_build) */
} bpt_point_mode;

/* out[i] = value of the first key equal to q[i], else the sentinel of the
 * value width. Results are in the caller's order for every mode.
 * Plain: replaces kernelSearchingNaiveOneStep (BPTreeExactSearch.cu:25-105).
 * Grouped: replaces genGrouping + kernelSearchingKeyGroupTwoStep
 * (grouping.cu:1677-1851, BPTreeExactSearch.cu:206-294); groups on the top
 * `group_bits` key bits (bpt_tree_set_group_bits, default 12).
 * The grouped mode uses a workspace owned by the tree: concurrent calls on
 * one tree must share one stream (or use bpt_point_search_ws). Plain and
 * learned need none (bpt_point_search_ws_size returns 0 for them). */
bpt_status bpt_point_search(bpt_tree *t, const void *d_q, uint64_t nq, void *d_out, int mode,
                            bpt_stream s);
/* Same, with a caller-provided workspace of at least bpt_point_search_ws_size bytes. */
uint64_t bpt_point_search_ws_size(const bpt_tree *t, uint64_t nq, int mode);
bpt_status bpt_point_search_ws(const bpt_tree *t, const void *d_q, uint64_t nq, void *d_out,
                               int mode, void *d_ws, uint64_t ws_bytes, bpt_stream s);
bpt_status bpt_tree_set_group_bits(bpt_tree *t, uint32_t bits);
/* No counterpart in the reference: the same search for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops). It calls no hipMalloc / hipFree. The grouped mode uses a static
 * device scratch of BPT_NOALLOC_SCRATCH_BYTES per GPU (held until the process
 * ends) instead of the tree's workspace; a batch whose workspace
 * (bpt_point_search_ws_size) is larger fails with BPT_ERR_CAPACITY (about 6M
 * queries on a 4-byte-key tree, 3M on an 8-byte one; batches below the
 * tree's "group_min" need no scratch, as they run the plain descent). All calls on one GPU
 * share its scratch, so calls that overlap on one GPU (other streams or
 * threads) race; calls in one stream are safe. The first call on a device
 * looks the scratch's address up. */
#define BPT_NOALLOC_SCRATCH_BYTES (64ull << 20)
bpt_status bpt_point_search_noalloc(const bpt_tree *t, const void *d_q, uint64_t nq, void *d_out,
                                    int mode, bpt_stream s);

/* pos[i] = lower_bound(keys, q[i]) (first slot with key >= q[i]; n if none).
 * Replaces the position half of kernelSearchingKeyRangeGroupStateTwo
 * (BtreeRangeSearch.cu:215-294) with a true lower bound. */
bpt_status bpt_lower_bound(const bpt_tree *t, const void *d_q, uint64_t nq, uint64_t *d_pos,
                           bpt_stream s);

/* ------------------------------------------------------------------ */
/* 1-D range search [lo, hi] inclusive                                  */
/* ------------------------------------------------------------------ */
/* first[i] = lower_bound(lo[i]); count[i] = |{k : lo[i] <= k <= hi[i]}|
 * (0 when lo > hi). Replaces SearchingBTreeGPURange's step 2 + step 3
 * count (BtreeRangeSearch.cu:786-925, :215-374). d_first may be NULL. */
bpt_status bpt_range_count(const bpt_tree *t, const void *d_lo, const void *d_hi, uint64_t nq,
                           uint64_t *d_first, uint64_t *d_count, bpt_stream s);
/* ids for query i (values of the matching slots, in key order) written at
 * d_ids[offsets[i] ...]; offsets = exclusive scan of counts (any
 * non-decreasing offsets with offsets[i] + count[i] <= offsets[i+1] also
 * work; nothing outside the [offsets[i], offsets[i] + count[i]) is written). */
bpt_status bpt_range_ids(const bpt_tree *t, const uint64_t *d_first, const uint64_t *d_count,
                         const uint64_t *d_offsets, uint64_t nq, void *d_ids, bpt_stream s);
/* d_out[i] = sum of d_in[0..i); *h_total (host, may be NULL) = sum of all.
 * Synchronizes `s` when h_total is not NULL. */
bpt_status bpt_exclusive_scan_u64(const uint64_t *d_in, uint64_t n, uint64_t *d_out,
                                  uint64_t *h_total, bpt_stream s);

/* Query-order grouping front ends for range queries (results unchanged). */
typedef enum bpt_group_mode {
    BPT_GROUP_NONE = 0,   /* SearchingBTreeGPURange */
    BPT_GROUP_GRID2D = 1, /* SearchingBTreeGPURangeWithGrouping (grouping2d.cu bucket) */
    BPT_GROUP_ZORDER = 2  /* SearchingBTreeGPURangeWithGroupingZorder (zOrder2d.cu key) */
} bpt_group_mode;
/* The grouped orders use the tree's workspace, like bpt_point_search:
 * concurrent calls on one tree must share one stream. */
bpt_status bpt_range_count_grouped(bpt_tree *t, const void *d_lo, const void *d_hi,
                                   uint64_t nq, int group_mode, uint64_t *d_first,
                                   uint64_t *d_count, bpt_stream s);

/* ------------------------------------------------------------------ */
/* Grouping building blocks                                             */
/* ------------------------------------------------------------------ */
/* Stable partition of n keys (key_bytes 4 or 8) by bucket
 * (key >> shift) & ((1 << bits) - 1), bits <= 16. d_perm[j] = original index
 * of the j-th element in grouped order; d_bucket_start (1 << bits) + 1
 * entries, may be NULL. Replaces the MSD passes of genGrouping
 * (grouping.cu:48-478,1677-1851) with an on-device stable radix partition;
 * synchronizes s. */
bpt_status bpt_radix_group(const void *d_keys, uint32_t key_bytes, uint64_t n, uint32_t shift,
                           uint32_t bits, uint32_t *d_perm, uint64_t *d_bucket_start,
                           bpt_stream s);
/* Stable sort permutation of n keys by bits [lo_bit, hi_bit): d_perm[j] =
 * index of the j-th element in sorted order (LSD radix, 8 bits a pass).
 * Synchronizes s. */
bpt_status bpt_sort_perm(const void *d_keys, uint32_t key_bytes, uint64_t n, uint32_t lo_bit,
                         uint32_t hi_bit, uint32_t *d_perm, bpt_stream s);
/* d_dst[i] = d_src[d_idx[i]] for 4- or 8-byte elements. */
bpt_status bpt_gather(const void *d_src, uint32_t elem_bytes, const uint32_t *d_idx, uint64_t n,
                      void *d_dst, bpt_stream s);
/* In-place stable sort of (key, value) pairs by key on the device, the input
 * preparation the reference does with cub::DeviceRadixSort::SortPairs
 * (main.cu:106-118). d_vals may be NULL. Synchronizes s. */
bpt_status bpt_sort_pairs(void *d_keys, void *d_vals, uint32_t key_bytes, uint32_t val_bytes,
                          uint64_t n, bpt_stream s);
/* 64-bit Morton code, x in even bits, y in odd bits (zOrder2d.cu:27-54). */
bpt_status bpt_morton_encode(const uint32_t *d_x, const uint32_t *d_y, uint64_t n,
                             uint64_t *d_z, bpt_stream s);
/* Z-order grouping key of (a, b) pairs: Morton(a / ix, b / iy) with
 * ix = ceil(max(a) / 2^32) in float, iy likewise (zOrder2d.cu:65-171,
 * genGrouping2dZ). Synchronizes s. */
bpt_status bpt_zorder2d(const uint64_t *d_a, const uint64_t *d_b, uint64_t n, uint64_t *d_out,
                        bpt_stream s);
/* 2-D grouping key of (a, b) pairs: the bucket of grouping2d.cu:24-31 for
 * four consecutive passes from the start pass of grouping2d.cu:610-635, one
 * byte each, most significant first. Synchronizes s. */
bpt_status bpt_group2d_keys(const uint64_t *d_a, const uint64_t *d_b, uint64_t n,
                            uint32_t *d_out, bpt_stream s);

/* ------------------------------------------------------------------ */
/* 2-D rectangle search over Morton-keyed points (u64 keys, u32 ids)    */
/* ------------------------------------------------------------------ */
/* rects: n x {x0, y0, x1, y1} uint32, inclusive. The tree holds
 * key = Morton(x, y), value = point id. counts[i] = points inside rect i;
 * bpt_rect_ids writes their ids ascending at ids[offsets[i] ...], with
 * offsets the exclusive scan of the counts (a rectangle whose offset equals
 * the next one's has no hits and is skipped).
 * group_mode orders the queries of the count (bpt_group_mode); results do
 * not change.
 * The grouped orders use the tree's workspace, like bpt_point_search:
 * concurrent calls on one tree must share one stream. */
bpt_status bpt_rect_count(bpt_tree *t, const uint32_t *d_rects, uint64_t nr, int group_mode,
                          uint64_t *d_counts, bpt_stream s);
bpt_status bpt_rect_ids(bpt_tree *t, const uint32_t *d_rects, uint64_t nr, int group_mode,
                        const uint64_t *d_offsets, uint32_t *d_ids, bpt_stream s);

/* ------------------------------------------------------------------ */
/* Learned position model (no reference counterpart)                    */
/* ------------------------------------------------------------------ */
/* Root: FP32 MLP 1-32-32-1 (ReLU) over x = log2(key + 1) / key_bits,
 * trained on the device by the library's forward/backward kernels (Adam)
 * to predict the rank fraction; it selects one of `leaf_models` linear
 * models with stored error bounds; the window lower bound answers exactly,
 * and models whose window exceeds `max_window` use the plain descent.
 * root: 0 = the MLP evaluated through its piecewise-linear form,
 *       1 = the MLP evaluated layer by layer,
 *       2 = no network, routing linear in the log-key feature,
 *       3 = no network, routing linear in the key. */
typedef struct bpt_learned_config {
    uint32_t leaf_models;   /* 0: 65536 */
    uint32_t max_window;    /* 0: 256 */
    uint32_t train_samples; /* keys sampled at uniform ranks (0: 2^20) */
    uint32_t train_steps;   /* Adam steps */
    float lr;               /* 0: 3e-3 */
    uint32_t seed;
    uint32_t root;
} bpt_learned_config;
typedef struct bpt_learned_stats {
    float initial_loss, final_loss; /* mean squared rank-fraction error */
    double train_ms, fit_ms;
    uint32_t breakpoints;           /* pieces - 1 of the compiled root */
    uint32_t models_used, fallback_models, max_window_seen;
    uint64_t fallback_keys;
    double mean_window;             /* window slots, averaged over keys */
} bpt_learned_stats;
/* Trains the root, fits the leaf models and attaches them to t (replacing
 * a previous model); then bpt_point_search(..., BPT_POINT_LEARNED, ...)
 * uses them. Synchronizes s. */
bpt_status bpt_learned_build(bpt_tree *t, const bpt_learned_config *cfg, bpt_learned_stats *stats,
                             bpt_stream s);
/* The trained root parameters: w1[32] b1[32] W2[32*32] b2[32] w3[32] b3. */
bpt_status bpt_learned_params(const bpt_tree *t, float *h_params, uint32_t count);
/* Mean squared error and its gradient for m samples on the device with the
 * training kernels (host arrays in/out; hidden must be 32). */
bpt_status bpt_mlp_loss_grad(const float *h_params, uint32_t hidden, const float *h_x, const float *h_t,
                             uint64_t m, float *h_grad, float *h_loss);

/* ------------------------------------------------------------------ */
/* Serving loop (no reference counterpart)                              */
/* ------------------------------------------------------------------ */
/* Continuous batching over a resident tree: `slots` streams, requests
 * admitted FIFO when a slot is free and their device bytes fit in a pool of
 * `mem_budget` bytes (they wait otherwise; one larger than the whole budget
 * is rejected with BPT_ERR_CAPACITY). Buffers stay the caller's until the
 * request completes. Point requests use `mode` (bpt_point_mode). The head
 * of the queue and the requests of its kind and buffer kind queued behind
 * it, in order (up to 512 requests and 2^18 queries, 2^20 for device
 * requests), are answered in one launch; each caller gets its own results.
 * flags = BPT_REQ_DEVICE: every buffer of the request is memory the tree's
 * device can address -- its device memory, or pinned host memory from
 * bpt_host_alloc -- aligned to its element width; there are no host copies
 * (a request alone is searched in place, a fused run is gathered into the
 * pool and its results scattered back by two copy kernels), and the results
 * are in place when wait/replay returns. flags = 0: host buffers, staged
 * through pinned memory. */
typedef struct bpt_server bpt_server;
enum { BPT_REQ_POINT = 0, BPT_REQ_RANGE = 1 };
enum { BPT_REQ_DEVICE = 1 };
typedef struct bpt_request {
    int32_t kind;   /* BPT_REQ_POINT or BPT_REQ_RANGE */
    int32_t flags;  /* 0 (host buffers) or BPT_REQ_DEVICE */
    uint64_t n;     /* queries in the request */
    const void *q;  /* point: keys */
    void *out;      /* point: values (sentinel on a miss) */
    const void *lo, *hi;     /* range: bounds (inclusive) */
    uint64_t *first, *count; /* range: lower bound (may be NULL) and count */
} bpt_request;
typedef struct bpt_server_stats {
    uint64_t completed, lookups;
    uint32_t max_inflight, pad_;
    uint64_t peak_bytes, budget_bytes;
    double p50_ms, p99_ms, max_ms, mean_ms; /* submit -> completion */
    double elapsed_s;                       /* first submit -> last completion */
} bpt_server_stats;
bpt_status bpt_server_create(bpt_tree *t, uint32_t slots, uint64_t mem_budget, int mode, bpt_server **out);
bpt_status bpt_server_submit(bpt_server *srv, const bpt_request *r, uint64_t *ticket);
/* Blocks until the request is done; returns its status. */
bpt_status bpt_server_wait(bpt_server *srv, uint64_t ticket, double *latency_ms);
bpt_status bpt_server_get_stats(bpt_server *srv, bpt_server_stats *st);
bpt_status bpt_server_reset_stats(bpt_server *srv);
bpt_status bpt_server_destroy(bpt_server *srv);
/* Open-loop replay: request i is submitted at arrival_s[i] seconds after
 * the call starts, regardless of completions; then all are waited for. */
bpt_status bpt_server_replay(bpt_server *srv, const bpt_request *reqs, uint32_t nreq, const double *arrival_s,
                             double *latency_ms);

/* ------------------------------------------------------------------ */
/* Multi-tree batches (no reference counterpart)                        */
/* ------------------------------------------------------------------ */
/* A forest: ntrees independent trees of 8-byte keys and values in one
 * device arena. Tree t's keys (sorted, unique or not) and values are
 * d_keys/d_vals[h_key_off[t] .. h_key_off[t+1]) (host offsets). */
typedef struct bpt_forest bpt_forest;
typedef struct bpt_batch bpt_batch;
typedef struct bpt_batch_config {
    uint32_t fanout;        /* 16, 32 or 64 (0: 16) */
    uint32_t leaf_cap;      /* 8 (one 128-byte line of keys + values) */
    uint32_t lds_bytes;     /* LDS budget for each tree's top levels (0: 32 KiB) */
    uint32_t chunk;         /* queries per work item (0: 8192) */
    uint32_t blocks_per_cu; /* persistent grid (0: 8) */
    uint32_t lanes;         /* lanes per query of the descent: 8 (a wavefront descends 8
                               queries at a time) or 4 (16 at a time); 0: 8 */
} bpt_batch_config;
bpt_status bpt_forest_build(const bpt_batch_config *cfg, uint32_t ntrees, const uint64_t *d_keys,
                            const uint64_t *d_vals, const uint64_t *h_key_off, bpt_forest **out,
                            bpt_stream s);
bpt_status bpt_forest_free(bpt_forest *f);
bpt_status bpt_forest_info(const bpt_forest *f, uint32_t *ntrees, uint32_t *fanout, uint64_t *leaf_bytes,
                           uint64_t *inner_bytes);
/* A batch over a forest: point queries of tree t at d_q[h_q_off[t] ..
 * h_q_off[t+1]) (results at the same positions of d_out; value or all-ones)
 * and range queries [d_lo, d_hi] of tree t at h_r_off[t] .. h_r_off[t+1]
 * (d_first = lower bound, d_count = inclusive count). Either part may be
 * absent (NULL offsets). The work list is built once. */
bpt_status bpt_batch_prepare(bpt_forest *f, uint32_t chunk, uint32_t blocks_per_cu, const uint64_t *d_q,
                             const uint64_t *h_q_off, uint64_t *d_out, const uint64_t *d_lo,
                             const uint64_t *d_hi, const uint64_t *h_r_off, uint64_t *d_first,
                             uint64_t *d_count, bpt_batch **out);
/* One persistent launch over every work item; use_graph != 0 replays a
 * hipGraph captured on the first call. */
bpt_status bpt_batch_run(bpt_batch *b, int use_graph, bpt_stream s);
bpt_status bpt_batch_free(bpt_batch *b);
bpt_status bpt_batch_info(const bpt_batch *b, uint32_t *nitems, uint64_t *npoints, uint64_t *nranges);
/* Times fanout {16,32,64} x LDS budget {16,32,64} KiB x queries per
 * wavefront {8,16} (lanes 8 or 4) x chunk {2K,8K,32K} on the given batch (reps timed runs after one warm-up each) and returns
 * the fastest configuration; `report` receives a JSON array of all timings. */
bpt_status bpt_forest_autotune(uint32_t ntrees, const uint64_t *d_keys, const uint64_t *d_vals,
                               const uint64_t *h_key_off, const uint64_t *d_q, const uint64_t *h_q_off,
                               uint64_t *d_out, const uint64_t *d_lo, const uint64_t *d_hi,
                               const uint64_t *h_r_off, uint64_t *d_first, uint64_t *d_count, int reps,
                               bpt_batch_config *best, char *report, uint64_t report_len);
/* Synthetic forest workload from a counter-based generator (splitmix64):
 * tree t's keys are n_t draws h(seed,1,t,i) uniform over [0, 2^64), sorted
 * on the device (a tree whose draws collide is redrawn from stream
 * 1 + 16 * attempt), value = rank; qpk point queries per key, hit_pct% of
 * them keys of the tree at a uniform rank, the rest uniform 64-bit values;
 * range j of tree t starts at a uniform rank and spans floor(2^u) ranks,
 * u uniform in [w_lo_exp, w_hi_exp] (log-uniform), clipped to the tree.
 * h_key_off[0] must be 0. Synchronizes s. */
bpt_status bpt_synth_forest(uint64_t seed, uint32_t ntrees, const uint64_t *h_key_off, uint32_t qpk,
                            uint32_t hit_pct, const uint64_t *h_r_off, uint32_t w_lo_exp, uint32_t w_hi_exp,
                            uint64_t *d_keys, uint64_t *d_vals, uint64_t *d_q, uint64_t *d_lo,
                            uint64_t *d_hi, bpt_stream s);

#ifdef __cplusplus
}
#endif
#endif
