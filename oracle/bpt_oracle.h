/*
 * CPU oracle for the B+tree index search path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests, the smoke
 * check and the CPU-baseline leg of the benchmark may call it; the product
 * library never links or calls it.
 *
 * Every function restates the algorithm of the reference file:line it cites
 * (paths relative to the sc25index/B-tree-index-searching tree), or a closed
 * form the reference's semantics reduce to. All arithmetic is integer
 * (uint64), keys of narrower width are passed widened.
 */
#ifndef BPT_ORACLE_H
#define BPT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint64_t u64;

/* ---------------- reference page tree (runtime PAGESIZE) ---------------- */
typedef struct ora_pt {
    u64 pagesize;      /* PAGESIZE, BPTree.h:8 (5000 there; runtime here)   */
    u64 npages;        /* pages actually used                               */
    u64 *key;          /* npages * pagesize, BTPage.key  (BPTree.h:15)      */
    u64 *data;         /* npages * pagesize, BTPage.data (BPTree.h:16)      */
    uint8_t *leaf;     /* BTPage.leafFlag                                   */
    int64_t *size;     /* BTPage.size                                       */
    u64 *max, *min;    /* BTPage.max / .min                                 */
    int64_t *next, *prev;
    u64 root;          /* root page id (genBPlusTree's dev_btree_out)       */
    u64 levels;        /* internal levels above the leaves                  */
} ora_pt;

/* Bulk build, BPTree.cu:73-125 (leaves) and BPTree.cu:28-70,218-243
 * (internal levels, separator = child max, child page id in data[]).
 * Leaves hold consecutive PAGESIZE chunks; the last leaf holds the remainder
 * (the reference pushes PAGESIZE for every leaf, main.cu:126-129).
 * Returns 0 on success, -1 on bad input (n == 0, pagesize < 2). */
int  ora_pt_build(const u64 *keys, const u64 *data, u64 n, u64 pagesize, ora_pt **out);
void ora_pt_free(ora_pt *pt);

/* Plain point search, BPTreeExactSearch.cu:25-105: per query, descend by
 * binary search over child maxima (break on equality, else lower bound,
 * clamp to the last child), then exact binary search in the leaf.
 * out[i] = data of the match, or `sentinel` (INVALID_DATA, BPTree.h:12).
 * `threads` > 0 sets the OpenMP thread count; 0 keeps the default. */
void ora_pt_point(const ora_pt *pt, const u64 *q, u64 nq, u64 *out, u64 sentinel, int threads);

/* Range step 2, BtreeRangeSearch.cu:215-294: (pos << 32) | leaf page id,
 * pos clamped against the parent page's size as the reference does. */
void ora_pt_range_locate(const ora_pt *pt, const u64 *q, u64 nq, u64 *out);

/* Range step 3 count, BtreeRangeSearch.cu:297-374, including the
 * reference's boundary behaviour (start leaf skips its last slot, end leaf
 * scans i < endPos && i < size-1, same-leaf case inclusive). Documentation
 * of the reference's numbers only; the canonical count is ora_range_closed. */
void ora_pt_range_count_refcompat(const ora_pt *pt, const u64 *lo, const u64 *hi, u64 nq, u64 *count);

/* ---------------- closed forms (canonical semantics) ---------------- */
/* pos[i] = lower_bound(keys, q[i]) */
void ora_lower_bound(const u64 *keys, u64 n, const u64 *q, u64 nq, u64 *pos);
/* out[i] = vals[lb] if keys[lb] == q[i] else sentinel (plain-search semantics
 * of BPTreeExactSearch.cu:25-105 for sorted keys). */
void ora_point_closed(const u64 *keys, const u64 *vals, u64 n, const u64 *q, u64 nq,
                      u64 *out, u64 sentinel);
/* Inclusive range [lo, hi]: first = lower_bound(lo), count = ub(hi) - lb(lo)
 * (0 when lo > hi). */
void ora_range_closed(const u64 *keys, u64 n, const u64 *lo, const u64 *hi, u64 nq,
                      u64 *first, u64 *count);

/* ---------------- Z-order / grouping keys ---------------- */
/* Morton code, x in even bits, y in odd bits (zOrder2d.cu:34-51), written as
 * the bit-by-bit definition rather than the mask cascade. */
u64  ora_morton2d(uint32_t x, uint32_t y);
void ora_demorton2d(u64 z, uint32_t *x, uint32_t *y);
/* Z-order grouping key of a (start, end) pair, zOrder2d.cu:27-54 with the
 * intervals of zOrder2d.cu:125-130 (float ceil of max / 2^32, as an int:
 * INT_MIN when the quotient does not fit, 0 when max == 0) and the divisor
 * that int becomes in the kernel (zOrder2d.cu:27-29,57). */
void ora_zorder_keys(const u64 *a, const u64 *b, u64 n, u64 *out);
int  ora_zorder_interval(u64 maxval);
u64  ora_zorder_divisor(u64 maxval);
/* grouping.cu:25-29 (8-bit MSD bucket of pass p) */
int  ora_bucket_identify(u64 x, int pass);
/* grouping2d.cu:24-31 (product of two 4-bit nibbles at a 3-bit stride) */
int  ora_bucket2d(u64 a, u64 b, int pass);
/* First pass of the 1D grouping, grouping.cu:1736-1744 */
int  ora_group_start_pass(u64 maxval);
/* First pass of the 2D grouping, grouping2d.cu:610-635 */
int  ora_group2d_start_pass(u64 maxa, u64 maxb);
/* Greedy merge of adjacent non-empty buckets while the running total stays
 * below cap, grouping.cu:855-872. Writes group sizes and bucket starts;
 * returns the number of groups. */
u64  ora_merge_groups(const u64 *hist, u64 nbuckets, u64 cap, u64 *sizes, u64 *first_bucket);

/* The grouping pass loop of genGrouping / genGroupingWithOnedata /
 * genGroupingWithTwodata (grouping.cu:1413-1460,1567-1612,1746-1791; kind 0,
 * 8-bit buckets, passes start_pass..7) and genGrouping2d
 * (grouping2d.cu:635-680; kind 1, nibble-product buckets, passes
 * start_pass..15), with both bookkeeping forms (grouping.cu:776-884 and
 * :1288-1363). perm (n entries, in/out) is the grouped order as indices into
 * a (and b for kind 1); list and prefix get the group sizes and starts and
 * need room for n + 1 entries. Returns the number of groups. */
u64  ora_group_lists(const u64 *a, const u64 *b, u64 n, u64 cap, int kind, int start_pass, u64 *perm,
                     u64 *list, u64 *prefix);

/* ---------------- rectangles (no reference counterpart) ---------------- */
/* Brute force: for each rect (x0,y0,x1,y1 inclusive) count points inside and,
 * when ids != NULL, write their ids ascending at offsets[i]. */
void ora_rect_bruteforce(const uint32_t *xs, const uint32_t *ys, const uint32_t *pid, u64 npts,
                         const uint32_t *rects, u64 nr, u64 *counts,
                         const u64 *offsets, uint32_t *ids, int threads);
/* Same answer through the sorted Morton keys: scan [Z(x0,y0), Z(x1,y1)],
 * filter by rectangle, sort ids. keys sorted ascending, ids parallel. */
void ora_rect_morton(const u64 *keys, const uint32_t *pid, u64 n,
                     const uint32_t *rects, u64 nr, u64 *counts,
                     const u64 *offsets, uint32_t *ids, int threads);
/* Tropf-Herzog BIGMIN: smallest Morton code > z inside [zmin, zmax]'s box. */
u64  ora_bigmin(u64 z, u64 zmin, u64 zmax);

/* ---------------- learned index (FP64 restatement) ---------------- */
/* 1-H-H-1 MLP with ReLU: y = w3 . relu(W2 relu(w1 x + b1) + b2) + b3.
 * params layout: w1[H] b1[H] W2[H*H] (row j = output j) b2[H] w3[H] b3[1].
 * Loss = mean over m samples of (y - t)^2. Gradient in grad (same layout). */
double ora_mlp_loss_grad(const double *params, int H, const double *x, const double *t, u64 m,
                         double *grad);
void   ora_mlp_forward(const double *params, int H, const double *x, u64 m, double *y);

int ora_omp_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
