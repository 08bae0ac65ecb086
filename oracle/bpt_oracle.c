/*
 * CPU oracle for the B+tree index search path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, and the CPU
 * baseline the benchmark times beside it. Never linked into the product.
 * Citations are reference file:line (sc25index/B-tree-index-searching).
 */
#include "bpt_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int ora_omp_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static void set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

/* ------------------------------------------------------------------------ */
/* Page tree build: BPTree.cu:73-125 (assignDataToLeaf), BPTree.cu:28-70     */
/* (kernelMaxIntermidatePage) and the host level loop BPTree.cu:218-243.     */
/* ------------------------------------------------------------------------ */
int ora_pt_build(const u64 *keys, const u64 *data, u64 n, u64 P, ora_pt **out) {
    if (!out || n == 0 || P < 2) return -1;
    u64 nleaves = (n + P - 1) / P;
    /* page count: leaves + every level until a single root page */
    u64 total = nleaves, lvl = nleaves;
    do { lvl = (lvl + P - 1) / P; total += lvl; } while (lvl > 1);

    ora_pt *pt = (ora_pt *)calloc(1, sizeof(ora_pt));
    pt->pagesize = P;
    pt->npages = total;
    /* the reference zero-fills its page pool (BPTree.cu:166) */
    pt->key = (u64 *)calloc(total * P, sizeof(u64));
    pt->data = (u64 *)calloc(total * P, sizeof(u64));
    pt->leaf = (uint8_t *)calloc(total, 1);
    pt->size = (int64_t *)calloc(total, sizeof(int64_t));
    pt->max = (u64 *)calloc(total, sizeof(u64));
    pt->min = (u64 *)calloc(total, sizeof(u64));
    pt->next = (int64_t *)calloc(total, sizeof(int64_t));
    pt->prev = (int64_t *)calloc(total, sizeof(int64_t));
    if (!pt->key || !pt->data || !pt->leaf || !pt->size || !pt->max || !pt->min ||
        !pt->next || !pt->prev) { ora_pt_free(pt); return -1; }

    /* leaves: page i holds [i*P, min((i+1)*P, n)); the last leaf holds the
     * remainder (the reference pushes P for every leaf, main.cu:126-129). */
    u64 *prefix = (u64 *)malloc(nleaves * sizeof(u64));
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)nleaves; i++) {
        u64 off = (u64)i * P;
        u64 cnt = (off + P <= n) ? P : n - off;
        memcpy(&pt->key[(u64)i * P], &keys[off], cnt * sizeof(u64));
        memcpy(&pt->data[(u64)i * P], &data[off], cnt * sizeof(u64));
        pt->size[i] = (int64_t)cnt;
        pt->max[i] = keys[off + cnt - 1];
        pt->min[i] = keys[off];
        pt->leaf[i] = 1;
        pt->next[i] = (i == (int64_t)nleaves - 1) ? -1 : i + 1;
        pt->prev[i] = (i == 0) ? -1 : i - 1;
        prefix[i] = (u64)i;
    }

    /* internal levels, bottom-up; separator = child max, data = child id.
     * The page counter advances by the level's node count (the reference
     * advances by min(nodes, 65535), BPTree.cu:224,232). */
    u64 current = nleaves;
    u64 prev_size = nleaves;
    u64 node_size = (nleaves + P - 1) / P;
    u64 levels = 0;
    u64 *inter = NULL;
    for (;;) {
        inter = (u64 *)malloc(node_size * sizeof(u64));
        #pragma omp parallel for schedule(static)
        for (int64_t g = 0; g < (int64_t)node_size; g++) {
            u64 items = ((u64)g != node_size - 1) ? P : prev_size % P;
            if (items == 0 && (u64)g == node_size - 1) items = P;
            u64 page = current + (u64)g;
            for (u64 t = 0; t < items; t++) {
                u64 child = prefix[(u64)g * P + t];
                pt->key[page * P + t] = pt->max[child];
                pt->data[page * P + t] = child;
            }
            inter[g] = page;
            pt->size[page] = (int64_t)items;
            pt->max[page] = pt->max[pt->data[page * P + items - 1]];
            pt->min[page] = pt->min[pt->data[page * P]];
            pt->next[page] = -1;
            pt->prev[page] = -1;
            pt->leaf[page] = 0;
        }
        current += node_size;
        levels++;
        free(prefix);
        prefix = inter;
        if (node_size == 1) break;
        prev_size = node_size;
        node_size = (node_size + P - 1) / P;
    }
    pt->root = prefix[0];
    pt->levels = levels;
    free(prefix);
    *out = pt;
    return 0;
}

void ora_pt_free(ora_pt *pt) {
    if (!pt) return;
    free(pt->key); free(pt->data); free(pt->leaf); free(pt->size);
    free(pt->max); free(pt->min); free(pt->next); free(pt->prev);
    free(pt);
}

/* Descent shared by the point and range kernels of the reference:
 * BPTreeExactSearch.cu:39-68 / BtreeRangeSearch.cu:228-256. Returns the leaf
 * page; *parent receives the last internal page visited. */
static u64 pt_descend(const ora_pt *pt, u64 q, u64 *parent) {
    const u64 P = pt->pagesize;
    u64 page = pt->root;
    for (;;) {
        int64_t found = -1;
        int start = 0, end = (int)pt->size[page] - 1;
        const u64 *k = &pt->key[page * P];
        while (start <= end) {
            int middle = start + (end - start) / 2;
            if (q == k[middle]) { found = middle; break; }
            else if (q < k[middle]) end = middle - 1;
            else start = middle + 1;
        }
        if (found < 0) found = start;
        if (found == pt->size[page]) found = start - 1;
        u64 child = pt->data[page * P + (u64)found];
        if (pt->leaf[child]) { if (parent) *parent = page; return child; }
        page = child;
    }
}

/* BPTreeExactSearch.cu:25-105 */
void ora_pt_point(const ora_pt *pt, const u64 *q, u64 nq, u64 *out, u64 sentinel, int threads) {
    set_threads(threads);
    const u64 P = pt->pagesize;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)nq; i++) {
        u64 key = q[i];
        u64 leaf = pt_descend(pt, key, NULL);
        const u64 *k = &pt->key[leaf * P];
        int64_t found = -1;
        int start = 0, end = (int)pt->size[leaf] - 1;
        while (start <= end) {
            int middle = start + (end - start) / 2;
            if (key == k[middle]) { found = middle; break; }
            else if (key < k[middle]) end = middle - 1;
            else start = middle + 1;
        }
        out[i] = (found < 0) ? sentinel : pt->data[leaf * P + (u64)found];
    }
}

/* BtreeRangeSearch.cu:215-294 */
void ora_pt_range_locate(const ora_pt *pt, const u64 *q, u64 nq, u64 *out) {
    const u64 P = pt->pagesize;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)nq; i++) {
        u64 key = q[i], parent = 0;
        u64 leaf = pt_descend(pt, key, &parent);
        const u64 *k = &pt->key[leaf * P];
        int64_t found = -1;
        int start = 0, end = (int)pt->size[leaf] - 1;
        while (start <= end) {
            int middle = start + (end - start) / 2;
            if (key == k[middle]) { found = middle; break; }
            else if (key < k[middle]) end = middle - 1;
            else start = middle + 1;
        }
        if (found < 0) found = start;
        /* compared against the parent page's size (BtreeRangeSearch.cu:284) */
        if (found == pt->size[parent]) found = start - 1;
        out[i] = ((u64)found << 32) | leaf;
    }
}

/* Slot i of a page as the reference reads it: past key[PAGESIZE-1] the
 * struct continues with data[] (BPTree.h:15-16). */
static u64 pt_key_at(const ora_pt *pt, u64 page, u64 i) {
    const u64 P = pt->pagesize;
    return (i < P) ? pt->key[page * P + i] : pt->data[page * P + (i - P)];
}

/* BtreeRangeSearch.cu:297-374 */
void ora_pt_range_count_refcompat(const ora_pt *pt, const u64 *lo, const u64 *hi, u64 nq, u64 *count) {
    u64 *ls = (u64 *)malloc(nq * sizeof(u64));
    u64 *le = (u64 *)malloc(nq * sizeof(u64));
    ora_pt_range_locate(pt, lo, nq, ls);
    ora_pt_range_locate(pt, hi, nq, le);
    #pragma omp parallel for schedule(static)
    for (int64_t h = 0; h < (int64_t)nq; h++) {
        u64 s = lo[h], e = hi[h], c = 0;
        u64 lstart = ls[h] & 0xffffffffULL, pstart = ls[h] >> 32;
        u64 lend = le[h] & 0xffffffffULL, pend = le[h] >> 32;
        if (lstart != lend) {
            for (int64_t i = (int64_t)pstart; i < pt->size[lstart] - 1; i++) {
                u64 k = pt_key_at(pt, lstart, (u64)i);
                if (k <= e && k >= s) c++;
            }
            int64_t cur = pt->next[lstart];
            while (cur != (int64_t)lend && cur >= 0) {
                for (int64_t i = 0; i < pt->size[cur]; i++) {
                    u64 k = pt_key_at(pt, (u64)cur, (u64)i);
                    if (k <= e && k >= s) c++;
                }
                cur = pt->next[cur];
            }
            for (int64_t i = 0; i < (int64_t)pend && i < pt->size[lend] - 1; i++) {
                u64 k = pt_key_at(pt, lend, (u64)i);
                if (k <= e && k >= s) c++;
            }
        } else {
            for (int64_t i = (int64_t)pstart; i <= (int64_t)pend; i++) {
                u64 k = pt_key_at(pt, lend, (u64)i);
                if (k >= s && k <= e) c++;
            }
        }
        count[h] = c;
    }
    free(ls); free(le);
}

/* ------------------------------------------------------------------------ */
/* Closed forms                                                              */
/* ------------------------------------------------------------------------ */
static u64 lb(const u64 *keys, u64 n, u64 q) {
    u64 lo = 0, hi = n;
    while (lo < hi) {
        u64 mid = lo + (hi - lo) / 2;
        if (keys[mid] < q) lo = mid + 1; else hi = mid;
    }
    return lo;
}
static u64 ub(const u64 *keys, u64 n, u64 q) {
    u64 lo = 0, hi = n;
    while (lo < hi) {
        u64 mid = lo + (hi - lo) / 2;
        if (keys[mid] <= q) lo = mid + 1; else hi = mid;
    }
    return lo;
}

void ora_lower_bound(const u64 *keys, u64 n, const u64 *q, u64 nq, u64 *pos) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)nq; i++) pos[i] = lb(keys, n, q[i]);
}

void ora_point_closed(const u64 *keys, const u64 *vals, u64 n, const u64 *q, u64 nq,
                      u64 *out, u64 sentinel) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)nq; i++) {
        u64 p = lb(keys, n, q[i]);
        out[i] = (p < n && keys[p] == q[i]) ? vals[p] : sentinel;
    }
}

void ora_range_closed(const u64 *keys, u64 n, const u64 *lo, const u64 *hi, u64 nq,
                      u64 *first, u64 *count) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)nq; i++) {
        u64 a = lb(keys, n, lo[i]);
        u64 b = ub(keys, n, hi[i]);
        if (first) first[i] = a;
        count[i] = (lo[i] > hi[i] || b < a) ? 0 : b - a;
    }
}

/* ------------------------------------------------------------------------ */
/* Z-order and grouping keys                                                 */
/* ------------------------------------------------------------------------ */
u64 ora_morton2d(uint32_t x, uint32_t y) {
    u64 z = 0;
    for (int i = 0; i < 32; i++) {
        z |= (u64)((x >> i) & 1u) << (2 * i);
        z |= (u64)((y >> i) & 1u) << (2 * i + 1);
    }
    return z;
}

void ora_demorton2d(u64 z, uint32_t *x, uint32_t *y) {
    uint32_t a = 0, b = 0;
    for (int i = 0; i < 32; i++) {
        a |= (uint32_t)((z >> (2 * i)) & 1u) << i;
        b |= (uint32_t)((z >> (2 * i + 1)) & 1u) << i;
    }
    *x = a; *y = b;
}

/* zOrder2d.cu:129-130: int interval = ceil((float)max / (float)2^32).
 * A quotient of 2^31 or more (max >= 2^63 - 2^38 after float rounding) does
 * not fit the int. The reference's host is x86, whose float-to-int
 * conversion (cvttss2si) then yields INT_MIN; that result is written out
 * here rather than left to the overflow. max == 0 gives 0. */
int ora_zorder_interval(u64 maxval) {
    float r = (float)maxval / 4294967296.0f;
    if (r >= 2147483648.0f) return INT32_MIN;
    return (int)ceilf(r);
}

/* The divisor the reference's kernel applies: the int interval travels as
 * ulli (zOrder2d.cu:57, sign-extended), back to int (:27, low 32 bits), and
 * :29 divides the ulli input by it, which converts the int to ulli again
 * (sign-extended). INT_MIN thus divides by 0xFFFFFFFF80000000, so every
 * quotient is 0 or 1. */
u64 ora_zorder_divisor(u64 maxval) {
    return (u64)(int64_t)ora_zorder_interval(maxval);
}

/* zOrder2d.cu:27-54: quotient truncated to int, then only its low 32 bits
 * enter the interleave (the mask cascade starts with & 0xFFFFFFFF).
 * A zero divisor (max == 0, a division by zero in the reference, whose
 * result CUDA leaves undefined) is taken as quotient 0: every coordinate of
 * that side is 0 then, so the order of the keys does not depend on it. */
void ora_zorder_keys(const u64 *a, const u64 *b, u64 n, u64 *out) {
    u64 ma = 0, mb = 0;
    for (u64 i = 0; i < n; i++) { if (a[i] > ma) ma = a[i]; if (b[i] > mb) mb = b[i]; }
    u64 ix = ora_zorder_divisor(ma), iy = ora_zorder_divisor(mb);
    for (u64 i = 0; i < n; i++)
        out[i] = ora_morton2d((uint32_t)(ix ? a[i] / ix : 0), (uint32_t)(iy ? b[i] / iy : 0));
}

/* grouping.cu:25-29 with BIT_PASS 8, BIT_MASK 0xFF (grouping.cu:15-17) */
int ora_bucket_identify(u64 x, int pass) {
    int maxPass = 64 / 8;
    int shift = 8 * (maxPass - 1 - pass);
    if (shift < 0 || shift >= 64) return 0;
    return (int)((x >> shift) & 0xFF);
}

/* grouping2d.cu:24-31 with BIT_PASS 3, BIT_MASK 0xF (grouping2d.cu:13-16).
 * A shift of 64 or more yields 0, as a 64-bit shift does on the device. */
int ora_bucket2d(u64 a, u64 b, int pass) {
    int maxPass = 64 / 3 + 1;
    int shift = 3 * (maxPass - 1 - pass);
    if (shift < 0 || shift >= 64) return 0;
    int n1 = (int)((a >> shift) & 0xF);
    int n2 = (int)((b >> shift) & 0xF);
    return n1 * n2;
}

static int bitwidth(u64 v) { int b = 0; while (v) { v >>= 1; b++; } return b; }

/* grouping.cu:1736-1744 */
int ora_group_start_pass(u64 maxval) {
    return 8 - (int)ceilf((float)bitwidth(maxval) / 8.0f);
}

/* grouping2d.cu:610-635 */
int ora_group2d_start_pass(u64 maxa, u64 maxb) {
    int a = 16 - (int)ceilf((float)bitwidth(maxa) / 3.0f);
    int b = 16 - (int)ceilf((float)bitwidth(maxb) / 3.0f);
    return a > b ? a : b;
}

/* grouping.cu:855-872 */
u64 ora_merge_groups(const u64 *hist, u64 nb, u64 cap, u64 *sizes, u64 *first_bucket) {
    u64 ng = 0;
    for (u64 j = 0; j < nb; j++) {
        if (hist[j] == 0) continue;
        u64 k = 0, cur = hist[j];
        while (cur < cap && (j + k + 1) < nb) {
            if (cur + hist[j + k + 1] < cap) { cur += hist[j + k + 1]; k++; }
            else break;
        }
        sizes[ng] = cur;
        first_bucket[ng] = j;
        ng++;
        j += k;
    }
    return ng;
}

/* ------------------------------------------------------------------------ */
/* The grouping pass loop and the group lists it returns                     */
/* ------------------------------------------------------------------------ */
/* Bucket of grouped element k at `pass`: 1-D grouping.cu:25-29, 2-D
 * grouping2d.cu:24-31. */
static int grp_bucket(const u64 *a, const u64 *b, const u64 *perm, u64 k, int kind, int pass) {
    u64 i = perm[k];
    return kind == 0 ? ora_bucket_identify(a[i], pass) : ora_bucket2d(a[i], b[i], pass);
}

/* Histogram of the range [lo, lo + cnt) of the grouped order, and a stable
 * counting sort of that range by bucket (the relocate kernels,
 * grouping.cu:115-478,513-770 and grouping2d.cu:106-343, move each group's
 * elements to their bucket's run; order inside a bucket is the reference's
 * atomics' and is not reproduced). */
static void grp_split(const u64 *a, const u64 *b, u64 *perm, u64 *tmp, u64 lo, u64 cnt, int kind, int pass,
                      u64 *hist) {
    u64 pos[256];
    memset(hist, 0, 256 * sizeof(u64));
    for (u64 k = lo; k < lo + cnt; k++) hist[grp_bucket(a, b, perm, k, kind, pass)]++;
    u64 acc = 0;
    for (int j = 0; j < 256; j++) { pos[j] = acc; acc += hist[j]; }
    for (u64 k = lo; k < lo + cnt; k++) tmp[lo + pos[grp_bucket(a, b, perm, k, kind, pass)]++] = perm[k];
    memcpy(perm + lo, tmp + lo, cnt * sizeof(u64));
}

/* The pass loop of genGrouping (grouping.cu:1746-1791), genGroupingWithOnedata
 * (:1567-1612), genGroupingWithTwodata (:1413-1460) and genGrouping2d
 * (grouping2d.cu:635-680). Each pass runs one of two bookkeeping forms:
 *  - fewer than 10 groups, radixGroupGPU (grouping.cu:776-884; 2-D
 *    grouping2d.cu:349-461): empty groups are dropped; groups below the cap
 *    pass through; a larger group is split into its 256 buckets, whose
 *    non-empty runs merge greedily while the running total stays below the
 *    cap (:855-872), with prefixes group start + bucket start (:869); a sole
 *    group whose elements all share one bucket stays as it is (:840-844).
 *  - 10 groups or more, radixGroupSingleKernelGPU (grouping.cu:1288-1363
 *    with the kernel :692-770; 2-D grouping2d.cu:464-542): every group above
 *    the cap is split into 256 bucket counts, every other group is one count
 *    in its first slot, and one greedy merge runs over all groups' slots in
 *    order (:1333-1352), so it may join the tail of one group with the head
 *    of the next; prefixes are running sums.
 * The loop ends after the pass whose largest group is at most the cap, or
 * after the last pass (8 for 1-D, 16 for 2-D). An empty list ends it too (the
 * reference reads groupList[0] of the empty list there, n == 0).
 * The small groups of the first form are copied from offset 0 in the
 * reference (grouping.cu:789), which loses data but not counts; the lists do
 * not depend on it and the data here moves correctly.
 * perm: in/out, n entries, the grouped order as indices into a (and b).
 * list, prefix: room for n + 1 entries. Returns the number of groups. */
u64 ora_group_lists(const u64 *a, const u64 *b, u64 n, u64 cap, int kind, int start_pass, u64 *perm, u64 *list,
                    u64 *prefix) {
    const int end_pass = kind == 0 ? 8 : 16;
    u64 ng = 1;
    list[0] = n;
    prefix[0] = 0;
    u64 *tmp = (u64 *)malloc((n ? n : 1) * sizeof(u64));
    u64 *nl = (u64 *)malloc((n + 1) * sizeof(u64));
    u64 *np = (u64 *)malloc((n + 1) * sizeof(u64));
    u64 hist[256];
    for (int pass = start_pass; pass < end_pass; pass++) {
        u64 m = 0;
        if (ng < 10) {
            for (u64 i = 0; i < ng; i++) {
                if (list[i] == 0) continue;
                if (list[i] < cap) { nl[m] = list[i]; np[m++] = prefix[i]; continue; }
                u64 h0[256];
                memset(h0, 0, sizeof h0);
                for (u64 k = prefix[i]; k < prefix[i] + list[i]; k++) h0[grp_bucket(a, b, perm, k, kind, pass)]++;
                int valid = 0;
                for (int j = 0; j < 256; j++) valid += h0[j] != 0;
                if (valid == 1 && ng == 1) { nl[m] = list[i]; np[m++] = prefix[i]; continue; }
                grp_split(a, b, perm, tmp, prefix[i], list[i], kind, pass, hist);
                u64 sizes[256], first[256], pre[256], acc = 0;
                for (int j = 0; j < 256; j++) { pre[j] = acc; acc += hist[j]; }
                u64 g = ora_merge_groups(hist, 256, cap, sizes, first);
                for (u64 t = 0; t < g; t++) { nl[m] = sizes[t]; np[m++] = prefix[i] + pre[first[t]]; }
            }
        } else {
            u64 *slots = (u64 *)calloc(256 * ng, sizeof(u64));
            for (u64 i = 0; i < ng; i++) {
                if (list[i] > cap) grp_split(a, b, perm, tmp, prefix[i], list[i], kind, pass, slots + 256 * i);
                else slots[256 * i] = list[i];
            }
            u64 *sizes = (u64 *)malloc(256 * ng * sizeof(u64)), *first = (u64 *)malloc(256 * ng * sizeof(u64));
            u64 g = ora_merge_groups(slots, 256 * ng, cap, sizes, first), run = 0;
            for (u64 t = 0; t < g; t++) { nl[m] = sizes[t]; np[m++] = run; run += sizes[t]; }
            free(slots); free(sizes); free(first);
        }
        memcpy(list, nl, m * sizeof(u64));
        memcpy(prefix, np, m * sizeof(u64));
        ng = m;
        if (ng == 0) break;
        u64 mx = 0;
        for (u64 i = 0; i < ng; i++) if (list[i] > mx) mx = list[i];
        if (mx <= cap) break;
    }
    free(tmp); free(nl); free(np);
    return ng;
}

/* ------------------------------------------------------------------------ */
/* Rectangles                                                                */
/* ------------------------------------------------------------------------ */
static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

void ora_rect_bruteforce(const uint32_t *xs, const uint32_t *ys, const uint32_t *pid, u64 npts,
                         const uint32_t *rects, u64 nr, u64 *counts,
                         const u64 *offsets, uint32_t *ids, int threads) {
    set_threads(threads);
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < (int64_t)nr; r++) {
        uint32_t x0 = rects[4 * r], y0 = rects[4 * r + 1], x1 = rects[4 * r + 2], y1 = rects[4 * r + 3];
        u64 c = 0;
        for (u64 p = 0; p < npts; p++) {
            if (xs[p] >= x0 && xs[p] <= x1 && ys[p] >= y0 && ys[p] <= y1) {
                if (ids) ids[offsets[r] + c] = pid[p];
                c++;
            }
        }
        counts[r] = c;
        if (ids && c > 1) qsort(&ids[offsets[r]], c, sizeof(uint32_t), cmp_u32);
    }
}

/* Quadtree decomposition of a rectangle into aligned Morton blocks: a block
 * of side 2^k at (cx, cy) (both multiples of 2^k) is the contiguous code
 * interval [Z(cx,cy), Z(cx,cy) + 4^k - 1]. Blocks inside the rectangle are
 * taken whole; straddling blocks are split, and below `min_side` scanned
 * with a coordinate filter. */
typedef struct { const u64 *keys; const uint32_t *pid; u64 n;
                 uint32_t x0, y0, x1, y1; u64 c; uint32_t *out; } rect_ctx;

static void rect_scan(rect_ctx *cx, u64 zlo, u64 zhi, int filter) {
    u64 p = lb(cx->keys, cx->n, zlo);
    for (; p < cx->n && cx->keys[p] <= zhi; p++) {
        if (filter) {
            uint32_t x, y;
            ora_demorton2d(cx->keys[p], &x, &y);
            if (x < cx->x0 || x > cx->x1 || y < cx->y0 || y > cx->y1) continue;
        }
        if (cx->out) cx->out[cx->c] = cx->pid[p];
        cx->c++;
    }
}

static void rect_rec(rect_ctx *cx, u64 bx, u64 by, int k) {
    u64 side = 1ULL << k;
    u64 ex = bx + side - 1, ey = by + side - 1;
    if (ex < cx->x0 || bx > cx->x1 || ey < cx->y0 || by > cx->y1) return;
    u64 zlo = ora_morton2d((uint32_t)bx, (uint32_t)by);
    u64 zhi = (k == 32) ? ~0ULL : zlo + ((1ULL << (2 * k)) - 1);
    int inside = bx >= cx->x0 && ex <= cx->x1 && by >= cx->y0 && ey <= cx->y1;
    if (inside) { rect_scan(cx, zlo, zhi, 0); return; }
    /* stop splitting when the block has few keys */
    u64 a = lb(cx->keys, cx->n, zlo);
    if (a >= cx->n || cx->keys[a] > zhi) return;
    u64 b = (zhi == ~0ULL) ? cx->n : lb(cx->keys, cx->n, zhi + 1);
    if (b - a <= 64 || k == 0) { rect_scan(cx, zlo, zhi, 1); return; }
    u64 h = side >> 1;
    rect_rec(cx, bx, by, k - 1);
    rect_rec(cx, bx + h, by, k - 1);
    rect_rec(cx, bx, by + h, k - 1);
    rect_rec(cx, bx + h, by + h, k - 1);
}

void ora_rect_morton(const u64 *keys, const uint32_t *pid, u64 n,
                     const uint32_t *rects, u64 nr, u64 *counts,
                     const u64 *offsets, uint32_t *ids, int threads) {
    set_threads(threads);
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < (int64_t)nr; r++) {
        rect_ctx cx;
        cx.keys = keys; cx.pid = pid; cx.n = n;
        cx.x0 = rects[4 * r]; cx.y0 = rects[4 * r + 1];
        cx.x1 = rects[4 * r + 2]; cx.y1 = rects[4 * r + 3];
        cx.c = 0;
        cx.out = ids ? &ids[offsets[r]] : NULL;
        if (cx.x0 <= cx.x1 && cx.y0 <= cx.y1) rect_rec(&cx, 0, 0, 32);
        counts[r] = cx.c;
        if (ids && cx.c > 1) qsort(cx.out, cx.c, sizeof(uint32_t), cmp_u32);
    }
}

/* Tropf & Herzog (1981), "Multidimensional range search in dynamically
 * balanced trees": BIGMIN for two interleaved dimensions. */
static u64 dim_lower(int i) {
    u64 m = (i & 1) ? 0xAAAAAAAAAAAAAAAAULL : 0x5555555555555555ULL;
    return (i == 0) ? 0 : (m & ((1ULL << i) - 1));
}
u64 ora_bigmin(u64 z, u64 zmin, u64 zmax) {
    u64 big = 0;
    for (int i = 63; i >= 0; i--) {
        u64 bit = 1ULL << i;
        int v = (z & bit) != 0, lo = (zmin & bit) != 0, hi = (zmax & bit) != 0;
        u64 low = dim_lower(i);
        if (!v && !lo && hi) {
            big = (zmin | bit) & ~low;
            zmax = (zmax & ~bit) | low;
        } else if (!v && lo && hi) {
            return zmin;
        } else if (v && !lo && !hi) {
            return big;
        } else if (v && !lo && hi) {
            zmin = (zmin | bit) & ~low;
        }
        /* (0,0,0) and (1,1,1): continue; (x,1,0) cannot occur for zmin <= zmax */
    }
    return big;
}

/* ------------------------------------------------------------------------ */
/* Learned-index MLP, FP64                                                   */
/* ------------------------------------------------------------------------ */
void ora_mlp_forward(const double *p, int H, const double *x, u64 m, double *y) {
    const double *w1 = p, *b1 = p + H, *W2 = p + 2 * H, *b2 = p + 2 * H + H * H;
    const double *w3 = b2 + H, *b3 = w3 + H;
    double *h1 = (double *)malloc(sizeof(double) * H);
    for (u64 s = 0; s < m; s++) {
        for (int j = 0; j < H; j++) { double a = w1[j] * x[s] + b1[j]; h1[j] = a > 0 ? a : 0; }
        double out = b3[0];
        for (int j = 0; j < H; j++) {
            double a = b2[j];
            for (int k = 0; k < H; k++) a += W2[j * H + k] * h1[k];
            out += w3[j] * (a > 0 ? a : 0);
        }
        y[s] = out;
    }
    free(h1);
}

double ora_mlp_loss_grad(const double *p, int H, const double *x, const double *t, u64 m,
                         double *g) {
    const double *w1 = p, *b1 = p + H, *W2 = p + 2 * H, *b2 = p + 2 * H + H * H;
    const double *w3 = b2 + H, *b3 = w3 + H;
    int np = 3 * H + H * H + H + 1;
    memset(g, 0, sizeof(double) * np);
    double *gw1 = g, *gb1 = g + H, *gW2 = g + 2 * H, *gb2 = g + 2 * H + H * H;
    double *gw3 = gb2 + H, *gb3 = gw3 + H;
    double *a1 = (double *)malloc(sizeof(double) * H), *h1 = (double *)malloc(sizeof(double) * H);
    double *a2 = (double *)malloc(sizeof(double) * H), *h2 = (double *)malloc(sizeof(double) * H);
    double *d2 = (double *)malloc(sizeof(double) * H);
    double loss = 0;
    for (u64 s = 0; s < m; s++) {
        for (int j = 0; j < H; j++) { a1[j] = w1[j] * x[s] + b1[j]; h1[j] = a1[j] > 0 ? a1[j] : 0; }
        double y = b3[0];
        for (int j = 0; j < H; j++) {
            double a = b2[j];
            for (int k = 0; k < H; k++) a += W2[j * H + k] * h1[k];
            a2[j] = a; h2[j] = a > 0 ? a : 0;
            y += w3[j] * h2[j];
        }
        double e = y - t[s];
        loss += e * e;
        double dy = 2.0 * e / (double)m;
        gb3[0] += dy;
        for (int j = 0; j < H; j++) {
            gw3[j] += dy * h2[j];
            d2[j] = (a2[j] > 0) ? dy * w3[j] : 0;
            gb2[j] += d2[j];
        }
        for (int k = 0; k < H; k++) {
            double d1 = 0;
            for (int j = 0; j < H; j++) { gW2[j * H + k] += d2[j] * h1[k]; d1 += d2[j] * W2[j * H + k]; }
            d1 = (a1[k] > 0) ? d1 : 0;
            gb1[k] += d1;
            gw1[k] += d1 * x[s];
        }
    }
    free(a1); free(h1); free(a2); free(h2); free(d2);
    return loss / (double)m;
}
