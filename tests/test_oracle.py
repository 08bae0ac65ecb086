"""CPU tests of the oracle itself: every restated function is pinned by a
closed form, brute force, an invariant or a committed golden vector, so that
a plausible mistake (dropped term, wrong index, swapped operand) fails."""
import itertools
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SENT64 = 0xFFFFFFFFFFFFFFFF


def _keys(rng, n, space=1 << 32):
    return np.unique(rng.integers(0, space, size=n, dtype=np.uint64))


@pytest.mark.parametrize("n,P", [(1, 2), (5, 2), (31, 32), (32, 32), (33, 32), (1000, 32),
                                 (1024, 32), (5000, 5000), (5001, 5000), (20000, 7)])
def test_page_tree_point_equals_closed_form(oracle, n, P):
    rng = np.random.default_rng(n * 7 + P)
    k = np.sort(rng.choice(1 << 20, size=n, replace=False).astype(np.uint64))
    v = rng.integers(0, 1 << 60, size=n, dtype=np.uint64)
    pt = oracle.PageTree(k, v, P)
    q = np.concatenate([k, rng.integers(0, 1 << 20, size=2000, dtype=np.uint64),
                        np.array([0, (1 << 20) + 5, SENT64], dtype=np.uint64)])
    got = pt.point(q, SENT64)
    # independent closed form in numpy
    pos = np.searchsorted(k, q, side="left")
    hit = (pos < n) & (k[np.minimum(pos, n - 1)] == q)
    want = np.where(hit, v[np.minimum(pos, n - 1)], np.uint64(SENT64))
    assert np.array_equal(got, want)
    assert np.array_equal(oracle.point_closed(k, v, q, SENT64), want)


@pytest.mark.parametrize("n,P", [(1, 32), (32, 32), (33, 32), (32 * 32, 32), (32 * 32 + 1, 32),
                                 (100000, 32), (12345, 5000)])
def test_page_tree_shape(oracle, n, P):
    k = np.arange(n, dtype=np.uint64) * 3
    pt = oracle.PageTree(k, k, P)
    # height = 1 + ceil(log_P(#leaves)) pages from root to leaf; levels above leaves:
    leaves = -(-n // P)
    lv, c = 0, leaves
    while True:
        c = -(-c // P)
        lv += 1
        if c == 1:
            break
    assert pt.levels == lv
    # pages = leaves + sum of level sizes; root is the last page
    assert pt.root == pt.npages - 1


def _refcompat_formula(k, P, lo, hi):
    """Set form of the reference's range count (BtreeRangeSearch.cu:297-374)
    for unique keys, lo <= hi, a two-level page tree: the count of keys in
    [lo, hi] among the slots it scans. Leaf positions are lower bounds
    clamped against the PARENT page's size (BtreeRangeSearch.cu:284, Q4):
    a position equal to the parent's child count becomes position - 1.
    Different leaves: start leaf slots [sp, size-1) (its last slot is
    skipped), every middle leaf whole, end leaf slots [0, min(ep, size-1)).
    Same leaf: slots [sp, ep] inclusive."""
    n = len(k)
    nleaves = -(-n // P)
    sizes = [min(P, n - j * P) for j in range(nleaves)]
    maxes = np.array([k[j * P + sizes[j] - 1] for j in range(nleaves)])

    def locate(q):
        leaf = min(int(np.searchsorted(maxes, q, "left")), nleaves - 1)
        keys = k[leaf * P: leaf * P + sizes[leaf]]
        pos = int(np.searchsorted(keys, q, "left"))
        parent = leaf // P
        psize = min(P, nleaves - parent * P)
        if pos == psize:
            pos -= 1
        return leaf, pos

    out = []
    for a, b in zip(lo, hi):
        ls, sp = locate(a)
        le, ep = locate(b)
        if ls != le:
            slots = list(range(ls * P + sp, ls * P + sizes[ls] - 1))
            for j in range(ls + 1, le):
                slots += list(range(j * P, j * P + sizes[j]))
            slots += list(range(le * P, le * P + min(ep, sizes[le] - 1)))
        else:
            slots = list(range(ls * P + sp, ls * P + ep + 1))
        vals = k[[s for s in slots if s < n]] if slots else np.empty(0, dtype=np.uint64)
        out.append(int(((vals >= a) & (vals <= b)).sum()))
    return np.array(out, dtype=np.uint64)


def test_refcompat_range_count_matches_formula(oracle):
    rng = np.random.default_rng(3)
    P = 8
    k = np.sort(rng.choice(10000, size=200, replace=False).astype(np.uint64))
    pt = oracle.PageTree(k, k, P)
    a = rng.integers(0, 10000, size=400, dtype=np.uint64)
    b = a + rng.integers(0, 3000, size=400, dtype=np.uint64)
    # keep hi inside the key range so the end-leaf clamp (Q4) never triggers
    b = np.minimum(b, k[-2])
    a = np.minimum(a, b)
    assert pt.levels == 2
    got = pt.range_count_refcompat(a, b)
    want = _refcompat_formula(k, P, a, b)
    assert np.array_equal(got, want)
    # and it differs from the canonical inclusive count where leaves differ
    _, canon = oracle.range_closed(k, a, b)
    assert (got <= canon).all() and (got < canon).any()


def test_refcompat_range_count_end_leaf_clamp(oracle):
    """The last leaf of a page tree of PAGESIZE 8, with lower bounds before it
    (so start and end leaves differ; with both in the last leaf the reference
    reads past the leaf). Counted against the canonical inclusive count, the
    reference's loops skip the start leaf's last slot (Q3) and stop before the
    end position, pulled back by one when it equals the parent page's child
    count (Q4, BtreeRangeSearch.cu:284); the two set restatements (C oracle and
    _refcompat_formula) must agree.
    * 147 keys: upper bounds past the largest key. The descent clamps to the
      last leaf (3 keys, its parent has 3 leaves): position 3, pulled back to
      2, and the end leaf's last slot is not scanned either way: canonical - 2.
    * 136 keys: the last parent has 1 leaf (8 keys). Upper bounds between that
      leaf's first and second keys land on position 1 = the parent's size,
      pulled back to 0: the leaf's first key, inside the range, is dropped:
      canonical - 2 again, one of them by Q4 alone (without it the end slots
      would be [0, 1))."""
    rng = np.random.default_rng(11)
    P = 8
    for n, hi_of in ((147, lambda k, m: k[-1] + rng.integers(1, 5000, size=m, dtype=np.uint64)),
                     (136, lambda k, m: rng.integers(int(k[128]) + 1, int(k[129]), size=m, dtype=np.uint64))):
        k = np.sort(rng.choice(100000, size=n, replace=False).astype(np.uint64))
        pt = oracle.PageTree(k, k, P)
        assert pt.levels == 2
        second_last_max = k[(len(k) - 1) // P * P - 1]
        a = rng.integers(0, int(second_last_max) + 1, size=300, dtype=np.uint64)
        b = hi_of(k, 300)
        got = pt.range_count_refcompat(a, b)
        assert np.array_equal(got, _refcompat_formula(k, P, a, b)), n
        _, canon = oracle.range_closed(k, a, b)
        assert np.array_equal(got, canon - 2), n


def test_range_closed_bruteforce(oracle):
    rng = np.random.default_rng(4)
    k = np.sort(rng.integers(0, 500, size=300, dtype=np.uint64))  # with duplicates
    lo = rng.integers(0, 520, size=500, dtype=np.uint64)
    hi = rng.integers(0, 520, size=500, dtype=np.uint64)
    first, cnt = oracle.range_closed(k, lo, hi)
    for i in range(len(lo)):
        sel = np.nonzero((k >= lo[i]) & (k <= hi[i]))[0]
        assert cnt[i] == len(sel)
        if len(sel):
            assert first[i] == sel[0]
    assert np.array_equal(oracle.lower_bound(k, lo), np.searchsorted(k, lo, "left").astype(np.uint64))


def test_morton_definition(oracle):
    rng = np.random.default_rng(5)
    assert oracle.morton2d(1, 0) == 1 and oracle.morton2d(0, 1) == 2
    assert oracle.morton2d(0xFFFFFFFF, 0) == 0x5555555555555555
    assert oracle.morton2d(0, 0xFFFFFFFF) == 0xAAAAAAAAAAAAAAAA
    for _ in range(200):
        x, y = (int(v) for v in rng.integers(0, 1 << 32, size=2))
        z = oracle.morton2d(x, y)
        # pdep-style reference: bit i of x at 2i, bit i of y at 2i+1
        want = sum(((x >> i) & 1) << (2 * i) | ((y >> i) & 1) << (2 * i + 1) for i in range(32))
        assert z == want
        assert oracle.demorton2d(z) == (x, y)


def test_morton_numpy_cascade_matches(oracle, wl):
    rng = np.random.default_rng(6)
    x = rng.integers(0, 1 << 32, size=1000, dtype=np.uint64).astype(np.uint32)
    y = rng.integers(0, 1 << 32, size=1000, dtype=np.uint64).astype(np.uint32)
    z = wl.morton_np(x, y)
    assert all(int(z[i]) == oracle.morton2d(int(x[i]), int(y[i])) for i in range(1000))


def test_zorder_keys_reference_semantics(oracle):
    # zOrder2d.cu:129-130: interval = ceil((float)max / 2^32), an int; values
    # below 2^32 give interval 1, so the key is Morton(a, b) exactly.
    assert oracle.zorder_interval(0xFFFFFFFF) == 1          # float(2^32-1) rounds to 2^32
    assert oracle.zorder_interval(1 << 33) == 2
    assert oracle.zorder_interval((1 << 33) + 1) == 2       # float rounding
    assert oracle.zorder_interval(3 << 40) == 3 << 8
    assert oracle.zorder_interval((1 << 62)) == 1 << 30
    # max == 0: 0 / 2^32 = 0, so the interval is 0 (a division by zero follows
    # in the reference's kernel; the oracle takes every quotient as 0)
    assert oracle.zorder_interval(0) == 0 and oracle.zorder_divisor(0) == 0
    # float(max) >= 2^63 makes the quotient >= 2^31, which the int cannot hold:
    # x86's cvttss2si gives INT_MIN. Floats in [2^62, 2^63) are 2^39 apart, so
    # 2^63 - 2^38 is the midpoint below 2^63 and rounds up to it (ties to even).
    INT_MIN = -(1 << 31)
    assert oracle.zorder_interval(1 << 63) == INT_MIN
    assert oracle.zorder_interval((1 << 64) - 1) == INT_MIN   # float rounds to 2^64
    assert oracle.zorder_interval((1 << 63) - (1 << 38)) == INT_MIN
    assert oracle.zorder_interval((1 << 63) - (1 << 38) - 1) == (1 << 31) - (1 << 7)
    # the int travels as ulli (zOrder2d.cu:57, sign-extended), back to int
    # (:27) and divides the ulli input (:29), which sign-extends it again
    assert oracle.zorder_divisor(1 << 63) == 0xFFFFFFFF80000000
    assert oracle.zorder_divisor(3 << 40) == 3 << 8
    a = np.array([5, 7, 0xFFFFFFFF], dtype=np.uint64)
    b = np.array([9, 0, 3], dtype=np.uint64)
    got = oracle.zorder_keys(a, b)
    assert [int(v) for v in got] == [oracle.morton2d(int(x), int(y)) for x, y in zip(a, b)]
    a2 = np.array([1 << 34, 3 << 33, 5], dtype=np.uint64)
    got2 = oracle.zorder_keys(a2, b)
    iv = oracle.zorder_interval(int(a2.max()))
    assert iv == 6  # max = 3 * 2^33 = 6 * 2^32
    assert [int(v) for v in got2] == [oracle.morton2d(int(x) // iv, int(y)) for x, y in zip(a2, b)]
    # max >= 2^63: every quotient is 1 at or above 0xFFFFFFFF80000000, else 0
    a3 = np.array([SENT64, 0xFFFFFFFF80000000, 0xFFFFFFFF7FFFFFFF, 5], dtype=np.uint64)
    b3 = np.array([9, 0, 3, 1], dtype=np.uint64)
    got3 = oracle.zorder_keys(a3, b3)
    assert [int(v) for v in got3] == [oracle.morton2d(1, 9), oracle.morton2d(1, 0),
                                      oracle.morton2d(0, 3), oracle.morton2d(0, 1)]
    # a side that is all zeros: coordinate 0 for every pair
    got4 = oracle.zorder_keys(np.zeros(3, dtype=np.uint64), b)
    assert [int(v) for v in got4] == [oracle.morton2d(0, int(y)) for y in b]


def test_bucket_functions(oracle):
    x = 0x0123456789ABCDEF
    assert [oracle.bucket_identify(x, p) for p in range(8)] == [0x01, 0x23, 0x45, 0x67, 0x89, 0xAB, 0xCD, 0xEF]
    # grouping2d.cu:24-31: nibble at shift 3*(21-pass), product of the two
    a, b = 0xF << 45, 0x3 << 45          # pass 6 -> shift 45
    assert oracle.bucket2d(a, b, 6) == 15 * 3
    assert oracle.bucket2d(a, b, 5) == 1 * 0            # shift 48: nibbles 1 and 0
    assert oracle.bucket2d(1 << 63, 1 << 63, 0) == 1   # shift 63: top bit
    assert oracle.bucket2d(~0 & SENT64, SENT64, -3) == 0  # shift >= 64 -> 0
    assert oracle.group_start_pass(0xFF) == 7 and oracle.group_start_pass(0x1FF) == 6
    assert oracle.group_start_pass(SENT64) == 0
    # 16 - ceil(64/3) = -6 for a 64-bit max, 16 - ceil(3/3) = 15 for 7: the larger wins
    assert oracle.group2d_start_pass(SENT64, 7) == 15
    assert oracle.group2d_start_pass(SENT64, SENT64) == -6


def test_merge_groups_closed_form(oracle):
    hist = np.array([0, 5, 5, 0, 30, 1, 1, 40, 0, 2], dtype=np.uint64)
    sizes, first = oracle.merge_groups(hist, 12)
    # greedy: from each non-empty bucket, absorb following buckets while the
    # running total stays below the cap
    assert sizes.tolist() == [10, 30, 2, 40, 2] and first.tolist() == [1, 4, 5, 7, 9]
    assert int(sizes.sum()) == int(hist.sum())


def test_bigmin_bruteforce(oracle):
    # 4-bit coordinates, 8-bit codes: brute force the next in-box code
    rng = np.random.default_rng(7)
    for _ in range(300):
        x0, x1 = sorted(int(v) for v in rng.integers(0, 16, size=2))
        y0, y1 = sorted(int(v) for v in rng.integers(0, 16, size=2))
        zmin, zmax = oracle.morton2d(x0, y0), oracle.morton2d(x1, y1)
        inbox = sorted(oracle.morton2d(x, y) for x in range(x0, x1 + 1) for y in range(y0, y1 + 1))
        for z in range(zmin, zmax + 1):
            x, y = oracle.demorton2d(z)
            if x0 <= x <= x1 and y0 <= y <= y1:
                continue
            nxt = [c for c in inbox if c > z]
            if nxt:
                assert oracle.bigmin(z, zmin, zmax) == nxt[0], (x0, y0, x1, y1, z)


def test_rect_morton_equals_bruteforce(oracle, wl):
    x, y, ids, keys, kid = wl.points_2d(20000, seed=1)
    r = wl.rects(300, seed=1, lo_exp=20, hi_exp=30)
    c1, o1, i1 = oracle.rect_bruteforce(x, y, ids, r)
    c2, o2, i2 = oracle.rect_morton(keys, kid, r)
    assert np.array_equal(c1, c2) and np.array_equal(i1, i2)
    assert c1.sum() > 0


def test_mlp_gradient_finite_differences(oracle):
    rng = np.random.default_rng(8)
    H = 6
    npar = 3 * H + H * H + H + 1
    p = rng.normal(0, 0.5, size=npar)
    x = rng.uniform(0, 1, size=40)
    t = rng.uniform(0, 1, size=40)
    loss, g = oracle.mlp_loss_grad(p, H, x, t)
    eps = 1e-6
    for j in range(npar):
        pp, pm = p.copy(), p.copy()
        pp[j] += eps
        pm[j] -= eps
        fd = (oracle.mlp_loss_grad(pp, H, x, t)[0] - oracle.mlp_loss_grad(pm, H, x, t)[0]) / (2 * eps)
        assert abs(fd - g[j]) <= 1e-6 + 1e-4 * abs(fd), j
    y = oracle.mlp_forward(p, H, x)
    assert abs(np.mean((y - t) ** 2) - loss) < 1e-12


def test_golden_point_fixture(oracle):
    """Committed vectors (tests/golden/gen_golden.py): C1 scaled to 4096 keys."""
    d = np.load(os.path.join(GOLDEN, "point_c1_small.npz"))
    pt = oracle.PageTree(d["keys"].astype(np.uint64), d["vals"].astype(np.uint64), 32)
    got = pt.point(d["queries"].astype(np.uint64), 0xFFFFFFFF)
    assert np.array_equal(got.astype(np.uint32), d["expected"])


def test_golden_range_fixture(oracle):
    d = np.load(os.path.join(GOLDEN, "range_small.npz"))
    first, cnt = oracle.range_closed(d["keys"].astype(np.uint64), d["lo"].astype(np.uint64),
                                     d["hi"].astype(np.uint64))
    assert np.array_equal(cnt, d["count"]) and np.array_equal(first, d["first"])


def test_golden_rect_fixture(oracle):
    d = np.load(os.path.join(GOLDEN, "rect_small.npz"))
    c, o, ids = oracle.rect_morton(d["keys"], d["kid"], d["rects"])
    assert np.array_equal(c, d["count"]) and np.array_equal(ids, d["ids"])
