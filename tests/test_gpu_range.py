"""GPU parity: 1-D inclusive range search (count, first, ids) in every
grouping mode, and 2-D rectangle search over Morton keys, against the oracle."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = (0, 1, 2)  # none, 2D grouping, Z-order


def _ids_expected(vals, first, cnt):
    if len(first) == 0:
        return np.empty(0, dtype=vals.dtype)
    return np.concatenate([vals[int(f):int(f) + int(c)] for f, c in zip(first, cnt)] + [vals[:0]])


@pytest.mark.parametrize("kb,vb", [(4, 4), (8, 8), (8, 4)])
@pytest.mark.parametrize("n", [0, 1, 5, 1000, 100003])
def test_range_matches_oracle(bpt, oracle, kb, vb, n):
    rng = np.random.default_rng(n + kb)
    kdt = np.uint32 if kb == 4 else np.uint64
    vdt = np.uint32 if vb == 4 else np.uint64
    space = (1 << 32) if kb == 4 else (1 << 64)
    keys = np.sort(rng.integers(0, space - 1, size=n, dtype=np.uint64)).astype(kdt)
    if n > 10:
        keys[3:8] = keys[3]       # duplicates spanning slots
        keys[-1] = kdt(space - 1)
    vals = rng.integers(0, (1 << (8 * vb)) - 1, size=n, dtype=np.uint64).astype(vdt)
    nq = 5000
    lo = rng.integers(0, space - 1, size=nq, dtype=np.uint64).astype(kdt)
    w = np.floor(np.exp2(rng.uniform(0, 8 * kb - 4, size=nq))).astype(np.uint64).astype(kdt)
    hi = (lo + w).astype(kdt)
    hi = np.where(hi < lo, kdt(space - 1), hi)   # clamp wrapped upper bounds
    lo[:4] = [0, 5, 0, space - 1]
    hi[:4] = [space - 1, 4, 0, space - 1]        # full range, lo > hi, point ranges
    if n:
        lo[4], hi[4] = keys[0], keys[-1]
    t = bpt.Tree(keys, vals, kb, vb)
    want_first, want_cnt = oracle.range_closed(keys, lo, hi)
    want_ids = _ids_expected(vals, want_first, want_cnt)
    for mode in MODES:
        first, cnt, offs, ids = t.range_search(lo, hi, group_mode=mode)
        assert np.array_equal(cnt, want_cnt), mode
        assert np.array_equal(first, want_first), mode
        assert np.array_equal(ids, want_ids), mode


def test_rect_golden_fixture(bpt):
    d = np.load(os.path.join(GOLDEN, "rect_small.npz"))
    t = bpt.Tree(d["keys"], d["kid"], 8, 4)
    for mode in MODES:
        c, o, ids = t.rect_search(d["rects"], group_mode=mode)
        assert np.array_equal(c, d["count"]) and np.array_equal(ids, d["ids"]), mode


@pytest.mark.parametrize("npts,nr,lo_exp,hi_exp", [(1 << 16, 3000, 8, 32), (1 << 20, 20000, 8, 28),
                                                    (1 << 20, 2000, 20, 32)])
def test_rect_matches_oracle(bpt, oracle, wl, npts, nr, lo_exp, hi_exp):
    x, y, ids, keys, kid = wl.points_2d(npts, seed=npts)
    r = wl.rects(nr, seed=nr, lo_exp=lo_exp, hi_exp=hi_exp)
    r[0] = [0, 0, 0xFFFFFFFF, 0xFFFFFFFF]        # everything
    r[1] = [10, 10, 5, 20]                         # x0 > x1: empty
    r[2] = [x[7], y[7], x[7], y[7]]                # a single point
    want_c, want_o, want_ids = oracle.rect_morton(keys, kid, r)
    # brute force pins the Morton oracle on a sample
    bc, _, bids = oracle.rect_bruteforce(x, y, ids, r[:50])
    assert np.array_equal(bc, want_c[:50])
    t = bpt.Tree(keys, kid, 8, 4)
    for mode in MODES:
        c, o, got = t.rect_search(r, group_mode=mode)
        assert np.array_equal(c, want_c), mode
        assert np.array_equal(got, want_ids), mode
    # the one-lane count and the 4-lane cooperative descent agree too
    for opt, val in (("lanes", 4), ("coop", 0)):
        t.set_option(opt, val)
        c, o, got = t.rect_search(r, group_mode=0)
        assert np.array_equal(c, want_c), opt
        assert np.array_equal(got, want_ids), opt


def test_rect_id_runs_at_sort_limits(bpt):
    """Runs of exactly c ids, c on both sides of each sort tier (lane
    registers up to 8, wave LDS up to 1024, wave global memory beyond),
    several in one wave and mixed with empty rectangles."""
    npts = 70_000
    rng = np.random.default_rng(5)
    x = np.arange(npts, dtype=np.uint32)
    y = np.zeros(npts, dtype=np.uint32)
    ids = rng.permutation(npts).astype(np.uint32)
    z = np.zeros(npts, dtype=np.uint64)
    for b in range(17):  # Morton code of (x, 0): x's bits spread to the even positions
        z |= ((x.astype(np.uint64) >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b)
    t = bpt.Tree(z, ids, 8, 4)
    sizes = [0, 1, 2, 7, 8, 9, 63, 64, 65, 1023, 1024, 1025, 2048, 4097, 65_537, 3, 0, 9, 1025]
    starts = rng.integers(0, npts - 65_537, len(sizes))
    rects = np.array([[s, 0, s + c - 1, 0] if c else [5, 1, 4, 1] for s, c in zip(starts, sizes)],
                     dtype=np.uint32)
    want_c = np.array(sizes, dtype=np.uint64)
    want_ids = np.concatenate([np.sort(ids[s:s + c]) for s, c in zip(starts, sizes)])
    for mode in MODES:
        c, o, got = t.rect_search(rects, group_mode=mode)
        assert np.array_equal(c, want_c), mode
        assert np.array_equal(got, want_ids), mode


def test_rect_long_walks_deferred(bpt):
    """Every rectangle holds 16 points of a 512 x 512 grid, more than the
    count's walk budget, so every one is deferred to the whole-wave count;
    with 4M of them each block's deferral queue overflows and its lanes
    finish the rest themselves. Counts and sorted ids against the grid's
    closed form."""
    side = 512
    g = np.arange(side * side, dtype=np.uint64)
    gx, gy = g % side, g // side
    z = np.zeros(len(g), dtype=np.uint64)
    for b in range(9):
        z |= ((gx >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b)
        z |= ((gy >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b + 1)
    rng = np.random.default_rng(21)
    pid = rng.permutation(len(g)).astype(np.uint32)   # id of the point at grid index g
    o = np.argsort(z)
    t = bpt.Tree(z[o], pid[o], 8, 4)
    nr = 1 << 22
    x0 = rng.integers(0, side - 3, nr, dtype=np.uint64)
    y0 = rng.integers(0, side - 3, nr, dtype=np.uint64)
    rects = np.stack([x0, y0, x0 + 3, y0 + 3], axis=1).astype(np.uint32)
    c, off, got = t.rect_search(rects, group_mode=0)
    assert np.all(c == 16)
    d = np.arange(4, dtype=np.uint64)
    cells = (y0[:, None, None] + d[None, :, None]) * side + (x0[:, None, None] + d[None, None, :])
    want = np.sort(pid[cells.reshape(nr, 16)], axis=1).reshape(-1)
    assert np.array_equal(got, want)
    t.close()


def test_rect_empty_tree(bpt):
    """On an empty tree every rectangle counts 0 and no id is written."""
    t = bpt.Tree(np.zeros(0, dtype=np.uint64), np.zeros(0, dtype=np.uint32), 8, 4)
    r = np.array([[0, 0, 0xFFFFFFFF, 0xFFFFFFFF], [5, 5, 9, 9], [7, 1, 3, 2]], dtype=np.uint32)
    for mode in MODES:
        c, o, got = t.rect_search(r, group_mode=mode)
        assert np.array_equal(c, np.zeros(3, dtype=np.uint64)) and len(got) == 0, mode
    t.close()


def test_rect_last_rectangle_large(bpt, oracle, wl):
    """The ids pass learns the last rectangle's count separately (it has no
    next offset): a last rectangle covering the whole domain (every point)
    or half of it is walked by a whole wave and returns every id, sorted."""
    x, y, ids, keys, kid = wl.points_2d(1 << 21, seed=4)
    t = bpt.Tree(keys, kid, 8, 4)
    r = wl.rects(1000, seed=4, lo_exp=10, hi_exp=24)
    for last in ([0, 0, 0xFFFFFFFF, 0xFFFFFFFF], [0, 0, 0x7FFFFFFF, 0xFFFFFFFF], [7, 7, 7, 7]):
        rr = np.concatenate([r, np.array([last], dtype=np.uint32)])
        want_c, want_o, want_ids = oracle.rect_morton(keys, kid, rr)
        for mode in MODES:
            c, o, got = t.rect_search(rr, group_mode=mode)
            assert np.array_equal(c, want_c), (last, mode)
            assert np.array_equal(got, want_ids), (last, mode)
    c, o, got = t.rect_search(np.array([[0, 0, 0xFFFFFFFF, 0xFFFFFFFF]], dtype=np.uint32))
    assert int(c[0]) == len(keys) and np.array_equal(got, np.sort(kid))
    t.close()


def test_rect_duplicate_points(bpt, oracle, wl):
    rng = np.random.default_rng(9)
    x = rng.integers(0, 64, size=5000, dtype=np.uint64).astype(np.uint32)   # many coincident points
    y = rng.integers(0, 64, size=5000, dtype=np.uint64).astype(np.uint32)
    ids = np.arange(5000, dtype=np.uint32)
    z = wl.morton_np(x, y)
    order = np.argsort(z, kind="stable")
    r = np.array([[0, 0, 63, 63], [3, 3, 3, 3], [10, 20, 30, 40], [63, 0, 63, 63]], dtype=np.uint32)
    want_c, _, want_ids = oracle.rect_bruteforce(x, y, ids, r)
    t = bpt.Tree(z[order], ids[order], 8, 4)
    c, o, got = t.rect_search(r)
    assert np.array_equal(c, want_c) and np.array_equal(got, want_ids)


@pytest.mark.parametrize("shift", [0, 28])
def test_rect_every_box_on_a_grid(bpt, oracle, wl, shift):
    """Half the cells of a 16 x 16 grid hold a point; every one of the 18496
    boxes on the grid is searched, so the walk's BIGMIN jumps run from every
    out-of-box key of every box. With shift 28 the grid sits in the top
    coordinate bits and the box bounds are not aligned to it."""
    rng = np.random.default_rng(11 + shift)
    cells = rng.permutation(256)[:128]
    gx, gy = (cells % 16).astype(np.uint64), (cells // 16).astype(np.uint64)
    low = (1 << shift) - 1
    x = ((gx << shift) | (rng.integers(0, low + 1, size=128, dtype=np.uint64) if shift else 0)).astype(np.uint32)
    y = ((gy << shift) | (rng.integers(0, low + 1, size=128, dtype=np.uint64) if shift else 0)).astype(np.uint32)
    ids = np.arange(128, dtype=np.uint32)
    z = wl.morton_np(x, y)
    order = np.argsort(z, kind="stable")
    b = np.array([(a, c) for a in range(16) for c in range(a, 16)], dtype=np.uint64)
    bx = np.repeat(b, len(b), axis=0)
    by = np.tile(b, (len(b), 1))
    r = np.stack([bx[:, 0] << shift, by[:, 0] << shift, (bx[:, 1] << shift) | low, (by[:, 1] << shift) | low],
                 axis=1).astype(np.uint32)
    want_c, _, want_ids = oracle.rect_bruteforce(x, y, ids, r)
    t = bpt.Tree(z[order], ids[order], 8, 4)
    for mode in MODES:
        c, o, got = t.rect_search(r, group_mode=mode)
        assert np.array_equal(c, want_c), mode
        assert np.array_equal(got, want_ids), mode
    t.set_option("coop", 0)
    c, o, got = t.rect_search(r)
    assert np.array_equal(c, want_c) and np.array_equal(got, want_ids)


def test_grouping_keys_match_reference_functions(bpt, oracle):
    rng = np.random.default_rng(10)
    a = rng.integers(0, 1 << 64, size=3000, dtype=np.uint64)
    b = rng.integers(0, 1 << 40, size=3000, dtype=np.uint64)
    # max(a) >= 2^63: the reference's INT_MIN interval; max(b) < 2^63: interval 256
    assert np.array_equal(bpt.zorder2d(a, b), oracle.zorder_keys(a, b))
    assert np.array_equal(bpt.zorder2d(b, a), oracle.zorder_keys(b, a))
    a2 = rng.integers(0, 1 << 62, size=3000, dtype=np.uint64)    # both below 2^63
    assert np.array_equal(bpt.zorder2d(a2, b), oracle.zorder_keys(a2, b))
    g = bpt.group2d_keys(a, b)
    p0 = max(0, oracle.group2d_start_pass(int(a.max()), int(b.max())))
    for i in range(0, 3000, 97):
        want = 0
        for k in range(4):
            want = (want << 8) | (oracle.bucket2d(int(a[i]), int(b[i]), p0 + k) if p0 + k <= 15 else 0)
        assert int(g[i]) == want
    x = rng.integers(0, 1 << 32, size=1000, dtype=np.uint64).astype(np.uint32)
    y = rng.integers(0, 1 << 32, size=1000, dtype=np.uint64).astype(np.uint32)
    z = bpt.morton_encode(x, y)
    assert all(int(z[i]) == oracle.morton2d(int(x[i]), int(y[i])) for i in range(0, 1000, 7))


@pytest.mark.parametrize("kb,shift,bits", [(4, 20, 12), (4, 0, 8), (8, 52, 12), (8, 10, 16)])
def test_radix_group_is_stable_partition(bpt, kb, shift, bits):
    rng = np.random.default_rng(bits + shift)
    kdt = np.uint32 if kb == 4 else np.uint64
    keys = rng.integers(0, (1 << (8 * kb)) - 1, size=70000, dtype=np.uint64).astype(kdt)
    keys[:1000] = keys[0]
    perm, starts = bpt.radix_group(keys, shift, bits)
    digit = ((keys.astype(np.uint64) >> np.uint64(shift)) & np.uint64((1 << bits) - 1)).astype(np.int64)
    want = np.argsort(digit, kind="stable")
    assert np.array_equal(perm.astype(np.int64), want)
    cnt = np.bincount(digit, minlength=1 << bits)
    assert np.array_equal(starts, np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint64))


@pytest.mark.parametrize("kb,vb", [(4, 4), (8, 8), (8, 4)])
@pytest.mark.parametrize("layout", ["gaps", "empty_runs", "one_wide", "scan"])
def test_range_ids_offsets(bpt, oracle, kb, vb, layout):
    """bpt_range_ids with caller-made offsets: gaps between the runs stay
    untouched, long runs of empty ranges and one range far wider than the
    rest come back whole."""
    rng = np.random.default_rng(len(layout) + kb + vb)
    kdt = np.uint32 if kb == 4 else np.uint64
    vdt = np.uint32 if vb == 4 else np.uint64
    n = 300_007
    keys = np.unique(rng.integers(0, 1 << (8 * kb - 1), size=n + n // 8, dtype=np.uint64))[:n].astype(kdt)
    n = len(keys)
    vals = rng.integers(0, (1 << (8 * vb)) - 1, size=n, dtype=np.uint64).astype(vdt)
    t = bpt.Tree(keys, vals, kb, vb)
    nq = 20_000
    cnt = np.floor(np.exp2(rng.uniform(0, 9, nq))).astype(np.uint64)
    gap = np.zeros(nq, np.uint64)
    if layout == "gaps":
        gap = rng.integers(0, 70, nq).astype(np.uint64)
    elif layout == "empty_runs":
        cnt[rng.random(nq) < 0.9] = 0
    elif layout == "one_wide":
        cnt[:] = rng.integers(0, 4, nq)
        cnt[nq // 3] = n // 2
    first = np.array([rng.integers(0, n - int(c) + 1) for c in cnt], dtype=np.uint64)
    offs = np.cumsum(gap + cnt).astype(np.uint64) - cnt + 17  # gap before each run, base 17
    total = int(offs[-1] + cnt[-1]) + 5
    fill = vdt(np.iinfo(vdt).max)
    want = np.full(total, fill, dtype=vdt)
    for f, c, o in zip(first, cnt, offs):
        want[int(o):int(o + c)] = vals[int(f):int(f + c)]
    dfirst, dcnt, doffs = (bpt.DeviceArray.from_host(a) for a in (first, cnt, offs))
    dids = bpt.DeviceArray.from_host(np.full(total, fill, dtype=vdt))
    bpt.check(bpt.lib().bpt_range_ids(t.h, dfirst.ptr, dcnt.ptr, doffs.ptr, nq, dids.ptr, None))
    got = dids.to_host()
    for a in (dfirst, dcnt, doffs, dids):
        a.free()
    assert np.array_equal(got, want)


def test_exclusive_scan(bpt):
    rng = np.random.default_rng(11)
    for n in (1, 7, 2048, 2049, 5_000_003):
        c = rng.integers(0, 1000, size=n, dtype=np.uint64)
        off, total = bpt.exclusive_scan(c)
        assert total == int(c.sum())
        assert np.array_equal(off, np.concatenate([[0], np.cumsum(c)[:-1]]).astype(np.uint64))


def test_workload_c3_full_size(bpt, oracle, wl):
    """32M Morton-keyed points, 4M rectangles: the three query orders return
    identical counts and ids, and the counts and ids of every rectangle equal
    the oracle's (Morton-range scan with a rectangle filter)."""
    npts, nr = 1 << 25, 1 << 22
    x, y, ids, keys, kid = wl.points_2d(npts)
    r = wl.rects(nr)
    print("c3 generated", flush=True)
    t = bpt.Tree(keys, kid, 8, 4)
    c0, o0, i0 = t.rect_search(r, group_mode=MODES[0])
    for mode in MODES[1:]:
        c, o, got = t.rect_search(r, group_mode=mode)
        assert np.array_equal(c, c0) and np.array_equal(got, i0), mode
    t.close()
    want_c, want_o, want_ids = oracle.rect_morton(keys, kid, r)
    assert np.array_equal(c0, want_c)
    assert np.array_equal(o0, want_o)
    assert np.array_equal(i0, want_ids)
    print(f"c3 checked: all {nr} rectangles, {len(want_ids)} ids", flush=True)


def test_workload_ranges_full_size(bpt, oracle, wl):
    """The 1-D range workload of the benchmark: 64M uint32 keys (value = rank)
    and 4M inclusive ranges of 1..1024 ranks. Counts and lower bounds equal
    the oracle for every range in every order; every id equals the oracle's
    (values are ranks, so range i's ids are first[i], first[i] + 1, ...)."""
    n, nr = 1 << 26, 1 << 22
    keys, vals = wl.point_keys(n, workload=2)
    rng = np.random.default_rng(11)
    st = rng.integers(0, n, nr)
    w = np.floor(np.exp2(rng.uniform(0, 10, nr))).astype(np.int64)
    lo, hi = keys[st], keys[np.minimum(st + w, n - 1)]
    print("ranges generated", flush=True)
    t = bpt.Tree(keys, vals)
    want_first, want_cnt = oracle.range_closed(keys, lo, hi)
    for mode in MODES:
        first, cnt, offs, ids = t.range_search(lo, hi, group_mode=mode, with_ids=mode == 0)
        assert np.array_equal(cnt, want_cnt), mode
        assert np.array_equal(first, want_first), mode
        if mode == 0:
            total = int(cnt.sum())
            assert len(ids) == total
            assert np.array_equal(offs, np.concatenate([[0], np.cumsum(want_cnt)[:-1]]).astype(np.uint64))
            wv = vals.astype(np.int64)
            for r0 in range(0, nr, 1 << 18):          # all ranges, a block of 2^18 at a time
                r1 = min(nr, r0 + (1 << 18))
                c = want_cnt[r0:r1].astype(np.int64)
                o = offs[r0:r1].astype(np.int64)
                pos = np.arange(int(o[0]), int(o[-1] + c[-1]), dtype=np.int64)
                start = np.repeat(want_first[r0:r1].astype(np.int64), c)
                want = wv[start + (pos - np.repeat(o, c))].astype(np.uint32)
                assert np.array_equal(ids[pos], want), r0
            del ids
    t.close()
