"""CPU tests of the oracle's grouping pass loop (ora_group_lists).

Each case below was worked out by hand from grouping.cu / grouping2d.cu and
its expected lists are written out, so the oracle is checked against the
hand computation, not against itself. Random inputs then check the
invariants any correct group list has (sizes sum to n, contiguous prefixes,
order between groups, caps).

Cases, with the reference lines they exercise:
  A  one pass, < 10 groups (grouping.cu:776-884): merge across empty
     buckets, stops when the sum reaches the cap (strict <).
  B  two passes, < 10 groups: a group equal to the cap is split (:786 keeps
     only groups below the cap), a smaller one passes through.
  C  two passes, the second with >= 10 groups (grouping.cu:1288-1363 and
     the kernel :692-770): only groups above the cap are split (:703), one
     greedy merge over all slots joins the tail of one group with the next.
  D  2-D (grouping2d.cu): nibble products, no-op passes above the data.
  E  2-D with a 64-bit maximum: negative start pass, shifts >= 64 give
     bucket 0, the loop ends at pass 15 with a group above the cap.
  F  genGrouping's start pass from the maximum (grouping.cu:1736-1744).
"""
import numpy as np
import pytest

SENT64 = 0xFFFFFFFFFFFFFFFF


def _top(*pairs):
    """Keys from (top byte, second byte, count) triples: top << 56 | second << 48."""
    out = []
    for top, second, count in pairs:
        out += [(top << 56) | (second << 48)] * count
    return np.array(out, dtype=np.uint64)


def _lists(oracle, keys, cap, start=0, b=None, kind=0):
    sizes, pre, perm = oracle.group_lists(keys, b, cap, kind, start)
    return sizes.tolist(), pre.tolist(), perm


def test_case_a_single_pass_merge(oracle):
    keys = _top((1, 0, 2), (2, 0, 1), (5, 0, 3), (7, 0, 1), (0xFF, 0, 1))
    sizes, pre, _ = _lists(oracle, keys, cap=4)
    # bucket 1 (2) + bucket 2 (1) = 3; adding bucket 5 (3) would reach 6.
    # bucket 5 (3) + bucket 7 (1) = 4 is not below the cap: separate.
    # bucket 7 (1) absorbs the empty buckets and bucket 255 (1).
    assert sizes == [3, 3, 2] and pre == [0, 3, 6]


def test_case_b_group_equal_to_cap_is_split(oracle):
    keys = _top((0x10, 0, 2), (0x10, 1, 1), (0x10, 2, 2), (0x20, 0, 1), (0x20, 5, 2), (0x20, 9, 1))
    sizes, pre, perm = _lists(oracle, keys, cap=4)
    # pass 0: [5 (top 0x10), 4 (top 0x20)]; pass 1 splits both (5 >= 4, 4 >= 4):
    # 0x10 -> {2+1}, {2}; 0x20 -> {1+2} (second bytes 0 and 5), {1} (9).
    assert sizes == [3, 2, 3, 1] and pre == [0, 3, 5, 8]
    grouped = keys[perm.astype(np.int64)]
    assert [int(k) >> 48 for k in grouped[5:8]] == [0x2000, 0x2005, 0x2005]


def test_case_c_single_kernel_form(oracle):
    keys = _top(*[(t, 0, 2) for t in range(0, 16, 2)], (16, 0, 1), (17, 0, 2),
                (20, 0, 3), (20, 0xFF, 1), (22, 0, 1), (24, 0, 2), (24, 0x80, 1))
    assert len(keys) == 27
    sizes, pre, perm = _lists(oracle, keys, cap=3)
    # pass 0 (one group): 13 groups [2]*8 + [1, 2, 4, 1, 3]; the 4 exceeds the
    # cap. pass 1 runs the single-kernel form: only the 4 (top 20) is split,
    # into 3 (second byte 0) and 1 (0xFF); that 1 merges with the next group
    # (top 22, size 1); the group of exactly 3 (top 24) is not split.
    assert sizes == [2] * 8 + [1, 2, 3, 2, 3]
    assert pre == [0, 2, 4, 6, 8, 10, 12, 14, 16, 17, 19, 22, 24]
    grouped = keys[perm.astype(np.int64)]
    assert sorted(int(k) >> 48 for k in grouped[22:24]) == [(20 << 8) | 0xFF, 22 << 8]


def test_case_d_2d_products(oracle):
    s21, s18 = 1 << 21, 1 << 18
    a = np.array([s21, 2 * s21, 3 * s21, s21 | s18, s21], dtype=np.uint64)
    b = np.array([s21, 3 * s21, 2 * s21, s21, s21 | 2 * s18], dtype=np.uint64)
    start = oracle.group2d_start_pass(int(a.max()), int(b.max()))
    assert start == 8  # 16 - ceil(23 / 3)
    sizes, pre, perm = _lists(oracle, a, cap=2, start=start, b=b, kind=1)
    # passes 8..13 read bits >= 24: all zero, nothing moves. pass 14 (bits
    # 21..24): products 1, 6, 6, 1, 1 -> [3 (product 1), 2 (product 6)].
    # pass 15 (bits 18..21): the 3 splits into products 64, 72, 80; the 2
    # (products 0, 0) is split too (2 >= cap) and stays one group.
    assert sizes == [1, 1, 1, 2] and pre == [0, 1, 2, 3]
    assert perm[:3].tolist() == [0, 3, 4] and sorted(perm[3:].tolist()) == [1, 2]


def test_case_e_2d_negative_start_and_exhausted_passes(oracle):
    h = 1 << 63
    a = np.array([h, h, 0], dtype=np.uint64)
    b = np.array([h, 0, 5], dtype=np.uint64)
    start = oracle.group2d_start_pass(int(a.max()), int(b.max()))
    assert start == -6
    sizes, pre, perm = _lists(oracle, a, cap=1, start=start, b=b, kind=1)
    # passes -6..-1 shift by >= 64: bucket 0. pass 0 (bit 63): products 1, 0, 0
    # -> [2, 1]. Passes 1..15 never separate (h, 0) from (0, 5); the loop ends
    # after pass 15 with a group above the cap.
    assert sizes == [2, 1] and pre == [0, 2]
    assert sorted(perm[:2].tolist()) == [1, 2] and perm[2] == 0


def test_case_f_start_pass_from_maximum(oracle):
    keys = np.array([0x0102, 0x0103, 0x0201, 0x0001], dtype=np.uint64)
    start = oracle.group_start_pass(int(keys.max()))
    assert start == 6  # 8 - ceil(10 / 8)
    sizes, pre, perm = _lists(oracle, keys, cap=2, start=start)
    # pass 6 reads byte 1: buckets 0 (1), 1 (2), 2 (1) -> no merge reaches
    # below 2 except the trailing 1, which absorbs nothing
    assert sizes == [1, 2, 1] and pre == [0, 1, 3]
    assert keys[perm.astype(np.int64)].tolist() == [0x0001, 0x0102, 0x0103, 0x0201]


def test_degenerate_inputs(oracle):
    # all-zero keys: max 0, start pass 8, no pass runs
    z = np.zeros(100, dtype=np.uint64)
    assert _lists(oracle, z, cap=10, start=oracle.group_start_pass(0))[:2] == ([100], [0])
    # n == 0 from pass 0: the only group is empty and is dropped
    assert _lists(oracle, np.zeros(0, dtype=np.uint64), cap=10)[:2] == ([], [])
    # everything below the cap: one pass, one merged group
    keys = _top((1, 0, 3), (9, 0, 3))
    assert _lists(oracle, keys, cap=100)[:2] == ([6], [0])


@pytest.mark.parametrize("n,cap,space,seed", [
    (5000, 200, 1 << 64, 1), (5000, 200, 1 << 20, 2), (20000, 33000, 1 << 64, 3),
    (100000, 1000, 1 << 64, 4), (100000, 1000, 1 << 36, 5), (3000, 50, 10, 6),
    (70000, 33000, 1 << 64, 7)])
def test_1d_invariants(oracle, n, cap, space, seed):
    rng = np.random.default_rng(seed)
    if space == 1 << 64:
        keys = rng.integers(0, 1 << 64, size=n, dtype=np.uint64)
        keys[: n // 10] = keys[n // 10: 2 * (n // 10)]  # duplicates
    else:
        keys = rng.integers(0, space, size=n, dtype=np.uint64)
    for start in (0, oracle.group_start_pass(int(keys.max()))):
        sizes, pre, perm = oracle.group_lists(keys, None, cap, 0, start)
        assert int(sizes.sum()) == n
        assert np.array_equal(pre, np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint64))
        assert (sizes > 0).all()
        assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint64))
        g = keys[perm.astype(np.int64)]
        bounds = np.concatenate([pre, [n]]).astype(np.int64)
        for i in range(len(sizes) - 1):
            # groups come from distinct byte prefixes, in order
            assert g[bounds[i]:bounds[i + 1]].max() < g[bounds[i + 1]:bounds[i + 2]].min()
        for i in np.nonzero(sizes > cap)[0]:
            # a group may stay above the cap only when no pass can split it:
            # all its keys are equal
            seg = g[bounds[i]:bounds[i + 1]]
            assert seg.min() == seg.max()


@pytest.mark.parametrize("n,cap,bits,seed", [(5000, 100, 64, 1), (20000, 500, 40, 2), (50000, 33000, 64, 3),
                                             (50000, 300, 30, 4)])
def test_2d_invariants(oracle, n, cap, bits, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 1 << bits, size=n, dtype=np.uint64)
    b = rng.integers(0, 1 << bits, size=n, dtype=np.uint64)
    start = oracle.group2d_start_pass(int(a.max()), int(b.max()))
    sizes, pre, perm = oracle.group_lists(a, b, cap, 1, start)
    assert int(sizes.sum()) == n and (sizes > 0).all()
    assert np.array_equal(pre, np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint64))
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint64))
    # every group above the cap is unsplittable: its pairs share the bucket of
    # every pass from the first effective one (pass 0 or later) to 15
    p = perm.astype(np.int64)
    bounds = np.concatenate([pre, [n]]).astype(np.int64)
    for i in np.nonzero(sizes > cap)[0]:
        idx = p[bounds[i]:bounds[i + 1]]
        for q in range(max(start, 0), 16):
            bk = {oracle.bucket2d(int(a[j]), int(b[j]), q) for j in idx[:50]}
            assert len(bk) == 1


def test_compat_bucket_functions_through_driver(tmp_path, oracle):
    """The drop-in's bucketIdentify / bucketIdentify2d (host functions of the
    compat layer; no GPU call) against the oracle, through bptree_groups."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "b-tree-index-searching_amd",
                       "bptree_groups")
    if not os.path.exists(exe):
        pytest.fail("bptree_groups is not built (run __graft_entry__.build())")
    rng = np.random.default_rng(11)
    a = np.concatenate([rng.integers(0, 1 << 64, size=500, dtype=np.uint64),
                        np.array([0, SENT64, 1 << 63, 0x0123456789ABCDEF], dtype=np.uint64)])
    b = np.concatenate([rng.integers(0, 1 << 64, size=500, dtype=np.uint64),
                        np.array([SENT64, 0, 1 << 63, 0xFEDCBA9876543210], dtype=np.uint64)])
    n = len(a)
    a.tofile(tmp_path / "a")
    b.tofile(tmp_path / "b")
    r = subprocess.run([exe, "bucketIdentify", str(n), "0", str(tmp_path / "o1"), str(tmp_path / "a")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "o1.0", dtype=np.uint64).reshape(8, n)
    for p in range(8):
        assert got[p].tolist() == [oracle.bucket_identify(int(x), p) for x in a]
    # closed form of grouping.cu:25-29 on a known key
    assert got[:, -1].tolist() == [0x01, 0x23, 0x45, 0x67, 0x89, 0xAB, 0xCD, 0xEF]
    r = subprocess.run([exe, "bucketIdentify2d", str(n), "0", str(tmp_path / "o2"), str(tmp_path / "a"),
                        str(tmp_path / "b")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "o2.0", dtype=np.uint64).reshape(22, n)
    for p in range(-6, 16):
        assert got[p + 6].tolist() == [oracle.bucket2d(int(x), int(y), p) for x, y in zip(a, b)]
    # pass 0 reads bit 63 only; pass 15 the nibble at bit 18
    assert got[6, -2] == 1 and got[6 + 15, -1] == ((0x0123456789ABCDEF >> 18) & 0xF) * ((0xFEDCBA9876543210 >> 18) & 0xF)
