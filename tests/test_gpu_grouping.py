"""GPU: the reference-named grouping functions (genGrouping,
genGroupingWithOnedata, genGroupingWithTwodata, genGrouping2d,
genGrouping2dZ) through the compiled driver bptree_groups, the way a caller
linked against the reference calls them. The group lists and prefixes must
equal the oracle's restatement of the reference's pass loop
(ora_group_lists), every group must hold the oracle's multiset of keys (the
order inside a group is unspecified in the reference), and payloads must
stay paired with their keys."""
import os
import subprocess

import numpy as np
import pytest

from test_oracle_grouping import _top

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "b-tree-index-searching_amd", "bptree_groups")
CAP = 33000  # GROUP_SEARCH_SIZE, BPTree.h:9


def _run(tmp_path, fn, cap, *arrays):
    n = len(arrays[0])
    paths = []
    for k, a in enumerate(arrays):
        p = str(tmp_path / f"in{k}")
        np.ascontiguousarray(a, dtype=np.uint64).tofile(p)
        paths.append(p)
    out = str(tmp_path / "out")
    print(f"[grouping] {fn} n={n} cap={cap}", flush=True)
    r = subprocess.run([EXE, fn, str(n), str(cap), out, *paths], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert f"{fn} ms=" in r.stdout, r.stdout + r.stderr
    res = {"stdout": r.stdout}
    for suffix in ("list", "prefix", "0", "1", "2"):
        if os.path.exists(f"{out}.{suffix}"):
            res[suffix] = np.fromfile(f"{out}.{suffix}", dtype=np.uint64)
    return res


def _group_ids(sizes):
    return np.repeat(np.arange(len(sizes)), sizes.astype(np.int64))


def _check_1d(oracle, res, keys, cap, start, payloads=()):
    sizes, pre, perm = oracle.group_lists(keys, None, cap, 0, start)
    assert np.array_equal(res["list"], sizes)
    assert np.array_equal(res["prefix"], pre)
    got = res["0"]
    want = keys[perm.astype(np.int64)]
    gid = _group_ids(sizes)
    # the same multiset of keys in every group
    assert np.array_equal(got[np.lexsort((got, gid))], want[np.lexsort((want, gid))])
    for k, f in enumerate(payloads):
        # payload k of the element with index i is f(i); index = payload 0
        idx = res["1"].astype(np.int64)
        assert np.array_equal(keys[idx], got)
        assert np.array_equal(res[str(k + 1)], f(idx).astype(np.uint64))


def test_hand_worked_cases(tmp_path, oracle):
    """The hand-worked cases of test_oracle_grouping.py, through the drop-in:
    the lists written there by hand must come back from the device path."""
    cases = [
        (_top((1, 0, 2), (2, 0, 1), (5, 0, 3), (7, 0, 1), (0xFF, 0, 1)), 4, [3, 3, 2], [0, 3, 6]),
        (_top((0x10, 0, 2), (0x10, 1, 1), (0x10, 2, 2), (0x20, 0, 1), (0x20, 5, 2), (0x20, 9, 1)), 4,
         [3, 2, 3, 1], [0, 3, 5, 8]),
        (_top(*[(t, 0, 2) for t in range(0, 16, 2)], (16, 0, 1), (17, 0, 2), (20, 0, 3), (20, 0xFF, 1),
              (22, 0, 1), (24, 0, 2), (24, 0x80, 1)), 3,
         [2] * 8 + [1, 2, 3, 2, 3], [0, 2, 4, 6, 8, 10, 12, 14, 16, 17, 19, 22, 24]),
    ]
    for keys, cap, sizes, pre in cases:
        idx = np.arange(len(keys), dtype=np.uint64)
        res = _run(tmp_path, "genGroupingWithTwodata", cap, keys, idx, idx * 3 + 1)
        assert res["list"].tolist() == sizes and res["prefix"].tolist() == pre
        _check_1d(oracle, res, keys, cap, 0, payloads=(lambda i: i, lambda i: i * 3 + 1))
    # case D (2-D products) and E (negative start pass)
    s21, s18 = 1 << 21, 1 << 18
    a = np.array([s21, 2 * s21, 3 * s21, s21 | s18, s21], dtype=np.uint64)
    b = np.array([s21, 3 * s21, 2 * s21, s21, s21 | 2 * s18], dtype=np.uint64)
    res = _run(tmp_path, "genGrouping2d", 2, a, b)
    assert res["list"].tolist() == [1, 1, 1, 2] and res["prefix"].tolist() == [0, 1, 2, 3]
    assert [(int(x), int(y)) for x, y in zip(res["0"][:3], res["1"][:3])] == [
        (s21, s21), (s21 | s18, s21), (s21, s21 | 2 * s18)]
    h = 1 << 63
    res = _run(tmp_path, "genGrouping2d", 1, np.array([h, h, 0], dtype=np.uint64),
               np.array([h, 0, 5], dtype=np.uint64))
    assert res["list"].tolist() == [2, 1] and res["prefix"].tolist() == [0, 2]
    assert (int(res["0"][2]), int(res["1"][2])) == (h, h)
    # case F: genGrouping's start pass from the maximum
    keys = np.array([0x0102, 0x0103, 0x0201, 0x0001], dtype=np.uint64)
    res = _run(tmp_path, "genGrouping", 2, keys)
    assert res["list"].tolist() == [1, 2, 1] and res["prefix"].tolist() == [0, 1, 3]
    assert res["0"].tolist() == [0x0001, 0x0102, 0x0103, 0x0201]


@pytest.mark.parametrize("n,cap,kind,seed", [
    (1 << 20, CAP, "uniform", 1), (1 << 22, CAP, "uniform", 2), (300_001, 1000, "uniform", 3),
    (1 << 20, CAP, "narrow", 4), (1 << 20, 500, "dups", 5), (200_000, CAP, "skew", 6), (1, CAP, "uniform", 7),
    (0, CAP, "uniform", 8)])
def test_gen_grouping_matches_oracle(tmp_path, oracle, n, cap, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        keys = rng.integers(0, 1 << 64, size=n, dtype=np.uint64)
    elif kind == "narrow":        # genGrouping starts at a later pass
        keys = rng.integers(0, 1 << 37, size=n, dtype=np.uint64)
    elif kind == "dups":          # groups that no pass can split
        keys = rng.integers(0, 700, size=n, dtype=np.uint64) << np.uint64(40)
    else:                         # a few heavy prefixes among uniform keys
        keys = rng.integers(0, 1 << 64, size=n, dtype=np.uint64)
        keys[: n // 2] = (rng.integers(0, 3, size=n // 2, dtype=np.uint64) << np.uint64(56)) | \
            rng.integers(0, 1 << 56, size=n // 2, dtype=np.uint64)
    res = _run(tmp_path, "genGrouping", cap, keys)
    start = oracle.group_start_pass(int(keys.max())) if n else 8
    if n == 0:
        assert res["list"].tolist() == [0] and res["prefix"].tolist() == [0]
        return
    _check_1d(oracle, res, keys, cap, start)
    idx = np.arange(n, dtype=np.uint64)
    res = _run(tmp_path, "genGroupingWithTwodata", cap, keys, idx, ~idx)
    _check_1d(oracle, res, keys, cap, 0, payloads=(lambda i: i, lambda i: ~i.astype(np.uint64)))
    res = _run(tmp_path, "genGroupingWithOnedata", cap, keys, idx)
    _check_1d(oracle, res, keys, cap, 0, payloads=(lambda i: i,))


@pytest.mark.parametrize("n,cap,bits,seed", [(1 << 20, CAP, 64, 1), (1 << 20, CAP, 40, 2), (300_000, 700, 64, 3),
                                             (500_000, 2000, 30, 4), (1 << 21, CAP, 50, 5)])
def test_gen_grouping_2d_matches_oracle(tmp_path, oracle, n, cap, bits, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 1 << bits, size=n, dtype=np.uint64)
    b = a + rng.integers(0, 1 << (bits - 8), size=n, dtype=np.uint64)  # range queries: end >= start
    res = _run(tmp_path, "genGrouping2d", cap, a, b)
    start = oracle.group2d_start_pass(int(a.max()), int(b.max()))
    sizes, pre, perm = oracle.group_lists(a, b, cap, 1, start)
    assert np.array_equal(res["list"], sizes) and np.array_equal(res["prefix"], pre)
    p = perm.astype(np.int64)
    gid = _group_ids(sizes)
    ga, gb = res["0"], res["1"]
    wa, wb = a[p], b[p]
    o1, o2 = np.lexsort((gb, ga, gid)), np.lexsort((wb, wa, gid))
    assert np.array_equal(ga[o1], wa[o2]) and np.array_equal(gb[o1], wb[o2])


def test_gen_grouping_2dz_both_interval_regimes(tmp_path, oracle):
    rng = np.random.default_rng(9)
    n = 1 << 18
    for a in (rng.integers(0, 1 << 40, size=n, dtype=np.uint64),                  # interval 256
              rng.integers(0, 1 << 64, size=n, dtype=np.uint64),                  # INT_MIN divisor
              np.full(n, (1 << 63) - (1 << 38), dtype=np.uint64),                 # at the threshold
              np.full(n, (1 << 63) - (1 << 38) - 1, dtype=np.uint64),             # just below it
              np.zeros(n, dtype=np.uint64)):                                      # interval 0
        a = a.copy()
        if a.any():
            a[: n // 2] = rng.integers(0, 1 << 32, size=n // 2, dtype=np.uint64)
        b = rng.integers(0, 1 << 64, size=n, dtype=np.uint64)
        res = _run(tmp_path, "genGrouping2dZ", 0, a, b)
        assert np.array_equal(res["0"], oracle.zorder_keys(a, b))
