"""GPU parity: bulk build + point search (plain, grouped) + lower bound,
through the C ABI, against the oracle on seeded inputs (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFGS = [  # key bytes, value bytes, fanout, leaf cap
    (4, 4, 32, 16), (4, 4, 16, 8), (4, 4, 64, 32), (8, 8, 16, 8), (8, 8, 32, 16),
    (8, 8, 64, 32), (8, 4, 32, 16), (8, 4, 16, 8),
]
SIZES = [0, 1, 2, 7, 8, 9, 16, 17, 1000, 4096, 65537, 300001]


def _case(n, kb, vb, seed):
    rng = np.random.default_rng(seed)
    kdt = np.uint32 if kb == 4 else np.uint64
    vdt = np.uint32 if vb == 4 else np.uint64
    space = (1 << 32) if kb == 4 else (1 << 64)
    keys = np.unique(rng.integers(0, space - 1, size=n, dtype=np.uint64)).astype(kdt)
    if len(keys) and n > 3:
        keys[0] = 0
        keys[-1] = kdt(space - 1)          # the all-ones key is a real key
    vals = rng.integers(0, (1 << (8 * vb)) - 1, size=len(keys), dtype=np.uint64).astype(vdt)
    q = np.concatenate([keys, rng.integers(0, space - 1, size=3000, dtype=np.uint64).astype(kdt),
                        np.array([0, space - 1, 1, space - 2], dtype=np.uint64).astype(kdt)])
    rng.shuffle(q)
    return keys, vals, q


def _expect(oracle, keys, vals, q, vb):
    sent = 0xFFFFFFFF if vb == 4 else 0xFFFFFFFFFFFFFFFF
    return oracle.point_closed(keys, vals, q, sent).astype(vals.dtype)


@pytest.mark.parametrize("cfg", CFGS, ids=lambda c: "k%dv%dF%dL%d" % c)
@pytest.mark.parametrize("n", SIZES)
def test_point_plain_and_grouped(bpt, oracle, cfg, n):
    kb, vb, F, LF = cfg
    keys, vals, q = _case(n, kb, vb, seed=n * 31 + F + LF)
    print(f"case n={n} cfg={cfg}", flush=True)
    t = bpt.Tree(keys, vals, kb, vb, fanout=F, leaf_cap=LF)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    want = _expect(oracle, keys, vals, q, vb)
    assert np.array_equal(t.point_search(q, bpt.PLAIN), want)
    t.set_option("lanes", 4)  # four lanes per query
    assert np.array_equal(t.point_search(q, bpt.PLAIN), want)
    t.set_option("addr32", 0)  # 64-bit addressing
    assert np.array_equal(t.point_search(q, bpt.PLAIN), want)
    t.set_option("lanes", 8)
    assert np.array_equal(t.point_search(q, bpt.PLAIN), want)
    t.set_option("leaf_nt", 0)
    assert np.array_equal(t.point_search(q, bpt.PLAIN), want)
    t.set_option("addr32", 1)
    t.set_option("leaf_nt", 1)
    t.set_option("coop", 0)   # one lane per query
    assert np.array_equal(t.point_search(q, bpt.PLAIN), want)
    t.set_option("coop", 1)
    if kb * 8 >= 12:
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want)
        t.set_option("group_sub", 0)  # leaf-maxima slices, whole leaves
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want)
        t.set_option("group_lanes", 8)
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want)
        t.set_option("group_lanes", 4)
        t.set_option("group_sub", 1)
    pos = t.lower_bound(q)
    assert np.array_equal(pos, oracle.lower_bound(keys, q))
    for opt, val in (("lanes", 4), ("coop", 0)):  # the cooperative lower bound with 4 lanes; one lane per query
        t.set_option(opt, val)
        assert np.array_equal(t.lower_bound(q), oracle.lower_bound(keys, q)), opt
    t.close()


def test_point_duplicates_first_occurrence(bpt, oracle):
    rng = np.random.default_rng(1)
    keys = np.sort(rng.integers(0, 300, size=5000, dtype=np.uint64).astype(np.uint32))
    vals = np.arange(len(keys), dtype=np.uint32)
    q = rng.integers(0, 320, size=4000, dtype=np.uint64).astype(np.uint32)
    t = bpt.Tree(keys, vals)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    want = _expect(oracle, keys, vals, q, 4)   # first occurrence = lower bound
    assert np.array_equal(t.point_search(q), want)
    assert np.array_equal(t.point_search(q, bpt.GROUPED), want)


def test_unsorted_build_rejected(bpt):
    keys = np.array([5, 3, 9], dtype=np.uint32)
    with pytest.raises(bpt.BptError) as e:
        bpt.Tree(keys, keys)
    assert e.value.status == 2


def test_grouped_small_batches_run_plain(bpt, oracle, wl):
    """Below the tree's group_min (2^24 by default) a grouped batch runs the
    plain descent: it needs no workspace and gives the same results; from
    group_min on, the grouped kernels run with their workspace."""
    keys, vals = wl.point_keys(1 << 20, workload=2)
    q = wl.uniform_hits(keys, 1 << 18)
    want = oracle.point_closed(keys, vals, q, 0xFFFFFFFF).astype(np.uint32)
    t = bpt.Tree(keys, vals)
    assert t.point_search_ws_size(1 << 18, bpt.GROUPED) == 0
    assert t.point_search_ws_size(1 << 24, bpt.GROUPED) > 0
    assert np.array_equal(t.point_search(q, bpt.GROUPED), want)
    t.set_option("group_min", 1 << 10)
    assert t.point_search_ws_size(1 << 18, bpt.GROUPED) > 0
    assert np.array_equal(t.point_search(q, bpt.GROUPED), want)
    with pytest.raises(bpt.BptError):
        t.set_option("group_min", -1)
    t.close()


@pytest.mark.parametrize("bits", [1, 8, 12, 13])
def test_grouped_bits_and_launch_shapes(bpt, oracle, bits):
    keys, vals, q = _case(200000, 4, 4, seed=bits)
    t = bpt.Tree(keys, vals)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    t.set_group_bits(bits)
    want = _expect(oracle, keys, vals, q, 4)
    assert np.array_equal(t.point_search(q, bpt.GROUPED), want)
    for block, bpc in ((64, 1), (128, 3), (256, 16)):
        t.set_launch(block, bpc)
        assert np.array_equal(t.point_search(q, bpt.PLAIN), want)


@pytest.mark.parametrize("kb,n,bits", [(4, 400000, 1), (8, 1200000, 1), (8, 300000, 4), (4, 6000000, 2)])
def test_grouped_slice_levels(bpt, oracle, kb, n, bits):
    """Few buckets allowed: the LDS slice moves up to level 2 or 3 and the
    search reads the levels below it from global memory."""
    keys, vals, q = _case(n, kb, kb, seed=n + bits)
    t = bpt.Tree(keys, vals, kb, kb)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    t.set_group_bits(bits)
    want = _expect(oracle, keys, vals, q, kb)
    for lanes in (4, 8):
        t.set_option("group_lanes", lanes)
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want), lanes
    t.set_option("group_lanes", 4)
    for slice_bytes in (16384, 32768, 65536):  # narrower buckets
        t.set_option("group_slice", slice_bytes)
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want), slice_bytes
    t.set_option("group_slice", 131072)
    t.set_option("coop", 0)
    assert np.array_equal(t.point_search(q, bpt.GROUPED), want)


@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("tail", [1, 3, 4, 5, 8, 9, 15, 16])
def test_grouped_sub_leaf_edges(bpt, oracle, kb, tail):
    """Sub-leaf slices at the edges: a last leaf holding `tail` keys (its
    later sub-leaves empty), runs of a duplicate key across sub-leaf and leaf
    boundaries, the all-ones key as a real key, and queries past the largest
    key, each against the leaf-maxima path and the oracle."""
    rng = np.random.default_rng(tail * 7 + kb)
    kdt = np.uint32 if kb == 4 else np.uint64
    lf = 16
    n = lf * 3000 + tail
    keys = np.sort(rng.integers(0, 40000, size=n, dtype=np.uint64)).astype(kdt)  # about 1.2 per key
    keys[2 * lf * 7 + 3: 2 * lf * 7 + 40] = keys[2 * lf * 7 + 3]  # a run over several sub-leaves
    keys = np.sort(keys)
    keys[-3:] = np.iinfo(kdt).max
    vals = np.arange(n, dtype=np.uint64).astype(kdt)
    t = bpt.Tree(keys, vals, kb, kb, fanout=32, leaf_cap=lf)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    q = np.concatenate([keys[::3], np.arange(0, 40100, 7, dtype=np.uint64).astype(kdt),
                        np.array([np.iinfo(kdt).max, np.iinfo(kdt).max - 1, 40000, 0], dtype=kdt)])
    rng.shuffle(q)
    want = _expect(oracle, keys, vals, q, kb)
    assert np.array_equal(t.point_search(q, bpt.GROUPED), want)
    t.set_option("group_sub", 0)
    assert np.array_equal(t.point_search(q, bpt.GROUPED), want)
    t.close()


@pytest.mark.parametrize("kb", [4, 8])
def test_grouped_keys_with_gaps(bpt, oracle, kb):
    """Keys in two dense clusters at both ends of the key space and a few in
    between: slices whose prefix tables jump over empty key ranges."""
    rng = np.random.default_rng(11 + kb)
    kdt = np.uint32 if kb == 4 else np.uint64
    top = int(np.iinfo(kdt).max)
    keys = np.unique(np.concatenate([
        np.arange(0, 300000, dtype=np.uint64), np.uint64(top) - np.arange(0, 300000, dtype=np.uint64),
        rng.integers(1 << 20, top >> 1, size=40, dtype=np.uint64)])).astype(kdt)
    vals = rng.integers(0, top, size=len(keys), dtype=np.uint64).astype(kdt)
    t = bpt.Tree(keys, vals, kb, kb)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    q = np.concatenate([keys[rng.integers(0, len(keys), size=200000)],
                        rng.integers(0, top, size=50000, dtype=np.uint64).astype(kdt),
                        np.array([0, top, 300000, top - 300000], dtype=np.uint64).astype(kdt)])
    want = _expect(oracle, keys, vals, q, kb)
    for sub in (1, 0):
        t.set_option("group_sub", sub)
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want), sub
    t.close()


def test_grouped_skewed_queries(bpt, oracle):
    """Queries piled into a few buckets, below the smallest and above the
    largest key, and repeated: runs of very uneven length."""
    rng = np.random.default_rng(5)
    keys = np.unique(rng.integers(1000, 1 << 31, size=2_000_000, dtype=np.uint64)).astype(np.uint32)
    vals = np.arange(len(keys), dtype=np.uint32)
    t = bpt.Tree(keys, vals)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    hot = keys[rng.integers(100000, 100400, size=900_000)]
    q = np.concatenate([hot, np.full(50_000, 7, np.uint32), np.full(50_000, 0xFFFFFFF0, np.uint32),
                        rng.integers(0, 1 << 32, size=100_000, dtype=np.uint64).astype(np.uint32)])
    rng.shuffle(q)
    want = _expect(oracle, keys, vals, q, 4)
    for chunk, xcd, slice_bytes, threads, blocks in ((0, 1, 131072, 1024, 1), (16384, 1, 131072, 1024, 1),
                                                     (256, 1, 131072, 512, 1), (4096, 0, 65536, 1024, 1),
                                                     (0, 1, 32768, 768, 2), (1024, 0, 16384, 512, 4)):
        t.set_option("group_chunk", chunk)
        t.set_option("group_xcd", xcd)
        t.set_option("group_slice", slice_bytes)
        t.set_option("group_threads", threads)
        t.set_option("group_blocks", blocks)
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want), (chunk, xcd, slice_bytes, threads, blocks)


def test_workload_c1_scaled_and_page_tree(bpt, oracle, wl):
    """1M uint32 keys, order-32 page tree of the reference restated on the
    CPU, against the flat GPU tree, on the CPU workload's query mix."""
    keys, vals = wl.point_keys(1 << 20, workload=1)
    q = wl.cpu_queries(keys, 1 << 20, workload=1)
    pt = oracle.PageTree(keys.astype(np.uint64), vals.astype(np.uint64), 32)
    want = pt.point(q.astype(np.uint64), 0xFFFFFFFF).astype(np.uint32)
    t = bpt.Tree(keys, vals)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    assert np.array_equal(t.point_search(q), want)
    assert np.array_equal(t.point_search(q, bpt.GROUPED), want)
    hits = (want != 0xFFFFFFFF).mean()
    assert 0.49 < hits < 0.52


@pytest.mark.parametrize("qset", ["uniform", "zipf"])
def test_workload_c2_full_size(bpt, oracle, wl, qset):
    """64M unique uint32 keys, 256M queries (uniform hits / Zipf 0.99 with
    10% misses): plain and grouped equal the oracle bit for bit."""
    n, nq = 1 << 26, 1 << 28
    keys, vals = wl.point_keys(n, workload=2)
    q = wl.uniform_hits(keys, nq) if qset == "uniform" else wl.zipf_queries(keys, nq)
    print(f"c2 {qset} generated", flush=True)
    t = bpt.Tree(keys, vals)
    dq = bpt.DeviceArray.from_host(q)
    out = bpt.DeviceArray(nq, np.uint32)
    want = oracle.point_closed(keys, vals, q, 0xFFFFFFFF).astype(np.uint32)
    for mode in (bpt.PLAIN, bpt.GROUPED):
        out.fill_bytes(0)
        t.point_search_dev(dq, out, mode)
        got = out.to_host()
        assert int((got != want).sum()) == 0, (qset, mode)
    if qset == "uniform":
        assert np.array_equal(want[:1000], np.searchsorted(keys, q[:1000]).astype(np.uint32))
