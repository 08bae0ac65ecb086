"""GPU: batched upserts and erases (bpt_tree_upsert / bpt_tree_erase) against
a numpy model of the same semantics; after every step the plain and grouped
searches, the lower bound and range counts equal the oracle on the model."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Model:
    """Sorted (key, value) arrays with the update semantics of the C ABI."""

    def __init__(self, keys, vals):
        o = np.argsort(keys, kind="stable")
        self.k, self.v = keys[o], vals[o]

    def upsert(self, k, v):
        rk = k[::-1]
        _, first = np.unique(rk, return_index=True)   # last occurrence of every key
        last = len(k) - 1 - first
        k, v = k[last], v[last]
        keep = ~np.isin(self.k, k)
        ak, av = np.concatenate([self.k[keep], k]), np.concatenate([self.v[keep], v])
        o = np.argsort(ak, kind="stable")
        self.k, self.v = ak[o], av[o]

    def erase(self, k):
        keep = ~np.isin(self.k, k)
        self.k, self.v = self.k[keep], self.v[keep]


def _check(bpt, oracle, t, m, rng, kb, vb, step):
    kdt, vdt = m.k.dtype.type, m.v.dtype.type
    top = (1 << (8 * kb)) - 1
    pool = m.k if len(m.k) else np.zeros(1, m.k.dtype)
    q = np.concatenate([pool[rng.integers(0, len(pool), 20000)],
                        rng.integers(0, top, 20000, dtype=np.uint64).astype(kdt),
                        np.array([0, top], dtype=np.uint64).astype(kdt)])
    sent = (1 << (8 * vb)) - 1
    want = oracle.point_closed(m.k, m.v, q, sent).astype(vdt)
    print(f"step {step}: n={len(m.k)} info={t.info()['n']}", flush=True)
    assert t.info()["n"] == len(m.k), step
    assert np.array_equal(t.point_search(q, bpt.PLAIN), want), step
    if kb == vb:
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want), step
    assert np.array_equal(t.lower_bound(q), oracle.lower_bound(m.k, q)), step
    lo = q[:5000]
    hi = (lo.astype(np.uint64) + rng.integers(0, 1 << 20, 5000, dtype=np.uint64)).astype(kdt)
    hi = np.maximum(hi, lo)
    _, cnt, _, _ = t.range_search(lo, hi, with_ids=False)
    _, want_c = oracle.range_closed(m.k, lo, hi)
    assert np.array_equal(cnt, want_c), step


@pytest.mark.parametrize("kb,vb,F,LF", [(4, 4, 32, 16), (8, 8, 16, 8), (8, 4, 32, 8)],
                         ids=lambda c: str(c))
def test_upsert_erase_sequence(bpt, oracle, kb, vb, F, LF):
    rng = np.random.default_rng(kb * 10 + vb + F)
    kdt = np.uint32 if kb == 4 else np.uint64
    vdt = np.uint32 if vb == 4 else np.uint64
    top = (1 << (8 * kb)) - 1
    n = 300_000
    keys = np.sort(rng.integers(0, top, n, dtype=np.uint64).astype(kdt))
    keys[1000:1040] = keys[1000]                      # a run of duplicate keys in the built tree
    vals = rng.integers(0, (1 << (8 * vb)) - 1, n, dtype=np.uint64).astype(vdt)
    t = bpt.Tree(keys, vals, kb, vb, fanout=F, leaf_cap=LF)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    m = Model(keys, vals)

    def batch(size, frac_existing):
        ne = int(size * frac_existing)
        k = np.concatenate([m.k[rng.integers(0, len(m.k), ne)] if len(m.k) else np.zeros(0, kdt),
                            rng.integers(0, top, size - ne, dtype=np.uint64).astype(kdt)])
        k = np.concatenate([k, k[: size // 10]])     # repeated keys in the batch
        rng.shuffle(k)
        return k, rng.integers(0, (1 << (8 * vb)) - 1, len(k), dtype=np.uint64).astype(vdt)

    k, v = batch(50_000, 0.5)
    k = np.concatenate([k, keys[1000:1001]])          # upsert the duplicated key: one slot remains
    v = np.concatenate([v, v[:1]])
    t.upsert(k, v); m.upsert(k, v)
    _check(bpt, oracle, t, m, rng, kb, vb, "upsert")
    assert int((m.k == keys[1000]).sum()) == 1

    k, _ = batch(40_000, 0.5)
    t.erase(k); m.erase(k)
    _check(bpt, oracle, t, m, rng, kb, vb, "erase")

    k2, v2 = k[:20_000], rng.integers(0, (1 << (8 * vb)) - 1, 20_000, dtype=np.uint64).astype(vdt)
    t.upsert(k2, v2); m.upsert(k2, v2)               # put back part of what was erased
    _check(bpt, oracle, t, m, rng, kb, vb, "re-upsert")

    t.upsert(k2[:1], v2[:1]); m.upsert(k2[:1], v2[:1])  # a batch of one
    _check(bpt, oracle, t, m, rng, kb, vb, "single")

    # many keys at one place: a run inside the widest gap and a run above the maximum
    gi = int(np.argmax(np.diff(m.k[:20_000].astype(np.float64))))
    g = int(min(int(m.k[gi + 1]) - int(m.k[gi]) - 1, 3000))
    inside = (np.uint64(m.k[gi]) + np.arange(1, g + 1, dtype=np.uint64)).astype(kdt)
    above = (np.uint64(top) - np.arange(3000, dtype=np.uint64)).astype(kdt)
    k3 = np.concatenate([inside, above])
    v3 = rng.integers(0, (1 << (8 * vb)) - 1, len(k3), dtype=np.uint64).astype(vdt)
    t.upsert(k3, v3); m.upsert(k3, v3)
    _check(bpt, oracle, t, m, rng, kb, vb, "clustered")

    allk = m.k.copy()
    t.erase(allk); m.erase(allk)                      # everything: an empty tree
    assert t.info()["n"] == 0
    assert np.all(t.point_search(allk[:1000]) == (1 << (8 * vb)) - 1)

    k, v = batch(30_000, 0.0)
    t.upsert(k, v); m.upsert(k, v)                    # an empty tree grows again
    _check(bpt, oracle, t, m, rng, kb, vb, "regrow")
    t.close()


def test_updates_keep_options_and_drop_learned(bpt, oracle):
    rng = np.random.default_rng(3)
    keys = np.unique(rng.integers(0, 1 << 62, 400_000, dtype=np.uint64))
    vals = np.arange(len(keys), dtype=np.uint64)
    t = bpt.Tree(keys, vals, 8, 8, fanout=16, leaf_cap=8)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    t.set_option("lanes", 4)
    t.learned_build(leaf_models=1024, train_steps=50, train_samples=1 << 14)
    m = Model(keys, vals)
    k = rng.integers(0, 1 << 62, 10_000, dtype=np.uint64)
    v = rng.integers(0, 1 << 62, 10_000, dtype=np.uint64)
    t.upsert(k, v); m.upsert(k, v)
    q = np.concatenate([k, keys[:1000]])
    assert np.array_equal(t.point_search(q), oracle.point_closed(m.k, m.v, q, (1 << 64) - 1))
    assert t.info()["fanout"] == 16 and t.info()["leaf_cap"] == 8
    with pytest.raises(bpt.BptError):                 # the model no longer matches the keys
        t.point_search(q, bpt.LEARNED)
    t.close()


def test_failed_rebuild_allocation_leaves_tree_unchanged(bpt, oracle):
    """An upsert whose rebuild needs a larger leaf buffer than the tree may
    allocate ("alloc_limit", forced from the host) fails with BPT_ERR_NOMEM
    and leaves every key, value and the learned model as they were; with the
    limit lifted the same upsert goes through."""
    rng = np.random.default_rng(5)
    keys = np.unique(rng.integers(0, 1 << 62, 200_000, dtype=np.uint64))
    vals = np.arange(len(keys), dtype=np.uint64)
    t = bpt.Tree(keys, vals, 8, 8, fanout=16, leaf_cap=8)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    t.learned_build(leaf_models=1024, train_steps=50, train_samples=1 << 14)
    t.set_option("alloc_limit", t.info()["leaf_bytes"])   # the current leaf buffer, no more
    k = rng.integers(0, 1 << 62, 100_000, dtype=np.uint64)   # grows the tree by half
    v = rng.integers(0, 1 << 62, 100_000, dtype=np.uint64)
    with pytest.raises(bpt.BptError) as e:
        t.upsert(k, v)
    assert e.value.status == 3                        # BPT_ERR_NOMEM
    q = np.concatenate([keys[rng.integers(0, len(keys), 50_000)], k[:20_000]])
    want = oracle.point_closed(keys, vals, q, (1 << 64) - 1)
    assert t.info()["n"] == len(keys)
    for mode in (bpt.PLAIN, bpt.GROUPED, bpt.LEARNED):
        assert np.array_equal(t.point_search(q, mode), want), mode
    t.set_option("alloc_limit", 0)
    t.upsert(k, v)
    m = Model(keys, vals)
    m.upsert(k, v)
    assert t.info()["n"] == len(m.k)
    assert np.array_equal(t.point_search(q), oracle.point_closed(m.k, m.v, q, (1 << 64) - 1))
    t.close()


def test_grouped_tables_follow_updates(bpt, oracle):
    """Grouped search reads the bucket tables the tree builds with itself and
    rebuilds after every update; a plan with other options builds its own per
    call. Both equal the oracle while updates change the number of buckets
    (uint32 tree: 262144 keys per bucket, so 1, then 3, then 1 bucket)."""
    rng = np.random.default_rng(11)
    keys = np.unique(rng.integers(0, 1 << 32, 100_000, dtype=np.uint64)).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, len(keys), dtype=np.uint64).astype(np.uint32)
    t = bpt.Tree(keys, vals, 4, 4)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    m = Model(keys, vals)

    def check(step):
        q = np.concatenate([m.k[rng.integers(0, len(m.k), 40_000)],
                            rng.integers(0, 1 << 32, 10_000, dtype=np.uint64).astype(np.uint32)])
        want = oracle.point_closed(m.k, m.v, q, (1 << 32) - 1).astype(np.uint32)
        print(f"step {step}: n={len(m.k)}", flush=True)
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want), step
        t.set_option("group_slice", 65536)            # another plan: tables built per call
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want), step
        t.set_option("group_slice", 131072)
        assert np.array_equal(t.point_search(q, bpt.GROUPED), want), step

    check("build")
    k = rng.integers(0, 1 << 32, 500_000, dtype=np.uint64).astype(np.uint32)
    v = rng.integers(0, 1 << 32, len(k), dtype=np.uint64).astype(np.uint32)
    t.upsert(k, v)
    m.upsert(k, v)
    check("upsert")
    e = m.k[rng.permutation(len(m.k))[:450_000]]
    t.erase(e)
    m.erase(e)
    check("erase")
    t.close()
