"""GPU: the continuous-batching serving loop. Each request's answer is the
same alone and in the mix (and equals the oracle); admission by device
memory queues requests instead of failing them; in-flight requests overlap."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SENT = 0xFFFFFFFFFFFFFFFF


def _requests(keys, nreq, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(nreq):
        n = int(np.floor(np.exp2(rng.uniform(0, 14))))
        if rng.random() < 0.9:
            q = np.where(rng.random(n) < 0.9, keys[rng.integers(0, len(keys), n)],
                         rng.integers(0, 1 << 50, n, dtype=np.uint64)).astype(np.uint64)
            reqs.append(("point", q))
        else:
            s = rng.integers(0, len(keys), n)
            w = np.floor(np.exp2(rng.uniform(4, 12, n))).astype(np.int64)
            lo = keys[s]
            hi = keys[np.minimum(s + w, len(keys) - 1)]
            reqs.append(("range", lo.astype(np.uint64), hi.astype(np.uint64)))
    return reqs


def _run(bpt, srv, reqs, arrival=None):
    outs, structs, keep = [], [], []
    for r in reqs:
        if r[0] == "point":
            out = np.zeros(len(r[1]), dtype=np.uint64)
            structs.append(bpt.Server.point_request(r[1], out))
            outs.append(out)
        else:
            first = np.zeros(len(r[1]), dtype=np.uint64)
            cnt = np.zeros(len(r[1]), dtype=np.uint64)
            structs.append(bpt.Server.range_request(r[1], r[2], first, cnt))
            outs.append((first, cnt))
    if arrival is None:  # one at a time
        for s in structs:
            srv.wait(srv.submit(s))
    else:
        srv.replay(structs, arrival)
    return outs


@pytest.fixture(scope="module")
def served_tree(bpt, wl):
    keys = wl.lognormal_keys(1 << 20, seed=8)
    vals = np.arange(len(keys), dtype=np.uint64)
    t = bpt.Tree(keys, vals, 8, 8)
    t.learned_build(leaf_models=8192, train_steps=300)
    return t, keys, vals


@pytest.mark.parametrize("mode", [0, 2])
def test_same_alone_and_in_the_mix(bpt, oracle, served_tree, mode):
    t, keys, vals = served_tree
    reqs = _requests(keys, 60, seed=mode)
    alone_srv = bpt.Server(t, slots=1, mem_budget=1 << 28, mode=mode)
    alone = _run(bpt, alone_srv, reqs)
    alone_srv.close()
    srv = bpt.Server(t, slots=6, mem_budget=1 << 28, mode=mode)
    rng = np.random.default_rng(1)
    arrival = np.cumsum(rng.exponential(20e-6, size=len(reqs)))   # Poisson, 50K requests/s
    mixed = _run(bpt, srv, reqs, arrival)
    st = srv.stats()
    srv.close()
    for r, a, m in zip(reqs, alone, mixed):
        if r[0] == "point":
            assert np.array_equal(a, m)
            assert np.array_equal(m, oracle.point_closed(keys, vals, r[1], SENT))
        else:
            assert np.array_equal(a[0], m[0]) and np.array_equal(a[1], m[1])
            f, c = oracle.range_closed(keys, r[1], r[2])
            assert np.array_equal(m[1], c) and np.array_equal(m[0], f)
    assert st["completed"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0


def test_memory_admission_queues_instead_of_failing(bpt, oracle, served_tree):
    t, keys, vals = served_tree
    rng = np.random.default_rng(2)
    qs = [keys[rng.integers(0, len(keys), 50_000)] for _ in range(12)]
    need = 50_000 * 16  # keys + values of one request
    srv = bpt.Server(t, slots=8, mem_budget=need + need // 2, mode=0)  # room for one at a time
    outs = [np.zeros(len(q), dtype=np.uint64) for q in qs]
    tickets = [srv.submit(bpt.Server.point_request(q, o)) for q, o in zip(qs, outs)]
    for tk in tickets:
        srv.wait(tk)
    st = srv.stats()
    assert st["completed"] == 12 and st["max_inflight"] == 1 and st["peak_bytes"] <= st["budget_bytes"]
    for q, o in zip(qs, outs):
        assert np.array_equal(o, oracle.point_closed(keys, vals, q, SENT))
    big = keys[:100_000]
    with pytest.raises(bpt.BptError) as e:
        srv.submit(bpt.Server.point_request(big, np.zeros(len(big), dtype=np.uint64)))
    assert e.value.status == 5
    srv.close()


def test_requests_overlap_on_the_device(bpt, served_tree):
    t, keys, vals = served_tree
    rng = np.random.default_rng(3)
    qs = [keys[rng.integers(0, len(keys), 200_000)] for _ in range(32)]
    outs = [np.zeros(len(q), dtype=np.uint64) for q in qs]
    srv = bpt.Server(t, slots=8, mem_budget=1 << 30, mode=0)
    lat = srv.replay([bpt.Server.point_request(q, o) for q, o in zip(qs, outs)], np.zeros(len(qs)))
    st = srv.stats()
    srv.close()
    assert st["max_inflight"] > 1
    assert (lat > 0).all()


@pytest.mark.parametrize("mode", [0, 2])
def test_small_point_requests_are_fused(bpt, oracle, served_tree, mode):
    """Hundreds of small requests arriving together are answered in fused
    launches of one kind (more requests in flight than slots): runs of point
    requests, and of range requests, between each other; every answer equals
    the oracle."""
    t, keys, vals = served_tree
    rng = np.random.default_rng(4 + mode)
    reqs = []
    for i in range(400):
        if i % 37 >= 33:  # runs of four range requests between the point requests
            s = rng.integers(0, len(keys), int(rng.integers(1, 60)))
            reqs.append(("range", keys[s].astype(np.uint64),
                         keys[np.minimum(s + 100, len(keys) - 1)].astype(np.uint64)))
        else:
            n = int(rng.integers(1, 200))
            reqs.append(("point", np.where(rng.random(n) < 0.8, keys[rng.integers(0, len(keys), n)],
                                           rng.integers(0, 1 << 50, n, dtype=np.uint64)).astype(np.uint64)))
    srv = bpt.Server(t, slots=2, mem_budget=1 << 26, mode=mode)
    outs = _run(bpt, srv, reqs, np.zeros(len(reqs)))
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["max_inflight"] > 2
    for r, m in zip(reqs, outs):
        if r[0] == "point":
            assert np.array_equal(m, oracle.point_closed(keys, vals, r[1], SENT))
        else:
            f, c = oracle.range_closed(keys, r[1], r[2])
            assert np.array_equal(m[1], c) and np.array_equal(m[0], f)


@pytest.fixture(scope="module")
def c5_tree(bpt, wl):
    """The learned-index configuration's tree: 256M unique lognormal uint64
    keys, value = rank, with the learned model trained as in the benchmark."""
    n = 1 << 28
    keys = wl.lognormal_keys(n, sort=lambda a: bpt.sort_pairs(a)[0])
    vals = np.arange(n, dtype=np.uint64)
    print("c5 serving: keys generated", flush=True)
    t = bpt.Tree(keys, vals, 8, 8)
    t.learned_build(leaf_models=65536, max_window=256, train_samples=1 << 20, train_steps=2000, root=0)
    yield t, keys, vals
    t.close()


@pytest.mark.parametrize("mode", [0, 2])
def test_config_scale_serving_same_alone_and_in_the_mix(bpt, oracle, c5_tree, mode):
    """At the configuration's scale: batches of floor(2^u) queries, u uniform
    in [0, 16] (1 to 64K), 90% point (10% misses) / 10% range of 2^4..2^12
    keys, over the 256M-key tree. Every answer is the same run alone (one
    slot, one request at a time) and in an open-loop Poisson mix on 8 slots,
    and equals the oracle."""
    t, keys, vals = c5_tree
    rng = np.random.default_rng(30 + mode)
    reqs = []
    for _ in range(300):
        m = int(np.floor(np.exp2(rng.uniform(0, 16))))
        if rng.random() < 0.9:
            q = np.where(rng.random(m) < 0.9, keys[rng.integers(0, len(keys), m)],
                         rng.integers(0, 1 << 50, m, dtype=np.uint64)).astype(np.uint64)
            reqs.append(("point", q))
        else:
            s = rng.integers(0, len(keys), m)
            w = np.floor(np.exp2(rng.uniform(4, 12, m))).astype(np.int64)
            reqs.append(("range", keys[s].astype(np.uint64), keys[np.minimum(s + w, len(keys) - 1)].astype(np.uint64)))
    alone_srv = bpt.Server(t, slots=1, mem_budget=1 << 30, mode=mode)
    alone = _run(bpt, alone_srv, reqs)
    alone_srv.close()
    srv = bpt.Server(t, slots=8, mem_budget=8 << 30, mode=mode)
    arrival = np.cumsum(rng.exponential(10e-6, size=len(reqs)))   # Poisson, 100K requests/s
    mixed = _run(bpt, srv, reqs, arrival)
    st = srv.stats()
    srv.close()
    print(f"c5 serving mode {mode}: {st}", flush=True)
    assert st["completed"] == len(reqs) and st["max_inflight"] > 1
    pq = [r[1] for r in reqs if r[0] == "point"]
    want_p = np.split(oracle.point_closed(keys, vals, np.concatenate(pq), SENT), np.cumsum([len(x) for x in pq])[:-1])
    k = 0
    for r, a, m in zip(reqs, alone, mixed):
        if r[0] == "point":
            assert np.array_equal(a, m) and np.array_equal(m, want_p[k])
            k += 1
        else:
            assert np.array_equal(a[0], m[0]) and np.array_equal(a[1], m[1])
            f, c = oracle.range_closed(keys, r[1], r[2])
            assert np.array_equal(m[1], c) and np.array_equal(m[0], f)
