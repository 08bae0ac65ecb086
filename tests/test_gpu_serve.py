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


def _device_requests(bpt, reqs, no_first=()):
    """The requests' inputs in device memory: one DeviceArray per buffer
    kind, each request a slice of it. Range requests listed in no_first ask
    for counts only."""
    pts = [r[1] for r in reqs if r[0] == "point"]
    rgs = [r for r in reqs if r[0] == "range"]
    cat = lambda xs: np.concatenate(xs) if xs else np.zeros(1, dtype=np.uint64)  # noqa: E731
    d = {"q": bpt.DeviceArray.from_host(cat(pts)), "lo": bpt.DeviceArray.from_host(cat([r[1] for r in rgs])),
         "hi": bpt.DeviceArray.from_host(cat([r[2] for r in rgs]))}
    d["out"] = bpt.DeviceArray(len(d["q"]), np.uint64)
    d["first"] = bpt.DeviceArray(len(d["lo"]), np.uint64)
    d["count"] = bpt.DeviceArray(len(d["lo"]), np.uint64)
    for a in (d["out"], d["first"], d["count"]):
        a.fill_bytes(0x5A)
    structs, where, pa, ra = [], [], 0, 0
    for i, r in enumerate(reqs):
        if r[0] == "point":
            structs.append(bpt.Server.point_request_dev(d["q"], d["out"], pa, len(r[1])))
            where.append(("point", pa, len(r[1])))
            pa += len(r[1])
        else:
            first = None if i in no_first else d["first"]
            structs.append(bpt.Server.range_request_dev(d["lo"], d["hi"], first, d["count"], ra, len(r[1])))
            where.append(("range", ra, len(r[1]), i in no_first))
            ra += len(r[1])
    return d, structs, where


def _device_results(d, where):
    out, first, count = d["out"].to_host(), d["first"].to_host(), d["count"].to_host()
    res = []
    for w in where:
        if w[0] == "point":
            res.append(out[w[1]:w[1] + w[2]])
        else:
            res.append((None if w[3] else first[w[1]:w[1] + w[2]], count[w[1]:w[1] + w[2]]))
    return res


@pytest.mark.parametrize("mode", [0, 2])
def test_device_requests_alone_and_fused(bpt, oracle, served_tree, mode):
    """Requests whose buffers are device memory: run alone (searched in
    place) and arriving together (fused runs gathered and scattered by the
    copy kernels), with some range requests asking for counts only. Every
    answer equals the oracle."""
    t, keys, vals = served_tree
    rng = np.random.default_rng(40 + mode)
    reqs = _requests(keys, 40, seed=50 + mode)
    for i in range(300):  # many small ones, so that runs of each kind fuse
        n = int(rng.integers(1, 300))
        if i % 29 >= 25:
            s_ = rng.integers(0, len(keys), n)
            reqs.append(("range", keys[s_].astype(np.uint64), keys[np.minimum(s_ + 77, len(keys) - 1)].astype(np.uint64)))
        else:
            reqs.append(("point", np.where(rng.random(n) < 0.8, keys[rng.integers(0, len(keys), n)],
                                           rng.integers(0, 1 << 50, n, dtype=np.uint64)).astype(np.uint64)))
    no_first = {i for i, r in enumerate(reqs) if r[0] == "range" and i % 3 == 0}
    for fused in (False, True):
        d, structs, where = _device_requests(bpt, reqs, no_first)
        srv = bpt.Server(t, slots=1 if not fused else 4, mem_budget=1 << 28, mode=mode)
        if fused:
            srv.replay(structs, np.zeros(len(structs)))
        else:
            for st_ in structs:
                srv.wait(srv.submit(st_))
        st = srv.stats()
        srv.close()
        got = _device_results(d, where)
        for a in d.values():
            a.free()
        assert st["completed"] == len(reqs)
        if fused:
            assert st["max_inflight"] > 4  # more requests in flight than slots: runs were fused
        for r, m in zip(reqs, got):
            if r[0] == "point":
                assert np.array_equal(m, oracle.point_closed(keys, vals, r[1], SENT))
            else:
                f, c = oracle.range_closed(keys, r[1], r[2])
                assert np.array_equal(m[1], c)
                if m[0] is None:
                    continue
                assert np.array_equal(m[0], f)


def test_device_requests_count_only_leave_first_untouched(bpt, served_tree):
    """Fused count-only range requests leave their slices of the lower-bound
    array as they were; the others fill theirs."""
    t, keys, vals = served_tree
    rng = np.random.default_rng(60)
    reqs = []
    for i in range(64):
        s_ = rng.integers(0, len(keys), 50)
        reqs.append(("range", keys[s_].astype(np.uint64), keys[np.minimum(s_ + 10, len(keys) - 1)].astype(np.uint64)))
    no_first = set(range(0, 64, 2))
    d, structs, where = _device_requests(bpt, reqs, no_first)
    srv = bpt.Server(t, slots=2, mem_budget=1 << 26, mode=0)
    srv.replay(structs, np.zeros(len(structs)))
    srv.close()
    first = d["first"].to_host()
    for w in where:
        sl = first[w[1]:w[1] + w[2]]
        if w[3]:
            assert (sl == 0x5A5A5A5A5A5A5A5A).all()
        else:
            assert (sl != 0x5A5A5A5A5A5A5A5A).all()
    for a in d.values():
        a.free()


def test_host_and_device_requests_in_one_stream(bpt, oracle, served_tree):
    """Host and device requests interleaved in one replay are never fused
    together; both kinds get the oracle's answers."""
    t, keys, vals = served_tree
    rng = np.random.default_rng(70)
    qs = [np.where(rng.random(n) < 0.9, keys[rng.integers(0, len(keys), n)],
                   rng.integers(0, 1 << 50, n, dtype=np.uint64)).astype(np.uint64)
          for n in rng.integers(1, 400, size=200)]
    d, dstructs, where = _device_requests(bpt, [("point", q) for q in qs[1::2]])
    houts = [np.zeros(len(q), dtype=np.uint64) for q in qs[0::2]]
    structs = []
    for i in range(len(qs)):
        structs.append(bpt.Server.point_request(qs[i], houts[i // 2]) if i % 2 == 0 else dstructs[i // 2])
    srv = bpt.Server(t, slots=2, mem_budget=1 << 26, mode=0)
    srv.replay(structs, np.zeros(len(structs)))
    assert srv.stats()["completed"] == len(qs)
    srv.close()
    dres = _device_results(d, where)
    for a in d.values():
        a.free()
    for i, q in enumerate(qs):
        got = houts[i // 2] if i % 2 == 0 else dres[i // 2]
        assert np.array_equal(got, oracle.point_closed(keys, vals, q, SENT))


@pytest.mark.parametrize("fused", [False, True])
def test_pinned_host_requests_without_staging(bpt, oracle, served_tree, fused):
    """BPT_REQ_DEVICE requests whose buffers are pinned host memory from
    bpt_host_alloc: the search (alone) or the gather and scatter kernels
    (fused) reach them over PCIe, with no staging copy. Every answer equals
    the oracle."""
    t, keys, vals = served_tree
    rng = np.random.default_rng(80 + fused)
    sizes = rng.integers(1, 3000, size=120)
    q = np.where(rng.random(int(sizes.sum())) < 0.85, keys[rng.integers(0, len(keys), int(sizes.sum()))],
                 rng.integers(0, 1 << 50, int(sizes.sum()), dtype=np.uint64)).astype(np.uint64)
    s_ = rng.integers(0, len(keys), 40 * 64)
    lo, hi = keys[s_].astype(np.uint64), keys[np.minimum(s_ + 500, len(keys) - 1)].astype(np.uint64)
    hq, hout = bpt.HostArray.from_host(q), bpt.HostArray(len(q), np.uint64)
    hlo, hhi = bpt.HostArray.from_host(lo), bpt.HostArray.from_host(hi)
    hfirst, hcount = bpt.HostArray(len(lo), np.uint64), bpt.HostArray(len(lo), np.uint64)
    hout.a[:] = 7
    structs, at = [], 0
    for m in sizes:
        structs.append(bpt.Server.point_request_dev(hq, hout, at, int(m)))
        at += int(m)
    for r in range(40):
        structs.insert(3 * r, bpt.Server.range_request_dev(hlo, hhi, hfirst, hcount, 64 * r, 64))
    srv = bpt.Server(t, slots=4 if fused else 1, mem_budget=1 << 28, mode=0)
    if fused:
        srv.replay(structs, np.zeros(len(structs)))
    else:
        for st_ in structs:
            srv.wait(srv.submit(st_))
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(structs)
    assert np.array_equal(hout.a, oracle.point_closed(keys, vals, q, SENT))
    f, c = oracle.range_closed(keys, lo, hi)
    assert np.array_equal(hcount.a, c) and np.array_equal(hfirst.a, f)
    for a in (hq, hout, hlo, hhi, hfirst, hcount):
        a.free()


def test_point_search_in_pinned_host_memory(bpt, oracle, served_tree):
    """bpt_point_search straight from and into pinned host memory (the
    kernel reads the queries and writes the results over PCIe)."""
    t, keys, vals = served_tree
    rng = np.random.default_rng(90)
    q = np.where(rng.random(100_003) < 0.9, keys[rng.integers(0, len(keys), 100_003)],
                 rng.integers(0, 1 << 50, 100_003, dtype=np.uint64)).astype(np.uint64)
    hq, hout = bpt.HostArray.from_host(q), bpt.HostArray(len(q), np.uint64)
    t.point_search_dev(hq, hout, bpt.PLAIN)
    bpt.sync()
    assert np.array_equal(hout.a, oracle.point_closed(keys, vals, q, SENT))
    hq.free()
    hout.free()


def test_device_request_checks(bpt, served_tree):
    """Bad flags and device buffers not aligned to their element width are
    rejected at submit, before any launch."""
    t, keys, vals = served_tree
    dq = bpt.DeviceArray.from_host(keys[:64])
    dout = bpt.DeviceArray(64, np.uint64)
    srv = bpt.Server(t, slots=1, mem_budget=1 << 24, mode=0)
    bad = bpt.Server.point_request_dev(dq, dout)
    bad.flags = 2
    with pytest.raises(bpt.BptError) as e:
        srv.submit(bad)
    assert e.value.status == 1 and "flags" in str(e.value)
    mis = bpt.Request(bpt.REQ_POINT, bpt.REQ_DEVICE, 8, dq.ptr + 4, dout.ptr, None, None, None, None)
    with pytest.raises(bpt.BptError) as e:
        srv.submit(mis)
    assert e.value.status == 1 and "aligned" in str(e.value)
    srv.close()
    dq.free()
    dout.free()


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
