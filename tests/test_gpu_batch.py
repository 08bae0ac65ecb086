"""GPU parity for multi-tree batches: one persistent launch over thousands of
trees (point + range), eager and hipGraph replays, every autotuned fanout,
against the oracle on the device-generated inputs copied back."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SENT = 0xFFFFFFFFFFFFFFFF


def _setup(bpt, wl, ntrees=300, lo=4, hi=14, seed=7):
    sizes = wl.forest_sizes(ntrees, lo, hi, seed=seed)
    key_off, r_off = wl.forest_offsets(sizes)
    dk, dv, dq, dlo, dhi = bpt.synth_forest(seed, key_off, qpk=4, hit_pct=80, r_off=r_off)
    q_off = key_off * np.uint64(4)
    return sizes, key_off, r_off, q_off, dk, dv, dq, dlo, dhi


def _expected(oracle, keys, vals, key_off, q, q_off, lo, hi, r_off, trees):
    out = {}
    for t in trees:
        k = keys[key_off[t]:key_off[t + 1]]
        v = vals[key_off[t]:key_off[t + 1]]
        pq = q[q_off[t]:q_off[t + 1]]
        f, c = oracle.range_closed(k, lo[r_off[t]:r_off[t + 1]], hi[r_off[t]:r_off[t + 1]])
        out[t] = (oracle.point_closed(k, v, pq, SENT), f, c)
    return out


def test_generator_matches_restatement(bpt, wl):
    sizes, key_off, r_off, q_off, dk, dv, dq, dlo, dhi = _setup(bpt, wl, ntrees=40)
    keys = dk.to_host()
    for t in range(40):
        want = wl.synth_tree_keys(7, t, int(sizes[t]))
        assert np.array_equal(keys[key_off[t]:key_off[t + 1]], want)
        assert (np.diff(want.astype(np.float64)) > 0).all()
    vals = dv.to_host()
    assert np.array_equal(vals[key_off[5]:key_off[6]], np.arange(int(sizes[5]), dtype=np.uint64))


def test_generator_distributions(bpt, wl):
    """Keys uniform over [0, 2^64), unique and sorted per tree; 80 % of the
    point queries are keys of their tree; range widths log-uniform in
    [2^4, 2^16] ranks (uniform log2 of the unclipped counts)."""
    sizes = np.full(8, 1 << 20, dtype=np.uint64)
    key_off, r_off = wl.forest_offsets(sizes, ranges_div=4)
    dk, dv, dq, dlo, dhi = bpt.synth_forest(3, key_off, qpk=4, hit_pct=80, r_off=r_off)
    keys, q, lo, hi = dk.to_host(), dq.to_host(), dlo.to_host(), dhi.to_host()
    u = []
    hits = 0
    logw = []
    for t in range(8):
        k = keys[key_off[t]:key_off[t + 1]]
        assert (k[1:] > k[:-1]).all()
        x = k.astype(np.float64) / 2.0 ** 64
        # sorted uniform draws: the i-th lies near (i + 1) / (n + 1) (max
        # deviation of the empirical CDF, Kolmogorov-Smirnov: above
        # 2.5 / sqrt(n) with probability about 2 exp(-12.5))
        n = len(k)
        assert np.max(np.abs(x - (np.arange(n) + 1) / (n + 1))) < 2.5 / np.sqrt(n)
        u.append(x.mean())
        tq = q[4 * key_off[t]:4 * key_off[t + 1]]
        hits += int(np.isin(tq, k).sum())
        a = np.searchsorted(k, lo[r_off[t]:r_off[t + 1]])
        b = np.searchsorted(k, hi[r_off[t]:r_off[t + 1]])
        w = (b - a + 1).astype(np.float64)
        logw.append(np.log2(w[a + w < n]))        # ranges not clipped by the tree's end
    assert abs(np.mean(u) - 0.5) < 0.01
    assert 0.79 < hits / (4 * int(key_off[-1])) < 0.81
    lw = np.concatenate(logw)
    assert lw.min() >= 4 and lw.max() <= 16
    # log2 of floor(2^u) lies within 1 of u, so the deciles of a uniform u on
    # [4, 16] are 5.2, 6.4, ... within about a unit
    for qt, want in zip(np.quantile(lw, [0.1, 0.5, 0.9]), (5.2, 10.0, 14.8)):
        assert abs(qt - want) < 0.6, (qt, want)


@pytest.mark.parametrize("fanout,lds,lanes", [(16, 16384, 8), (32, 16384, 8), (64, 16384, 8), (16, 65536, 8),
                                              (32, 65536, 8), (64, 65536, 8), (16, 32768, 4), (32, 16384, 4),
                                              (64, 65536, 4)])
def test_batch_matches_oracle(bpt, oracle, wl, fanout, lds, lanes):
    sizes, key_off, r_off, q_off, dk, dv, dq, dlo, dhi = _setup(bpt, wl)
    nq, nr = int(q_off[-1]), int(r_off[-1])
    f = bpt.Forest(dk, dv, key_off, fanout=fanout, lds_bytes=lds, lanes=lanes)
    dout = bpt.DeviceArray(nq, np.uint64)
    dfirst = bpt.DeviceArray(nr, np.uint64)
    dcount = bpt.DeviceArray(nr, np.uint64)
    keys, vals, q, lo, hi = dk.to_host(), dv.to_host(), dq.to_host(), dlo.to_host(), dhi.to_host()
    exp = _expected(oracle, keys, vals, key_off, q, q_off, lo, hi, r_off, range(len(sizes)))
    for chunk in (512, 8192):
        b = bpt.Batch(f, dq, q_off, dout, dlo, dhi, r_off, dfirst, dcount, chunk=chunk)
        for graph in (False, True, True, True):  # eager, then three replays of one graph
            dout.fill_bytes(0)
            dcount.fill_bytes(0xAB)
            b.run(graph=graph)
            bpt.sync()
            out, first, count = dout.to_host(), dfirst.to_host(), dcount.to_host()
            for t, (po, fe, ce) in exp.items():
                assert np.array_equal(out[q_off[t]:q_off[t + 1]], po), (t, chunk, graph)
                assert np.array_equal(first[r_off[t]:r_off[t + 1]], fe), (t, chunk, graph)
                assert np.array_equal(count[r_off[t]:r_off[t + 1]], ce), (t, chunk, graph)
        b.close()
    hits = np.mean(np.concatenate([exp[t][0] for t in exp]) != SENT)
    assert 0.75 < hits < 0.85


def test_autotune_returns_fastest_valid_config(bpt, oracle, wl):
    sizes, key_off, r_off, q_off, dk, dv, dq, dlo, dhi = _setup(bpt, wl, ntrees=64, lo=8, hi=16)
    nq, nr = int(q_off[-1]), int(r_off[-1])
    dout = bpt.DeviceArray(nq, np.uint64)
    dfirst = bpt.DeviceArray(nr, np.uint64)
    dcount = bpt.DeviceArray(nr, np.uint64)
    best, report = bpt.forest_autotune(dk, dv, key_off, dq, q_off, dout, dlo, dhi, r_off, dfirst, dcount, reps=1)
    assert len(report) == 54
    assert {r["queries_per_wave"] for r in report} == {8, 16}
    assert best["fanout"] in (16, 32, 64) and best["lanes"] in (4, 8) and min(r["ms"] for r in report) > 0
    f = bpt.Forest(dk, dv, key_off, fanout=best["fanout"], lds_bytes=best["lds_bytes"], lanes=best["lanes"])
    b = bpt.Batch(f, dq, q_off, dout, dlo, dhi, r_off, dfirst, dcount, chunk=best["chunk"])
    b.run(graph=True)
    bpt.sync()
    keys, vals, q = dk.to_host(), dv.to_host(), dq.to_host()
    exp = _expected(oracle, keys, vals, key_off, q, q_off, dlo.to_host(), dhi.to_host(), r_off, range(64))
    out, count = dout.to_host(), dcount.to_host()
    for t, (po, fe, ce) in exp.items():
        assert np.array_equal(out[q_off[t]:q_off[t + 1]], po)
        assert np.array_equal(count[r_off[t]:r_off[t + 1]], ce)


def test_workload_c4_full_size(bpt, oracle, wl):
    """The benchmarked forest: 4096 uint64 trees of 2^10..2^20 keys (632M in
    all), keys uniform over [0, 2^64), 4 point queries per key (80 % hits)
    and size/16 ranges of log-uniform width per tree, one graph-captured
    launch; every tree's points and ranges against the oracle, its slices of
    the device inputs streamed to the host a group of trees at a time."""
    sizes = wl.forest_sizes(4096, 10, 20)
    key_off, r_off = wl.forest_offsets(sizes)
    dk, dv, dq, dlo, dhi = bpt.synth_forest(0, key_off, qpk=4, hit_pct=80, r_off=r_off)
    q_off = key_off * np.uint64(4)
    nq, nr = int(q_off[-1]), int(r_off[-1])
    print(f"c4 generated: {int(key_off[-1])} keys, {nq} points, {nr} ranges", flush=True)
    f = bpt.Forest(dk, dv, key_off, fanout=16, lds_bytes=32768)
    dout = bpt.DeviceArray(nq, np.uint64)
    dfirst = bpt.DeviceArray(nr, np.uint64)
    dcount = bpt.DeviceArray(nr, np.uint64)
    dout.fill_bytes(0)
    dcount.fill_bytes(0xAB)
    b = bpt.Batch(f, dq, q_off, dout, dlo, dhi, r_off, dfirst, dcount, chunk=2048)
    b.run(graph=True)
    bpt.sync()
    b.close()
    f.close()
    hits = 0
    t0 = 0
    ko, qo, ro = key_off.astype(np.int64), q_off.astype(np.int64), r_off.astype(np.int64)
    while t0 < 4096:
        t1 = t0 + 1
        while t1 < 4096 and ko[t1 + 1] - ko[t0] <= (1 << 24):
            t1 += 1
        k = dk.to_host(offset=int(ko[t0]), count=int(ko[t1] - ko[t0]))
        v = dv.to_host(offset=int(ko[t0]), count=int(ko[t1] - ko[t0]))
        q = dq.to_host(offset=int(qo[t0]), count=int(qo[t1] - qo[t0]))
        out = dout.to_host(offset=int(qo[t0]), count=int(qo[t1] - qo[t0]))
        lo = dlo.to_host(offset=int(ro[t0]), count=int(ro[t1] - ro[t0]))
        hi = dhi.to_host(offset=int(ro[t0]), count=int(ro[t1] - ro[t0]))
        first = dfirst.to_host(offset=int(ro[t0]), count=int(ro[t1] - ro[t0]))
        count = dcount.to_host(offset=int(ro[t0]), count=int(ro[t1] - ro[t0]))
        for t in range(t0, t1):
            ks = slice(ko[t] - ko[t0], ko[t + 1] - ko[t0])
            qs = slice(qo[t] - qo[t0], qo[t + 1] - qo[t0])
            rs = slice(ro[t] - ro[t0], ro[t + 1] - ro[t0])
            want = oracle.point_closed(k[ks], v[ks], q[qs], SENT)
            assert np.array_equal(out[qs], want), t
            fe, ce = oracle.range_closed(k[ks], lo[rs], hi[rs])
            assert np.array_equal(first[rs], fe), t
            assert np.array_equal(count[rs], ce), t
            hits += int((want != SENT).sum())
        t0 = t1
    print(f"c4 checked: all 4096 trees, {nq} points, {nr} ranges", flush=True)
    assert 0.79 < hits / nq < 0.81
