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


@pytest.mark.parametrize("fanout", [16, 32, 64])
@pytest.mark.parametrize("lds", [16384, 65536])
def test_batch_matches_oracle(bpt, oracle, wl, fanout, lds):
    sizes, key_off, r_off, q_off, dk, dv, dq, dlo, dhi = _setup(bpt, wl)
    nq, nr = int(q_off[-1]), int(r_off[-1])
    f = bpt.Forest(dk, dv, key_off, fanout=fanout, lds_bytes=lds)
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
    assert len(report) == 27
    assert best["fanout"] in (16, 32, 64) and min(r["ms"] for r in report) > 0
    f = bpt.Forest(dk, dv, key_off, fanout=best["fanout"], lds_bytes=best["lds_bytes"])
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
    all), 4 point queries per key and size/16 ranges per tree, one
    graph-captured launch; 32 trees (the smallest, the largest and a spread)
    against the oracle on their slices of the device inputs."""
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
    rng = np.random.default_rng(11)
    trees = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), *rng.choice(len(sizes), 30, replace=False).tolist()})
    hits = []
    for t in trees:
        k0, k1 = int(key_off[t]), int(key_off[t + 1])
        q0, q1 = int(q_off[t]), int(q_off[t + 1])
        r0, r1 = int(r_off[t]), int(r_off[t + 1])
        k, v = dk.to_host(offset=k0, count=k1 - k0), dv.to_host(offset=k0, count=k1 - k0)
        q = dq.to_host(offset=q0, count=q1 - q0)
        want = oracle.point_closed(k, v, q, SENT)
        assert np.array_equal(dout.to_host(offset=q0, count=q1 - q0), want), t
        fe, ce = oracle.range_closed(k, dlo.to_host(offset=r0, count=r1 - r0), dhi.to_host(offset=r0, count=r1 - r0))
        assert np.array_equal(dfirst.to_host(offset=r0, count=r1 - r0), fe), t
        assert np.array_equal(dcount.to_host(offset=r0, count=r1 - r0), ce), t
        hits.append(np.mean(want != SENT))
    assert 0.75 < float(np.mean(hits)) < 0.85
    b.close()
