"""GPU: bpt_point_search_noalloc (the point search that allocates nothing and
uses the library's static per-GPU scratch) through the C ABI, against
bpt_point_search on the same tree and queries, for every mode, on two calls
in a row, the second with a smaller batch; and the capacity error for a
batch whose workspace exceeds the scratch (rejected on the host)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20  # BPT_NOALLOC_SCRATCH_BYTES


def _run_both(bpt, t, q, mode, kdt, vdt):
    dq = bpt.DeviceArray.from_host(q.astype(kdt))
    out_ref = bpt.DeviceArray(len(q), vdt)
    out_na = bpt.DeviceArray(len(q), vdt)
    t.point_search_dev(dq, out_ref, mode)
    t.point_search_noalloc_dev(dq, out_na, mode)
    ref, got = out_ref.to_host(), out_na.to_host()
    for a in (dq, out_ref, out_na):
        a.free()
    return ref, got


@pytest.mark.parametrize("kb", [4, 8])
def test_noalloc_matches_point_search(bpt, oracle, kb):
    rng = np.random.default_rng(40 + kb)
    kdt = np.uint32 if kb == 4 else np.uint64
    n = 1 << 20
    keys = np.unique(rng.integers(0, (1 << (8 * kb)) - 1, n, dtype=np.uint64)).astype(kdt)
    vals = np.arange(len(keys), dtype=kdt)
    t = bpt.Tree(keys, vals, kb, kb)
    t.set_option("group_min", 0)  # the grouped kernels at every batch size
    modes = [bpt.PLAIN, bpt.GROUPED]
    if kb == 8:
        t.learned_build(leaf_models=4096, train_steps=100, train_samples=1 << 16)
        modes.append(bpt.LEARNED)
    sent = (1 << (8 * kb)) - 1
    for nq in (1 << 21, 100_003):          # two calls in a row, the second smaller
        q = np.concatenate([keys[rng.integers(0, len(keys), nq - nq // 4)],
                            rng.integers(0, sent, nq // 4, dtype=np.uint64).astype(kdt)])
        want = oracle.point_closed(keys, vals, q, sent).astype(kdt)
        for mode in modes:
            print(f"[noalloc] kb={kb} nq={nq} mode={mode}", flush=True)
            ref, got = _run_both(bpt, t, q, mode, kdt, kdt)
            assert np.array_equal(got, ref), (nq, mode)
            assert np.array_equal(got, want), (nq, mode)
    # a batch whose grouped workspace exceeds the scratch is refused before any launch
    big = 1 << 26 if kb == 4 else 1 << 25
    assert t.point_search_ws_size(big, bpt.GROUPED) > SCRATCH
    dq = bpt.DeviceArray(big, kdt)
    out = bpt.DeviceArray(big, kdt)
    with pytest.raises(bpt.BptError) as e:
        t.point_search_noalloc_dev(dq, out, bpt.GROUPED)
    assert e.value.status == 5  # BPT_ERR_CAPACITY
    dq.free()
    out.free()
    t.close()
