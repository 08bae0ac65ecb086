"""GPU: learned position model. Gradients of the training kernels against
the oracle's FP64 backward; exact answers (== plain descent == oracle) for
every root mode and with forced fallback; what the trained network gains
over routing without it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SENT = 0xFFFFFFFFFFFFFFFF


def test_mlp_gradient_matches_oracle(bpt, oracle):
    rng = np.random.default_rng(3)
    H = 32
    npar = 3 * H + H * H + H + 1
    for trial in range(3):
        p = rng.normal(0, 0.4, size=npar).astype(np.float32)
        x = rng.uniform(0, 1, size=5000).astype(np.float32)
        t = np.sort(rng.uniform(0, 1, size=5000)).astype(np.float32)
        loss, g = bpt.mlp_loss_grad(p, x, t)
        rl, rg = oracle.mlp_loss_grad(p.astype(np.float64), H, x.astype(np.float64), t.astype(np.float64))
        # FP32 sums over 5000 samples against an FP64 reference
        assert abs(loss - rl) <= 1e-4 * abs(rl) + 1e-7
        assert np.max(np.abs(g - rg)) <= 2e-4 * np.max(np.abs(rg)) + 1e-6, trial


@pytest.fixture(scope="module")
def lognormal_tree(bpt, wl):
    keys = wl.lognormal_keys(1 << 21, seed=4)
    vals = np.arange(len(keys), dtype=np.uint64)
    rng = np.random.default_rng(5)
    q = np.concatenate([keys[rng.integers(0, len(keys), 300_000)],
                        rng.integers(0, 1 << 50, 60_000, dtype=np.uint64),
                        np.array([0, 1, keys[0], keys[-1], keys[-1] + 1, SENT], dtype=np.uint64)])
    t = bpt.Tree(keys, vals, 8, 8)
    return t, keys, vals, q


@pytest.mark.parametrize("root", [0, 1, 2, 3])
def test_learned_exact_every_root(bpt, oracle, lognormal_tree, root):
    t, keys, vals, q = lognormal_tree
    st = t.learned_build(leaf_models=16384, train_steps=300, root=root)
    want = oracle.point_closed(keys, vals, q, SENT)
    assert np.array_equal(t.point_search(q, bpt.LEARNED), want), st
    t.set_option("coop", 0)   # one lane per query, binary search in the window
    assert np.array_equal(t.point_search(q, bpt.LEARNED), want), st
    t.set_option("coop", 1)
    assert np.array_equal(t.point_search(q, bpt.PLAIN), want)


def test_learned_forced_fallback_still_exact(bpt, oracle, lognormal_tree):
    t, keys, vals, q = lognormal_tree
    st = t.learned_build(leaf_models=4096, max_window=4, train_steps=200)
    assert st["fallback_models"] > 0 and st["fallback_keys"] > 0
    want = oracle.point_closed(keys, vals, q, SENT)
    assert np.array_equal(t.point_search(q, bpt.LEARNED), want)
    t.set_option("coop", 0)
    assert np.array_equal(t.point_search(q, bpt.LEARNED), want)
    t.set_option("coop", 1)
    # windows wider than 32 leaves take several 32-leaf ranking steps
    st = t.learned_build(leaf_models=256, max_window=1 << 20, train_steps=200)
    assert st["max_window_seen"] > 8 * 64
    assert np.array_equal(t.point_search(q, bpt.LEARNED), want)


@pytest.mark.parametrize("block", [64, 128, 256])
def test_learned_fallback_queue_shapes(bpt, oracle, lognormal_tree, block):
    """Queries of fallback models wait in a per-wave LDS queue and are answered
    64 at a time, the rest at the wave's end. With max_window 4 almost every
    model falls back, so queues fill and spill over every batch; batch sizes
    around one and two waves leave partial queues; 1, 2 and 4 waves per block
    give each block 1, 2 or 4 queues. Every answer equals the oracle, and the
    learned mode needs no workspace."""
    t, keys, vals, q = lognormal_tree
    rng = np.random.default_rng(17)
    t.set_launch(block=block)
    try:
        for max_window in (4, 256):
            t.learned_build(leaf_models=4096, max_window=max_window, train_steps=100)
            assert t.point_search_ws_size(1 << 20, bpt.LEARNED) == 0
            for nq in (1, 63, 64, 65, 129, 1000, 100_003):
                qq = q[rng.integers(0, len(q), nq)]
                want = oracle.point_closed(keys, vals, qq, SENT)
                got = t.point_search(qq, bpt.LEARNED)
                assert np.array_equal(got, want), (block, max_window, nq)
    finally:
        t.set_launch(block=256)


def test_trained_network_gains_over_no_network(bpt, lognormal_tree):
    t, keys, vals, q = lognormal_tree
    trained = t.learned_build(leaf_models=16384, train_steps=600, root=0)
    assert trained["final_loss"] < 0.1 * trained["initial_loss"]
    linear_key = t.learned_build(leaf_models=16384, root=3)
    # without the network, lognormal keys pile into few leaf models
    assert trained["mean_window"] < 0.25 * linear_key["mean_window"]
    assert trained["fallback_keys"] < linear_key["fallback_keys"]
    assert trained["models_used"] > 4 * linear_key["models_used"]


def test_learned_u32_tree(bpt, oracle, wl):
    keys, vals = wl.point_keys(1 << 20, workload=9)
    q = wl.cpu_queries(keys, 1 << 18, workload=9)
    t = bpt.Tree(keys, vals)
    t.learned_build(leaf_models=8192, train_steps=200)
    assert np.array_equal(t.point_search(q, bpt.LEARNED), oracle.point_closed(keys, vals, q, 0xFFFFFFFF).astype(np.uint32))


def test_learned_requires_model(bpt):
    keys = np.arange(100, dtype=np.uint64)
    t = bpt.Tree(keys, keys, 8, 8)
    with pytest.raises(bpt.BptError):
        t.point_search(keys, bpt.LEARNED)


def test_workload_c5_full_size(bpt, oracle, wl):
    """256M unique lognormal uint64 keys, the learned model trained as in the
    benchmark, 2^27 queries (90 % hits): learned equals plain on every query,
    and both equal the oracle on every query."""
    n, nq = 1 << 28, 1 << 27
    keys = wl.lognormal_keys(n, sort=lambda a: bpt.sort_pairs(a)[0])
    vals = np.arange(n, dtype=np.uint64)
    print("c5 keys generated", flush=True)
    t = bpt.Tree(keys, vals, 8, 8)
    st = t.learned_build(leaf_models=65536, max_window=256, train_samples=1 << 20, train_steps=2000, root=0)
    assert st["models_used"] > 60000 and st["mean_window"] < 400
    rng = wl.rng_for(0, 5, 1)
    q = np.where(rng.random(nq) < 0.9, keys[rng.integers(0, n, nq)],
                 rng.integers(0, 1 << 50, nq, dtype=np.uint64)).astype(np.uint64)
    dq = bpt.DeviceArray.from_host(q)
    out = bpt.DeviceArray(nq, np.uint64)
    t.point_search_dev(dq, out, bpt.PLAIN)
    plain = out.to_host()
    t.point_search_dev(dq, out, bpt.LEARNED)
    learned = out.to_host()
    assert int((plain != learned).sum()) == 0
    assert np.array_equal(plain, oracle.point_closed(keys, vals, q, (1 << 64) - 1))
    print(f"c5 checked: all {nq} queries", flush=True)
    dq.free()
    out.free()
    t.close()


def test_learned_clustered_models_with_far_keys(bpt, oracle):
    """Models (routing linear in the key, 4096 of them) that each hold a
    cluster of 2000 keys and one key 2^51 further: windows of about 2000
    slots spanning many level-1 nodes, predictions flat over the cluster;
    then 16 models, whose windows span most of the tree. Every answer equals
    the oracle."""
    rng = np.random.default_rng(12)
    parts = []
    for j in range(0, 4096, 41):
        base = j << 52
        parts.append(base + np.arange(2000, dtype=np.uint64) * np.uint64(3))
        parts.append(np.array([base + (1 << 51)], dtype=np.uint64))
    keys = np.concatenate(parts).astype(np.uint64)
    vals = np.arange(len(keys), dtype=np.uint64)
    t = bpt.Tree(keys, vals, 8, 8)
    q = np.concatenate([keys, keys + np.uint64(1), keys[rng.integers(0, len(keys), 50_000)],
                        rng.integers(0, 1 << 64, 50_000, dtype=np.uint64)])
    want = oracle.point_closed(keys, vals, q, SENT)
    for models in (4096, 16):
        st = t.learned_build(leaf_models=models, max_window=1 << 20, train_steps=50, root=3)
        assert st["fallback_models"] == 0, st
        assert np.array_equal(t.point_search(q, bpt.LEARNED), want), models
    t.close()
