"""CPU: the seeded input generators (examples/bptree_gen.cpp) that stand in
for the reference's randomgenator/ tools. The reference seeds from the clock,
so its exact files cannot be reproduced (parity unpinned for the sequences);
these tests pin the distributions it draws from (randgen.cu:58-246,
main.cpp:29-56) by their closed-form statistics, and determinism per seed."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "b-tree-index-searching_amd")
SRC = os.path.join(PKG, "examples", "bptree_gen.cpp")


@pytest.fixture(scope="module")
def gen(tmp_path_factory):
    exe = os.path.join(PKG, "bptree_gen")
    if os.path.exists(exe) and os.path.getmtime(exe) >= os.path.getmtime(SRC):
        return exe
    if shutil.which("g++") is None:
        pytest.skip("g++ not available to build the generator")
    exe = str(tmp_path_factory.mktemp("gen") / "bptree_gen")
    subprocess.run(["g++", "-O2", "-std=c++17", SRC, "-o", exe], check=True)
    return exe


def _run(gen, *args, cwd=None):
    r = subprocess.run([gen, *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _u64(path):
    return np.fromfile(path, dtype=np.uint64)


def test_keys_distinct_uniform_and_seeded(gen, tmp_path):
    n, top = 200_000, 10_000_000
    _run(gen, "keys", n, tmp_path / "a", 1, top)
    _run(gen, "keys", n, tmp_path / "b", 1, top)
    _run(gen, "keys", n, tmp_path / "c", 2, top)
    a = _u64(tmp_path / "a")
    assert len(a) == n and len(np.unique(a)) == n
    assert a.min() >= 1 and a.max() <= top
    assert not np.all(a[1:] > a[:-1])               # written in random order, like the reservoir
    assert np.array_equal(a, _u64(tmp_path / "b"))  # same seed, same file
    assert not np.array_equal(a, _u64(tmp_path / "c"))
    # a uniform n-subset of [1, top]: mean top/2, and each tenth holds n/10
    assert abs(a.astype(np.float64).mean() / (top / 2) - 1) < 0.01
    h = np.histogram(a, bins=10, range=(1, top + 1))[0]
    assert np.all(np.abs(h - n / 10) < 5 * np.sqrt(n / 10))
    # every key of a full range when n == top
    _run(gen, "keys", 1000, tmp_path / "d", 3, 1000)
    assert np.array_equal(np.sort(_u64(tmp_path / "d")), np.arange(1, 1001, dtype=np.uint64))


@pytest.fixture(scope="module")
def keyfile(gen, tmp_path_factory):
    d = tmp_path_factory.mktemp("keys")
    _run(gen, "keys", 100_000, d / "k", 5, 10**9)
    return d / "k"


def test_queries_uniform_and_normal(gen, keyfile, tmp_path):
    M = int(_u64(keyfile).max())
    nq = 400_000
    _run(gen, "queries", keyfile, nq, 0, tmp_path / "u", 9)
    u = _u64(tmp_path / "u").astype(np.float64)
    assert u.min() >= 0 and u.max() <= M
    assert abs(u.mean() / (M / 2) - 1) < 0.01
    assert abs(u.std() / (M / np.sqrt(12)) - 1) < 0.01
    for dist, rel in ((1, 0.35), (4, 0.63)):
        _run(gen, "queries", keyfile, nq, dist, tmp_path / f"n{dist}", 9)
        x = _u64(tmp_path / f"n{dist}")
        mean = M // 2
        sd = rel * mean
        assert x.max() <= M
        # draws outside [0, M] are replaced by the mean: P(|z| > 1 / rel)
        from math import erfc, sqrt
        p_out = erfc(1 / rel / sqrt(2))
        frac = float((x == mean).mean())
        assert abs(frac - p_out) < 5 * np.sqrt(p_out * (1 - p_out) / nq) + 1e-4, (dist, frac, p_out)
        inside = x[x != mean].astype(np.float64)
        assert abs(inside.mean() / mean - 1) < 0.01
        if dist == 1:  # nearly untruncated: the sample sd is the normal's
            assert abs(inside.std() / sd - 1) < 0.02


def test_queries_zipf_frequencies(gen, keyfile, tmp_path):
    M = int(_u64(keyfile).max())
    nq = 1_000_000
    _run(gen, "queries", keyfile, nq, 2, tmp_path / "z", 4)
    z = _u64(tmp_path / "z")
    assert z.min() >= 1 and z.max() <= M
    H = np.log(M) + 0.5772156649015329 + 1 / (2 * M)  # harmonic number H_M
    for k in (1, 2, 3, 5, 10, 100):
        p = 1 / (k * H)
        got = float((z == k).mean())
        assert abs(got - p) < 5 * np.sqrt(p * (1 - p) / nq), (k, got, p)
    # P(k <= m) = H_m / H_M for a larger m
    m = 10**6
    Hm = np.log(m) + 0.5772156649015329 + 1 / (2 * m)
    p = Hm / H
    assert abs(float((z <= m).mean()) - p) < 5 * np.sqrt(p * (1 - p) / nq)


def test_queries_five_gaussians(gen, keyfile, tmp_path):
    M = int(_u64(keyfile).max())
    nq = 100_003
    _run(gen, "queries", keyfile, nq, 3, tmp_path / "g", 2)
    g = _u64(tmp_path / "g")
    per = nq // 5
    for i in range(5):
        part = g[i * per:(i + 1) * per] if i < 4 else g[4 * per:]
        # sd 0.1 around an integer centre c, truncated toward zero: c - 1 or c
        vals = np.unique(part)
        assert len(vals) <= 2 and (len(vals) == 1 or vals[1] - vals[0] == 1), vals
        assert vals.max() < M


def test_default_names_and_ranges(gen, keyfile, tmp_path):
    _run(gen, "queries", keyfile, 3_000_000, 1, cwd=tmp_path)
    assert os.path.exists(tmp_path / "exp-normal-0-Q3")
    _run(gen, "ranges", keyfile, 50_000, 0, tmp_path / "s", tmp_path / "e", 6, 16)
    s, e = _u64(tmp_path / "s"), _u64(tmp_path / "e")
    w = e - s
    assert np.all(e >= s) and w.min() >= 1 and w.max() <= 1 << 16
    # log-uniform widths: log2 w is uniform on [0, 16)
    assert abs(np.log2(w.astype(np.float64)).mean() - 8) < 0.1
