"""Seeded synthetic inputs for the five workloads (numpy PCG64; seed 0).

Each workload uses an independent stream derived from (seed, workload id,
stream id), so scaled-down test sizes and the full benchmark sizes come from
the same code. The reference ships no data (its run scripts point at files
that are not in the repository), so every input here is synthetic.

  point_keys      unique uniform keys, sorted, value = rank
  cpu_queries     1M-key CPU workload: even i hit, odd i uniform (mostly miss)
  uniform_hits    queries uniform over the key set
  zipf_queries    Zipf(s) over popularity ranks mapped to key ranks through a
                  seeded permutation, plus a fraction of uniform misses
  points_2d       half uniform (x, y), half from a 16-centre Gaussian mixture
  rects           centres uniform, sides log-uniform
"""
from __future__ import annotations

import numpy as np

U32_SPACE = 1 << 32


def rng_for(seed: int, workload: int, stream: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, workload, stream])))


def unique_uniform(rng: np.random.Generator, n: int, bits: int) -> np.ndarray:
    """n distinct integers uniform over [0, 2^bits), sorted ascending."""
    dt = np.uint32 if bits <= 32 else np.uint64
    space = 1 << bits
    if n > space:
        raise ValueError("more keys than the key space")
    have = np.empty(0, dtype=dt)
    while len(have) < n:
        extra = max(1024, (n - len(have)) + (n - len(have)) // 8)
        draw = rng.integers(0, space, size=extra, dtype=np.uint64 if bits > 32 else np.uint64)
        have = np.unique(np.concatenate([have, draw.astype(dt)]))
    if len(have) > n:
        drop = rng.choice(len(have), size=len(have) - n, replace=False)
        keep = np.ones(len(have), dtype=bool)
        keep[drop] = False
        have = have[keep]
    return have


def point_keys(n: int, seed: int = 0, workload: int = 2, bits: int = 32):
    """Sorted unique keys and value = rank (uint32 for 32-bit keys)."""
    rng = rng_for(seed, workload, 0)
    keys = unique_uniform(rng, n, bits)
    vdt = np.uint32 if bits <= 32 else np.uint64
    vals = np.arange(n, dtype=vdt)
    return keys, vals


def cpu_queries(keys: np.ndarray, nq: int, seed: int = 0, workload: int = 1) -> np.ndarray:
    """Even i: a key drawn uniformly from the key set; odd i: uniform uint32."""
    rng = rng_for(seed, workload, 1)
    q = np.empty(nq, dtype=keys.dtype)
    nh = (nq + 1) // 2
    q[0::2] = keys[rng.integers(0, len(keys), size=nh)]
    q[1::2] = rng.integers(0, U32_SPACE, size=nq // 2, dtype=np.uint64).astype(keys.dtype)
    return q


def uniform_hits(keys: np.ndarray, nq: int, seed: int = 0, workload: int = 2, stream: int = 1):
    rng = rng_for(seed, workload, stream)
    return keys[rng.integers(0, len(keys), size=nq)]


def zipf_ranks(rng: np.random.Generator, n: int, size: int, s: float) -> np.ndarray:
    """Ranks in [0, n) with P(rank r) ~ (r + 1)^-s, by inverting the
    continuous approximation of the generalized harmonic number
    H(x) = (x^(1-s) - 1) / (1 - s) (s != 1)."""
    a = 1.0 - s
    hmax = ((n + 1.0) ** a - 1.0) / a
    u = rng.random(size) * hmax
    x = np.power(1.0 + u * a, 1.0 / a)
    r = np.floor(x).astype(np.int64) - 1
    np.clip(r, 0, n - 1, out=r)
    return r


def zipf_queries(keys: np.ndarray, nq: int, s: float = 0.99, miss_frac: float = 0.10,
                 seed: int = 0, workload: int = 2, stream: int = 2) -> np.ndarray:
    rng = rng_for(seed, workload, stream)
    n = len(keys)
    perm = rng.permutation(n)
    q = keys[perm[zipf_ranks(rng, n, nq, s)]]
    miss = rng.random(nq) < miss_frac
    nm = int(miss.sum())
    q[miss] = rng.integers(0, U32_SPACE if keys.dtype == np.uint32 else 1 << 63, size=nm,
                           dtype=np.uint64).astype(keys.dtype)
    return q


def morton_np(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Morton code, x in even bits (magic-mask interleave in numpy)."""
    def spread(v):
        v = v.astype(np.uint64)
        v = (v | (v << np.uint64(16))) & np.uint64(0x0000FFFF0000FFFF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF00FF00FF)
        v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
        v = (v | (v << np.uint64(2))) & np.uint64(0x3333333333333333)
        v = (v | (v << np.uint64(1))) & np.uint64(0x5555555555555555)
        return v
    return spread(x) | (spread(y) << np.uint64(1))


def points_2d(n: int, seed: int = 0, workload: int = 3, centres: int = 16, sigma: float = 2.0 ** 24):
    """Returns x, y (uint32), ids (uint32 generation index), keys (Morton,
    sorted) and the ids in key order."""
    rng = rng_for(seed, workload, 0)
    nu = n // 2
    ng = n - nu
    x = np.empty(n, dtype=np.uint32)
    y = np.empty(n, dtype=np.uint32)
    x[:nu] = rng.integers(0, U32_SPACE, size=nu, dtype=np.uint64).astype(np.uint32)
    y[:nu] = rng.integers(0, U32_SPACE, size=nu, dtype=np.uint64).astype(np.uint32)
    c = rng.integers(0, U32_SPACE, size=(centres, 2), dtype=np.uint64).astype(np.float64)
    comp = rng.integers(0, centres, size=ng)
    gx = np.empty(ng)
    gy = np.empty(ng)
    todo = np.arange(ng)
    while len(todo):  # rejection of samples outside [0, 2^32)
        sx = c[comp[todo], 0] + rng.normal(0.0, sigma, size=len(todo))
        sy = c[comp[todo], 1] + rng.normal(0.0, sigma, size=len(todo))
        ok = (sx >= 0) & (sx < U32_SPACE) & (sy >= 0) & (sy < U32_SPACE)
        gx[todo[ok]] = sx[ok]
        gy[todo[ok]] = sy[ok]
        todo = todo[~ok]
    x[nu:] = gx.astype(np.uint64).astype(np.uint32)
    y[nu:] = gy.astype(np.uint64).astype(np.uint32)
    ids = np.arange(n, dtype=np.uint32)
    z = morton_np(x, y)
    order = np.argsort(z, kind="stable")
    return x, y, ids, z[order], ids[order]


def rects(nr: int, seed: int = 0, workload: int = 3, lo_exp: float = 8, hi_exp: float = 20):
    """(nr, 4) uint32 [x0, y0, x1, y1], inclusive; sides log-uniform in
    [2^lo_exp, 2^hi_exp], centres uniform, clamped to the coordinate range."""
    rng = rng_for(seed, workload, 1)
    cx = rng.integers(0, U32_SPACE, size=nr, dtype=np.uint64).astype(np.int64)
    cy = rng.integers(0, U32_SPACE, size=nr, dtype=np.uint64).astype(np.int64)
    w = np.floor(np.exp2(rng.uniform(lo_exp, hi_exp, size=nr))).astype(np.int64)
    h = np.floor(np.exp2(rng.uniform(lo_exp, hi_exp, size=nr))).astype(np.int64)
    x0 = np.clip(cx - w // 2, 0, U32_SPACE - 1)
    y0 = np.clip(cy - h // 2, 0, U32_SPACE - 1)
    x1 = np.clip(x0 + w - 1, 0, U32_SPACE - 1)
    y1 = np.clip(y0 + h - 1, 0, U32_SPACE - 1)
    return np.stack([x0, y0, x1, y1], axis=1).astype(np.uint32)


# ---------------------------------------------------------------- forest workload
# numpy restatement of the device generator (bpt_synth_forest): splitmix64
# (Steele, Lea, Flood 2014) in a counter-based chain h(seed, stream, tree, i).
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def hmix(seed, stream, tree, i):
    a = splitmix64(np.uint64(seed) ^ splitmix64(np.uint64(stream)))
    return splitmix64(splitmix64(a ^ np.uint64(tree)) ^ np.asarray(i, dtype=np.uint64))


def synth_tree_keys(seed, tree, n):
    """Keys of one tree: n draws uniform over [0, 2^64), sorted; a tree whose
    draws collide is redrawn from stream 1 + 16 * attempt."""
    i = np.arange(n, dtype=np.uint64)
    for attempt in range(9):
        k = np.sort(hmix(seed, 1 + 16 * attempt, tree, i))
        if n < 2 or (k[1:] != k[:-1]).all():
            return k
    raise RuntimeError("keys collide after 8 redraws")


def forest_sizes(ntrees, lo_exp=10, hi_exp=20, seed=0, workload=4):
    """Tree sizes floor(2^U), U uniform in [lo_exp, hi_exp] (log-uniform)."""
    rng = rng_for(seed, workload, 0)
    return np.floor(np.exp2(rng.uniform(lo_exp, hi_exp, size=ntrees))).astype(np.uint64)


def forest_offsets(sizes, per_key=1, ranges_div=16):
    key_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    rc = np.maximum(np.asarray(sizes, dtype=np.uint64) // np.uint64(ranges_div), np.uint64(1))
    r_off = np.concatenate([[0], np.cumsum(rc)]).astype(np.uint64)
    return key_off, r_off


# ---------------------------------------------------------------- learned-index workload
def lognormal_keys(n, mu=0.0, sigma=2.0, bits=50, seed=0, workload=5, sort=None):
    """n unique uint64 keys floor(s / s_max * 2^bits), s ~ LogNormal(mu, sigma),
    sorted; duplicates are dropped and replaced by fresh draws (top-up).
    `sort` may be a callable sorting a uint64 array (e.g. on the device)."""
    rng = rng_for(seed, workload, 0)
    s = np.exp(rng.normal(mu, sigma, size=n))
    smax = s.max()
    scale = float(1 << bits) / smax
    keys = np.floor(s * scale).astype(np.uint64)
    del s
    while True:
        keys = sort(keys) if sort is not None else np.sort(keys)
        keep = np.ones(len(keys), dtype=bool)
        keep[1:] = keys[1:] != keys[:-1]
        keys = keys[keep]
        if len(keys) >= n:
            break
        extra = np.exp(rng.normal(mu, sigma, size=n - len(keys) + 16))
        extra = np.floor(np.minimum(extra, smax) * scale).astype(np.uint64)
        keys = np.concatenate([keys, extra])
    if len(keys) > n:
        drop = rng.choice(len(keys), size=len(keys) - n, replace=False)
        keep = np.ones(len(keys), dtype=bool)
        keep[drop] = False
        keys = keys[keep]
    return keys
