#!/usr/bin/env python3
"""Benchmark of the B+tree point-search hot path on one MI355X.

Workload: 64M (2^26) unique uint32 keys uniform over [0, 2^32), value = rank,
bulk-built into the flat gfx950 tree (fanout 32, 16-slot leaves); one step =
one plain point-search pass over 256M (2^28) queries drawn uniformly from the
key set, with queries, tree and results resident in HBM. Also measured
beside it (not the headline value), on the same tree: the radix-grouped and
learned forms on the uniform and the Zipf(0.99) (10% misses) query sets, 1-D
ranges in the three query orders, the serving loop's p50/p99 latency at 0.5x
and 0.8x of its saturation, and batched upserts; then, at their full sizes,
the 2-D rectangle (32M points, 4M rectangles), forest (4096 trees) and
learned-index + serving (256M lognormal keys) configurations, summarised
under "configs" (each also runs alone with --config c3 / c4 / c5).

Prints ONE JSON line. `roofline.achieved` uses the algorithmic bytes of a
lookup (4 B query + 4 B result + one 128 B leaf line = 136 B) over the plain
kernel's average launch time from HIP events on its stream. `cpu_baseline`
is the oracle's restatement of the reference's page tree search
(PAGESIZE 5000, OpenMP over queries) timed on this host on a bounded sample;
its results are also compared with the GPU's for the same queries.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "b-tree-index-searching_amd")
sys.path.insert(0, PKG)

import bptree  # noqa: E402
import workloads  # noqa: E402

METRIC = "point lookups/s @64M keys plain vs grouped vs learned; HBM GB/s; p99 ms"
BYTES_PER_LOOKUP = 4 + 4 + 128
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_form(tree, dq, dout, mode, steps, warmup):
    """Returns (wall seconds per step, list of per-launch event ms)."""
    for _ in range(warmup):
        tree.point_search_dev(dq, dout, mode)
    bptree.sync()
    timers = [bptree.Timer() for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        timers[k].start()
        tree.point_search_dev(dq, dout, mode)
        timers[k].stop()
    bptree.sync()
    wall = (time.perf_counter() - t0) / steps
    ev = [t.ms() for t in timers]
    for t in timers:
        t.close()
    return wall, ev


def batch_sweep(tree, q):
    """Wall ms per call of the plain descent, the grouped kernels (group_min
    0) and the grouped mode with the tree's default group_min, for the first
    n queries, n = 2^16 .. 2^26; every form's results are compared with
    plain's."""
    rows = []
    for lg in (16, 18, 20, 21, 22, 23, 24, 25, 26):
        n = 1 << lg
        dq = bptree.DeviceArray.from_host(q[:n])
        dout = bptree.DeviceArray(n, np.uint32)
        steps = 50 if n <= (1 << 20) else 10
        row = {"queries": n}
        ref = None
        for name, mode, gmin in (("plain", bptree.PLAIN, None), ("grouped_kernels", bptree.GROUPED, 0),
                                 ("grouped_mode", bptree.GROUPED, 1 << 24)):
            if gmin is not None:
                tree.set_option("group_min", gmin)
            wall, _ = time_form(tree, dq, dout, mode, steps, 3)
            got = dout.to_host()
            if ref is None:
                ref = got
            row[f"{name}_ms"] = wall * 1e3
            row[f"{name}_mismatches"] = int((got != ref).sum())
        rows.append(row)
        log(f"batch sweep: {row}")
        dq.free()
        dout.free()
    return rows


def host_info(threads):
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"threads_used": threads, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "cpu_model": cpu_model}


def cpu_baseline(keys, vals, q, gpu_out, target_s=10.0, zipf=None):
    """Oracle page tree (reference semantics, PAGESIZE 5000) on host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker; only this leg of the benchmark uses it

    threads = min(16, len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    pt = oracle.PageTree(keys.astype(np.uint64), vals.astype(np.uint64), 5000)
    build_s = time.perf_counter() - t0
    probe = q[: 1 << 20].astype(np.uint64)
    pt.point(probe, 0xFFFFFFFF, threads)  # warm-up
    t0 = time.perf_counter()
    pt.point(probe, 0xFFFFFFFF, threads)
    per_q = (time.perf_counter() - t0) / len(probe)
    s = int(min(len(q), max(1 << 20, target_s / max(per_q, 1e-12))))
    s = 1 << int(np.log2(s))
    sample = q[:s].astype(np.uint64)
    t0 = time.perf_counter()
    res = pt.point(sample, 0xFFFFFFFF, threads)
    dt = time.perf_counter() - t0
    mism = int((res.astype(np.uint32) != gpu_out[:s]).sum())
    one = q[: 1 << 20].astype(np.uint64)  # the same search on one core, a smaller sample
    t0 = time.perf_counter()
    res1 = pt.point(one, 0xFFFFFFFF, 1)
    dt1 = time.perf_counter() - t0
    mism += int((res1.astype(np.uint32) != gpu_out[: len(one)]).sum())
    zipf_rate = None
    if zipf is not None:  # the second query set: a sample of the Zipf queries and the GPU's results
        zq, zgpu = zipf
        t0 = time.perf_counter()
        resz = pt.point(zq.astype(np.uint64), 0xFFFFFFFF, threads)
        dtz = time.perf_counter() - t0
        mism += int((resz.astype(np.uint32) != zgpu).sum())
        zipf_rate = {"value": len(zq) / dtz, "sample": f"first {len(zq)} Zipf(0.99) queries, {dtz:.1f} s"}
    pt.close()
    return {
        "value": s / dt, "unit": "lookups/s", "cores": threads, "kind": "port",
        "host": host_info(threads),
        "sample": f"first {s} of the {len(q)} uniform-hit queries against the full 64M-key page "
                  f"tree (PAGESIZE 5000, reference default), OpenMP static schedule, "
                  f"{dt:.1f} s; page-tree build {build_s:.1f} s excluded",
        "one_core": {"value": len(one) / dt1, "sample": f"first {len(one)} queries, {dt1:.1f} s"},
        "zipf": zipf_rate,
        "mismatches_vs_gpu": mism,
    }


def common_line(metric, value, unit, args, ms, dtype, config, roofline, extra):
    line = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded)", "config": config,
            "roofline": roofline}
    line.update(extra)
    return line


def bench_c3(args):
    """2-D range search over 32M Morton-keyed points, 4M rectangles, three
    query orders (basic / 2-D grouping / Z-order); count + sorted ids."""
    npts, nr = 1 << 25, 1 << 22
    t0 = time.perf_counter()
    x, y, ids, keys, kid = workloads.points_2d(npts)
    rects = workloads.rects(nr)
    log(f"generated {npts} points, {nr} rectangles in {time.perf_counter() - t0:.1f} s")
    tree = bptree.Tree(keys, kid, 8, 4)
    dr = bptree.DeviceArray.from_host(rects.reshape(-1))
    dcount = bptree.DeviceArray(nr, np.uint64)
    doff = bptree.DeviceArray(nr, np.uint64)
    L = bptree.lib()
    forms, results = {}, {}
    # the caller-owned id buffer, kept across calls and grown when a step needs more
    ids_buf = [bptree.DeviceArray(1 << 20, np.uint32)]
    for mode, name in ((0, "basic"), (1, "grid2d"), (2, "zorder")):
        def step():
            bptree.check(L.bpt_rect_count(tree.h, dr.ptr, nr, mode, dcount.ptr, None))
            tot = bptree.C.c_uint64(0)
            bptree.check(L.bpt_exclusive_scan_u64(dcount.ptr, nr, doff.ptr, bptree.C.byref(tot), None))
            if tot.value > len(ids_buf[0]):
                ids_buf[0].free()
                ids_buf[0] = bptree.DeviceArray(tot.value, np.uint32)
            bptree.check(L.bpt_rect_ids(tree.h, dr.ptr, nr, mode, doff.ptr, ids_buf[0].ptr, None))
            return tot.value
        for _ in range(args.warmup):
            step()
        bptree.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
            bptree.sync()
        dt = (time.perf_counter() - t0) / args.steps
        tot = step()
        results[name] = (dcount.to_host(), ids_buf[0].to_host(count=tot))
        forms[name] = {"rects_per_s": nr / dt, "ms": dt * 1e3}
        log(f"c3 {name}: {dt * 1e3:.2f} ms, {nr / dt / 1e6:.1f} M rects/s")
    base = results["basic"]
    for name in ("grid2d", "zorder"):
        forms[name]["identical_to_basic"] = bool(np.array_equal(base[0], results[name][0]) and
                                                 np.array_equal(base[1], results[name][1]))
    cnt = base[0]
    extra = {"forms": forms, "hits_total": int(cnt.sum()), "hits_mean": float(cnt.mean())}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = min(16, len(os.sched_getaffinity(0)))
        sample = 1 << 16
        t0 = time.perf_counter()
        c, o, sids = oracle.rect_morton(keys, kid, rects[:sample], threads=threads)
        dt = time.perf_counter() - t0
        offs = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
        mism = int((c != cnt[:sample]).sum()) + int(not np.array_equal(sids, base[1][:offs[sample]]))
        extra["cpu_baseline"] = {"value": sample / dt, "unit": "rects/s", "cores": threads, "kind": "port",
                                 "sample": f"first {sample} rectangles, quadtree Morton decomposition + filter "
                                           f"+ sort (oracle), {dt:.1f} s", "mismatches_vs_gpu": mism}
    best = forms["basic"]
    line = common_line("2-D rectangle queries/s @32M points (count + sorted ids)", nr * 1e3 / best["ms"],
                       "rects/s", args, best["ms"], "u64",
                       {"workload": "2-D range search: 32M points (half uniform, half 16-centre Gaussian "
                                    "mixture, sigma 2^24) keyed by 64-bit Morton code; 4M rectangles, centres "
                                    "uniform, sides log-uniform in [2^8, 2^20]", "points": npts, "rects": nr},
                       {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                        "traffic": None, "note": "reported, not gated: bytes depend on the decomposition"},
                       extra)
    tree.close()
    for a in (dr, dcount, doff, ids_buf[0]):
        a.free()
    return line


def bench_c1(args):
    """The reference's CPU-runnable case: 1M unique uint32 keys, value = rank,
    1M queries (half from the key set, half uniform uint32). The page tree of
    the reference's search restated on the CPU (the oracle) at order 32 and at
    the reference default PAGESIZE 5000, on all host cores and on one (median
    of 5 after a warm-up, tree build excluded); the GPU plain search on the same
    inputs is the parity check and is reported beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the checker; this configuration measures it as the CPU baseline

    n = nq = 1 << 20
    keys, vals = workloads.point_keys(n, workload=1)
    q = workloads.cpu_queries(keys, nq, workload=1)
    q64 = q.astype(np.uint64)
    threads = min(16, len(os.sched_getaffinity(0)))
    tree = bptree.Tree(keys, vals)
    dq = bptree.DeviceArray.from_host(q)
    dout = bptree.DeviceArray(nq, np.uint32)
    wall_g, ev = time_form(tree, dq, dout, bptree.PLAIN, max(args.steps, 10), args.warmup)
    gpu = dout.to_host()
    forms, mism = {}, 0
    for ps in (32, 5000):
        pt = oracle.PageTree(keys.astype(np.uint64), vals.astype(np.uint64), ps)
        for cores in (threads, 1):
            pt.point(q64, 0xFFFFFFFF, cores)  # warm-up
            ts = []
            for _ in range(5):
                t0 = time.perf_counter()
                res = pt.point(q64, 0xFFFFFFFF, cores)
                ts.append(time.perf_counter() - t0)
            med = float(np.median(ts))
            forms[f"cpu_pagesize{ps}_{cores}core"] = {"lookups_per_s": nq / med, "ms": med * 1e3, "cores": cores}
            mism += int((res.astype(np.uint32) != gpu).sum())
        pt.close()
    forms["gpu_plain"] = {"lookups_per_s": nq / (float(np.median(ev)) * 1e-3), "kernel_ms": float(np.median(ev))}
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    head = forms[f"cpu_pagesize32_{threads}core"]
    line = common_line("CPU point lookups/s @1M uint32 keys, order-32 page tree (reference search semantics)",
                       head["lookups_per_s"], "lookups/s", args, head["ms"], "u32",
                       {"workload": "plain CPU point search: 1M unique uint32 keys uniform in [0,2^32), value = "
                                    "rank, B+tree order 32; 1M queries, half from the key set, half uniform uint32",
                        "keys": n, "queries": nq},
                       {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                        "traffic": None, "note": "CPU configuration; the GPU plain search of the same inputs is "
                                                 "in forms.gpu_plain"},
                       {"forms": forms, "mismatches_cpu_vs_gpu": mism, "hit_fraction": float((gpu != 0xFFFFFFFF).mean()),
                        "host": {"cores_used": threads, "nproc": os.cpu_count(), "cpu_model": cpu_model},
                        "cpu_baseline": {"value": head["lookups_per_s"], "unit": "lookups/s", "cores": threads,
                                         "kind": "port", "sample": f"all {nq} queries, median of 5"}})
    dq.free()
    dout.free()
    tree.close()
    return line


def bench_c4(args):
    """4096 independent uint64 trees (sizes log-uniform in [1K, 1M]), 4x
    size point queries per tree (80% hits) + size/16 ranges per tree, one
    persistent hipGraph-captured launch, autotuned configuration."""
    ntrees = 4096
    sizes = workloads.forest_sizes(ntrees, 10, 20)
    key_off, r_off = workloads.forest_offsets(sizes)
    q_off = key_off * np.uint64(4)
    t0 = time.perf_counter()
    dk, dv, dq, dlo, dhi = bptree.synth_forest(0, key_off, qpk=4, hit_pct=80, r_off=r_off)
    nq, nr = int(q_off[-1]), int(r_off[-1])
    log(f"c4: {ntrees} trees, {int(key_off[-1])} keys, {nq} point + {nr} range queries generated in "
        f"{time.perf_counter() - t0:.1f} s")
    dout = bptree.DeviceArray(nq, np.uint64)
    dfirst = bptree.DeviceArray(nr, np.uint64)
    dcount = bptree.DeviceArray(nr, np.uint64)
    t0 = time.perf_counter()
    if args.no_autotune:  # the configuration the autotuner picks on MI355X
        best = {"fanout": 16, "leaf_cap": 8, "lds_bytes": 32768, "chunk": 8192, "blocks_per_cu": 8, "lanes": 8}
        report = []
    else:
        best, report = bptree.forest_autotune(dk, dv, key_off, dq, q_off, dout, dlo, dhi, r_off, dfirst, dcount,
                                              reps=1)
    log(f"autotune {time.perf_counter() - t0:.1f} s -> {best}")
    forest = bptree.Forest(dk, dv, key_off, fanout=best["fanout"], lds_bytes=best["lds_bytes"], lanes=best["lanes"])
    batch = bptree.Batch(forest, dq, q_off, dout, dlo, dhi, r_off, dfirst, dcount, chunk=best["chunk"])
    for _ in range(args.warmup):
        batch.run(graph=True)
    bptree.sync()
    tm = [bptree.Timer() for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        tm[k].start()
        batch.run(graph=True)
        tm[k].stop()
    bptree.sync()
    wall = (time.perf_counter() - t0) / args.steps
    kms = float(np.mean([t.ms() for t in tm]))
    # points only, for the roofline of the point lookups
    batch_p = bptree.Batch(forest, dq, q_off, dout, chunk=best["chunk"])
    batch_p.run(graph=True)
    bptree.sync()
    tp = bptree.Timer()
    tp.start()
    batch_p.run(graph=True)
    tp.stop()
    pms = tp.ms()
    # baseline: one plain launch per tree (same trees, separate handles)
    trees = []
    for t in range(ntrees):
        a, b = int(key_off[t]), int(key_off[t + 1])
        d = bptree.BuildDesc(8, 8, 16, 8, 0, 0)
        h = bptree.vp()
        bptree.check(bptree.lib().bpt_tree_build(bptree.C.byref(d), dk.ptr + 8 * a, dv.ptr + 8 * a, b - a,
                                                 bptree.C.byref(h), None))
        trees.append(h.value)
    L = bptree.lib()
    bptree.sync()
    t0 = time.perf_counter()
    for t in range(ntrees):
        qa, qb = int(q_off[t]), int(q_off[t + 1])
        L.bpt_point_search(trees[t], dq.ptr + 8 * qa, qb - qa, dout.ptr + 8 * qa, 0, None)
        ra, rb = int(r_off[t]), int(r_off[t + 1])
        L.bpt_range_count(trees[t], dlo.ptr + 8 * ra, dhi.ptr + 8 * ra, rb - ra, dfirst.ptr + 8 * ra,
                          dcount.ptr + 8 * ra, None)
    bptree.sync()
    per_tree_s = time.perf_counter() - t0
    for h in trees:
        L.bpt_tree_free(h)
    batch.run(graph=True)
    bptree.sync()
    extra = {"autotune": {"best": best, "configs_timed": len(report)},
             "per_tree_launches_s": per_tree_s, "speedup_vs_per_tree_launches": per_tree_s / wall,
             "points": nq, "ranges": nr, "batch_ms": kms, "points_only_ms": pms}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        keys, vals = dk.to_host(), dv.to_host()
        out, cnt = dout.to_host(), dcount.to_host()
        q, lo, hi = dq.to_host(), dlo.to_host(), dhi.to_host()
        sample = list(range(0, ntrees, 64)) + [int(np.argmax(sizes))]
        mism, nsq = 0, 0
        t0 = time.perf_counter()
        for t in sample:
            k = keys[key_off[t]:key_off[t + 1]]
            v = vals[key_off[t]:key_off[t + 1]]
            pq = q[q_off[t]:q_off[t + 1]]
            pt = oracle.PageTree(k, v, 5000)
            want = pt.point(pq, 0xFFFFFFFFFFFFFFFF, min(16, len(os.sched_getaffinity(0))))
            pt.close()
            _, wc = oracle.range_closed(k, lo[r_off[t]:r_off[t + 1]], hi[r_off[t]:r_off[t + 1]])
            mism += int((want != out[q_off[t]:q_off[t + 1]]).sum()) + int((wc != cnt[r_off[t]:r_off[t + 1]]).sum())
            nsq += len(pq)
        dt = time.perf_counter() - t0
        extra["cpu_baseline"] = {"value": nsq / dt, "unit": "lookups/s", "cores": min(16, len(os.sched_getaffinity(0))),
                                 "kind": "port", "sample": f"{len(sample)} trees (every 64th + the largest), "
                                 f"page tree PAGESIZE 5000 point search + closed-form range counts, {dt:.1f} s",
                                 "mismatches_vs_gpu": mism}
    bytes_pts = 144 * nq
    line = common_line("batched point lookups/s @4096 uint64 trees (one persistent hipGraph launch)", nq / wall,
                       "lookups/s", args, wall * 1e3, "u64",
                       {"workload": "4096 uint64 trees, sizes log-uniform in [1K, 1M] (keys uniform in [0, 2^64), "
                                    "unique, value = rank); 4x size point queries per tree (80% hits) + size/16 "
                                    "range queries per tree (widths log-uniform in [2^4, 2^16] ranks) in the same "
                                    "launch",
                        "trees": ntrees, "keys": int(key_off[-1])},
                       {"bound": "hbm", "achieved": bytes_pts / (pms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": bytes_pts / (pms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "k_forest_batch (points only)", "kernel_ms": pms, "bytes_per_lookup": 144},
                       extra)
    batch.close()
    batch_p.close()
    forest.close()
    for a in (dk, dv, dq, dlo, dhi, dout, dfirst, dcount):
        a.free()
    return line


def compat_grouping_run(n):
    """Runs bptree_groups genGrouping on n uniform uint64 keys and returns its
    wall time (the call: copy in, the device pass loop, the group lists)."""
    import subprocess
    import tempfile
    exe = os.path.join(PKG, "bptree_groups")
    keys = np.random.default_rng(12).integers(0, 1 << 64, size=n, dtype=np.uint64)
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        keys.tofile(os.path.join(d, "k"))
        del keys
        r = subprocess.run([exe, "genGrouping", str(n), "33000", os.path.join(d, "o"), os.path.join(d, "k")],
                           capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            return {"error": r.stderr[-300:]}
        sizes = np.fromfile(os.path.join(d, "o.list"), dtype=np.uint64)
    fields = dict(kv.split("=") for kv in r.stdout.split() if "=" in kv)
    return {"ms": float(fields["ms"]), "groups": int(fields["groups"]), "keys": n,
            "sizes_sum_to_n": bool(int(sizes.sum()) == n), "max_group": int(sizes.max()),
            "note": "wall time of the call from host memory: H2D 512 MB, device histograms and segmented "
                    "reorders per pass, the merge on the host; grouped keys stay on the device"}


def serving_run(tree, keys, n, miss_space, nreq, modes, workload):
    """The serving loop over a resident tree: an open-loop stream of nreq mixed
    batches (sizes floor(2^u), u uniform in [0, 16]; 90% point batches with
    10% misses uniform in [0, miss_space), 10% range batches of widths
    log-uniform in [2^4, 2^12] keys). For each mode: saturation (all requests
    at once, median of three replays), then Poisson arrivals at 0.5x and 0.8x
    of it, p50/p99 latency.
    Every answer is checked against the tree's own search of the same keys
    in one batch (the serving tests check them against the oracle)."""
    kdt = keys.dtype.type
    vdt = np.uint32 if tree.info()["val_bytes"] == 4 else np.uint64
    rr = workloads.rng_for(0, workload, 2)
    reqs, keep = [], []
    for _ in range(nreq):
        m = int(np.floor(np.exp2(rr.uniform(0, 16))))
        if rr.random() < 0.9:
            qq = np.where(rr.random(m) < 0.9, keys[rr.integers(0, n, m)],
                          rr.integers(0, miss_space, m, dtype=np.uint64).astype(kdt)).astype(kdt)
            out = np.zeros(m, dtype=vdt)
            reqs.append(bptree.Server.point_request(qq, out))
            keep.append(("p", qq, out))
        else:
            st_ = rr.integers(0, n, m)
            w = np.floor(np.exp2(rr.uniform(4, 12, m))).astype(np.int64)
            lo, hi = keys[st_], keys[np.minimum(st_ + w, n - 1)]
            first, cnt = np.zeros(m, dtype=np.uint64), np.zeros(m, dtype=np.uint64)
            reqs.append(bptree.Server.range_request(lo, hi, first, cnt))
            keep.append(("r", lo, hi, first, cnt))
    total_lookups = sum(int(r.n) for r in reqs)
    pq = np.concatenate([k[1] for k in keep if k[0] == "p"])
    want_p = tree.point_search(pq, bptree.PLAIN)
    # the same requests with device buffers (slices of one device array per
    # buffer kind): no host staging, results left on the device
    rk = [k for k in keep if k[0] == "r"]
    dq = bptree.DeviceArray.from_host(pq)
    dout = bptree.DeviceArray(len(pq), vdt)
    dlo = bptree.DeviceArray.from_host(np.concatenate([k[1] for k in rk]))
    dhi = bptree.DeviceArray.from_host(np.concatenate([k[2] for k in rk]))
    dfirst, dcount = bptree.DeviceArray(len(dlo), np.uint64), bptree.DeviceArray(len(dlo), np.uint64)
    # and with pinned host buffers that the device addresses directly
    # (bpt_host_alloc): no staging copies, the data crosses PCIe in the kernels
    hq, hout = bptree.HostArray.from_host(pq), bptree.HostArray(len(pq), vdt)
    hlo = bptree.HostArray.from_host(np.concatenate([k[1] for k in rk]))
    hhi = bptree.HostArray.from_host(np.concatenate([k[2] for k in rk]))
    hfirst, hcount = bptree.HostArray(len(hlo), np.uint64), bptree.HostArray(len(hlo), np.uint64)
    dreqs, hreqs, pa, ra = [], [], 0, 0
    for k in keep:
        if k[0] == "p":
            dreqs.append(bptree.Server.point_request_dev(dq, dout, pa, len(k[1])))
            hreqs.append(bptree.Server.point_request_dev(hq, hout, pa, len(k[1])))
            pa += len(k[1])
        else:
            dreqs.append(bptree.Server.range_request_dev(dlo, dhi, dfirst, dcount, ra, len(k[1])))
            hreqs.append(bptree.Server.range_request_dev(hlo, hhi, hfirst, hcount, ra, len(k[1])))
            ra += len(k[1])

    def measure(srv, rq):
        srv.replay(rq[:200], np.zeros(200))  # warm-up
        # saturation: every request due at once, the median of three replays
        # (one replay takes a few ms, short enough to vary between runs); the
        # ctypes request array is built once, outside the timed call
        arr, due = bptree.Server.request_array(rq), np.zeros(nreq)
        sats, sat_stats = [], []
        for _ in range(3):
            srv.reset_stats()
            t0 = time.perf_counter()
            srv.replay(arr, due)
            sats.append(nreq / (time.perf_counter() - t0))
            sat_stats.append(srv.stats())
        mid = int(np.argsort(sats)[1])
        sat = sats[mid]
        res = {"saturation_requests_per_s": sat, "saturation_runs_requests_per_s": sats,
               "saturation": sat_stats[mid]}
        for frac in (0.5, 0.8):
            srv.reset_stats()
            srv.replay(arr, np.cumsum(rr.exponential(1.0 / (frac * sat), size=nreq)))
            stt = srv.stats()
            res[f"load_{frac}"] = {"rate_requests_per_s": frac * sat, "p50_ms": stt["p50_ms"], "p99_ms": stt["p99_ms"],
                                   "lookups_per_s": stt["lookups"] / stt["elapsed_s"] if stt["elapsed_s"] else None,
                                   "max_inflight": stt["max_inflight"]}
        return res

    serving = {}
    for mode in modes:
        name = {bptree.PLAIN: "plain", bptree.GROUPED: "grouped", bptree.LEARNED: "learned"}[mode]
        srv = bptree.Server(tree, slots=8, mem_budget=8 << 30, mode=mode)
        res = measure(srv, reqs)
        got_p = np.concatenate([k[2] for k in keep if k[0] == "p"])
        res["point_mismatches_vs_one_batch"] = int((got_p != want_p).sum())
        dout.fill_bytes(0)
        res_d = measure(srv, dreqs)
        res_d["point_mismatches_vs_one_batch"] = int((dout.to_host() != want_p).sum())
        res_d["note"] = "requests with device buffers (BPT_REQ_DEVICE): no host staging, results left on the device"
        hout.a[:] = 0
        res_h = measure(srv, hreqs)
        res_h["point_mismatches_vs_one_batch"] = int((hout.a != want_p).sum())
        res_h["note"] = ("requests in pinned host memory (bpt_host_alloc) passed as BPT_REQ_DEVICE: no staging "
                         "copies, the kernels read and write them over PCIe")
        srv.close()
        serving[name] = res
        serving[name + "_device"] = res_d
        serving[name + "_pinned_host"] = res_h
        log(f"serving {name}: " + json.dumps(res))
        log(f"serving {name}, device buffers: " + json.dumps(res_d))
        log(f"serving {name}, pinned host buffers: " + json.dumps(res_h))
    for a in (dq, dout, dlo, dhi, dfirst, dcount, hq, hout, hlo, hhi, hfirst, hcount):
        a.free()
    return serving, nreq, total_lookups


def bench_c5(args):
    """256M unique uint64 lognormal keys: learned position model vs plain
    descent (throughput), then an open-loop Poisson stream of mixed batches
    through the serving loop at 0.5x and 0.8x of its saturation rate."""
    n = 1 << 28

    def dev_sort(a):
        return bptree.sort_pairs(a)[0]

    t0 = time.perf_counter()
    keys = workloads.lognormal_keys(n, sort=dev_sort)
    vals = np.arange(n, dtype=np.uint64)
    log(f"c5: {n} lognormal keys in {time.perf_counter() - t0:.1f} s")
    tree = bptree.Tree(keys, vals, 8, 8)
    lst = tree.learned_build(leaf_models=65536, max_window=256, train_samples=1 << 20, train_steps=2000, root=0)
    log(f"learned: {lst}")
    rng = workloads.rng_for(0, 5, 1)
    nq = 1 << 27
    q = np.where(rng.random(nq) < 0.9, keys[rng.integers(0, n, nq)],
                 rng.integers(0, 1 << 50, nq, dtype=np.uint64)).astype(np.uint64)
    dq = bptree.DeviceArray.from_host(q)
    dout = bptree.DeviceArray(nq, np.uint64)
    forms = {}
    outs = {}
    for mode, name in ((bptree.PLAIN, "plain"), (bptree.LEARNED, "learned")):
        wall, ev = time_form(tree, dq, dout, mode, args.steps, args.warmup)
        forms[name] = {"lookups_per_s": nq / wall, "kernel_ms": float(np.mean(ev))}
        outs[name] = dout.to_host()
        log(f"c5 {name}: {wall * 1e3:.2f} ms, {nq / wall / 1e9:.2f} G lookups/s")
    forms["learned"]["speedup_vs_plain"] = forms["plain"]["kernel_ms"] / forms["learned"]["kernel_ms"]
    forms["learned"]["mismatches_vs_plain"] = int((outs["plain"] != outs["learned"]).sum())
    if args.learned_only:
        print(json.dumps({"form": "c5_lookups", "forms": forms, "learned_model": lst, "steps": args.steps}),
              flush=True)
        tree.close()
        return
    # layer-by-layer root evaluation, for comparison with the piecewise-linear form
    tree.learned_build(leaf_models=65536, max_window=256, train_samples=1 << 20, train_steps=2000, root=1)
    wall, ev = time_form(tree, dq, dout, bptree.LEARNED, max(2, args.steps // 2), 1)
    forms["learned_layerwise_root"] = {"lookups_per_s": nq / wall, "kernel_ms": float(np.mean(ev)),
                                       "mismatches_vs_plain": int((dout.to_host() != outs["plain"]).sum())}
    tree.learned_build(leaf_models=65536, max_window=256, train_samples=1 << 20, train_steps=2000, root=0)
    dq.free()
    dout.free()

    serving, nreq, total_lookups = serving_run(tree, keys, n, 1 << 50, 3000, (bptree.LEARNED, bptree.PLAIN), 5)
    extra = {"forms": forms, "learned_model": lst, "serving": serving, "serving_requests": nreq,
             "serving_lookups": total_lookups}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = min(16, len(os.sched_getaffinity(0)))
        pt = oracle.PageTree(keys, vals, 5000)
        sample = q[: 1 << 24]
        t0 = time.perf_counter()
        want = pt.point(sample, 0xFFFFFFFFFFFFFFFF, threads)
        dt = time.perf_counter() - t0
        pt.close()
        extra["cpu_baseline"] = {"value": len(sample) / dt, "unit": "lookups/s", "cores": threads, "kind": "port",
                                 "sample": f"first {len(sample)} queries, page tree PAGESIZE 5000, {dt:.1f} s",
                                 "mismatches_vs_gpu": int((want != outs["learned"][: len(sample)]).sum())}
    lk = forms["learned"]
    line = common_line("learned-index point lookups/s @256M lognormal uint64 keys; p50/p99 ms serving",
                       lk["lookups_per_s"], "lookups/s", args, lk["kernel_ms"], "u64",
                       {"workload": "256M unique uint64 keys, lognormal(0, 2) scaled to [0, 2^50], value = rank; "
                                    "MLP 1-32-32-1 root over 65536 linear leaves, window cap 256; 128M point "
                                    "queries (90% hits); serving: 3000 mixed batches (sizes log-uniform 1..64K, "
                                    "90% point / 10% range of 2^4..2^12 keys), open-loop Poisson",
                        "keys": n, "queries": nq},
                       {"bound": "hbm", "achieved": 144 * nq / (lk["kernel_ms"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": 144 * nq / (lk["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "traffic": None, "kernel": "k_point_learned", "kernel_ms": lk["kernel_ms"],
                        "bytes_per_lookup": 144},
                       extra)
    tree.close()
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--keys-log2", type=int, default=26)
    ap.add_argument("--queries-log2", type=int, default=28)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip grouped / Zipf side measurements")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 2-D rectangle, forest and learned-index configurations after the headline")
    ap.add_argument("--grouped-only", action="store_true",
                    help="time only the grouped form on the uniform set (for profiling its kernels)")
    ap.add_argument("--batch-sweep", action="store_true",
                    help="plain vs grouped kernels vs the grouped mode's dispatch over batch sizes 2^16..2^26 "
                         "on the 64M-key tree (wall time per call)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run the 1 GiB device copy kernel (HBM counter calibration)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="with --config c4: skip the autotuner and use its usual choice (for profiling)")
    ap.add_argument("--learned-only", action="store_true",
                    help="with --config c5: only the plain and learned lookups (for profiling)")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"],
                    help="c2 (default): 64M-key point search; c1: 1M-key CPU page tree (the reference's CPU "
                         "case); c3: 2-D rectangles; c4: 4096-tree batch; c5: learned index + serving")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is "
                                   "designed (range-shard or replicate) but not built"}))
        sys.exit(2)
    if args.config in ("c3", "c4", "c5"):
        line = {"c3": bench_c3, "c4": bench_c4, "c5": bench_c5}[args.config](args)
        print(json.dumps(line), flush=True)
        return
    if args.config == "c1":
        return bench_c1(args)

    n, nq = 1 << args.keys_log2, 1 << args.queries_log2
    dev = bptree.device_info(0)
    log(f"device {dev['name']} cus={dev['cus']}")
    t0 = time.perf_counter()
    keys, vals = workloads.point_keys(n, workload=2)
    q = workloads.uniform_hits(keys, nq)
    log(f"generated {n} keys, {nq} queries in {time.perf_counter() - t0:.1f} s")

    dk = bptree.DeviceArray.from_host(keys)
    dv = bptree.DeviceArray.from_host(vals)
    timer = bptree.Timer()
    timer.start()
    tree = bptree.Tree(dk, dv, 4, 4)
    timer.stop()
    build_ms = timer.ms()
    info = tree.info()
    dk.free()
    dv.free()
    log(f"built tree: {info['height']} internal levels {info['level_nodes']} lds_levels={info['lds_levels']} "
        f"in {build_ms:.1f} ms")

    if args.calibrate_copy:
        a = bptree.DeviceArray(1 << 28, np.uint32)
        b = bptree.DeviceArray(1 << 28, np.uint32)
        for _ in range(3):
            bptree.check(bptree.lib().bpt_copy_d2d(b.ptr, a.ptr, a.nbytes, None))
        a.free()
        b.free()
        log("ran the 1 GiB copy kernel 3 times (reads 1 GiB, writes 1 GiB each)")

    # stream-copy bandwidth of this GPU (1 GiB device-to-device, read + write),
    # the practical ceiling the roofline fraction is also compared with
    a = bptree.DeviceArray(1 << 28, np.uint32)
    b = bptree.DeviceArray(1 << 28, np.uint32)
    copy_ms = []
    for _ in range(4):
        ct = bptree.Timer()
        ct.start()
        bptree.check(bptree.lib().bpt_copy_d2d(b.ptr, a.ptr, a.nbytes, None))
        ct.stop()
        copy_ms.append(ct.ms())
        ct.close()
    a.free()
    b.free()
    copy_gbps = 2 * (1 << 30) / (float(np.median(copy_ms[1:])) * 1e-3) / 1e9
    log(f"1 GiB device copy (k_copy_u4, read + write): {copy_gbps:.0f} GB/s")

    if args.batch_sweep:
        print(json.dumps({"batch_sweep": batch_sweep(tree, q)}), flush=True)
        tree.close()
        return
    dq = bptree.DeviceArray.from_host(q)
    dout = bptree.DeviceArray(nq, np.uint32)
    if args.grouped_only:
        wall_g, _ = time_form(tree, dq, dout, bptree.GROUPED, args.steps, args.warmup)
        print(json.dumps({"form": "grouped_uniform", "lookups_per_s": nq / wall_g, "ms": wall_g * 1e3,
                          "steps": args.steps}), flush=True)
        tree.close()
        return
    wall, ev = time_form(tree, dq, dout, bptree.PLAIN, args.steps, args.warmup)
    kernel_ms = float(np.mean(ev))
    value = nq / wall
    achieved = BYTES_PER_LOOKUP * nq / (kernel_ms * 1e-3) / 1e9
    log(f"plain: {wall * 1e3:.3f} ms/step wall, kernel avg {kernel_ms:.3f} ms, "
        f"{nq / (kernel_ms * 1e-3) / 1e9:.2f} G lookups/s")
    gpu_out = dout.to_host()

    forms = {"plain_uniform": {"lookups_per_s": value, "kernel_ms": kernel_ms}}
    zipf_sample = None
    if not args.no_forms:
        # end to end from host memory: queries H2D, plain search, results D2H
        # (pageable numpy buffers, synchronous copies), as a host caller sees it
        host_out = np.empty(nq, np.uint32)
        e2e = []
        for _ in range(3):
            t0 = time.perf_counter()
            bptree.check(bptree.lib().bpt_copy_h2d(dq.ptr, q.ctypes.data, q.nbytes, None))
            tree.point_search_dev(dq, dout, bptree.PLAIN)
            bptree.check(bptree.lib().bpt_copy_d2h(host_out.ctypes.data, dout.ptr, host_out.nbytes, None))
            e2e.append(time.perf_counter() - t0)
        wall_e = float(np.median(e2e))
        forms["plain_end_to_end"] = {"lookups_per_s": nq / wall_e, "ms": wall_e * 1e3,
                                     "identical_to_resident": bool(np.array_equal(host_out, gpu_out)),
                                     "note": "pageable host buffers; H2D 1 GiB + search + D2H 1 GiB"}
        # the same from pinned host buffers (bpt_host_alloc): DMA copies in and
        # out, and the search reading and writing the host buffers in place
        hq, hout = bptree.HostArray.from_host(q), bptree.HostArray(nq, np.uint32)
        e2e_dma, e2e_zc = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            bptree.check(bptree.lib().bpt_copy_h2d(dq.ptr, hq.ptr, hq.nbytes, None))
            tree.point_search_dev(dq, dout, bptree.PLAIN)
            bptree.check(bptree.lib().bpt_copy_d2h(hout.ptr, dout.ptr, hout.nbytes, None))
            e2e_dma.append(time.perf_counter() - t0)
        same_dma = bool(np.array_equal(hout.a, gpu_out))
        hout.a[:] = 0
        for _ in range(3):
            t0 = time.perf_counter()
            tree.point_search_dev(hq, hout, bptree.PLAIN)
            bptree.sync()
            e2e_zc.append(time.perf_counter() - t0)
        same_zc = bool(np.array_equal(hout.a, gpu_out))
        hq.free()
        hout.free()
        w_dma, w_zc = float(np.median(e2e_dma)), float(np.median(e2e_zc))
        forms["plain_end_to_end_pinned"] = {
            "dma": {"lookups_per_s": nq / w_dma, "ms": w_dma * 1e3, "identical_to_resident": same_dma},
            "in_place": {"lookups_per_s": nq / w_zc, "ms": w_zc * 1e3, "identical_to_resident": same_zc},
            "note": "pinned host buffers (bpt_host_alloc): H2D 1 GiB + search + D2H 1 GiB by DMA, or the search "
                    "reading the queries and writing the results in host memory over PCIe"}
        log("end to end, pinned: " + json.dumps(forms["plain_end_to_end_pinned"]))
        del host_out
        wall_g, ev_g = time_form(tree, dq, dout, bptree.GROUPED, max(3, args.steps // 2), 1)
        # the grouped form's own algorithmic bytes per lookup: the histogram
        # reads the query (4), the scatter reads it and writes it with its
        # 2-byte tile index (10), the search reads it and writes the result
        # (8) and reads the leaves once (leaf bytes / queries), the unpermute
        # reads the result and the index and writes the result (10)
        own = 4 + 10 + 8 + info["leaf_bytes"] / nq + 10
        forms["grouped_uniform"] = {"lookups_per_s": nq / wall_g, "ms": wall_g * 1e3,
                                    "speedup_vs_plain": wall / wall_g,
                                    "algorithmic_GBps": BYTES_PER_LOOKUP * nq / wall_g / 1e9,
                                    "own_bytes_per_lookup": own,
                                    "own_roofline": {"achieved": own * nq / wall_g / 1e9, "peak": HBM_PEAK_GBS,
                                                     "unit": "GB/s", "frac": own * nq / wall_g / 1e9 / HBM_PEAK_GBS},
                                    "mismatches_vs_plain": int((dout.to_host() != gpu_out).sum())}
        # the learned position model on the same tree: MLP 1-32-32-1 root over
        # 65536 linear leaf models, windows of at most 256 slots, trained on
        # the GPU; the plain descent answers the queries whose window is wider
        t0 = time.perf_counter()
        lst = tree.learned_build(leaf_models=65536, max_window=256, train_samples=1 << 20, train_steps=2000,
                                 root=0)
        train_s = time.perf_counter() - t0
        wall_l, ev_l = time_form(tree, dq, dout, bptree.LEARNED, max(3, args.steps // 2), 1)
        forms["learned_uniform"] = {"lookups_per_s": nq / wall_l, "ms": wall_l * 1e3,
                                    "kernel_ms": float(np.mean(ev_l)), "speedup_vs_plain": wall / wall_l,
                                    "mismatches_vs_plain": int((dout.to_host() != gpu_out).sum()),
                                    "build_s": train_s, "model": lst}
        log(f"learned: {wall_l * 1e3:.3f} ms ({wall / wall_l:.2f}x plain), model {lst}")
        qz = workloads.zipf_queries(keys, nq)
        dqz = bptree.DeviceArray.from_host(qz)
        wall_pz, _ = time_form(tree, dqz, dout, bptree.PLAIN, max(3, args.steps // 2), 1)
        plain_z = dout.to_host()
        zipf_sample = (qz[: 1 << 24].copy(), plain_z[: 1 << 24].copy())
        wall_gz, _ = time_form(tree, dqz, dout, bptree.GROUPED, max(3, args.steps // 2), 1)
        forms["plain_zipf"] = {"lookups_per_s": nq / wall_pz, "ms": wall_pz * 1e3}
        forms["grouped_zipf"] = {"lookups_per_s": nq / wall_gz, "ms": wall_gz * 1e3,
                                 "speedup_vs_plain": wall_pz / wall_gz,
                                 "mismatches_vs_plain": int((dout.to_host() != plain_z).sum())}
        wall_lz, _ = time_form(tree, dqz, dout, bptree.LEARNED, max(3, args.steps // 2), 1)
        forms["learned_zipf"] = {"lookups_per_s": nq / wall_lz, "ms": wall_lz * 1e3,
                                 "speedup_vs_plain": wall_pz / wall_lz,
                                 "mismatches_vs_plain": int((dout.to_host() != plain_z).sum())}
        dqz.free()
        del qz
        # 1-D ranges over the same tree: 2^22 [lo, hi] with widths log-uniform in
        # [1, 2^10] ranks; one step = counts, their scan, and the ids, in each order
        rng_r = np.random.default_rng(11)
        nrq = 1 << 22
        st_ = rng_r.integers(0, n, nrq)
        wr = np.floor(np.exp2(rng_r.uniform(0, 10, nrq))).astype(np.int64)
        dlo = bptree.DeviceArray.from_host(keys[st_])
        dhi = bptree.DeviceArray.from_host(keys[np.minimum(st_ + wr, n - 1)])
        dfirst = bptree.DeviceArray(nrq, np.uint64)
        dcnt = bptree.DeviceArray(nrq, np.uint64)
        doff = bptree.DeviceArray(nrq, np.uint64)
        L = bptree.lib()
        base = None
        # the caller owns the ids buffer: kept across steps, grown when a
        # step's total exceeds it (the C ABI writes into caller memory)
        ids_buf = [None]

        def ids_for(total):
            if ids_buf[0] is None or ids_buf[0].n < total:
                if ids_buf[0] is not None:
                    ids_buf[0].free()
                ids_buf[0] = bptree.DeviceArray(max(total, 1), np.uint32)
            return ids_buf[0]
        for gm, gname in ((bptree.GROUP_NONE, "basic"), (bptree.GROUP_GRID2D, "grid2d"),
                          (bptree.GROUP_ZORDER, "zorder")):
            def rstep():
                tree.range_count_dev(dlo, dhi, dfirst, dcnt, gm)
                tot = bptree.C.c_uint64(0)
                bptree.check(L.bpt_exclusive_scan_u64(dcnt.ptr, nrq, doff.ptr, bptree.C.byref(tot), None))
                dids = ids_for(tot.value)
                bptree.check(L.bpt_range_ids(tree.h, dfirst.ptr, dcnt.ptr, doff.ptr, nrq, dids.ptr, None))
                return tot.value
            rstep()
            bptree.sync()
            t0 = time.perf_counter()
            reps = 5
            for _ in range(reps):
                rstep()
                bptree.sync()
            dt = (time.perf_counter() - t0) / reps
            tot_ids = rstep()
            cnt_h = dcnt.to_host()
            ids_h = ids_buf[0].to_host(count=min(tot_ids, 1 << 24))
            if base is None:
                base = (cnt_h, ids_h)
            forms[f"range_{gname}"] = {"ranges_per_s": nrq / dt, "ms": dt * 1e3, "ids": int(cnt_h.sum()),
                                       "ids_per_s": float(cnt_h.sum()) / dt,
                                       "identical_to_basic": bool(np.array_equal(cnt_h, base[0])
                                                                  and np.array_equal(ids_h, base[1]))}
        for a in (dlo, dhi, dfirst, dcnt, doff, ids_buf[0]):
            a.free()
        # the reference-named genGrouping (host keys in, grouped keys on the
        # device out) over 2^26 uniform uint64 keys with the reference's group
        # cap of 33000, through the compiled drop-in driver
        forms["compat_genGrouping_2^26"] = compat_grouping_run(1 << 26)
        log("genGrouping: " + json.dumps(forms["compat_genGrouping_2^26"]))
        # the serving loop over this tree (p50 / p99 latency)
        serving, nreq, sl = serving_run(tree, keys, n, 1 << 32, 3000, (bptree.PLAIN, bptree.LEARNED), 2)
        forms["serving"] = {"requests": nreq, "lookups": sl, **serving}
        # batched updates: batches of 2^20 upserts (half of them existing keys)
        # merged into the tree, which is rebuilt; wall time of each whole call.
        # The first call also allocates the merge's workspace and regrows the
        # leaves; the later ones reuse both.
        rng_u = np.random.default_rng(7)
        nu = 1 << 20
        walls = []
        for _ in range(4):
            ku = np.concatenate([keys[rng_u.integers(0, n, nu // 2)],
                                 rng_u.integers(0, 1 << 32, nu - nu // 2, dtype=np.uint64).astype(np.uint32)])
            vu = rng_u.integers(0, 1 << 32, nu, dtype=np.uint64).astype(np.uint32)
            dku, dvu = bptree.DeviceArray.from_host(ku), bptree.DeviceArray.from_host(vu)
            bptree.sync()
            t0 = time.perf_counter()
            tree.upsert(dku, dvu)
            walls.append(time.perf_counter() - t0)
            dku.free()
            dvu.free()
        wall_u = float(np.median(walls[1:]))
        forms["upsert_2^20"] = {"ms": wall_u * 1e3, "pairs_per_s": nu / wall_u, "ms_first_call": walls[0] * 1e3,
                                "keys_before": n, "keys_after": tree.info()["n"]}
        log("forms: " + json.dumps(forms))

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "point_plain_64M_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("l2_egress_bytes_per_launch")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "lookups/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded numpy PCG64, seed 0)",
        "config": {
            "workload": "GPU point search, plain descent: 64M unique uint32 keys uniform in [0,2^32), "
                        "value = rank; 256M queries uniform over the key set",
            "keys": n, "queries": nq, "fanout": info["fanout"], "leaf_cap": info["leaf_cap"],
            "height": info["height"], "build_ms": build_ms,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_counts": "bytes leaving the L2 (TCC FETCH_SIZE + WRITE_SIZE per launch, scaled by the "
                              "same counters on a copy kernel of known size); gfx950 has no counter behind the "
                              "Infinity Cache, so Infinity Cache hits are included: an upper bound on HBM bytes",
            "kernel": "k_point_coop", "kernel_ms": kernel_ms,
            "bytes_per_lookup": BYTES_PER_LOOKUP,
            "stream_copy_GBps": copy_gbps, "frac_of_copy": achieved / copy_gbps,
        },
        "forms": forms,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(keys, vals, q, gpu_out, zipf=zipf_sample)
    tree.close()
    dq.free()
    dout.free()
    if not args.no_configs:
        # the other configurations at their full sizes, summarised beside the
        # headline (each also runs alone with --config)
        del keys, vals, q, gpu_out, zipf_sample
        line["configs"] = {}
        for name, fn in (("c3", bench_c3), ("c4", bench_c4), ("c5", bench_c5)):
            t0 = time.perf_counter()
            sub = fn(args)
            line["configs"][name] = summarize(name, sub)
            line["configs"][name]["bench_s"] = time.perf_counter() - t0
            log(f"{name}: " + json.dumps(line["configs"][name]))
    print(json.dumps(line), flush=True)


def summarize(name, sub):
    """The headline figures of one configuration's line."""
    keep = {"metric": sub["metric"], "value": sub["value"], "unit": sub["unit"], "ms_per_step": sub["ms_per_step"],
            "workload": sub["config"]["workload"]}
    rf = sub.get("roofline") or {}
    if rf.get("frac") is not None:
        keep["roofline"] = {k: rf.get(k) for k in ("achieved", "peak", "unit", "frac", "kernel", "kernel_ms",
                                                   "bytes_per_lookup")}
    cb = sub.get("cpu_baseline")
    if cb:
        keep["cpu_baseline"] = {k: cb.get(k) for k in ("value", "unit", "cores", "kind", "sample",
                                                       "mismatches_vs_gpu")}
    if name == "c3":
        keep["forms"] = sub["forms"]
        keep["hits_total"] = sub["hits_total"]
    elif name == "c4":
        keep.update({k: sub[k] for k in ("autotune", "speedup_vs_per_tree_launches", "points", "ranges",
                                        "batch_ms", "points_only_ms")})
    elif name == "c5":
        keep["forms"] = {k: {kk: vv for kk, vv in v.items() if kk != "model"} for k, v in sub["forms"].items()}
        keep["learned_model"] = sub["learned_model"]
        keep["serving"] = {m: {k: v for k, v in r.items() if k != "saturation"} for m, r in sub["serving"].items()}
    return keep


if __name__ == "__main__":
    main()
