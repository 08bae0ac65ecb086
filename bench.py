#!/usr/bin/env python3
"""Benchmark of the B+tree point-search hot path on one MI355X.

Workload: 64M (2^26) unique uint32 keys uniform over [0, 2^32), value = rank,
bulk-built into the flat gfx950 tree (fanout 32, 16-slot leaves); one step =
one plain point-search pass over 256M (2^28) queries drawn uniformly from the
key set, with queries, tree and results resident in HBM. Also measured
beside it (not the headline value): the radix-grouped form and the Zipf(0.99)
query set with 10% misses.

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


def cpu_baseline(keys, vals, q, gpu_out, target_s=10.0):
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
    pt.close()
    return {
        "value": s / dt, "unit": "lookups/s", "cores": threads, "kind": "port",
        "sample": f"first {s} of the {len(q)} uniform-hit queries against the full 64M-key page "
                  f"tree (PAGESIZE 5000, reference default), OpenMP static schedule, "
                  f"{dt:.1f} s; page-tree build {build_s:.1f} s excluded",
        "mismatches_vs_gpu": mism,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--keys-log2", type=int, default=26)
    ap.add_argument("--queries-log2", type=int, default=28)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip grouped / Zipf side measurements")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run the 1 GiB device copy kernel (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is "
                                   "designed (range-shard or replicate) but not built"}))
        sys.exit(2)

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

    dq = bptree.DeviceArray.from_host(q)
    dout = bptree.DeviceArray(nq, np.uint32)
    wall, ev = time_form(tree, dq, dout, bptree.PLAIN, args.steps, args.warmup)
    kernel_ms = float(np.mean(ev))
    value = nq / wall
    achieved = BYTES_PER_LOOKUP * nq / (kernel_ms * 1e-3) / 1e9
    log(f"plain: {wall * 1e3:.3f} ms/step wall, kernel avg {kernel_ms:.3f} ms, "
        f"{nq / (kernel_ms * 1e-3) / 1e9:.2f} G lookups/s")
    gpu_out = dout.to_host()

    forms = {"plain_uniform": {"lookups_per_s": value, "kernel_ms": kernel_ms}}
    if not args.no_forms:
        wall_g, ev_g = time_form(tree, dq, dout, bptree.GROUPED, max(3, args.steps // 2), 1)
        forms["grouped_uniform"] = {"lookups_per_s": nq / wall_g, "ms": wall_g * 1e3,
                                    "speedup_vs_plain": wall / wall_g,
                                    "mismatches_vs_plain": int((dout.to_host() != gpu_out).sum())}
        qz = workloads.zipf_queries(keys, nq)
        dqz = bptree.DeviceArray.from_host(qz)
        del qz
        wall_pz, _ = time_form(tree, dqz, dout, bptree.PLAIN, max(3, args.steps // 2), 1)
        plain_z = dout.to_host()
        wall_gz, _ = time_form(tree, dqz, dout, bptree.GROUPED, max(3, args.steps // 2), 1)
        forms["plain_zipf"] = {"lookups_per_s": nq / wall_pz, "ms": wall_pz * 1e3}
        forms["grouped_zipf"] = {"lookups_per_s": nq / wall_gz, "ms": wall_gz * 1e3,
                                 "speedup_vs_plain": wall_pz / wall_gz,
                                 "mismatches_vs_plain": int((dout.to_host() != plain_z).sum())}
        dqz.free()
        log("forms: " + json.dumps(forms))

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "point_plain_64M_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

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
            "kernel": "k_point_coop", "kernel_ms": kernel_ms,
            "bytes_per_lookup": BYTES_PER_LOOKUP,
        },
        "forms": forms,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(keys, vals, q, gpu_out)
    tree.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
