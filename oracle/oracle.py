"""ctypes front end of the CPU oracle (``bpt_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke``
and by the CPU-baseline leg of ``bench.py``. The product library never calls
it. Each wrapper names the reference file:line its C function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbpt_oracle.so")
_lib = None

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)


def build() -> str:
    """Compile the oracle with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.ora_pt_build.argtypes = [u64p, u64p, C.c_uint64, C.c_uint64, C.POINTER(C.c_void_p)]
        L.ora_pt_build.restype = C.c_int
        L.ora_pt_free.argtypes = [C.c_void_p]
        L.ora_pt_point.argtypes = [C.c_void_p, u64p, C.c_uint64, u64p, C.c_uint64, C.c_int]
        L.ora_pt_range_locate.argtypes = [C.c_void_p, u64p, C.c_uint64, u64p]
        L.ora_pt_range_count_refcompat.argtypes = [C.c_void_p, u64p, u64p, C.c_uint64, u64p]
        L.ora_lower_bound.argtypes = [u64p, C.c_uint64, u64p, C.c_uint64, u64p]
        L.ora_point_closed.argtypes = [u64p, u64p, C.c_uint64, u64p, C.c_uint64, u64p, C.c_uint64]
        L.ora_range_closed.argtypes = [u64p, C.c_uint64, u64p, u64p, C.c_uint64, u64p, u64p]
        L.ora_morton2d.argtypes = [C.c_uint32, C.c_uint32]
        L.ora_morton2d.restype = C.c_uint64
        L.ora_demorton2d.argtypes = [C.c_uint64, u32p, u32p]
        L.ora_zorder_keys.argtypes = [u64p, u64p, C.c_uint64, u64p]
        L.ora_zorder_interval.argtypes = [C.c_uint64]
        L.ora_zorder_interval.restype = C.c_int
        L.ora_zorder_divisor.argtypes = [C.c_uint64]
        L.ora_zorder_divisor.restype = C.c_uint64
        L.ora_group_lists.argtypes = [u64p, u64p, C.c_uint64, C.c_uint64, C.c_int, C.c_int, u64p, u64p, u64p]
        L.ora_group_lists.restype = C.c_uint64
        L.ora_bucket_identify.argtypes = [C.c_uint64, C.c_int]
        L.ora_bucket2d.argtypes = [C.c_uint64, C.c_uint64, C.c_int]
        L.ora_group_start_pass.argtypes = [C.c_uint64]
        L.ora_group2d_start_pass.argtypes = [C.c_uint64, C.c_uint64]
        L.ora_merge_groups.argtypes = [u64p, C.c_uint64, C.c_uint64, u64p, u64p]
        L.ora_merge_groups.restype = C.c_uint64
        L.ora_rect_bruteforce.argtypes = [u32p, u32p, u32p, C.c_uint64, u32p, C.c_uint64, u64p,
                                          u64p, u32p, C.c_int]
        L.ora_rect_morton.argtypes = [u64p, u32p, C.c_uint64, u32p, C.c_uint64, u64p, u64p, u32p,
                                      C.c_int]
        L.ora_bigmin.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.ora_bigmin.restype = C.c_uint64
        L.ora_mlp_loss_grad.argtypes = [f64p, C.c_int, f64p, f64p, C.c_uint64, f64p]
        L.ora_mlp_loss_grad.restype = C.c_double
        L.ora_mlp_forward.argtypes = [f64p, C.c_int, f64p, C.c_uint64, f64p]
        L.ora_omp_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(u64p)


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(u32p)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(f64p)


class PageTree:
    """Reference page tree (BPTree.cu:28-125,218-243) with runtime PAGESIZE."""

    def __init__(self, keys, data, pagesize):
        self._k, kp = _u64(keys)
        self._d, dp = _u64(data)
        h = C.c_void_p()
        rc = lib().ora_pt_build(kp, dp, len(self._k), int(pagesize), C.byref(h))
        if rc != 0:
            raise ValueError("ora_pt_build failed (n == 0 or pagesize < 2)")
        self.h = h
        self.pagesize = int(pagesize)

    def close(self):
        if self.h:
            lib().ora_pt_free(self.h)
            self.h = None

    __del__ = close

    @property
    def levels(self):
        # ora_pt fields: pagesize, npages, 8 pointers, root, levels
        raw = C.cast(self.h, C.POINTER(C.c_uint64))
        return int(raw[11])

    @property
    def root(self):
        raw = C.cast(self.h, C.POINTER(C.c_uint64))
        return int(raw[10])

    @property
    def npages(self):
        raw = C.cast(self.h, C.POINTER(C.c_uint64))
        return int(raw[1])

    def point(self, q, sentinel=0xFFFFFFFFFFFFFFFF, threads=0):
        """BPTreeExactSearch.cu:25-105."""
        q, qp = _u64(q)
        out = np.empty(len(q), dtype=np.uint64)
        lib().ora_pt_point(self.h, qp, len(q), out.ctypes.data_as(u64p), sentinel, threads)
        return out

    def range_locate(self, q):
        """BtreeRangeSearch.cu:215-294: (pos << 32) | leaf."""
        q, qp = _u64(q)
        out = np.empty(len(q), dtype=np.uint64)
        lib().ora_pt_range_locate(self.h, qp, len(q), out.ctypes.data_as(u64p))
        return out

    def range_count_refcompat(self, lo, hi):
        """BtreeRangeSearch.cu:297-374 (reference boundary behaviour)."""
        lo, lp = _u64(lo)
        hi, hp = _u64(hi)
        out = np.empty(len(lo), dtype=np.uint64)
        lib().ora_pt_range_count_refcompat(self.h, lp, hp, len(lo), out.ctypes.data_as(u64p))
        return out


def lower_bound(keys, q):
    keys, kp = _u64(keys)
    q, qp = _u64(q)
    out = np.empty(len(q), dtype=np.uint64)
    lib().ora_lower_bound(kp, len(keys), qp, len(q), out.ctypes.data_as(u64p))
    return out


def point_closed(keys, vals, q, sentinel):
    keys, kp = _u64(keys)
    vals, vp = _u64(vals)
    q, qp = _u64(q)
    out = np.empty(len(q), dtype=np.uint64)
    lib().ora_point_closed(kp, vp, len(keys), qp, len(q), out.ctypes.data_as(u64p), sentinel)
    return out


def range_closed(keys, lo, hi):
    keys, kp = _u64(keys)
    lo, lp = _u64(lo)
    hi, hp = _u64(hi)
    first = np.empty(len(lo), dtype=np.uint64)
    cnt = np.empty(len(lo), dtype=np.uint64)
    lib().ora_range_closed(kp, len(keys), lp, hp, len(lo), first.ctypes.data_as(u64p),
                           cnt.ctypes.data_as(u64p))
    return first, cnt


def morton2d(x, y):
    return int(lib().ora_morton2d(int(x), int(y)))


def demorton2d(z):
    a, b = C.c_uint32(), C.c_uint32()
    lib().ora_demorton2d(int(z), C.byref(a), C.byref(b))
    return a.value, b.value


def zorder_keys(a, b):
    """zOrder2d.cu:27-54,125-130."""
    a, ap = _u64(a)
    b, bp = _u64(b)
    out = np.empty(len(a), dtype=np.uint64)
    lib().ora_zorder_keys(ap, bp, len(a), out.ctypes.data_as(u64p))
    return out


def zorder_interval(m):
    """zOrder2d.cu:129-130: the int interval (INT_MIN when it overflows)."""
    return int(lib().ora_zorder_interval(int(m)))


def zorder_divisor(m):
    """The 64-bit divisor that interval becomes in zOrder2d.cu:27-29,57."""
    return int(lib().ora_zorder_divisor(int(m)))


def group_lists(a, b, cap, kind, start_pass):
    """Group lists of the reference's grouping pass loop (ora_group_lists).

    kind 0: genGrouping / genGroupingWithOnedata / genGroupingWithTwodata
    (grouping.cu); kind 1: genGrouping2d (grouping2d.cu), b required.
    Returns (sizes, prefixes, perm) with perm the grouped order of indices."""
    a, ap = _u64(a)
    n = len(a)
    if kind == 1:
        b, bp = _u64(b)
    else:
        b, bp = _u64(np.zeros(1, dtype=np.uint64))
    perm = np.arange(n, dtype=np.uint64)
    sizes = np.zeros(n + 1, dtype=np.uint64)
    pre = np.zeros(n + 1, dtype=np.uint64)
    g = lib().ora_group_lists(ap, bp, n, int(cap), int(kind), int(start_pass), perm.ctypes.data_as(u64p),
                              sizes.ctypes.data_as(u64p), pre.ctypes.data_as(u64p))
    return sizes[:g].copy(), pre[:g].copy(), perm


def bucket_identify(x, p):
    return int(lib().ora_bucket_identify(int(x), int(p)))


def bucket2d(a, b, p):
    return int(lib().ora_bucket2d(int(a), int(b), int(p)))


def group_start_pass(m):
    return int(lib().ora_group_start_pass(int(m)))


def group2d_start_pass(a, b):
    return int(lib().ora_group2d_start_pass(int(a), int(b)))


def merge_groups(hist, cap):
    hist, hp = _u64(hist)
    sizes = np.empty(len(hist), dtype=np.uint64)
    first = np.empty(len(hist), dtype=np.uint64)
    ng = lib().ora_merge_groups(hp, len(hist), int(cap), sizes.ctypes.data_as(u64p),
                                first.ctypes.data_as(u64p))
    return sizes[:ng].copy(), first[:ng].copy()


def rect_bruteforce(xs, ys, pid, rects, with_ids=True, threads=0):
    xs, xp = _u32(xs)
    ys, yp = _u32(ys)
    pid, pp = _u32(pid)
    rects, rp = _u32(np.asarray(rects).reshape(-1))
    nr = len(rects) // 4
    counts = np.empty(nr, dtype=np.uint64)
    lib().ora_rect_bruteforce(xp, yp, pp, len(xs), rp, nr, counts.ctypes.data_as(u64p),
                              None, None, threads)
    if not with_ids:
        return counts, None, None
    offsets = np.zeros(nr, dtype=np.uint64)
    if nr:
        offsets[1:] = np.cumsum(counts)[:-1]
    ids = np.empty(int(counts.sum()), dtype=np.uint32)
    lib().ora_rect_bruteforce(xp, yp, pp, len(xs), rp, nr, counts.ctypes.data_as(u64p),
                              offsets.ctypes.data_as(u64p), ids.ctypes.data_as(u32p), threads)
    return counts, offsets, ids


def rect_morton(keys, pid, rects, with_ids=True, threads=0):
    keys, kp = _u64(keys)
    pid, pp = _u32(pid)
    rects, rp = _u32(np.asarray(rects).reshape(-1))
    nr = len(rects) // 4
    counts = np.empty(nr, dtype=np.uint64)
    lib().ora_rect_morton(kp, pp, len(keys), rp, nr, counts.ctypes.data_as(u64p), None, None,
                          threads)
    if not with_ids:
        return counts, None, None
    offsets = np.zeros(nr, dtype=np.uint64)
    if nr:
        offsets[1:] = np.cumsum(counts)[:-1]
    ids = np.empty(int(counts.sum()), dtype=np.uint32)
    lib().ora_rect_morton(kp, pp, len(keys), rp, nr, counts.ctypes.data_as(u64p),
                          offsets.ctypes.data_as(u64p), ids.ctypes.data_as(u32p), threads)
    return counts, offsets, ids


def bigmin(z, zmin, zmax):
    return int(lib().ora_bigmin(int(z), int(zmin), int(zmax)))


def mlp_loss_grad(params, H, x, t):
    p, pp = _f64(params)
    x, xp = _f64(x)
    t, tp = _f64(t)
    g = np.empty_like(p)
    loss = lib().ora_mlp_loss_grad(pp, int(H), xp, tp, len(x), g.ctypes.data_as(f64p))
    return float(loss), g


def mlp_forward(params, H, x):
    p, pp = _f64(params)
    x, xp = _f64(x)
    y = np.empty(len(x), dtype=np.float64)
    lib().ora_mlp_forward(pp, int(H), xp, len(x), y.ctypes.data_as(f64p))
    return y


def omp_max_threads():
    return int(lib().ora_omp_max_threads())
