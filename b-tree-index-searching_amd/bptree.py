"""Python host side of libbptree_amd: ctypes bindings over the C ABI in
include/bptree.h, plus thin classes that mirror the reference's call surface
(BPTree.h:26-55): build a tree from sorted keys, point search (plain /
grouped / learned), range search (plain / 2D-grouping / Z-order), rectangle
search over Morton keys, multi-tree batches and the serving loop.

Device memory is owned through the library's own allocator (no PyTorch): a
``DeviceArray`` wraps one allocation. The library is loaded from this
directory; if it is missing the import fails loudly — there is no CPU
fallback on this path.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbptree_amd.so")

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
vp = C.c_void_p

BPT_OK = 0
STATUS = {0: "OK", 1: "INVALID", 2: "UNSORTED", 3: "NOMEM", 4: "HIP", 5: "CAPACITY",
          6: "UNSUPPORTED", 7: "NOT_FOUND"}
PLAIN, GROUPED, LEARNED = 0, 1, 2
GROUP_NONE, GROUP_GRID2D, GROUP_ZORDER = 0, 1, 2
SENTINEL = {4: 0xFFFFFFFF, 8: 0xFFFFFFFFFFFFFFFF}
MAX_LEVELS = 16


class BptError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"bpt status {STATUS.get(status, status)}: {msg}")
        self.status = status


class BuildDesc(C.Structure):
    _fields_ = [("key_bytes", C.c_uint32), ("val_bytes", C.c_uint32), ("fanout", C.c_uint32),
                ("leaf_cap", C.c_uint32), ("check_sorted", C.c_uint32), ("lds_bytes", C.c_uint32)]


class TreeInfo(C.Structure):
    _fields_ = [("n", C.c_uint64), ("key_bytes", C.c_uint32), ("val_bytes", C.c_uint32),
                ("fanout", C.c_uint32), ("leaf_cap", C.c_uint32), ("height", C.c_uint32),
                ("lds_levels", C.c_uint32), ("lds_bytes", C.c_uint64), ("leaf_stride", C.c_uint64),
                ("leaf_bytes", C.c_uint64), ("inner_bytes", C.c_uint64),
                ("level_nodes", C.c_uint64 * MAX_LEVELS)]


# (name, restype, argtypes) of every exported symbol; include/bptree.h order.
SIGNATURES = [
    ("bpt_last_error", C.c_char_p, []),
    ("bpt_version", C.c_char_p, []),
    ("bpt_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("bpt_device_info", C.c_int, [C.c_int, C.POINTER(C.c_int), u64p, C.c_char_p, C.c_int]),
    ("bpt_device_alloc", C.c_int, [C.c_uint64, C.POINTER(vp)]),
    ("bpt_device_free", C.c_int, [vp]),
    ("bpt_host_alloc", C.c_int, [C.c_uint64, C.POINTER(vp)]),
    ("bpt_host_free", C.c_int, [vp]),
    ("bpt_memset", C.c_int, [vp, C.c_int, C.c_uint64, vp]),
    ("bpt_copy_h2d", C.c_int, [vp, vp, C.c_uint64, vp]),
    ("bpt_copy_d2h", C.c_int, [vp, vp, C.c_uint64, vp]),
    ("bpt_copy_d2d", C.c_int, [vp, vp, C.c_uint64, vp]),
    ("bpt_stream_create", C.c_int, [C.POINTER(vp)]),
    ("bpt_stream_destroy", C.c_int, [vp]),
    ("bpt_stream_sync", C.c_int, [vp]),
    ("bpt_device_sync", C.c_int, []),
    ("bpt_timer_create", C.c_int, [C.POINTER(vp)]),
    ("bpt_timer_destroy", C.c_int, [vp]),
    ("bpt_timer_start", C.c_int, [vp, vp]),
    ("bpt_timer_stop", C.c_int, [vp, vp]),
    ("bpt_timer_ms", C.c_int, [vp, C.POINTER(C.c_float)]),
    ("bpt_tree_build", C.c_int, [C.POINTER(BuildDesc), vp, vp, C.c_uint64, C.POINTER(vp), vp]),
    ("bpt_tree_free", C.c_int, [vp]),
    ("bpt_tree_get_info", C.c_int, [vp, C.POINTER(TreeInfo)]),
    ("bpt_tree_get_layout", C.c_int, [vp, C.POINTER(vp), C.POINTER(vp)]),
    ("bpt_tree_set_launch", C.c_int, [vp, C.c_uint32, C.c_uint32]),
    ("bpt_tree_upsert", C.c_int, [vp, vp, vp, C.c_uint64, vp]),
    ("bpt_tree_erase", C.c_int, [vp, vp, C.c_uint64, vp]),
    ("bpt_tree_set_option", C.c_int, [vp, C.c_char_p, C.c_int64]),
    ("bpt_point_search", C.c_int, [vp, vp, C.c_uint64, vp, C.c_int, vp]),
    ("bpt_point_search_ws_size", C.c_uint64, [vp, C.c_uint64, C.c_int]),
    ("bpt_point_search_ws", C.c_int, [vp, vp, C.c_uint64, vp, C.c_int, vp, C.c_uint64, vp]),
    ("bpt_point_search_noalloc", C.c_int, [vp, vp, C.c_uint64, vp, C.c_int, vp]),
    ("bpt_tree_set_group_bits", C.c_int, [vp, C.c_uint32]),
    ("bpt_lower_bound", C.c_int, [vp, vp, C.c_uint64, vp, vp]),
    ("bpt_range_count", C.c_int, [vp, vp, vp, C.c_uint64, vp, vp, vp]),
    ("bpt_range_ids", C.c_int, [vp, vp, vp, vp, C.c_uint64, vp, vp]),
    ("bpt_exclusive_scan_u64", C.c_int, [vp, C.c_uint64, vp, u64p, vp]),
    ("bpt_range_count_grouped", C.c_int, [vp, vp, vp, C.c_uint64, C.c_int, vp, vp, vp]),
    ("bpt_radix_group", C.c_int, [vp, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, vp, vp, vp]),
    ("bpt_sort_perm", C.c_int, [vp, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, vp, vp]),
    ("bpt_gather", C.c_int, [vp, C.c_uint32, vp, C.c_uint64, vp, vp]),
    ("bpt_sort_pairs", C.c_int, [vp, vp, C.c_uint32, C.c_uint32, C.c_uint64, vp]),
    ("bpt_morton_encode", C.c_int, [vp, vp, C.c_uint64, vp, vp]),
    ("bpt_zorder2d", C.c_int, [vp, vp, C.c_uint64, vp, vp]),
    ("bpt_group2d_keys", C.c_int, [vp, vp, C.c_uint64, vp, vp]),
    ("bpt_learned_build", C.c_int, [vp, vp, vp, vp]),
    ("bpt_learned_params", C.c_int, [vp, C.POINTER(C.c_float), C.c_uint32]),
    ("bpt_mlp_loss_grad", C.c_int, [C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), C.c_uint64, C.POINTER(C.c_float),
                                    C.POINTER(C.c_float)]),
    ("bpt_server_create", C.c_int, [vp, C.c_uint32, C.c_uint64, C.c_int, C.POINTER(vp)]),
    ("bpt_server_submit", C.c_int, [vp, vp, u64p]),
    ("bpt_server_wait", C.c_int, [vp, C.c_uint64, C.POINTER(C.c_double)]),
    ("bpt_server_get_stats", C.c_int, [vp, vp]),
    ("bpt_server_reset_stats", C.c_int, [vp]),
    ("bpt_server_destroy", C.c_int, [vp]),
    ("bpt_server_replay", C.c_int, [vp, vp, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("bpt_forest_build", C.c_int, [vp, C.c_uint32, vp, vp, u64p, C.POINTER(vp), vp]),
    ("bpt_forest_free", C.c_int, [vp]),
    ("bpt_forest_info", C.c_int, [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), u64p, u64p]),
    ("bpt_batch_prepare", C.c_int, [vp, C.c_uint32, C.c_uint32, vp, u64p, vp, vp, vp, u64p, vp, vp,
                                    C.POINTER(vp)]),
    ("bpt_batch_run", C.c_int, [vp, C.c_int, vp]),
    ("bpt_batch_free", C.c_int, [vp]),
    ("bpt_batch_info", C.c_int, [vp, C.POINTER(C.c_uint32), u64p, u64p]),
    ("bpt_forest_autotune", C.c_int, [C.c_uint32, vp, vp, u64p, vp, u64p, vp, vp, vp, u64p, vp, vp, C.c_int,
                                      vp, C.c_char_p, C.c_uint64]),
    ("bpt_synth_forest", C.c_int, [C.c_uint64, C.c_uint32, u64p, C.c_uint32, C.c_uint32, u64p, C.c_uint32,
                                   C.c_uint32, vp, vp, vp, vp, vp, vp]),
    ("bpt_rect_count", C.c_int, [vp, vp, C.c_uint64, C.c_int, vp, vp]),
    ("bpt_rect_ids", C.c_int, [vp, vp, C.c_uint64, C.c_int, vp, vp, vp]),
]

_lib = None
_exiting = False


def _mark_exit():
    global _exiting
    _exiting = True


atexit.register(_mark_exit)


def lib():
    """Loads libbptree_amd.so (built by build.py); raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                              "(there is no CPU fallback for the HIP path)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != BPT_OK:
        raise BptError(status, lib().bpt_last_error().decode(errors="replace"))


def device_count():
    n = C.c_int(0)
    check(lib().bpt_device_count(C.byref(n)))
    return n.value


def device_info(dev=0):
    cu = C.c_int(0)
    mem = C.c_uint64(0)
    name = C.create_string_buffer(256)
    check(lib().bpt_device_info(dev, C.byref(cu), C.byref(mem), name, 256))
    return {"cus": cu.value, "hbm_bytes": mem.value, "name": name.value.decode()}


def sync():
    check(lib().bpt_device_sync())


class DeviceArray:
    """One device allocation holding a 1-D array of a numpy dtype."""

    def __init__(self, n, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = vp()
        check(lib().bpt_device_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value or 0

    @classmethod
    def from_host(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            check(lib().bpt_copy_h2d(d.ptr, a.ctypes.data, d.nbytes, stream))
        return d

    def to_host(self, stream=None, count=None, offset=0):
        """Elements [offset, offset + count) (all by default) as a numpy array."""
        offset = int(offset)
        n = self.n - offset if count is None else int(count)
        if offset < 0 or n < 0 or offset + n > self.n:
            raise ValueError(f"range [{offset}, {offset + n}) outside a device array of {self.n}")
        out = np.empty(n, dtype=self.dtype)
        if n:
            check(lib().bpt_copy_d2h(out.ctypes.data, self.ptr + offset * self.dtype.itemsize,
                                     n * self.dtype.itemsize, stream))
        return out

    def fill_bytes(self, value, stream=None):
        check(lib().bpt_memset(self.ptr, value, self.nbytes, stream))

    def free(self):
        if getattr(self, "ptr", 0) and not _exiting:
            lib().bpt_device_free(self.ptr)
        self.ptr = 0

    __del__ = free

    def __len__(self):
        return self.n


class HostArray:
    """A 1-D array in pinned host memory that the device addresses at the
    same pointer (bpt_host_alloc); `a` is its numpy view, valid until
    free() (a view kept past it points at released memory). It can stand
    in for a DeviceArray in BPT_REQ_DEVICE serving requests and in
    point_search_dev."""

    def __init__(self, n, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = vp()
        check(lib().bpt_host_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value or 0
        buf = (C.c_char * self.nbytes).from_address(self.ptr) if self.nbytes else bytearray(0)
        self.a = np.frombuffer(buf, dtype=self.dtype, count=self.n)

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        h = cls(a.size, a.dtype)
        h.a[:] = a
        return h

    def free(self):
        self.a = None
        if getattr(self, "ptr", 0) and not _exiting:
            lib().bpt_host_free(self.ptr)
        self.ptr = 0

    __del__ = free

    def __len__(self):
        return self.n


class Timer:
    """HIP events recorded on the stream the measured kernels run on."""

    def __init__(self):
        p = vp()
        check(lib().bpt_timer_create(C.byref(p)))
        self.h = p.value

    def start(self, stream=None):
        check(lib().bpt_timer_start(self.h, stream))

    def stop(self, stream=None):
        check(lib().bpt_timer_stop(self.h, stream))

    def ms(self):
        v = C.c_float(0)
        check(lib().bpt_timer_ms(self.h, C.byref(v)))
        return float(v.value)

    def close(self):
        if getattr(self, "h", None) and not _exiting:
            lib().bpt_timer_destroy(self.h)
        self.h = None

    __del__ = close


class Stream:
    def __init__(self):
        p = vp()
        check(lib().bpt_stream_create(C.byref(p)))
        self.h = p.value

    def sync(self):
        check(lib().bpt_stream_sync(self.h))

    def close(self):
        if getattr(self, "h", None) and not _exiting:
            lib().bpt_stream_destroy(self.h)
        self.h = None

    __del__ = close


def _dev(a, dtype):
    """Returns (DeviceArray, owned) for a host array or DeviceArray."""
    if isinstance(a, DeviceArray):
        if a.dtype != np.dtype(dtype):
            raise TypeError(f"device array dtype {a.dtype} != {np.dtype(dtype)}")
        return a, False
    return DeviceArray.from_host(np.asarray(a, dtype=dtype)), True


_KDT = {4: np.uint32, 8: np.uint64}


MLP_HIDDEN = 32
MLP_PARAMS = 3 * MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + 1


class LearnedConfig(C.Structure):
    _fields_ = [("leaf_models", C.c_uint32), ("max_window", C.c_uint32), ("train_samples", C.c_uint32),
                ("train_steps", C.c_uint32), ("lr", C.c_float), ("seed", C.c_uint32), ("root", C.c_uint32)]


class LearnedStats(C.Structure):
    _fields_ = [("initial_loss", C.c_float), ("final_loss", C.c_float), ("train_ms", C.c_double),
                ("fit_ms", C.c_double), ("breakpoints", C.c_uint32), ("models_used", C.c_uint32),
                ("fallback_models", C.c_uint32), ("max_window_seen", C.c_uint32),
                ("fallback_keys", C.c_uint64), ("mean_window", C.c_double)]


def mlp_loss_grad(params, x, t):
    """Loss and gradient of the 1-32-32-1 MLP from the training kernels."""
    fp = C.POINTER(C.c_float)
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    g = np.empty(MLP_PARAMS, dtype=np.float32)
    loss = C.c_float(0)
    check(lib().bpt_mlp_loss_grad(p.ctypes.data_as(fp), MLP_HIDDEN, x.ctypes.data_as(fp), t.ctypes.data_as(fp),
                                  len(x), g.ctypes.data_as(fp), C.byref(loss)))
    return float(loss.value), g


class Tree:
    """A bulk-built B+tree resident on the device (genBPlusTree, BPTree.cu:127-306)."""

    def __init__(self, keys, vals, key_bytes=None, val_bytes=None, fanout=0, leaf_cap=0,
                 check_sorted=True, lds_bytes=0, stream=None):
        if key_bytes is None:
            key_bytes = np.dtype(keys.dtype).itemsize
        if val_bytes is None:
            val_bytes = np.dtype(vals.dtype).itemsize
        self.key_bytes, self.val_bytes = int(key_bytes), int(val_bytes)
        self.kdt, self.vdt = _KDT[self.key_bytes], _KDT[self.val_bytes]
        dk, ok = _dev(keys, self.kdt)
        dv, ov = _dev(vals, self.vdt)
        d = BuildDesc(self.key_bytes, self.val_bytes, fanout, leaf_cap, 1 if check_sorted else 0,
                      lds_bytes)
        h = vp()
        try:
            check(lib().bpt_tree_build(C.byref(d), dk.ptr, dv.ptr, dk.n, C.byref(h), stream))
        finally:
            if ok:
                dk.free()
            if ov:
                dv.free()
        self.h = h.value
        self.n = dk.n
        self.sentinel = SENTINEL[self.val_bytes]

    def close(self):
        if getattr(self, "h", None) and not _exiting:
            lib().bpt_tree_free(self.h)
        self.h = None

    __del__ = close

    def info(self):
        i = TreeInfo()
        check(lib().bpt_tree_get_info(self.h, C.byref(i)))
        d = {k: getattr(i, k) for k, _ in TreeInfo._fields_ if k != "level_nodes"}
        d["level_nodes"] = [int(i.level_nodes[k]) for k in range(i.height + 1)]
        return d

    def set_launch(self, block=0, blocks_per_cu=0):
        check(lib().bpt_tree_set_launch(self.h, block, blocks_per_cu))

    def set_option(self, key, value):
        check(lib().bpt_tree_set_option(self.h, key.encode(), int(value)))

    def set_group_bits(self, bits):
        check(lib().bpt_tree_set_group_bits(self.h, bits))

    # ---- point search ------------------------------------------------------
    def point_search_dev(self, dq, dout, mode=PLAIN, stream=None):
        check(lib().bpt_point_search(self.h, dq.ptr, dq.n, dout.ptr, mode, stream))

    def point_search_ws_size(self, nq, mode):
        return int(lib().bpt_point_search_ws_size(self.h, nq, mode))

    def point_search_noalloc_dev(self, dq, dout, mode=PLAIN, stream=None):
        """bpt_point_search_noalloc: the library's static scratch, no allocation."""
        check(lib().bpt_point_search_noalloc(self.h, dq.ptr, dq.n, dout.ptr, mode, stream))

    def point_search(self, q, mode=PLAIN, stream=None):
        dq, own = _dev(q, self.kdt)
        out = DeviceArray(dq.n, self.vdt)
        self.point_search_dev(dq, out, mode, stream)
        r = out.to_host(stream)
        out.free()
        if own:
            dq.free()
        return r

    def upsert(self, keys, vals, stream=None):
        """Each key takes its value or is added (last of equal keys wins)."""
        dk, ok = _dev(keys, self.kdt)
        dv, ov = _dev(vals, self.vdt)
        try:
            check(lib().bpt_tree_upsert(self.h, dk.ptr, dv.ptr, dk.n, stream))
        finally:
            if ok:
                dk.free()
            if ov:
                dv.free()

    def erase(self, keys, stream=None):
        """Removes every slot whose key is one of `keys`."""
        dk, ok = _dev(keys, self.kdt)
        try:
            check(lib().bpt_tree_erase(self.h, dk.ptr, dk.n, stream))
        finally:
            if ok:
                dk.free()

    def lower_bound(self, q, stream=None):
        dq, own = _dev(q, self.kdt)
        pos = DeviceArray(dq.n, np.uint64)
        check(lib().bpt_lower_bound(self.h, dq.ptr, dq.n, pos.ptr, stream))
        r = pos.to_host(stream)
        if own:
            dq.free()
        return r

    # ---- learned position model --------------------------------------------
    def learned_build(self, leaf_models=65536, max_window=256, train_samples=1 << 20, train_steps=2000,
                      lr=3e-3, seed=1, root=0, stream=None):
        cfg = LearnedConfig(leaf_models, max_window, train_samples, train_steps, lr, seed, root)
        st = LearnedStats()
        check(lib().bpt_learned_build(self.h, C.byref(cfg), C.byref(st), stream))
        return {k: getattr(st, k) for k, _ in LearnedStats._fields_}

    def learned_params(self):
        p = np.empty(MLP_PARAMS, dtype=np.float32)
        check(lib().bpt_learned_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), MLP_PARAMS))
        return p

    # ---- range search ------------------------------------------------------
    def range_count_dev(self, dlo, dhi, dfirst, dcount, group_mode=GROUP_NONE, stream=None):
        fp = dfirst.ptr if dfirst is not None else None
        if group_mode == GROUP_NONE:
            check(lib().bpt_range_count(self.h, dlo.ptr, dhi.ptr, dlo.n, fp, dcount.ptr, stream))
        else:
            check(lib().bpt_range_count_grouped(self.h, dlo.ptr, dhi.ptr, dlo.n, group_mode, fp,
                                                dcount.ptr, stream))

    def range_search(self, lo, hi, group_mode=GROUP_NONE, with_ids=True, stream=None):
        """Returns (first, count, offsets, ids) in the caller's query order."""
        dlo, o1 = _dev(lo, self.kdt)
        dhi, o2 = _dev(hi, self.kdt)
        n = dlo.n
        dfirst = DeviceArray(n, np.uint64)
        dcount = DeviceArray(n, np.uint64)
        self.range_count_dev(dlo, dhi, dfirst, dcount, group_mode, stream)
        first, count = dfirst.to_host(stream), dcount.to_host(stream)
        offsets = ids = None
        if with_ids:
            doff = DeviceArray(n, np.uint64)
            total = C.c_uint64(0)
            check(lib().bpt_exclusive_scan_u64(dcount.ptr, n, doff.ptr, C.byref(total), stream))
            dids = DeviceArray(total.value, self.vdt)
            check(lib().bpt_range_ids(self.h, dfirst.ptr, dcount.ptr, doff.ptr, n, dids.ptr, stream))
            offsets, ids = doff.to_host(stream), dids.to_host(stream)
        for d, own in ((dlo, o1), (dhi, o2)):
            if own:
                d.free()
        return first, count, offsets, ids

    # ---- rectangles --------------------------------------------------------
    def rect_search(self, rects, group_mode=GROUP_NONE, with_ids=True, stream=None):
        """rects: (n, 4) uint32 [x0, y0, x1, y1]; returns (counts, offsets, ids)."""
        dr, own = _dev(np.asarray(rects, dtype=np.uint32).reshape(-1), np.uint32)
        nr = dr.n // 4
        dcount = DeviceArray(nr, np.uint64)
        check(lib().bpt_rect_count(self.h, dr.ptr, nr, group_mode, dcount.ptr, stream))
        counts = dcount.to_host(stream)
        offsets = ids = None
        if with_ids:
            doff = DeviceArray(nr, np.uint64)
            total = C.c_uint64(0)
            check(lib().bpt_exclusive_scan_u64(dcount.ptr, nr, doff.ptr, C.byref(total), stream))
            dids = DeviceArray(total.value, np.uint32)
            if total.value:  # no ids, no buffer to write them to
                check(lib().bpt_rect_ids(self.h, dr.ptr, nr, group_mode, doff.ptr, dids.ptr, stream))
            offsets, ids = doff.to_host(stream), dids.to_host(stream)
        if own:
            dr.free()
        return counts, offsets, ids


def exclusive_scan(counts):
    d = DeviceArray.from_host(np.asarray(counts, dtype=np.uint64))
    out = DeviceArray(d.n, np.uint64)
    total = C.c_uint64(0)
    check(lib().bpt_exclusive_scan_u64(d.ptr, d.n, out.ptr, C.byref(total), None))
    return out.to_host(), int(total.value)


def radix_group(keys, shift, bits):
    keys = np.ascontiguousarray(keys)
    kb = keys.dtype.itemsize
    d = DeviceArray.from_host(keys)
    perm = DeviceArray(d.n, np.uint32)
    starts = DeviceArray((1 << bits) + 1, np.uint64)
    check(lib().bpt_radix_group(d.ptr, kb, d.n, shift, bits, perm.ptr, starts.ptr, None))
    return perm.to_host(), starts.to_host()


def morton_encode(x, y):
    dx = DeviceArray.from_host(np.asarray(x, dtype=np.uint32))
    dy = DeviceArray.from_host(np.asarray(y, dtype=np.uint32))
    dz = DeviceArray(dx.n, np.uint64)
    check(lib().bpt_morton_encode(dx.ptr, dy.ptr, dx.n, dz.ptr, None))
    return dz.to_host()


def zorder2d(a, b):
    da = DeviceArray.from_host(np.asarray(a, dtype=np.uint64))
    db = DeviceArray.from_host(np.asarray(b, dtype=np.uint64))
    out = DeviceArray(da.n, np.uint64)
    check(lib().bpt_zorder2d(da.ptr, db.ptr, da.n, out.ptr, None))
    return out.to_host()


def group2d_keys(a, b):
    da = DeviceArray.from_host(np.asarray(a, dtype=np.uint64))
    db = DeviceArray.from_host(np.asarray(b, dtype=np.uint64))
    out = DeviceArray(da.n, np.uint32)
    check(lib().bpt_group2d_keys(da.ptr, db.ptr, da.n, out.ptr, None))
    return out.to_host()


def sort_pairs(keys, vals=None):
    """Sorts (key, value) pairs by key on the device; returns host arrays."""
    keys = np.ascontiguousarray(keys)
    dk = DeviceArray.from_host(keys)
    dv = DeviceArray.from_host(np.ascontiguousarray(vals)) if vals is not None else None
    vb = dv.dtype.itemsize if dv is not None else 4
    check(lib().bpt_sort_pairs(dk.ptr, dv.ptr if dv is not None else None, keys.dtype.itemsize, vb,
                               dk.n, None))
    return dk.to_host(), (dv.to_host() if dv is not None else None)


class BatchConfig(C.Structure):
    _fields_ = [("fanout", C.c_uint32), ("leaf_cap", C.c_uint32), ("lds_bytes", C.c_uint32),
                ("chunk", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("lanes", C.c_uint32)]


def _offp(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint64).ctypes.data_as(u64p)


class Forest:
    """Thousands of 8-byte-key trees in one device arena (bpt_forest_*)."""

    def __init__(self, dkeys, dvals, key_off, fanout=16, lds_bytes=32768, lanes=8, stream=None):
        self.key_off = np.ascontiguousarray(key_off, dtype=np.uint64)
        self.ntrees = len(self.key_off) - 1
        cfg = BatchConfig(fanout, 8, lds_bytes, 0, 0, lanes)
        h = vp()
        check(lib().bpt_forest_build(C.byref(cfg), self.ntrees, dkeys.ptr, dvals.ptr,
                                     _offp(self.key_off), C.byref(h), stream))
        self.h = h.value
        self.fanout = fanout

    def close(self):
        if getattr(self, "h", None) and not _exiting:
            lib().bpt_forest_free(self.h)
        self.h = None

    __del__ = close


class Batch:
    """A prepared persistent-launch batch over a forest (bpt_batch_*)."""

    def __init__(self, forest, dq=None, q_off=None, dout=None, dlo=None, dhi=None, r_off=None,
                 dfirst=None, dcount=None, chunk=0, blocks_per_cu=0):
        self._keep = (np.ascontiguousarray(q_off, dtype=np.uint64) if q_off is not None else None,
                      np.ascontiguousarray(r_off, dtype=np.uint64) if r_off is not None else None)
        h = vp()
        P = lambda d: d.ptr if d is not None else None  # noqa: E731
        check(lib().bpt_batch_prepare(forest.h, chunk, blocks_per_cu, P(dq), _offp(self._keep[0]), P(dout),
                                      P(dlo), P(dhi), _offp(self._keep[1]), P(dfirst), P(dcount), C.byref(h)))
        self.h = h.value
        self.forest = forest

    def run(self, graph=False, stream=None):
        check(lib().bpt_batch_run(self.h, 1 if graph else 0, stream))

    def info(self):
        ni, npt, nr = C.c_uint32(), C.c_uint64(), C.c_uint64()
        check(lib().bpt_batch_info(self.h, C.byref(ni), C.byref(npt), C.byref(nr)))
        return {"items": ni.value, "points": npt.value, "ranges": nr.value}

    def close(self):
        if getattr(self, "h", None) and not _exiting:
            lib().bpt_batch_free(self.h)
        self.h = None

    __del__ = close


def synth_forest(seed, key_off, qpk=4, hit_pct=80, r_off=None, w_lo_exp=4, w_hi_exp=16):
    """Device-generated forest workload; returns device arrays."""
    key_off = np.ascontiguousarray(key_off, dtype=np.uint64)
    nt = len(key_off) - 1
    total = int(key_off[-1])
    dk = DeviceArray(total, np.uint64)
    dv = DeviceArray(total, np.uint64)
    dq = DeviceArray(total * qpk, np.uint64) if qpk else None
    nr = int(r_off[-1]) if r_off is not None else 0
    dlo = DeviceArray(nr, np.uint64) if nr else None
    dhi = DeviceArray(nr, np.uint64) if nr else None
    check(lib().bpt_synth_forest(seed, nt, _offp(key_off), qpk, hit_pct, _offp(r_off), w_lo_exp, w_hi_exp,
                                 dk.ptr, dv.ptr, dq.ptr if dq else None, dlo.ptr if dlo else None,
                                 dhi.ptr if dhi else None, None))
    return dk, dv, dq, dlo, dhi


def forest_autotune(dk, dv, key_off, dq, q_off, dout, dlo=None, dhi=None, r_off=None, dfirst=None,
                    dcount=None, reps=2):
    best = BatchConfig()
    rep = C.create_string_buffer(1 << 16)
    key_off = np.ascontiguousarray(key_off, dtype=np.uint64)
    P = lambda d: d.ptr if d is not None else None  # noqa: E731
    check(lib().bpt_forest_autotune(len(key_off) - 1, dk.ptr, dv.ptr, _offp(key_off), P(dq), _offp(q_off),
                                    P(dout), P(dlo), P(dhi), _offp(r_off), P(dfirst), P(dcount), reps,
                                    C.byref(best), rep, len(rep)))
    import json as _json
    return ({k: getattr(best, k) for k, _ in BatchConfig._fields_}, _json.loads(rep.value.decode()))


REQ_POINT, REQ_RANGE = 0, 1
REQ_DEVICE = 1  # Request.flags: the request's buffers are device memory


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("flags", C.c_int32), ("n", C.c_uint64), ("q", vp), ("out", vp),
                ("lo", vp), ("hi", vp), ("first", vp), ("count", vp)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("lookups", C.c_uint64), ("max_inflight", C.c_uint32),
                ("pad_", C.c_uint32), ("peak_bytes", C.c_uint64), ("budget_bytes", C.c_uint64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double), ("mean_ms", C.c_double),
                ("elapsed_s", C.c_double)]


class Server:
    """Continuous-batching server over a resident tree (bpt_server_*).
    Host arrays passed to submit() must stay alive until wait()."""

    def __init__(self, tree, slots=4, mem_budget=1 << 30, mode=PLAIN):
        h = vp()
        check(lib().bpt_server_create(tree.h, slots, mem_budget, mode, C.byref(h)))
        self.h = h.value
        self.tree = tree
        self._live = {}

    @staticmethod
    def point_request(q, out):
        return Request(REQ_POINT, 0, len(q), q.ctypes.data, out.ctypes.data, None, None, None, None)

    @staticmethod
    def range_request(lo, hi, first, count):
        return Request(REQ_RANGE, 0, len(lo), None, None, lo.ctypes.data, hi.ctypes.data,
                       first.ctypes.data if first is not None else None, count.ctypes.data)

    @staticmethod
    def point_request_dev(dq, dout, offset=0, n=None):
        """A point request over DeviceArrays: queries dq[offset:offset+n],
        values written to dout[offset:offset+n] (the whole arrays by default)."""
        n = len(dq) - offset if n is None else n
        return Request(REQ_POINT, REQ_DEVICE, n, dq.ptr + offset * dq.dtype.itemsize,
                       dout.ptr + offset * dout.dtype.itemsize, None, None, None, None)

    @staticmethod
    def range_request_dev(dlo, dhi, dfirst, dcount, offset=0, n=None):
        """A range request over DeviceArrays (dfirst may be None), elements
        [offset, offset + n) of each (the whole arrays by default)."""
        n = len(dlo) - offset if n is None else n
        at = lambda a: a.ptr + offset * a.dtype.itemsize if a is not None else None  # noqa: E731
        return Request(REQ_RANGE, REQ_DEVICE, n, None, None, at(dlo), at(dhi), at(dfirst), at(dcount))

    def submit(self, req, keep=()):
        t = C.c_uint64(0)
        check(lib().bpt_server_submit(self.h, C.byref(req), C.byref(t)))
        self._live[t.value] = keep
        return t.value

    def wait(self, ticket):
        ms = C.c_double(0)
        st = lib().bpt_server_wait(self.h, ticket, C.byref(ms))
        self._live.pop(ticket, None)
        check(st)
        return ms.value

    @staticmethod
    def request_array(reqs):
        """The requests as one ctypes array, built once for repeated replays
        (building it costs about 0.7 us per request in Python)."""
        return (Request * len(reqs))(*reqs)

    def replay(self, reqs, arrival_s):
        """Open-loop replay of a list of Requests or a request_array()."""
        arr = reqs if isinstance(reqs, C.Array) else (Request * len(reqs))(*reqs)
        a = np.ascontiguousarray(arrival_s, dtype=np.float64)
        lat = np.zeros(len(reqs), dtype=np.float64)
        dp = C.POINTER(C.c_double)
        check(lib().bpt_server_replay(self.h, arr, len(reqs), a.ctypes.data_as(dp), lat.ctypes.data_as(dp)))
        return lat

    def stats(self):
        s = ServerStats()
        check(lib().bpt_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_ if k != "pad_"}

    def reset_stats(self):
        check(lib().bpt_server_reset_stats(self.h))

    def close(self):
        if getattr(self, "h", None) and not _exiting:
            lib().bpt_server_destroy(self.h)
        self.h = None

    __del__ = close
