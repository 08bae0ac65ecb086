"""CPU checks of the C-ABI boundary: the library loads and exports every
function include/*.h declares (no device calls)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bpt_\w+)\s*\(", re.M)


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(DECL.findall(text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("bpt_tree_build", "bpt_point_search", "bpt_range_count", "bpt_range_ids",
                 "bpt_rect_count", "bpt_radix_group", "bpt_zorder2d", "bpt_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import bptree
    lib = ctypes.CDLL(bptree.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def compat_declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "compat", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names.update(re.findall(r"^\s*[A-Za-z_][\w\s\*&]*?\b([A-Za-z_]\w*)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_every_compat_function():
    """The reference-named C++ surface (include/compat/*.h) is defined in the
    library under its C++ names."""
    import shutil
    import subprocess
    import bptree
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-DC", "--defined-only", bptree.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split(" ", 2)[2].split("(")[0] for line in out.splitlines() if " T " in line}
    declared = compat_declared()
    assert "genBPlusTree" in declared and "SearchingBTreeGPURangeResults" in declared
    missing = [n for n in declared if n not in exported]
    assert not missing, missing


def test_python_bindings_cover_header():
    import bptree
    bound = {name for name, _, _ in bptree.SIGNATURES}
    missing = [n for n in declared_symbols() if n not in bound]
    assert not missing, missing


def test_binding_structs_match_header_layout(tmp_path):
    """Every struct the ctypes binding passes across the ABI has the size,
    field offsets and field sizes of its C declaration (gcc on the header)."""
    import subprocess
    import bptree
    structs = {"BuildDesc": "bpt_build_desc", "TreeInfo": "bpt_tree_info", "LearnedConfig": "bpt_learned_config",
               "LearnedStats": "bpt_learned_stats", "BatchConfig": "bpt_batch_config", "Request": "bpt_request",
               "ServerStats": "bpt_server_stats"}
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "bptree.h"', "int main(void) {"]
    want = []
    for py, c in structs.items():
        cls = getattr(bptree, py)
        src.append(f'printf("{py} %zu\\n", sizeof({c}));')
        want.append(f"{py} {ctypes.sizeof(cls)}")
        for f in cls._fields_:
            src.append(f'printf("{py}.{f[0]} %zu %zu\\n", offsetof({c}, {f[0]}), sizeof((({c} *)0)->{f[0]}));')
            fd = getattr(cls, f[0])
            want.append(f"{py}.{f[0]} {fd.offset} {fd.size}")
    src.append("return 0; }")
    (tmp_path / "layout.c").write_text("\n".join(src))
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(tmp_path / "layout.c"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    assert [g for g in got if g] == want


def test_errors_without_device_are_reported_not_fatal():
    import bptree
    L = bptree.lib()
    assert L.bpt_version().startswith(b"bptree_amd")
    # invalid arguments are rejected before any device work
    d = bptree.BuildDesc(4, 8, 32, 16, 0, 0)   # 4/8 widths unsupported
    h = ctypes.c_void_p()
    st = L.bpt_tree_build(ctypes.byref(d), None, None, 0, ctypes.byref(h), None)
    assert st == 1 and b"not supported" in L.bpt_last_error()
    d = bptree.BuildDesc(4, 4, 48, 16, 0, 0)
    assert L.bpt_tree_build(ctypes.byref(d), None, None, 0, ctypes.byref(h), None) == 1
    assert L.bpt_tree_free(None) == 0
    # pinned host memory: NULL output rejected, empty and NULL frees are no-ops
    p = ctypes.c_void_p(1)
    assert L.bpt_host_alloc(16, None) == 1 and b"NULL" in L.bpt_last_error()
    assert L.bpt_host_alloc(0, ctypes.byref(p)) == 0 and not p.value
    assert L.bpt_host_free(None) == 0


def test_bench_refuses_more_than_one_gpu():
    """The benchmark measures one GPU; asking for more exits with a message
    (and status 2) before touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout + r.stderr
