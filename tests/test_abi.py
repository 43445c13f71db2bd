"""CPU tests of the C ABI boundary: libmgv.so loads, exports every function
declared in include/mgv.h, and its host-only logic and error paths behave
(no kernel is launched here)."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

import mgv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header: str):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mgv_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = mgv.load()
    names = declared_functions(os.path.join(ROOT, "include", "mgv.h"))
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each
    assert not [n for n in names if n not in mgv.SIGNATURES], [n for n in names if n not in mgv.SIGNATURES]


def test_struct_sizes_match_header():
    assert C.sizeof(mgv.Result) == 64
    assert C.sizeof(mgv.Launch) == 16
    assert mgv.load().mgv_abi_version() == 1


def test_ht_plan():
    nb, nbytes = mgv.ht_plan(1 << 24, 0.5)
    assert nb == 1 << 22 and nbytes == 256 << 20          # C2: 2^25 slots x 8 B
    nb, _ = mgv.ht_plan(1 << 22, 0.5)
    assert nb == 1 << 20
    nb, _ = mgv.ht_plan(1000, 0.75)
    assert nb * 8 >= 1000 / 0.75 and nb * 8 > 1000
    nb, _ = mgv.ht_plan(8, 0.9)
    assert nb * 8 > 8                                     # at least one empty slot
    nb, _ = mgv.ht_plan(0, 0.5)
    assert nb == 1
    with pytest.raises(mgv.MgvError) as e:
        mgv.ht_plan(10, 0.0)
    assert e.value.code == mgv.E_INVALID
    with pytest.raises(mgv.MgvError):
        mgv.ht_plan(10, 0.95)


def test_calls_fail_cleanly_without_init_or_gpu():
    lib = mgv.load()
    p = C.c_void_p()
    rc = lib.mgv_malloc(C.byref(p), 16)
    # without a GPU (this container) or before mgv_init: a status, never exit()
    assert rc in (0, mgv.E_NOT_INIT)
    if rc == 0:
        lib.mgv_free(p)
    assert isinstance(lib.mgv_last_error_string(), bytes)


def test_ops_library_exports_every_declared_symbol():
    lib = mgv.load_ops()
    names = declared_functions(os.path.join(ROOT, "include", "mgv_ops.h"))
    assert len(names) >= 7
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert not [n for n in names if n not in mgv.OPS_SIGNATURES]


def test_ops_calls_fail_cleanly_before_init():
    mgv.load_ops()
    s = C.c_int64()
    with pytest.raises(mgv.MgvError) as e:
        mgv.ops_call("mgv_ops_sum_select", None, None, 10, 5, 4096, 0, 1, 256, None, C.byref(s))
    assert e.value.code == mgv.E_NOT_INIT
    assert "mgv_ops_init" in mgv.load_ops().mgv_ops_last_error().decode()


def test_struct_layouts_match_the_headers(tmp_path):
    """sizeof of every struct the binding mirrors, as gcc lays out
    include/mgv.h and include/mgv_ops.h, equals the ctypes layout."""
    import subprocess
    structs = {"mgv_serve_request": mgv.ServeRequest, "mgv_serve_result": mgv.ServeResult,
               "mgv_serve_config": mgv.ServeConfig, "mgv_serve_stats": mgv.ServeStats, "mgv_query": mgv.Query,
               "mgv_work_item": mgv.WorkItem, "mgv_result": mgv.Result, "mgv_launch": mgv.Launch,
               "mgv_ring_entry": mgv.RingEntry, "mgv_device_props": mgv.DeviceProps}
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "mgv_ops.h"\nint main(void) {\n' +
                   "".join(f'  printf("%zu\\n", sizeof({k}));\n' for k in structs) + "  return 0;\n}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [C.sizeof(t) for t in structs.values()], dict(zip(structs, got))
