"""The C-ABI library: loads, exports every function include/*.h declares, agrees with the
ctypes struct layouts, and rejects bad arguments before touching a device."""
import ctypes as C
import os
import re
import subprocess

import pytest

import accelgraph_amd as ag
from accelgraph_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in os.listdir(INCLUDE):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{)]*\)\s*;", text):
            name = m.group(1)
            if name not in ("if", "while", "return", "sizeof"):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    lib = ag.lib()
    names = declared_functions()
    assert len(names) > 40
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = sorted(names - exported)
    assert not missing, missing
    for n in names:
        assert n in _abi.SIGNATURES, f"{n} has no ctypes signature"
        getattr(lib, n)


def test_struct_layouts_match_headers(tmp_path):
    structs = {"struct PageRankStats": _abi.PageRankStats, "struct SPMVStats": _abi.SPMVStats,
               "struct Arguments": _abi.Arguments, "struct GraphCSR": _abi.GraphCSR,
               "struct EdgeList": _abi.EdgeList, "struct Vertex": _abi.Vertex,
               "ag_plan_params": _abi.PlanParams, "ag_graph_info": _abi.GraphInfo,
               "ag_pr_options": _abi.PrOptions, "ag_pr_result": _abi.PrResult}
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "accelgraph.h"',
           '#include "pageRank.h"', '#include "SPMV.h"', 'int main(void){']
    for cname in structs:
        src.append(f'printf("%zu\\n", sizeof({cname}));')
    src.append('printf("%zu %zu\\n", offsetof(struct Arguments, epsilon), '
               'offsetof(struct Arguments, afu_config));')
    src.append("return 0;}")
    c = tmp_path / "sizes.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", f"-I{INCLUDE}", str(c), "-o", str(exe)], check=True)
    vals = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True,
                          check=True).stdout.split()
    for (cname, py), v in zip(structs.items(), vals):
        assert C.sizeof(py) == int(v), cname
    assert int(vals[-2]) == _abi.Arguments.epsilon.offset
    assert int(vals[-1]) == _abi.Arguments.afu_config.offset


def test_status_strings_and_version():
    lib = ag.lib()
    assert lib.ag_abi_version() == 1
    assert lib.ag_status_string(0) == b"ok"
    assert lib.ag_status_string(5) == b"device wait timed out"


def test_null_arguments_rejected_without_device():
    lib = ag.lib()
    h = C.c_void_p()
    assert lib.ag_graph_upload(None, 1, 0, None, None, None, None, None, None,
                               C.byref(h)) == _abi.AG_ERR_INVALID
    assert lib.ag_device_sync(None) == _abi.AG_ERR_INVALID
    assert lib.ag_pr_begin(None, None, C.byref(h)) == _abi.AG_ERR_INVALID
    assert lib.ag_spmv_pull(None, None, None, 1, None) == _abi.AG_ERR_INVALID
    assert lib.ag_device_open(0, None) == _abi.AG_ERR_INVALID
    assert lib.ag_device_set_watchdog(None, 1, 1) == _abi.AG_ERR_INVALID


def test_no_device_reports_cleanly():
    if ag.Device.count() > 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    st = ag.lib().ag_device_open(0, C.byref(h))
    assert st == _abi.AG_ERR_NO_DEVICE and not h.value
    assert b"device" in ag.lib().ag_last_error()
    with pytest.raises(ag.AGError) as e:
        ag.Device(0)
    assert e.value.status == _abi.AG_ERR_NO_DEVICE


def test_missing_library_fails_loudly(tmp_path):
    """There is no fallback implementation: a library that is not built is an error, not a
    silent switch to another path."""
    with pytest.raises(RuntimeError, match="not built"):
        _abi.load(str(tmp_path / "libaccelgraph.so"))
