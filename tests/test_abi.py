"""The C-ABI library loads, exports every entry point include/cgp/*.h
declares, and reports errors the way the header documents (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import cgp_amd
from cgp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cgp", "cgp.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(cgp_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["cgp_graph_create_csr", "cgp_engine_create", "cgp_run_batch", "cgp_serve_submit",
                 "cgp_serve_step", "cgp_last_error", "cgp_engine_autotune", "cgp_load_gr"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_abi_version():
    assert _lib.lib.cgp_abi_version() == 1


def test_device_entry_points_fail_cleanly_without_gpu():
    if cgp_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = cgp_amd.HostCSR.rmat(6, 4, seed=1)
    with pytest.raises(cgp_amd.CgpError) as e:
        cgp_amd.Graph.from_host(h)
    assert e.value.status == cgp_amd.CGP_EDEVICE
    assert "HIP device" in e.value.message


def test_bad_csr_rejected_before_touching_device():
    with pytest.raises(cgp_amd.CgpError) as e:
        cgp_amd.Graph.from_csr([0, 2, 1], [0, 1])
    assert e.value.status == cgp_amd.CGP_EINVAL
    with pytest.raises(cgp_amd.CgpError) as e:
        cgp_amd.Graph.from_csr([0, 1, 2], [0, 5])
    assert e.value.status == cgp_amd.CGP_EINVAL


def test_last_error_is_thread_local():
    import threading
    msgs = {}

    def worker(key, scale):
        try:
            cgp_amd.HostCSR.rmat(scale, 4)
        except cgp_amd.CgpError as e:
            msgs[key] = e.message

    t = threading.Thread(target=worker, args=("t", 0))
    t.start()
    t.join()
    assert "scale" in msgs["t"]


def test_cpp_facade_compiles():
    src = os.path.join(ROOT, "include", "cgp", "cgp.hpp")
    if not os.path.exists(src):
        pytest.skip("no C++ facade")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", src], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr


STRUCTS = ["cgp_query", "cgp_engine_opts", "cgp_run_stats", "cgp_iteration_record", "cgp_serve_opts",
           "cgp_completion"]


def test_ctypes_structs_match_the_header(tmp_path):
    # Size and every field offset of the ctypes mirrors equal what a C
    # compiler lays out from include/cgp/cgp.h.
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "cgp/cgp.h"', "int main(void) {"]
    for s in STRUCTS:
        py = getattr(_lib, s)
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (s, s))
        for name, _ in py._fields_:
            cname = "pass" if name == "pass_" else name
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (s, name, s, cname))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True, timeout=120)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        s, name, val = line.split()
        got[(s, name)] = int(val)
    for s in STRUCTS:
        py = getattr(_lib, s)
        assert got[(s, "size")] == C.sizeof(py), s
        for name, _ in py._fields_:
            assert got[(s, name)] == getattr(py, name).offset, (s, name)


def test_engine_option_defaults():
    o = _lib.cgp_engine_opts()
    _lib.lib.cgp_engine_opts_default(C.byref(o))
    assert o.dense_threshold == 0  # automatic: resolved from sssp_delta at engine creation
    assert o.dense_unroll == 16
    assert o.sssp_delta == -1
    assert o.dense_wide_addr == 0
    assert o.concurrent_advance == 0
    assert o.block_size == 256


def test_renumber_choice_constants_and_validation():
    # The header's CGP_RENUMBER_* values, their ctypes mirror, the Python
    # argument mapping, and the rejection of an unknown choice, which comes
    # before any device call.
    text = open(HEADER).read()
    for name, v in [("CGP_RENUMBER_AUTO", 0), ("CGP_RENUMBER_ON", 1), ("CGP_RENUMBER_OFF", 2)]:
        assert re.search(r"#define %s %d\b" % (name, v), text), name
    assert (_lib.RENUMBER_AUTO, _lib.RENUMBER_ON, _lib.RENUMBER_OFF) == (0, 1, 2)
    assert [cgp_amd._renumber(r) for r in (None, True, False)] == [0, 1, 2]
    off = (C.c_int32 * 3)(0, 1, 2)
    idx = (C.c_int32 * 2)(1, 0)
    out = C.c_void_p()
    with pytest.raises(cgp_amd.CgpError) as e:
        _lib.check(_lib.lib.cgp_graph_create_csr_ex(2, 2, off, idx, None, 7, C.byref(out)))
    assert e.value.status == cgp_amd.CGP_EINVAL
    assert "renumber" in e.value.message
    assert not out.value
