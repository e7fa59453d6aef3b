"""The C ABI: libddx.so loads and exports every function include/ddx.h declares."""
import ctypes
import os
import re
import subprocess

import pytest
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ddx.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ddx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ["ddx_cpu_solve_strain", "ddx_cpu_solve_strain_guess", "ddx_cpu_ns_tricks",
                 "ddx_gpu_open", "ddx_gpu_solve", "ddx_gpu_request_stop", "ddx_last_error",
                 "ddx_stream_open", "ddx_stream_push", "ddx_stream_poll", "ddx_stream_close"]:
        assert must in names, must


@pytest.mark.parametrize("variant", ["", "sig", "shared"])
def test_library_exports_every_declared_symbol(ddx, variant):
    """Every build (exact TT entries, 12-byte signature entries, the shared TT) exports the whole
    ABI."""
    lib = ddx.lib(variant)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", ddx.LIB_PATHS[variant]], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\bT (ddx_\w+)", out))
    assert set(declared()) <= exported


def test_no_gpu_calls_without_a_device(ddx):
    """Without a GPU the GPU entry points fail cleanly (no crash, an error message)."""
    import torch
    if torch.cuda.is_available():
        return
    cfg = ddx.GpuConfig()
    ddx.lib().ddx_gpu_default_config(ctypes.byref(cfg))
    h = ctypes.c_void_p()
    rc = ddx.lib().ddx_gpu_open(ctypes.byref(cfg), ctypes.byref(h))
    assert rc != 0 and ddx.lib().ddx_last_error()


def test_headers_compile_on_their_own(tmp_path):
    """Each header of the header-only core, and the C ABI header (as C99 and as C++), compiles
    with nothing included before it."""
    inc = os.path.join(ROOT, "include")
    for h in sorted(os.listdir(os.path.join(inc, "ddx"))):
        src = tmp_path / ("use_%s.cpp" % h.replace(".", "_"))
        src.write_text('#include "ddx/%s"\n' % h)
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wno-unknown-pragmas", "-I" + inc, str(src)],
                           capture_output=True, text=True)
        assert r.returncode == 0, h + "\n" + r.stderr
    for lang, cc, ext in [("c99", "gcc", "c"), ("c++17", "g++", "cpp")]:
        src = tmp_path / ("use_ddx_h." + ext)
        src.write_text('#include "ddx.h"\n')
        r = subprocess.run([cc, "-std=" + lang, "-fsyntax-only", "-I" + inc, str(src)], capture_output=True, text=True)
        assert r.returncode == 0, lang + "\n" + r.stderr
