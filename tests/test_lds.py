"""The LDS tag-line cache (dds_gpu_amd/csrc/ddx_lds.hpp), its logic run on the host: a lane
that reads and writes its tag lines through the cache must produce the plain solver's tables
with exactly the same nodes, and the cache must actually serve lines."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "support", "lds_cpu.cpp")


@pytest.fixture(scope="module")
def lds(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("lds") / "liblds.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                           "-I" + os.path.join(ROOT, "include"), SRC, "-o", out])
    lib = ctypes.CDLL(out)
    P = ctypes.c_void_p
    lib.ddx_lds_run.argtypes = [P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P]
    lib.ddx_lds_run.restype = ctypes.c_uint64

    def run(hands, tt_log2, jpd, cache):
        n = len(hands)
        out = np.zeros((n, 20), np.uint8)
        hits = ctypes.c_uint64()
        nodes = lib.ddx_lds_run(hands.ctypes.data, n, tt_log2, jpd, int(cache), out.ctypes.data, ctypes.byref(hits))
        return out, nodes, hits.value
    return run


@pytest.mark.parametrize("K,n,tt_log2,jpd", [(8, 300, 10, 5), (10, 40, 9, 20), (13, 2, 11, 5)])
def test_cache_is_transparent(ddx, lds, K, n, tt_log2, jpd):
    hands = ddx.random_deals(5 + K, n, K)
    out0, nodes0, _ = lds(hands, tt_log2, jpd, False)
    out1, nodes1, hits = lds(hands, tt_log2, jpd, True)
    assert np.array_equal(out0, out1) and nodes0 == nodes1
    assert hits > 0.05 * nodes0  # a few tag lines per node; a real share comes from the cache
    ref, _ = ddx.cpu_solve(hands, threads=8, tt_log2=12)
    assert np.array_equal(out0, ref)


def test_cache_across_generation_wraps(ddx, lds):
    """More than 63 jobs on one lane wrap the TT generation (tags cleared): the cached lines
    must be dropped with them."""
    hands = ddx.random_deals(77, 40, 5)
    out0, nodes0, _ = lds(hands, 9, 20, False)
    out1, nodes1, _ = lds(hands, 9, 20, True)
    assert np.array_equal(out0, out1) and nodes0 == nodes1
