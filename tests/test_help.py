"""Tail helpers (dds_gpu_amd/csrc/ddx_help.hpp): many solver lanes stepped round-robin on the
CPU with the same machine and helper code as the GPU kernel; the tables must equal the plain
solver's whatever the split threshold, and the helpers must shorten the batch's tail."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "support", "help_sim.cpp")


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("help") / "libhelp_sim.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                           "-I" + os.path.join(ROOT, "include"), SRC, "-o", out])
    lib = ctypes.CDLL(out)
    P = ctypes.c_void_p
    lib.ddx_help_sim.argtypes = [P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_uint64, ctypes.c_int64, P, P]

    def run(hands, lanes, jpd, spawn, tt_log2=9, pool=1 << 16):
        n = len(hands)
        out = np.zeros((n, 20), np.uint8)
        st = np.zeros(8, np.int64)
        rc = lib.ddx_help_sim(hands.ctypes.data, n, lanes, tt_log2, jpd, spawn, pool, out.ctypes.data,
                              st.ctypes.data)
        assert rc == 0
        return out, {"rounds": int(st[0]), "records": int(st[1]), "taken": int(st[2]), "nodes": int(st[3]),
                     "answers": int(st[4]), "drained_at": int(st[6])}
    return run


@pytest.mark.parametrize("K,n,lanes,jpd", [(7, 200, 16, 5), (9, 80, 32, 20), (10, 40, 64, 5)])
def test_helpers_exact(ddx, sim, K, n, lanes, jpd):
    hands = ddx.random_deals(11 + K, n, K)
    ref, _ = ddx.cpu_solve(hands, threads=8, tt_log2=12)
    for spawn in (1 << 40, 2000, 100):
        out, st = sim(hands, lanes, jpd, spawn)
        assert np.array_equal(out, ref), (spawn, st)
        assert st["answers"] == 20 * n
        if spawn < 1 << 40:
            assert st["records"] > 0 and st["taken"] <= st["records"]


def test_helpers_exact_with_a_tiny_pool(ddx, sim):
    """Pool overflow: spawns that do not fit are dropped, answers stay exact."""
    hands = ddx.random_deals(19, 60, 9)
    ref, _ = ddx.cpu_solve(hands, threads=8, tt_log2=12)
    out, st = sim(hands, 32, 5, 100, pool=40)
    assert np.array_equal(out, ref) and st["records"] > 40


def test_helpers_shorten_the_tail(ddx, sim):
    """One heavy 11-card deal, one job per lead, 128 lanes: the last answer comes in fewer
    rounds with helpers than without."""
    hands = ddx.random_deals(99, 300, 11)[64:65]
    ref, _ = ddx.cpu_solve(hands, tt_log2=9)
    out0, st0 = sim(hands, 128, 20, 1 << 40)
    out1, st1 = sim(hands, 128, 20, 20000, pool=1 << 18)
    assert np.array_equal(out0, ref) and np.array_equal(out1, ref)
    print("rounds without helpers %d, with %d (%.2fx); nodes %d -> %d" %
          (st0["rounds"], st1["rounds"], st0["rounds"] / st1["rounds"], st0["nodes"], st1["nodes"]))
    assert st1["rounds"] < 0.6 * st0["rounds"]
