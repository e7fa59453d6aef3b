"""The TT-hit verifier (DDX_VERIFY, the reference's -DDD_VERIFY, src/dd.h:651-663): every TT hit
is re-checked by a search of the same position and target that uses no TT. On the solver's own
TT it must find no contradiction; on a TT whose entries were overwritten with false bounds it
must report them. Plus the dd_verify command line (the reference's tests/verify_one.cpp)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "support", "verify_corrupt.cpp")
TOOL = os.path.join(ROOT, "dds_gpu_amd", "bin", "dd_verify")


@pytest.fixture(scope="module")
def vlib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("verify") / "libverify.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                           "-I" + os.path.join(ROOT, "include"), SRC, "-o", out])
    lib = ctypes.CDLL(out)
    P = ctypes.c_void_p
    lib.ddx_verify_run.argtypes = [P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, P, P, P]
    lib.ddx_verify_run.restype = ctypes.c_int64

    def run(hands, tt_log2, corrupt):
        hands = np.ascontiguousarray(hands, np.uint64)
        out = np.zeros((len(hands), 20), np.uint8)
        ver, bad = ctypes.c_uint64(), ctypes.c_uint64()
        lib.ddx_verify_run(hands.ctypes.data, len(hands), tt_log2, int(corrupt), out.ctypes.data,
                           ctypes.byref(ver), ctypes.byref(bad))
        return out, ver.value, bad.value
    return run


@pytest.mark.parametrize("K,n", [(5, 200), (7, 60)])
def test_verifier_silent_on_solver_tt(ddx, oracle, vlib, K, n):
    hands = ddx.random_deals(31 + K, n, K)
    out, ver, bad = vlib(hands, 10, False)
    assert ver > 1000 and bad == 0
    assert np.array_equal(out, oracle.tables(hands))  # verification changes no answer


def test_verifier_reports_lying_entries(ddx, vlib):
    hands = ddx.random_deals(77, 60, 7)
    _, ver, bad = vlib(hands, 10, True)
    assert ver > 0 and bad > 0


def test_dd_verify_tool(ddx, oracle):
    r = subprocess.run([TOOL, "6", "100", "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "100 deals of 6 cards: verified" in r.stdout and ", 0 bad" in r.stdout
    h = ddx.random_deals(9, 1, 8)[0]
    want = oracle.tables(h[None])[0]
    r = subprocess.run([TOOL] + ["%x" % int(x) for x in h] + ["4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = [int(x) for x in r.stdout.split("tricks N E S W:")[1].split("\n")[0].split()]
    assert got == [int(x) for x in want[16:20]]
    assert subprocess.run([TOOL, "1", "2"] + ["0"] * 2 + ["9"], capture_output=True).returncode == 2
