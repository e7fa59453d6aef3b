"""Full search vs the plain search (the reference's check, tests/check.sh:17-24): the core
compiled with every pruning switched off (DDX_NO_TT, DDX_NO_QT, DDX_NO_LT, DDX_NO_QT2,
DDX_NO_SMALL) must give the same tables as the full solver on K = 5..8-card endings, all five
strains (the reference's own check leaves NT out), and both must equal the oracle."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "support", "plain_build.cpp")
PLAIN = ["-DDDX_NO_TT=1", "-DDDX_NO_QT=1", "-DDDX_NO_LT=1", "-DDDX_NO_QT2=1", "-DDDX_NO_SMALL=1"]


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    d = tmp_path_factory.mktemp("plain")
    libs = {}
    for name, flags in [("full", []), ("plain", PLAIN)]:
        out = str(d / ("lib%s.so" % name))
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                               "-I" + os.path.join(ROOT, "include")] + flags + [SRC, "-o", out])
        lib = ctypes.CDLL(out)
        lib.ddx_plain_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
        lib.ddx_plain_run.restype = ctypes.c_uint64
        libs[name] = lib

    def run(name, hands, tt_log2=12):
        hands = np.ascontiguousarray(hands, np.uint64)
        out = np.zeros((len(hands), 20), np.uint8)
        nodes = libs[name].ddx_plain_run(hands.ctypes.data, len(hands), tt_log2, out.ctypes.data)
        return out, nodes
    return run


@pytest.mark.parametrize("K,n", [(5, 300), (6, 300), (7, 100), (8, 20)])
def test_full_equals_plain(ddx, oracle, builds, K, n):
    hands = ddx.random_deals(99 + K, n, K)
    full, nf = builds("full", hands)
    plain, np_ = builds("plain", hands)
    assert np.array_equal(full, plain)
    assert np.array_equal(full, oracle.tables(hands))
    assert np_ > nf  # the pruning is switched off: the plain search enters more nodes


def test_per_leader_switch_matches_the_reference(tmp_path):
    """DDX_PER_LEADER_TT (the reference's -DDD_PER_LEADER_TT measurement switch: the four leads
    of a strain share no TT entries) enters exactly the reference's nodes with the same switch
    on the first 400 C1 endings (tests/golden/per_leader_tt.npz), 1.85x the shared-TT count."""
    from conftest import golden
    g = golden("per_leader_tt.npz")
    out_so = str(tmp_path / "libperlead.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                           "-I" + os.path.join(ROOT, "include"), "-DDDX_PER_LEADER_TT=1", SRC, "-o", out_so])
    lib = ctypes.CDLL(out_so)
    lib.ddx_plain_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
    lib.ddx_plain_run.restype = ctypes.c_uint64
    hands = np.ascontiguousarray(g["hands"], np.uint64)
    out = np.zeros((len(hands), 20), np.uint8)
    nodes = lib.ddx_plain_run(hands.ctypes.data, len(hands), 10, out.ctypes.data)
    assert np.array_equal(out, g["tables"])
    assert nodes == int(g["nodes"].sum())
    assert nodes > 1.5 * int(golden("c1_endings.npz")["nodes"][:len(hands)].sum())
