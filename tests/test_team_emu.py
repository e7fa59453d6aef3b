"""The 16-lane team ops of the GPU kernel (dds_gpu_amd/csrc/ddx_team.hpp), restated lane by lane
on the host (tests/support/team_emu.cpp) and run through the same machine as the sequential ops:

  * their logic is the sequential ops': every job of the first 256 C1 endings enters exactly the
    reference's nodes;
  * a job's nodes do not depend on what the search ran before (each job starts on a fresh TT
    generation): with the jobs run last-first, every job still enters the reference's nodes,
    under both sets of ops. This is what lets the GPU, where any search may run any job, be
    compared node for node with the reference's one-context order."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT, golden

N = 256


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("team") / "libteam_emu.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                           "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "support", "team_emu.cpp"), "-o", out])
    lib = ctypes.CDLL(out)
    P = ctypes.c_void_p
    lib.ddx_team_emu_solve.argtypes = [P, ctypes.c_int64, ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int]
    return lib


@pytest.mark.parametrize("ops,order", [(1, 0), (1, 1), (0, 1)])
def test_job_nodes_are_the_references(emu, ops, order):
    c1 = golden("c1_endings.npz")
    hands = np.ascontiguousarray(c1["hands"][:N], np.uint64)
    out = np.zeros((N, 20), np.uint8)
    nodes = np.zeros(N * 5, np.uint64)
    emu.ddx_team_emu_solve(hands.ctypes.data, N, 10, out.ctypes.data, nodes.ctypes.data, ops, order)
    assert np.array_equal(out, c1["tables"][:N])
    assert np.array_equal(nodes.reshape(-1, 5), c1["nodes"][:N])
