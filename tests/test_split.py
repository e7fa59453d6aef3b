"""Splitting long searches over idle lanes (include/ddx/split.hpp), run on the host: many lanes
advance round-robin through the same machine the GPU runs (tests/support/split_sim.cpp), lanes
without a job take published tasks, and owners wait for their joins. Every table must equal the
oracle's (endings) or the reference solver's (full deals) whatever the split threshold, lane
count and job split; with splitting off the simulation must enter exactly the plain solver's
nodes; and with it on, tasks must actually be published and taken."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT, golden

SRC = os.path.join(ROOT, "tests", "support", "split_sim.cpp")


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("split") / "libsplit_sim.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                           "-I" + os.path.join(ROOT, "include"), SRC, "-o", out])
    lib = ctypes.CDLL(out)
    P = ctypes.c_void_p
    lib.ddx_split_sim.argtypes = [P, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int64, P, P]

    def run(hands, lanes, tt_log2=10, jpd=5, min_left=2, min_nodes=64, task_cap=1 << 20):
        hands = np.ascontiguousarray(hands, np.uint64)
        n = len(hands)
        out = np.zeros((n, 20), np.uint8)
        st = np.zeros(9, np.int64)
        lib.ddx_split_sim(hands.ctypes.data, n, lanes, tt_log2, jpd, min_left, min_nodes, task_cap,
                          out.ctypes.data, st.ctypes.data)
        keys = ["rounds", "nodes", "splits", "taken", "dropped", "slots", "joins", "empty_at", "abandoned"]
        return out, dict(zip(keys, st.tolist()))
    return run


def test_split_off_is_the_plain_machine(ddx, sim):
    """With splitting off, one lane enters exactly the CPU solver's nodes, job by job."""
    hands = ddx.random_deals(31, 64, 8)
    out, st = sim(hands, lanes=1, min_left=-1)
    want = ddx.cpu_solve(hands, threads=1, tt_log2=10, jobs_per_deal=5)
    assert np.array_equal(out, want[0])
    assert st["nodes"] == want[1].nodes and st["splits"] == 0


@pytest.mark.parametrize("K,n,lanes,jpd,min_nodes", [(7, 48, 64, 5, 16), (9, 24, 96, 20, 32),
                                                     (10, 12, 128, 5, 64), (8, 40, 16, 20, 16)])
def test_split_endings_equal_the_oracle(ddx, oracle, sim, K, n, lanes, jpd, min_nodes):
    """Mixed endings with a low split threshold (many splits, nested joins, cut-offs that drop
    running tasks): every answer equals the oracle's."""
    hands = ddx.random_deals(700 + K, n, K)
    out, st = sim(hands, lanes=lanes, jpd=jpd, min_nodes=min_nodes)
    assert np.array_equal(out, oracle.tables(hands))
    assert st["splits"] > 0 and st["taken"] > 0


def test_split_full_deals_equal_the_reference(sim):
    """Full deals of the reference-solved fixture, more lanes than jobs: splits start at once."""
    g = golden("full_2000.npz")
    idx = np.argsort(g["nodes"].astype(np.int64).sum(axis=1))[200:203]  # moderately heavy, cheap enough here
    out, st = sim(g["hands"][idx], lanes=64, jpd=20, min_nodes=4096)
    assert np.array_equal(out, g["tables"][idx])
    assert st["splits"] > 0


def test_split_pool_overflow_is_harmless(ddx, oracle, sim):
    """A task pool far too small: publishers that run out of slots mark them skipped, lanes holding
    tickets for them move on, and the searches finish on their own lanes -- still exact."""
    hands = ddx.random_deals(909, 32, 8)
    out, st = sim(hands, lanes=48, jpd=5, min_nodes=16, task_cap=40)
    assert np.array_equal(out, oracle.tables(hands))
    assert st["slots"] == 40
