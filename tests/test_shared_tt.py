"""The shared-TT build (libddx_shared.so: one lock-free TT for every search, 32-byte entries
with their whole key and a checksum, valid across deals with the same trump; reference
-DDD_SHARED_TT, src/dd.h:66-73, 223-227, 246-323) against the reference built the same way
(tests/golden/shared_tt.npz, written by oracle/_ref/ref_harness_shared).

Pinned three ways:
  * node for node: one CPU context of the shared build spends exactly the reference's nodes on
    every job of 8- and 11-card endings and full deals, with the TT kept across the deals of a
    group (so entries written for one deal serve later ones);
  * against the exact build: the tables equal the exact reference's;
  * under contention: CPU threads sharing one TT, and on the GPU 65,536 lanes sharing one small
    TT (stores overwrite each other's ways constantly, and readers meet entries being written),
    give every table of the fixtures."""
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT, golden

SH = golden("shared_tt.npz")


@pytest.fixture(scope="module")
def ddx_shared(ddx):
    ddx.lib("shared")
    return ddx


def test_fixture_tables_are_the_exact_reference_tables():
    c1, mixed, full = golden("c1_endings.npz"), golden("mixed_endings.npz"), golden("full_2000.npz")
    assert np.array_equal(SH["c1_hands"], c1["hands"][:len(SH["c1_hands"])])
    assert np.array_equal(SH["c1_tables"], c1["tables"][:len(SH["c1_tables"])])
    k11 = mixed["k"] == 11
    assert np.array_equal(SH["k11_hands"], mixed["hands"][k11][:len(SH["k11_hands"])])
    assert np.array_equal(SH["k11_tables"], mixed["tables"][k11][:len(SH["k11_tables"])])
    assert np.array_equal(SH["full_hands"], full["hands"][SH["full_rows"]])
    assert np.array_equal(SH["full_tables"], full["tables"][SH["full_rows"]])


@pytest.mark.parametrize("group", ["c1", "k11", "full"])
def test_cpu_shared_build_node_for_node(ddx_shared, group):
    hands = SH[group + "_hands"]
    out, st, nodes = ddx_shared.cpu_solve(hands, tt_log2=10, job_nodes=True, variant="shared")
    assert np.array_equal(out, SH[group + "_tables"])
    assert np.array_equal(nodes.reshape(-1, 5), SH[group + "_nodes"])


def test_entries_serve_later_deals(ddx_shared):
    """The same endings solved twice in one context: the second pass finds the first pass's
    entries (the shared TT keeps them across deals) and spends fewer nodes."""
    c = ddx_shared.CpuContext(tt_log2=16, variant="shared")
    hands = SH["c1_hands"][:50]
    spent = []
    for _ in range(2):
        before = c.nodes
        for h in hands:
            for strain in range(5):
                c.solve_strain(h, strain)
        spent.append(c.nodes - before)
    c.close()
    assert spent[1] < spent[0] / 2


def test_torn_entries_are_misses(tmp_path):
    """A stored entry answers; with one of its words taken from another entry (a read that met
    a half-written way), or its bounds changed behind the checksum, it is a miss
    (tests/support/shared_tt_check.cpp)."""
    exe = str(tmp_path / "shared_tt_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wno-unknown-pragmas", "-DDDX_SHARED_TT=1",
                           "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "support", "shared_tt_check.cpp"), "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stdout


def test_cpu_threads_share_one_tt(ddx_shared):
    hands = SH["c1_hands"]
    out, st = ddx_shared.cpu_solve(hands, threads=6, tt_log2=9, variant="shared")
    assert np.array_equal(out, SH["c1_tables"])


def test_shared_build_rejects_what_it_lacks(ddx_shared):
    for kw in ({"team": 16}, {"split_min_nodes": 1024, "schedule": ddx_shared.SCHED_SPLIT},
               {"lds_lines": 3, "schedule": ddx_shared.SCHED_SPLIT}):
        with pytest.raises(ddx_shared.DdxError, match="shared-TT build"):
            ddx_shared.GpuSolver(lanes=4096, tt_log2=10, variant="shared", **kw)


@pytest.mark.gpu
def test_gpu_shared_tt_tables(ddx_shared):
    for group in ("c1", "k11", "full"):
        hands = SH[group + "_hands"]
        with ddx_shared.GpuSolver(lanes=4096, tt_log2=18, jobs_per_deal=5, schedule=ddx_shared.SCHED_SPLIT,
                                  variant="shared") as s:
            out, st = s.solve(hands)
        print("shared TT (2^18 buckets), %s: %d deals, %d nodes" % (group, len(hands), st.nodes), flush=True)
        assert st.jobs_done == 5 * len(hands)
        assert np.array_equal(out, SH[group + "_tables"])


@pytest.mark.gpu
def test_gpu_shared_tt_under_contention(ddx_shared):
    """65,536 lanes on one TT of 2^12 buckets (2.4 MB): 4,096 C1 endings, then the 12 full deals,
    then the endings again on the TT the first two batches left."""
    c1, full = golden("c1_endings.npz"), SH
    with ddx_shared.GpuSolver(lanes=65536, tt_log2=12, jobs_per_deal=5, schedule=ddx_shared.SCHED_SPLIT,
                              variant="shared") as s:
        for hands, want in ((c1["hands"], c1["tables"]), (full["full_hands"], full["full_tables"]),
                            (c1["hands"], c1["tables"])):
            out, st = s.solve(hands)
            print("shared TT (2^12 buckets), 65,536 lanes: %d deals, %d nodes" % (len(hands), st.nodes), flush=True)
            assert np.array_equal(out, want)
