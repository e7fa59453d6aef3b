"""The ddx core on the CPU vs the reference solver's golden vectors and the oracle."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT, golden

SUPPORT_SRC = os.path.join(ROOT, "tests", "support", "ddx_funcs.cpp")


@pytest.fixture(scope="module")
def funcs_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("support") / "libddx_funcs.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                           "-I" + os.path.join(ROOT, "include"), SUPPORT_SRC, "-o", out])
    lib = ctypes.CDLL(out)
    lib.ddx_test_funcs.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    return lib


FIELD_GROUPS = {
    "suit codes": range(14, 18), "lengths": [18], "hash": [19],
    "quick tricks": [20, 21, 22], "later tricks": [23, 24], "2nd-hand quick tricks": [25, 26],
    "move count": [27], "move order": range(28, 41), "move weights": range(41, 54),
}


def test_helpers_match_reference_vectors(funcs_lib):
    """Keys, QuickTricks, LaterTricks, 2nd-hand QT, move order and weights on 6,000 random
    positions == the reference's outputs (tests/golden/funcs.npz)."""
    rec = golden("funcs.npz")["records"]
    out = np.zeros_like(rec)
    funcs_lib.ddx_test_funcs(rec.ctypes.data, len(rec), out.ctypes.data)
    assert (rec[:, 6] == 0).sum() > 1000 and (rec[:, 6] == 1).sum() > 1000
    for name, cols in FIELD_GROUPS.items():
        cols = list(cols)
        bad = np.nonzero((out[:, cols] != rec[:, cols]).any(axis=1))[0]
        assert len(bad) == 0, "%s differs at %d positions, first %s" % (name, len(bad), bad[:5])


def test_c1_endings_tables_and_nodes(ddx):
    """4,096 8-card endings, TT_LOG2 10, jobs in index order on one context: every table and
    the node count of every job equal the reference's."""
    g = golden("c1_endings.npz")
    t, st, nodes = ddx.cpu_solve(g["hands"], threads=1, tt_log2=10, job_nodes=True)
    assert np.array_equal(t, g["tables"])
    assert np.array_equal(nodes.reshape(-1, 5), g["nodes"])


def test_c1_endings_step_from_guess(ddx):
    """dd_solve_strain_guess semantics: same tables, reference node counts job by job."""
    g = golden("c1_endings.npz")
    t, st, nodes = ddx.cpu_solve(g["hands"], threads=1, tt_log2=10, seed_mode=ddx.SEED_STEP,
                                 seeds=g["seeds"], job_nodes=True)
    assert np.array_equal(t, g["tables"])
    assert np.array_equal(nodes.reshape(-1, 5), g["nodes_seeded"])


def test_full_deals_nodes_match_reference(ddx):
    """First 3 full deals of the bench stream on a fresh context: identical node counts."""
    g = golden("full_2000.npz")
    t, st, nodes = ddx.cpu_solve(g["hands"][:3], threads=1, tt_log2=10, job_nodes=True)
    assert np.array_equal(t, g["tables"][:3])
    assert np.array_equal(nodes.reshape(-1, 5), g["nodes"][:3])


def test_full_deals_vs_dds_tables(ddx, oracle):
    """16 full deals of the reference's DDS fixture: tables equal DDS's."""
    hands, tables = oracle.pgx(golden("dds_4000.npy")[:, 100:116, :])
    t, _ = ddx.cpu_solve(hands, threads=min(8, os.cpu_count()), tt_log2=12)
    assert np.array_equal(t, tables)


@pytest.mark.parametrize("mode", ["step", "cert", "certp"])
def test_seed_policies_are_exact(ddx, oracle, mode):
    """Any seeds, even bad ones, give the exact tables (the seed only changes the work)."""
    h = ddx.random_deals(55, 300, 7)
    want = oracle.tables(h)
    rng = np.random.default_rng(3)
    seeds = rng.integers(0, 14, (300, 20)).astype(np.uint8)
    m = {"step": ddx.SEED_STEP, "cert": ddx.SEED_CERT, "certp": ddx.SEED_CERTP}[mode]
    t, _ = ddx.cpu_solve(h, threads=4, tt_log2=10, seed_mode=m,
                         seeds=None if mode == "certp" else seeds)
    assert np.array_equal(t, want)


def test_exact_seeds_need_two_searches_per_lead(ddx):
    """CERT with the true NS tricks as seeds: at most 2 searches per lead, fewer nodes."""
    g = golden("c1_endings.npz")
    h, tabs = g["hands"][:400], g["tables"][:400]
    seeds = ddx.ns_tricks_of(tabs).clip(0, 13).astype(np.uint8)
    t0, s0 = ddx.cpu_solve(h, threads=1, tt_log2=10)
    t1, s1 = ddx.cpu_solve(h, threads=1, tt_log2=10, seed_mode=ddx.SEED_CERT, seeds=seeds)
    assert np.array_equal(t0, t1)
    assert s1.nodes < s0.nodes


def test_jobs_per_lead_same_tables(ddx):
    g = golden("c1_endings.npz")
    t, _ = ddx.cpu_solve(g["hands"][:500], threads=4, tt_log2=10, jobs_per_deal=20)
    assert np.array_equal(t, g["tables"][:500])


def test_generation_wrap_keeps_results_exact(ddx):
    """More than 63 solves on one context wrap the TT generation (tags cleared)."""
    g = golden("c1_endings.npz")
    h = g["hands"][:40]  # 200 jobs > 63 generations
    t, _ = ddx.cpu_solve(h, threads=1, tt_log2=9)
    assert np.array_equal(t, g["tables"][:40])


def test_cpu_context_api(ddx, oracle):
    h = ddx.random_deals(9, 20, 6)
    want = oracle.tables(h).reshape(-1, 5, 4)
    ctx = ddx.CpuContext(tt_log2=12)
    for i in range(len(h)):
        for st in range(5):
            assert list(ctx.solve_strain(h[i], st)) == list(want[i, st])
            ns = [int(want[i, st, d]) if d % 2 == 0 else 13 - int(want[i, st, d]) for d in range(4)]
            guess = [max(0, x - 1) for x in ns]
            assert list(ctx.solve_strain(h[i], st, guess)) == list(want[i, st])
            for leader in range(4):
                d = (leader + 3) & 3
                assert ctx.ns_tricks(h[i], st, leader, guess=3) == ns[d]
    assert ctx.nodes > 0
    ctx.close()


def test_errors_are_reported(ddx):
    bad = np.array([[1, 1, 2, 4]], np.uint64)  # overlapping hands
    with pytest.raises(ddx.DdxError) as e:
        ddx.cpu_solve(bad)
    assert e.value.rc == -1 and "not a valid deal" in str(e.value)
    with pytest.raises(ddx.DdxError):
        ddx.CpuContext(tt_log2=8)
    with pytest.raises(ddx.DdxError):
        ddx.cpu_solve(ddx.random_deals(1, 2, 5), tt_log2=8)


def test_mixed_endings_fixture(ddx):
    """Endings of 5, 7, 9 and 11 cards: the CPU machine gives the reference solver's tables."""
    m = golden("mixed_endings.npz")
    out, _ = ddx.cpu_solve(m["hands"], threads=8, tt_log2=10)
    assert np.array_equal(out, m["tables"])
