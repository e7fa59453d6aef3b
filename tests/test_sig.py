"""The signature-entry build (libddx_sig.so: 12-byte TT entries holding a 64-bit signature of
the key; reference -DDD_TT_SIG, src/dd.h:74-85, 352-437) against the reference built the same
way (tests/golden/sig_entries.npz, written by oracle/_ref/ref_harness_sig).

Pinned three ways:
  * node for node: the CPU machine of the signature build spends exactly the reference's
    nodes on every job (one context, jobs in index order), on 8- and 11-card endings and
    full deals;
  * against the exact build: the same deals' tables equal the exact reference's (c1, mixed
    and full fixtures); on the 8-card endings the two builds spend the same nodes, while on
    full deals the signature build, which also stores positions with more than 29 relevant
    cards, spends fewer;
  * on the GPU: every table of the signature kernel equals the fixture's, and its node total
    equals the reference's."""
import numpy as np
import pytest
from conftest import golden

SIG = golden("sig_entries.npz")


@pytest.fixture(scope="module")
def ddx_sig(ddx):
    ddx.lib("sig")
    return ddx


def test_fixture_deals_are_the_exact_fixtures_deals(ddx):
    c1, mixed, full = golden("c1_endings.npz"), golden("mixed_endings.npz"), golden("full_2000.npz")
    assert np.array_equal(SIG["c1_hands"], c1["hands"][:len(SIG["c1_hands"])])
    k11 = mixed["hands"][mixed["k"] == 11]
    assert np.array_equal(SIG["k11_hands"], k11[:len(SIG["k11_hands"])])
    assert np.array_equal(SIG["full_hands"], full["hands"][SIG["full_rows"]])
    # No signature collision on these deals: the tables are the exact solver's.
    assert np.array_equal(SIG["c1_tables"], c1["tables"][:len(SIG["c1_tables"])])
    assert np.array_equal(SIG["k11_tables"], mixed["tables"][mixed["k"] == 11][:len(SIG["k11_tables"])])
    assert np.array_equal(SIG["full_tables"], full["tables"][SIG["full_rows"]])
    # On the 8-card endings both builds spend the same nodes; on full deals, where positions
    # with more than 29 relevant cards are common, the signature build stores those too and
    # spends fewer.
    assert np.array_equal(SIG["c1_nodes"], c1["nodes"][:len(SIG["c1_nodes"])])
    assert SIG["full_nodes"].sum() < full["nodes"][SIG["full_rows"]].sum()


@pytest.mark.parametrize("group", ["c1", "k11", "full"])
def test_cpu_signature_build_node_for_node(ddx_sig, group):
    hands = SIG[group + "_hands"]
    out, st, nodes = ddx_sig.cpu_solve(hands, tt_log2=10, job_nodes=True, variant="sig")
    assert np.array_equal(out, SIG[group + "_tables"])
    assert np.array_equal(nodes.reshape(-1, 5), SIG[group + "_nodes"])


def test_exact_build_differs_where_entries_would_exceed_29_cards(ddx):
    hands = SIG["full_hands"][:4]
    out, st, nodes = ddx.cpu_solve(hands, tt_log2=10, job_nodes=True)
    assert np.array_equal(out, SIG["full_tables"][:4])
    assert not np.array_equal(nodes.reshape(-1, 5), SIG["full_nodes"][:4])


def test_signature_build_has_no_team_kernels(ddx_sig):
    for open_ in (lambda: ddx_sig.GpuSolver(lanes=4096, tt_log2=10, team=16, variant="sig"),
                  lambda: ddx_sig.Stream(lanes=4096, tt_log2=10, team=16, variant="sig")):
        with pytest.raises(ddx_sig.DdxError, match="one lane per search"):
            open_()


@pytest.mark.gpu
def test_gpu_signature_kernel_tables(ddx_sig):
    for group in ("c1", "k11", "full"):
        hands = SIG[group + "_hands"]
        with ddx_sig.GpuSolver(lanes=4096, tt_log2=10, jobs_per_deal=5, schedule=ddx_sig.SCHED_SPLIT,
                               variant="sig") as s:
            out, st = s.solve(hands)
        print("signature kernel, %s: %d deals, %d nodes (reference one-context total %d)" %
              (group, len(hands), st.nodes, int(SIG[group + "_nodes"].sum())), flush=True)
        assert st.jobs_done == 5 * len(hands)
        assert np.array_equal(out, SIG[group + "_tables"])
        assert st.nodes == int(SIG[group + "_nodes"].sum())  # every job on a fresh generation
