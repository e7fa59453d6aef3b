"""The trained trick-count net (dds_gpu_amd/models/tricknet.bin) as a seed of the exact search:
what it predicts, that every seeded answer stays exact, and what it saves over the plain
binary search. The deals are the first 2,000 of the bench's stream (2026), which the net never
saw (it is trained on stream 7001)."""
import os
import sys

import numpy as np
import pytest
from conftest import ROOT, golden

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_ref  # noqa: E402

MODEL = os.path.join(ROOT, "dds_gpu_amd", "models", "tricknet.bin")


@pytest.fixture(scope="module")
def net_params():
    if not os.path.exists(MODEL):
        pytest.skip("no trained net in dds_gpu_amd/models")
    return mlp_ref.load(MODEL)


@pytest.fixture(scope="module")
def easy_deals():
    """Fixture deals whose whole table costs the reference solver under 1.5M nodes at TT_LOG2
    10, so that CPU and GPU runs finish in seconds (the first 64 such deals)."""
    g = golden("full_2000.npz")
    cost = g["nodes"].astype(np.int64).sum(axis=1)
    idx = np.nonzero(cost < 1_500_000)[0][:64]
    return g["hands"][idx], g["tables"][idx]


def test_net_predicts_better_than_chance(net_params):
    """Held-out accuracy of the oracle forward on 200 fixture deals: the exact NS count more
    often than the best constant guess, and within one trick for most entries."""
    g = golden("full_2000.npz")
    X, y = mlp_ref.encode(g["hands"][:200], g["tables"][:200])
    pred, _ = mlp_ref.predict(net_params, X)
    exact = float((pred == y).mean())
    within1 = float((np.abs(pred - y) <= 1).mean())
    const = float((y == np.bincount(y).argmax()).mean())
    print("held out: exact %.3f, within one %.3f, constant %.3f" % (exact, within1, const))
    assert exact > const + 0.1 and within1 > 0.8


def test_learned_seeds_exact_and_cheaper_on_cpu(ddx, net_params, easy_deals):
    """CPU solver, same TT and jobs: seeded by the net (DDX_SEED_CERT) the tables equal the
    reference's and the search visits fewer nodes than the plain binary search."""
    hands, tables = easy_deals
    X, _ = mlp_ref.encode(hands, tables)
    seeds = mlp_ref.predict(net_params, X)[0].astype(np.uint8).reshape(-1, 20)
    plain, sp = ddx.cpu_solve(hands, threads=8, tt_log2=12)
    learned, sl = ddx.cpu_solve(hands, threads=8, tt_log2=12, seed_mode=ddx.SEED_CERT, seeds=seeds)
    print("nodes: plain %d, learned %d (%.3f)" % (sp.nodes, sl.nodes, sl.nodes / sp.nodes))
    assert np.array_equal(plain, tables) and np.array_equal(learned, tables)
    assert sl.nodes < sp.nodes


@pytest.mark.gpu
def test_gpu_predictions_match_oracle(ddx, net_params):
    """The GPU forward (BF16 MFMA kernels) picks the oracle's class except at near-ties."""
    g = golden("full_2000.npz")
    hands = g["hands"][:500]
    net = ddx.TrickNet(path=MODEL)
    got = net.predict(hands).reshape(-1)
    net.close()
    X, _ = mlp_ref.encode(hands, g["tables"][:500])
    want, Z = mlp_ref.predict(net_params, X)
    top2 = np.sort(Z, axis=1)[:, -2:]
    clear = (top2[:, 1] - top2[:, 0]) > 0.05
    assert (got[clear] == want[clear]).all()
    assert (got == want).mean() > 0.99


@pytest.mark.gpu
def test_learned_seeds_exact_and_cheaper_on_gpu(ddx, easy_deals):
    """GPU batch solver (16-lane teams), the net's GPU predictions as seeds: exact tables and
    fewer nodes than the plain binary search and than the no-net seed (previous lead's answer,
    DDX_SEED_CERTP)."""
    hands, tables = easy_deals
    net = ddx.TrickNet(path=MODEL)
    seeds = net.predict(hands)
    net.close()
    res = {}
    for name, mode, s in [("plain", ddx.SEED_PLAIN, None), ("certp", ddx.SEED_CERTP, None),
                          ("learned", ddx.SEED_CERT, seeds)]:
        with ddx.GpuSolver(lanes=len(hands) * 5 * 16 // 256 * 256 + 256, team=16, tt_log2=12,
                           jobs_per_deal=5, schedule=2, seed_mode=mode) as g:
            out, st = g.solve(hands, seeds=s)
        assert np.array_equal(out, tables), name
        res[name] = st.nodes
    print("nodes:", res, "learned/plain %.3f" % (res["learned"] / res["plain"]))
    assert res["learned"] < res["plain"]
