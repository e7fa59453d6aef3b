"""GPU parity of the split kernels (include/ddx/split.hpp): once the job queue is empty, long
searches are split over idle lanes. Low thresholds force many splits, nested joins and
cut-off tasks; every table must still equal the oracle's or the reference solver's.

Each solve runs under a host timer that requests a stop: a search that never finished would
end the kernel at its next stop poll and fail the test instead of holding the GPU."""
import threading
import time

import numpy as np
import pytest
from conftest import golden

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def solve_limited(g, hands, seconds, seeds=None):
    timer = threading.Timer(seconds, g.request_stop)
    timer.start()
    try:
        out, st = g.solve(hands, seeds=seeds, allow_stop=True)
    finally:
        timer.cancel()
    assert st.jobs_done == g.cfg.jobs_per_deal * len(hands), "stopped after %.0f s" % seconds
    return out, st


@pytest.fixture(scope="module")
def endings(ddx, oracle):
    hands = np.concatenate([ddx.random_deals(3100 + k, 32, k) for k in range(6, 10)])
    return hands, oracle.tables(hands)


@pytest.mark.parametrize("waves", [0, 3])
@pytest.mark.parametrize("jpd", [5, 20])
def test_split_endings_vs_oracle(ddx, endings, waves, jpd):
    """Endings of 6..9 cards, many more lanes than jobs, a 64-node split threshold."""
    hands, want = endings
    log("split endings: waves", waves, "jobs per deal", jpd)
    with ddx.GpuSolver(lanes=16384, tt_log2=10, jobs_per_deal=jpd, schedule=ddx.SCHED_SPLIT,
                       waves_per_simd=waves, split_min_nodes=64) as g:
        t0 = time.time()
        out, st = solve_limited(g, hands, 120)
        log("  %.1f s, %d nodes, %d splits, %d tasks taken" % (time.time() - t0, st.nodes, st.splits, st.tasks_taken))
    assert np.array_equal(out, want)
    assert st.splits > 0 and st.tasks_taken > 0


def test_split_full_deals_vs_reference(ddx):
    """Full deals of the bench stream (the 24 with the most nodes under 4e6 in the reference
    solver's count) on 16,384 lanes: all their 480 answers equal the reference's tables."""
    g = golden("full_2000.npz")
    tot = g["nodes"].astype(np.int64).sum(axis=1)
    idx = np.nonzero(tot < 4_000_000)[0]
    idx = idx[np.argsort(tot[idx])[-24:]]
    with ddx.GpuSolver(lanes=16384, tt_log2=10, jobs_per_deal=5, schedule=ddx.SCHED_SPLIT,
                       split_min_nodes=4096) as s:
        t0 = time.time()
        out, st = solve_limited(s, g["hands"][idx], 150)
        log("split full deals: %.1f s, %d nodes (reference %d), %d splits, %d tasks taken" %
            (time.time() - t0, st.nodes, int(tot[idx].sum()), st.splits, st.tasks_taken))
    assert np.array_equal(out, g["tables"][idx])
    assert st.splits > 0


def test_split_learned_seeds(ddx):
    """Seeded searches (CERT) with splitting: exact on the C1 endings with random seeds."""
    c1 = golden("c1_endings.npz")
    rng = np.random.default_rng(5)
    seeds = rng.integers(0, 9, (512, 20)).astype(np.uint8)
    with ddx.GpuSolver(lanes=8192, tt_log2=10, schedule=ddx.SCHED_SPLIT, seed_mode=ddx.SEED_CERT,
                       split_min_nodes=32) as g:
        out, st = solve_limited(g, c1["hands"][:512], 120, seeds=seeds)
    assert np.array_equal(out, c1["tables"][:512])
    assert st.splits > 0
