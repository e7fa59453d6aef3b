"""GPU parity: the MI355X kernels through the C ABI vs the golden vectors and the oracle."""
import os
import threading
import time

import numpy as np
import pytest
from conftest import golden

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c1():
    return golden("c1_endings.npz")


@pytest.mark.parametrize("team", [1, 16])
@pytest.mark.parametrize("schedule", [0, 1, 2, 3])
def test_c1_endings_all_schedules(ddx, c1, schedule, team):
    """4,096 8-card endings (C1), TT_LOG2 10: all 81,920 answers equal the reference's, with one
    lane per search and with 16-lane teams, under every wave schedule and the per-thread one.
    Each job starts on a fresh TT generation, so with one lane per search the GPU enters exactly
    the reference's nodes, whichever search runs which job."""
    log("c1 endings, schedule", schedule, "team", team)
    with ddx.GpuSolver(lanes=8192 * team, tt_log2=10, schedule=schedule, team=team) as g:
        out, st = g.solve(c1["hands"])
    assert np.array_equal(out, c1["tables"])
    assert st.jobs_done == 5 * len(c1["hands"]) and st.nodes > 0
    log("nodes gpu %d, reference one-context total %d" % (st.nodes, int(c1["nodes"].sum())))
    if team == 1:
        assert st.nodes == int(c1["nodes"].sum())


def test_per_lead_jobs(ddx, c1):
    with ddx.GpuSolver(lanes=16384, tt_log2=10, jobs_per_deal=20) as g:
        out, _ = g.solve(c1["hands"][:2048])
    assert np.array_equal(out, c1["tables"][:2048])


def test_team_stream_of_mixed_endings(ddx, oracle):
    hands = np.concatenate([ddx.random_deals(700 + k, 50, k) for k in range(1, 11)])
    with ddx.GpuSolver(lanes=16384, tt_log2=9, team=16, jobs_per_deal=20) as g:
        out, _ = g.solve(hands)
    assert np.array_equal(out, oracle.tables(hands))


@pytest.mark.parametrize("mode", ["step", "cert", "certp"])
def test_seed_modes_exact(ddx, c1, mode):
    """Seeded searches (noisy seeds, even random ones) give the exact tables."""
    m = {"step": ddx.SEED_STEP, "cert": ddx.SEED_CERT, "certp": ddx.SEED_CERTP}[mode]
    rng = np.random.default_rng(11)
    seeds = rng.integers(0, 14, (2048, 20)).astype(np.uint8)
    seeds[:1024] = c1["seeds"][:1024]
    with ddx.GpuSolver(lanes=8192, tt_log2=10, seed_mode=m) as g:
        out, _ = g.solve(c1["hands"][:2048], seeds=None if mode == "certp" else seeds)
    assert np.array_equal(out, c1["tables"][:2048])


def test_endings_vs_oracle_mixed_sizes(ddx, oracle):
    """Endings of 1..9 cards in one batch (uneven jobs) vs the oracle."""
    hands = np.concatenate([ddx.random_deals(900 + k, 60, k) for k in range(1, 10)])
    want = oracle.tables(hands)
    with ddx.GpuSolver(lanes=4096, tt_log2=9, schedule=2) as g:
        out, _ = g.solve(hands)
    assert np.array_equal(out, want)


def test_full_deals_vs_dds_tables(ddx, oracle):
    """Full 13-card deals of the reference's DDS fixture, one job per lead."""
    hands, tables = oracle.pgx(golden("dds_4000.npy")[:, :2, :])
    t0 = time.time()
    with ddx.GpuSolver(lanes=4096, tt_log2=12, jobs_per_deal=20, team=16) as g:
        out, st = g.solve(hands)
    log("2 full deals: %.1fs, %.3g nodes" % (time.time() - t0, st.nodes))
    assert np.array_equal(out, tables)


# The eight bench-stream deals with the fewest reference nodes (a full deal's cost spans
# 3e4 to 1e9 nodes; these keep the test within seconds).
EASY_FULL = [1386, 707, 1081, 1897, 1468, 1032, 356, 1069]


def test_full_deals_vs_reference_fixture(ddx):
    """Deals of the bench stream vs the reference solver's tables."""
    g = golden("full_2000.npz")
    for team in (1, 16):
        with ddx.GpuSolver(lanes=256 * team, tt_log2=10, jobs_per_deal=20, team=team) as s:
            out, _ = s.solve(g["hands"][EASY_FULL])
        assert np.array_equal(out, g["tables"][EASY_FULL])




def test_repeated_calls_keep_generations_apart(ddx, oracle):
    """TT generations persist across calls: entries of an earlier batch never match."""
    a = ddx.random_deals(71, 300, 6)
    b = ddx.random_deals(72, 300, 6)
    wa, wb = oracle.tables(a), oracle.tables(b)
    with ddx.GpuSolver(lanes=1024, tt_log2=9) as g:
        for _ in range(14):  # > 63 generations per lane: exercises the tag clear
            assert np.array_equal(g.solve(a)[0], wa)
            assert np.array_equal(g.solve(b)[0], wb)


def test_stop_leaves_unsolved_marked(ddx, oracle):
    """A stop request ends the run; unfinished answers are 0xFF, finished ones exact."""
    hands = ddx.random_deals(81, 3000, 10)
    with ddx.GpuSolver(lanes=1024, tt_log2=10) as g:
        timer = threading.Timer(1.5, g.request_stop)
        timer.start()
        out, st = g.solve(hands, allow_stop=True)
        timer.cancel()
    done = (out != 255).all(axis=1)
    assert 0 < st.jobs_done < 5 * len(hands)
    want = oracle.tables(hands[done][:200])
    assert np.array_equal(out[done][:200], want)
    # answers of a strain are written lead by lead: every byte is either final or 0xFF
    part = out[~done]
    assert ((part == 255) | (part <= 13)).all()


@pytest.mark.parametrize("team,tt_log2,waves", [(1, 10, 0), (16, 12, 0), (16, 12, 3)])
def test_stream_matches_batch(ddx, c1, team, tt_log2, waves):
    """The persistent streaming kernel: all deals come back, tables exact, ids intact."""
    n = 1500
    hands = c1["hands"][:n]
    got = {}
    with ddx.Stream(lanes=4096, tt_log2=tt_log2, jobs_per_deal=20, ring_slots=256, team=team,
                    waves_per_simd=waves) as s:
        i = 0
        t0 = time.time()
        while len(got) < n and time.time() - t0 < 60:
            while i < n and s.push(hands[i], i):
                i += 1
            ids, tabs = s.poll()
            for k, t in zip(ids, tabs):
                got[int(k)] = t
            if i == n:
                s.finish()
            time.sleep(0.001)
        st = s.close()
    assert len(got) == n
    out = np.stack([got[k] for k in range(n)])
    assert np.array_equal(out, c1["tables"][:n])
    assert st.jobs_done == 20 * n


@pytest.mark.parametrize("waves,lds", [(3, 0), (4, 0), (0, 3), (3, 3)])
def test_occupancy_and_lds_variants_exact(ddx, c1, waves, lds):
    """Kernels held to 3 or 4 waves per SIMD, and the LDS tag-line cache: C1 tables exact; the
    cache serves a real share of the tag lines."""
    hands = c1["hands"]
    with ddx.GpuSolver(lanes=8192, tt_log2=10, jobs_per_deal=5, schedule=2, waves_per_simd=waves,
                       lds_lines=lds) as g:
        out, st = g.solve(hands)
    log("waves %d lds %d: nodes %d, line cache hits %d" % (waves, lds, st.nodes, st.tt_line_cache_hits))
    assert np.array_equal(out, c1["tables"])
    if lds:
        assert st.tt_line_cache_hits > 0.05 * st.nodes
    else:
        assert st.tt_line_cache_hits == 0


def test_stream_with_net_predictions_while_running(ddx, c1):
    """Seeds predicted one deal at a time on the host while the persistent stream kernel runs
    (the net's copies and kernels must not wait for that kernel): tables exact."""
    model = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dds_gpu_amd",
                         "models", "tricknet.bin")
    if not os.path.exists(model):
        pytest.skip("no trained net")
    n = 300
    hands = c1["hands"][:n]
    net = ddx.TrickNet(path=model)
    net.predict(hands[:1])  # buffers exist before the stream starts
    got = {}
    with ddx.Stream(lanes=4096, tt_log2=10, jobs_per_deal=20, seed_mode=ddx.SEED_CERT, ring_slots=256) as s:
        i = 0
        t0 = time.time()
        while len(got) < n and time.time() - t0 < 60:
            while i < n:
                seeds = net.predict(hands[i:i + 1])[0]
                if not s.push(hands[i], i, seeds):
                    break
                i += 1
            ids, tabs = s.poll()
            for k, t in zip(ids, tabs):
                got[int(k)] = t
            if i == n:
                s.finish()
            time.sleep(0.001)
        s.close()
    net.close()
    assert len(got) == n
    assert np.array_equal(np.stack([got[k] for k in range(n)]), c1["tables"][:n])


def test_dd_pbn_gpu_mode(c1):
    """dd_pbn with GPU=1 (one batch on the GPU): leader-side tricks of the C1 endings."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    R = "23456789TJQKA"
    hands, tables = c1["hands"][:200], c1["tables"][:200]
    text = "\n".join("N:" + " ".join(".".join("".join(R[r] for r in range(12, -1, -1) if (int(h[p]) >> (16 * s + r)) & 1)
                                             for s in range(4)) for p in range(4)) for h in hands) + "\n"
    r = subprocess.run([os.path.join(root, "dds_gpu_amd", "bin", "dd_pbn")], input=text, capture_output=True,
                       text=True, timeout=200, env=dict(os.environ, GPU="1"))
    assert r.returncode == 0, r.stderr
    got = np.array([list(map(int, l.split())) for l in r.stdout.strip().split("\n")]).reshape(-1, 5, 4)
    t = tables.reshape(-1, 5, 4).astype(int)
    for leader in range(4):
        decl = (leader + 3) % 4
        ns = t[:, :, decl] if decl % 2 == 0 else 13 - t[:, :, decl]
        assert np.array_equal(got[:, :, leader], ns if leader % 2 == 0 else 8 - ns)


def test_bad_configs_fail_cleanly(ddx):
    with pytest.raises(ddx.DdxError):
        ddx.GpuSolver(lanes=100)
    with pytest.raises(ddx.DdxError):
        ddx.GpuSolver(tt_log2=8)
    with ddx.GpuSolver(lanes=256) as g:
        with pytest.raises(ddx.DdxError):
            g.solve(np.array([[1, 1, 2, 4]], np.uint64))


def test_graph_launch_replays_match_eager(ddx, c1, oracle):
    """Batches launched as a captured HIP graph, replayed several times with eager work in
    between: every replay's tables are exact."""
    hs = [c1["hands"][i * 500:(i + 1) * 500] for i in range(3)]
    with ddx.GpuSolver(lanes=4096, tt_log2=10, graph=True) as g, \
            ddx.GpuSolver(lanes=2048, tt_log2=10) as eager:
        for rep in range(4):
            i = rep % 3
            out, st = g.solve(hs[i])
            assert np.array_equal(out, c1["tables"][i * 500:(i + 1) * 500]), rep
            assert st.jobs_done == 2500
            e, _ = eager.solve(hs[(i + 1) % 3])
            assert np.array_equal(e, c1["tables"][((i + 1) % 3) * 500:((i + 1) % 3 + 1) * 500])


def test_stream_close_after_finish_stops_heavy_searches(ddx):
    """close() after finish() stops searches still running: the four heaviest deals of the bench
    stream (jobs of 1e8-5e8 nodes, hours on one search) are pushed, finish() marks the end of
    input, and close() one second later returns within seconds -- the kernel keeps reading the
    stop bit after the end of input -- with none of those deals reported."""
    g = golden("full_2000.npz")
    heavy = np.argsort(g["nodes"].astype(np.int64).max(axis=1))[-4:]
    with ddx.Stream(lanes=4096, tt_log2=10, jobs_per_deal=5, ring_slots=64) as s:
        for k, d in enumerate(heavy):
            assert s.push(g["hands"][d], k)
        s.finish()
        time.sleep(1.0)
        ids, _ = s.poll()
        t0 = time.time()
        s.close()
        took = time.time() - t0
    log("close after finish took %.2f s" % took)
    assert took < 10 and len(ids) == 0


@pytest.mark.parametrize("schedule", [0, 2])
def test_stage_profile_counts_are_consistent(ddx, c1, schedule):
    """The per-stage profile (config.profile; the reference's -DDD_PROFILE output,
    gpu/gpu_main.cu:282-301): every node entry is counted at exactly one exit point, so the exit
    counts sum to the solve's nodes; a round's active lanes never exceed the wave's 64; cycles
    are counted for every kind that ran; and the profiling build's tables stay exact."""
    with ddx.GpuSolver(lanes=8192, tt_log2=10, schedule=schedule, profile=True) as g:
        g.profile(reset=True)
        out, st = g.solve(c1["hands"])
        v = g.profile(reset=True).astype(np.int64)
    assert np.array_equal(out, c1["tables"])
    rounds, lanes, exits, cycles = v[0:8], v[8:16], v[16:23], v[24:32]
    log("schedule %d: rounds %s, lanes %s, exits %s" % (schedule, rounds.tolist(), lanes.tolist(), exits.tolist()))
    assert exits.sum() == st.nodes
    assert (lanes <= 64 * rounds).all() and (lanes[rounds > 0] > 0).all()
    assert (cycles[rounds > 0] > 0).all() and rounds[2] > 0  # ENTER (trick start) always runs
