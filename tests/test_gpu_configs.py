"""The workload configurations on the GPU, each within a test-sized time budget:

  C2  the plain wavefront kernel (65,536 lanes, TT_LOG2 10, one job per strain) on the first
      2,000 deals of the bench stream: every answer it finishes within the budget equals the
      reference solver's table; the finished fraction is printed and must stay high. The
      heaviest jobs of these deals cost up to 5.3e8 nodes, over an hour on one lane of this
      configuration; tests/test_zy_c2_to_completion.py solves all 40,000 answers with 16-lane
      teams and larger TTs.
  C3  an uneven batch (endings of 5, 7, 9 and 11 cards and full deals, shuffled) launched as a
      captured HIP graph on the tuned 3-waves-per-SIMD kernel, replayed: every answer equals
      the reference solver's, on every replay.
  C4  the same 2,000 deals seeded by the trained net (CERT) on the tuned configuration: every
      finished answer exact.
  C5  dd_stream on full deals with Poisson arrivals and the 4,096-slot ring: the CPU re-check
      of sampled deals finds no mismatch, and latencies are reported.

Each batch solve runs under a host timer that requests a stop, so a budget overrun ends the
kernel at its next stop poll."""
import os
import re
import subprocess
import threading
import time

import numpy as np
import pytest
from conftest import ROOT, golden

pytestmark = pytest.mark.gpu

MODEL = os.path.join(ROOT, "dds_gpu_amd", "models", "tricknet.bin")


def log(*a):
    print(*a, flush=True)


def solve_budget(g, hands, seconds, seeds=None):
    timer = threading.Timer(seconds, g.request_stop)
    timer.start()
    t0 = time.time()
    try:
        out, st = g.solve(hands, seeds=seeds, allow_stop=True)
    finally:
        timer.cancel()
    return out, st, time.time() - t0


def check_finished(out, want, label):
    done = out != 255
    wrong = int(((out != want) & done).sum())
    frac = done.mean()
    whole = int(done.all(axis=1).sum())
    log("%s: %d of %d answers finished (%.2f%%), %d whole tables, %d wrong" %
        (label, int(done.sum()), out.size, 100 * frac, whole, wrong))
    assert wrong == 0
    return frac


def test_c2_subset_plain_kernel_budget(ddx):
    g = golden("full_2000.npz")
    with ddx.GpuSolver(lanes=65536, tt_log2=10, jobs_per_deal=5, schedule=ddx.SCHED_SPLIT) as s:
        out, st, secs = solve_budget(s, g["hands"], 90)
    frac = check_finished(out, g["tables"], "C2 2,000 deals, 65,536 lanes, TT_LOG2 10, %.0f s" % secs)
    assert frac > 0.9


def test_c3_mixed_batch_graph_replays(ddx):
    mixed = golden("mixed_endings.npz")
    full = golden("full_2000.npz")
    cheap = np.argsort(full["nodes"].astype(np.int64).sum(axis=1))[:300]
    hands = np.concatenate([mixed["hands"], full["hands"][cheap]])
    want = np.concatenate([mixed["tables"], full["tables"][cheap]])
    order = np.random.default_rng(33).permutation(len(hands))
    hands, want = hands[order], want[order]
    with ddx.GpuSolver(lanes=65536, tt_log2=10, jobs_per_deal=5, schedule=ddx.SCHED_SPLIT,
                       waves_per_simd=3, graph=True) as s:
        for replay in range(3):
            out, st, secs = solve_budget(s, hands, 120)
            log("C3 replay %d: %d deals (K = 5, 7, 9, 11, 13), %.1f s, %d nodes" % (replay, len(hands), secs, st.nodes))
            assert st.jobs_done == 5 * len(hands)
            assert np.array_equal(out, want)


def test_c4_learned_subset_budget(ddx):
    g = golden("full_2000.npz")
    net = ddx.TrickNet(path=MODEL)
    seeds = net.predict(g["hands"])
    net.close()
    with ddx.GpuSolver(lanes=196608, tt_log2=11, jobs_per_deal=5, schedule=ddx.SCHED_SPLIT,
                       waves_per_simd=3, seed_mode=ddx.SEED_CERT) as s:
        out, st, secs = solve_budget(s, g["hands"], 90, seeds=seeds)
    frac = check_finished(out, g["tables"], "C4 2,000 deals seeded by the net, 196,608 lanes, TT_LOG2 11, %.0f s" % secs)
    assert frac > 0.9


def test_c5_stream_full_deals_poisson():
    exe = os.path.join(ROOT, "dds_gpu_amd", "bin", "dd_stream")
    env = dict(os.environ, THREADS="32768", RATE="10", RING="4096", SEED="77")
    r = subprocess.run([exe, "bench", "40", "10", "4"], capture_output=True, text=True, timeout=300, env=env)
    log(r.stderr.strip().splitlines()[-3:])
    assert r.returncode == 0, r.stderr
    m = re.search(r"CPU check: (\d+) wrong of (\d+)", r.stderr)
    assert m and m.group(1) == "0" and int(m.group(2)) > 0
    assert re.search(r"latency ms: p50 \S+ p99 \S+", r.stderr)
