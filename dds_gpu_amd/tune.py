"""Autotuner for the batch solver on the local GPU.

    python -m dds_gpu_amd.tune [--jobs 200000] [--full-share 0.2] [--quick] [--out tuned.json]

Workload (the uneven-batch configuration): jobs mixing random endings of 5, 7, 9 and 11 cards
per hand and full 13-card deals in equal proportions (seeded, shuffled), 5 jobs per deal
(200,000 jobs = 40,000 deals by default). Every candidate configuration solves the same batch
to completion (tail included) through a captured HIP graph; its score is whole tables per
second, and every candidate's tables must equal the first one's (all are exact solvers, so
any difference is a bug and aborts the sweep).

Swept: threads per CU (lanes = 256 CUs x threads/CU) and lanes per search (1 or a 16-lane
team), TT_LOG2 in {9, 10, 11} for one lane per search and the same HBM per TT for teams
({13, 14, 15}), the stage schedule (priority / chain / split) and jobs per deal (5, 20).
The best configuration is written to dds_gpu_amd/tuned.json, which bench.py and the tools read.
"""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from dds_gpu_amd import ddx  # noqa: E402

CUS = 256


def workload(jobs, seed=3):
    deals = jobs // 5
    per = deals // 5
    parts = [ddx.random_deals(seed + k, per, k) for k in (5, 7, 9, 11, 13)]
    hands = np.concatenate(parts)
    rng = np.random.default_rng(seed)
    return hands[rng.permutation(len(hands))]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def candidates(quick):
    out = []
    for team, tpc in [(1, 256), (1, 512), (16, 256), (16, 512)]:
        tts = (9, 10, 11) if team == 1 else (13, 14, 15)
        for tt, sched, jpd in itertools.product(tts, (0, 1, 2), (5, 20)):
            if quick and (sched != 2 or tt not in (10, 14)):
                continue
            out.append({"lanes": CUS * tpc, "team": team, "tt_log2": tt, "schedule": sched,
                        "jobs_per_deal": jpd})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=200000)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--limit", type=float, default=600.0, help="seconds per candidate before it is dropped")
    ap.add_argument("--out", default=os.path.join(HERE, "tuned.json"))
    a = ap.parse_args()
    hands = workload(a.jobs)
    n = len(hands)
    log("workload: %d deals (%d jobs), endings of 5/7/9/11 cards and full deals" % (n, 5 * n))
    ref = None
    results = []
    for cfg in candidates(a.quick):
        g = ddx.GpuSolver(graph=True, **cfg)
        import threading
        timer = threading.Timer(a.limit, g.request_stop)
        timer.start()
        t0 = time.time()
        out, st = g.solve(hands, allow_stop=True)
        timer.cancel()
        secs = g.last_kernel_ms / 1e3
        g.close()
        done = bool((out != 255).all())
        if done and ref is None:
            ref = out
        if done and not np.array_equal(out, ref):
            raise SystemExit("tables differ between configurations: %s" % cfg)
        r = dict(cfg, seconds=round(secs, 2), tables_per_s=round(n / secs, 1) if done else 0.0,
                 finished=done, nodes=st.nodes)
        results.append(r)
        log(json.dumps(r))
    best = max((r for r in results if r["finished"]), key=lambda r: r["tables_per_s"])
    json.dump({"best": best, "all": results, "workload_deals": n}, open(a.out, "w"), indent=1)
    print(json.dumps(best))


if __name__ == "__main__":
    main()
