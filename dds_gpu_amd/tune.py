"""Autotuner for the persistent batch solver on the local GPU.

    python -m dds_gpu_amd.tune [--workload full|mixed] [--deals N] [--warmup-s 10] [--measure-s 10]
                               [--quick] [--out dds_gpu_amd/tuned.json]

Workloads:
  full   N (default 100,000) seeded full 52-card deals (the bench's deal stream)
  mixed  the uneven batch: 200,000 jobs (40,000 deals at 5 jobs per deal) mixing random endings
         of 5, 7, 9 and 11 cards per hand and full deals in equal proportions, seeded, shuffled
Each candidate launches one persistent solve over the whole batch (atomic job queue, captured
HIP graph launch), runs warmup-s seconds untimed, then measure-s seconds timed from the live
counters, then stops. Three scores, as tables per second:
  whole   whole tables (all 20 answers written) when the stop is requested, over the seconds
          from the launch -- the bench's value; the default
  window  jobs finished in the timed window, a job counting 1/jobs_per_deal of a table; it
          decays as heavy jobs pile up on the lanes, so it favours draining easy jobs
  launch  jobs finished from the launch to the end of the window, over that time Every answer a candidate finishes is compared with the
first candidate's answers for the same deals (all candidates are exact solvers; a difference
is a bug and aborts the sweep).

Swept: lanes (threads per CU x 256 CUs, with the kernel held to 3 or 4 waves per SIMD where
the threads need it), lanes per search (one lane, or a 16-lane team), TT_LOG2
per search (one lane: 9..12, i.e. 0.16-1.3 MB of TT per search and up to 258 GB in all of the
288 GB; a team: 13..14), jobs per deal (5 = one job per strain, 20 = one per lead) and the stage schedule
(priority / chain / split, or none: the per-thread kernel). The best is written to `out` under the workload's name; bench.py
and the tools read the "full" entry.
"""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np
import torch  # before the ddx library: one HIP runtime for both (torch holds the device buffers)

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from dds_gpu_amd import ddx  # noqa: E402

CUS = 256


def workload(kind, deals, seed=3):
    if kind == "full":
        return ddx.random_deals(2026, deals, 13)
    per = deals // 5
    parts = [ddx.random_deals(seed + k, per, k) for k in (5, 7, 9, 11, 13)]
    hands = np.concatenate(parts)
    rng = np.random.default_rng(seed)
    return hands[rng.permutation(len(hands))]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def candidates(quick):
    """(threads per CU, waves per SIMD it needs) x team x TT_LOG2 x schedule x jobs per deal.
    256 threads per CU is one wave per SIMD; 512 two (the compiler's allocation); 768 and
    1024 need the kernels held to 3 or 4 waves per SIMD."""
    out = []
    for team, tpc, w in [(1, 256, 0), (1, 512, 0), (1, 768, 3), (1, 1024, 4), (16, 512, 0), (16, 768, 3)]:
        tts = (9, 10, 11, 12) if team == 1 else (13, 14)
        for tt, sched, jpd in itertools.product(tts, (0, 1, 2, 3), (5, 20)):
            if (quick or w) and sched != 2:
                continue
            if quick and (tpc == 256 or team == 16 or jpd == 20):
                continue  # measured far behind on full deals: one wave per SIMD; teams; jobs per lead
            if team == 1 and tpc * CUS * 320 * (1 << tt) > 260e9:
                continue  # keep the TTs within ~260 GB of the 288 GB (the rest: stacks, batch, net)
            out.append({"lanes": CUS * tpc, "team": team, "tt_log2": tt, "schedule": sched,
                        "jobs_per_deal": jpd, "waves_per_simd": w})
    return out


def measure(cfg, hands, warmup_s, measure_s):
    """Sustained tables/s of one candidate; returns (score, answers, details)."""
    n = len(hands)
    dev = torch.device("cuda:0")
    d_hands = torch.from_numpy(hands.view(np.int64)).to(dev)
    d_out = torch.empty((n, 20), dtype=torch.uint8, device=dev)
    g = ddx.GpuSolver(graph=True, **cfg)
    torch.cuda.synchronize()
    g.solve_async(d_hands.data_ptr(), n, d_out.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    t0 = time.perf_counter()
    time.sleep(warmup_s)
    j0, n0 = g.progress_nodes()
    t1 = time.perf_counter()
    time.sleep(measure_s)
    j1, n1 = g.progress_nodes()
    t2 = time.perf_counter()
    g.request_stop()
    t_stop = time.perf_counter()
    g.wait(allow_stop=True)
    g.close()
    out = d_out.cpu().numpy()
    jpd = cfg["jobs_per_deal"]
    window = (j1 - j0) / jpd / (t2 - t1)
    launch = j1 / jpd / (t2 - t0)
    whole = int((out != 255).all(axis=1).sum()) / (t_stop - t0)
    return window, out, {"whole_tables_per_s": round(whole, 2), "tables_per_s_from_launch": round(launch, 2),
                         "nodes_per_s": round((n1 - n0) / (t2 - t1)), "warm_jobs": int(j0),
                         "timed_jobs": int(j1 - j0), "startup_s": round(t1 - t0 - warmup_s, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="full", choices=["full", "mixed"])
    ap.add_argument("--deals", type=int, default=0, help="default: 100,000 full / 40,000 mixed")
    ap.add_argument("--warmup-s", type=float, default=10.0)
    ap.add_argument("--measure-s", type=float, default=10.0)
    ap.add_argument("--quick", action="store_true", help="split schedule only")
    ap.add_argument("--candidates", default="",
                    help="a JSON list of configurations to measure instead of the built-in sets")
    ap.add_argument("--score", default="whole", choices=["whole", "window", "launch"])
    ap.add_argument("--out", default=os.path.join(HERE, "tuned.json"))
    a = ap.parse_args()
    deals = a.deals or (100000 if a.workload == "full" else 40000)
    hands = workload(a.workload, deals)
    log("workload %s: %d deals" % (a.workload, len(hands)))
    ref = None
    results = []
    for cfg in (json.loads(a.candidates) if a.candidates else candidates(a.quick)):
        score, out, det = measure(cfg, hands, a.warmup_s, a.measure_s)
        if ref is None:
            ref = out
        both = (out != 255) & (ref != 255)
        if (out[both] != ref[both]).any():
            raise SystemExit("answers differ between configurations: %s" % cfg)
        ref = np.where(ref == 255, out, ref)
        r = dict(cfg, tables_per_s=round(score, 2), **det)
        results.append(r)
        log(json.dumps(r))
    score = a.score
    key = {"whole": "whole_tables_per_s", "window": "tables_per_s", "launch": "tables_per_s_from_launch"}[score]
    best = max(results, key=lambda r: r[key])
    doc = json.load(open(a.out)) if os.path.exists(a.out) else {}
    doc[a.workload] = {"best": best, "score": score, "all": results, "deals": len(hands),
                       "warmup_s": a.warmup_s, "measure_s": a.measure_s}
    if a.workload == "full":
        doc["best"] = best
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps(best))


if __name__ == "__main__":
    main()
