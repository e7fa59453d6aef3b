"""Per-stage profile of the wavefront kernel on the local GPU: active-lane fraction and share of
shader cycles for each stage of the search, and where node entries decide.

    python -m dds_gpu_amd.stage_profile [--seconds 20] [--lanes 65536] [--tt-log2 10]
                                        [--schedule 2] [--jobs-per-deal 5] [--out FILE.json]

Runs the profiling build of the batch kernel (config.profile = 1: wave-uniform counters kept
per stage, at some cost in speed) over the bench's deal stream for `seconds`, then stops it.
Per stage kind: rounds (a round = one pass of a wave over that stage), the mean number of
active lanes per round out of 64 (divergence: the lanes of a wave that run the stage while
the others wait), and the share of the kernel's shader cycles spent in it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from dds_gpu_amd import ddx  # noqa: E402

KINDS = {0: "JOB", 1: "ROOT", 2: "ENTER (trick start)", 3: "NEXT", 4: "RET",
         6: "cheap-stage chain (with follow-node ENTER under SPLIT)"}
EXITS = ["trivial bound", "last trick", "TT", "quick tricks", "later tricks", "2nd-hand quick tricks",
         "moves generated"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seconds", type=float, default=20)
    ap.add_argument("--deals", type=int, default=100000)
    ap.add_argument("--lanes", type=int, default=65536)
    ap.add_argument("--tt-log2", type=int, default=10)
    ap.add_argument("--schedule", type=int, default=2)
    ap.add_argument("--jobs-per-deal", type=int, default=5)
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    hands = ddx.random_deals(2026, a.deals, 13)
    g = ddx.GpuSolver(lanes=a.lanes, tt_log2=a.tt_log2, schedule=a.schedule, jobs_per_deal=a.jobs_per_deal,
                      profile=True)
    g.profile(reset=True)
    import threading
    timer = threading.Timer(a.seconds, g.request_stop)
    timer.start()
    t0 = time.time()
    _, st = g.solve(hands, allow_stop=True)
    timer.cancel()
    v = g.profile(reset=True).astype(np.float64)
    g.close()
    cycles = v[24:32].sum()
    stages = {}
    for k, name in KINDS.items():
        if v[k] == 0:
            continue
        stages[name] = {"rounds": int(v[k]), "active_lanes_per_round": round(v[8 + k] / v[k], 2),
                        "active_fraction": round(v[8 + k] / v[k] / 64, 4),
                        "cycle_share": round(v[24 + k] / cycles, 4) if cycles else None}
    exits = {EXITS[i]: int(v[16 + i]) for i in range(len(EXITS))}
    doc = {"config": {"lanes": a.lanes, "tt_log2": a.tt_log2, "schedule": a.schedule,
                      "jobs_per_deal": a.jobs_per_deal, "seconds": round(time.time() - t0, 1)},
           "nodes": st.nodes, "jobs_done": st.jobs_done, "stages": stages, "node_entries_decided_at": exits}
    print(json.dumps(doc, indent=1))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(doc, f, indent=1)


if __name__ == "__main__":
    main()
