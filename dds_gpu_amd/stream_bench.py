"""Streaming latency on the local GPU: dd_stream bench at Poisson arrival rates.

    python -m dds_gpu_amd.stream_bench --capacity C [--loads 0.5,0.8,1.0,1.2] [--seconds 60]
        [--lanes 65536] [--team 16] [--tt-log2 12] [--waves 0|3] [--model dds_gpu_amd/models/tricknet.bin]
        [--out FILE.json]

For each load L, dd_stream runs its bench for `seconds` with seeded random full deals arriving
as a Poisson process at L x C deals/s into the 4,096-slot pinned ring (32-byte records in,
28-byte records out of order), once with the plain search and once seeded by the trained net
(skipped when there is no model). Recorded per run: solved deals/s, the p50 and p99 latency
from arrival to result over the deals that arrived in the first half of the run (">=" when the
percentile falls on a deal still unsolved at the end), and the CPU re-check of every 50th deal.
"""
import argparse
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TOOL = os.path.join(HERE, "bin", "dd_stream")


def run(seconds, tt_log2, env):
    e = dict(os.environ, **{k: str(v) for k, v in env.items()})
    p = subprocess.Popen([TOOL, "bench", str(seconds), str(tt_log2), "50"], env=e, stderr=subprocess.PIPE,
                         stdout=subprocess.DEVNULL, text=True)
    lines = []
    for line in p.stderr:  # pass the tool's progress lines through
        print("   ", line.rstrip(), file=sys.stderr, flush=True)
        lines.append(line)
    rc = p.wait()
    text = "".join(lines)
    m = re.search(r"(\d+) deals solved in ([\d.]+) s \(([\d.]+) deals/s\), CPU check: (\d+) wrong of (\d+)", text)
    l = re.search(r"latency ms: p50 (\S+) p99 (\S+) \(arrival to result, (\d+) deals arrived by (\d+) s, "
                  r"(\d+) unsolved\)", text)
    if rc not in (0, 1) or not m or not l:
        raise SystemExit("dd_stream failed (rc %d):\n%s" % (rc, text[-2000:]))
    return {"deals_per_s": float(m.group(3)), "solved": int(m.group(1)), "seconds": float(m.group(2)),
            "checked": int(m.group(5)), "wrong": int(m.group(4)), "p50_ms": l.group(1), "p99_ms": l.group(2),
            "latency_deals": int(l.group(3)), "unsolved_at_end": int(l.group(5)), "rc": rc}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--capacity", type=float, required=True, help="deals/s the solver sustains")
    ap.add_argument("--loads", default="0.5,0.8,1.0,1.2")
    ap.add_argument("--seconds", type=float, default=60)
    ap.add_argument("--lanes", type=int, default=65536)
    ap.add_argument("--team", type=int, default=16)
    ap.add_argument("--tt-log2", type=int, default=12)
    ap.add_argument("--waves", type=int, default=0, help="3: kernel held to 3 waves per SIMD")
    ap.add_argument("--model", default=os.path.join(HERE, "models", "tricknet.bin"))
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    results = []
    for load in [float(x) for x in a.loads.split(",")]:
        for seeded in ([False, True] if os.path.exists(a.model) else [False]):
            env = {"THREADS": a.lanes, "TEAM": a.team, "RATE": load * a.capacity, "PER_LEAD": 1, "WAVES": a.waves}
            if seeded:
                env["MODEL"] = a.model
            print("== load %.2f (%.1f deals/s), %s" % (load, load * a.capacity, "learned" if seeded else "plain"),
                  file=sys.stderr, flush=True)
            r = dict(load=load, rate=round(load * a.capacity, 2), learned=seeded, **run(a.seconds, a.tt_log2, env))
            print(json.dumps(r), flush=True)
            results.append(r)
    doc = {"capacity": a.capacity, "lanes": a.lanes, "team": a.team, "tt_log2": a.tt_log2, "waves_per_simd": a.waves,
           "seconds": a.seconds, "runs": results}
    if a.out:
        with open(a.out, "w") as f:
            json.dump(doc, f, indent=1)
    if any(r["wrong"] for r in results):
        sys.exit(1)


if __name__ == "__main__":
    main()
