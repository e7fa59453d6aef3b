"""Solved deals for training the trick-count net.

    python -m dds_gpu_amd.make_labels OUT.bin --seed 7001 --first 0 --deals 20000 [--gpu]
                                      [--time-limit S] [--partial] [--from-npz DIR]
                                      [--lanes L] [--waves 0|3] [--tt-log2 T] [--jobs-per-deal 5|20]

Deals come from the seeded generator (stream --seed, indices [first, first + deals)); tables
from the exact solver (the GPU, or the CPU). With --time-limit the solve stops after S seconds:
only whole tables are written, or with --partial every deal with at least one answer, the
unsolved answers left as 255 (the trainer skips them). --from-npz converts chunks written by
an earlier run (npz with `hands`, `tables`). Records: 4 x uint64 hands + 20 uint8 tables
(52 bytes), appended to OUT.bin.
"""
import argparse
import glob
import os
import sys
import threading
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from dds_gpu_amd import ddx  # noqa: E402


def write(path, hands, tables):
    rec = np.zeros(len(hands), dtype=[("h", "<u8", (4,)), ("t", "u1", (20,))])
    rec["h"], rec["t"] = hands, tables
    with open(path, "ab") as f:
        rec.tofile(f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--seed", type=int, default=7001)
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--deals", type=int, default=20000)
    ap.add_argument("--gpu", action="store_true")
    ap.add_argument("--time-limit", type=float, default=0)
    ap.add_argument("--from-npz", default="")
    ap.add_argument("--partial", action="store_true")
    ap.add_argument("--lanes", type=int, default=65536)
    ap.add_argument("--team", type=int, default=1)
    ap.add_argument("--tt-log2", type=int, default=12)
    ap.add_argument("--jobs-per-deal", type=int, default=20)
    ap.add_argument("--waves", type=int, default=0, help="3: the kernel held to 3 waves per SIMD")
    a = ap.parse_args()
    if a.from_npz:
        files = sorted(glob.glob(os.path.join(a.from_npz, "*.npz")))
        for fn in files:
            z = np.load(fn, allow_pickle=False)
            write(a.out, z["hands"], z["tables"])
        print("converted %d chunks" % len(files))
        return
    hands = ddx.random_deals(a.seed, a.deals, 13, a.first)
    if a.gpu:
        with ddx.GpuSolver(lanes=a.lanes, tt_log2=a.tt_log2, team=a.team, jobs_per_deal=a.jobs_per_deal,
                           schedule=2, waves_per_simd=a.waves) as g:
            res = {}
            worker = threading.Thread(target=lambda: res.update(r=g.solve(hands, allow_stop=True)))
            worker.start()
            t0 = time.time()
            stopped = False
            while worker.is_alive():  # progress lines while the solve runs
                worker.join(timeout=30)
                el = time.time() - t0
                print("  %.0f s: %d jobs done" % (el, g.progress()), flush=True)
                if a.time_limit > 0 and el > a.time_limit and not stopped:
                    g.request_stop()
                    stopped = True
            tables, st = res["r"]
    else:
        tables, st = ddx.cpu_solve(hands, threads=os.cpu_count(), tt_log2=14)
    done = (tables != 255).any(axis=1) if a.partial else (tables != 255).all(axis=1)
    write(a.out, hands[done], tables[done])
    print("%d of %d deals written, %d answers (%.3g nodes)" % (done.sum(), a.deals,
                                                               (tables[done] != 255).sum(), st.nodes))


if __name__ == "__main__":
    main()
