"""Solved deals for training the trick-count net.

    python -m dds_gpu_amd.make_labels OUT.bin --seed 7001 --first 0 --deals 20000 [--gpu]
                                      [--time-limit S] [--from-npz DIR]

Deals come from the seeded generator (stream --seed, indices [first, first + deals)); tables
from the exact solver (the GPU with 16-lane teams, or the CPU). With --time-limit only deals
solved within the limit are written, so a batch's slowest deals can be left out. --from-npz
converts chunks written by an earlier run (npz with `hands`, `tables`). Records: 4 x uint64
hands + 20 uint8 tables (52 bytes), appended to OUT.bin.
"""
import argparse
import glob
import os
import sys
import threading

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
        with ddx.GpuSolver(lanes=65536, tt_log2=14, team=16, jobs_per_deal=20, schedule=2) as g:
            timer = threading.Timer(a.time_limit, g.request_stop) if a.time_limit > 0 else None
            if timer:
                timer.start()
            tables, st = g.solve(hands, allow_stop=True)
            if timer:
                timer.cancel()
    else:
        tables, st = ddx.cpu_solve(hands, threads=os.cpu_count(), tt_log2=14)
    done = (tables != 255).all(axis=1)
    write(a.out, hands[done], tables[done])
    print("%d of %d deals solved and written (%.3g nodes)" % (done.sum(), a.deals, st.nodes))


if __name__ == "__main__":
    main()
