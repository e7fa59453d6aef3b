#!/usr/bin/env python3
"""Benchmark: full-deal double-dummy tables per second on one MI355X, learned-seeded vs plain.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--step-seconds S] [--deals D]
                    [--legs plain,tuned,learned] [--model PATH] [--no-cpu-baseline]
                    [--variant sig|shared] [--tt-log2 N]

Workload: the seeded stream 2026 of full 52-card deals (uniform Fisher-Yates, 13 cards per
hand), D = 100,000 deals resident in HBM, 20 results per deal (uint8). Every leg launches
one persistent solve over the whole batch; the lanes pull (deal, strain) or (deal, lead) jobs
from an atomic queue. A "step" is S seconds of that running solve, S = min(--step-seconds,
--leg-seconds / (W + K)): W untimed warmup steps, then K timed steps; the live counters
(finished jobs, nodes) are read at the step boundaries, then the solve is stopped and
synchronized. A 100,000-deal batch holds ~1e12 search nodes with a heavy tail (single jobs of
up to ~1e9 nodes), far more than a bench can run to completion, so:
  value        whole tables -- deals whose 20 answers are all written -- when the stop is
               requested, over the seconds from the launch to the stop (the ramp from a cold
               start included; the tail still running at the stop not counted; each leg also
               reports the jobs unfinished at the stop)
  window_rate  jobs finished in the K timed steps, 1/jobs-per-deal of a table each (the
               reference's own measure, README.md:166-176), which decays as heavy jobs pile up

Legs (the same deals and queue order in each):
  plain    the plain wavefront kernel: 65,536 lanes, one lane per search, TT_LOG2 10 per lane,
           5 jobs per deal, binary search from 7 (DDX_SEED_PLAIN)
  tuned    the autotuned persistent configuration (dds_gpu_amd/tuned.json), plain seeds
  learned  the tuned configuration seeded by the trick-count net (DDX_SEED_CERT): the net's
           GPU prediction for the whole batch is timed and its cost charged pro rata
`value` is the learned leg when a trained net exists (else the tuned / plain leg);
`speedup_vs_plain` = value / plain. --variant runs the legs on another TT build of the library
(libddx_sig.so: 12-byte signature entries; libddx_shared.so: one TT shared by every search, whose
TT_LOG2, set with --tt-log2, is the size of that one TT); the default is the exact build.

Checks and side measurements in the same JSON line:
  bit_exact     every solved answer of the first 2,000 deals, in every leg, against the
                reference CPU solver's tables (tests/golden/full_2000.npz); a mismatch exits 1
  roofline      the headline leg's solve kernel: algorithmic TT bytes per launch (64 B per probed
                tag line, 16 B per entry read, 144 B per store, 52 B of deal I/O per finished
                table) over the kernel's HIP-event duration, against the 8 TB/s HBM peak
  cpu_baseline  the reference's own CPU solver (oracle/_ref/dd_cpu, built from the reference
                sources) on a bounded sample of the same deals, all host cores, TT_LOG2 18
                (its default); it also re-checks that sample's fixture tables
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "full-deal DD tables/sec on 1 MI355X, learned-seeded vs plain, bit-exact; p99 ms"
STREAM_SEED = 2026
HBM_PEAK_GBS = 8000.0
# Algorithmic TT bytes by build: per entry a probe reads, and per store (the exact and signature
# stores read and rewrite the 64 B tag line and write one entry; the shared store reads the tag
# line and writes one 4 B tag and one entry).
TT_BYTES = {"": (16.0, 144.0), "sig": (12.0, 140.0), "shared": (32.0, 100.0)}
PLAIN = {"lanes": 65536, "team": 1, "tt_log2": 10, "jobs_per_deal": 5, "schedule": 2, "waves_per_simd": 0,
         "block": 256}
DEFAULT_MODEL = os.path.join(ROOT, "dds_gpu_amd", "models", "tricknet.bin")
DEFAULT_TUNED = os.path.join(ROOT, "dds_gpu_amd", "tuned.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--step-seconds", type=float, default=10.0,
                    help="longest step; a step is shortened so that one leg fits --leg-seconds")
    ap.add_argument("--leg-seconds", type=float, default=80.0,
                    help="wall budget of one leg's solve, warmup and timed steps together")
    ap.add_argument("--deals", type=int, default=100000)
    ap.add_argument("--legs", default="plain,tuned,learned")
    ap.add_argument("--model", default=DEFAULT_MODEL)
    ap.add_argument("--tuned", default=DEFAULT_TUNED)
    ap.add_argument("--cpu-sample", type=int, default=-1, help="deals for the CPU baseline (-1: auto)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variant", default="", choices=["", "sig", "shared"],
                    help="TT build of the library: exact (default), sig, shared")
    ap.add_argument("--tt-log2", type=int, default=0, help="TT_LOG2 of every leg (0: the legs' own)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 5 plain 1 GiB device-to-device copies (a known-size kernel "
                         "against which HBM byte counters are scaled)")
    return ap.parse_args()


def copy_calibration(torch):
    """Five 1 GiB device copies (torch's elementwise copy kernel): 1 GiB read + 1 GiB written
    per dispatch. Under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE they give the counters' scale."""
    x = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    log("copy calibration: 1 GiB copy %.3f ms = %.0f GB/s (read + write)" % (ms, 2 * (1 << 30) / ms / 1e6))
    del x, y
    return {"bytes_read": 1 << 30, "bytes_written": 1 << 30, "ms": round(ms, 4)}


def cpu_baseline(ddx, hands, tables, sample):
    """Times the reference's dd_cpu on `sample` deals; returns the JSON object or None. `tables`
    are written into the sample file's value field, so dd_cpu also checks them (exit 0 = all
    equal)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "dd_cpu")
    if not os.path.exists(exe):
        log("cpu_baseline: oracle/_ref/dd_cpu is not built; skipped")
        return None
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "sample.npy")
        ddx.pgx_write(path, hands[:sample], tables[:sample])
        t0 = time.time()
        r = subprocess.run([exe, path, "0", str(sample), str(cores), "18"], capture_output=True,
                           text=True, timeout=600)
        wall = time.time() - t0
    line = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else ""
    log("cpu_baseline:", line)
    try:
        secs = float(line.split(":")[1].split("s,")[0])
    except (IndexError, ValueError):
        secs = wall
    return {"value": round(sample / secs, 3), "unit": "tables/s", "cores": cores, "kind": "reference",
            "sample": "first %d deals of the same stream, run to completion by the reference dd_cpu "
                      "(cpu/cpu_main.cpp) built from the reference sources, %d threads, TT_LOG2 18 "
                      "(its default); %.1f s; its tables agree with the fixture's: %s" %
                      (sample, cores, secs, "yes" if r.returncode == 0 else "NO rc=%d" % r.returncode),
            "verified": r.returncode == 0}


def step_seconds(a):
    """Length of one step: --step-seconds, shortened so that warmup and timed steps of one leg
    fit --leg-seconds."""
    return min(a.step_seconds, a.leg_seconds / max(1, a.warmup + a.steps))


def run_leg(ddx, torch, name, cfg, d_hands, n, a, gold, seeds=None, predict_s=0.0):
    """One persistent solve over the batch: W warmup steps, K timed steps of S seconds."""
    dev = d_hands.device
    d_out = torch.empty((n, 20), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    g = ddx.GpuSolver(lanes=cfg["lanes"], tt_log2=cfg["tt_log2"], jobs_per_deal=cfg["jobs_per_deal"],
                      schedule=cfg["schedule"], team=cfg["team"], waves_per_simd=cfg["waves_per_simd"],
                      block=cfg.get("block", 256), variant=cfg.get("variant", ""),
                      seed_mode=ddx.SEED_CERT if seeds is not None else ddx.SEED_PLAIN)
    jpd = cfg["jobs_per_deal"]
    log("[%s] %s, %.1f GB of lane state" % (name, json.dumps(cfg), g.device_bytes / 1e9))
    torch.cuda.synchronize()
    g.solve_async(d_hands.data_ptr(), n, d_out.data_ptr(),
                  d_seeds=seeds.data_ptr() if seeds is not None else 0, stream=stream)
    t_launch = time.perf_counter()
    marks = [(t_launch, 0, 0)]
    step_s = step_seconds(a)
    for i in range(a.warmup + a.steps):
        target = t_launch + (i + 1) * step_s
        while time.perf_counter() < target:
            time.sleep(min(0.05, max(0.0, target - time.perf_counter())))
        j, nodes = g.progress_nodes()
        marks.append((time.perf_counter(), j, nodes))
        t, j0, n0 = marks[-2]
        log("[%s] %s step %d: %.1f tables/s, %.3g nodes/s (%d jobs so far)" %
            (name, "warmup" if i < a.warmup else "timed", i, (j - j0) / jpd / (marks[-1][0] - t),
             (nodes - n0) / (marks[-1][0] - t), j))
    g.request_stop()
    t_stop = time.perf_counter()
    st = g.wait(allow_stop=True)
    torch.cuda.synchronize()
    kms = g.last_kernel_ms
    g.close()
    (t0, j0, nd0), (t1, j1, nd1) = marks[a.warmup], marks[-1]
    window_tables = (j1 - j0) / jpd
    secs = t1 - t0
    full = d_out.cpu().numpy()
    # Whole tables: deals whose 20 answers were all written when the stop was requested
    # (answers written during the few ms the stop takes to reach every wave count too).
    whole = int((full != 255).all(axis=1).sum())
    from_launch = t_stop - t_launch
    charge = predict_s * whole / n  # the net's share for the tables counted
    value = whole / (from_launch + charge)
    window_rate = window_tables / (secs + predict_s * window_tables / n)
    out = full[:len(gold["tables"])]
    solved = out != 255
    wrong = int(((out != gold["tables"]) & solved).sum())
    ent_b, store_b = TT_BYTES[cfg.get("variant", "")]
    alg = 64.0 * st.tt_probes + ent_b * st.tt_entry_reads + store_b * st.tt_stores + 52.0 * st.jobs_done / jpd
    achieved = alg / (kms / 1e3) / 1e9
    res = {"value": round(value, 2), "config": dict(cfg, seed_mode="cert" if seeds is not None else "plain"),
           "whole_tables": whole, "seconds_from_launch": round(from_launch, 3),
           "jobs_unfinished_at_stop": int(n * jpd - st.jobs_done),
           "window_rate": round(window_rate, 2), "window_s": round(secs, 3), "window_tables": window_tables,
           "nodes_per_s": round((nd1 - nd0) / secs, 0),
           "nodes_per_table": round((nd1 - nd0) / max(window_tables, 1e-9), 0),
           "predict_s_charged": round(charge, 5),
           "bit_exact": {"checked_answers": int(solved.sum()), "wrong": wrong},
           "kernel": {"ms": round(kms, 1), "alg_bytes": alg, "achieved_gbs": round(achieved, 3),
                      "nodes": st.nodes, "tt_probes": st.tt_probes, "tt_entry_reads": st.tt_entry_reads,
                      "tt_stores": st.tt_stores, "jobs_done": st.jobs_done}}
    log("[%s] %d whole tables in %.1f s from launch: %.2f tables/s (window rate %.2f); %d answers "
        "checked, %d wrong" % (name, whole, from_launch, value, window_rate, solved.sum(), wrong))
    return res


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: deals shard across GPUs with no exchange (each GPU "
              "would take its own range of the batch), but only --gpus 1 is run here", file=sys.stderr)
        sys.exit(2)
    import torch
    from dds_gpu_amd import ddx

    if not torch.cuda.is_available():
        log("bench.py needs a GPU")
        sys.exit(1)
    n = a.deals
    hands = ddx.random_deals(STREAM_SEED, n, 13)
    gold = np.load(os.path.join(ROOT, "tests", "golden", "full_2000.npz"))
    m = min(n, len(gold["tables"]))
    assert np.array_equal(gold["hands"][:m], hands[:m]), "deal stream differs from the fixture"
    gold = {"tables": gold["tables"][:m]}
    dev = torch.device("cuda:0")
    d_hands = torch.from_numpy(hands.view(np.int64)).to(dev)

    tuned = dict(PLAIN)
    if os.path.exists(a.tuned):
        b = json.load(open(a.tuned))["best"]
        tuned = {k: b.get(k, PLAIN[k]) for k in PLAIN}
    plain = dict(PLAIN)
    for cfg in (plain, tuned):
        if a.variant:
            cfg["variant"] = a.variant
        if a.tt_log2:
            cfg["tt_log2"] = a.tt_log2
    legs = a.legs.split(",")
    have_model = os.path.exists(a.model)
    if "learned" in legs and not have_model:
        log("no trained net at %s: learned leg skipped" % a.model)
        legs.remove("learned")
    calib = copy_calibration(torch) if a.copy_calibration else None
    results = {}
    for name in legs:
        if name == "plain":
            results[name] = run_leg(ddx, torch, name, plain, d_hands, n, a, gold)
        elif name == "tuned":
            results[name] = run_leg(ddx, torch, name, tuned, d_hands, n, a, gold)
        elif name == "learned":
            net = ddx.TrickNet(path=a.model)
            d_seeds = torch.empty((n, 20), dtype=torch.uint8, device=dev)
            stream = torch.cuda.current_stream().cuda_stream
            net.predict_device(d_hands.data_ptr(), n, d_seeds.data_ptr(), stream=stream)  # warm
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            net.predict_device(d_hands.data_ptr(), n, d_seeds.data_ptr(), stream=stream)
            e1.record()
            torch.cuda.synchronize()
            predict_s = e0.elapsed_time(e1) / 1e3
            log("[learned] net prediction for %d deals: %.1f ms" % (n, predict_s * 1e3))
            r = run_leg(ddx, torch, name, tuned, d_hands, n, a, gold, seeds=d_seeds, predict_s=predict_s)
            r["predict_ms_whole_batch"] = round(predict_s * 1e3, 2)
            results[name] = r
            net.close()
    head = next(k for k in ("learned", "tuned", "plain") if k in results)
    H = results[head]
    kern = H["kernel"]
    # HBM bytes per launch of this leg's kernel, measured by rocprofv3 --pmc FETCH_SIZE /
    # WRITE_SIZE over this command's learned leg and scaled by a 1 GiB copy's counts (the
    # committed profile names the command); used only when its configuration is this leg's, and
    # scaled by this launch's duration over the profiled one's (1.0 at --steps 20 --warmup 5).
    traffic, extra = None, {}
    tpath = os.path.join(ROOT, "profiles", "solve_kernel_hbm_traffic.json")
    if os.path.exists(tpath):
        tr = json.load(open(tpath))
        if tr.get("config") == H["config"] and tr.get("deals") == n and tr.get("launch_ms"):
            traffic = round(tr["bytes_per_launch"] * kern["ms"] / tr["launch_ms"])
            extra = {k: tr[k] for k in ("l2_hit_rate", "valu_busy", "traffic_source") if k in tr}
            extra["traffic_profiled_launch_ms"] = tr["launch_ms"]
    roofline = {"bound": "hbm", "achieved": kern["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(kern["achieved_gbs"] / HBM_PEAK_GBS, 6), "traffic": traffic,
                "kernel": "solve_kernel", "leg": head, **extra,
                "note": "integer tree search: TT probes are dependent random 64 B reads and the "
                        "stage code diverges across lanes; the kernel is bound by instruction "
                        "issue (VALU busy) under divergence, far below the HBM peak"}
    cpu = None
    if not a.no_cpu_baseline:
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        # ~24 full deals per host thread: 10-15 s of the reference solver's work on its cores
        sample = a.cpu_sample if a.cpu_sample > 0 else min(m, 24 * cores)
        cpu = cpu_baseline(ddx, hands, gold["tables"], sample)
    wrong = sum(r["bit_exact"]["wrong"] for r in results.values())
    checked = sum(r["bit_exact"]["checked_answers"] for r in results.values())
    res = {
        "metric": METRIC, "value": H["value"], "unit": "tables/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(step_seconds(a) * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic: seeded uniform full deals (stream %d)" % STREAM_SEED,
        "config": {"workload": "%d full deals resident in HBM; one persistent solve per leg, stopped "
                               "after %d warmup + %d timed steps of %.2f s; value = whole tables (all 20 "
                               "answers) finished from launch to the stop / those seconds; headline "
                               "leg '%s'" % (n, a.warmup, a.steps, step_seconds(a), head),
                   "deals": n, "legs": {k: r["config"] for k, r in results.items()}},
        "legs": {k: {kk: v for kk, v in r.items() if kk != "config"} for k, r in results.items()},
        "speedup_vs_plain": round(H["value"] / results["plain"]["value"], 3) if "plain" in results else None,
        "bit_exact": {"checked_answers": checked, "wrong": wrong,
                      "against": "reference CPU solver tables of the first %d deals" % m},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if calib:
        res["copy_calibration"] = calib
    print(json.dumps(res), flush=True)
    if wrong or (cpu and not cpu["verified"]):
        sys.exit(1)


if __name__ == "__main__":
    main()
