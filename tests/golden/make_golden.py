"""Writes the golden fixtures of tests/golden/ from the reference's own CPU solver.

Run in a container that has the reference tree (default /root/reference):

    python tests/golden/make_golden.py [--ref /root/reference] [--only NAME]

It builds oracle/_ref (oracle/Makefile compiles the reference sources where they lie) and
runs oracle/_ref/ref_harness, a harness of ours around the reference's header-only API. The
fixtures hold data only -- deals, tables, node counts, per-function outputs -- never
reference code. Deals come from this project's seeded generator (ddx_random_deals), which
the tests regenerate and compare before use.

Fixtures:
  c1_endings.npz   4,096 random 8-card endings (seed 1): reference tables at TT_LOG2 10 and
                   the nodes of every (deal, strain) job on one context in index order; plus
                   the same with step-from-guess seeds (dd_solve_strain_guess).
  full_2000.npz    the first 2,000 full deals of the bench stream (seed 2026): reference
                   tables at TT_LOG2 10 (8 chunks of 250 deals, one context per chunk, job
                   nodes per chunk).
  funcs.npz        6,000 random positions with the outputs of the reference's key, bound and
                   move-ordering helpers (layout in oracle/ref_harness.cpp).
  dds_4000.npy     rows 0..3999 of the reference fixture data/deals_20k.npy (Pgx DDS tables;
                   Apache-2.0, Pgx authors), kept as data.
  mixed_endings.npz  300 random endings each of 5, 7, 9 and 11 cards per hand (seed 3000 + K),
                   reference tables at TT_LOG2 10: the endings of the uneven (mixed) batch.
  sig_entries.npz  the reference built with 12-byte signature TT entries (-DDD_TT_SIG,
                   oracle/_ref/ref_harness_sig), TT_LOG2 10, one context per group, jobs in
                   index order: tables and job nodes of the first 400 C1 endings, the first
                   100 11-card mixed endings and 12 full deals of full_2000.npz.
  shared_tt.npz    the same deals and layout from the reference built with its lock-free
                   shared TT (-DDD_SHARED_TT, oracle/_ref/ref_harness_shared): one TT of 2^10
                   buckets per group, kept across the group's deals.
  per_leader_tt.npz  the first 400 C1 endings solved by the reference with its measurement
                   switch -DDD_PER_LEADER_TT (no TT entries shared between the leads of a
                   strain; oracle/_ref/ref_harness_perlead), TT_LOG2 10: tables and job nodes.
"""
import argparse
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

C1_SEED, C1_K, C1_N = 1, 8, 4096
C2_SEED, FULL_N, FULL_CHUNK = 2026, 2000, 250
FUNCS_N, FUNCS_SEED = 6000, 77
MIXED_KS, MIXED_N, MIXED_SEED = (5, 7, 9, 11), 300, 3000
SIG_C1_N, SIG_K11_N = 400, 100
SIG_FULL = (0, 1, 2, 4, 5, 6, 7, 8, 10, 12, 14, 15)  # full_2000 rows below 3e6 reference nodes


def harness(ref, *args, exe="ref_harness"):
    exe = os.path.join(ROOT, "oracle", "_ref", exe)
    subprocess.check_call([exe] + [str(a) for a in args])


def deals(seed, first, n, k):
    from dds_gpu_amd import ddx
    return ddx.random_deals(seed, n, k, first)


def ref_tables(ref, hands, tt_log2, seeds=None, exe="ref_harness"):
    with tempfile.TemporaryDirectory() as d:
        hp, tp, npth = (os.path.join(d, x) for x in ("h.bin", "t.bin", "n.bin"))
        np.ascontiguousarray(hands, np.uint64).tofile(hp)
        args = ["tables", hp, len(hands), tt_log2, tp, npth]
        if seeds is not None:
            sp = os.path.join(d, "s.bin")
            np.ascontiguousarray(seeds, np.uint8).tofile(sp)
            args.append(sp)
        harness(ref, *args, exe=exe)
        t = np.fromfile(tp, np.uint8).reshape(-1, 20)
        nodes = np.fromfile(npth, np.uint64).reshape(-1, 5)
    return t, nodes


def noisy_seeds(tables, rng_seed):
    """NS-trick guesses = the true NS tricks +-1 (clipped), deterministic."""
    t = tables.reshape(-1, 5, 4).astype(np.int64)
    ns = t.copy()
    ns[:, :, 1::2] = 13 - t[:, :, 1::2]
    rng = np.random.default_rng(rng_seed)
    ns = np.clip(ns + rng.integers(-1, 2, ns.shape), 0, 13)
    return ns.reshape(-1, 20).astype(np.uint8)


def make_c1(ref):
    h = deals(C1_SEED, 0, C1_N, C1_K)
    t, nodes = ref_tables(ref, h, 10)
    seeds = noisy_seeds(t, 5)
    ts, nodes_s = ref_tables(ref, h, 10, seeds)
    assert np.array_equal(t, ts)
    np.savez_compressed(os.path.join(HERE, "c1_endings.npz"), hands=h, tables=t, nodes=nodes,
                        seeds=seeds, nodes_seeded=nodes_s, seed=C1_SEED, k=C1_K, tt_log2=10)


def make_full(ref):
    h = deals(C2_SEED, 0, FULL_N, 13)
    chunks = [h[i:i + FULL_CHUNK] for i in range(0, FULL_N, FULL_CHUNK)]
    with ThreadPoolExecutor(max_workers=len(chunks)) as ex:
        res = list(ex.map(lambda c: ref_tables(ref, c, 10), chunks))
    t = np.concatenate([r[0] for r in res])
    nodes = np.concatenate([r[1] for r in res])
    np.savez_compressed(os.path.join(HERE, "full_2000.npz"), hands=h, tables=t, nodes=nodes,
                        seed=C2_SEED, chunk=FULL_CHUNK, tt_log2=10)


def make_mixed(ref):
    hands, ks = [], []
    for k in MIXED_KS:
        hands.append(deals(MIXED_SEED + k, 0, MIXED_N, k))
        ks += [k] * MIXED_N
    h = np.concatenate(hands)
    chunks = [h[i:i + MIXED_N] for i in range(0, len(h), MIXED_N)]
    with ThreadPoolExecutor(max_workers=len(chunks)) as ex:
        res = list(ex.map(lambda c: ref_tables(ref, c, 10), chunks))
    t = np.concatenate([r[0] for r in res])
    np.savez_compressed(os.path.join(HERE, "mixed_endings.npz"), hands=h, tables=t, k=np.array(ks, np.int8),
                        seed=MIXED_SEED, tt_log2=10)


def make_variant(ref, exe, fixture):
    groups = {"c1": deals(C1_SEED, 0, SIG_C1_N, C1_K),
              "k11": deals(MIXED_SEED + 11, 0, SIG_K11_N, 11),
              "full": deals(C2_SEED, 0, FULL_N, 13)[list(SIG_FULL)]}
    with ThreadPoolExecutor(max_workers=len(groups)) as ex:
        res = dict(zip(groups, ex.map(lambda h: ref_tables(ref, h, 10, exe=exe), groups.values())))
    out = {}
    for name, h in groups.items():
        out[name + "_hands"] = h
        out[name + "_tables"], out[name + "_nodes"] = res[name]
    np.savez_compressed(os.path.join(HERE, fixture), full_rows=np.array(SIG_FULL), tt_log2=10, **out)


def make_sig(ref):
    make_variant(ref, "ref_harness_sig", "sig_entries.npz")


def make_shared(ref):
    make_variant(ref, "ref_harness_shared", "shared_tt.npz")


def make_perlead(ref):
    h = deals(C1_SEED, 0, SIG_C1_N, C1_K)
    t, nodes = ref_tables(ref, h, 10, exe="ref_harness_perlead")
    np.savez_compressed(os.path.join(HERE, "per_leader_tt.npz"), hands=h, tables=t, nodes=nodes, tt_log2=10)


def make_funcs(ref):
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.bin")
        harness(ref, "funcs", FUNCS_N, FUNCS_SEED, p)
        rec = np.fromfile(p, np.int64).reshape(FUNCS_N, -1)
    np.savez_compressed(os.path.join(HERE, "funcs.npz"), records=rec)


def make_dds(ref):
    a = np.load(os.path.join(ref, "data", "deals_20k.npy"), allow_pickle=False)
    np.save(os.path.join(HERE, "dds_4000.npy"), np.ascontiguousarray(a[:, :4000, :]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "REF=" + a.ref])
    jobs = {"c1": make_c1, "funcs": make_funcs, "dds": make_dds, "full": make_full, "mixed": make_mixed,
            "sig": make_sig, "shared": make_shared, "perlead": make_perlead}
    for name, fn in jobs.items():
        if a.only and name not in a.only.split(","):
            continue
        print("making", name, flush=True)
        fn(a.ref)


if __name__ == "__main__":
    main()
