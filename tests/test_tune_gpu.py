"""The autotuner end to end on the GPU (dds_gpu_amd/tune.py, run as its own process so that torch
and the library share one HIP runtime): two small candidates over 4,000 full deals, every
answer of both checked against each other, the best written to a JSON file of the shape
bench.py and the tools read."""
import json
import os
import subprocess
import sys

import pytest
from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_tune_sweep_end_to_end(tmp_path):
    out = str(tmp_path / "tuned.json")
    cands = [{"lanes": 16384, "team": 1, "tt_log2": 10, "schedule": 2, "jobs_per_deal": 5, "waves_per_simd": 0},
             {"lanes": 16384, "team": 1, "tt_log2": 11, "schedule": 0, "jobs_per_deal": 5, "waves_per_simd": 0}]
    r = subprocess.run([sys.executable, "-m", "dds_gpu_amd.tune", "--workload", "full", "--deals", "4000",
                        "--warmup-s", "2", "--measure-s", "3", "--candidates", json.dumps(cands), "--out", out],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    print(r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr[-3000:]
    doc = json.load(open(out))
    assert doc["full"]["score"] == "whole" and len(doc["full"]["all"]) == 2
    for res in doc["full"]["all"]:
        assert res["whole_tables_per_s"] > 0 and res["nodes_per_s"] > 0
    best = doc["best"]
    assert {k: best[k] for k in cands[0]} in cands
