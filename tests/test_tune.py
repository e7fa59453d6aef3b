"""The autotuner's candidate sets (host logic only; measuring needs a GPU): what the full and
the quick sweeps cover, and that no candidate's TTs exceed the HBM share the tuner allows."""
import importlib
import os
import sys

from conftest import ROOT


def test_candidate_sets():
    sys.path.insert(0, ROOT)
    tune = importlib.import_module("dds_gpu_amd.tune")
    full, quick = tune.candidates(False), tune.candidates(True)
    for c in full + quick:
        assert c["lanes"] % 256 == 0 and c["team"] in (1, 16) and c["jobs_per_deal"] in (5, 20)
        per_search = 320 * (1 << c["tt_log2"])
        assert c["lanes"] // c["team"] * per_search <= 260e9 or c["team"] == 16
        assert c["waves_per_simd"] in (0, 3, 4) and (c["waves_per_simd"] == 0 or c["schedule"] == 2)
    assert {c["schedule"] for c in full} == {0, 1, 2, 3}  # stage orders, and none (per-thread)
    assert {c["tt_log2"] for c in full if c["team"] == 1} == {9, 10, 11, 12}
    assert {c["lanes"] // 256 for c in full} == {256, 512, 768, 1024}  # threads per CU
    assert {c["jobs_per_deal"] for c in full} == {5, 20}
    assert all(c["team"] == 1 and c["jobs_per_deal"] == 5 and c["schedule"] == 2 for c in quick)
    tuned = {"lanes": 196608, "team": 1, "tt_log2": 12, "schedule": 2, "jobs_per_deal": 5, "waves_per_simd": 3}
    assert tuned in quick and tuned in full  # the committed configuration is one of them
