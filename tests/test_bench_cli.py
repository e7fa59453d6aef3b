"""bench.py's command line on the CPU: more than one GPU is refused with a message, and the
options the measurements use are there."""
import os
import subprocess
import sys

from conftest import ROOT


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + list(args), capture_output=True,
                          text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2 and "only --gpus 1" in r.stderr and not r.stdout.strip()


def test_options():
    out = run("--help").stdout
    for opt in ("--steps", "--warmup", "--leg-seconds", "--legs", "--variant", "--tt-log2", "--no-cpu-baseline",
                "--copy-calibration"):
        assert opt in out, opt
