"""The search machine reads no memory it did not write: built with the memory sanitizer (host
code only, tests/support/msan_machine.cpp), with the TT entries and the frame stack left
uninitialized, it solves 200 C1 endings with no report and enters the reference's nodes."""
import os
import shutil
import subprocess

import numpy as np
import pytest
from conftest import ROOT, golden

CLANG = "/opt/rocm/llvm/bin/clang++" if os.path.exists("/opt/rocm/llvm/bin/clang++") else shutil.which("clang++")


@pytest.mark.skipif(not CLANG, reason="no clang++ with the memory sanitizer")
def test_machine_under_memory_sanitizer(tmp_path):
    exe = str(tmp_path / "msan_machine")
    subprocess.check_call([CLANG, "-O1", "-g", "-std=c++17", "-fsanitize=memory", "-fno-omit-frame-pointer",
                           "-Wno-unknown-pragmas", "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "support", "msan_machine.cpp"), "-o", exe])
    c1 = golden("c1_endings.npz")
    n = 200
    hp = str(tmp_path / "hands.bin")
    np.ascontiguousarray(c1["hands"][:n], np.uint64).tofile(hp)
    r = subprocess.run([exe, hp, str(n)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "MemorySanitizer" not in r.stderr, r.stderr[-3000:]
    nodes, checksum = map(int, r.stdout.split())
    assert nodes == int(c1["nodes"][:n].sum())
    assert checksum == int(c1["tables"][:n].astype(np.int64).sum())
