"""Host-code sanitizers (CPU only):

  * the search machine reads no memory it did not write: built with the memory sanitizer
    (tests/support/msan_machine.cpp), with the TT entries and the frame stack left
    uninitialized, it solves 200 C1 endings with no report and enters the reference's nodes;
  * the shared TT has no data race: four threads on one TT, built with the thread sanitizer
    (tests/support/tsan_shared_tt.cpp), with no report and the reference's tables;
  * no undefined behaviour in any TT build's search (-fsanitize=undefined);
  * the host side of the C ABI runs clean under the address sanitizer."""
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


def test_threads_on_one_tt_under_thread_sanitizer(tmp_path):
    """Four threads on one shared TT, built with the thread sanitizer (host code only,
    tests/support/tsan_shared_tt.cpp): no race reported, and the tables are the reference's."""
    exe = str(tmp_path / "tsan_shared_tt")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=thread", "-Wno-unknown-pragmas",
                           "-DDDX_SHARED_TT=1", "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "support", "tsan_shared_tt.cpp"), "-o", exe, "-pthread"])
    c1 = golden("c1_endings.npz")
    hp = str(tmp_path / "hands.bin")
    np.ascontiguousarray(c1["hands"][:200], np.uint64).tofile(hp)
    r = subprocess.run([exe, hp, "200"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ThreadSanitizer" not in r.stderr, r.stderr[-3000:]
    assert r.stdout.split() == ["sum", str(int(c1["tables"][:200].astype(np.int64).sum()))]


@pytest.mark.parametrize("build", ["", "-DDDX_TT_SIG=1", "-DDDX_SHARED_TT=1"])
def test_machine_under_undefined_behaviour_sanitizer(tmp_path, build):
    """No undefined behaviour (shifts, overflows, misaligned accesses) in any TT build's search
    on 200 C1 endings and 4 full deals (g++ -fsanitize=undefined, aborting on the first report)."""
    exe = str(tmp_path / "ubsan_machine")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=undefined", "-fno-sanitize-recover=undefined",
                           "-Wno-unknown-pragmas", "-I" + os.path.join(ROOT, "include")] + ([build] if build else []) +
                          [os.path.join(ROOT, "tests", "support", "msan_machine.cpp"), "-o", exe])
    c1, full = golden("c1_endings.npz"), golden("full_2000.npz")
    rows = [1386, 707, 1081, 1897]
    for hands, tables in ((c1["hands"][:200], c1["tables"][:200]), (full["hands"][rows], full["tables"][rows])):
        hp = str(tmp_path / "hands.bin")
        np.ascontiguousarray(hands, np.uint64).tofile(hp)
        r = subprocess.run([exe, hp, str(len(hands))], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0 and "runtime error" not in r.stderr, r.stderr[-3000:]
        assert int(r.stdout.split()[1]) == int(tables.astype(np.int64).sum())


def test_host_abi_under_address_sanitizer(tmp_path):
    """The host side of the C ABI (tests/support/asan_abi.cpp with dds_gpu_amd/csrc/ddx_host.cpp,
    g++ -fsanitize=address,undefined): CPU batch solves with one and four threads, strain
    solves, a Pgx round trip and argument errors, with no report and no leak."""
    exe = str(tmp_path / "asan_abi")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-Wno-unknown-pragmas",
                           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "dds_gpu_amd", "csrc"),
                           os.path.join(ROOT, "tests", "support", "asan_abi.cpp"),
                           os.path.join(ROOT, "dds_gpu_amd", "csrc", "ddx_host.cpp"), "-o", exe, "-pthread"])
    r = subprocess.run([exe, str(tmp_path / "t.npy"), str(tmp_path / "missing.npy")], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), (r.returncode, r.stderr[-3000:])
