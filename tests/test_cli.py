"""The command-line boundary: dd_cpu, dd_gpu and dd_stream keep the reference's argv forms,
output lines, records and exit codes (0 = every checked answer equal, 1 = a mismatch, a load
failure or a GPU error, 2 = usage or TT_LOG2 < 9; reference cpu/cpu_main.cpp:12-23,85-87,
gpu/gpu_main.cu:165-200,340-341, gpu/stream_main.cu:4-12,124-132,289-314).

A mismatch is provoked by changing one stored table entry of the Pgx file the tool checks
against -- the data, never the solver."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest
from conftest import ROOT, golden

BIN = os.path.join(ROOT, "dds_gpu_amd", "bin")

# Fixture deals with the fewest reference nodes (cheap full deals).
EASY_FULL = [1386, 707, 1081, 1897, 1468, 1032, 356, 1069]


def tool(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip("%s is not built" % name)
    return path


def run(args, env=None, stdin=None, timeout=300):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run(args, input=stdin, capture_output=True, timeout=timeout, env=e)


@pytest.fixture(scope="module")
def easy_file(ddx, tmp_path_factory):
    """A Pgx file of eight cheap fixture deals with the reference solver's tables, and a copy with
    one stored answer changed."""
    g = golden("full_2000.npz")
    d = tmp_path_factory.mktemp("cli")
    good, bad = str(d / "easy.npy"), str(d / "easy_bad.npy")
    ddx.pgx_write(good, g["hands"][EASY_FULL], g["tables"][EASY_FULL])
    t = g["tables"][EASY_FULL].copy()
    t[3, 7] = (t[3, 7] + 1) % 14
    ddx.pgx_write(bad, g["hands"][EASY_FULL], t)
    return good, bad


# ---- usage errors and load failures: no solve is started --------------------------------

@pytest.mark.parametrize("name,args", [
    ("dd_cpu", []), ("dd_cpu", ["f.npy", "0"]),
    ("dd_gpu", []), ("dd_gpu", ["7"]), ("dd_gpu", ["0", "10"]), ("dd_gpu", ["14", "10"]),
])
def test_usage_errors_exit_2(name, args):
    r = run([tool(name)] + args)
    assert r.returncode == 2, r.stderr


def test_tt_log2_below_9_exits_2(easy_file):
    good, _ = easy_file
    assert run([tool("dd_cpu"), good, "0", "2", "1", "8"]).returncode == 2
    assert run([tool("dd_gpu"), "7", "10", "8"]).returncode == 2
    assert run([tool("dd_gpu"), good, "0", "2", "8"]).returncode == 2
    assert run([tool("dd_stream"), "bench", "1", "8"]).returncode == 2


def test_load_failure_exits_1(tmp_path):
    missing = str(tmp_path / "missing.npy")
    assert run([tool("dd_cpu"), missing, "0", "2"]).returncode == 1
    assert run([tool("dd_gpu"), missing, "0", "2"]).returncode == 1


# ---- dd_cpu -------------------------------------------------------------------------------

DD_CPU_LINE = re.compile(r"^(\d+) deals, (\d+) threads: [\d.]+s, [\d.]+ deals/s, \d+ nodes/deal, "
                         r"(\d+) wrong of (\d+)$")


def test_dd_cpu_line_and_exit_codes(easy_file):
    good, bad = easy_file
    r = run([tool("dd_cpu"), good, "0", "8", "4", "12"])
    assert r.returncode == 0, r.stdout + r.stderr
    m = DD_CPU_LINE.match(r.stdout.decode().strip().splitlines()[-1])
    assert m and m.groups() == ("8", "4", "0", "160")
    r = run([tool("dd_cpu"), bad, "0", "8", "4", "12"])
    assert r.returncode == 1
    lines = r.stdout.decode().strip().splitlines()
    assert lines[0] == "deal 3 strain 1 seat 3: want %d got %d" % (
        (golden("full_2000.npz")["tables"][EASY_FULL[3], 7] + 1) % 14, golden("full_2000.npz")["tables"][EASY_FULL[3], 7])
    assert DD_CPU_LINE.match(lines[-1]).group(3) == "1"


def test_dd_cpu_start_offset(easy_file):
    """START picks the rows: rows 4..7 of the good file all agree."""
    good, _ = easy_file
    r = run([tool("dd_cpu"), good, "4", "4", "2", "11"])
    assert r.returncode == 0 and DD_CPU_LINE.match(r.stdout.decode().strip().splitlines()[-1]).group(4) == "80"


# ---- dd_gpu and dd_stream on the GPU -----------------------------------------------------

@pytest.mark.gpu
def test_dd_gpu_endings_mode():
    """dd_gpu K DEALS [TT_LOG2] [CHECK]: random 6-card endings, the first 300 re-solved on the CPU."""
    r = run([tool("dd_gpu"), "6", "600", "10", "300"], env={"THREADS": "4096"})
    out = r.stdout.decode()
    assert r.returncode == 0, out + r.stderr.decode()
    assert out.strip().splitlines()[-1] == "0 wrong of 6000 checked"
    assert re.search(r"GPU: [\d.]+s, \d+ deals/s .*600 whole tables", out)


@pytest.mark.gpu
def test_dd_gpu_file_mode(easy_file):
    """dd_gpu FILE START DEALS [TT_LOG2]: checked against the stored tables; one changed entry in the
    file gives exit 1 and "1 wrong of 160 checked"."""
    good, bad = easy_file
    r = run([tool("dd_gpu"), good, "0", "8", "10"], env={"THREADS": "256"})
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()
    assert r.stdout.decode().strip().splitlines()[-1] == "0 wrong of 160 checked"
    r = run([tool("dd_gpu"), bad, "0", "8", "10"], env={"THREADS": "256"})
    assert r.returncode == 1
    assert r.stdout.decode().strip().splitlines()[-1] == "1 wrong of 160 checked"


@pytest.mark.gpu
def test_dd_stream_records():
    """dd_stream: 32-byte deals on stdin, one 28-byte record per deal on stdout (int64 input index,
    then tricks[strain][declarer]), every index once, every table the reference's."""
    g = golden("full_2000.npz")
    hands = g["hands"][EASY_FULL]
    stdin = b"".join(struct.pack("<4Q", *[int(x) for x in h]) for h in hands)
    r = run([tool("dd_stream")], env={"THREADS": "4096"}, stdin=stdin, timeout=300)
    assert r.returncode == 0, r.stderr.decode()
    data = r.stdout
    assert len(data) == 28 * len(hands)
    seen = set()
    for k in range(len(hands)):
        idx = struct.unpack_from("<q", data, 28 * k)[0]
        table = np.frombuffer(data, np.uint8, 20, 28 * k + 8)
        assert 0 <= idx < len(hands) and idx not in seen
        seen.add(idx)
        assert np.array_equal(table, g["tables"][EASY_FULL[idx]]), idx


# ---- the other TT builds' tools (bin/sig, bin/shared) ---------------------------------------

VARIANT_TT = {"sig": "12", "shared": "16"}  # shared: TT_LOG2 is the size of the one TT


@pytest.mark.parametrize("variant", ["sig", "shared"])
def test_variant_dd_cpu(easy_file, variant):
    good, bad = easy_file
    r = run([tool(os.path.join(variant, "dd_cpu")), good, "0", "8", "4", VARIANT_TT[variant]])
    assert r.returncode == 0, r.stdout + r.stderr
    assert DD_CPU_LINE.match(r.stdout.decode().strip().splitlines()[-1]).groups() == ("8", "4", "0", "160")
    assert run([tool(os.path.join(variant, "dd_cpu")), bad, "0", "8", "4", VARIANT_TT[variant]]).returncode == 1


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["sig", "shared"])
def test_variant_dd_gpu_and_stream(easy_file, variant):
    good, _ = easy_file
    r = run([tool(os.path.join(variant, "dd_gpu")), good, "0", "8", VARIANT_TT[variant]], env={"THREADS": "256"})
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()
    assert r.stdout.decode().strip().splitlines()[-1] == "0 wrong of 160 checked"
    g = golden("full_2000.npz")
    stdin = b"".join(struct.pack("<4Q", *[int(x) for x in h]) for h in g["hands"][EASY_FULL])
    r = run([tool(os.path.join(variant, "dd_stream"))], env={"THREADS": "4096"}, stdin=stdin)
    assert r.returncode == 0, r.stderr.decode()
    assert len(r.stdout) == 28 * len(EASY_FULL)
    for k in range(len(EASY_FULL)):
        idx = struct.unpack_from("<q", r.stdout, 28 * k)[0]
        assert np.array_equal(np.frombuffer(r.stdout, np.uint8, 20, 28 * k + 8), g["tables"][EASY_FULL[idx]])
