"""Command-line tools: dd_pbn (PBN in, leader-side tricks out, the reference's DDS cross-check
interface) and pbn (a Pgx deal printed as PBN with its stored table), checked against the
reference solver's tables, the oracle and each other. CPU only (dd_pbn's GPU mode is covered
in test_gpu.py)."""
import os
import subprocess

import numpy as np
import pytest
from conftest import ROOT, golden

BIN = os.path.join(ROOT, "dds_gpu_amd", "bin")
R = "23456789TJQKA"


def tool(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail("%s is not built (python -m dds_gpu_amd.build)" % path)
    return path


def to_pbn(h):
    return "N:" + " ".join(".".join("".join(R[r] for r in range(12, -1, -1) if (int(h[p]) >> (16 * s + r)) & 1)
                                    for s in range(4)) for p in range(4))


def leader_side(tables, k):
    """Leader-side tricks [deal][strain][leader] from declarer tables (E/W entries 13 - NS)."""
    t = tables.reshape(-1, 5, 4).astype(int)
    out = np.zeros_like(t)
    for leader in range(4):
        decl = (leader + 3) % 4
        ns = t[:, :, decl] if decl % 2 == 0 else 13 - t[:, :, decl]
        out[:, :, leader] = ns if leader % 2 == 0 else k - ns
    return out


def test_dd_pbn_full_deals_vs_reference_tables():
    g = golden("full_2000.npz")
    idx = np.nonzero(g["nodes"].astype(np.int64).sum(axis=1) < 1_500_000)[0][:5]
    text = "\n".join(["% a comment line is skipped"] + [to_pbn(h) for h in g["hands"][idx]]) + "\n"
    r = subprocess.run([tool("dd_pbn")], input=text, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = np.array([list(map(int, l.split())) for l in r.stdout.strip().split("\n")]).reshape(-1, 5, 4)
    assert np.array_equal(got, leader_side(g["tables"][idx], 13))


def test_dd_pbn_endings_vs_oracle(ddx, oracle):
    hands = ddx.random_deals(31, 40, 6)
    text = "\n".join(to_pbn(h) for h in hands) + "\n"
    r = subprocess.run([tool("dd_pbn")], input=text, capture_output=True, text=True, timeout=300)
    got = np.array([list(map(int, l.split())) for l in r.stdout.strip().split("\n")]).reshape(-1, 5, 4)
    assert np.array_equal(got, leader_side(oracle.tables(hands), 6))


def test_pbn_prints_pgx_deal_and_table(ddx, tmp_path):
    g = golden("full_2000.npz")
    path = str(tmp_path / "d.npy")
    ddx.pgx_write(path, g["hands"][:3], g["tables"][:3])
    for i in range(3):
        r = subprocess.run([tool("pbn"), path, str(i)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0
        line1, line2 = r.stdout.strip().split("\n")
        assert line1 == to_pbn(g["hands"][i])
        want = [g["tables"][i][st * 4 + seat] for seat in range(4) for st in range(5)]
        assert list(map(int, line2.split(","))) == want
    assert subprocess.run([tool("pbn"), path, "7"], capture_output=True).returncode == 1
