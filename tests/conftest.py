import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def golden(name):
    path = os.path.join(GOLDEN, name)
    if name.endswith(".npz"):
        return dict(np.load(path, allow_pickle=False))
    return np.load(path, allow_pickle=False)


@pytest.fixture(scope="session")
def ddx():
    from dds_gpu_amd import ddx as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def oracle():
    """The oracle's C library (oracle/bin/libddo.so), built by oracle/Makefile if missing."""
    so = os.path.join(ROOT, "oracle", "bin", "libddo.so")
    if not os.path.exists(so):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "bin/libddo.so"])
    return Oracle(ctypes.CDLL(so))


class Oracle:
    def __init__(self, lib):
        self.lib = lib
        P = ctypes.c_void_p
        lib.ddo_random_deal.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int, P]
        lib.ddo_pgx_hands.argtypes = [P, P]
        lib.ddo_pgx_tricks.argtypes = [P, ctypes.c_int, ctypes.c_int]
        lib.ddo_ns_tricks.argtypes = [P, ctypes.c_int, ctypes.c_int]
        lib.ddo_ns_tricks_brute.argtypes = [P, ctypes.c_int, ctypes.c_int]
        lib.ddo_table.argtypes = [P, ctypes.c_int, P]
        lib.ddo_tables.argtypes = [P, ctypes.c_int64, ctypes.c_int, P]

    def deals(self, seed, n, k, first=0):
        h = np.zeros((n, 4), np.uint64)
        for i in range(n):
            self.lib.ddo_random_deal(seed, first + i, k, h[i].ctypes.data_as(ctypes.c_void_p))
        return h

    def tables(self, hands, threads=None, brute=False):
        hands = np.ascontiguousarray(hands, np.uint64)
        out = np.zeros((len(hands), 20), np.uint8)
        if brute:
            for i in range(len(hands)):
                assert self.lib.ddo_table(hands[i].ctypes.data_as(ctypes.c_void_p), 1,
                                          out[i].ctypes.data_as(ctypes.c_void_p)) == 0
            return out
        threads = threads or min(8, os.cpu_count() or 1)
        rc = self.lib.ddo_tables(hands.ctypes.data_as(ctypes.c_void_p), len(hands), threads,
                                 out.ctypes.data_as(ctypes.c_void_p))
        assert rc == 0
        return out

    def pgx(self, arr):
        """Hands (n, 4) and tables (n, 20) from a Pgx array (2, n, 4)."""
        keys = np.ascontiguousarray(arr[0], np.int32)
        vals = np.ascontiguousarray(arr[1], np.int32)
        n = keys.shape[0]
        hands = np.zeros((n, 4), np.uint64)
        tables = np.zeros((n, 20), np.uint8)
        for i in range(n):
            self.lib.ddo_pgx_hands(keys[i].ctypes.data_as(ctypes.c_void_p),
                                   hands[i].ctypes.data_as(ctypes.c_void_p))
            for st in range(5):
                for seat in range(4):
                    tables[i, st * 4 + seat] = self.lib.ddo_pgx_tricks(
                        vals[i].ctypes.data_as(ctypes.c_void_p), seat, st)
        return hands, tables
