"""The TT probe and store read the matching entries of a tag line in batches (four, two) with
their loads issued together. They must return exactly what a one-at-a-time read in tag order
returns (the reference's dd_tt_probe / dd_tt_store, src/dd.h:458-532): the bound, the relevant
cards, the move hint, the count of entries examined, and the TT contents after a store -- on
random lines with every number of matching tags from 0 to 16."""
import ctypes
import os
import subprocess

import numpy as np
from conftest import ROOT


def test_batched_probe_and_store_equal_one_by_one(tmp_path):
    out = str(tmp_path / "libttbatch.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas",
                           "-I" + os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "support", "tt_batch.cpp"), "-o", out])
    lib = ctypes.CDLL(out)
    lib.ddx_tt_batch_check.argtypes = [ctypes.c_int64, ctypes.c_uint, ctypes.c_void_p]
    lib.ddx_tt_batch_check.restype = ctypes.c_int64
    covered = np.zeros(17, np.int64)
    bad = lib.ddx_tt_batch_check(200000, 12345, covered.ctypes.data)
    assert bad == 0
    assert (covered[[0, 1, 4, 5, 8, 9, 12, 13, 16]] > 0).all(), covered  # every batch boundary
