"""The first 2,000 deals of the bench stream solved to completion on the GPU: all 40,000 answers
equal the reference solver's tables.

The heaviest jobs of these deals cost up to 5.3e8 nodes (single null-window searches of up to
2.3e8), more than an hour on one lane of the plain configuration. Sixteen-lane teams search
several times faster per search, and a 21 MB TT per search (TT_LOG2 16) halves the heaviest
jobs' nodes, so this configuration -- 131,072 lanes = 8,192 team searches, one job per lead --
finishes the set in ~11 minutes. The answers are written to mapped host memory, so the test
sees each one arrive and prints the completion curve; progress lines go straight to the
terminal every 30 s (past pytest's capture), so a long quiet run is never taken for a hang. A
host-side limit stops the kernel at 840 s, which fails the test rather than holding the GPU.

At ~11 minutes it runs only when DDX_LONG_GPU_TESTS=1 is set, so the default GPU suite stays
short:

    DDX_LONG_GPU_TESTS=1 python -u -m pytest tests/test_zy_c2_to_completion.py -m gpu -v -s"""
import ctypes
import os
import time

import numpy as np
import pytest
from conftest import golden

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("DDX_LONG_GPU_TESTS") != "1",
                                 reason="~11 min run to completion; set DDX_LONG_GPU_TESTS=1")]

H2D, MAPPED = 1, 2  # hipMemcpyHostToDevice, hipHostMallocMapped


def test_c2_subset_all_answers_to_completion(ddx, capsys):
    def say(msg):
        with capsys.disabled():
            print(msg, flush=True)

    g = golden("full_2000.npz")
    n = len(g["hands"])
    ddx.lib()
    hip = ctypes.CDLL("libamdhip64.so")  # the runtime libddx.so already loaded
    host, dev_out, d_hands = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipHostMalloc(ctypes.byref(host), ctypes.c_size_t(n * 20), ctypes.c_uint(MAPPED)) == 0
    assert hip.hipHostGetDevicePointer(ctypes.byref(dev_out), host, ctypes.c_uint(0)) == 0
    assert hip.hipMalloc(ctypes.byref(d_hands), ctypes.c_size_t(n * 32)) == 0
    try:
        hands = np.ascontiguousarray(g["hands"])
        assert hip.hipMemcpy(d_hands, hands.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(n * 32), H2D) == 0
        ctypes.memset(host, 0xFF, n * 20)  # unsolved until the kernel writes an answer
        out = np.ctypeslib.as_array((ctypes.c_uint8 * (n * 20)).from_address(host.value)).reshape(n, 20)
        first = np.full((n, 20), -1.0)
        with ddx.GpuSolver(lanes=131072, team=16, tt_log2=16, jobs_per_deal=20, schedule=ddx.SCHED_SPLIT) as s:
            s.solve_async(d_hands.value, n, dev_out.value)
            t0, shown = time.time(), 0.0
            while True:
                time.sleep(0.2)
                t = time.time() - t0
                done = out != 255
                first[(first < 0) & done] = t
                if t - shown >= 30:
                    say("  C2 to completion, %.0f s: %d of %d answers" % (t, int(done.sum()), n * 20))
                    shown = t
                if done.all() or t > 840:
                    break
            if not done.all():
                s.request_stop()
            st = s.wait(allow_stop=True)
        answers = np.sort(first.ravel())
        whole = np.sort(first.max(axis=1))
        say("2,000 deals: last answer at %.1f s, %d nodes; answers 50%% / 90%% / 99%% / 99.9%% by "
            "%.1f / %.1f / %.1f / %.1f s; 1,000 / 1,900 / 1,990 whole tables by %.1f / %.1f / %.1f s" %
            (answers[-1], st.nodes, answers[n * 10 - 1], answers[n * 18 - 1], answers[int(n * 19.8) - 1],
             answers[int(n * 19.98) - 1], whole[999], whole[1899], whole[1989]))
        assert (first >= 0).all(), "stopped at the time limit"
        assert st.jobs_done == n * 20
        assert np.array_equal(out, g["tables"])
    finally:
        hip.hipFree(d_hands)
        hip.hipHostFree(host)
