"""ddx_mlp_predict_noalloc (no allocation; activations in the library's fixed device scratch)
against ddx_mlp_predict on the same deals: identical seeds on two calls in a row, the second
with fewer deals, so leftovers of the first call in the shared scratch must not leak into the
second. Sizes beyond the scratch are refused. Device buffers come from the HIP runtime the
library itself uses."""
import ctypes
import os

import numpy as np
import pytest
from conftest import ROOT

pytestmark = pytest.mark.gpu

MODEL = os.path.join(ROOT, "dds_gpu_amd", "models", "tricknet.bin")
H2D, D2H = 1, 2


def test_noalloc_predict_equals_predict(ddx):
    ddx.lib()
    hip = ctypes.CDLL("libamdhip64.so")  # the runtime libddx.so already loaded

    def dmalloc(nbytes):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        return p

    net = ddx.TrickNet(path=MODEL)
    bufs = []
    try:
        for n, seed in ((256, 17), (37, 18)):  # the full scratch, then a smaller call with padding rows
            hands = np.ascontiguousarray(ddx.random_deals(seed, n, 13))
            d_hands, d_want, d_got = dmalloc(n * 32), dmalloc(n * 20), dmalloc(n * 20)
            bufs += [d_hands, d_want, d_got]
            assert hip.hipMemcpy(d_hands, hands.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(n * 32), H2D) == 0
            net.predict_device(d_hands.value, n, d_want.value)  # no stream: returns when done
            net.predict_device_noalloc(d_hands.value, n, d_got.value)  # null stream, waits too
            w, g = np.zeros((n, 20), np.uint8), np.zeros((n, 20), np.uint8)
            for host, dev in ((w, d_want), (g, d_got)):
                assert hip.hipMemcpy(host.ctypes.data_as(ctypes.c_void_p), dev, ctypes.c_size_t(n * 20), D2H) == 0
            print("n=%d: %d seeds, %d differ" % (n, w.size, int((w != g).sum())), flush=True)
            assert (w <= 13).all() and np.array_equal(w, g)
        # Beyond the scratch: a clear error, nothing launched.
        rc = ddx.lib().ddx_mlp_predict_noalloc(net.h, d_hands, 257, d_got, None)
        assert rc == ddx.E_ARG and b"exceed" in ddx.lib().ddx_last_error()
    finally:
        for b in bufs:
            hip.hipFree(b)
        net.close()
