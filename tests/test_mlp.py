"""The trick-count net: the CPU oracle pinned by finite differences, then the GPU kernels
(forward, backward, Adam training, prediction) checked against it."""
import os
import sys

import numpy as np
import pytest
from conftest import ROOT, golden

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_ref  # noqa: E402


def random_params(seed=0, scale=0.08):
    rng = np.random.default_rng(seed)
    p = mlp_ref.unpack(rng.normal(0, scale, sum(int(np.prod(s)) for _, s in mlp_ref.SIZES)).astype(np.float32))
    p["W1"][:, mlp_ref.IN_USED:] = 0
    p["W3"][mlp_ref.OUT_USED:] = 0
    p["b3"][mlp_ref.OUT_USED:] = 0
    return mlp_ref.pack(p)


@pytest.fixture(scope="module")
def batch():
    g = golden("full_2000.npz")
    return g["hands"][:8], g["tables"][:8]


def test_oracle_backward_matches_finite_differences(batch):
    """FP64, no rounding: analytic gradients == central differences on sampled parameters
    of every tensor (a dropped term, wrong sign or transposed operand fails)."""
    X, y = mlp_ref.encode(*batch)
    X, y = X[:60], y[:60]
    p = random_params(1).astype(np.float64)
    _, _, g = mlp_ref.loss_and_grads(p, X, y, rounding=False, dtype=np.float64)
    rng = np.random.default_rng(2)
    offs, off = [], 0
    for name, shape in mlp_ref.SIZES:
        n = int(np.prod(shape))
        cand = np.arange(off, off + n)
        if name == "W1":
            cand = cand[(cand - off) % mlp_ref.IN < 20]  # columns with some active inputs
        offs += list(rng.choice(cand, 6, replace=False))
        off += n
    eps = 1e-6
    for i in offs:
        q = p.copy()
        q[i] += eps
        lp, _, _ = mlp_ref.loss_and_grads(q, X, y, rounding=False, dtype=np.float64)
        q[i] -= 2 * eps
        lm, _, _ = mlp_ref.loss_and_grads(q, X, y, rounding=False, dtype=np.float64)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_oracle_encoding(batch):
    X, y = mlp_ref.encode(*batch)
    assert X.shape == (160, 224) and (X.sum(axis=1) == 54).all()  # 52 cards + strain + leader
    assert (X[:, 217:] == 0).all() and y.min() >= 0 and y.max() <= 13


def test_oracle_unknown_labels_are_skipped(batch):
    """Unsolved table bytes (255) drop out: the loss and gradients with masked rows equal those
    of the known rows alone, rescaled by known / all (the mean runs over every row)."""
    hands, tables = batch
    t = tables.copy()
    t[:, ::3] = 255
    X, y = mlp_ref.encode(hands, t)
    assert (y.reshape(-1, 20)[:, ::3] == -1).all()
    p = random_params(6).astype(np.float64)
    l_all, _, g_all = mlp_ref.loss_and_grads(p, X, y, rounding=False, dtype=np.float64)
    k = y >= 0
    l_k, _, g_k = mlp_ref.loss_and_grads(p, X[k], y[k], rounding=False, dtype=np.float64)
    f = k.sum() / len(y)
    assert abs(l_all - f * l_k) < 1e-12
    assert np.allclose(g_all, f * g_k, rtol=1e-10, atol=1e-14)


def test_oracle_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.1415926], np.float32)
    assert list(mlp_ref.bf16(x)) == [1.0, 1.0, 1.0078125, -2.5, 3.140625]


@pytest.mark.gpu
@pytest.mark.parametrize("masked", [False, True])
def test_gpu_gradients_match_oracle(ddx, batch, masked):
    """Loss and all 127,504 gradient entries of the GPU kernels vs the oracle with the same
    BF16 rounding points; tolerance covers FP32 summation order (and its rare one-ulp BF16
    flips): relative L2 error below 2e-2 per tensor. masked: a third of the table bytes are
    255 (unsolved) and must drop out of the loss."""
    hands, tables = batch
    if masked:
        tables = tables.copy()
        tables[:, ::3] = 255
    net = ddx.TrickNet(seed=3)
    p = random_params(4)
    net.set_params(p)
    g_gpu, loss_gpu, acc_gpu = net.grads(hands, tables)
    X, y = mlp_ref.encode(hands, tables)
    loss, acc, g = mlp_ref.loss_and_grads(p, X, y)
    assert abs(loss_gpu - loss) < 1e-3 * max(1, loss)
    gu, ou = mlp_ref.unpack(g_gpu), mlp_ref.unpack(g)
    for name, _ in mlp_ref.SIZES:
        err = np.linalg.norm(gu[name] - ou[name]) / max(np.linalg.norm(ou[name]), 1e-12)
        assert err < 2e-2, (name, err)
    net.close()


@pytest.mark.gpu
def test_gpu_training_learns_and_predicts(ddx, tmp_path):
    """A few Adam epochs on 1,600 solved deals lower the loss; predictions on held-out deals
    beat the constant predictor and the saved net reloads to the same seeds."""
    g = golden("full_2000.npz")
    h, t = g["hands"], g["tables"]
    net = ddx.TrickNet(seed=5)
    l0, _ = net.train(h[:1600], t[:1600], epochs=1, batch_deals=200, lr=1e-3)
    l1, a1 = net.train(h[:1600], t[:1600], epochs=6, batch_deals=200, lr=1e-3)
    assert l1 < l0
    seeds = net.predict(h[1600:])
    ns = ddx.ns_tricks_of(t[1600:])
    exact = float((seeds == ns).mean())
    within1 = float((np.abs(seeds.astype(int) - ns) <= 1).mean())
    const = float((ns == np.bincount(ns.reshape(-1)).argmax()).mean())
    print("held out: exact %.3f, within one %.3f, constant %.3f" % (exact, within1, const))
    assert exact > const
    path = str(tmp_path / "mlp_test.bin")
    net.save(path)
    net2 = ddx.TrickNet(seed=9, path=path)
    assert np.array_equal(net2.predict(h[1600:1700]), seeds[:100])
    net.close()
    net2.close()
