"""CPU oracle of the trick-count MLP (test infrastructure only; numpy FP32).

Only tests/ use it, as the checker of the GPU kernels in dds_gpu_amd/csrc/ddx_mlp.hip. The net
has no counterpart in the reference; what it restates is the math: one-hot encoding, two ReLU
layers, softmax cross-entropy over 14 classes, and the backward pass. With bf16=True it rounds
to BF16 at the points where the GPU path stores BF16 (weights, hidden activations, output and
hidden gradients), so GPU and oracle differ only by FP32 summation order. Its own backward
pass is pinned by central finite differences in FP64 (tests/test_mlp.py).
"""
import numpy as np

IN, IN_USED, H, OUT, OUT_USED = 224, 217, 256, 16, 14
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)), ("b3", (OUT,))]


def unpack(p):
    out, off = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[off:off + n].reshape(shape)
        off += n
    return out


def pack(d):
    return np.concatenate([d[name].reshape(-1) for name, _ in SIZES])


def bf16(x):
    """Round float32 to the nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def encode(hands, tables):
    """X (20n, 224) one-hot features and labels (NS tricks) for deals hands (n, 4); a table
    byte of 255 (unsolved) gives label -1, which the loss skips."""
    n = len(hands)
    X = np.zeros((n * 20, IN), np.float32)
    y = np.zeros(n * 20, np.int64)
    for d in range(n):
        for strain in range(5):
            for decl in range(4):
                s = d * 20 + strain * 4 + decl
                for c in range(52):
                    bitpos = 16 * (c // 13) + c % 13
                    for o in range(4):
                        if (int(hands[d, o]) >> bitpos) & 1:
                            X[s, 4 * c + o] = 1.0
                X[s, 208 + strain] = 1.0
                X[s, 213 + ((decl + 1) & 3)] = 1.0
                t = int(tables[d, strain * 4 + decl])
                y[s] = -1 if t == 255 else t if decl % 2 == 0 else 13 - t
    return X, y


def loss_and_grads(p, X, y, rounding=True, dtype=np.float32):
    """Mean cross-entropy, accuracy and the gradient vector of the parameters p."""
    r = bf16 if rounding else (lambda a: a)
    w = unpack(p.astype(dtype))
    W1, W2, W3 = r(w["W1"]).astype(dtype), r(w["W2"]).astype(dtype), r(w["W3"]).astype(dtype)
    X = X.astype(dtype)
    M = len(X)
    H1 = r(np.maximum(X @ W1.T + w["b1"], 0)).astype(dtype)
    H2 = r(np.maximum(H1 @ W2.T + w["b2"], 0)).astype(dtype)
    Z = (H2 @ W3.T + w["b3"])[:, :OUT_USED]
    Z = Z - Z.max(axis=1, keepdims=True)
    E = np.exp(Z)
    P = E / E.sum(axis=1, keepdims=True)
    known = y >= 0
    yk = np.where(known, y, 0)
    loss = float(np.sum(np.where(known, -np.log(P[np.arange(M), yk]), 0.0)) / M)
    acc = float(np.sum(known & (P.argmax(axis=1) == y)) / M)
    dZ = np.zeros((M, OUT), dtype)
    dZ[:, :OUT_USED] = P
    dZ[np.arange(M), yk] -= 1
    dZ[~known] = 0
    dZ = r(dZ / M).astype(dtype)
    g = {"W3": dZ.T @ H2, "b3": dZ.sum(0)}
    dH2 = r((dZ @ W3) * (H2 != 0)).astype(dtype)
    g["W2"], g["b2"] = dH2.T @ H1, dH2.sum(0)
    dH1 = r((dH2 @ W2) * (H1 != 0)).astype(dtype)
    g["W1"], g["b1"] = dH1.T @ X, dH1.sum(0)
    g["W1"][:, IN_USED:] = 0
    g["W3"][OUT_USED:] = 0
    g["b3"][OUT_USED:] = 0
    return loss, acc, pack(g)


def load(path):
    """Parameters from a file written by ddx_mlp_save: "DDXMLP1\\0", int64 count, fp32 values."""
    with open(path, "rb") as f:
        magic = f.read(8)
        cnt = int(np.frombuffer(f.read(8), np.int64)[0])
        p = np.frombuffer(f.read(4 * cnt), np.float32).copy()
    assert magic == b"DDXMLP1\0" and cnt == sum(int(np.prod(s)) for _, s in SIZES), "not a net file"
    return p


def predict(p, X):
    """Argmax NS-trick predictions (BF16 weights and activations, as the GPU forward)."""
    w = unpack(p.astype(np.float32))
    H1 = bf16(np.maximum(X @ bf16(w["W1"]).T + w["b1"], 0))
    H2 = bf16(np.maximum(H1 @ bf16(w["W2"]).T + w["b2"], 0))
    Z = (H2 @ bf16(w["W3"]).T + w["b3"])[:, :OUT_USED]
    return Z.argmax(axis=1), Z
