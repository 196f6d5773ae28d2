"""FP64 numpy restatement of the tail model (32 -> 64 -> 64 -> 2, tanh) and
its loss gradient.  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The model has no counterpart in the reference.  The analytic gradient here is
pinned by central finite differences (``numeric_grad``), and the GPU kernels
are pinned against it.  Parameter layout: W1[64][32] b1[64] W2[64][64] b2[64]
W3[2][64] b3[2].
"""
import numpy as np

IN, H, OUT = 32, 64, 2
SHAPES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)), ("b3", (OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def forward(p, X):
    w = unpack(p)
    h1 = np.tanh(X @ w["W1"].T + w["b1"])
    h2 = np.tanh(h1 @ w["W2"].T + w["b2"])
    return h2 @ w["W3"].T + w["b3"], (h1, h2)


def loss(p, X, Y):
    y, _ = forward(p, X)
    return np.mean((y - Y) ** 2)


def grad(p, X, Y):
    w = unpack(p)
    y, (h1, h2) = forward(p, X)
    d3 = 2.0 * (y - Y) / y.size
    g = {"W3": d3.T @ h2, "b3": d3.sum(0)}
    d2 = (d3 @ w["W3"]) * (1 - h2 ** 2)
    g["W2"], g["b2"] = d2.T @ h1, d2.sum(0)
    d1 = (d2 @ w["W2"]) * (1 - h1 ** 2)
    g["W1"], g["b1"] = d1.T @ X, d1.sum(0)
    return np.concatenate([g[name].ravel() for name, _ in SHAPES])


def numeric_grad(p, X, Y, idx, h=1e-6):
    p = np.asarray(p, dtype=np.float64)
    out = np.zeros(len(idx))
    for k, i in enumerate(idx):
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        out[k] = (loss(a, X, Y) - loss(b, X, Y)) / (2 * h)
    return out
