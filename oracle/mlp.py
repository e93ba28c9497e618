"""TEST INFRASTRUCTURE ONLY: float64 numpy restatement of the warm-start MLP
(48 -> 256 -> 256 -> 256 -> 3, exact-erf GELU) used to check the HIP
forward / backward / loss / Adam kernels. The network has no counterpart in
the reference; its gradients are pinned by finite differences
(tests/test_oracle_mlp.py)."""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

IN, HID, OUT = 48, 256, 3
SHAPES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)), ("b3", (HID,)),
          ("W4", (OUT, HID)), ("b4", (OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, x):
    W = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    Z1 = x @ W["W1"].T + W["b1"]
    H1 = gelu(Z1)
    Z2 = H1 @ W["W2"].T + W["b2"]
    H2 = gelu(Z2)
    Z3 = H2 @ W["W3"].T + W["b3"]
    H3 = gelu(Z3)
    y = H3 @ W["W4"].T + W["b4"]
    return y, (x, Z1, H1, Z2, H2, Z3, H3)


def backward(p, cache, dy):
    W = unpack(p)
    x, Z1, H1, Z2, H2, Z3, H3 = cache
    g = {}
    g["W4"] = dy.T @ H3
    g["b4"] = dy.sum(0)
    dZ3 = (dy @ W["W4"]) * gelu_grad(Z3)
    g["W3"] = dZ3.T @ H2
    g["b3"] = dZ3.sum(0)
    dZ2 = (dZ3 @ W["W3"]) * gelu_grad(Z2)
    g["W2"] = dZ2.T @ H1
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ W["W2"]) * gelu_grad(Z1)
    g["W1"] = dZ1.T @ x
    g["b1"] = dZ1.sum(0)
    return pack(g)


def mse(y, t, w=None):
    w = np.ones(len(y)) if w is None else np.asarray(w, dtype=np.float64)
    W = w.sum()
    d = y - t
    return float((w[:, None] * d * d).sum() / W), 2.0 * w[:, None] * d / W


def adam(p, g, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v


def init(seed=0):
    """He-style initialisation used by the warm-start trainer."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            d[name] = rng.standard_normal(shape) * math.sqrt(2.0 / shape[1])
        else:
            d[name] = np.zeros(shape)
    d["W4"] *= 0.1
    return pack(d)
