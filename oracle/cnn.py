"""CPU oracle of the SVF surrogate network -- TEST INFRASTRUCTURE ONLY.

A plain numpy restatement (float32 or float64) of the 4-level
encoder-decoder the device runs: 3x3 same-padded convolutions with bias and
ReLU, 2x2 average pooling, nearest upsampling, channel concatenation of the
skip, a 1x1 sigmoid head and the mean squared error over the core window;
forward and hand-written backward. The reference has no surrogate, so the
oracle is pinned by mathematics instead: finite differences of the loss in
float64 and brute-force loops for the convolution (``tests/test_cnn_oracle.py``).

Layouts follow the device: activations NHWC, weights [cout][ks * ks][cin].
"""
from __future__ import annotations

import numpy as np


def conv(x, w, b, ks):
    """Same-padded convolution: x (n, h, w, cin), w (cout, ks*ks, cin), b (cout,) -> (n, h, w, cout)."""
    n, h, wd, cin = x.shape
    r = ks // 2
    xp = np.pad(x, ((0, 0), (r, r), (r, r), (0, 0)))
    out = np.zeros((n, h, wd, w.shape[0]), dtype=np.result_type(x, w))
    for t in range(ks * ks):
        dy, dx = t // ks, t % ks
        out += xp[:, dy:dy + h, dx:dx + wd, :] @ w[:, t, :].T
    return out + b


def conv_bwd(x, w, dz, ks):
    """Gradients of conv: (dx, dw, db)."""
    n, h, wd, cin = x.shape
    r = ks // 2
    xp = np.pad(x, ((0, 0), (r, r), (r, r), (0, 0)))
    dxp = np.zeros_like(xp, dtype=np.result_type(x, dz))
    dw = np.zeros_like(w, dtype=np.result_type(w, dz))
    for t in range(ks * ks):
        dy, dx = t // ks, t % ks
        win = xp[:, dy:dy + h, dx:dx + wd, :]
        dw[:, t, :] = np.einsum("nhwo,nhwi->oi", dz, win)
        dxp[:, dy:dy + h, dx:dx + wd, :] += dz @ w[:, t, :]
    db = dz.sum(axis=(0, 1, 2))
    return dxp[:, r:r + h, r:r + wd, :], dw, db


def pool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_bwd(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(d):
    n, h, w, c = d.shape
    return d.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


LAYERS = ("e1a", "e1", "e2a", "e2", "e3a", "e3", "e4a", "e4", "d3a", "d3", "d2a", "d2", "d1a", "d1", "head")


def layer_shapes(ch, cin0=16, cout_head=16):
    """(cout, ks, cin) of every conv layer for level channels ch = (c1, c2, c3, c4)."""
    c1, c2, c3, c4 = ch
    return {"e1a": (c1, 3, cin0), "e1": (c1, 3, c1), "e2a": (c2, 3, c1), "e2": (c2, 3, c2),
            "e3a": (c3, 3, c2), "e3": (c3, 3, c3), "e4a": (c4, 3, c3), "e4": (c4, 3, c4),
            "d3a": (c3, 3, c4 + c3), "d3": (c3, 3, c3), "d2a": (c2, 3, c3 + c2), "d2": (c2, 3, c2),
            "d1a": (c1, 3, c2 + c1), "d1": (c1, 3, c1), "head": (cout_head, 1, c1)}


def bf16_round(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def forward(params, x, store=None):
    """params: {name: (w, b)}; x (n, h, w, 16) -> (sigmoid outputs, cache).

    ``store`` is applied to every tensor the device keeps in memory between
    ops (identity by default; ``bf16_round`` mirrors the device's bf16
    activations)."""
    st = store or (lambda v: v)
    relu = lambda v: np.maximum(v, 0)  # noqa: E731
    a = {"x": st(x)}

    def cv(name, inp):
        w, b = params[name]
        return conv(inp, w, b, w.shape[1] == 9 and 3 or 1)
    a["e1a"] = st(relu(cv("e1a", a["x"])))
    a["e1"] = st(relu(cv("e1", a["e1a"])))
    a["p1"] = st(pool(a["e1"]))
    a["e2a"] = st(relu(cv("e2a", a["p1"])))
    a["e2"] = st(relu(cv("e2", a["e2a"])))
    a["p2"] = st(pool(a["e2"]))
    a["e3a"] = st(relu(cv("e3a", a["p2"])))
    a["e3"] = st(relu(cv("e3", a["e3a"])))
    a["p3"] = st(pool(a["e3"]))
    a["e4a"] = st(relu(cv("e4a", a["p3"])))
    a["e4"] = st(relu(cv("e4", a["e4a"])))
    a["c3"] = np.concatenate([up(a["e4"]), a["e3"]], axis=3)
    a["d3a"] = st(relu(cv("d3a", a["c3"])))
    a["d3"] = st(relu(cv("d3", a["d3a"])))
    a["c2"] = np.concatenate([up(a["d3"]), a["e2"]], axis=3)
    a["d2a"] = st(relu(cv("d2a", a["c2"])))
    a["d2"] = st(relu(cv("d2", a["d2a"])))
    a["c1"] = np.concatenate([up(a["d2"]), a["e1"]], axis=3)
    a["d1a"] = st(relu(cv("d1a", a["c1"])))
    a["d1"] = st(relu(cv("d1", a["d1a"])))
    z = cv("head", a["d1"])
    a["y"] = 1.0 / (1.0 + np.exp(-z))
    return a["y"], a


def loss(y, labels, c0, core, nl):
    d = y[:, c0:c0 + core, c0:c0 + core, :nl] - labels
    return float((d * d).mean())


def backward(params, a, labels, c0, core, nl, store=None):
    """Gradients {name: (dw, db)} of the mean squared error over the core window and first nl channels
    (``store`` as in :func:`forward`, applied to the stored gradient tensors)."""
    st = store or (lambda v: v)
    y = a["y"]
    count = labels.size
    dy = np.zeros_like(y)
    dy[:, c0:c0 + core, c0:c0 + core, :nl] = 2.0 * (y[:, c0:c0 + core, c0:c0 + core, :nl] - labels) / count
    dz = st(dy * y * (1 - y))
    g = {}

    def back(name, inp, dz_):
        w, _ = params[name]
        dx, dw, db = conv_bwd(inp, w, dz_, 3 if w.shape[1] == 9 else 1)
        g[name] = (dw, db)
        return st(dx)

    relu_b = lambda d, out: d * (out > 0)  # noqa: E731
    d = back("head", a["d1"], dz)
    d = back("d1", a["d1a"], relu_b(d, a["d1"]))
    d = back("d1a", a["c1"], relu_b(d, a["d1a"]))
    c = a["d2"].shape[3]
    d_up, d_e1 = d[..., :c], d[..., c:]
    d = st(up_bwd(d_up))
    d = back("d2", a["d2a"], relu_b(d, a["d2"]))
    d = back("d2a", a["c2"], relu_b(d, a["d2a"]))
    c = a["d3"].shape[3]
    d_up, d_e2 = d[..., :c], d[..., c:]
    d = st(up_bwd(d_up))
    d = back("d3", a["d3a"], relu_b(d, a["d3"]))
    d = back("d3a", a["c3"], relu_b(d, a["d3a"]))
    c = a["e4"].shape[3]
    d_up, d_e3 = d[..., :c], d[..., c:]
    d = st(up_bwd(d_up))
    d = back("e4", a["e4a"], relu_b(d, a["e4"]))
    d = back("e4a", a["p3"], relu_b(d, a["e4a"]))
    d = st(d_e3 + pool_bwd(d))
    d = back("e3", a["e3a"], relu_b(d, a["e3"]))
    d = back("e3a", a["p2"], relu_b(d, a["e3a"]))
    d = st(d_e2 + pool_bwd(d))
    d = back("e2", a["e2a"], relu_b(d, a["e2"]))
    d = back("e2a", a["p1"], relu_b(d, a["e2a"]))
    d = st(d_e1 + pool_bwd(d))
    d = back("e1", a["e1a"], relu_b(d, a["e1"]))
    back("e1a", a["x"], relu_b(d, a["e1a"]))
    return g
