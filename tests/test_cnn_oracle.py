"""The surrogate network's CPU oracle, pinned by mathematics (the reference has
no surrogate): brute-force convolution loops, and finite differences of the
loss in float64 for every layer's weights and biases."""
import numpy as np
import pytest

from oracle import cnn as C


def test_conv_matches_brute_force_loops():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 6, 3))
    for ks in (1, 3):
        w = rng.standard_normal((4, ks * ks, 3))
        b = rng.standard_normal(4)
        got = C.conv(x, w, b, ks)
        ref = np.zeros((2, 5, 6, 4))
        r = ks // 2
        for n in range(2):
            for i in range(5):
                for j in range(6):
                    for o in range(4):
                        s = b[o]
                        for t in range(ks * ks):
                            ii, jj = i + t // ks - r, j + t % ks - r
                            if 0 <= ii < 5 and 0 <= jj < 6:
                                s += x[n, ii, jj] @ w[o, t]
                        ref[n, i, j, o] = s
        assert np.allclose(got, ref, atol=1e-12)


def test_pool_up_are_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 4, 6, 2))
    d = rng.standard_normal((1, 2, 3, 2))
    # <pool(x), d> = <x, pool_bwd(d)> and <up(d), x> = <d, up_bwd(x)>
    assert np.isclose((C.pool(x) * d).sum(), (x * C.pool_bwd(d)).sum())
    assert np.isclose((C.up(d) * x).sum(), (d * C.up_bwd(x)).sum())


def _tiny(seed=2):
    rng = np.random.default_rng(seed)
    ch = (2, 3, 4, 5)
    params = {}
    for name, (co, ks, ci) in C.layer_shapes(ch, cin0=3, cout_head=3).items():
        params[name] = (rng.standard_normal((co, ks * ks, ci)) * np.sqrt(2.0 / (ks * ks * ci)),
                        rng.standard_normal(co) * 0.1)
    x = rng.uniform(0, 1, (2, 16, 16, 3))
    lab = rng.uniform(0, 1, (2, 8, 8, 2))
    return params, x, lab


def test_backward_matches_finite_differences():
    params, x, lab = _tiny()
    c0, core, nl = 4, 8, 2
    y, a = C.forward(params, x)
    g = C.backward(params, a, lab, c0, core, nl)
    rng = np.random.default_rng(3)
    h = 1e-6
    for name in C.LAYERS:
        w, b = params[name]
        for which, arr, grad in (("w", w, g[name][0]), ("b", b, g[name][1])):
            for _ in range(3):
                idx = tuple(rng.integers(0, s) for s in arr.shape)
                old = arr[idx]
                arr[idx] = old + h
                lp = C.loss(C.forward(params, x)[0], lab, c0, core, nl)
                arr[idx] = old - h
                lm = C.loss(C.forward(params, x)[0], lab, c0, core, nl)
                arr[idx] = old
                num = (lp - lm) / (2 * h)
                assert abs(num - grad[idx]) <= 1e-6 + 1e-4 * abs(num), (name, which, idx, num, grad[idx])


def test_planted_backward_mistake_is_caught(monkeypatch):
    """A dropped skip gradient (a plausible mistake) fails the finite-difference check."""
    params, x, lab = _tiny(5)
    c0, core, nl = 4, 8, 2
    orig = C.pool_bwd
    monkeypatch.setattr(C, "pool_bwd", lambda d: orig(d) * 0.5)
    y, a = C.forward(params, x)
    g = C.backward(params, a, lab, c0, core, nl)
    w, b = params["e1a"]
    h = 1e-6
    diffs = []
    for k in range(w.size):
        idx = np.unravel_index(k, w.shape)
        old = w[idx]
        w[idx] = old + h
        lp = C.loss(C.forward(params, x)[0], lab, c0, core, nl)
        w[idx] = old - h
        lm = C.loss(C.forward(params, x)[0], lab, c0, core, nl)
        w[idx] = old
        diffs.append(abs((lp - lm) / (2 * h) - g["e1a"][0][idx]))
    assert max(diffs) > 1e-5
