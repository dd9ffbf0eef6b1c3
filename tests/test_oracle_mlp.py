"""Pins the oracle's MLP (the learned correction's network) on the CPU:
backward pass against central finite differences in FP64, BF16 rounding
against the definition, and the forward against a numpy restatement."""
import numpy as np

import oracle as O


def _params(seed=0, inp=12, out=3, scale=0.2):
    rng = np.random.default_rng(seed)
    return (scale * rng.normal(size=O.mlp_nparams(inp, out))).astype(np.float32)


def _numpy_forward(w, x, inp=12, out=3):
    H = O.MLP_H
    o = 0
    mats = []
    for shape in [(inp, H), (H,), (H, H), (H,), (H, H), (H,), (H, out), (out,)]:
        n = int(np.prod(shape))
        mats.append(w[o:o + n].astype(np.float64).reshape(shape))
        o += n
    h = x.astype(np.float64)
    for L in range(3):
        h = np.maximum(h @ mats[2 * L] + mats[2 * L + 1], 0)
    return h @ mats[6] + mats[7]


def test_forward_matches_numpy_fp64():
    w = _params()
    x = np.random.default_rng(1).normal(size=(37, 12)).astype(np.float32)
    assert np.allclose(O.mlp_forward(w, x, bf16=False), _numpy_forward(w, x), rtol=1e-12, atol=1e-12)


def test_bf16_rounding_is_nearest_even():
    w = _params()
    x = np.random.default_rng(2).normal(size=(8, 12)).astype(np.float32)
    # round the weights and inputs to bf16 in numpy; bf16 mode must equal fp64 on them
    def bf(a):
        u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        return u.astype(np.uint32).view(np.float32)
    y_bf = O.mlp_forward(w, x, bf16=True)
    y_ref = O.mlp_forward(w, x, bf16=False)
    assert np.abs(y_bf - y_ref).max() > 0  # rounding does something
    assert np.abs(y_bf - y_ref).max() < 0.05 * np.abs(y_ref).max()
    assert bf(np.array([1.0 + 2 ** -8], np.float32))[0] == 1.0  # tie -> even
    assert bf(np.array([1.0 + 3 * 2 ** -8], np.float32))[0] == 1.0 + 2 ** -6


def test_gradient_matches_finite_differences():
    w = _params(3, scale=0.3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(6, 12)).astype(np.float32)
    t = rng.normal(size=(6, 3)).astype(np.float32)
    g, loss = O.mlp_grad(w, x, t, bf16=False)
    idx = rng.choice(w.size, 60, replace=False)
    # perturb in float32 parameter space; central differences in FP64 loss
    for i in idx:
        h = 1e-3 * max(1.0, abs(float(w[i])))
        wp = w.copy(); wp[i] += h
        wm = w.copy(); wm[i] -= h
        hp = float(wp[i]) - float(w[i])
        hm = float(w[i]) - float(wm[i])
        lp = O.mlp_grad(wp, x, t, bf16=False)[1]
        lm = O.mlp_grad(wm, x, t, bf16=False)[1]
        fd = (lp - lm) / (hp + hm)
        assert abs(fd - g[i]) <= 1e-3 * max(1e-4, abs(g[i])) + 1e-6, (i, fd, g[i])
