"""Pins the MLP oracle: analytic gradients vs FP64 central differences; bf16 rounding vs exhaustive properties."""
import numpy as np

from oracle import mlp as om
from adock_amd.mlp import SHAPES


def _params(seed=0, scale=0.3):
    g = np.random.default_rng(seed)
    return {k: g.standard_normal(s) * (scale if k.startswith("W") else 0.1) for k, s in SHAPES.items()}


def test_gradients_vs_central_differences():
    p = _params()
    g = np.random.default_rng(1)
    X = g.standard_normal((6, 128))
    T = g.standard_normal((6, 23))
    M = (g.random((6, 23)) < 0.8).astype(float)
    loss, gr = om.loss_and_grads(p, X, T, M)
    h = 1e-6
    for k in SHAPES:
        flat = p[k].reshape(-1)
        for idx in g.choice(flat.size, 8, replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = om.loss_and_grads(p, X, T, M)
            flat[idx] = old - h
            lm, _ = om.loss_and_grads(p, X, T, M)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(gr[k].reshape(-1)[idx] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, idx)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.1416, 65504.0, 1e-20, 0.0], np.float32)
    y = om.bf16(x)
    assert y[0] == 1.0 and y[1] == 1.0          # 1 + 2^-8 is a tie: rounds to even (1.0)
    assert y[2] == np.float32(1.0078125)          # 1 + 1.5 * 2^-8 rounds up
    r = np.random.default_rng(0).standard_normal(10000).astype(np.float32)
    e = np.abs(om.bf16(r) - r) / np.abs(r)
    assert e.max() <= 2.0 ** -8
    assert np.all(om.bf16(om.bf16(r)) == om.bf16(r))
