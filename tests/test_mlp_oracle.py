"""Pins of the FP64 reference perceptron (oracle/mlp_fp64.py) that the GPU
training kernels are checked against: its gradient equals central finite
differences of its loss, its GELU matches the erf form closely and has the
right derivative, and its bfloat16 rounding is round-to-nearest-even."""
import math

import numpy as np
import pytest

from oracle import mlp_fp64 as M


def _params(dims, seed):
    n_in, hidden, layers, n_out = dims
    rng = np.random.default_rng(seed)
    fan = [n_in] + [hidden] * layers + [n_out]
    parts = []
    for i in range(layers + 1):
        parts.append(rng.normal(0, 1 / math.sqrt(fan[i]), fan[i] * fan[i + 1]))
        parts.append(rng.normal(0, 0.1, fan[i + 1]))
    return np.concatenate(parts)


def test_gradient_equals_finite_differences():
    dims = (5, 8, 3, 4)
    p = _params(dims, 1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(7, 5))
    t = rng.normal(size=(7, 4))
    w = np.array([1.0, 0.5, 2.0, 1.5])
    _, g, _ = M.loss_and_grad(p, dims, x, t, w)
    h = 1e-6
    idx = rng.choice(len(p), size=60, replace=False)
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, dims, x, t, w)[0] - M.loss_and_grad(pm, dims, x, t, w)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-6, abs=1e-9), i


def test_gelu_tanh_form():
    from scipy.special import erf
    z = np.linspace(-6, 6, 1001)
    exact = 0.5 * z * (1 + erf(z / math.sqrt(2)))
    assert np.abs(M.gelu(z) - exact).max() < 1e-3          # tanh approximation error
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(M.gelu_grad(z), fd, rtol=1e-6, atol=1e-8)


def test_bf16_round_to_nearest_even():
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    assert M.bf16_round(np.array([1.0 + 2 ** -8]))[0] == 1.0
    assert M.bf16_round(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6
    assert M.bf16_round(np.array([1.0 + 2 ** -8 + 2 ** -20]))[0] == 1.0 + 2 ** -7
    x = np.random.default_rng(0).normal(size=1000) * 1e3
    r = M.bf16_round(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)
    assert np.all((r.astype(np.float32).view(np.uint32) & 0xFFFF) == 0)
