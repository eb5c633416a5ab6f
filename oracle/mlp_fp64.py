"""FP64 reference of the surrogate perceptron (numpy).  TEST INFRASTRUCTURE ONLY.

Same network and loss as include/nrg_mlp.h: layers W_l (fan_in x fan_out),
b_l; GELU in tanh form (Hendrycks & Gimpel 2016) on hidden layers; loss
L = sum_rows sum_c w_c (y_c - t_c)^2 / n.  With `bf16=True` it rounds to
bfloat16 (round to nearest even) exactly where the GPU kernels store bf16:
the weights, the input, every hidden activation, the output gradient and
every back-propagated pre-activation gradient; everything else is FP64.
The backward pass is pinned by central finite differences in FP64
(tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

C_GELU = math.sqrt(2.0 / math.pi)
A_GELU = 0.044715


def bf16_round(x):
    """Round float64 values to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    nan = np.isnan(f)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + A_GELU * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + A_GELU * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * A_GELU * z * z)


def unpack(params, dims):
    """Packed FP32/FP64 parameter vector -> [(W, b)] per layer."""
    n_in, hidden, layers, n_out = dims
    fan = [n_in] + [hidden] * layers + [n_out]
    out, q = [], 0
    for i in range(layers + 1):
        W = np.asarray(params[q:q + fan[i] * fan[i + 1]], dtype=np.float64).reshape(fan[i], fan[i + 1])
        q += fan[i] * fan[i + 1]
        b = np.asarray(params[q:q + fan[i + 1]], dtype=np.float64)
        q += fan[i + 1]
        out.append((W, b))
    assert q == len(params)
    return out


def forward(params, dims, x, bf16=False):
    """Returns (y, cache)."""
    layers = unpack(params, dims)
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    h = rnd(x)
    cache = [h]
    zs = []
    for i, (W, b) in enumerate(layers):
        z = h @ rnd(W) + b
        if i < len(layers) - 1:
            zs.append(z)
            h = rnd(gelu(z))
            cache.append(h)
        else:
            y = z
    return y, (cache, zs)


def loss_and_grad(params, dims, x, t, w=None, bf16=False):
    """(loss, gradient in the packed order, y)."""
    layers = unpack(params, dims)
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, (hs, zs) = forward(params, dims, x, bf16)
    n = x.shape[0]
    wv = np.ones(y.shape[1]) if w is None else np.asarray(w, dtype=np.float64)
    d = y - t
    loss = float((wv * d * d).sum() / n)
    dz = rnd(2.0 * wv * d / n)
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        gW = hs[i].T @ dz
        gb = dz.sum(axis=0)
        grads[i] = (gW, gb)
        if i > 0:
            dz = rnd((dz @ rnd(W).T) * gelu_grad(zs[i - 1]))
    flat = np.concatenate([np.concatenate([g.ravel(), b]) for g, b in grads])
    return loss, flat, y
