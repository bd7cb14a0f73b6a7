"""CPU oracle of the tile-skip network (TEST INFRASTRUCTURE ONLY).

Restates, in float64 numpy, the network the library evaluates and trains on
the device (include/m1march.h, m1_mlp_*): 11 inputs, three hidden layers of
64 ReLU units, one logit; weights applied through their bfloat16 values
(round to nearest even from float32), biases in float32; the loss is the
mean binary cross-entropy of sigmoid(logit).  No reference counterpart: the
reference has no learned component.  Only the tests import this module.

Pinned in tests/test_mlp_oracle.py: the gradient against central finite
differences (with bf16 rounding off, so the loss is smooth), bf16 rounding
against hand-worked ties, and the loss against a direct evaluation.
"""

from __future__ import annotations

import numpy as np

NIN = 11
NH = 64
SIZES = [(NIN, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH,), ()]
NPARAM = sum(int(np.prod(s)) if s else 1 for s in SIZES)


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    nan = (u & 0x7F800000) == 0x7F800000
    nan &= (u & 0x7FFFFF) != 0
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, round_weights=True):
    """[W1, b1, W2, b2, W3, b3, W4, b4] from the flat parameter vector."""
    p = np.asarray(params, dtype=np.float32)
    out, k = [], 0
    for idx, shape in enumerate(SIZES):
        n = int(np.prod(shape)) if shape else 1
        v = p[k:k + n]
        k += n
        is_weight = idx % 2 == 0
        v = bf16(v) if (is_weight and round_weights) else v.astype(np.float64)
        out.append(v.reshape(shape) if shape else v[0])
    return out


def forward(params, x, round_weights=True):
    """Logits (n,) and post-ReLU activations [h1, h2, h3]."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params, round_weights)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    h3 = np.maximum(h2 @ W3 + b3, 0.0)
    z = h3 @ W4 + b4
    return z, (h1, h2, h3)


def probabilities(params, x, round_weights=True):
    z, _ = forward(params, x, round_weights)
    return 1.0 / (1.0 + np.exp(-z))


def loss(params, x, y, round_weights=True):
    z, _ = forward(params, x, round_weights)
    y = np.asarray(y, dtype=np.float64)
    return float(np.mean(np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z)))))


def loss_grad(params, x, y, round_weights=True):
    """Mean loss and its gradient with respect to the flat parameters
    (weights enter through their bf16 values when round_weights)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params, round_weights)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = len(x)
    z, (h1, h2, h3) = forward(params, x, round_weights)
    L = float(np.mean(np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z)))))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / n
    d3 = np.outer(dz, W4) * (h3 > 0)
    d2 = (d3 @ W3.T) * (h2 > 0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    grads = [x.T @ d1, d1.sum(0), h1.T @ d2, d2.sum(0), h2.T @ d3, d3.sum(0), h3.T @ dz,
             np.array(dz.sum())]
    return L, np.concatenate([np.ravel(g) for g in grads])


def init_params(rng, scale=1.0):
    """He-style initial parameters as float32."""
    parts = []
    for idx, shape in enumerate(SIZES):
        if idx % 2 == 0:
            fan_in = shape[0] if shape else 1
            parts.append(rng.normal(0.0, scale * np.sqrt(2.0 / fan_in), size=shape).ravel())
        else:
            parts.append(np.zeros(int(np.prod(shape)) if shape else 1))
    return np.concatenate(parts).astype(np.float32)
