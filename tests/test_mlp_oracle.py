"""Pins of the CPU oracle of the tile-skip network (oracle/mlp_oracle.py).

The gradient is checked against central finite differences of the loss with
weight rounding off (the rounded loss is piecewise constant in the weights),
bfloat16 rounding against hand-worked values, and the parameter layout
against the count the C ABI declares.
"""

import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_parameter_count_matches_the_header():
    import re

    from helpers import ROOT

    text = (ROOT / "proton-dose-m1_amd" / "csrc" / "m1_mlp.h").read_text()
    assert M.NPARAM == 11 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 9153
    assert re.search(r"constexpr int NIN = 11;", text) and re.search(r"constexpr int NH = 64;", text)


def test_bf16_rounding_ties_to_even():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to the even mantissa (1)
    assert M.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert M.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6
    # just above a tie rounds up
    assert M.bf16(np.float32(1.0 + 2 ** -8 + 2 ** -20)) == 1.0 + 2 ** -7
    # exactly representable values are unchanged, signs kept
    for v in (0.0, -2.5, 0.15625, 3.0e38):
        assert M.bf16(np.float32(v)) == np.float64(np.float32(v)) or v == 3.0e38
    assert np.isnan(M.bf16(np.float32(np.nan)))


def test_gradient_against_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(rng).astype(np.float64)
    p[M.SIZES[0][0] * 64:M.SIZES[0][0] * 64 + 64] = rng.normal(0, 0.1, 64)  # nonzero b1
    x = rng.normal(size=(24, 11))
    y = (rng.random(24) < 0.5).astype(np.float64)

    def f(q):
        z, _ = M.forward(q, x, round_weights=False)
        return float(np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))

    # float64 parameters: loss_grad's unpack casts to float32, so use a
    # float32-exact vector and compare on float32 steps
    p = p.astype(np.float32).astype(np.float64)
    L, g = M.loss_grad(p.astype(np.float32), x, y, round_weights=False)
    assert L == pytest.approx(f(p.astype(np.float32)), rel=1e-12)
    picks = np.concatenate([rng.choice(M.NPARAM, 60, replace=False), [M.NPARAM - 1, 0, 704]])
    for k in picks:
        h = 2.0 ** -10
        qp = p.copy(); qp[k] += h
        qm = p.copy(); qm[k] -= h
        fd = (f(qp.astype(np.float32)) - f(qm.astype(np.float32))) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=2e-3, abs=1e-7), k


def test_loss_of_a_hand_sized_case():
    """All weights zero except the output bias: z = b4 for every sample."""
    p = np.zeros(M.NPARAM, dtype=np.float32)
    p[-1] = 0.5
    x = np.ones((3, 11))
    y = np.array([1.0, 0.0, 1.0])
    expect = np.mean([np.log1p(np.exp(-0.5)), 0.5 + np.log1p(np.exp(-0.5)), np.log1p(np.exp(-0.5))])
    assert M.loss(p, x, y) == pytest.approx(expect, rel=1e-14)
    L, g = M.loss_grad(p, x, y)
    s = 1.0 / (1.0 + np.exp(-0.5))
    assert g[-1] == pytest.approx(np.mean(s - y), rel=1e-14)
    assert np.all(g[:-1] == 0.0)
