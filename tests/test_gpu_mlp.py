"""The tile-skip network on an MI355X (m1_mlp_*) against its CPU oracle.

Forward probabilities, loss and parameter gradients of the device kernels
(bf16 weights, float32 accumulation) are compared with the float64 oracle
evaluated through the same bf16 weights: tolerances are float32-accumulation
level.  One Adam step is checked against its closed form, and training on a
synthetic rule must learn it.
"""

import numpy as np
import pytest

from oracle import mlp_oracle as M
from proton_solver import _native
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    p = M.init_params(rng)
    p[704:768] = rng.normal(0, 0.1, 64).astype(np.float32)  # nonzero b1
    x = rng.normal(size=(n, 11)).astype(np.float32)
    y = (rng.random(n) < 0.4).astype(np.float32)
    return p, x, y


def test_forward_matches_oracle():
    p, x, _ = _problem(1000, 1)
    net = _native.Mlp(0, p)
    assert np.array_equal(net.get_params(), p)
    got = net.forward(x)
    ref = M.probabilities(p, x)
    err = np.abs(got - ref).max()
    say("mlp forward max |dp|", err)
    assert err < 2e-6


def test_loss_and_gradient_match_oracle():
    p, x, y = _problem(300, 2)
    net = _native.Mlp(0, p)
    L, g = net.loss_grad(x, y)
    Lr, gr = M.loss_grad(p, x, y)
    rel = np.abs(g - gr).max() / np.abs(gr).max()
    say("mlp loss", L, Lr, "grad max rel err", rel)
    assert L == pytest.approx(Lr, rel=1e-6)
    assert rel < 1e-5
    # every layer is reached (no silently zero block)
    for a, b in [(0, 704), (704, 768), (768, 4864), (4864, 4928), (4928, 9024), (9024, 9088),
                 (9088, 9152), (9152, 9153)]:
        assert np.abs(gr[a:b]).max() > 0 and np.abs(g[a:b]).max() > 0


def test_one_adam_step_closed_form():
    p, x, y = _problem(512, 3)
    net = _native.Mlp(0, p)
    _, g = M.loss_grad(p, x, y)
    lr = 1e-3
    net.train(x, y, epochs=1, batch=512, lr=lr, seed=0)  # one minibatch of all samples
    got = net.get_params()
    # step 1 of Adam: m_hat = g, v_hat = g^2  ->  p - lr * g / (|g| + 1e-8)
    expect = (p.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)).astype(np.float32)
    big = np.abs(g) > 1e-5
    assert np.allclose(got[big], expect[big], rtol=0, atol=2e-6)


def test_training_learns_a_rule():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(20000, 11)).astype(np.float32)
    y = ((x[:, 0] + 0.5 * x[:, 1] - x[:, 2]) > 0.3).astype(np.float32)
    net = _native.Mlp(0, M.init_params(rng))
    L0, _ = net.loss_grad(x[:4096], y[:4096])
    L = net.train(x[:16000], y[:16000], epochs=15, batch=512, lr=2e-3, seed=7)
    acc = np.mean((net.forward(x[16000:]) > 0.5) == (y[16000:] > 0.5))
    say("mlp training loss", L0, "->", L, "held-out accuracy", acc)
    assert L < 0.25 * L0 and acc > 0.97
