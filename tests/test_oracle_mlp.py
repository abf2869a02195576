"""Pins of the MLP surrogate oracle (oracle/mlp_oracle.py) by independent
checks: GELU (tanh form) against its logistic form and x*Phi(x) from scipy, its derivative and the backward
pass against central finite differences in FP64, the packed layout against
a hand-built network, and BF16 rounding against torch's bfloat16 cast."""
import math

import numpy as np
import pytest
import torch
from scipy.stats import norm

import mlp_oracle as mo


def test_gelu_closed_form():
    x = np.linspace(-8, 8, 1601)
    # logistic form used by the kernels: x / (1 + exp(-2u))
    u = np.sqrt(2 / np.pi) * (x + 0.044715 * x ** 3)
    assert np.allclose(mo.gelu(x), x / (1 + np.exp(-2 * u)), rtol=1e-14, atol=1e-15)
    # it approximates x Phi(x) to a few 1e-4
    assert np.max(np.abs(mo.gelu(x) - x * norm.cdf(x))) < 1e-3
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    assert mo.gelu(np.array([30.0]))[0] == 30.0


def test_gelu_grad_finite_difference():
    x = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.gelu_grad(x))) < 1e-8


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-8, 8, 20000),
                        [1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(x), ref)
    # ties to even and exactly representable values
    assert mo.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    assert mo.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    assert mo.bf16(np.array([1.5]))[0] == 1.5


def test_layout_round_trip_and_size():
    assert mo.N_PARAMS == 256 * 24 + 2 * 256 * 264 + 16 * 264
    p = mo.init_params(3)
    assert np.array_equal(mo.pack(mo.unpack(p)), p)
    # padding is zero
    for (din, kpad, ld, dout, op), o in zip(mo.LAYERS, mo.offsets()):
        M = p[o:o + op * ld].reshape(op, ld)
        mask = np.ones_like(M, dtype=bool)
        mask[:dout, :din] = False
        mask[:dout, kpad] = False
        assert not M[mask].any()


def test_hand_built_network():
    # a single path x0 -> GELU -> GELU -> GELU -> y_j = c_j * h + b_j
    layers = [(np.zeros((o, i)), np.zeros(o)) for (i, _, _, o, _) in mo.LAYERS]
    layers[0][0][0, 0] = 1.0
    layers[1][0][0, 0] = 1.0
    layers[2][0][0, 0] = 1.0
    c = np.arange(1, 9, dtype=float)
    layers[3][0][:, 0] = c
    layers[3][1][:] = -0.5
    p = mo.pack(layers)
    x = np.zeros((5, 11))
    x[:, 0] = np.linspace(-2, 2, 5)
    x[:, 3] = 7.0  # ignored input
    y, _ = mo.forward(p, x, bf=False)
    g = lambda v: 0.5 * v * (1 + np.tanh(np.sqrt(2 / np.pi) * (v + 0.044715 * v ** 3)))
    h = g(g(g(x[:, 0])))
    assert np.allclose(y, h[:, None] * c[None, :] - 0.5, rtol=0, atol=1e-14)


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_finite_difference(seed):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed + 10)
    # non-zero biases so their gradients are exercised
    layers = mo.unpack(p)
    layers = [(W, rng.standard_normal(b.shape) * 0.1) for W, b in layers]
    p = mo.pack(layers)
    x = rng.standard_normal((6, 11))
    t = rng.standard_normal((6, 8))
    loss, g = mo.loss_and_grad(p, x, t, bf=False)
    live = np.flatnonzero(mo.pack([(np.ones_like(W), np.ones_like(b)) for W, b in layers]))
    pick = rng.choice(live, 40, replace=False)
    # every layer's weights and biases
    for (din, kpad, ld, dout, op), o in zip(mo.LAYERS, mo.offsets()):
        pick = np.append(pick, [o + 0 * ld + 0, o + (dout - 1) * ld + kpad])
    for i in pick:
        h = 1e-5 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = mo.loss_and_grad(pp, x, t, bf=False)[0]
        lm = mo.loss_and_grad(pm, x, t, bf=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
    # padding gets no gradient
    assert not g[np.setdiff1d(np.arange(mo.N_PARAMS), live)].any()


def test_loss_is_mean_squared_error():
    p = mo.init_params(4)
    rng = np.random.default_rng(9)
    x = rng.standard_normal((8, 11))
    y, _ = mo.forward(p, x, bf=False)
    t = y + 0.5
    loss, _ = mo.loss_and_grad(p, x, t, bf=False)
    assert math.isclose(loss, 0.25, rel_tol=1e-12)
