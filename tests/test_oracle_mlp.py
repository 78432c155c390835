"""Pins of the closure network's CPU oracle (oracle/mlp.py) by what the
mathematics fixes: GELU's closed-form values, the backward pass against
central finite differences of the forward pass (FP64, no FP16 rounding), and
Adam's first step, which the textbook update reduces to lr * g / (|g| + eps)."""
import math

import numpy as np
import pytest

import mlp as M


def test_gelu_closed_form():
    z = np.array([0.0, 1.0, -1.0, 8.0, -8.0])
    g = M.gelu(z)
    assert g[0] == 0.0
    assert abs(g[1] - 0.8413447460685429) < 1e-15          # Phi(1)
    assert abs(g[2] + 0.15865525393145707) < 1e-15         # -1 * Phi(-1)
    assert abs(g[3] - 8.0) < 1e-12 and abs(g[4]) < 1e-12
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.abs(fd - M.gelu_grad(z)).max() < 1e-8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(M.NPAR) * 0.3
    x = rng.standard_normal((7, M.IN))
    t = rng.standard_normal((7, M.OUT))
    loss, g = M.loss_grad(p, x, t, half=False)
    idx = np.concatenate([rng.choice(np.arange(M.OB1), 6, replace=False), [M.OB1 + 3, M.OW2 + 777, M.OB2 + 5, M.OW3 + 11, M.OB3 + 24]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, x, t, half=False)[0] - M.loss_grad(pm, x, t, half=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    assert loss == pytest.approx(float(((M.forward(p, x, half=False)[0] - t) ** 2).mean()), rel=1e-14)


def test_layout_offsets():
    assert M.NPAR == 23449
    W1, b1, W2, b2, W3, b3 = M.unpack(np.arange(M.NPAR, dtype=float))
    assert W1[0, 1] == 1 and W1[1, 0] == M.IN          # row-major [out][in]
    assert b3[-1] == M.NPAR - 1
    assert np.array_equal(M.pack(W1, b1, W2, b2, W3, b3), np.arange(M.NPAR))


def test_half_rounding_is_nearest_even():
    assert M.f16(1.0 + 2.0 ** -11) == 1.0                  # tie -> even
    assert M.f16(1.0 + 3 * 2.0 ** -11) == 1.0 + 2.0 ** -9  # tie -> even (up)
    assert M.f16(65504.0) == 65504.0


def test_adam_first_step():
    g = np.array([1e-3, -2.0, 0.5, 0.0])
    p = np.zeros(4)
    p1, m, v = M.adam_step(p, g, np.zeros(4), np.zeros(4), 1, lr=0.01)
    expect = -0.01 * g / (np.abs(g) + 1e-8)
    assert np.allclose(p1, expect, rtol=1e-12, atol=0)
    assert math.isclose(m[1], -0.2) and math.isclose(v[1], 0.004)
