"""Pins the closure oracle (oracle/mlp_oracle.py) on the CPU: its gradient
against finite differences, GELU and Adam against closed forms, its feature
restatement against fields with known answers, its BF16 rounding against
hand-computed bit patterns."""

import math

import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _data(n=29, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, mo.IN)), rng.standard_normal((n, mo.OUT))


def test_gradient_matches_central_differences():
    p = mo.init_params(3).astype(np.float64)
    p[mo.O_B1:mo.O_W2] = 0.1  # non-zero biases exercise their gradients too
    X, T = _data()
    n = len(X)
    _, g = mo.loss_and_grad(p, X, T, bf16_round=False)

    def mse(q):
        return mo.loss_and_grad(q, X, T, bf16_round=False)[0] / (3 * n)

    rng = np.random.default_rng(1)
    # some entries of every block (padding entries have zero gradient)
    blocks = [(mo.O_W1, mo.O_B1), (mo.O_B1, mo.O_W2), (mo.O_W2, mo.O_B2), (mo.O_B2, mo.O_W3),
              (mo.O_W3, mo.O_W3 + mo.OUT * mo.HID), (mo.O_B3, mo.O_B3 + mo.OUT)]
    for lo, hi in blocks:
        for i in rng.choice(np.arange(lo, hi), min(6, hi - lo), replace=False):
            h = 1e-6
            q1, q2 = p.copy(), p.copy()
            q1[i] += h
            q2[i] -= h
            fd = (mse(q1) - mse(q2)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # padded inputs and outputs carry no gradient
    W1g = g[mo.O_W1:mo.O_B1].reshape(mo.HID, mo.IN_PAD)
    assert np.all(W1g[:, mo.IN:] == 0)
    assert np.all(g[mo.O_W3 + mo.OUT * mo.HID:mo.O_B3] == 0) and np.all(g[mo.O_B3 + mo.OUT:] == 0)


def test_gelu_closed_forms():
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    assert mo.gelu_grad(np.array([0.0]))[0] == 0.5
    x = np.linspace(-3, 3, 13)
    np.testing.assert_allclose(mo.gelu(x) - mo.gelu(-x), x, atol=1e-15)
    # derivative of the closed form by central differences
    h = 1e-6
    np.testing.assert_allclose(mo.gelu_grad(x), (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h), atol=1e-8)
    assert abs(mo.gelu(np.array([1.0]))[0] - 0.5 * (1 + math.erf(1 / math.sqrt(2)))) < 1e-15


def test_adam_first_step_moves_each_parameter_by_lr():
    rng = np.random.default_rng(2)
    p = rng.standard_normal(50)
    g = rng.standard_normal(50)
    p1, m, v = mo.adam(p, g, np.zeros(50), np.zeros(50), 1, lr=1e-3, eps=0.0)
    np.testing.assert_allclose(np.abs(p1 - p), 1e-3, rtol=1e-12)
    assert np.all(np.sign(p - p1) == np.sign(g))


def test_bf16_rounding_bits():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3*2^-9 rounds up
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.140625])
    np.testing.assert_array_equal(mo.bf16(vals), [1.0, 1.0, 1.0 + 2 ** -7, -2.5, 3.140625])


def test_features_on_known_fields():
    nx, ny, nz = 12, 12, 12
    rng = np.random.default_rng(4)
    x = rng.uniform(4.0, 8.0, (20, 3))
    v = rng.standard_normal((20, 3)) * 0.01
    solid = np.zeros((nx, ny, nz), np.int32)
    idx = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).astype(float)
    B = np.array([[0.01, -0.02, 0.005], [0.0, 0.03, 0.01], [-0.01, 0.0, 0.02]])
    a = np.array([0.02, -0.01, 0.005])
    u = a + idx @ B.T
    phi = 0.1 + idx @ np.array([0.02, -0.04, 0.04])
    f = mo.closure_features(x, v, u, solid, phi)
    # linear field: interpolated value and gradient reproduced exactly
    np.testing.assert_allclose(f[:, 0:3], a + x @ B.T - v, atol=1e-12)
    np.testing.assert_allclose(f[:, 3:12], np.broadcast_to(B.reshape(-1), (20, 9)), atol=1e-12)
    np.testing.assert_allclose(f[:, 12], 0.1 + x @ np.array([0.02, -0.04, 0.04]), atol=1e-12)
    g = np.array([0.02, -0.04, 0.04])
    np.testing.assert_allclose(f[:, 13:16], np.broadcast_to(-g / np.linalg.norm(g), (20, 3)), atol=1e-12)
    np.testing.assert_allclose(f[:, 16], 1.0, atol=1e-12)
    np.testing.assert_allclose(f[:, 17], np.linalg.norm(f[:, 0:3], axis=1), atol=1e-15)
    # a solid node drops out of the support and the velocity, not out of phi
    solid[int(x[0, 0]), int(x[0, 1]), int(x[0, 2])] = 1
    f2 = mo.closure_features(x[:1], v[:1], u, solid, phi)
    assert f2[0, 16] < 1.0 and f2[0, 12] == pytest.approx(f[0, 12], abs=1e-12)
