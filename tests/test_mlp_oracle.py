"""Pins of the network checker (oracle/mlp_oracle.py): its backpropagation
against finite differences of its loss, its BF16 rounding against the
library's host rounding and the IEEE definition, and its Adam step against
the closed form of the first step."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _net(dims, seed):
    rng = np.random.default_rng(seed)
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    return rng.standard_normal(n) * 0.5


def test_backprop_matches_finite_differences():
    dims = (3, 8, 8, 5)
    flat = _net(dims, 1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((6, 3))
    t = rng.standard_normal((6, 4))
    lab = (rng.random(6) > 0.5).astype(np.float64)
    loss, grads = mo.loss_and_grads(mo.split(flat, dims), x, t, lab, 0.7)
    g = mo.flatten(grads)
    h = 1e-6
    fd = np.empty_like(flat)
    for i in range(flat.size):
        e = np.zeros_like(flat)
        e[i] = h
        lp, _ = mo.loss_and_grads(mo.split(flat + e, dims), x, t, lab, 0.7)
        lm, _ = mo.loss_and_grads(mo.split(flat - e, dims), x, t, lab, 0.7)
        fd[i] = (lp - lm) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_gelu_derivative():
    z = np.linspace(-5, 5, 41)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(z), fd, rtol=1e-7, atol=1e-9)
    # GELU(x) = x Phi(x): Phi(0) = 1/2, Phi(1.96) ~ 0.975
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    assert mo.gelu(np.array([1.96]))[0] / 1.96 == pytest.approx(0.9750021, abs=1e-6)


def test_bf16_rounding():
    # exact values, halfway cases to even, and the float32 definition
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5,
                  3.0e38, np.inf, -np.inf], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5,
                     2.9999e38, np.inf, -np.inf])
    got = mo.bf16(x)
    np.testing.assert_allclose(got[:5], want[:5], rtol=0)
    assert abs(got[5] - 3.0e38) / 3.0e38 < 2 ** -8
    assert np.isinf(got[6]) and np.isinf(got[7])
    assert np.isnan(mo.bf16(np.array([np.nan], np.float32)))[0]
    rng = np.random.default_rng(3)
    r = rng.standard_normal(10000).astype(np.float32)
    err = np.abs(mo.bf16(r) - r) / np.abs(r)
    assert err.max() <= 2.0 ** -8  # half an ulp of an 8-bit significand


def test_bf16_rounding_matches_library():
    import inovesa_amd as V
    rng = np.random.default_rng(4)
    r = (rng.standard_normal(5000) * 10.0 ** rng.integers(-20, 20, 5000)
         ).astype(np.float32)
    assert np.array_equal(V.round_bf16(r).astype(np.float64), mo.bf16(r))


def test_adam_first_step_closed_form():
    # after one step m^ = g and v^ = g^2: p1 = p0 - lr g / (|g| + eps)
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(50)
    g = rng.standard_normal(50)
    p1, m, v = mo.adam(p0, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    np.testing.assert_allclose(p1, p0 - 1e-3 * g / (np.abs(g) + 1e-8),
                               rtol=1e-14)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-14)
    np.testing.assert_allclose(v, 0.001 * g * g, rtol=1e-14)
