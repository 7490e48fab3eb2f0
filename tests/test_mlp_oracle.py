"""Pins the warm-start MLP oracle (oracle/mlp_ref.py) on the CPU."""
import numpy as np
import pytest
from scipy.special import erf

from oracle import mlp_ref as M


def test_bf16_rounding_is_round_to_nearest_even():
    # exactly representable values pass through
    for v in (0.0, 1.0, -2.5, 0.15625, 3.0e38):
        assert M.bf16(v) == np.float32(v) or abs(M.bf16(v) / v - 1) < 2 ** -8
    # ties: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7 -> even (1.0)
    assert M.bf16(1.0 + 2 ** -8) == 1.0
    # 1 + 3 * 2^-8 lies halfway between 1 + 2^-7 and 1 + 2^-6 -> even mantissa (1 + 2^-6)
    assert M.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    # above the tie rounds up
    assert M.bf16(1.0 + 2 ** -8 + 2 ** -12) == 1.0 + 2 ** -7
    # relative error bound 2^-8 on random values
    x = np.random.default_rng(0).normal(size=10000) * 10.0 ** np.random.default_rng(1).uniform(-5, 5, 10000)
    assert np.max(np.abs(M.bf16(x) / x - 1)) <= 2 ** -8


def test_gelu_closed_form_and_values():
    z = np.linspace(-6, 6, 101)
    np.testing.assert_allclose(M.gelu(z), z * 0.5 * (1 + erf(z / np.sqrt(2))), rtol=1e-15)
    assert M.gelu(0.0) == 0.0
    # gelu(1) = Phi(1) = 0.8413447460685429
    assert abs(M.gelu(1.0) - 0.8413447460685429) < 1e-15
    # derivative against central differences
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), atol=1e-8)


def test_backprop_equals_central_differences():
    rng = np.random.default_rng(3)
    n_in, H, n_out, m = 6, 5, 4, 7
    p = M.init_params(n_in, H, n_out, seed=2)
    X = rng.normal(size=(m, n_in))
    T = rng.normal(size=(m, n_out))
    L, g = M.grad(p, X, T)
    assert abs(L - M.loss(p, X, T)) < 1e-14
    h = 1e-6
    for name in M.NAMES:
        num = np.zeros_like(p[name])
        it = np.nditer(p[name], flags=["multi_index"])
        for _ in it:
            i = it.multi_index
            q = {k: v.copy() for k, v in p.items()}
            q[name][i] += h
            lp = M.loss(q, X, T)
            q[name][i] -= 2 * h
            lm = M.loss(q, X, T)
            num[i] = (lp - lm) / (2 * h)
        np.testing.assert_allclose(g[name], num, rtol=1e-6, atol=1e-9, err_msg=name)


def test_bf16_forward_is_close_to_float64():
    rng = np.random.default_rng(4)
    p = M.init_params(32, 24, 16, seed=5)
    X = rng.normal(size=(50, 32))
    Y64, _ = M.forward(p, X)
    Yb, _ = M.forward(p, X, bf16_ops=True)
    # three layers of 2^-8 relative rounding on operands
    assert np.max(np.abs(Yb - Y64)) < 0.05 * np.max(np.abs(Y64))


def test_features_targets_warm_v_round_trip():
    rng = np.random.default_rng(6)
    n = 5
    V = (1 + rng.normal(0, 0.02, (3, n))) * np.exp(1j * rng.normal(0, 0.1, (3, n)))
    T = M.targets(V)
    kind = np.array([2, 0, 1, 1, 1])
    vfix = np.abs(V[0])
    W = M.warm_v(T, np.zeros(2 * n), np.ones(2 * n), kind, vfix)
    # PQ buses reproduce V exactly; PV keep the set-point magnitude; slack angle 0
    np.testing.assert_allclose(W[:, 2:], V[:, 2:], rtol=1e-14)
    np.testing.assert_allclose(np.abs(W[:, 1]), vfix[1])
    np.testing.assert_allclose(np.angle(W[:, 0]), 0.0)
    S = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    F = M.features(S, np.zeros(2 * n), np.ones(2 * n), mask=np.ones((3, 2)))
    assert F.shape == (3, 2 * n + 2)
    np.testing.assert_array_equal(F[:, 0:2 * n:2], S.real)
    np.testing.assert_array_equal(F[:, 1:2 * n:2], S.imag)


def test_adam_sign_update_and_clipping():
    rng = np.random.default_rng(4)
    n, lr = 50, 1e-3
    w0 = rng.normal(size=n)
    g = rng.normal(size=n) * 10.0 ** rng.uniform(-3, 3, size=n)
    z = np.zeros(n)
    # step 1: the bias-corrected moments are g and g^2
    w1, m1, v1 = M.adam_step(w0, z, z, g, 1, lr)
    np.testing.assert_allclose(w0 - w1, lr * g / (np.abs(g) + 1e-8), rtol=1e-12)
    # the same gradient again: still the sign update
    w2, _, _ = M.adam_step(w1, m1, v1, g, 2, lr)
    np.testing.assert_allclose(w1 - w2, lr * g / (np.abs(g) + 1e-8), rtol=1e-9)
    # clipping keeps the direction and caps the norm; at or above the norm it
    # is a no-op
    nrm = np.linalg.norm(g)
    _, mc, _ = M.adam_step(w0, z, z, g, 1, lr, clip=0.25 * nrm)
    np.testing.assert_allclose(mc / 0.1, 0.25 * g, rtol=1e-12)
    _, mn, _ = M.adam_step(w0, z, z, g, 1, lr, clip=nrm * (1 + 1e-12))
    np.testing.assert_array_equal(mn, m1)
    # a second, different gradient: b1 weighs the first moment, b2 the second
    h = rng.normal(size=n)
    _, m2, v2 = M.adam_step(w1, m1, v1, h, 2, lr)
    np.testing.assert_allclose(m2, 0.9 * 0.1 * g + 0.1 * h, rtol=1e-12)
    np.testing.assert_allclose(v2, 0.999 * 0.001 * g * g + 0.001 * h * h, rtol=1e-12)
