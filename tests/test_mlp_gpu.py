"""The warm-start MLP kernels (through the C ABI) against the float64 oracle."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import _capi
from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
@pytest.mark.parametrize("m,n,k", [(128, 128, 32), (200, 136, 72), (64, 520, 1000), (300, 40, 20), (1, 8, 8)])
def test_gemm_bf16_against_float64(ta, tb, m, n, k):
    _say(f"gemm ta={ta} tb={tb} m={m} n={n} k={k}")
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = rng.normal(size=(k, m) if ta else (m, k))
    B = rng.normal(size=(n, k) if tb else (k, n))
    Ab, Bb = M.bf16_bits(A), M.bf16_bits(B)
    C = _capi.gemm_bf16(Ab, Bb, ta, tb)
    Aop = M.bf16_value(Ab).T if ta else M.bf16_value(Ab)
    Bop = M.bf16_value(Bb).T if tb else M.bf16_value(Bb)
    ref = Aop @ Bop
    # bf16 x bf16 products are exact in fp32; only the fp32 accumulation over
    # K rounds: |err| <= K u32 (|A| |B|), u32 = 2^-24
    bound = k * 2.0 ** -24 * (np.abs(Aop) @ np.abs(Bop)) + 1e-30
    assert np.all(np.abs(C - ref) <= bound), np.max(np.abs(C - ref) / bound)
    # and the float64 product of the unrounded inputs is within the bf16
    # unit-roundoff bound (2 x 2^-8 per product)
    full = (A.T if ta else A) @ (B.T if tb else B)
    assert np.all(np.abs(C - full) <= 2 * 2.0 ** -8 * (np.abs(Aop) @ np.abs(Bop)) + bound)


def _mlp(n_in, H, n_out, seed=0):
    net = _capi.Mlp(n_in, H, n_out, seed=seed)
    p = M.init_params(n_in, H, n_out, seed=seed)
    net.set_params(p)
    back = net.get_params()
    for k in M.NAMES:
        np.testing.assert_array_equal(back[k], np.asarray(p[k], np.float32))
    return net, {k: np.asarray(v, np.float32).astype(np.float64) for k, v in p.items()}


@pytest.mark.parametrize("n_in,H,n_out,m", [(48, 64, 40, 200), (136, 256, 96, 70), (236, 128, 236, 33)])
def test_mlp_forward_against_oracle(n_in, H, n_out, m):
    _say(f"mlp forward {n_in}-{H}-{n_out} m={m}")
    net, p = _mlp(n_in, H, n_out)
    X = np.random.default_rng(1).normal(size=(m, n_in)).astype(np.float32)
    Y = net.forward(X)
    Yb, _ = M.forward(p, X.astype(np.float64), bf16_ops=True)
    Y64, _ = M.forward(p, X.astype(np.float64))
    scale = np.max(np.abs(Y64))
    # same bf16 roundings as the device: fp32-accumulation differences plus
    # rare one-ulp flips of a rounded hidden value
    assert np.max(np.abs(Y - Yb)) < 5e-3 * scale
    assert np.max(np.abs(Y - Y64)) < 5e-2 * scale


@pytest.mark.parametrize("n_in,H,n_out,m", [(48, 64, 40, 200), (136, 256, 96, 130)])
def test_mlp_gradients_against_oracle(n_in, H, n_out, m):
    _say(f"mlp grad {n_in}-{H}-{n_out} m={m}")
    net, p = _mlp(n_in, H, n_out, seed=3)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(m, n_in)).astype(np.float32)
    T = rng.normal(size=(m, n_out)).astype(np.float32)
    L, g = net.loss_grad(X, T)
    Lref, gref = M.grad(p, X.astype(np.float64), T.astype(np.float64), bf16_ops=True)
    assert abs(L - Lref) < 1e-3 * Lref
    for k in M.NAMES:
        err = np.linalg.norm(g[k] - gref[k]) / np.linalg.norm(gref[k])
        assert err < 2e-2, (k, err)


def test_mlp_training_reduces_loss():
    _say("mlp train")
    rng = np.random.default_rng(5)
    n_in, H, n_out, m = 64, 128, 32, 512
    Wt = rng.normal(size=(n_in, n_out)) / np.sqrt(n_in)
    X = rng.normal(size=(m, n_in)).astype(np.float32)
    T = (np.tanh(X @ Wt)).astype(np.float32)
    net = _capi.Mlp(n_in, H, n_out, seed=1)
    l0 = net.train_step(X, T, 1e-3)
    for _ in range(300):
        l1 = net.train_step(X, T, 1e-3)
    assert np.isfinite(l1) and l1 < 0.1 * l0, (l0, l1)
    # the trained parameters reproduce the loss through the oracle
    p = {k: v.astype(np.float64) for k, v in net.get_params().items()}
    Lref = M.loss(p, X.astype(np.float64), T.astype(np.float64))
    Ldev, _ = net.loss_grad(X, T)
    assert abs(Ldev - Lref) < 0.05 * Lref + 1e-6


def _flat(net, d):
    return net.flatten(d).astype(np.float64)


@pytest.mark.parametrize("clip_frac", [0.0, 0.3])
def test_adam_steps_with_clipping_against_oracle(clip_frac):
    """Two training steps equal the oracle's Adam applied to the device's own
    gradients (fp32 update arithmetic); the clip norm sits below both
    gradients' norms, which differ, so an unclipped update would not match."""
    _say(f"mlp adam clip_frac={clip_frac}")
    net, p = _mlp(48, 64, 40, seed=4)
    rng = np.random.default_rng(9)
    X1 = rng.normal(size=(96, 48)).astype(np.float32)
    T1 = rng.normal(size=(96, 40)).astype(np.float32)
    X2 = rng.normal(size=(96, 48)).astype(np.float32)
    T2 = (4.0 * rng.normal(size=(96, 40))).astype(np.float32)
    lr = 1e-3
    w0 = _flat(net, p)
    _, g1 = net.loss_grad(X1, T1)
    g1 = _flat(net, g1)
    clip = clip_frac * np.linalg.norm(g1)
    net.set_grad_clip(clip)
    net.train_step(X1, T1, lr)
    w1 = _flat(net, net.get_params())
    z = np.zeros_like(w0)
    r1, m1, v1 = M.adam_step(w0, z, z, g1, 1, lr, clip=clip)
    d, dr = w1 - w0, r1 - w0
    assert np.linalg.norm(d - dr) < 1e-3 * np.linalg.norm(dr)
    _, g2 = net.loss_grad(X2, T2)
    g2 = _flat(net, g2)
    if clip_frac:
        assert np.linalg.norm(g2) > 1.5 * np.linalg.norm(g1)
    net.train_step(X2, T2, lr)
    w2 = _flat(net, net.get_params())
    r2, _, _ = M.adam_step(w1, m1, v1, g2, 2, lr, clip=clip)
    d, dr = w2 - w1, r2 - w1
    assert np.linalg.norm(d - dr) < 1e-3 * np.linalg.norm(dr)
    if clip_frac:  # the unclipped update is measurably different
        _, mu, vu = M.adam_step(w0, z, z, g1, 1, lr)
        ru, _, _ = M.adam_step(w1, mu, vu, g2, 2, lr)
        assert np.linalg.norm((ru - w1) - dr) > 0.05 * np.linalg.norm(dr)


def test_mlp_rejects_bad_widths():
    with pytest.raises(ValueError):
        _capi.Mlp(0, 64, 8)
    with pytest.raises(ValueError):
        _capi.Mlp(16, -1, 8)
