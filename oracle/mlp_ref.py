"""The oracle of the warm-start MLP: float64 numpy forward and backward
(test infrastructure only; the product never imports it).

The reference has no learned component (its ML scripts time random-init
"PFdelta" models, 20260831_neurips_paper/scripts/ml_injection.py), so this is
a restatement of the textbook model the device kernels implement:

    Z1 = X W1 + b1,  H1 = gelu(Z1)
    Z2 = H1 W2 + b2, H2 = gelu(Z2)
    Y  = H2 W3 + b3
    loss = mean((Y - T)^2) over all M x n_out entries

with the exact (erf) GELU.  ``forward(..., bf16=True)`` rounds what the
device stores in bf16 (X, the weights' GEMM copies, H1, H2 and the backward
signals) to bf16, round-to-nearest-even, so the device result can be checked
against it with fp32-accumulation tolerances.

Pinning (tests/test_mlp_oracle.py): ``grad`` equals central differences of
``loss``; ``gelu`` equals scipy's ``erf`` form and its known values; ``bf16``
rounding equals the IEEE round-to-nearest-even of the upper 16 bits;
``adam_step`` reduces to the sign update lr g / (|g| + eps) whenever the first
and second moments are those of one repeated gradient (Kingma & Ba 2015,
Algorithm 1), and its clipping keeps the direction and caps the norm.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

NAMES = ("W1", "b1", "W2", "b2", "W3", "b3")


def bf16_bits(x):
    """float -> bf16 bit pattern (uint16), round to nearest even."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    nan = np.isnan(f)
    if nan.any():
        r = np.where(nan, np.uint16(0x7FC0), r)
    return r


def bf16_value(bits):
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16(x):
    """x rounded to bf16, as float64."""
    return bf16_value(bf16_bits(x))


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def init_params(n_in, H, n_out, seed=0):
    rng = np.random.default_rng(seed)

    def glorot(a, b):
        lim = np.sqrt(6.0 / (a + b))
        return rng.uniform(-lim, lim, size=(a, b))

    return {"W1": glorot(n_in, H), "b1": rng.normal(0, 0.1, H), "W2": glorot(H, H), "b2": rng.normal(0, 0.1, H),
            "W3": glorot(H, n_out), "b3": rng.normal(0, 0.1, n_out)}


def forward(p, X, bf16_ops=False):
    """Returns Y and the cache (Z1, H1, Z2, H2) used by ``grad``."""
    q = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    X = q(X)
    Z1 = X @ q(p["W1"]) + p["b1"]
    H1 = q(gelu(Z1))
    Z2 = H1 @ q(p["W2"]) + p["b2"]
    H2 = q(gelu(Z2))
    Y = H2 @ q(p["W3"]) + p["b3"]
    return Y, (X, Z1, H1, Z2, H2)


def loss(p, X, T):
    Y, _ = forward(p, X)
    return float(np.mean((Y - T) ** 2))


def grad(p, X, T, bf16_ops=False):
    """(loss, {name: dloss/dname}) by back-propagation."""
    q = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    Y, (Xq, Z1, H1, Z2, H2) = forward(p, X, bf16_ops)
    n_el = Y.size
    L = float(np.mean((Y - T) ** 2))
    dY = 2.0 * (Y - T) / n_el
    dYq = q(dY)
    g = {"W3": H2.T @ dYq, "b3": dY.sum(0)}
    dZ2 = (dYq @ q(p["W3"]).T) * gelu_grad(Z2)
    dZ2q = q(dZ2)
    g["W2"] = H1.T @ dZ2q
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2q @ q(p["W2"]).T) * gelu_grad(Z1)
    dZ1q = q(dZ1)
    g["W1"] = Xq.T @ dZ1q
    g["b1"] = dZ1.sum(0)
    return L, g


def features(S, mu, sigma, mask=None):
    """Per-bus (P, Q) interleaved, normalised; then the 0/1 outage mask."""
    S = np.atleast_2d(S)
    F = np.empty((S.shape[0], 2 * S.shape[1]))
    F[:, 0::2] = S.real
    F[:, 1::2] = S.imag
    F = (F - mu) / sigma
    if mask is not None:
        F = np.concatenate([F, np.asarray(mask, np.float64)], axis=1)
    return F


def targets(V):
    """Per-bus (Va, Vm) interleaved."""
    V = np.atleast_2d(V)
    T = np.empty((V.shape[0], 2 * V.shape[1]))
    T[:, 0::2] = np.angle(V)
    T[:, 1::2] = np.abs(V)
    return T


def warm_v(Y, mu, sigma, kind, vfix):
    """Start voltages from normalised outputs: |V| fixed at PV/slack (kind 0/2),
    angle 0 at the slack (kind 2)."""
    Yd = Y * sigma + mu
    va = Yd[:, 0::2].copy()
    vm = Yd[:, 1::2].copy()
    kind = np.asarray(kind)
    vm[:, kind != 1] = np.asarray(vfix)[kind != 1]
    va[:, kind == 2] = 0.0
    return vm * np.exp(1j * va)


def adam_step(w, m, v, g, step, lr, clip=0.0, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) of the flat parameter
    vector w with moments m, v at step number ``step`` (1-based), after
    scaling g to norm ``clip`` when its norm exceeds it (0: no clipping).
    Returns (w, m, v), new arrays."""
    g = np.asarray(g, np.float64)
    if clip > 0.0:
        nrm = float(np.sqrt(np.sum(g * g)))
        if nrm > clip:
            g = g * (clip / nrm)
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
