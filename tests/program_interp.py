"""CPU interpreter of the factorization / substitution programs.

It executes, row-vectorised in numpy, what the device kernels ``k_factor`` and
``k_backward`` execute per lane: the same program records, the same A/U
stream positions and the same LDS slot traffic (including the zero slot a
missing source reads and the trash slot an unneeded result goes to).  Tests
use it to check the host symbolic layer on the CPU against scipy,
independently of the GPU.
"""
from __future__ import annotations

import numpy as np

REF_SLOT = 0x7FFF
REF_FIRST = 0x8000


def frec_words(c):
    return (1 + ((c + 1) * (c + 2) + 1) // 2 + 3) // 4 * 4


def brec_words(c):
    return (2 + (c + 2) // 2 + 3) // 4 * 4


def _refs(prog, p, head, words):
    w = np.asarray(prog[p + head:p + words], dtype=np.int64) & 0xFFFFFFFF
    return np.stack([w & 0xFFFF, w >> 16], axis=1).ravel()


def factor_forward(info, prog, A, pivot_eps=0.0, scale=None):
    """A: [rows, a_len] dense stream -> (U stream [rows, u_len], clamped [rows])."""
    rows = A.shape[0]
    S = info["slots_f"]
    sl = np.full((S + 2, rows), np.nan)
    sl[S] = 0.0  # zero slot
    U = np.zeros((rows, info["u_len"]))
    eps = pivot_eps * (np.ones(rows) if scale is None else scale)
    clamped = np.zeros(rows, np.int64)
    p = ap = up = 0
    for _k in range(info["d"]):
        c = prog[p] & 0xFFFF
        words = frec_words(c)
        r = _refs(prog, p, 1, words)
        fk = A[:, ap] + sl[r[1] & REF_SLOT]
        dv = A[:, ap + 1] + sl[r[0] & REF_SLOT]
        Lv = [A[:, ap + 2 + q] + sl[r[2 + q] & REF_SLOT] for q in range(c)]
        Uv = [A[:, ap + 2 + c + q] + sl[r[2 + c + q] & REF_SLOT] for q in range(c)]
        ap += 2 + 2 * c
        small = np.abs(dv) < eps
        dv = np.where(small, np.where(dv < 0, -eps, eps), dv)
        clamped += small
        lv = [v / dv for v in Lv]
        U[:, up] = fk
        U[:, up + 1] = dv
        for q in range(c):
            U[:, up + 2 + q] = Uv[q]
        up += c + 2
        base = 2 + 2 * c
        for a in range(c):
            for b in range(c):
                ref = r[base + a * c + b]
                old = 0.0 if ref & REF_FIRST else sl[ref & REF_SLOT]
                sl[ref & REF_SLOT] = old - lv[a] * Uv[b]
        base += c * c
        for a in range(c):
            ref = r[base + a]
            old = 0.0 if ref & REF_FIRST else sl[ref & REF_SLOT]
            sl[ref & REF_SLOT] = old - lv[a] * fk
        assert np.all(sl[S] == 0.0)
        p += words
    assert p == len(prog) and ap == info["a_len"] and up == info["u_len"]
    return U, clamped


def backward(info, prog, U):
    """U stream -> x [rows, d] in input labels."""
    rows = U.shape[0]
    d = info["d"]
    S = info["slots_b"]
    sl = np.full((S + 2, rows), np.nan)
    x_out = np.zeros((rows, d))
    p = 0
    up = info["u_len"]
    for _q in range(d):
        c = prog[p] & 0xFFFF
        words = brec_words(c)
        perm_k = prog[p + 1]
        r = _refs(prog, p, 2, words)
        up -= c + 2
        s = U[:, up].copy()
        for q in range(c):
            s -= U[:, up + 2 + q] * sl[r[q]]
        x = s / U[:, up + 1]
        x_out[:, perm_k] = x
        sl[r[c]] = x
        p += words
    assert p == len(prog) and up == 0
    return x_out


def polar_update(V, dx, pvpq, pq):
    """k_update's formula: V <- V ((|V| + dVm) / |V|) exp(j dVa) per bus."""
    V = np.array(V, dtype=np.complex128)
    npvpq = len(pvpq)
    dva = np.zeros(V.shape)
    dvm = np.zeros(V.shape)
    dva[..., pvpq] = dx[..., :npvpq]
    dvm[..., pq] = dx[..., npvpq:]
    m = np.abs(V)
    return V * ((m + dvm) / m) * np.exp(1j * dva)


def astream_from_csr(info, apos, fpos, vals, rhs):
    """Raw A stream from CSR values [rows, nnz] and right-hand sides [rows, d];
    fill positions stay 0."""
    vals = np.atleast_2d(vals)
    rhs = np.atleast_2d(rhs)
    A = np.zeros((vals.shape[0], info["a_len"]))
    m = apos >= 0
    A[:, apos[m]] = vals[:, m]
    A[:, fpos] = rhs
    return A
