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

REF_UNIT = 256  # a ref is slot * 256 (the byte offset of a 64-lane FP32 slot row)


def _r4(v):
    return (v + 3) // 4 * 4


def frec_src_words(c):
    return 3 + 2 * c


def frec_row_words(c):
    return c


def frec_words(c):
    return _r4(frec_src_words(c) + (c + 1) * frec_row_words(c))


def brec_words(c):
    return _r4(3 + c)


def _slot(ref):
    assert ref % REF_UNIT == 0
    return ref // REF_UNIT


def factor_forward(info, prog, A, pivot_eps=0.0, scale=None):
    """A: [rows, a_len] dense stream -> (U stream [rows, u_len], clamped [rows]).

    Slot semantics of the device kernel: a source slot is read and then
    cleared; an update target accumulates (slots start at zero)."""
    rows = A.shape[0]
    S = info["slots_f"]
    sl = np.zeros((S + 2, rows))
    U = np.zeros((rows, info["u_len"]))
    eps = pivot_eps * (np.ones(rows) if scale is None else scale)
    clamped = np.zeros(rows, np.int64)
    p = ap = up = 0
    prog = np.asarray(prog, dtype=np.int64)
    for _k in range(info["d"]):
        c = int(prog[p] & 0xFFFF)
        src = [_slot(int(v)) for v in prog[p + 1:p + 3 + 2 * c]]

        def take(q):
            v = sl[src[q]].copy()
            sl[src[q]] = 0.0
            return v
        dv = A[:, ap + 1] + take(0)
        fk = A[:, ap] + take(1)
        Lv = [A[:, ap + 2 + q] + take(2 + q) for q in range(c)]
        Uv = [A[:, ap + 2 + c + q] + take(2 + c + q) for q in range(c)]
        assert np.all(sl[S] == 0.0)
        ap += 2 + 2 * c
        small = np.abs(dv) < eps
        dv = np.where(small, np.where(dv < 0, -eps, eps), dv)
        clamped += small
        inv = 1.0 / dv
        lv = [v * inv for v in Lv]
        U[:, up] = fk
        U[:, up + 1] = inv
        for q in range(c):
            U[:, up + 2 + q] = Uv[q]
        up += c + 2
        base = p + frec_src_words(c)
        rw = frec_row_words(c)
        for a in range(c):
            for b in range(c):
                sl[_slot(int(prog[base + a * rw + b]))] += -lv[a] * Uv[b]
        base += c * rw
        for a in range(c):
            sl[_slot(int(prog[base + a]))] += -lv[a] * fk
        p += frec_words(c)
    assert p == len(prog) and ap == info["a_len"] and up == info["u_len"]
    return U, clamped


def backward(info, prog, U):
    """U stream -> x [rows, d] in input labels (U(k,k) is stored as its reciprocal)."""
    rows = U.shape[0]
    d = info["d"]
    S = info["slots_b"]
    sl = np.full((S + 2, rows), np.nan)
    x_out = np.zeros((rows, d))
    p = 0
    up = info["u_len"]
    prog = np.asarray(prog, dtype=np.int64)
    for _q in range(d):
        c = int(prog[p] & 0xFFFF)
        perm_k = int(prog[p + 1])
        r = [_slot(int(v)) for v in prog[p + 2:p + 3 + c]]
        up -= c + 2
        s = U[:, up].copy()
        for q in range(c):
            s -= U[:, up + 2 + q] * sl[r[q]]
        x = s * U[:, up + 1]
        x_out[:, perm_k] = x
        sl[r[c]] = x
        p += brec_words(c)
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
