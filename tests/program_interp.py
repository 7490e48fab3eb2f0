"""CPU interpreter of the factorization / substitution programs.

It executes, row-vectorised in numpy, exactly what the device kernels
``k_factor`` and ``k_backward`` execute per lane: the same program words, the
same A/U stream positions and the same LDS slot traffic.  Tests use it to check
the host symbolic layer on the CPU (against scipy), independently of the GPU.
"""
from __future__ import annotations

import numpy as np

NOSLOT = 0xFFFF
SLOT = 0xFFFF
FIRST = 1 << 16
INA = 1 << 17


def factor_forward(info, prog, A, pivot_eps=0.0, scale=None):
    """A: [rows, a_len] stream -> (U stream [rows, u_len], clamped [rows])."""
    rows = A.shape[0]
    d = info["d"]
    sl = np.full((info["slots_f"] + 1, rows), np.nan)
    U = np.zeros((rows, info["u_len"]))
    eps = pivot_eps * (np.ones(rows) if scale is None else scale)
    clamped = np.zeros(rows, np.int64)
    p = ap = up = 0
    for _k in range(d):
        c = prog[p]
        fk = A[:, ap].copy()
        ap += 1
        bref = prog[p + 1]
        if bref != NOSLOT:
            fk += sl[bref]
        dref = prog[p + 2]
        dv = np.zeros(rows)
        if dref & INA:
            dv = A[:, ap].copy()
            ap += 1
        if dref & SLOT != NOSLOT:
            dv += sl[dref & SLOT]
        small = np.abs(dv) < eps
        dv = np.where(small, np.where(dv < 0, -eps, eps), dv)
        clamped += small
        p += 3
        lv, uv = [], []
        for store in (lv, uv):
            for q in range(c):
                ref = prog[p + q]
                v = np.zeros(rows)
                if ref & INA:
                    v = A[:, ap].copy()
                    ap += 1
                if ref & SLOT != NOSLOT:
                    v = v + sl[ref & SLOT]
                store.append(v)
            p += c
        lv = [v / dv for v in lv]
        U[:, up] = fk
        U[:, up + 1] = dv
        for q in range(c):
            U[:, up + 2 + q] = uv[q]
        up += c + 2
        for r in range(c):
            for q in range(c):
                ref = prog[p + q]
                v = -lv[r] * uv[q]
                if ref & FIRST:
                    sl[ref & SLOT] = v
                else:
                    sl[ref & SLOT] += v
            p += c
        for r in range(c):
            ref = prog[p + r]
            v = -lv[r] * fk
            if ref & FIRST:
                sl[ref & SLOT] = v
            else:
                sl[ref & SLOT] += v
        p += c
    assert p == len(prog) and ap == info["a_len"] and up == info["u_len"]
    return U, clamped


def backward(info, prog, U, perm=None, V=None):
    """U stream -> x [rows, d] in input labels (raw mode, V is None) or, with
    V ([rows, n] complex), the fused polar bus update applied to a copy of V."""
    rows = U.shape[0]
    d = info["d"]
    sl = np.full((max(info["slots_b"], 1), rows), np.nan)
    x_out = np.zeros((rows, d))
    Vn = None if V is None else V.copy()
    p = 0
    up = info["u_len"]
    for k in range(d - 1, -1, -1):
        c = prog[p]
        up -= c + 2
        s = U[:, up].copy()
        piv = U[:, up + 1]
        for q in range(c):
            s -= U[:, up + 2 + q] * sl[prog[p + 1 + q]]
        p += 1 + c
        x = s / piv
        out, op, bus, ps = prog[p:p + 4]
        p += 4
        if Vn is not None and op != 0:
            if op == 1:
                dva, dvm = x, 0.0
            elif op == 2:
                dva, dvm = x, sl[ps]
            else:
                dva, dvm = sl[ps], x
            v = Vn[:, bus]
            m = np.abs(v)
            Vn[:, bus] = v * ((m + dvm) / m) * np.exp(1j * dva)
        if perm is not None:
            x_out[:, perm[k]] = x
        if out != NOSLOT:
            sl[out] = x
    assert p == len(prog) and up == 0
    return Vn if V is not None else x_out


def astream_from_csr(info, apos, fpos, vals, rhs):
    """Raw A stream from CSR values [rows, nnz] and right-hand sides [rows, d]."""
    vals = np.atleast_2d(vals)
    rhs = np.atleast_2d(rhs)
    A = np.zeros((vals.shape[0], info["a_len"]))
    m = apos >= 0
    A[:, apos[m]] = vals[:, m]
    A[:, fpos] = rhs
    return A
