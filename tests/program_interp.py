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


def frec2_src_words(c):
    return 7 + 4 * c


def frec2_words(c):
    return _r4(frec2_src_words(c) + (c + 1) * c)


def brec_words(c):
    return _r4(3 + c)


def brec2_words(c):
    return _r4(5 + c)


def _slot(ref):
    assert ref % REF_UNIT == 0
    return ref // REF_UNIT


MERGE_HEAD = 0x7FFE
DUMP_HEAD = 0x7FFD
GATHER_HEAD = 0x7FFC
PRELOAD_HEAD = 0x7FFB


def mrec_words(n):
    return _r4(2 + 2 * n)


def _segs(plan):
    return plan.array("segs").reshape(-1, 9)


def factor_forward(plan, A, pivot_eps=0.0, scale=None, split=False):
    """A: [rows, a_len] dense stream -> (U stream [rows, u_len], clamped [rows]).

    Slot semantics of the device kernel: a source slot is read and then
    cleared; an update target accumulates (slots start at zero).  A record
    of width 2 is a supernode of pivots k, k+1: pivot k is applied to pivot
    k+1's values in registers and struct(k+1) gets one rank-2 update.  Each
    walker runs its own program segment; walker 1 runs first (the kernel's
    walkers are independent until walker 0's first merge record), then
    walker 0 with its merges (dst += src, src = 0) and the top part."""
    info = plan.info()
    prog = np.asarray(plan.array("prog_fs" if split else "prog_f"), dtype=np.int64)
    segs = plan.array("split_segs").reshape(-1, 9) if split else _segs(plan)
    rows = A.shape[0]
    S = info["slots_fs" if split else "slots_f"]
    u_len = info["split_u_len" if split else "u_len"]
    sl = np.zeros((S + 2, rows))
    pend = np.full((info["pend_len"], rows), np.nan)
    U = np.zeros((rows, u_len))
    eps = pivot_eps * (np.ones(rows) if scale is None else scale)
    clamped = np.zeros(rows, np.int64)

    def clamp(dv):
        nonlocal clamped
        small = np.abs(dv) < eps
        clamped += small
        return np.where(small, np.where(dv < 0, -eps, eps), dv)

    def run(sg):
        p, ap, up, k = int(sg[0]), int(sg[4]), int(sg[6]), 0
        while k < sg[8]:
            h = int(prog[p] & 0xFFFF)
            if h == MERGE_HEAD:
                n = int(prog[p + 1])
                for i in range(n):
                    dst, src = _slot(int(prog[p + 2 + 2 * i])), _slot(int(prog[p + 3 + 2 * i]))
                    sl[dst] += sl[src]
                    sl[src] = 0.0
                p += mrec_words(n)
                continue
            if h == GATHER_HEAD:  # split walk: dst += pend[index]
                n = int(prog[p + 1])
                for i in range(n):
                    dst, src = _slot(int(prog[p + 2 + 2 * i])), int(prog[p + 3 + 2 * i])
                    sl[dst] += pend[src]
                p += mrec_words(n)
                continue
            c, wdt = h & 0x7FFF, 1 + (h >> 15)
            if wdt == 1:
                src = [_slot(int(v)) for v in prog[p + 1:p + 3 + 2 * c]]
                vals = []
                for q in range(2 + 2 * c):
                    vals.append(sl[src[q]].copy())
                    sl[src[q]] = 0.0
                dv = A[:, ap + 1] + vals[0]
                fk = A[:, ap] + vals[1]
                Lv = [A[:, ap + 2 + q] + vals[2 + q] for q in range(c)]
                Uv = [A[:, ap + 2 + c + q] + vals[2 + c + q] for q in range(c)]
                ap += 2 + 2 * c
                dv = clamp(dv)
                inv = 1.0 / dv
                lv = [v * inv for v in Lv]
                U[:, up] = fk
                U[:, up + 1] = inv
                for q in range(c):
                    U[:, up + 2 + q] = Uv[q]
                up += c + 2
                base = p + frec_src_words(c)
                upd = [(lv, Uv, fk)]
                p += frec_words(c)
                k += 1
            else:
                nv = 6 + 4 * c
                k1 = 4 + 2 * c
                src = [_slot(int(v)) for v in prog[p + 1:p + 1 + nv]]
                vals = []
                for q in range(nv):
                    vals.append(sl[src[q]].copy())
                    sl[src[q]] = 0.0
                sw = {0: 1, 1: 0, k1: k1 + 1, k1 + 1: k1}
                x = [A[:, ap + q] + vals[sw.get(q, q)] for q in range(nv)]
                ap += nv
                fk, d1, e, b = x[0], clamp(x[1]), x[2], x[3 + c]
                inv1 = 1.0 / d1
                l = e * inv1
                mk = [x[3 + i] * inv1 for i in range(c)]
                uk = [x[4 + c + j] for j in range(c)]
                U[:, up] = fk
                U[:, up + 1] = inv1
                U[:, up + 2] = b
                for j in range(c):
                    U[:, up + 3 + j] = uk[j]
                fk1 = x[k1] - l * fk
                d2 = clamp(x[k1 + 1] - l * b)
                inv2 = 1.0 / d2
                mk1 = [(x[k1 + 2 + i] - mk[i] * b) * inv2 for i in range(c)]
                uk1 = [x[k1 + 2 + c + j] - l * uk[j] for j in range(c)]
                u1 = up + 3 + c
                U[:, u1] = fk1
                U[:, u1 + 1] = inv2
                for j in range(c):
                    U[:, u1 + 2 + j] = uk1[j]
                up += 5 + 2 * c
                base = p + frec2_src_words(c)
                upd = [(mk, uk, fk), (mk1, uk1, fk1)]
                p += frec2_words(c)
                k += 2
            assert np.all(sl[S] == 0.0)
            rw = c
            for a in range(c):
                for b2 in range(c):
                    t = _slot(int(prog[base + a * rw + b2]))
                    for lv, uv, _ in upd:
                        sl[t] += -lv[a] * uv[b2]
            base += c * rw
            for a in range(c):
                t = _slot(int(prog[base + a]))
                for lv, _, f in upd:
                    sl[t] += -lv[a] * f
        # split walk: the set's pending contributions go to pend (dump records)
        while p < sg[0] + sg[1] and int(prog[p] & 0xFFFF) == DUMP_HEAD:
            n, first = int(prog[p + 1]), int(prog[p + 2])
            for i in range(n):
                t = _slot(int(prog[p + 3 + i]))
                pend[first + i] = sl[t]
                sl[t] = 0.0
            p += _r4(3 + n)
        assert k == sg[8] and p <= sg[0] + sg[1]
        assert ap == sg[4] + sg[5] and up == sg[6] + sg[7]

    if split:  # every program on fresh slots: fine sets, middle parts, the top part
        for w in range(len(segs)):
            sl[:] = 0.0
            run(segs[w])
            assert np.all(sl[:S] == 0.0)
        assert not np.isnan(pend).any()
        return U, clamped
    for w in range(len(segs) - 1, -1, -1):
        run(segs[w])
    assert np.all(sl[:S] == 0.0)  # every slot consumed
    return U, clamped


def backward(plan, U, split=False):
    """U stream -> x [rows, d] in input labels (U(k,k) is stored as its
    reciprocal; width-2 records solve x(k+1) then x(k)).  Walker 0 runs
    first (its segment starts with the top part), then the others."""
    info = plan.info()
    prog = np.asarray(plan.array("prog_bs" if split else "prog_b"), dtype=np.int64)
    segs = plan.array("split_bsegs" if split else "bsegs").reshape(-1, 9)
    rows = U.shape[0]
    d = info["d"]
    S = info["slots_bs" if split else "slots_b"]
    sl = np.full((S + 2, rows), np.nan)
    x_out = np.full((rows, d), np.nan) if split else np.zeros((rows, d))

    def run(sg):
        p, up, q = int(sg[2]), int(sg[6] + sg[7]), 0
        while q < sg[8]:
            h = int(prog[p] & 0xFFFF)
            if h == PRELOAD_HEAD:  # split walk: a top x the set reads
                n = int(prog[p + 1])
                for i in range(n):
                    v = x_out[:, int(prog[p + 3 + 2 * i])]
                    assert not np.isnan(v).any()
                    sl[_slot(int(prog[p + 2 + 2 * i]))] = v
                p += mrec_words(n)
                continue
            c, wdt = h & 0x7FFF, 1 + (h >> 15)
            if wdt == 1:
                perm_k = int(prog[p + 1])
                r = [_slot(int(v)) for v in prog[p + 2:p + 3 + c]]
                up -= c + 2
                s = U[:, up].copy()
                for i in range(c):
                    s -= U[:, up + 2 + i] * sl[r[i]]
                x = s * U[:, up + 1]
                x_out[:, perm_k] = x
                sl[r[c]] = x
                p += brec_words(c)
                q += 1
            else:
                pk1, pk = int(prog[p + 1]), int(prog[p + 2])
                r = [_slot(int(v)) for v in prog[p + 3:p + 5 + c]]
                up -= 5 + 2 * c
                xs = [sl[r[i]] for i in range(c)]
                s2 = U[:, up + 3 + c].copy()
                for i in range(c):
                    s2 -= U[:, up + 5 + c + i] * xs[i]
                x2 = s2 * U[:, up + 4 + c]
                s1 = U[:, up] - U[:, up + 2] * x2
                for i in range(c):
                    s1 -= U[:, up + 3 + i] * xs[i]
                x1 = s1 * U[:, up + 1]
                x_out[:, pk1] = x2
                x_out[:, pk] = x1
                sl[r[c]] = x2
                sl[r[c + 1]] = x1
                p += brec2_words(c)
                q += 2
        assert q == sg[8] and p <= sg[2] + sg[3] and up == sg[6]

    if split:  # the top part first (its x in x_out), then the middles, then the fine sets
        for w in range(len(segs) - 1, -1, -1):
            sl[:] = np.nan
            run(segs[w])
        return x_out
    for w in range(len(segs)):
        run(segs[w])
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
