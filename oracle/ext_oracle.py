"""The oracle for the extension rows (test infrastructure only).

Only ``tests/`` import this module, as the checker.  It is a numpy/scipy
restatement, in FP64, of the Newton system with the reference's three
features of its augmented ledger:

* distributed slack (``src/_cpp/acpf_nr_kernels.cu:793-862``: the residual
  gains ``-sa w`` on participants' P rows, ``J(P_i, sa) = w_i``,
  ``sa += dx[sa]``, ``sa0 = sum Re(Sbus)``);
* HVDC angle droop (``acpf_nr_kernels.cu:871-969``, flow formulas
  ``acpf_nr_kernels.cuh:412-473``);
* voltage control, bordered (``acpf_nr_kernels.cu:974-1068``,
  ``acpf_nr_kernels.cuh:517-532``),

in the layout the library defines (``include/gpusim2grid_amd.h``,
``g2g_extensions``): rows ``[P(pvpq) | Q(pq) | P(slack) if distributed |
per group: voltage row, sharing rows]`` and columns ``[theta(pvpq) | Vm(pq) |
sa | Q per controller]``.  The Jacobian is assembled here in closed form
(MATPOWER's dS/dVa, dS/dVm, ``tests/python/_jacobian_ref.py:26-55``, plus the
feature entries) and the linear solves are scipy's SuperLU.

Pinning (tests/test_ext_oracle.py): the Jacobian equals central finite
differences of F in every unknown, extension ones included; each feature
reduces, in a special case, to the bare system solved by the C oracle
(a distributed slack of weight 1 on the slack is the plain slack; one
controller regulating its own bus with slope 0 is a PV bus; a droop line
with k = 0 and no losses is a pair of constant injections); and converged
rows satisfy the features' own equations.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class Layout:
    """Row and column indices of every equation and unknown."""

    def __init__(self, n, pv, pq, slack, ext):
        pv = np.asarray(pv, np.int64)
        pq = np.asarray(pq, np.int64)
        self.n = n
        self.pvpq = np.sort(np.concatenate([pv, pq]))
        self.pq = pq
        npvpq, npq = self.pvpq.size, pq.size
        self.row_p = np.full(n, -1)
        self.row_p[self.pvpq] = np.arange(npvpq)
        self.col_th = self.row_p.copy()
        self.col_vm = np.full(n, -1)
        self.col_vm[pq] = npvpq + np.arange(npq)
        nxt = npvpq + npq
        self.sa = None
        self.ext = ext
        self.ds = ext is not None and ext.slack_weights is not None
        if self.ds:
            self.row_p[int(np.atleast_1d(slack)[0])] = self.sa = nxt
            nxt += 1
        self.q_col, self.v_row, self.sh_rows = [], [], []
        for g in (ext.voltage if ext is not None else []):
            cols = list(range(nxt, nxt + len(g.buses)))
            nxt += len(g.buses)
            self.q_col += cols
            self.v_row.append(cols[0])
            self.sh_rows.append(cols[1:])
        self.d = nxt
        # state vector: [sa] + [Q per controller]
        self.n_xs = int(self.ds) + len(self.q_col)


def hvdc_received(p_abs, lf_send, lf_recv, r):
    line = (1.0 - lf_send) * p_abs
    return (1.0 - lf_recv) * (line - r * line * line)


def hvdc_flows(h, th1, th2):
    """Active power leaving bus1 and bus2 into the line (acpf_nr_kernels.cuh:450-473)."""
    if h.status == 2:
        return 0.0, 0.0
    if h.status > 0:
        return h.pmax12, -hvdc_received(h.pmax12, h.lf1, h.lf2, h.r)
    if h.status < 0:
        return -hvdc_received(h.pmax21, h.lf2, h.lf1, h.r), h.pmax21
    raw = h.p0 + h.k * (th1 - th2)
    if raw >= 0:
        return raw, -hvdc_received(raw, h.lf1, h.lf2, h.r)
    return -hvdc_received(-raw, h.lf2, h.lf1, h.r), -raw


def hvdc_slopes(h, th1, th2):
    """d p1 / d theta1 and d p2 / d theta1 in the linear regime (acpf_nr_kernels.cu:940-964)."""
    raw = h.p0 + h.k * (th1 - th2)
    from1 = raw >= 0
    line = (1.0 - (h.lf1 if from1 else h.lf2)) * abs(raw)
    rs = (1.0 - h.lf1) * (1.0 - h.lf2) * (1.0 - 2.0 * h.r * line)
    return (h.k, -h.k * rs) if from1 else (h.k * rs, -h.k)


def f_ext(Ybus, V, Sbus, xs, L):
    """F (= -mismatch, the reference's sign) of the augmented system."""
    Y = sp.csr_matrix(Ybus)
    mis = V * np.conj(Y @ V) - Sbus
    F = np.zeros(L.d)
    rp = L.row_p >= 0
    F[L.row_p[rp]] = -mis.real[rp]
    F[L.col_vm[L.pq]] = -mis.imag[L.pq]
    ext = L.ext
    if ext is None:
        return F
    k = 0
    if L.ds:
        sa = xs[0]
        k = 1
        for i in np.nonzero(ext.slack_weights)[0]:
            F[L.row_p[i]] -= sa * ext.slack_weights[i]
    th = np.angle(V)
    for h in ext.hvdc:
        p1, p2 = hvdc_flows(h, th[h.bus1], th[h.bus2])
        if L.row_p[h.bus1] >= 0:
            F[L.row_p[h.bus1]] -= p1
        if L.row_p[h.bus2] >= 0:
            F[L.row_p[h.bus2]] -= p2
    c = 0
    for g, vr, shr in zip(ext.voltage, L.v_row, L.sh_rows):
        Q = xs[k + c:k + c + len(g.buses)]
        for b, q in zip(g.buses, Q):
            F[L.col_vm[b]] += q
        F[vr] = -(abs(V[g.reg_bus]) + np.dot(g.slopes, Q) - g.v_set)
        for j, r in enumerate(shr, start=1):
            F[r] = -(g.weights[0] * Q[j] - g.weights[j] * Q[0])
        c += len(g.buses)
    return F


def j_ext(Ybus, V, xs, L):
    """The augmented Jacobian d(mismatch)/dx (dense), MATPOWER form plus the features."""
    Y = sp.csr_matrix(Ybus).toarray()
    I = Y @ V
    Vn = V / np.abs(V)
    dS_dVa = 1j * np.diag(V) @ np.conj(np.diag(I) - Y @ np.diag(V))
    dS_dVm = np.diag(V) @ np.conj(Y @ np.diag(Vn)) + np.conj(np.diag(I)) @ np.diag(Vn)
    J = np.zeros((L.d, L.d))
    pr = np.nonzero(L.row_p >= 0)[0]
    th = np.nonzero(L.col_th >= 0)[0]
    vm = L.pq
    J[np.ix_(L.row_p[pr], L.col_th[th])] = dS_dVa.real[np.ix_(pr, th)]
    J[np.ix_(L.row_p[pr], L.col_vm[vm])] = dS_dVm.real[np.ix_(pr, vm)]
    J[np.ix_(L.col_vm[vm], L.col_th[th])] = dS_dVa.imag[np.ix_(vm, th)]
    J[np.ix_(L.col_vm[vm], L.col_vm[vm])] = dS_dVm.imag[np.ix_(vm, vm)]
    ext = L.ext
    if ext is None:
        return J
    if L.ds:
        for i in np.nonzero(ext.slack_weights)[0]:
            J[L.row_p[i], L.sa] += ext.slack_weights[i]
    ang = np.angle(V)
    for h in ext.hvdc:
        if h.status != 0:
            continue
        s1, s2 = hvdc_slopes(h, ang[h.bus1], ang[h.bus2])
        for bus, s in ((h.bus1, s1), (h.bus2, s2)):
            r = L.row_p[bus]
            if r < 0:
                continue
            if L.col_th[h.bus1] >= 0:
                J[r, L.col_th[h.bus1]] += s
            if L.col_th[h.bus2] >= 0:
                J[r, L.col_th[h.bus2]] -= s
    c = 0
    for g, vr, shr in zip(ext.voltage, L.v_row, L.sh_rows):
        cols = L.q_col[c:c + len(g.buses)]
        for b, qc, s in zip(g.buses, cols, g.slopes):
            J[L.col_vm[b], qc] -= 1.0
            J[vr, qc] += s
        J[vr, L.col_vm[g.reg_bus]] += 1.0
        for j, r in enumerate(shr, start=1):
            J[r, cols[j]] += g.weights[0]
            J[r, cols[0]] -= g.weights[j]
        c += len(g.buses)
    return J


def xs0(Sbus, L):
    """The state a solve starts from: sa = sum Re(Sbus), Q = 0."""
    x = np.zeros(L.n_xs)
    if L.ds:
        x[0] = np.sum(np.real(Sbus))
    return x


def nr_ext(Ybus, V0, Sbus, L, max_iter=10, tol=1e-8):
    """Plain FP64 NR of one row on the augmented system: (V, xs, status,
    iterations, residual), with the status rule of row_status.hpp:28-50
    (strict r < tol; non-finite -> 2; budget -> 1)."""
    V = np.array(V0, dtype=np.complex128)
    xs = xs0(Sbus, L)
    k0 = int(L.ds)
    it = 0
    while True:
        F = f_ext(Ybus, V, Sbus, xs, L)
        r = np.max(np.abs(F)) if F.size else 0.0
        if not np.isfinite(r):
            return V, xs, 2, it, r
        if r < tol:
            return V, xs, 0, it, r
        if it >= max_iter:
            return V, xs, 1, it, r
        J = j_ext(Ybus, V, xs, L)
        dx = spla.spsolve(sp.csc_matrix(J), F)
        th = np.angle(V)
        vm = np.abs(V)
        m = L.col_th >= 0
        th[m] += dx[L.col_th[m]]
        vm[L.pq] += dx[L.col_vm[L.pq]]
        V = vm * np.exp(1j * th)
        if L.ds:
            xs[0] += dx[L.sa]
        for c, qc in enumerate(L.q_col):
            xs[k0 + c] += dx[qc]
        it += 1
