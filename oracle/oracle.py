"""The oracle: CPU FP64 Newton-Raphson power flow (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this module, as the checker.  The product package never does.

Two independent restatements of the reference's bare ``[pvpq | pq]``
Newton-Raphson live here:

* ``libnr_oracle.so`` (``nr_oracle.c``): the J layout of
  ``src/_cpp/acpf_nr.cu:145-236``, mismatch/J/update of
  ``src/_cpp/acpf_nr_kernels.cu:493-788``, status rules of
  ``src/_cpp/contingency/row_status.hpp:28-50``, with its own sparse LU
  (minimum degree + partial pivoting) and POSIX threads over rows;
* numpy/scipy helpers in MATPOWER form (the reference's own test helpers
  ``tests/python/_jacobian_ref.py:26-55`` and ``tests/python/conftest.py:326-338``),
  used to pin the C oracle.

Pinning (tests/test_oracle.py): the C mismatch equals the numpy closed form;
the C Jacobian equals the MATPOWER-form Jacobian and central finite
differences of F; the C sparse solve equals scipy's SuperLU; the NR converges
quadratically and its solutions satisfy the power-flow equations.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile ``libnr_oracle.so`` with gcc (in place)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libnr_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_jpattern_size.restype = ctypes.c_int
        L.orc_jpattern_size.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int, _i32p,
                                        ctypes.c_int, _i32p, ctypes.POINTER(ctypes.c_int)]
        L.orc_jpattern.restype = None
        L.orc_jpattern.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int, _i32p, ctypes.c_int,
                                   _i32p, _i32p, _i32p, _i32p]
        L.orc_f_and_j.restype = ctypes.c_double
        L.orc_f_and_j.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_int, _i32p,
                                  ctypes.c_int, _i32p, _f64p, _f64p, _f64p, ctypes.c_void_p]
        L.orc_sparse_solve.restype = ctypes.c_int
        L.orc_sparse_solve.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_sparse_refactor_solve.restype = ctypes.c_int
        L.orc_sparse_refactor_solve.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]
        L.orc_nr_batch.restype = None
        L.orc_nr_batch.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_longlong,
                                   ctypes.c_int, _i32p, ctypes.c_int, _i32p, ctypes.c_int, _f64p,
                                   _f64p, ctypes.c_longlong, ctypes.c_int, ctypes.c_double, _i32p,
                                   _i32p, _f64p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        _LIB = L
    return _LIB


def _csr(Y):
    Y = sp.csr_matrix(Y)
    Y.sort_indices()
    return (Y.indptr.astype(np.int32), Y.indices.astype(np.int32),
            np.ascontiguousarray(Y.data.astype(np.complex128)))


def _cplx(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).view(np.float64)


def jpattern(Ybus, pv, pq):
    """Reference-layout J CSR (Jp, Ji) and the four scatter maps [4, nY]."""
    Yp, Yi, _ = _csr(Ybus)
    n = Yp.size - 1
    pv = np.ascontiguousarray(pv, dtype=np.int32)
    pq = np.ascontiguousarray(pq, dtype=np.int32)
    d = ctypes.c_int()
    nJ = lib().orc_jpattern_size(n, Yp, Yi, pv.size, pv, pq.size, pq, ctypes.byref(d))
    Jp = np.empty(d.value + 1, np.int32)
    Ji = np.empty(nJ, np.int32)
    maps = np.empty(4 * Yi.size, np.int32)
    lib().orc_jpattern(n, Yp, Yi, pv.size, pv, pq.size, pq, Jp, Ji, maps)
    return Jp, Ji, maps.reshape(4, -1)


def f_and_j(Ybus, V, Sbus, pv, pq, with_j=True):
    """(F, J csr, ||F||inf) at V, in the reference layout."""
    Yp, Yi, Yx = _csr(Ybus)
    n = Yp.size - 1
    pv = np.ascontiguousarray(pv, dtype=np.int32)
    pq = np.ascontiguousarray(pq, dtype=np.int32)
    Jp, Ji, _ = jpattern(Ybus, pv, pq)
    d = Jp.size - 1
    F = np.empty(d)
    Jx = np.empty(Ji.size)
    r = lib().orc_f_and_j(n, Yp, Yi, _cplx(Yx), pv.size, pv, pq.size, pq, _cplx(V).copy(),
                          _cplx(Sbus).copy(), F, Jx.ctypes.data if with_j else None)
    J = sp.csr_matrix((Jx, Ji, Jp), shape=(d, d)) if with_j else None
    return F, J, r


def sparse_solve(A, b):
    """x = A^-1 b with the oracle's own sparse LU."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    x = np.empty(A.shape[0])
    rc = lib().orc_sparse_solve(A.shape[0], A.indptr.astype(np.int32), A.indices.astype(np.int32),
                                np.ascontiguousarray(A.data, dtype=np.float64),
                                np.ascontiguousarray(b, dtype=np.float64), x)
    if rc != 0:
        raise ArithmeticError("oracle LU: structurally or numerically singular")
    return x


def sparse_refactor_solve(A1, A2, b):
    """Factorizes A1, refactorizes on its pivots with A2's values (same
    pattern), returns A2^-1 b; None when the refactorization meets a weak
    pivot (the NR then factorizes afresh)."""
    A1 = sp.csr_matrix(A1)
    A1.sort_indices()
    A2 = sp.csr_matrix(A2)
    A2.sort_indices()
    if not (np.array_equal(A1.indptr, A2.indptr) and np.array_equal(A1.indices, A2.indices)):
        raise ValueError("A1 and A2 must share the pattern")
    x = np.empty(A1.shape[0])
    rc = lib().orc_sparse_refactor_solve(A1.shape[0], A1.indptr.astype(np.int32), A1.indices.astype(np.int32),
                                         np.ascontiguousarray(A1.data, dtype=np.float64),
                                         np.ascontiguousarray(A2.data, dtype=np.float64),
                                         np.ascontiguousarray(b, dtype=np.float64), x)
    return x if rc == 0 else None


def nr_batch(Ybus, V0, Sbus, pv, pq, max_iter=10, tol=1e-8, n_threads=1, Yx_rows=None, max_dva=0.0,
             max_dvm=0.0, refactor=False):
    """Plain FP64 NR over rows (step scaling when max_dva > 0; refactor: a
    row's later iterations refactorize on its first factorization's pivots).

    V0 : (n,) shared start or (rows, n); Sbus : (n,) shared or (rows, n);
    Yx_rows : optional (rows, nY) per-row Ybus values on Ybus's pattern.
    Returns (V [rows, n], status [rows], iters [rows], residual [rows]).
    Status: 0 Converged, 1 MaxIter, 2 Diverged.
    """
    Yp, Yi, Yx = _csr(Ybus)
    n = Yp.size - 1
    S = np.atleast_2d(np.asarray(Sbus, dtype=np.complex128))
    rows = max(S.shape[0], np.atleast_2d(V0).shape[0],
               1 if Yx_rows is None else np.atleast_2d(Yx_rows).shape[0])
    V = np.array(np.broadcast_to(np.atleast_2d(np.asarray(V0, dtype=np.complex128)), (rows, n)))
    s_stride = n if S.shape[0] > 1 else 0
    if Yx_rows is None:
        Yall, y_stride = Yx, 0
    else:
        Yall = np.ascontiguousarray(np.atleast_2d(Yx_rows), dtype=np.complex128)
        y_stride = Yall.shape[1]
    pv = np.ascontiguousarray(pv, dtype=np.int32)
    pq = np.ascontiguousarray(pq, dtype=np.int32)
    status = np.empty(rows, np.int32)
    iters = np.empty(rows, np.int32)
    res = np.empty(rows)
    Vf = np.ascontiguousarray(V).view(np.float64)
    lib().orc_nr_batch(n, Yp, Yi, _cplx(Yall), y_stride, pv.size, pv, pq.size, pq, rows, Vf,
                       _cplx(np.ascontiguousarray(S)), s_stride, int(max_iter), float(tol), status,
                       iters, res, int(n_threads), float(max_dva), float(max_dvm), int(bool(refactor)))
    return Vf.view(np.complex128).reshape(rows, n), status, iters, res


def nr_masked(Ybus, V0, Sbus, pv, pq, masks, max_iter=10, tol=1e-8, Yx_rows=None, n_threads=1):
    """NR of rows whose ``masks[r]`` buses are islanded (handle_disconnected_grid,
    ``src/_cpp/contingency_analysis_helper.cpp:657-692`` and
    ``src/_cpp/acpf_nr_kernels.cu:219-307``): a masked bus's equations are
    identity rows with zero mismatch, so its voltage never moves, it is left
    out of the residual, and the row's other buses solve as the grid restricted
    to them (the trips that island it leave no coupling).  Each row is solved
    as that restricted grid with ``nr_batch``; masked voltages come back NaN.
    Rows with an empty mask solve on the full grid.  Same returns as nr_batch.
    """
    Y = sp.csr_matrix(Ybus)
    Y.sort_indices()
    n = Y.shape[0]
    S = np.atleast_2d(np.asarray(Sbus, dtype=np.complex128))
    rows = len(masks)
    V0 = np.broadcast_to(np.atleast_2d(np.asarray(V0, dtype=np.complex128)), (rows, n))
    S = np.broadcast_to(S, (rows, n))
    Vout = np.empty((rows, n), np.complex128)
    st = np.empty(rows, np.int32)
    it = np.empty(rows, np.int32)
    res = np.empty(rows)
    for r in range(rows):
        Yr = Y if Yx_rows is None else sp.csr_matrix((np.asarray(Yx_rows[r]), Y.indices, Y.indptr), shape=Y.shape)
        live = np.setdiff1d(np.arange(n), np.asarray(masks[r], dtype=np.int64))
        new = np.full(n, -1, np.int64)
        new[live] = np.arange(live.size)
        sub = sp.csr_matrix(Yr[live][:, live])
        pvs, pqs = new[np.asarray(pv)], new[np.asarray(pq)]
        v, s_, i_, q_ = nr_batch(sub, V0[r, live], S[r, live], pvs[pvs >= 0], pqs[pqs >= 0], max_iter, tol,
                                 n_threads=n_threads)
        Vout[r] = np.nan
        Vout[r, live] = v[0]
        st[r], it[r], res[r] = s_[0], i_[0], q_[0]
    return Vout, st, it, res


def limit_violations(V, res, status, table, trips=None, tol=1e-6, rel_tol=1e-9, K=16):
    """Limit checks of solved rows, record by record as the reference's
    check_limit_violations_kernel defines them
    (``src/_cpp/contingency/violation_kernels.cuh:80-160``,
    ``violation_kernels.cu:196-276``) and get_violations orders them
    (``src/gpusim2grid/contingency_analysis/__init__.py:616-673``).

    V : (rows, n) complex; res, status : (rows,); table : dict as
    ``_capi.Batch.limit_violations`` takes; trips : per row, tripped branches.
    Returns per row a dict: count, truncated, n_low, n_high, n_current and
    ``records``, a list of (element_type, element_id, side, type, value, limit).
    """
    f = np.asarray(table["from"])
    t = np.asarray(table["to"])
    m = f.size
    nl = int(table["n_lines"])
    base_or, base_ex = np.asarray(table["base_or"], float), np.asarray(table["base_ex"], float)
    l1, l2 = np.asarray(table["limit1_ka"], float), np.asarray(table["limit2_ka"], float)
    vn, vmin, vmax = (np.asarray(table[k], float) for k in ("vn_kv", "vmin_kv", "vmax_kv"))
    Y = [np.asarray(table[k], np.complex128) for k in ("yff", "yft", "ytf", "ytt")]
    up, dn = 1.0 + rel_tol, 1.0 - rel_tol
    out = []
    for r in range(V.shape[0]):
        if status[r] == 3:
            out.append(dict(count=-1, truncated=0, n_low=-1, n_high=-1, n_current=-1, records=[]))
            continue
        if np.isnan(res[r]) or res[r] > tol:
            out.append(dict(count=1, truncated=0, n_low=0, n_high=0, n_current=0,
                            records=[(3, -1, 0, 4, float(res[r]), float(tol))]))
            continue
        v = V[r]
        vf = np.where(f >= 0, v[np.maximum(f, 0)], 0)
        vt = np.where(t >= 0, v[np.maximum(t, 0)], 0)
        # I = y1 va + y2 vb, in the (re, im) operation order of a complex product
        cur = lambda y1, va, y2, vb: ((y1.real * va.real - y1.imag * va.imag)  # noqa: E731
                                      + (y2.real * vb.real - y2.imag * vb.imag),
                                      (y1.real * va.imag + y1.imag * va.real)
                                      + (y2.real * vb.imag + y2.imag * vb.real))
        with np.errstate(invalid="ignore", divide="ignore"):
            ior, ioi = cur(Y[0], vf, Y[1], vt)
            ier, iei = cur(Y[2], vf, Y[3], vt)
            ior, ioi = np.where(f >= 0, ior, 0.0), np.where(f >= 0, ioi, 0.0)
            ier, iei = np.where(t >= 0, ier, 0.0), np.where(t >= 0, iei, 0.0)
            ko = np.sqrt(ior * ior + ioi * ioi) * base_or * 0.001
            ke = np.sqrt(ier * ier + iei * iei) * base_ex * 0.001
            vm = np.sqrt(v.real * v.real + v.imag * v.imag) * vn
            tripped = np.zeros(m, bool)
            if trips is not None:
                tripped[list(trips[r])] = True
            groups = {0: [], 1: [], 2: []}  # CURRENT, LOW_VOLTAGE, HIGH_VOLTAGE
            for side, k, lim in ((1, ko, l1), (2, ke, l2)):
                hit = ~tripped & ~np.isnan(lim) & (k > lim * up)
                for l in np.flatnonzero(hit):
                    et, eid = (1, l) if l < nl else (2, l - nl)
                    groups[0].append((abs(k[l] / lim[l] - 1.0), 2 * l + side - 1,
                                      (et, int(eid), side, 2, float(k[l]), float(lim[l]))))
            low = ~np.isnan(vmin) & (vm < vmin * dn)
            high = ~low & ~np.isnan(vmax) & (vm > vmax * up)
            for g, hit, lim, code in ((1, low, vmin, 0), (2, high, vmax, 1)):
                for b in np.flatnonzero(hit):
                    groups[g].append((abs(vm[b] / lim[b] - 1.0), 2 * m + b,
                                      (0, int(b), 0, code, float(vm[b]), float(lim[b]))))
        recs = []
        for g in (0, 1, 2):
            kept = sorted(groups[g], key=lambda x: (-x[0], x[1]))[:K]
            recs.extend(x[2] for x in kept)
        out.append(dict(count=len(recs), truncated=int(any(len(groups[g]) > K for g in groups)),
                        n_low=len(groups[1]), n_high=len(groups[2]), n_current=len(groups[0]), records=recs))
    return out


# ---------------------------------------------------------------------------
# numpy / scipy restatements (MATPOWER form) used to pin the C oracle
# ---------------------------------------------------------------------------

def mismatch_np(Ybus, V, Sbus, pv, pq):
    """F = -[Re(mis)(pvpq sorted) ; Im(mis)(pq)], mis = V conj(Y V) - Sbus."""
    pvpq = np.sort(np.concatenate([pv, pq]))
    mis = V * np.conj(Ybus @ V) - Sbus
    return -np.concatenate([mis[pvpq].real, mis[np.asarray(pq)].imag])


def jacobian_np(Ybus, V, pv, pq):
    """MATPOWER dS/dVa, dS/dVm restricted to [pvpq | pq] (dense)."""
    Ybus = sp.csr_matrix(Ybus)
    pvpq = np.sort(np.concatenate([pv, pq]))
    pq = np.asarray(pq)
    I = Ybus @ V
    dV = sp.diags(V)
    dSa = (1j * dV @ (sp.diags(I) - Ybus @ dV).conj()).toarray()
    dSm = (dV @ (Ybus @ sp.diags(V / np.abs(V))).conj() + sp.diags(np.conj(I) * V / np.abs(V))).toarray()
    return np.block([[dSa[np.ix_(pvpq, pvpq)].real, dSm[np.ix_(pvpq, pq)].real],
                     [dSa[np.ix_(pq, pvpq)].imag, dSm[np.ix_(pq, pq)].imag]])


def polar_update_np(V, dx, pv, pq):
    """theta += dx (pvpq), then |V| += dx (pq): the reference's polar update."""
    pvpq = np.sort(np.concatenate([pv, pq]))
    Va = np.angle(V)
    Vm = np.abs(V)
    Va[pvpq] += dx[:pvpq.size]
    Vm[np.asarray(pq)] += dx[pvpq.size:]
    return Vm * np.exp(1j * Va)


def outage_ybus(grid, branch_ids):
    """Ybus rebuilt from the branch list without ``branch_ids`` (closed form)."""
    from gpusim2grid_amd.synthetic import build_ybus
    keep = np.ones(grid.n_branch, dtype=bool)
    keep[list(branch_ids)] = False
    Y = build_ybus(grid.n_bus, grid.branch_from[keep], grid.branch_to[keep], grid.yff[keep],
                   grid.yft[keep], grid.ytf[keep], grid.ytt[keep])
    return Y


def _pattern_keys(Y):
    """Sorted CSR Y -> int64 keys row * ncols + col of its entries."""
    rows = np.repeat(np.arange(Y.shape[0], dtype=np.int64), np.diff(Y.indptr))
    return rows * Y.shape[1] + Y.indices.astype(np.int64)


def ybus_values_on_pattern(Ybase, Y):
    """Values of Y on the (superset) pattern of Ybase, zeros where Y has none."""
    Ybase = sp.csr_matrix(Ybase)
    Ybase.sort_indices()
    Yd = sp.csr_matrix(Y)
    Yd.sum_duplicates()
    Yd.sort_indices()
    kb, ky = _pattern_keys(Ybase), _pattern_keys(Yd)
    out = np.zeros(Ybase.nnz, np.complex128)
    if ky.size:
        pos = np.minimum(np.searchsorted(ky, kb), ky.size - 1)
        hit = ky[pos] == kb
        out[hit] = Yd.data[pos[hit]]
    return out


def outage_ybus_rows(grid, outages):
    """Per-row Ybus values on grid.Ybus's pattern with each row's branches
    tripped: the base values minus each tripped branch's four terms (yff at
    (f, f), yft at (f, t), ytf at (t, f), ytt at (t, t)), the closed form
    outage_ybus rebuilds from the branch list.  [rows, nY] complex."""
    Yb = sp.csr_matrix(grid.Ybus)
    Yb.sort_indices()
    kb = _pattern_keys(Yb)
    n = Yb.shape[1]
    f = np.asarray(grid.branch_from, np.int64)
    t = np.asarray(grid.branch_to, np.int64)
    at = [np.searchsorted(kb, a * n + b) for a, b in ((f, f), (f, t), (t, f), (t, t))]
    val = [grid.yff, grid.yft, grid.ytf, grid.ytt]
    Y = np.tile(Yb.data.astype(np.complex128), (len(outages), 1))
    for r, out in enumerate(outages):
        for l in out:
            for a, v in zip(at, val):
                Y[r, a[l]] -= v[l]
    return Y
