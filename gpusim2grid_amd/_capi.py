"""ctypes binding of the C ABI (``include/gpusim2grid_amd.h``).

This is the binding a Python caller without the pybind11 module would add;
the parity tests drive the kernels through it.  Errors raise ``ValueError``
(invalid argument) or ``RuntimeError`` (runtime / HIP failure), as the
reference maps ``std::invalid_argument`` / ``std::runtime_error``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpusim2grid_amd.so")

G2G_OK, G2G_ERR_INVALID, G2G_ERR_RUNTIME, G2G_ERR_HIP = 0, 1, 2, 3
FP64, MIXED = 0, 1
CHUNKED, CONTINUOUS = 0, 1
CONVERGED, MAXITER, DIVERGED, NOTSIMULATED = 0, 1, 2, 3
ARR = dict(jp=0, ji=1, maps=2, pvpq=3, perm=4, colp=5, coli=6, prog_f=7, prog_b=8, apos=9, fpos=10,
           uoff=11, segs=12, bsegs=13, prog_fs=14, prog_bs=15, split_segs=16, split_bsegs=17,
           split_apos=18, split_fpos=19)


class PlanInfo(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in (
        "n", "nY", "d", "nJ", "npvpq", "npq", "slack", "nnz", "nLU", "n_updates", "max_c", "height",
        "slots_f", "slots_b", "a_len", "u_len", "n_steps", "walkers", "n_top", "bwalkers",
        "u_stride_f64", "u_stride_f32", "n_ext", "n_xs", "sa_col", "n_split", "slots_fs", "slots_bs",
        "pend_len", "split_a_len", "split_u_len", "a_stride_f64", "a_stride_f32")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LimitTable(ctypes.Structure):
    """g2g_limit_table"""
    _fields_ = [("m", ctypes.c_int32), ("from_", ctypes.c_void_p), ("to", ctypes.c_void_p),
                ("y4", ctypes.c_void_p), ("base_or", ctypes.c_void_p), ("base_ex", ctypes.c_void_p),
                ("limit1_ka", ctypes.c_void_p), ("limit2_ka", ctypes.c_void_p), ("n_lines", ctypes.c_int32),
                ("vn_kv", ctypes.c_void_p), ("vmin_kv", ctypes.c_void_p), ("vmax_kv", ctypes.c_void_p)]


class Violations(ctypes.Structure):
    """g2g_violations"""
    _fields_ = [(k, ctypes.c_void_p) for k in ("count", "truncated", "n_low", "n_high", "n_current",
                                                "element_type", "element_id", "side", "type", "value",
                                                "limit")]


MAX_VIOLATION_CAPACITY = 1024


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("n_converged", ctypes.c_int32),
                ("n_maxiter", ctypes.c_int32), ("n_diverged", ctypes.c_int32),
                ("ms_total", ctypes.c_double), ("ms_mismatch", ctypes.c_double),
                ("ms_factor", ctypes.c_double), ("ms_backward", ctypes.c_double),
                ("bytes_factor", ctypes.c_int64), ("rows_factored", ctypes.c_int64),
                ("n_factor_launches", ctypes.c_int32), ("n_requeued", ctypes.c_int32),
                ("ms_warm_start", ctypes.c_double), ("ms_factor_main", ctypes.c_double),
                ("rows_factored_main", ctypes.c_int64), ("n_factor_main", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIB = None
_P = ctypes.c_void_p


def lib():
    """Loads ``libgpusim2grid_amd.so`` (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m gpusim2grid_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "g2g_last_error": (ctypes.c_char_p, []),
        "g2g_version": (ctypes.c_int, []),
        "g2g_set_device": (ctypes.c_int, [i32]),
        "g2g_plan_create": (ctypes.c_int, [i32, _P, _P, _P, i32, _P, i32, _P, i32, ctypes.POINTER(_P)]),
        "g2g_plan_create_ext": (ctypes.c_int, [i32, _P, _P, _P, i32, _P, i32, _P, i32, _P, ctypes.POINTER(_P)]),
        "g2g_batch_get_ext_state": (ctypes.c_int, [_P, _P]),
        "g2g_batch_set_ext_state": (ctypes.c_int, [_P, _P]),
        "g2g_lu_plan_create": (ctypes.c_int, [i32, _P, _P, _P, ctypes.POINTER(_P)]),
        "g2g_plan_destroy": (None, [_P]),
        "g2g_plan_get_info": (ctypes.c_int, [_P, ctypes.POINTER(PlanInfo)]),
        "g2g_plan_get_array": (ctypes.c_int, [_P, i32, _P, i64, ctypes.POINTER(i64)]),
        "g2g_batch_create": (ctypes.c_int, [_P, i32, i32, i32, _P, ctypes.POINTER(_P)]),
        "g2g_batch_destroy": (None, [_P]),
        "g2g_batch_set_v": (ctypes.c_int, [_P, _P, i64, i32, i32, i32]),
        "g2g_batch_set_sbus": (ctypes.c_int, [_P, _P, i64, i32, i32, i32]),
        "g2g_batch_set_ybus_edits": (ctypes.c_int, [_P, _P, _P, _P, i32]),
        "g2g_batch_set_bus_mask": (ctypes.c_int, [_P, _P, _P, i32]),
        "g2g_batch_set_active_rows": (ctypes.c_int, [_P, i32]),
        "g2g_batch_solve": (ctypes.c_int, [_P, i32, dbl, i32, ctypes.POINTER(SolveStats)]),
        "g2g_batch_get_v": (ctypes.c_int, [_P, _P, i64, i32]),
        "g2g_batch_get_results": (ctypes.c_int, [_P, _P, _P, _P]),
        "g2g_batch_eval": (ctypes.c_int, [_P, _P, _P, _P]),
        "g2g_lu_solve_batch": (ctypes.c_int, [_P, i32, i32, _P, _P, _P, dbl, i32, _P, _P]),
        "g2g_lu_solve_batch_noalloc": (ctypes.c_int, [_P, i32, i32, _P, _P, _P, dbl, i32, _P, _P]),
        "g2g_batch_set_step_control": (ctypes.c_int, [_P, i32, i32, dbl, dbl]),
        "g2g_batch_stream_begin": (ctypes.c_int, [_P]),
        "g2g_batch_solve_adjoint": (ctypes.c_int, [_P, _P, _P]),
        "g2g_batch_stream_fill": (ctypes.c_int, [_P, i32, _P, _P, i64, _P, i64, i32, _P, _P, _P]),
        "g2g_batch_stream_round": (ctypes.c_int, [_P, i32, i32, dbl]),
        "g2g_batch_stream_take": (ctypes.c_int, [_P, i32, _P, _P, i64, i32]),
        "g2g_batch_stream_warm": (ctypes.c_int, [_P, _P, i32, _P, _P, i32]),
        "g2g_batch_device_arrays": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
        "g2g_batch_warm_start": (ctypes.c_int, [_P, _P, _P, i32]),
        "g2g_batch_set_warm_contraction": (ctypes.c_int, [_P, dbl]),
        "g2g_batch_set_warm_mask": (ctypes.c_int, [_P, _P, i32]),
        "g2g_batch_set_stream_graph": (ctypes.c_int, [_P, i32]),
        "g2g_batch_set_refinement_residual": (ctypes.c_int, [_P, i32]),
        "g2g_batch_stream_graph_info": (ctypes.c_int, [_P, ctypes.POINTER(i32), ctypes.POINTER(i64)]),
        "g2g_batch_branch_flows": (ctypes.c_int, [_P, i32, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
        "g2g_batch_limit_violations": (ctypes.c_int, [_P, ctypes.POINTER(LimitTable), _P, _P, ctypes.c_double,
                                                      ctypes.c_double, i32, ctypes.POINTER(Violations)]),
        "g2g_batch_solve_warm": (ctypes.c_int, [_P, _P, _P, i32, i32, dbl, i32, dbl, ctypes.POINTER(SolveStats)]),
        "g2g_mlp_create": (ctypes.c_int, [i32, i32, i32, ctypes.c_uint64, _P, ctypes.POINTER(_P)]),
        "g2g_mlp_destroy": (None, [_P]),
        "g2g_mlp_num_params": (i64, [_P]),
        "g2g_mlp_param_offsets": (ctypes.c_int, [_P, _P]),
        "g2g_mlp_get_params": (ctypes.c_int, [_P, _P, i64]),
        "g2g_mlp_set_params": (ctypes.c_int, [_P, _P, i64]),
        "g2g_mlp_set_normalization": (ctypes.c_int, [_P, _P, _P, _P, _P]),
        "g2g_mlp_set_buses": (ctypes.c_int, [_P, _P, _P, i32]),
        "g2g_mlp_forward": (ctypes.c_int, [_P, _P, i32, _P, i32]),
        "g2g_mlp_loss_grad": (ctypes.c_int, [_P, _P, _P, i32, _P, _P, i32]),
        "g2g_mlp_train_step": (ctypes.c_int, [_P, _P, _P, i32, ctypes.c_float, _P, i32]),
        "g2g_mlp_set_grad_clip": (ctypes.c_int, [_P, ctypes.c_float]),
        "g2g_gemm_bf16": (ctypes.c_int, [i32, i32, i32, i32, i32, _P, i64, _P, i64, _P, i64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def exported_symbols():
    """Names of the g2g_* entry points the header declares."""
    hdr = os.path.join(os.path.dirname(HERE), "include", "gpusim2grid_amd.h")
    import re
    with open(hdr) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(g2g_[a-z0-9_]+)\s*\(", text)))


def _check(rc):
    if rc != G2G_OK:
        msg = lib().g2g_last_error().decode()
        if rc == G2G_ERR_INVALID:
            raise ValueError(msg)
        raise RuntimeError(msg)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class Plan:
    """Row-independent tables of one grid (or one raw matrix pattern)."""

    def __init__(self, handle):
        self._h = handle
        self._destroy = lib().g2g_plan_destroy

    @classmethod
    def from_grid(cls, Ybus, pv, pq, slack, ext=None):
        """ext: an ``extensions.Extensions`` (extension rows) or None."""
        import scipy.sparse as sp
        Y = sp.csr_matrix(Ybus)
        Y.sort_indices()
        Yp, Yi, Yx = _i32(Y.indptr), _i32(Y.indices), _c128(Y.data)
        pv, pq = _i32(pv), _i32(pq)
        h = _P()
        s0 = int(np.atleast_1d(slack)[0])
        if ext is None or ext.empty():
            _check(lib().g2g_plan_create(Y.shape[0], _ptr(Yp), _ptr(Yi), _ptr(Yx), pv.size, _ptr(pv), pq.size,
                                         _ptr(pq), s0, ctypes.byref(h)))
        else:
            from .extensions import to_c
            c, keep = to_c(ext, Y.shape[0])
            _check(lib().g2g_plan_create_ext(Y.shape[0], _ptr(Yp), _ptr(Yi), _ptr(Yx), pv.size, _ptr(pv), pq.size,
                                             _ptr(pq), s0, ctypes.byref(c), ctypes.byref(h)))
            del keep
        return cls(h)

    @classmethod
    def from_matrix(cls, A, order=None):
        import scipy.sparse as sp
        A = sp.csr_matrix(A)
        A.sort_indices()
        Ap, Ai = _i32(A.indptr), _i32(A.indices)
        o = None if order is None else _i32(order)
        h = _P()
        _check(lib().g2g_lu_plan_create(A.shape[0], _ptr(Ap), _ptr(Ai), _ptr(o), ctypes.byref(h)))
        return cls(h)

    def __del__(self):
        # the destroy function is looked up while the module is alive (at
        # interpreter exit its globals may already be gone)
        h, self._h = getattr(self, "_h", None), None
        destroy = getattr(self, "_destroy", None)
        if h and destroy:
            destroy(h)

    def info(self):
        inf = PlanInfo()
        _check(lib().g2g_plan_get_info(self._h, ctypes.byref(inf)))
        return inf.as_dict()

    def array(self, name):
        n = ctypes.c_int64()
        _check(lib().g2g_plan_get_array(self._h, ARR[name], None, 0, ctypes.byref(n)))
        out = np.empty(n.value, np.int32)
        _check(lib().g2g_plan_get_array(self._h, ARR[name], _ptr(out), out.size, ctypes.byref(n)))
        return out

    def lu_solve(self, vals, rhs, precision=FP64, pivot_eps=None, ir_steps=0, stream=None):
        vals = np.ascontiguousarray(np.atleast_2d(vals), dtype=np.float64)
        rhs = np.ascontiguousarray(np.atleast_2d(rhs), dtype=np.float64)
        rows = vals.shape[0]
        x = np.empty_like(rhs)
        cl = np.empty(rows, np.int32)
        eps = (1e-5 if precision == MIXED else 1e-12) if pivot_eps is None else pivot_eps
        _check(lib().g2g_lu_solve_batch(self._h, rows, precision, _ptr(vals), _ptr(rhs), _ptr(x), eps,
                                        int(ir_steps), _ptr(cl), stream))
        return x, cl


def lu_solve_noalloc(plan, rows, precision, d_vals, d_rhs, d_x, pivot_eps=None, ir_steps=0, d_clamped=None,
                     stream=None):
    """g2g_lu_solve_batch_noalloc on device pointers (ints): nothing is allocated."""
    eps = (1e-5 if precision == MIXED else 1e-12) if pivot_eps is None else pivot_eps
    _check(lib().g2g_lu_solve_batch_noalloc(plan._h, int(rows), precision, _P(d_vals), _P(d_rhs), _P(d_x), eps,
                                            int(ir_steps), None if d_clamped is None else _P(d_clamped), stream))


class Batch:
    """Device state of many power flows of one plan."""

    def __init__(self, plan: Plan, rows: int, precision=FP64, per_row_ybus=False, stream=None):
        self.plan = plan
        self.rows = int(rows)
        self.n = plan.info()["n"]
        h = _P()
        _check(lib().g2g_batch_create(plan._h, self.rows, precision, int(bool(per_row_ybus)), stream,
                                      ctypes.byref(h)))
        self._h = h
        self._destroy = lib().g2g_batch_destroy

    def __del__(self):
        # the destroy function is looked up while the module is alive (at
        # interpreter exit its globals may already be gone)
        h, self._h = getattr(self, "_h", None), None
        destroy = getattr(self, "_destroy", None)
        if h and destroy:
            destroy(h)

    def set_v(self, V, row0=0):
        V = _c128(V)
        rows = 1 if V.ndim == 1 else V.shape[0]
        ld = 0 if V.ndim == 1 else V.shape[1]
        if V.ndim == 1:
            rows = self.rows - row0
        _check(lib().g2g_batch_set_v(self._h, _ptr(V), ld, rows, row0, 0))

    def set_sbus(self, S, row0=0):
        S = _c128(S)
        ld = 0 if S.ndim == 1 else S.shape[1]
        rows = self.rows - row0 if S.ndim == 1 else S.shape[0]
        _check(lib().g2g_batch_set_sbus(self._h, _ptr(S), ld, rows, row0, 0))

    def set_v_device(self, ptr, ld, rows, row0=0):
        _check(lib().g2g_batch_set_v(self._h, _P(ptr), ld, rows, row0, 1))

    def set_sbus_device(self, ptr, ld, rows, row0=0):
        _check(lib().g2g_batch_set_sbus(self._h, _P(ptr), ld, rows, row0, 1))

    def set_ybus_edits(self, row_ptr, entry, delta):
        row_ptr, entry, delta = _i32(row_ptr), _i32(entry), _c128(delta)
        _check(lib().g2g_batch_set_ybus_edits(self._h, _ptr(row_ptr), _ptr(entry), _ptr(delta),
                                              row_ptr.size - 1))

    def set_bus_mask(self, row_ptr, buses=None):
        """Islanded buses per row (CSR row_ptr / buses); row_ptr None clears the mask."""
        if row_ptr is None:
            _check(lib().g2g_batch_set_bus_mask(self._h, None, None, 0))
            return
        row_ptr, buses = _i32(row_ptr), _i32(buses if buses is not None else [])
        _check(lib().g2g_batch_set_bus_mask(self._h, _ptr(row_ptr), _ptr(buses), row_ptr.size - 1))

    def set_active_rows(self, k):
        _check(lib().g2g_batch_set_active_rows(self._h, int(k)))

    def solve_adjoint(self, rhs):
        """x with J^T x = rhs per active row (J at the batch's V); rhs [rows, d]."""
        rhs = np.ascontiguousarray(np.atleast_2d(rhs), dtype=np.float64)
        x = np.empty_like(rhs)
        _check(lib().g2g_batch_solve_adjoint(self._h, _ptr(rhs), _ptr(x)))
        return x

    # -- streamed rows (continuous batching) --------------------------------
    def stream_begin(self):
        """Every slot free; the batch's iteration clock at 0."""
        _check(lib().g2g_batch_stream_begin(self._h))

    def stream_fill(self, slots, S, V0, edits=None):
        """Rows S [k, n] (complex host arrays; V0 [k, n] or one shared row) into
        the given free slots, each with its outage edits (row_ptr, entry, delta)."""
        slots = _i32(slots)
        k = slots.size
        S = np.ascontiguousarray(np.atleast_2d(S), dtype=np.complex128)
        V0 = np.ascontiguousarray(np.atleast_2d(V0), dtype=np.complex128)
        ldv = 0 if V0.shape[0] == 1 and k > 1 else V0.shape[1]
        ep = en = ed = None
        if edits is not None:
            ep, en, ed = _i32(edits[0]), _i32(edits[1]), _c128(edits[2])
        _check(lib().g2g_batch_stream_fill(self._h, k, _ptr(slots), _ptr(S), S.shape[1], _ptr(V0), ldv, 0,
                                           _ptr(ep), _ptr(en), _ptr(ed)))

    def stream_warm(self, mlp, slots, mask=None):
        """Start voltages of the given (just filled) slots from the MLP, on the device."""
        slots = _i32(slots)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        _check(lib().g2g_batch_stream_warm(self._h, mlp._h, slots.size, _ptr(slots), _ptr(m),
                                           0 if m is None else m.shape[1]))

    def stream_round(self, k=1, max_iter=10, tol=1e-8):
        _check(lib().g2g_batch_stream_round(self._h, int(k), int(max_iter), float(tol)))

    def set_stream_graph(self, on):
        """Streamed rounds captured as a HIP graph and replayed (default on)."""
        _check(lib().g2g_batch_set_stream_graph(self._h, int(bool(on))))

    def stream_graph_info(self):
        on, rep = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().g2g_batch_stream_graph_info(self._h, ctypes.byref(on), ctypes.byref(rep)))
        return bool(on.value), int(rep.value)

    def stream_take(self, slots, out=None):
        """V of the given slots (host [k, n], or into `out`, a C-contiguous
        complex128 [>= k, n] array such as a pinned buffer); the slots become free."""
        slots = _i32(slots)
        if out is None:
            V = np.empty((slots.size, self.n), np.complex128)
        else:
            if out.dtype != np.complex128 or not out.flags.c_contiguous or out.shape[1] != self.n \
                    or out.shape[0] < slots.size:
                raise ValueError("out must be a C-contiguous complex128 [>= k, n] array")
            V = out[:slots.size]
        _check(lib().g2g_batch_stream_take(self._h, slots.size, _ptr(slots), _ptr(V), self.n, 0))
        return V

    def set_step_control(self, ir_steps=0, scaling=False, max_dva=0.5, max_dvm=0.1):
        """Refinement rounds per Newton step and per-row step scaling of the next solves."""
        _check(lib().g2g_batch_set_step_control(self._h, int(ir_steps), 1 if scaling else 0, float(max_dva),
                                                float(max_dvm)))

    def solve(self, max_iter=10, tol=1e-8, scheduling=CONTINUOUS):
        st = SolveStats()
        _check(lib().g2g_batch_solve(self._h, int(max_iter), float(tol), int(scheduling), ctypes.byref(st)))
        return st.as_dict()

    def set_refinement_residual(self, fp64):
        """Refinement rounds' residual: FP64 from V (True) or the solve's precision."""
        _check(lib().g2g_batch_set_refinement_residual(self._h, int(bool(fp64))))

    def set_warm_mask(self, mask):
        """Resident outage mask for warm starts (rows x n_mask, 0/1), used
        when solve_warm / warm_start get no mask; None drops it."""
        if mask is None:
            _check(lib().g2g_batch_set_warm_mask(self._h, None, 0))
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        if m.ndim != 2 or m.shape[0] != self.rows:
            raise ValueError("the mask needs one row per batch row")
        _check(lib().g2g_batch_set_warm_mask(self._h, _ptr(m), m.shape[1]))

    def set_warm_contraction(self, q):
        _check(lib().g2g_batch_set_warm_contraction(self._h, float(q)))

    def solve_warm(self, mlp, k_warm=3, growth=2.0, max_iter=10, tol=1e-8, mask=None):
        """Warm-started solve with flat-start fallback (g2g_batch_solve_warm);
        the batch's V (set_v) is the fallback start."""
        st = SolveStats()
        n_mask = 0
        m = None
        if mask is not None:
            m = np.ascontiguousarray(mask, dtype=np.uint8)
            n_mask = m.shape[1]
        _check(lib().g2g_batch_solve_warm(self._h, mlp._h, _ptr(m), n_mask, int(k_warm), float(growth),
                                          int(max_iter), float(tol), ctypes.byref(st)))
        return st.as_dict()

    def branch_flows(self, from_bus, to_bus, yff, yft, ytf, ytt, base_or, base_ex, trips=None):
        """(or_amps, ex_amps), active rows x branches, at the batch's V."""
        f, t = _i32(from_bus), _i32(to_bus)
        m = f.size
        y4 = np.ascontiguousarray(np.stack([yff, yft, ytf, ytt], axis=1), dtype=np.complex128)
        bo = np.ascontiguousarray(base_or, dtype=np.float64)
        be = np.ascontiguousarray(base_ex, dtype=np.float64)
        tp = tb = None
        if trips is not None:
            tp = _i32(np.concatenate([[0], np.cumsum([len(x) for x in trips])]))
            tb = _i32(np.concatenate([np.asarray(x, np.int64) for x in trips]) if len(trips) else [])
        rows = self.rows
        o = np.empty((rows, m))
        e = np.empty((rows, m))
        _check(lib().g2g_batch_branch_flows(self._h, m, _ptr(f), _ptr(t), _ptr(y4), _ptr(bo), _ptr(be), _ptr(tp),
                                            _ptr(tb), _ptr(o), _ptr(e)))
        return o, e

    def limit_violations(self, table, trips=None, tol=1e-6, rel_tol=1e-9, capacity=16, rows=None):
        """Limit checks of the active rows at the batch's V.  ``table``: dict with
        from, to, yff, yft, ytf, ytt, base_or, base_ex, limit1_ka, limit2_ka,
        n_lines, vn_kv, vmin_kv, vmax_kv.  Returns a dict of the C ABI's outputs
        (records shaped [rows, 3 capacity])."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        keep = dict(fr=_i32(table["from"]), to=_i32(table["to"]),
                    y4=np.ascontiguousarray(np.stack([table["yff"], table["yft"], table["ytf"], table["ytt"]],
                                                     axis=1), dtype=np.complex128),
                    bo=f(table["base_or"]), be=f(table["base_ex"]), l1=f(table["limit1_ka"]),
                    l2=f(table["limit2_ka"]), vn=f(table["vn_kv"]), vmin=f(table["vmin_kv"]),
                    vmax=f(table["vmax_kv"]))
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        t = LimitTable(keep["fr"].size, vp(keep["fr"]), vp(keep["to"]), vp(keep["y4"]), vp(keep["bo"]),
                       vp(keep["be"]), vp(keep["l1"]), vp(keep["l2"]), int(table["n_lines"]), vp(keep["vn"]),
                       vp(keep["vmin"]), vp(keep["vmax"]))
        tp = tb = None
        if trips is not None:
            tp = _i32(np.concatenate([[0], np.cumsum([len(x) for x in trips])]))
            tb = _i32(np.concatenate([np.asarray(x, np.int64) for x in trips]) if len(trips) else [])
        R = self.rows if rows is None else int(rows)
        K = int(capacity)
        out = {k: np.zeros(R, np.int32) for k in ("count", "truncated", "n_low", "n_high", "n_current")}
        for k in ("element_type", "element_id", "side", "type"):
            out[k] = np.zeros((R, 3 * K), np.int32)
        for k in ("value", "limit"):
            out[k] = np.full((R, 3 * K), np.nan)
        v = Violations(*[vp(out[k]) for k, _ in Violations._fields_])
        _check(lib().g2g_batch_limit_violations(self._h, ctypes.byref(t), _ptr(tp), _ptr(tb), float(tol),
                                                float(rel_tol), K, ctypes.byref(v)))
        return out

    def warm_start(self, mlp, mask=None):
        """Writes the MLP's start voltages as the batch's V."""
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        _check(lib().g2g_batch_warm_start(self._h, mlp._h, _ptr(m), 0 if m is None else m.shape[1]))

    def get_v(self):
        V = np.empty((self.rows, self.n), np.complex128)
        _check(lib().g2g_batch_get_v(self._h, _ptr(V), self.n, 0))
        return V

    def get_v_device(self, ptr, ld):
        _check(lib().g2g_batch_get_v(self._h, _P(ptr), ld, 1))

    def results(self):
        st = np.empty(self.rows, np.int32)
        it = np.empty(self.rows, np.int32)
        res = np.empty(self.rows)
        _check(lib().g2g_batch_get_results(self._h, _ptr(st), _ptr(it), _ptr(res)))
        return st, it, res

    def ext_state(self):
        """[rows, n_xs] extension state (sa, then each controller's Q)."""
        k = self.plan.info()["n_xs"]
        out = np.zeros((self.rows, k))
        if k:
            _check(lib().g2g_batch_get_ext_state(self._h, _ptr(out)))
        return out

    def set_ext_state(self, xs):
        xs = np.ascontiguousarray(xs, dtype=np.float64)
        _check(lib().g2g_batch_set_ext_state(self._h, _ptr(xs)))

    def eval(self, with_j=True):
        inf = self.plan.info()
        F = np.empty((self.rows, inf["d"]))
        J = np.empty((self.rows, inf["nJ"])) if with_j else None
        r = np.empty(self.rows)
        _check(lib().g2g_batch_eval(self._h, _ptr(F), _ptr(J), _ptr(r)))
        return F, J, r

    def device_arrays(self):
        v, s = _P(), _P()
        _check(lib().g2g_batch_device_arrays(self._h, ctypes.byref(v), ctypes.byref(s)))
        return v.value, s.value


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Mlp:
    """The warm-start MLP (in -> H -> H -> out, GELU; bf16 MFMA, fp32 master
    weights and Adam state on the device)."""

    NAMES = ("W1", "b1", "W2", "b2", "W3", "b3")

    def __init__(self, n_in, n_hidden, n_out, seed=0, stream=None):
        self.n_in, self.H, self.n_out = int(n_in), int(n_hidden), int(n_out)
        h = _P()
        _check(lib().g2g_mlp_create(self.n_in, self.H, self.n_out, int(seed), stream, ctypes.byref(h)))
        self._h = h
        self._destroy = lib().g2g_mlp_destroy
        self.n_params = int(lib().g2g_mlp_num_params(h))
        off = np.empty(6, np.int64)
        _check(lib().g2g_mlp_param_offsets(h, _ptr(off)))
        self.offsets = off

    def __del__(self):
        # the destroy function is looked up while the module is alive (at
        # interpreter exit its globals may already be gone)
        h, self._h = getattr(self, "_h", None), None
        destroy = getattr(self, "_destroy", None)
        if h and destroy:
            destroy(h)

    def shapes(self):
        n_in, H, n_out = self.n_in, self.H, self.n_out
        return {"W1": (n_in, H), "b1": (H,), "W2": (H, H), "b2": (H,), "W3": (H, n_out), "b3": (n_out,)}

    def unflatten(self, flat):
        sh = self.shapes()
        return {k: flat[o:o + int(np.prod(sh[k]))].reshape(sh[k]) for k, o in zip(self.NAMES, self.offsets)}

    def flatten(self, params):
        flat = np.zeros(self.n_params, np.float32)
        for k, o in zip(self.NAMES, self.offsets):
            v = np.asarray(params[k], np.float32).ravel()
            flat[o:o + v.size] = v
        return flat

    def get_params(self):
        out = np.empty(self.n_params, np.float32)
        _check(lib().g2g_mlp_get_params(self._h, _ptr(out), out.size))
        return self.unflatten(out)

    def set_params(self, params):
        flat = self.flatten(params)
        _check(lib().g2g_mlp_set_params(self._h, _ptr(flat), flat.size))

    def set_normalization(self, in_mu, in_sigma, out_mu, out_sigma):
        a = [_f32(x) for x in (in_mu, in_sigma, out_mu, out_sigma)]
        _check(lib().g2g_mlp_set_normalization(self._h, *[_ptr(x) for x in a]))

    def set_buses(self, kind, vfix):
        kind = _i32(kind)
        vfix = np.ascontiguousarray(vfix, dtype=np.float64)
        _check(lib().g2g_mlp_set_buses(self._h, _ptr(kind), _ptr(vfix), kind.size))

    def forward(self, X):
        X = _f32(X)
        Y = np.empty((X.shape[0], self.n_out), np.float32)
        _check(lib().g2g_mlp_forward(self._h, _ptr(X), X.shape[0], _ptr(Y), 0))
        return Y

    def forward_device(self, x_ptr, M, y_ptr):
        _check(lib().g2g_mlp_forward(self._h, _P(x_ptr), int(M), _P(y_ptr), 1))

    def loss_grad(self, X, T):
        X, T = _f32(X), _f32(T)
        loss = np.empty(1, np.float32)
        g = np.empty(self.n_params, np.float32)
        _check(lib().g2g_mlp_loss_grad(self._h, _ptr(X), _ptr(T), X.shape[0], _ptr(loss), _ptr(g), 0))
        return float(loss[0]), self.unflatten(g)

    def train_step(self, X, T, lr):
        X, T = _f32(X), _f32(T)
        loss = np.empty(1, np.float32)
        _check(lib().g2g_mlp_train_step(self._h, _ptr(X), _ptr(T), X.shape[0], float(lr), _ptr(loss), 0))
        return float(loss[0])

    def set_grad_clip(self, max_norm):
        _check(lib().g2g_mlp_set_grad_clip(self._h, float(max_norm)))

    def train_step_device(self, x_ptr, t_ptr, M, lr):
        loss = np.empty(1, np.float32)
        _check(lib().g2g_mlp_train_step(self._h, _P(x_ptr), _P(t_ptr), int(M), float(lr), _ptr(loss), 1))
        return float(loss[0])


def gemm_bf16(A, B, trans_a=False, trans_b=False):
    """C = op(A) op(B) on the MLP's bf16 MFMA kernel; A and B are uint16 arrays
    of bf16 bit patterns (2-D, row-major, as stored)."""
    A = np.ascontiguousarray(A, dtype=np.uint16)
    B = np.ascontiguousarray(B, dtype=np.uint16)
    M = A.shape[1] if trans_a else A.shape[0]
    K = A.shape[0] if trans_a else A.shape[1]
    N = B.shape[0] if trans_b else B.shape[1]
    kb = B.shape[1] if trans_b else B.shape[0]
    if kb != K:
        raise ValueError("inner dimensions differ")
    C = np.empty((M, N), np.float32)
    _check(lib().g2g_gemm_bf16(int(trans_a), int(trans_b), M, N, K, _ptr(A), A.shape[1], _ptr(B), B.shape[1],
                               _ptr(C), N))
    return C
