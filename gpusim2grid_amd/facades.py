"""The reference's Python facades in array ("tuple") mode.

``AcPfGPU`` (``src/gpusim2grid/acpf_nr/gpu_facade.py``), ``InjectionSweepGPU``
(``injection_sweep/gpu_facade.py``), ``ContingencyAnalysisGPU``
(``contingency_analysis/gpu_facade.py``) and ``ScenarioSweepGPU``
(``scenario_sweep/gpu_facade.py``) take ``grid=(Ybus, Vinit, Sbus, slack_ids,
slack_weights, pv, pq)`` and drive the ``_gpusim2grid`` sessions, whose solves
run the HIP kernels.  Settings take effect on the next run, as in the
reference (``CLAUDE.md:130-138``).  A lightsim2grid grid object is not
accepted: the bridge is out of scope.

``compute()`` returns the voltages as a DLPack capsule aliasing the session's
device results (the reference's ``compute`` contract,
``injection_sweep/gpu_facade.py:515-524``): the next run overwrites them, so
clone a tensor to keep it.  With ``as_numpy=True`` it returns a host array.

``extensions=`` (an ``extensions.Extensions``) adds the distributed-slack,
HVDC-droop and voltage-control rows to the Newton system; the reference takes
these from a lightsim2grid grid, which array mode does not carry.
"""
from __future__ import annotations


import numpy as np

from . import _gpusim2grid as _cpp

__all__ = ["AcPfGPU", "InjectionSweepGPU", "ContingencyAnalysisGPU", "ScenarioSweepGPU", "RowStatus"]

RowStatus = _cpp.RowStatus

_SCHED = {"chunked": _cpp.BatchScheduling.Chunked, "continuous": _cpp.BatchScheduling.Continuous}


def _precision(precision):
    """'fp64' (default) or 'mixed' (FP32 Jacobian / LU / solves, FP64 mismatch,
    test and state).  The reference validates the name against its compiled
    real type (``_ls2g_utils._validate_precision``); here precision is chosen
    per session."""
    if precision is None or precision == "fp64":
        return _cpp.Precision.Fp64
    if precision == "mixed":
        return _cpp.Precision.Mixed
    if precision == "fp32":
        raise ValueError("precision='fp32': this build keeps the state in FP64; use 'mixed' for FP32 linear algebra")
    raise ValueError(f"unknown precision {precision!r} (expected 'fp64' or 'mixed')")


def _device(device):
    """None | int | 'cuda' | 'cuda:N' -> ordinal (-1: current device)."""
    if device is None or device == "cuda":
        return -1
    if isinstance(device, str):
        if device.startswith("cuda:"):
            return int(device[5:])
        raise ValueError(f"unknown device {device!r}")
    return int(device)


def _tuple(grid):
    if not isinstance(grid, (tuple, list)) or len(grid) != 7:
        raise TypeError("grid must be the array tuple (Ybus, Vinit, Sbus, slack_ids, slack_weights, pv, pq); "
                        "lightsim2grid grid objects need the bridge, which this build does not have")
    Ybus, Vinit, Sbus, slack_ids, slack_weights, pv, pq = grid
    return (Ybus, np.ascontiguousarray(Vinit, dtype=complex), np.ascontiguousarray(Sbus, dtype=complex),
            np.ascontiguousarray(np.atleast_1d(slack_ids), dtype=np.int32),
            np.ascontiguousarray(np.atleast_1d(slack_weights), dtype=np.float64),
            np.ascontiguousarray(pv, dtype=np.int32), np.ascontiguousarray(pq, dtype=np.int32))


def _count(name, value):
    """A non-negative integer option (a bool is not one)."""
    if type(value) is bool or not hasattr(value, "__index__"):
        raise TypeError(f"{name} must be an int, got {type(value).__name__}")
    value = value.__index__()
    if value < 0:
        raise ValueError(f"{name} must be >= 0, got {value}")
    return value


def _check_common(ir_n_steps, scaling_max_voltage_change, max_dVa, max_dVm, use_bridge):
    """The step options as session keywords.  Messages and exception types are
    the reference's (contingency_analysis/__init__.py:157-167).  An array
    tuple carries no grid configuration, so ``scaling_max_voltage_change=None``
    (inherit) means off; the thresholds default to 0.5 rad / 0.1 pu
    (acpf_nr/gpu_facade.py:196-198)."""
    if use_bridge:
        raise ValueError("use_bridge=True requires a lightsim2grid grid object, not an explicit-array tuple.")
    opts = dict(ir_n_steps=_count("ir_n_steps", ir_n_steps),
                scaling_max_voltage_change=bool(scaling_max_voltage_change), max_dVa=0.5, max_dVm=0.1)
    for key, val in (("max_dVa", max_dVa), ("max_dVm", max_dVm)):
        if val is not None:
            opts[key] = float(val)
    return opts


def _ext(extensions):
    if extensions is None:
        return None
    return extensions.as_dict() if hasattr(extensions, "as_dict") else extensions


class _HostResult:
    """Lazy host view of a result (the reference's ``.to_numpy()`` holders)."""

    def __init__(self, a):
        self._a = a

    def to_numpy(self):
        return self._a


class AcPfGPU:
    """Single-system AC Newton-Raphson on the GPU (the reference's
    ``acpf_nr/gpu_facade.py`` ``AcPfGPU``, tuple mode)."""

    def __init__(self, grid, *, precision="fp64", max_iter=10, tol=1e-8, device=None, init_from_n_powerflow=True,
                 use_bridge=None, reordering_alg="default", matching_alg="none", pivot_epsilon_alg="default",
                 ir_n_steps=0, debug_base_case=False, scaling_max_voltage_change=None, max_dVa=None, max_dVm=None,
                 use_distributed_slack=True, extensions=None):
        step = _check_common(ir_n_steps, scaling_max_voltage_change, max_dVa, max_dVm, use_bridge)
        Ybus, Vinit, Sbus, slack_ids, slack_weights, pv, pq = _tuple(grid)
        self._s = _cpp.AcPfNrSession(Ybus, Vinit, Sbus, slack_ids, slack_weights, pv, pq, int(max_iter), float(tol),
                                     _device(device), presolved_v=bool(init_from_n_powerflow),
                                     precision=_precision(precision), extensions=_ext(extensions), **step)

    def solve(self):
        """The solved complex voltage vector (host copy)."""
        return self._s.get_v()

    def resolve(self, Vinit, Sbus=None, *, max_iter=None, tol=None):
        """Solve again from ``Vinit`` (and new ``Sbus``) reusing the plan;
        ``max_iter`` / ``tol`` default to the constructor's."""
        opt = (-1 if max_iter is None else int(max_iter), -1.0 if tol is None else float(tol))
        S = None if Sbus is None else np.ascontiguousarray(Sbus, dtype=complex)
        return self._s.resolve(np.ascontiguousarray(Vinit, dtype=complex), S, *opt)

    def ac_pf(self, Vinit, max_iter, tol):
        """``LSGrid.ac_pf``'s contract: the session solves from ``Vinit`` with
        the PV and fixed buses' |V| reset to their set-points (taken from the
        first start); the solved V, or an empty array when the row did not
        converge."""
        start = np.ascontiguousarray(Vinit, dtype=complex)
        if not self._s.solve_at_setpoints(start, int(max_iter), float(tol)):
            return np.empty(0, dtype=complex)
        return self._s.get_v()

    def get_F(self):
        return self._s.get_F()

    def v_dlpack(self):
        """DLPack (n_bus,) complex128 capsule aliasing the session's device V
        (the next resolve overwrites it)."""
        return self._s.v_dlpack()

    def ext_state(self):
        """Extension state of the solve: sa (distributed slack), then each
        voltage controller's Q (per unit)."""
        return self._s.get_ext_state()

    def solve_JT_dlpack(self, rhs):
        """Solve J^T x = rhs at the solved V (the adjoint) -> DLPack capsule of
        a device float64 vector."""
        import torch
        from torch.utils import dlpack
        r = rhs.cpu().numpy() if hasattr(rhs, "cpu") else np.asarray(rhs)
        x = self._s.solve_JT(np.ascontiguousarray(r, dtype=np.float64).ravel())
        return dlpack.to_dlpack(torch.from_numpy(x).to("cuda"))

    def get_J(self):
        """The Jacobian at the current V as a scipy CSR matrix."""
        import scipy.sparse as sp
        indptr, indices, data = self._s.get_J()
        d = self._s.dim_J
        return sp.csr_matrix((data, indices, indptr), shape=(d, d))

    @property
    def timings(self):
        return self._s.timings

    @property
    def session(self):
        return self._s


class _BatchFacade:
    _session_cls = None

    def __init__(self, grid, *, init_from_n_powerflow=True, precision="fp64", nb_iter=4, max_iter_base=10,
                 tol_base=1e-8, device=None, use_bridge=None, ir_n_steps=0, scaling_max_voltage_change=None,
                 max_dVa=None, max_dVm=None, handle_disconnected_grid=False, compute_limit_violations=False,
                 scheduling=None, nb_iter_per_round=1, tol=None, linear_solver="own_lu", extensions=None,
                 **_ignored):
        step = _check_common(ir_n_steps, scaling_max_voltage_change, max_dVa, max_dVm, use_bridge)
        if nb_iter_per_round != 1:
            raise ValueError("nb_iter_per_round: rows are checked after every iteration in this build (1)")
        if linear_solver not in ("cudss", "own_lu"):
            raise ValueError(f"unknown linear_solver {linear_solver!r}")
        Ybus, Vinit, Sbus, slack_ids, slack_weights, pv, pq = _tuple(grid)
        self._s = self._session_cls(Ybus, Vinit, Sbus, slack_ids, slack_weights, pv, pq, 100, int(nb_iter),
                                    int(max_iter_base), float(tol_base), _device(device),
                                    presolved_v=bool(init_from_n_powerflow), precision=_precision(precision),
                                    extensions=_ext(extensions), **step)
        self.scheduling = scheduling or "chunked"
        self._s.handle_disconnected_grid = bool(handle_disconnected_grid)
        self._s.compute_limit_violations = bool(compute_limit_violations)
        if tol is not None:
            self.tol = tol
        self._last_residuals = None
        self._warm = None

    # settings ---------------------------------------------------------------
    @property
    def handle_disconnected_grid(self):
        """bool: solve a row whose trips split the grid on its largest connected
        component (the other buses frozen in J and reported as NaN) instead of
        skipping it; a split that strands the slack is still skipped
        (``contingency_analysis/__init__.py:496-508``).  Next run."""
        return self._s.handle_disconnected_grid

    @handle_disconnected_grid.setter
    def handle_disconnected_grid(self, value):
        self._s.handle_disconnected_grid = bool(value)

    def get_disconnected(self):
        """Per row, True when its trips split the grid and it was not simulated."""
        return self._s.get_disconnected().astype(bool)

    # limit violations (contingency_analysis/__init__.py:589-735) ----------------
    @property
    def compute_limit_violations(self):
        """bool: check every solved row against the limits of :meth:`set_limits`
        right after its chunk's solve.  Next run."""
        return self._s.compute_limit_violations

    @compute_limit_violations.setter
    def compute_limit_violations(self, value):
        self._s.compute_limit_violations = bool(value)

    @property
    def violation_tol(self):
        """Residual above which a solved row gets a DIVERGENCE record (1e-6)."""
        return self._s.violation_tol

    @violation_tol.setter
    def violation_tol(self, value):
        self._s.violation_tol = float(value)

    @property
    def violation_rel_tol(self):
        """Relative margin a value must clear past its limit (1e-9)."""
        return self._s.violation_rel_tol

    @violation_rel_tol.setter
    def violation_rel_tol(self, value):
        self._s.violation_rel_tol = float(value)

    @property
    def violation_capacity(self):
        """Records kept per row and per violation type, the most severe (16)."""
        return self._s.violation_capacity

    @violation_capacity.setter
    def violation_capacity(self, value):
        self._s.violation_capacity = int(value)

    def set_limits(self, bus_vmin_kv, bus_vmax_kv, branch_limit_a1_ka, branch_limit_a2_ka, n_lines):
        """Per-bus voltage limits (kV) and per-branch current limits (kA, lines
        then transformers); NaN = none.  ``n_lines`` splits the branch ids into
        line and transformer ids of the records.  Needs set_branch_data()."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        self._s.set_limits(f(bus_vmin_kv), f(bus_vmax_kv), f(branch_limit_a1_ka), f(branch_limit_a2_ka),
                           int(n_lines))

    def get_violations(self):
        """Per row, its list of :class:`LimitViolation`: CURRENT, then LOW_VOLTAGE,
        then HIGH_VOLTAGE records, each type its most severe first."""
        from .violations import violations_from_session
        if not self._s.compute_limit_violations:
            raise RuntimeError("get_violations() requires compute_limit_violations=True")
        return violations_from_session(self._s)

    def get_violations_truncated(self):
        """Per row, True where a type had more violations than violation_capacity."""
        return self._s.get_violation_truncated().astype(bool)

    def get_violation_counts(self):
        """The true counts per row ('low_voltage', 'high_voltage', 'current'), -1
        for a row that was not simulated."""
        return {"low_voltage": np.asarray(self._s.get_violation_count_low_voltage()),
                "high_voltage": np.asarray(self._s.get_violation_count_high_voltage()),
                "current": np.asarray(self._s.get_violation_count_current())}

    @property
    def nb_iter(self):
        return self._s.nb_iter

    @nb_iter.setter
    def nb_iter(self, v):
        self._s.nb_iter = int(v)

    @property
    def scheduling(self):
        return "chunked" if self._s.scheduling == _cpp.BatchScheduling.Chunked else "continuous"

    @scheduling.setter
    def scheduling(self, v):
        if v not in _SCHED:
            raise ValueError(f"unknown scheduling {v!r} (expected 'chunked' or 'continuous')")
        self._s.scheduling = _SCHED[v]

    @property
    def tol(self):
        return self._s.tol

    @tol.setter
    def tol(self, v):
        self._s.tol = float(v)

    @property
    def precision(self):
        return "fp64" if self._s.precision == _cpp.Precision.Fp64 else "mixed"

    @precision.setter
    def precision(self, v):
        self._s.precision = _precision(v)

    def set_branch_data(self, branch_from, branch_to, yff_eff, yft_eff, ytf_eff, ytt_eff, bus_vn_kv=None,
                        sn_mva=100.0):
        """Per-branch pi-model admittances (from / to bus, yff, yft, ytf, ytt),
        bus nominal voltages (kV; default 1) and the system base (MVA), for
        outages and :meth:`compute_flows`."""
        n = self._s.n_bus
        vn = np.ones(n) if bus_vn_kv is None else np.ascontiguousarray(bus_vn_kv, dtype=float)
        self._s.set_branch_data(np.ascontiguousarray(branch_from, dtype=np.int32),
                                np.ascontiguousarray(branch_to, dtype=np.int32),
                                *[np.ascontiguousarray(y, dtype=complex) for y in (yff_eff, yft_eff, ytf_eff,
                                                                                   ytt_eff)],
                                vn, float(sn_mva))

    def compute_flows(self):
        """Branch currents (A) of every row at its result, on the device;
        tripped branches carry 0 and rows not simulated NaN."""
        self._s.compute_flows()

    @property
    def or_amps(self):
        return self._s.get_or_amps()

    @property
    def ex_amps(self):
        return self._s.get_ex_amps()

    @property
    def n_branches(self):
        return self._s.n_branches

    def set_warm_start(self, warm, k_warm=4, growth=2.0):
        """Learned warm start (``gpusim2grid_amd.warmstart.WarmStart``) with
        flat-start fallback on the next runs; None switches it off."""
        self._warm = warm
        if warm is None:
            self._s.clear_warm_start()
        else:
            self._s.set_warm_start(warm.mlp._h.value, int(k_warm), float(growth))

    # running ----------------------------------------------------------------
    def compute(self, batch_size=512, as_numpy=False):
        """Solve every row; a DLPack (n_rows, n_bus) complex128 capsule aliasing
        the session's device results, which the next run overwrites
        (``as_numpy``: a host copy).  Residuals go to :meth:`last_residuals`."""
        self._s.batch_size = int(batch_size)
        self._s.run()
        self._last_residuals = self._s.get_residuals()
        if as_numpy:
            return self._s.get_V_results().reshape(-1, self._s.n_bus)
        return self._s.v_results_dlpack()

    def ext_state(self):
        """[n_rows, n_state] extension state of the last run (sa, then each
        voltage controller's Q; NaN for rows not simulated)."""
        return self._s.get_ext_state()

    def last_residuals(self):
        if self._last_residuals is None:
            raise RuntimeError("Call compute() before last_residuals().")
        return self._last_residuals

    @property
    def V_results(self):
        return _HostResult(self._s.get_V_results().reshape(-1, self._s.n_bus))

    def get_row_status(self):
        return self._s.get_row_status()

    def get_row_iterations(self):
        return self._s.get_row_iterations()

    def converged(self, tol=None):
        r = self.last_residuals()
        return r < (self.tol if tol is None else tol)

    @property
    def timings(self):
        return self._s.timings

    @property
    def n_bus(self):
        return self._s.n_bus

    @property
    def solver(self):
        return self._s


class InjectionSweepGPU(_BatchFacade):
    """Many injection profiles on one grid (``injection_sweep/gpu_facade.py``)."""

    _session_cls = _cpp.InjectionSweepSession

    def set_injections(self, p_mw, q_mvar, sn_mva):
        """(n_scenarios, n_bus) MW / MVAr per bus; Sbus = (p + jq) / sn_mva."""
        self._s.set_injections(np.ascontiguousarray(p_mw, dtype=float), np.ascontiguousarray(q_mvar, dtype=float),
                               float(sn_mva))

    def set_v_init_from(self, V):
        """Start voltages: one row, one per scenario, or None (the base case)."""
        self._s.set_v_init_from(None if V is None else np.ascontiguousarray(V, dtype=complex))

    def set_gen_v(self, gen_v, gen_bus):
        """Per-scenario generator voltage set-points (n_scenarios, n_gen) in per
        unit, generator g holding bus gen_bus[g]: the start |V| of that bus when
        it is PV or slack (``injection_sweep/gpu_facade.py:457-511``; array mode
        names the buses, which the reference reads from its lightsim2grid grid)."""
        self._s.set_gen_v(np.ascontiguousarray(gen_v, dtype=np.float64),
                          np.ascontiguousarray(gen_bus, dtype=np.int32))

    @property
    def n_scenarios(self):
        return self._s.n_scenarios


class ContingencyAnalysisGPU(_BatchFacade):
    """N-k branch outages as value edits on the shared Ybus pattern
    (``contingency_analysis/gpu_facade.py``)."""

    _session_cls = _cpp.ContingencyAnalysisSession

    def __init__(self, grid, **kw):
        super().__init__(grid, **kw)
        self._ctg = []

    def add_contingencies_by_branch_id(self, branch_ids_per_ctg):
        """One contingency per list of tripped branch ids."""
        self._ctg.extend([list(map(int, c)) for c in branch_ids_per_ctg])
        self._s.build_contingencies(self._ctg)

    def add_contingency(self, branch_ids=()):
        self.add_contingencies_by_branch_id([list(branch_ids)])

    def get_contingencies(self):
        return [list(c) for c in self._ctg]

    @property
    def n_contingencies(self):
        return self._s.n_contingencies


class ScenarioSweepGPU(InjectionSweepGPU):
    """Per-row injections and per-row branch trips together
    (``scenario_sweep/gpu_facade.py``)."""

    _session_cls = _cpp.ScenarioSweepSession

    def set_topology(self, branch_ids_per_scenario):
        self._s.set_topology([list(map(int, c)) for c in branch_ids_per_scenario])
