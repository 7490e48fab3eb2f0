"""Extension rows of the Newton system: distributed slack, HVDC lines in
angle droop, and voltage control (remote-voltage generators and SVCs).

The reference reads these from a lightsim2grid grid model into its augmented
ledger (``src/_cpp/ledger_data.hpp:61-155``; kernels
``acpf_nr_kernels.cu:793-1068``); array mode has no such source, so here the
caller describes them directly.  The row/column layout is this library's own
(``include/gpusim2grid_amd.h``, ``g2g_extensions``): extension rows and
columns follow the bare ``[P(pvpq) | Q(pq)]`` system, the distributed slack
first, then each voltage-control group.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

HVDC_LINEAR, HVDC_SAT_12, HVDC_SAT_21, HVDC_TRIPPED = 0, 1, -1, 2


@dataclass
class HvdcLine:
    """p = p0 + k (theta1 - theta2) per unit from bus1 to bus2, converter loss
    factors lf1 / lf2, line resistance r (pu), send limits pmax12 / pmax21;
    status 0 linear, +1 / -1 saturated at pmax12 / pmax21, 2 tripped."""
    bus1: int
    bus2: int
    p0: float
    k: float
    lf1: float = 0.0
    lf2: float = 0.0
    r: float = 0.0
    pmax12: float = 1e9
    pmax21: float = 1e9
    status: int = HVDC_LINEAR


@dataclass
class VoltageGroup:
    """Controllers at buses ``buses`` (pq) hold |V(reg_bus)| + sum slope_c Q_c
    at v_set, sharing reactive power in proportion to ``weights``."""
    reg_bus: int
    v_set: float
    buses: list
    slopes: list = None
    weights: list = None

    def __post_init__(self):
        k = len(self.buses)
        self.slopes = [0.0] * k if self.slopes is None else list(self.slopes)
        self.weights = [1.0] * k if self.weights is None else list(self.weights)


@dataclass
class Extensions:
    slack_weights: np.ndarray = None   # [n] distributed-slack weights, or None
    hvdc: list = field(default_factory=list)
    voltage: list = field(default_factory=list)

    def empty(self):
        return self.slack_weights is None and not self.hvdc and not self.voltage

    def n_state(self):
        return (self.slack_weights is not None) + sum(len(g.buses) for g in self.voltage)

    def as_dict(self):
        """The ``extensions=`` argument of the ``_gpusim2grid`` sessions."""
        d = {"slack_weights": None if self.slack_weights is None
             else np.ascontiguousarray(self.slack_weights, dtype=np.float64), "hvdc": None, "vc": None}
        if self.hvdc:
            h = self.hvdc
            d["hvdc"] = {"bus1": np.array([x.bus1 for x in h], np.int32),
                         "bus2": np.array([x.bus2 for x in h], np.int32),
                         "status": np.array([x.status for x in h], np.int32)}
            for name in ("p0", "k", "lf1", "lf2", "r", "pmax12", "pmax21"):
                d["hvdc"][name] = np.array([getattr(x, name) for x in h], np.float64)
        if self.voltage:
            g = self.voltage
            d["vc"] = {"reg_bus": np.array([x.reg_bus for x in g], np.int32),
                       "v_set": np.array([x.v_set for x in g], np.float64),
                       "ptr": np.concatenate([[0], np.cumsum([len(x.buses) for x in g])]).astype(np.int32),
                       "bus": np.array([b for x in g for b in x.buses], np.int32),
                       "slope": np.array([v for x in g for v in x.slopes], np.float64),
                       "weight": np.array([v for x in g for v in x.weights], np.float64)}
        return d


class CExtensions(ctypes.Structure):
    """g2g_extensions"""
    _P = ctypes.c_void_p
    _fields_ = [("slack_weights", _P), ("n_hvdc", ctypes.c_int32), ("hvdc_bus1", _P), ("hvdc_bus2", _P),
                ("hvdc_status", _P), ("hvdc_p0", _P), ("hvdc_k", _P), ("hvdc_lf1", _P), ("hvdc_lf2", _P),
                ("hvdc_r", _P), ("hvdc_pmax12", _P), ("hvdc_pmax21", _P), ("n_vc_groups", ctypes.c_int32),
                ("vc_reg_bus", _P), ("vc_v_set", _P), ("vc_ptr", _P), ("vc_bus", _P), ("vc_slope", _P),
                ("vc_weight", _P)]


def to_c(ext: Extensions, n_bus: int):
    """(CExtensions, keep-alive list of the arrays it points to)."""
    keep = []
    if ext.slack_weights is not None and np.size(ext.slack_weights) != n_bus:
        raise ValueError(f"slack_weights has {np.size(ext.slack_weights)} entries, the grid {n_bus} buses")

    def arr(v, dt):
        a = np.ascontiguousarray(v, dtype=dt)
        keep.append(a)
        return a.ctypes.data if a.size else None

    c = CExtensions()
    if ext.slack_weights is not None:
        c.slack_weights = arr(ext.slack_weights, np.float64)
    h = ext.hvdc
    c.n_hvdc = len(h)
    if h:
        c.hvdc_bus1 = arr([x.bus1 for x in h], np.int32)
        c.hvdc_bus2 = arr([x.bus2 for x in h], np.int32)
        c.hvdc_status = arr([x.status for x in h], np.int32)
        for name in ("p0", "k", "lf1", "lf2", "r", "pmax12", "pmax21"):
            setattr(c, "hvdc_" + name, arr([getattr(x, name) for x in h], np.float64))
    g = ext.voltage
    c.n_vc_groups = len(g)
    if g:
        c.vc_reg_bus = arr([x.reg_bus for x in g], np.int32)
        c.vc_v_set = arr([x.v_set for x in g], np.float64)
        c.vc_ptr = arr(np.concatenate([[0], np.cumsum([len(x.buses) for x in g])]), np.int32)
        c.vc_bus = arr([b for x in g for b in x.buses], np.int32)
        c.vc_slope = arr([s for x in g for s in x.slopes], np.float64)
        c.vc_weight = arr([w for x in g for w in x.weights], np.float64)
    return c, keep
