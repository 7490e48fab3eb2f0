"""Limit-violation records of ``compute_limit_violations``: the int codes and
the record type of gpusim2grid's ``contingency_analysis/_limit_violations.py``
(``ViolationElementType``, ``LimitViolationType``, ``LimitViolation``; codes of
``src/_cpp/contingency/limit_violation_types.hpp:46-52``), and the unpacking of
a session's flat per-row record buffers into one list per row
(``contingency_analysis/__init__.py:616-673``)."""
from dataclasses import dataclass
from enum import IntEnum

import numpy as np

__all__ = ["ViolationElementType", "LimitViolationType", "LimitViolation", "violations_from_session"]


class ViolationElementType(IntEnum):
    BUS = 0
    LINE = 1
    TRAFO = 2
    GRID = 3  # the whole row (NOT_SIMULATED / DIVERGENCE)


class LimitViolationType(IntEnum):
    LOW_VOLTAGE = 0
    HIGH_VOLTAGE = 1
    CURRENT = 2
    NOT_SIMULATED = 3  # the row was skipped before any solve
    DIVERGENCE = 4     # solved, but its residual is NaN or above violation_tol


@dataclass(frozen=True)
class LimitViolation:
    """One record: element_id is the bus for BUS, the own-type (line or
    transformer) index for LINE / TRAFO, -1 for GRID; side 0 for buses and
    GRID, 1 (origin) or 2 (extremity) for branches; value / limit in kV
    (voltages), kA (currents), or residual / tol (DIVERGENCE)."""
    element_type: ViolationElementType
    element_id: int
    side: int
    violation_type: LimitViolationType
    value: float
    limit: float


def violations_from_session(s):
    """Per row, its records: a not-simulated row (count -1) gets one GRID /
    NOT_SIMULATED record with NaN value and limit."""
    counts = np.asarray(s.get_violation_count())
    et = np.asarray(s.get_violation_element_type())
    ei = np.asarray(s.get_violation_element_id())
    sd = np.asarray(s.get_violation_side())
    ty = np.asarray(s.get_violation_type())
    va = np.asarray(s.get_violation_value())
    li = np.asarray(s.get_violation_limit())
    stride = len(et) // max(len(counts), 1)
    out = []
    for r, c in enumerate(counts):
        if c < 0:
            out.append([LimitViolation(ViolationElementType.GRID, -1, 0, LimitViolationType.NOT_SIMULATED,
                                       float("nan"), float("nan"))])
            continue
        b = r * stride
        out.append([LimitViolation(ViolationElementType(int(et[b + i])), int(ei[b + i]), int(sd[b + i]),
                                   LimitViolationType(int(ty[b + i])), float(va[b + i]), float(li[b + i]))
                    for i in range(int(c))])
    return out
