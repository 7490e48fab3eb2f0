"""MI355X-native batched Newton-Raphson AC power flow.

The facades ``AcPfGPU``, ``InjectionSweepGPU``, ``ContingencyAnalysisGPU`` and
``ScenarioSweepGPU`` and ``warmup`` carry gpusim2grid's names
(``src/gpusim2grid/__init__.py:21-38``); they load the native module on first
use, so ``python -m gpusim2grid_amd.build`` works before anything is built.
"""

__version__ = "0.1.0"

_FACADES = ("AcPfGPU", "InjectionSweepGPU", "ContingencyAnalysisGPU", "ScenarioSweepGPU", "RowStatus")

_VIOLATIONS = ("ViolationElementType", "LimitViolationType", "LimitViolation")

__all__ = list(_FACADES) + list(_VIOLATIONS) + ["warmup"]


def __getattr__(name):
    if name in _FACADES:
        from . import facades
        return getattr(facades, name)
    if name in _VIOLATIONS:
        from . import violations
        return getattr(violations, name)
    if name == "warmup":
        from ._gpusim2grid import warmup
        return warmup
    raise AttributeError(name)
