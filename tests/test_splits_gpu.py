"""The mismatch and update kernels split each tile's buses over several
workgroups (per-split residual partials combined by a one-wave status kernel).
Only the order of an exact maximum changes, so a split batch gives bitwise the
voltages, residuals, statuses and iteration counts of an unsplit one: with
and without step scaling (which keeps the update unsplit), in both
precisions, and in both scheduling modes."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _solve(g, S, prec, splits, monkeypatch, scaling, scheduling):
    monkeypatch.setenv("G2G_MIS_SPLITS", str(splits))
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(P, S.shape[0], precision=prec)
    if scaling:
        B.set_step_control(scaling=True, max_dva=0.2, max_dvm=0.05)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.solve(max_iter=10, tol=1e-8, scheduling=scheduling)
    st, it, res = B.results()
    return B.get_v(), st, it, res


@pytest.mark.parametrize("name,prec,scaling,scheduling", [
    ("c1", _capi.FP64, False, _capi.CONTINUOUS),
    ("c1", _capi.MIXED, False, _capi.CONTINUOUS),
    ("c1", _capi.FP64, True, _capi.CONTINUOUS),
    ("c1", _capi.MIXED, False, _capi.CHUNKED),
    ("c2", _capi.MIXED, False, _capi.CONTINUOUS),
])
def test_split_kernels_match_one_workgroup_per_tile(name, prec, scaling, scheduling, monkeypatch):
    g = {"c1": syn.grid_c1, "c2": syn.grid_c2}[name]()
    rows = 200  # four tiles, the last one partial
    p, q = syn.injection_scenarios(g, rows, seed=7)
    S = (p + 1j * q) / g.sn_mva
    _say(f"splits {name} prec={prec} scaling={scaling} scheduling={scheduling} rows={rows}")
    V1, st1, it1, r1 = _solve(g, S, prec, 1, monkeypatch, scaling, scheduling)
    if not scaling:  # damped steps may need more than the budget; only equality matters there
        assert (st1 == 0).mean() > 0.9
    for splits in (2, 4):
        V, st, it, r = _solve(g, S, prec, splits, monkeypatch, scaling, scheduling)
        np.testing.assert_array_equal(st, st1)
        np.testing.assert_array_equal(it, it1)
        np.testing.assert_array_equal(r, r1)
        np.testing.assert_array_equal(V, V1)
