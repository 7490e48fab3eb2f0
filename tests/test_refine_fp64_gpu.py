"""Refinement rounds with an FP64 residual in mixed precision: one Newton step
from the same start lands next to the FP64 step, where the plain mixed step
is visibly off; full solves match the oracle row by row."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _batch(plan, S, V0, prec, ir=0, fp64_res=False, outages=None, g=None):
    B = _capi.Batch(plan, S.shape[0], precision=prec, per_row_ybus=outages is not None)
    B.set_sbus(S)
    B.set_v(V0)
    if outages is not None:
        B.set_ybus_edits(*syn.outage_edits(g, outages))
    B.set_step_control(ir_steps=ir)
    B.set_refinement_residual(fp64_res)
    return B


@pytest.mark.parametrize("outages", [False, True])
def test_one_refined_step_lands_on_the_fp64_step(outages):
    g = syn.grid_c1()
    rows = 192
    _say(f"refine fp64 one step outages={outages}")
    p, q = syn.injection_scenarios(g, rows, seed=31)
    S = (p + 1j * q) / g.sn_mva
    outs = syn.random_outages(g, rows, seed=32) if outages else None
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    V1 = {}
    for name, prec, ir, f64 in (("fp64", _capi.FP64, 0, False), ("mixed", _capi.MIXED, 0, False),
                                ("mixed_ir", _capi.MIXED, 1, True), ("mixed_ir2", _capi.MIXED, 2, True)):
        B = _batch(plan, S, g.flat_start(), prec, ir, f64, outs, g)
        B.solve(max_iter=1, tol=1e-300)  # exactly one Newton update
        V1[name] = B.get_v()
    e_mixed = np.abs(V1["mixed"] - V1["fp64"]).max()
    e_ir = np.abs(V1["mixed_ir"] - V1["fp64"]).max()
    e_ir2 = np.abs(V1["mixed_ir2"] - V1["fp64"]).max()
    _say(f"step error vs FP64: mixed {e_mixed:.2e}, one FP64-residual round {e_ir:.2e}, two {e_ir2:.2e}")
    assert e_mixed > 0.0
    assert e_ir < 0.05 * e_mixed
    assert e_ir2 <= e_ir * 1.0001 + 1e-15


def test_refined_mixed_solves_match_the_oracle():
    g = syn.grid_118()
    rows = 300
    _say("refine fp64 full solves")
    p, q = syn.injection_scenarios(g, rows, seed=33)
    S = (p + 1j * q) / g.sn_mva
    outs = syn.random_outages(g, rows, seed=34)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _batch(plan, S, g.flat_start(), _capi.MIXED, 1, True, outs, g)
    B.solve(max_iter=10, tol=1e-8)
    st, it, res = B.results()
    V = B.get_v()
    Yrows = orc.outage_ybus_rows(g, outs)
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8, n_threads=16, Yx_rows=Yrows)
    np.testing.assert_array_equal(st, sto)
    ok = sto == 0
    np.testing.assert_array_equal(it[ok], ito[ok])  # the refined step is the FP64 step to rounding
    assert np.abs(V[ok] - Vo[ok]).max() < 1e-9
    assert (res[ok] < 1e-8).all()


def test_fp64_residual_rejected_with_extension_rows():
    from ext_cases import features
    g = syn.grid_118()
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack, ext=features(g, ("ds",)))
    B = _capi.Batch(P, 4)
    with pytest.raises(ValueError):
        B.set_refinement_residual(True)


@pytest.mark.parametrize("grid,ir", [("c1", 1), ("c1", 2), ("118", 1)])
def test_forward_walk_refinement_equals_a_second_factorization(monkeypatch, grid, ir):
    """The refinement's correction solve walks the factorization's program
    forward only, over the multipliers the factorization stored in the A
    stream; a second full factorization (G2G_IR_REFACTOR=1) gives bitwise the
    same solves.  The 2,000-bus plan has two walkers and columns longer than
    the kernels' unrolled cases; the 118-bus plan one walker.  (The split walk,
    which a batch of a few tiles would take, refactorizes.)"""
    g = syn.grid_c1() if grid == "c1" else syn.grid_118()
    rows = 192
    _say(f"forward-walk refinement {grid} ir={ir}")
    p, q = syn.injection_scenarios(g, rows, seed=35)
    S = (p + 1j * q) / g.sn_mva
    outs = syn.random_outages(g, rows, seed=36)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    if grid == "c1":
        assert plan.info()["walkers"] == 2 and plan.info()["max_c"] > 7
    monkeypatch.setenv("G2G_SPLIT", "0")  # the main walk (a few tiles would take the split walk)
    got = {}
    for refactor in ("0", "1"):
        monkeypatch.setenv("G2G_IR_REFACTOR", refactor)
        B = _batch(plan, S, g.flat_start(), _capi.MIXED, ir, True, outs, g)
        B.solve(max_iter=10, tol=1e-8)
        got[refactor] = (B.get_v(), *B.results())
    V0, st0, it0, r0 = got["0"]
    V1, st1, it1, r1 = got["1"]
    np.testing.assert_array_equal(st0, st1)
    np.testing.assert_array_equal(it0, it1)
    assert (st0 == 0).sum() > rows // 2
    np.testing.assert_array_equal(V0, V1)
    np.testing.assert_array_equal(r0, r1)
