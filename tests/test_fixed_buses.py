"""Several fixed buses: every bus in neither pv nor pq keeps its start V, as
the reference's array mode treats them (acpf_nr.cu:644-647,1732-1733)."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _two_more_fixed(g):
    """The 118-bus grid with two of its PV buses turned into fixed buses."""
    pv = np.asarray(g.pv)
    extra = pv[[3, len(pv) // 2]]
    return np.setdiff1d(pv, extra), extra


def test_plan_accepts_several_fixed_buses():
    g = syn.grid_118()
    pv, extra = _two_more_fixed(g)
    P = _capi.Plan.from_grid(g.Ybus, pv, g.pq, g.slack)
    inf = P.info()
    assert inf["npvpq"] == len(pv) + len(g.pq) and inf["d"] == inf["npvpq"] + len(g.pq)
    pvpq = P.array("pvpq")
    assert not np.isin(extra, pvpq).any() and g.slack not in pvpq
    # the pattern is the oracle's (the fixed buses have no rows or columns)
    _, Jo, _ = orc.f_and_j(g.Ybus, g.flat_start(), g.Sbus, pv, g.pq)
    np.testing.assert_array_equal(P.array("jp"), Jo.indptr)
    np.testing.assert_array_equal(P.array("ji"), Jo.indices)


def test_plan_rejects_a_slack_listed_in_pv():
    g = syn.grid_118()
    with pytest.raises(ValueError):
        _capi.Plan.from_grid(g.Ybus, np.append(g.pv, g.slack), g.pq, g.slack)


def test_session_rejects_a_grid_without_fixed_bus():
    from gpusim2grid_amd import _gpusim2grid as m
    g = syn.grid_118()
    pv = np.append(g.pv, g.slack)
    with pytest.raises(ValueError):
        m.AcPfNrSession(g.Ybus, g.flat_start(), g.Sbus, np.array([g.slack], np.int32), np.ones(1), pv.astype(np.int32),
                        np.asarray(g.pq, np.int32), 10, 1e-8)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [_capi.FP64, _capi.MIXED])
def test_gpu_three_fixed_buses_match_the_oracle(prec):
    g = syn.grid_118()
    pv, extra = _two_more_fixed(g)
    rows = 80
    _say(f"three fixed buses 118 prec={prec} rows={rows}")
    p, q = syn.injection_scenarios(g, rows, seed=7)
    S = (p + 1j * q) / g.sn_mva
    rng = np.random.default_rng(3)
    V0 = np.tile(g.flat_start(), (rows, 1))
    # the fixed buses hold angles of their own in each row
    V0[:, extra] *= np.exp(1j * rng.uniform(-0.05, 0.05, (rows, extra.size)))
    P = _capi.Plan.from_grid(g.Ybus, pv, g.pq, g.slack)
    B = _capi.Batch(P, rows, precision=prec)
    B.set_sbus(S)
    B.set_v(V0)
    B.solve(max_iter=10, tol=1e-8)
    st, it, _ = B.results()
    V = B.get_v()
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, V0, S, pv, g.pq, 10, 1e-8, n_threads=8)
    np.testing.assert_array_equal(st, sto)
    assert (st == 0).all()
    if prec == _capi.FP64:
        np.testing.assert_array_equal(it, ito)
    assert np.abs(V - Vo).max() < (1e-9 if prec == _capi.FP64 else 1e-7)
    np.testing.assert_array_equal(V[:, extra], V0[:, extra])  # fixed buses keep their start V


@pytest.mark.gpu
def test_gpu_facade_takes_several_slack_ids():
    """AcPfGPU with slack_ids listing every fixed bus (slack_weights are not
    read in array mode)."""
    from gpusim2grid_amd import AcPfGPU
    g = syn.grid_118()
    pv, extra = _two_more_fixed(g)
    _say("facade several slack ids 118")
    slack_ids = np.concatenate([np.atleast_1d(g.slack), extra])
    grid = (g.Ybus, g.flat_start(), g.Sbus, slack_ids, np.full(slack_ids.size, 1.0 / slack_ids.size), pv, g.pq)
    ac = AcPfGPU(grid, init_from_n_powerflow=False, max_iter=10)
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), g.Sbus, pv, g.pq, 10, 1e-8)
    assert ac.timings.converged and ac.timings.nb_iter == ito[0]
    assert np.abs(ac.solve() - Vo[0]).max() < 1e-9
