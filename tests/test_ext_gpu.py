"""Extension rows (distributed slack, HVDC droop, voltage control) on the
HIP path against the extension oracle."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn
from oracle import ext_oracle as xo
from ext_cases import features

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rows(g, rows, seed=1):
    p, q = syn.injection_scenarios(g, rows, seed=seed)
    return (p + 1j * q) / g.sn_mva


@pytest.mark.parametrize("which", [("ds",), ("hvdc",), ("vc",), ("ds", "hvdc", "vc")])
def test_eval_matches_the_oracle(which):
    g = syn.grid_118()
    ext = features(g, which)
    L = xo.Layout(g.n_bus, g.pv, g.pq, g.slack, ext)
    rows = 70
    _say(f"ext eval {which} rows={rows}")
    rng = np.random.default_rng(2)
    V = np.array([g.flat_start() * np.exp(1j * rng.normal(0, 0.05, g.n_bus)) * (1 + rng.normal(0, 0.02, g.n_bus))
                  for _ in range(rows)])
    S = _rows(g, rows)
    xs = rng.normal(0, 0.05, (rows, L.n_xs))
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack, ext=ext)
    B = _capi.Batch(P, rows)
    B.set_v(V)
    B.set_sbus(S)
    B.set_ext_state(xs)
    F, J, r = B.eval(with_j=True)
    np.testing.assert_array_equal(B.ext_state(), xs)
    jp, ji = P.array("jp"), P.array("ji")
    for k in (0, 33, 69):
        Fo = xo.f_ext(g.Ybus, V[k], S[k], xs[k], L)
        Jo = xo.j_ext(g.Ybus, V[k], xs[k], L)
        np.testing.assert_allclose(F[k], Fo, rtol=0, atol=1e-12)
        Jg = np.zeros_like(Jo)
        for i in range(L.d):
            Jg[i, ji[jp[i]:jp[i + 1]]] = J[k, jp[i]:jp[i + 1]]
        np.testing.assert_allclose(Jg, Jo, rtol=1e-12, atol=1e-11)
        assert r[k] == pytest.approx(np.abs(Fo).max(), rel=1e-12)


@pytest.mark.parametrize("which,prec", [(("ds",), _capi.FP64), (("hvdc",), _capi.FP64), (("vc",), _capi.FP64),
                                        (("ds", "hvdc", "vc"), _capi.FP64), (("ds", "hvdc", "vc"), _capi.MIXED)])
def test_solve_matches_the_oracle(which, prec):
    g = syn.grid_118()
    ext = features(g, which)
    L = xo.Layout(g.n_bus, g.pv, g.pq, g.slack, ext)
    rows = 80  # two tiles, the second partly filled
    _say(f"ext solve {which} prec={prec} rows={rows}")
    S = _rows(g, rows, seed=3)
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack, ext=ext)
    B = _capi.Batch(P, rows, precision=prec)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.solve(max_iter=15, tol=1e-8)
    st, it, res = B.results()
    V = B.get_v()
    xs = B.ext_state()
    for k in range(rows):
        Vo, xso, sto, ito, _ = xo.nr_ext(g.Ybus, g.flat_start(), S[k], L, max_iter=15, tol=1e-8)
        assert st[k] == sto, k
        if prec == _capi.FP64:
            assert it[k] == ito, (k, it[k], ito)
            assert np.abs(V[k] - Vo).max() < 1e-9
            np.testing.assert_allclose(xs[k], xso, rtol=0, atol=1e-8)
        else:
            assert np.abs(V[k] - Vo).max() < 1e-7
            np.testing.assert_allclose(xs[k], xso, rtol=0, atol=1e-6)
    assert (st == 0).all()


def test_distributed_slack_on_the_2000_bus_grid():
    """Every PV bus shares the slack (an extension column with hundreds of
    entries, eliminated last) on the 2,000-bus grid, with outage rows."""
    g = syn.grid_c1()
    w = np.zeros(g.n_bus)
    w[np.asarray(g.pv)] = 1.0 / len(g.pv)
    from gpusim2grid_amd.extensions import Extensions
    ext = Extensions(slack_weights=w)
    L = xo.Layout(g.n_bus, g.pv, g.pq, g.slack, ext)
    rows = 12
    _say(f"ext solve ds c1 rows={rows}")
    S = _rows(g, rows, seed=5)
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack, ext=ext)
    B = _capi.Batch(P, rows)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.solve(max_iter=10, tol=1e-8)
    st, it, _ = B.results()
    V = B.get_v()
    xs = B.ext_state()
    for k in range(rows):
        Vo, xso, sto, ito, _ = xo.nr_ext(g.Ybus, g.flat_start(), S[k], L)
        assert st[k] == sto == 0 and it[k] == ito
        assert np.abs(V[k] - Vo).max() < 1e-9
        assert abs(xs[k, 0] - xso[0]) < 1e-8


def test_batch_size_does_not_change_the_rows(monkeypatch):
    """Rows solved in a batch of 320 (tail compaction moves the extension
    state with its rows) equal the same rows solved alone."""
    monkeypatch.setenv("G2G_COMPACT", "1")  # compact even this small a batch
    g = syn.grid_118()
    ext = features(g)
    rows = 320
    _say(f"ext compaction rows={rows}")
    S = _rows(g, rows, seed=8)
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack, ext=ext)
    B = _capi.Batch(P, rows)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.solve(max_iter=15, tol=1e-8)
    V, xs = B.get_v(), B.ext_state()
    for k in (0, 77, 319):
        B1 = _capi.Batch(P, 1)
        B1.set_sbus(S[k:k + 1])
        B1.set_v(g.flat_start())
        B1.solve(max_iter=15, tol=1e-8)
        np.testing.assert_array_equal(B1.get_v()[0], V[k])
        np.testing.assert_array_equal(B1.ext_state()[0], xs[k])


def test_unsupported_combinations_raise():
    g = syn.grid_118()
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack, ext=features(g, ("ds",)))
    B = _capi.Batch(P, 4)
    _say("ext unsupported combinations")
    with pytest.raises(ValueError):
        B.set_step_control(0, True, 0.5, 0.1)
    with pytest.raises(ValueError):
        B.stream_begin()
