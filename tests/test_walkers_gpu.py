"""The factorization and backward substitution split over two walker
wavefronts (independent elimination subtrees, then the top part) give the
oracle's statuses, iteration counts and voltages, in both precisions; a plan
whose two-walker kernels do not fit the LDS falls back to one walker."""
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


def _sbus_rows(g, rows, seed=1):
    p, q = syn.injection_scenarios(g, rows, seed=seed)
    return (p + 1j * q) / g.sn_mva


@pytest.mark.parametrize("name,prec", [("118", _capi.FP64), ("118", _capi.MIXED), ("c1", _capi.FP64),
                                       ("c1", _capi.MIXED), ("c2", _capi.MIXED)])
def test_two_walkers_match_oracle(name, prec, monkeypatch):
    monkeypatch.setenv("G2G_WALKERS", "2")
    g = {"118": syn.grid_118, "c1": syn.grid_c1, "c2": syn.grid_c2}[name]()
    rows = 130
    _say(f"two walkers {name} prec={prec} rows={rows}")
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    assert P.info()["walkers"] == 2
    S = _sbus_rows(g, rows)
    B = _capi.Batch(P, rows, precision=prec)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.solve(max_iter=10, tol=1e-8)
    st, it, res = B.results()
    V = B.get_v()
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8, n_threads=16)
    conv = sto == 0
    assert conv.mean() > 0.9
    if prec == _capi.FP64:
        np.testing.assert_array_equal(st, sto)
        np.testing.assert_array_equal(it, ito)
        assert np.abs(V[conv] - Vo[conv]).max() < 1e-8
    else:
        assert (st[conv] == 0).all() and (res[conv] < 1e-8).all()
        assert (it[conv] <= ito[conv] + 2).all()
        assert np.abs(V[conv] - Vo[conv]).max() < 1e-6


@pytest.mark.parametrize("prec", [_capi.FP64, _capi.MIXED])
def test_two_walker_raw_solve_matches_one_walker(prec, monkeypatch):
    """Raw batched solves of the 118-bus Jacobian: the two-walker program
    solves the same systems as the one-walker program."""
    g = syn.grid_118()
    rng = np.random.default_rng(4)
    rows = 70
    Js, Fs = [], []
    for _ in range(rows):
        V = g.flat_start() * np.exp(1j * rng.normal(0, 0.05, g.n_bus))
        F, J, _ = orc.f_and_j(g.Ybus, V, g.Sbus, g.pv, g.pq)
        Js.append(J)
        Fs.append(F)
    vals = np.stack([J.data for J in Js])
    rhs = np.stack(Fs)
    out = {}
    for w in (1, 2):
        monkeypatch.setenv("G2G_WALKERS", str(w))
        P = _capi.Plan.from_matrix(Js[0])
        assert P.info()["walkers"] == w
        _say(f"raw solve walkers={w} prec={prec}")
        out[w], cl = P.lu_solve(vals, rhs, precision=prec)
        assert (cl == 0).all()
    tol = 1e-10 if prec == _capi.FP64 else 1e-3
    for r in range(rows):
        scale = np.abs(out[1][r]).max()
        assert np.abs(out[2][r] - out[1][r]).max() <= tol * scale
        assert np.abs(Js[r] @ out[2][r] - rhs[r]).max() <= (1e-9 if prec == _capi.FP64 else 1e-3) * np.abs(rhs[r]).max()
