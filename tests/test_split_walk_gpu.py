"""The split walk on the GPU (G2G_SPLIT=1 forces it from the first
iteration): statuses, iteration counts and voltages of the CPU oracle."""
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


def _solve(g, S, rows, prec, edits=None):
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(P, rows, precision=prec, per_row_ybus=edits is not None)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    if edits is not None:
        B.set_ybus_edits(*edits)
    stats = B.solve(max_iter=10, tol=1e-8)
    st, it, res = B.results()
    return B.get_v(), st, it, stats


@pytest.mark.parametrize("prec", [_capi.FP64, _capi.MIXED])
def test_split_walk_matches_oracle(monkeypatch, prec):
    g = syn.grid_c1()
    rows = 150  # three tiles, the last partly filled
    p, q = syn.injection_scenarios(g, rows, seed=31)
    S = (p + 1j * q) / g.sn_mva
    _say(f"split walk c1 prec={prec} rows={rows}")
    monkeypatch.setenv("G2G_SPLIT", "1")
    V, st, it, stats = _solve(g, S, rows, prec)
    monkeypatch.setenv("G2G_SPLIT", "0")
    Vm, stm, itm, stats_m = _solve(g, S, rows, prec)
    # the solve stats split the factorizations by walk: none is a main-layout
    # walk when the split walk is forced, all are when it is off
    assert stats["n_factor_main"] == 0 and stats["ms_factor_main"] == 0.0 and stats["rows_factored_main"] == 0
    assert stats_m["n_factor_main"] == stats_m["n_factor_launches"] > 0
    assert stats_m["rows_factored_main"] == stats_m["rows_factored"]
    assert 0.0 < stats_m["ms_factor_main"] <= stats_m["ms_factor"] * 1.0001
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8, n_threads=16)
    np.testing.assert_array_equal(st, sto)
    np.testing.assert_array_equal(stm, sto)
    assert (st == 0).all()
    if prec == _capi.FP64:
        np.testing.assert_array_equal(it, ito)
        assert np.abs(V - Vo).max() < 1e-9
        assert np.abs(V - Vm).max() < 1e-11  # the two walks differ in summation order only
    else:
        assert np.abs(V - Vo).max() < 1e-7


def test_split_walk_on_outage_rows_and_compaction(monkeypatch):
    """Outage rows (per-row Ybus) in a batch large enough to compact its tail,
    the split walk taken at the tail (default rule) and forced."""
    g = syn.grid_c1()
    rows = 400
    p, q = syn.injection_scenarios(g, rows, seed=32)
    S = (p + 1j * q) / g.sn_mva
    outs = syn.random_outages(g, rows, max_k=2, seed=33)
    edits = syn.outage_edits(g, outs)
    _say(f"split walk c1 outages rows={rows}")
    # per-row Ybus values: the base values less each row's edits (the edits
    # themselves are pinned against rebuilt Ybus in test_gpu_parity)
    Y = g.Ybus.tocsr()
    Y.sort_indices()
    row_ptr, ent, delta = edits
    Yx = np.tile(Y.data, (rows, 1))
    for r in range(rows):
        for k in range(row_ptr[r], row_ptr[r + 1]):
            Yx[r, ent[k]] -= delta[k]
    Vo, sto, ito, _ = orc.nr_batch(Y, g.flat_start(), S, g.pv, g.pq, 10, 1e-8, n_threads=16, Yx_rows=Yx)
    for mode in ("1", None):
        if mode is None:
            monkeypatch.delenv("G2G_SPLIT", raising=False)
        else:
            monkeypatch.setenv("G2G_SPLIT", mode)
        V, st, it, _ = _solve(g, S, rows, _capi.FP64, edits=edits)
        np.testing.assert_array_equal(st, sto)
        np.testing.assert_array_equal(it, ito)
        assert np.abs(V[st == 0] - Vo[st == 0]).max() < 1e-9
