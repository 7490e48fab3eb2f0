"""Streamed rows (continuous batching with slot refill) and the serving loop,
through the C ABI, against the CPU oracle."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import _capi
from gpusim2grid_amd import serving
from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec,outages", [(_capi.FP64, False), (_capi.FP64, True), (_capi.MIXED, True)])
def test_streamed_rows_match_oracle(prec, outages):
    """600 rows through 128 slots, refilled as rows finish: every row's
    status, update count and V are the oracle's for that row alone."""
    g = syn.grid_118()
    rows, slots = 600, 128
    _say(f"stream prec={prec} outages={outages}")
    p, q = syn.injection_scenarios(g, rows, seed=8)
    S = (p + 1j * q) / g.sn_mva
    outs = syn.random_outages(g, rows, seed=6) if outages else [[] for _ in range(rows)]
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, slots, precision=prec, per_row_ybus=True)
    B.stream_begin()
    free = list(range(slots))
    busy = {}
    nxt = 0
    got = {}
    rounds = 0
    while len(got) < rows:
        k = min(len(free), rows - nxt)
        if k:
            sl = [free.pop() for _ in range(k)]
            ids = list(range(nxt, nxt + k))
            B.stream_fill(sl, S[ids], g.flat_start()[None], edits=syn.outage_edits(g, [outs[i] for i in ids]))
            busy.update(zip(sl, ids))
            nxt += k
        B.stream_round(k=1, max_iter=10, tol=1e-8)
        rounds += 1
        assert rounds < 200
        st, it, _ = B.results()
        done = [s for s in busy if st[s] not in (-1, _capi.NOTSIMULATED)]
        if done:
            V = B.stream_take(done)
            for q_i, s in enumerate(done):
                got[busy.pop(s)] = (int(st[s]), int(it[s]), V[q_i])
                free.append(s)
    Yrows = np.stack([orc.ybus_values_on_pattern(g.Ybus, orc.outage_ybus(g, o)) for o in outs])
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8, n_threads=16, Yx_rows=Yrows)
    st = np.array([got[i][0] for i in range(rows)])
    it = np.array([got[i][1] for i in range(rows)])
    V = np.stack([got[i][2] for i in range(rows)])
    np.testing.assert_array_equal(st, sto)
    conv = sto == 0
    if prec == _capi.FP64:
        np.testing.assert_array_equal(it, ito)
        assert np.abs(V[conv] - Vo[conv]).max() < 1e-8
    else:
        assert (it[conv] <= ito[conv] + 2).all() and np.abs(V[conv] - Vo[conv]).max() < 1e-6
    # the slots were reused: far fewer rounds than rows
    assert rounds < rows // 8


def test_serving_loop_two_grids():
    """Poisson requests over two resident grids: every request comes back
    converged, with the oracle's solution, and latencies are reported."""
    grids = {"g118": syn.grid_118(), "c0": syn.grid_c0()}
    reqs = serving.poisson_requests(grids, rate=20000.0, n=1500, max_outages=2, seed=5)
    _say(f"serving {len(reqs)} requests")
    srv = serving.PowerFlowServer(grids, slots=256, precision=_capi.MIXED, hbm_budget=8e9)
    done, rep = srv.run(reqs, timeout=120.0)
    _say(str(rep))
    assert rep["requests"] == len(reqs)
    by = {}
    for r in done:
        by.setdefault(r.grid, []).append(r)
    for name, rs in by.items():
        g = grids[name]
        S = np.stack([r.S for r in rs])
        Yrows = np.stack([orc.ybus_values_on_pattern(g.Ybus, orc.outage_ybus(g, r.outages)) for r in rs])
        Vo, sto, _, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8, n_threads=16, Yx_rows=Yrows)
        st = np.array([r.status for r in rs])
        np.testing.assert_array_equal(st, sto)
        ok = sto == 0
        V = np.stack([r.V for r in rs])
        assert np.abs(V[ok] - Vo[ok]).max() < 1e-6
    assert rep["p50_ms"] > 0 and rep["p99_ms"] >= rep["p50_ms"]


def _train_warm(g, rows=4096, epochs=8, hidden=256):
    """A warm start for grid g trained on outage rows solved by the device
    (0-2 non-islanding outages, the mask among the inputs)."""
    import torch
    from gpusim2grid_amd.warmstart import WarmStart
    p, q = syn.injection_scenarios(g, rows, seed=2)
    S = (p + 1j * q) / g.sn_mva
    outs = syn.random_outages(g, rows, max_k=2, seed=4)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, rows, precision=_capi.MIXED, per_row_ybus=True)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.set_ybus_edits(*syn.outage_edits(g, outs))
    B.solve(max_iter=10, tol=1e-8)
    st, _, _ = B.results()
    V = B.get_v()
    del B
    ok = st == 0
    ws = WarmStart(g.n_bus, g.pv, g.pq, g.slack, np.abs(g.flat_start()), n_hidden=hidden, seed=2,
                   n_mask=g.n_branch)
    ws.train(torch.as_tensor(S[ok], device="cuda"), torch.as_tensor(V[ok], device="cuda"),
             mask=torch.as_tensor(syn.outage_mask(g, outs)[ok], device="cuda"), epochs=epochs, batch=256, lr=1e-3)
    return ws


def test_serving_three_grids_outages_and_warm_start_against_oracle():
    """The serving workload in small: Poisson requests over the 118-, 2,000-
    and 6,500-bus grids, 0-3 tripped branches each, resident per-grid batches
    with slot refill, learned warm starts with the flat-start fallback, mixed
    precision.  Every request comes back; per grid, a sample of requests is
    solved by the oracle from the flat start on the request's own outage
    Ybus: where the oracle converges the server converged to the same V, and
    every converged answer satisfies the oracle's mismatch test."""
    grids = {"g118": syn.grid_118(), "c1": syn.grid_c1(), "c2": syn.grid_c2()}
    warm = {}
    for name, g in grids.items():
        _say(f"serving c4: training the warm start of {name}")
        warm[name] = _train_warm(g)
    reqs = serving.poisson_requests(grids, rate=30000.0, n=3000, max_outages=3, seed=7)
    _say(f"serving c4: {len(reqs)} requests")
    srv = serving.PowerFlowServer(grids, slots={"g118": 512, "c1": 1024, "c2": 1024}, precision=_capi.MIXED,
                                  warm=warm, hbm_budget=64e9)
    done, rep = srv.run(reqs, timeout=300.0)
    _say(str(rep))
    assert rep["requests"] == len(reqs) and len({r.rid for r in done}) == len(reqs)
    by = {}
    for r in done:
        by.setdefault(r.grid, []).append(r)
    assert set(by) == set(grids)
    for name, rs in by.items():
        g = grids[name]
        rs = sorted(rs, key=lambda r: r.rid)[:200]
        _say(f"serving c4: oracle on {len(rs)} requests of {name}")
        S = np.stack([r.S for r in rs])
        Yrows = np.stack([orc.ybus_values_on_pattern(g.Ybus, orc.outage_ybus(g, r.outages)) for r in rs])
        Vo, sto, _, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8, n_threads=16, Yx_rows=Yrows)
        st = np.array([r.status for r in rs])
        V = np.stack([r.V for r in rs])
        ok = sto == 0
        assert ok.mean() > 0.95
        assert (st[ok] == 0).all()
        assert np.abs(V[ok] - Vo[ok]).max() < 1e-6
        for k in np.flatnonzero(st == 0):
            F = orc.mismatch_np(orc.outage_ybus(g, rs[k].outages), V[k], S[k], g.pv, g.pq)
            assert np.max(np.abs(F)) < 2e-8  # the device's own test was < 1e-8
    assert rep["p50_ms"] > 0 and rep["p99_ms"] >= rep["p50_ms"]
