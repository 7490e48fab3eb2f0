"""Limit violations of solved rows (compute_limit_violations): the HIP check
(through the C ABI and the facades) against the oracle, and the oracle
against hand-computed records -- modelled on the reference's
tests/python/test_limit_violations.py cases (divergence record, tripped
branches skipped, capacity truncation with true counts, ties, rel_tol)."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tuple(g):
    return (g.Ybus, g.flat_start(), g.Sbus, np.array([g.slack]), np.array([1.0]), g.pv, g.pq)


# ---- CPU: the oracle against hand-computed records ----------------------------
def _hand_case():
    """3 buses, 4 series branches (y = 1 / (0.01 + 0.1j)): 0->1 and 1->2 are
    lines, 0->2 and a copy of 0->1 are transformers (n_lines = 2)."""
    V = np.array([1.0, 0.98 * np.exp(-0.05j), 0.97 * np.exp(-0.08j)])
    y = 1.0 / (0.01 + 0.1j)
    f = np.array([0, 1, 0, 0])
    t = np.array([1, 2, 2, 1])
    vn = np.array([138.0, 138.0, 69.0])
    base = 100e6 / (np.sqrt(3) * vn[f] * 1e3)
    # |I_or| = |I_ex| = |y| |V_f - V_t| for a series branch
    ka = np.abs(y) * np.abs(V[f] - V[t]) * base * 1e-3
    table = dict(**{"from": f}, to=t, yff=np.full(4, y), yft=np.full(4, -y), ytf=np.full(4, -y),
                 ytt=np.full(4, y), base_or=base, base_ex=base,
                 limit1_ka=np.array([0.5 * ka[0], np.nan, 2.0 * ka[2], 0.5 * ka[3]]),
                 limit2_ka=np.array([np.nan, 0.8 * ka[1], np.nan, np.nan]), n_lines=2, vn_kv=vn,
                 vmin_kv=np.array([np.nan, np.nan, 69.0]), vmax_kv=np.array([np.nan, 0.9 * 138.0, np.nan]))
    return V, ka, table


def _close(recs, want):
    assert len(recs) == len(want), (recs, want)
    for a, b in zip(recs, want):
        assert a[:4] == b[:4], (a, b)
        np.testing.assert_allclose(a[4:], b[4:], rtol=1e-13)


def test_oracle_hand_case():
    V, ka, table = _hand_case()
    vm = np.abs(V) * table["vn_kv"]
    cur = [(1, 0, 1, 2, ka[0], 0.5 * ka[0]),   # severity 1.0
           (2, 1, 1, 2, ka[3], 0.5 * ka[3]),   # the same severity: after, by scan order
           (1, 1, 2, 2, ka[1], 0.8 * ka[1])]   # severity 0.25
    low = [(0, 2, 0, 0, vm[2], 69.0)]
    high = [(0, 1, 0, 1, vm[1], 0.9 * 138.0)]
    ok = np.array([0, 0, 0])
    res = np.full(3, 1e-10)
    (o,) = orc.limit_violations(V[None], res[:1], ok[:1], table)
    _close(o["records"], cur + low + high)
    assert (o["count"], o["truncated"], o["n_low"], o["n_high"], o["n_current"]) == (5, 0, 1, 1, 3)
    # capacity 1: the most severe of each type, true counts kept
    (o,) = orc.limit_violations(V[None], res[:1], ok[:1], table, K=1)
    _close(o["records"], cur[:1] + low + high)
    assert (o["count"], o["truncated"], o["n_current"]) == (3, 1, 3)
    # a tripped branch is skipped
    (o,) = orc.limit_violations(V[None], res[:1], ok[:1], table, trips=[[0]])
    _close(o["records"], cur[1:] + low + high)
    # the relative margin: 2.0x clears 1.9, 1.25x does not; no voltage is 90 % off
    (o,) = orc.limit_violations(V[None], res[:1], ok[:1], table, rel_tol=0.9)
    _close(o["records"], cur[:2])
    # divergence and not simulated
    o = orc.limit_violations(np.stack([V, V, V]), np.array([1e-3, np.nan, 1e-10]), np.array([1, 2, 3]), table)
    assert o[0]["records"] == [(3, -1, 0, 4, 1e-3, 1e-6)] and o[0]["n_current"] == 0
    assert o[1]["count"] == 1 and np.isnan(o[1]["records"][0][4])
    assert o[2]["count"] == -1 and o[2]["n_low"] == -1
    # a NaN (islanded) voltage never violates
    V2 = V.copy()
    V2[2] = np.nan
    (o,) = orc.limit_violations(V2[None], res[:1], ok[:1], table)
    assert [r[:4] for r in o["records"]] == [(1, 0, 1, 2), (2, 1, 1, 2), (0, 1, 0, 1)]


# ---- GPU ----------------------------------------------------------------------
def _table(g, V_ref, seed):
    """Limits around one solved row's currents and voltages: 15 % unset, a few
    zero current limits (infinite severities, so exact ties)."""
    rng = np.random.default_rng(seed)
    m, n = g.n_branch, g.n_bus
    vn = rng.choice([69.0, 138.0, 345.0], size=n)
    base_or = g.sn_mva * 1e6 / (np.sqrt(3) * vn[g.branch_from] * 1e3)
    base_ex = g.sn_mva * 1e6 / (np.sqrt(3) * vn[g.branch_to] * 1e3)
    io = np.abs(g.yff * V_ref[g.branch_from] + g.yft * V_ref[g.branch_to]) * base_or * 1e-3
    ie = np.abs(g.ytf * V_ref[g.branch_from] + g.ytt * V_ref[g.branch_to]) * base_ex * 1e-3
    l1 = io * rng.uniform(0.7, 1.5, m)
    l2 = ie * rng.uniform(0.7, 1.5, m)
    l1[rng.random(m) < 0.15] = np.nan
    l2[rng.random(m) < 0.15] = np.nan
    l1[rng.choice(m, 3, replace=False)] = 0.0
    vm = np.abs(V_ref) * vn
    vmin = vm * rng.uniform(0.97, 1.01, n)
    vmax = vm * rng.uniform(0.99, 1.03, n)
    vmin[rng.random(n) < 0.15] = np.nan
    vmax[rng.random(n) < 0.15] = np.nan
    return dict(**{"from": g.branch_from, "to": g.branch_to}, yff=g.yff, yft=g.yft, ytf=g.ytf, ytt=g.ytt,
                base_or=base_or, base_ex=base_ex, limit1_ka=l1, limit2_ka=l2, n_lines=int(0.8 * m), vn_kv=vn,
                vmin_kv=vmin, vmax_kv=vmax)


def _compare(out, want, K):
    rows = len(want)
    for key in ("count", "truncated", "n_low", "n_high", "n_current"):
        np.testing.assert_array_equal(out[key][:rows], [w[key] for w in want], err_msg=key)
    for r, w in enumerate(want):
        c = w["count"]
        got = list(zip(out["element_type"][r, :c], out["element_id"][r, :c], out["side"][r, :c],
                       out["type"][r, :c]))
        assert got == [x[:4] for x in w["records"]], r
        v = np.array([x[4] for x in w["records"]], dtype=float).reshape(-1)
        lim = np.array([x[5] for x in w["records"]], dtype=float).reshape(-1)
        np.testing.assert_allclose(out["value"][r, :c], v, rtol=1e-15, equal_nan=True)
        np.testing.assert_allclose(out["limit"][r, :c], lim, rtol=1e-15, equal_nan=True)


@pytest.mark.gpu
@pytest.mark.parametrize("K", [2, 64])
def test_kernel_matches_oracle(K):
    from gpusim2grid_amd import _capi
    g = syn.grid_118()
    rows = 130
    _say(f"limit violations 118 rows={rows} K={K}")
    rng = np.random.default_rng(31)
    nb = np.flatnonzero(~syn.bridges(g.n_bus, g.branch_from, g.branch_to))
    trips = [sorted(rng.choice(nb, size=r % 3, replace=False).tolist()) for r in range(rows)]
    p, q = syn.injection_scenarios(g, rows, seed=32)
    S = (p + 1j * q) / g.sn_mva
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, rows, precision=_capi.FP64, per_row_ybus=True)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.set_ybus_edits(*syn.outage_edits(g, trips))
    B.solve(max_iter=10, tol=1e-8, scheduling=_capi.CONTINUOUS)
    st, _, res = B.results()
    V = B.get_v()
    table = _table(g, V[0], seed=33)
    tol = float(np.quantile(res, 0.8))  # a fifth of the rows get a DIVERGENCE record
    out = B.limit_violations(table, trips=trips, tol=tol, rel_tol=1e-9, capacity=K)
    want = orc.limit_violations(V, res, st, table, trips=trips, tol=tol, rel_tol=1e-9, K=K)
    _compare(out, want, K)
    n_div = sum(1 for w in want if w["records"] and w["records"][0][3] == 4)
    assert 0 < n_div < rows
    if K == 2:
        assert out["truncated"].sum() > 10  # the radix-select path ran
    # fewer active rows: only those are checked
    B.set_active_rows(70)
    out2 = B.limit_violations(table, trips=trips[:70], tol=tol, capacity=K)
    _compare(out2, want[:70], K)


@pytest.mark.gpu
def test_capacity_and_margin_arguments():
    from gpusim2grid_amd import _capi
    g = syn.grid_118()
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, 4, precision=_capi.FP64)
    B.set_sbus(g.Sbus)
    B.set_v(g.flat_start())
    B.solve(max_iter=10, tol=1e-8)
    table = _table(g, B.get_v()[0], seed=1)
    with pytest.raises(ValueError, match="capacity"):
        B.limit_violations(table, capacity=0)
    with pytest.raises(ValueError, match="capacity"):
        B.limit_violations(table, capacity=_capi.MAX_VIOLATION_CAPACITY + 1)
    with pytest.raises(ValueError, match="rel_tol"):
        B.limit_violations(table, rel_tol=1.0)
    with pytest.raises(ValueError, match="tripped"):
        B.limit_violations(table, trips=[[g.n_branch], [], [], []])
    out = B.limit_violations(table, capacity=_capi.MAX_VIOLATION_CAPACITY)
    assert (out["count"] >= 0).all()


@pytest.mark.gpu
def test_contingency_analysis_get_violations():
    from gpusim2grid_amd import ContingencyAnalysisGPU, LimitViolation, LimitViolationType, ViolationElementType
    g = syn.grid_c1()
    _say("ContingencyAnalysisGPU c1 compute_limit_violations")
    br = syn.bridges(g.n_bus, g.branch_from, g.branch_to)
    ok = np.flatnonzero(~br)[:60]
    bad = np.flatnonzero(br)[:2]
    ctg = [[int(l)] for l in ok[:30]] + [[int(l)] for l in bad] + [[int(a), int(b)] for a, b in zip(ok[30::2],
                                                                                                     ok[31::2])]
    ca = ContingencyAnalysisGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=10, scheduling="continuous",
                                compute_limit_violations=True)
    with pytest.raises(RuntimeError):
        ca.set_limits(np.zeros(g.n_bus), np.zeros(g.n_bus), np.zeros(g.n_branch), np.zeros(g.n_branch), 0)
    Vb = ca.solver.get_v_base()
    table = _table(g, Vb, seed=41)
    vn = table["vn_kv"]
    ca.set_branch_data(g.branch_from, g.branch_to, g.yff, g.yft, g.ytf, g.ytt, bus_vn_kv=vn, sn_mva=g.sn_mva)
    ca.set_limits(table["vmin_kv"], table["vmax_kv"], table["limit1_ka"], table["limit2_ka"], table["n_lines"])
    ca.violation_capacity = 5
    ca.add_contingencies_by_branch_id(ctg)
    V = ca.compute(batch_size=16, as_numpy=True)
    st = ca.get_row_status()
    res = ca.last_residuals()
    viol = ca.get_violations()
    want = orc.limit_violations(V, res, st, table, trips=ctg, tol=ca.violation_tol,
                                rel_tol=ca.violation_rel_tol, K=5)
    assert len(viol) == len(ctg)
    for r, (v, w) in enumerate(zip(viol, want)):
        if w["count"] < 0:
            assert r in (30, 31)
            assert v == [LimitViolation(ViolationElementType.GRID, -1, 0, LimitViolationType.NOT_SIMULATED,
                                        v[0].value, v[0].limit)]
            continue
        assert [(int(x.element_type), x.element_id, x.side, int(x.violation_type)) for x in v] == \
            [x[:4] for x in w["records"]], r
        np.testing.assert_allclose([x.value for x in v], [x[4] for x in w["records"]], rtol=1e-15)
    counts = ca.get_violation_counts()
    np.testing.assert_array_equal(counts["current"], [w["n_current"] for w in want])
    np.testing.assert_array_equal(counts["low_voltage"], [w["n_low"] for w in want])
    np.testing.assert_array_equal(ca.get_violations_truncated(), [bool(w["truncated"]) for w in want])
    assert any(len(v) > 1 for v in viol)
    # switched off: no records
    ca.compute_limit_violations = False
    ca.compute(batch_size=16)
    with pytest.raises(RuntimeError):
        ca.get_violations()
