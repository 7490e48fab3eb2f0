"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _perturbed(g, seed=0, rows=1):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(rows):
        out.append(g.flat_start() * np.exp(1j * rng.normal(0, 0.05, g.n_bus)) * (1 + rng.normal(0, 0.02, g.n_bus)))
    return np.array(out)


def _sbus_rows(g, rows, seed=1):
    p, q = syn.injection_scenarios(g, rows, seed=seed)
    return (p + 1j * q) / g.sn_mva


def _gpu_solve(g, S, V0, rows, precision=_capi.FP64, max_iter=10, tol=1e-8, sched=_capi.CONTINUOUS,
               edits=None):
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(P, rows, precision=precision, per_row_ybus=edits is not None)
    B.set_sbus(S)
    B.set_v(V0)
    if edits is not None:
        B.set_ybus_edits(*edits)
    stats = B.solve(max_iter=max_iter, tol=tol, scheduling=sched)
    st, it, res = B.results()
    return B.get_v(), st, it, res, stats


@pytest.mark.parametrize("name,prec", [("c0", _capi.FP64), ("118", _capi.FP64), ("c1", _capi.FP64),
                                       ("118", _capi.MIXED)])
def test_mismatch_and_jacobian_kernel(name, prec):
    g = {"c0": syn.grid_c0, "118": syn.grid_118, "c1": syn.grid_c1}[name]()
    rows = 70  # two tiles, the second partly filled
    _say(f"eval {name} prec={prec} rows={rows}")
    V = _perturbed(g, 0, rows)
    S = _sbus_rows(g, rows)
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(P, rows, precision=prec)
    B.set_v(V)
    B.set_sbus(S)
    F, J, r = B.eval(with_j=True)
    for k in (0, 37, 69):
        Fo, Jo, ro = orc.f_and_j(g.Ybus, V[k], S[k], g.pv, g.pq)
        if prec == _capi.FP64:
            np.testing.assert_allclose(F[k], Fo, rtol=0, atol=1e-12)
            np.testing.assert_allclose(J[k], Jo.data, rtol=1e-13, atol=1e-11)
        else:
            # F stored in FP32 in the factorization stream; J in FP32
            np.testing.assert_allclose(F[k], Fo, rtol=1e-6, atol=1e-6)
            np.testing.assert_allclose(J[k], Jo.data, rtol=1e-6, atol=1e-5)
        assert r[k] == pytest.approx(ro, rel=1e-13)  # the residual itself is FP64 in both


def test_single_solve_c0_matches_oracle():
    g = syn.grid_c0()
    _say("c0 single solve")
    V, st, it, res, stats = _gpu_solve(g, g.Sbus, g.flat_start(), 1)
    Vo, sto, ito, reso = orc.nr_batch(g.Ybus, g.flat_start(), g.Sbus, g.pv, g.pq)
    assert st[0] == sto[0] == 0
    assert it[0] == ito[0]
    assert np.abs(V[0] - Vo[0]).max() < 1e-8
    assert res[0] < 1e-8


@pytest.mark.parametrize("name,rows", [("118", 200), ("c1", 130), ("c2", 96)])
def test_injection_sweep_matches_oracle(name, rows):
    g = {"118": syn.grid_118, "c1": syn.grid_c1, "c2": syn.grid_c2}[name]()
    _say(f"sweep {name} rows={rows}")
    S = _sbus_rows(g, rows)
    V, st, it, res, stats = _gpu_solve(g, S, g.flat_start(), rows)
    Vo, sto, ito, reso = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, n_threads=ORACLE_THREADS)
    np.testing.assert_array_equal(st, sto)
    np.testing.assert_array_equal(it, ito)
    conv = sto == 0
    assert conv.mean() > 0.9
    assert np.abs(V[conv] - Vo[conv]).max() < 1e-8
    assert stats["n_converged"] == conv.sum()


def test_status_rules_nan_row_and_budget():
    g = syn.grid_118()
    rows = 66
    S = _sbus_rows(g, rows)
    S[3] = np.nan  # a non-finite row diverges (test_injection_batch.py:447)
    V, st, it, res, _ = _gpu_solve(g, S, g.flat_start(), rows, max_iter=10)
    assert st[3] == 2 and np.isnan(res[3]) and it[3] == 0
    assert (np.delete(st, 3) == 0).all()
    # a budget of 1 leaves every row at MaxIter (test_continuous_batching.py:316-332)
    V1, st1, it1, res1, _ = _gpu_solve(g, np.delete(S, 3, axis=0), g.flat_start(), rows - 1, max_iter=1)
    assert (st1 == 1).all() and (it1 == 1).all() and (res1 >= 1e-8).all()
    # already converged at V0: zero updates
    V2, st2, it2, _, _ = _gpu_solve(g, np.delete(S, 3, axis=0)[:5], np.delete(V, 3, axis=0)[:5], 5)
    assert (st2 == 0).all() and (it2 == 0).all()


def test_chunked_schedule_runs_the_fixed_count():
    g = syn.grid_118()
    rows = 64
    S = _sbus_rows(g, rows)
    V, st, it, res, _ = _gpu_solve(g, S, g.flat_start(), rows, max_iter=6, sched=_capi.CHUNKED)
    Vc, stc, itc, resc, _ = _gpu_solve(g, S, g.flat_start(), rows, max_iter=10, sched=_capi.CONTINUOUS)
    assert (it == 6).all() and (st == 0).all()
    assert np.abs(V - Vc).max() < 1e-9  # extra Newton steps past convergence move V by rounding only


def test_batch_size_invariance():
    """Rows are independent lanes: the same row gives bit-identical results in
    any batch (test_contingency_batch.py:174-222)."""
    g = syn.grid_c1()
    S = _sbus_rows(g, 100)
    Va, sta, ita, resa, _ = _gpu_solve(g, S, g.flat_start(), 100)
    Vb, stb, itb, resb, _ = _gpu_solve(g, S[37:41], g.flat_start(), 4)
    np.testing.assert_array_equal(Va[37:41], Vb)
    np.testing.assert_array_equal(ita[37:41], itb)


def _outage_edits(g, rows_branches):
    Y = sp.csr_matrix(g.Ybus)
    Y.sort_indices()

    def k_of(i, j):
        lo, hi = Y.indptr[i], Y.indptr[i + 1]
        return lo + int(np.searchsorted(Y.indices[lo:hi], j))

    row_ptr, ent, delta = [0], [], []
    for brs in rows_branches:
        for l in brs:
            f, t = int(g.branch_from[l]), int(g.branch_to[l])
            for (i, j, v) in ((f, f, g.yff[l]), (t, t, g.ytt[l]), (f, t, g.yft[l]), (t, f, g.ytf[l])):
                ent.append(k_of(i, j))
                delta.append(v)
        row_ptr.append(len(ent))
    return np.array(row_ptr), np.array(ent), np.array(delta)


def test_n1_outages_match_oracle():
    g = syn.grid_c1()
    bridge = syn.bridges(g.n_bus, g.branch_from, g.branch_to)
    cand = np.flatnonzero(~bridge)[:130]
    rows = cand.size
    _say(f"N-1 c1 rows={rows}")
    edits = _outage_edits(g, [[l] for l in cand])
    V, st, it, res, _ = _gpu_solve(g, g.Sbus, g.flat_start(), rows, edits=edits)
    Yrows = np.stack([orc.ybus_values_on_pattern(g.Ybus, orc.outage_ybus(g, [l])) for l in cand])
    Vo, sto, ito, reso = orc.nr_batch(g.Ybus, g.flat_start(), g.Sbus, g.pv, g.pq, Yx_rows=Yrows,
                                      n_threads=ORACLE_THREADS)
    np.testing.assert_array_equal(st, sto)
    np.testing.assert_array_equal(it, ito)
    conv = sto == 0
    assert np.abs(V[conv] - Vo[conv]).max() < 1e-8


def test_mixed_precision_reaches_fp64_tolerance():
    g = syn.grid_c2()
    rows = 96
    _say(f"mixed c2 rows={rows}")
    S = _sbus_rows(g, rows)
    V, st, it, res, _ = _gpu_solve(g, S, g.flat_start(), rows, precision=_capi.MIXED)
    Vo, sto, ito, reso = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, n_threads=ORACLE_THREADS)
    ok = (sto == 0)
    assert (st[ok] == 0).all()
    assert (res[ok] < 1e-8).all()
    assert np.abs(V[ok] - Vo[ok]).max() < 1e-6
    assert (it[ok] <= ito[ok] + 2).all()


@pytest.mark.parametrize("prec", [_capi.FP64, _capi.MIXED])
def test_raw_lu_solve_matches_scipy(prec):
    g = syn.grid_118()
    V = _perturbed(g, 3)[0]
    F, J, _ = orc.f_and_j(g.Ybus, V, g.Sbus, g.pv, g.pq)
    P = _capi.Plan.from_matrix(J)
    rows = 20
    vals = np.stack([J.data * (1 + 0.01 * k) for k in range(rows)])
    rhs = np.stack([F * (k + 1) for k in range(rows)])
    x, cl = P.lu_solve(vals, rhs, precision=prec)
    tol = 1e-10 if prec == _capi.FP64 else 1e-3
    for k in range(rows):
        Jk = sp.csr_matrix((vals[k], J.indices, J.indptr), shape=J.shape)
        r = np.abs(Jk @ x[k] - rhs[k]).max() / np.abs(rhs[k]).max()
        assert r < tol, (k, r)
    assert (cl == 0).all()


@pytest.mark.parametrize("prec", [_capi.FP64, _capi.MIXED])
def test_tail_compaction_keeps_every_row(prec, monkeypatch):
    """A batch whose rows mostly start converged: after the first update few
    rows are left, they are compacted into dense tiles and scattered back;
    statuses, iteration counts and voltages are the oracle's, row by row."""
    monkeypatch.setenv("G2G_COMPACT", "1")  # compact even this small a batch
    g = syn.grid_118()
    rows = 1000
    _say(f"tail compaction prec={prec} rows={rows}")
    S = _sbus_rows(g, rows, seed=11)
    Vs, _, _, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-12, n_threads=ORACLE_THREADS)
    V0 = Vs.copy()
    rng = np.random.default_rng(12)
    hard = rng.choice(rows, size=90, replace=False)
    V0[hard] = g.flat_start()
    V0[hard[:30]] *= np.exp(1j * rng.normal(0, 0.02, (30, g.n_bus)))
    V, st, it, res, _ = _gpu_solve(g, S, V0, rows, precision=prec)
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, V0, S, g.pv, g.pq, 10, 1e-8, n_threads=ORACLE_THREADS)
    np.testing.assert_array_equal(st, sto)
    if prec == _capi.FP64:
        np.testing.assert_array_equal(it, ito)
        assert np.abs(V - Vo).max() < 1e-8
    else:
        assert (it <= ito + 2).all() and np.abs(V - Vo).max() < 1e-6
    assert (it[hard] > 0).all() and (np.delete(it, hard) == 0).all()
