"""Islanded buses (handle_disconnected_grid): rows whose trips split the grid
solve on the largest connected component, the other buses frozen as identity
rows of J and reported NaN -- the reference's compute_component_masks,
apply_bus_mask_kernel and mask_V_nan_kernel
(src/_cpp/contingency_analysis_helper.cpp:657-692, src/_cpp/acpf_nr_kernels.cu:219-307),
modelled on its tests/python/test_contingency_batch.py islanding cases."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tuple(g):
    return (g.Ybus, g.flat_start(), g.Sbus, np.array([g.slack]), np.array([1.0]), g.pv, g.pq)


def _off_main(g, ids):
    """Buses outside the largest component once `ids` are tripped (a tie goes
    to the component met first in bus order)."""
    keep = np.ones(g.n_branch, dtype=bool)
    keep[list(ids)] = False
    A = sp.coo_matrix((np.ones(int(keep.sum())), (g.branch_from[keep], g.branch_to[keep])),
                      shape=(g.n_bus, g.n_bus))
    nc, lab = sp.csgraph.connected_components(A, directed=False)
    if nc == 1:
        return np.zeros(0, np.int64)
    first = {}
    for i, c in enumerate(lab):
        first.setdefault(c, len(first))
    order = np.array([first[c] for c in lab])
    cnt = np.bincount(order)
    return np.flatnonzero(order != int(np.argmax(cnt)))


def _islanding_trips(g, limit):
    """Single bridge trips that strand a set of buses without the slack."""
    br = np.flatnonzero(syn.bridges(g.n_bus, g.branch_from, g.branch_to))
    out = []
    for l in br:
        off = _off_main(g, [int(l)])
        if off.size and g.slack not in off:
            out.append(([int(l)], off))
        if len(out) == limit:
            break
    return out


def _disjoint_union(g):
    """g plus a copy of g's branches on n more buses, all of them PQ (no slack)."""
    n = g.n_bus
    f = np.concatenate([g.branch_from, g.branch_from + n])
    t = np.concatenate([g.branch_to, g.branch_to + n])
    cat = lambda a: np.concatenate([a, a])  # noqa: E731
    Y = syn.build_ybus(2 * n, f, t, cat(g.yff), cat(g.yft), cat(g.ytf), cat(g.ytt))
    pq = np.concatenate([g.pq, np.arange(n, 2 * n)]).astype(np.int32)
    return Y, pq


# ---- CPU: the masked oracle -------------------------------------------------
def test_masked_oracle_solves_the_kept_component_alone():
    """A grid joined by a disjoint copy of itself (all PQ, so it has no slack):
    masking the copy must give the grid's own solution, status and iteration
    count exactly, and NaN on the copy -- a wrong bus index in the
    restriction, or a masked bus left in the residual, fails this."""
    g = syn.grid_118()
    Y2, pq2 = _disjoint_union(g)
    n = g.n_bus
    S = syn.injection_scenarios(g, 3, seed=5)
    S = (S[0] + 1j * S[1]) / g.sn_mva
    S2 = np.concatenate([S, S], axis=1)
    V2 = np.concatenate([g.flat_start(), g.flat_start()])
    Vm, stm, itm, rm = orc.nr_masked(Y2, V2, S2, g.pv, pq2, [np.arange(n, 2 * n)] * 3)
    Vo, sto, ito, ro = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8)
    np.testing.assert_array_equal(stm, sto)
    np.testing.assert_array_equal(itm, ito)
    np.testing.assert_allclose(Vm[:, :n], Vo, rtol=0, atol=1e-13)
    np.testing.assert_allclose(rm, ro, rtol=1e-6, atol=1e-15)
    assert np.isnan(Vm[:, n:]).all()
    # an empty mask is the plain solve
    Ve, ste, _, _ = orc.nr_masked(g.Ybus, g.flat_start(), S, g.pv, g.pq, [[], [], []])
    np.testing.assert_array_equal(Ve, Vo)


def test_off_main_component_rule():
    g = syn.grid_118()
    trips = _islanding_trips(g, 3)
    assert trips, "the 118-bus grid has islanding bridges"
    for ids, off in trips:
        assert 0 < off.size < g.n_bus // 2
    # tripping every branch of the slack strands it as a one-bus component
    ids = np.flatnonzero((g.branch_from == g.slack) | (g.branch_to == g.slack))
    assert g.slack in _off_main(g, ids)


# ---- GPU ----------------------------------------------------------------------
def _rows_with_islands(g, rows, seed):
    """Per row: trips (every third an islanding bridge, others a non-bridge
    pair or nothing), the masked buses, Sbus and a finite start V."""
    rng = np.random.default_rng(seed)
    isl = _islanding_trips(g, 6)
    nb = np.flatnonzero(~syn.bridges(g.n_bus, g.branch_from, g.branch_to))
    trips, masks = [], []
    for r in range(rows):
        if r % 3 == 0:
            ids, off = isl[r // 3 % len(isl)]
            trips.append(ids)
            masks.append(off)
        else:
            ids = sorted(rng.choice(nb, size=r % 2, replace=False).tolist())
            trips.append(ids)
            masks.append(np.zeros(0, np.int64))
    p, q = syn.injection_scenarios(g, rows, seed=seed)
    return trips, masks, (p + 1j * q) / g.sn_mva


def _csr_of(masks):
    ptr = np.zeros(len(masks) + 1, np.int32)
    ptr[1:] = np.cumsum([len(m) for m in masks])
    return ptr, np.concatenate([np.asarray(m, np.int32) for m in masks]) if ptr[-1] else np.zeros(0, np.int32)


def _yx_rows(g, trips):
    Y = sp.csr_matrix(g.Ybus)
    Y.sort_indices()
    memo = {}
    for t in trips:
        if tuple(t) not in memo:
            memo[tuple(t)] = orc.ybus_values_on_pattern(Y, orc.outage_ybus(g, t))
    return np.array([memo[tuple(t)] for t in trips])


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_masked_rows_match_oracle(prec):
    from gpusim2grid_amd import _capi
    P = {"fp64": _capi.FP64, "mixed": _capi.MIXED}[prec]
    g = syn.grid_118()
    rows = 150
    _say(f"masked rows 118 prec={prec} rows={rows}")
    trips, masks, S = _rows_with_islands(g, rows, seed=21)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, rows, precision=P, per_row_ybus=True)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.set_ybus_edits(*syn.outage_edits(g, trips))
    B.set_bus_mask(*_csr_of(masks))
    B.solve(max_iter=10, tol=1e-8, scheduling=_capi.CONTINUOUS)
    st, it, res = B.results()
    V = B.get_v()
    Vo, sto, ito, ro = orc.nr_masked(g.Ybus, g.flat_start(), S, g.pv, g.pq, masks, 10, 1e-8,
                                     Yx_rows=_yx_rows(g, trips))
    np.testing.assert_array_equal(st, sto)
    assert (st == 0).all()
    live = ~np.isnan(Vo)
    np.testing.assert_array_equal(np.isnan(V), ~live)
    if P == _capi.FP64:
        np.testing.assert_array_equal(it, ito)
        assert np.abs(V[live] - Vo[live]).max() < 1e-9
    else:
        assert (it <= ito + 2).all() and np.abs(V[live] - Vo[live]).max() < 1e-6
    # the residual leaves the masked equations out
    assert (res < 1e-8).all()
    # clearing the mask: the islanding rows no longer solve (an island with no slack)
    B.set_bus_mask(None)
    B.set_v(g.flat_start())
    B.solve(max_iter=10, tol=1e-8, scheduling=_capi.CONTINUOUS)
    st2, _, _ = B.results()
    assert (st2[0::3] != 0).all() and (st2[1::3] == 0).all()


@pytest.mark.gpu
def test_masked_rows_through_tail_compaction(monkeypatch):
    """Most rows start solved, so the few still iterating are compacted into
    dense tiles; their masks travel with them."""
    monkeypatch.setenv("G2G_COMPACT", "1")  # compact even this small a batch
    from gpusim2grid_amd import _capi
    g = syn.grid_118()
    rows = 900
    _say(f"masked rows, tail compaction, rows={rows}")
    trips, masks, S = _rows_with_islands(g, rows, seed=23)
    yx = _yx_rows(g, trips)
    Vs, _, _, _ = orc.nr_masked(g.Ybus, g.flat_start(), S, g.pv, g.pq, masks, 10, 1e-12, Yx_rows=yx)
    V0 = np.where(np.isnan(Vs), g.flat_start()[None, :], Vs)
    rng = np.random.default_rng(24)
    hard = rng.choice(rows, size=80, replace=False)
    V0[hard] = g.flat_start()
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, rows, precision=_capi.FP64, per_row_ybus=True)
    B.set_sbus(S)
    B.set_v(V0)
    B.set_ybus_edits(*syn.outage_edits(g, trips))
    B.set_bus_mask(*_csr_of(masks))
    B.solve(max_iter=10, tol=1e-8, scheduling=_capi.CONTINUOUS)
    st, it, _ = B.results()
    V = B.get_v()
    Vo, sto, ito, _ = orc.nr_masked(g.Ybus, V0, S, g.pv, g.pq, masks, 10, 1e-8, Yx_rows=yx)
    np.testing.assert_array_equal(st, sto)
    np.testing.assert_array_equal(it, ito)
    live = ~np.isnan(Vo)
    np.testing.assert_array_equal(np.isnan(V), ~live)
    assert np.abs(V[live] - Vo[live]).max() < 1e-8
    assert (it[hard] > 0).all() and (np.delete(it, hard) == 0).all()


@pytest.mark.gpu
def test_bus_mask_argument_checks():
    from gpusim2grid_amd import _capi
    g = syn.grid_118()
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, 4, precision=_capi.FP64)
    with pytest.raises(Exception, match="slack"):
        B.set_bus_mask([0, 1, 1, 1, 1], [int(np.ravel(g.slack)[0])])
    with pytest.raises(Exception, match="range"):
        B.set_bus_mask([0, 1, 1, 1, 1], [g.n_bus])
    B.set_bus_mask([0, 0, 0, 0, 0], [])
    B.set_bus_mask(None)
    B.stream_begin()
    with pytest.raises(Exception, match="stream"):
        B.set_bus_mask([0, 1, 1, 1, 1], [1])


@pytest.mark.gpu
def test_contingency_analysis_handle_disconnected_grid():
    from gpusim2grid_amd import ContingencyAnalysisGPU, RowStatus
    g = syn.grid_118()
    _say("ContingencyAnalysisGPU 118 handle_disconnected_grid")
    isl = _islanding_trips(g, 4)
    nb = np.flatnonzero(~syn.bridges(g.n_bus, g.branch_from, g.branch_to))
    slack_cut = np.flatnonzero((g.branch_from == g.slack) | (g.branch_to == g.slack)).tolist()
    ctg = [ids for ids, _ in isl] + [[int(nb[0])], slack_cut, [int(nb[1]), int(nb[2])]]
    ca = ContingencyAnalysisGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=10, scheduling="continuous",
                                handle_disconnected_grid=True)
    ca.set_branch_data(g.branch_from, g.branch_to, g.yff, g.yft, g.ytf, g.ytt)
    ca.add_contingencies_by_branch_id(ctg)
    V = ca.compute(batch_size=16, as_numpy=True)
    st = ca.get_row_status()
    it = ca.get_row_iterations()
    disc = ca.get_disconnected()
    Vb = ca.solver.get_v_base()
    k = len(isl)
    np.testing.assert_array_equal(disc, [False] * (k + 1) + [True, False])
    assert st[k + 1] == int(RowStatus.NotSimulated) and np.isnan(V[k + 1]).all()
    masks = [off for _, off in isl] + [[], [], []]
    rows = [r for r in range(len(ctg)) if r != k + 1]
    Vo, sto, ito, _ = orc.nr_masked(g.Ybus, Vb, g.Sbus, g.pv, g.pq, [masks[r] for r in rows], 10, 1e-8,
                                    Yx_rows=_yx_rows(g, [ctg[r] for r in rows]))
    for q, r in enumerate(rows):
        assert st[r] == sto[q] and it[r] == ito[q], r
        live = ~np.isnan(Vo[q])
        np.testing.assert_array_equal(np.isnan(V[r]), ~live)
        assert np.abs(V[r][live] - Vo[q][live]).max() < 1e-9, r
    # switched off: the islanding rows are skipped again
    ca.handle_disconnected_grid = False
    ca.compute(batch_size=16)
    d = ca.get_disconnected()
    assert d[:k].all() and not d[k] and d[k + 1] and not d[k + 2]
