"""Learned warm start with flat-start fallback, through the C ABI, against
the CPU oracle's flat-start solutions."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn
from gpusim2grid_amd.warmstart import WarmStart
from oracle import mlp_ref as M
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sbus(g, rows, seed):
    p, q = syn.injection_scenarios(g, rows, seed=seed)
    return (p + 1j * q) / g.sn_mva


@pytest.fixture(scope="module")
def trained():
    import torch
    g = syn.grid_118()
    rows = 4096
    _say("warm: solving the training set")
    S = _sbus(g, rows, seed=2)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, rows)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.solve(max_iter=10, tol=1e-8)
    st, _, _ = B.results()
    V = B.get_v()
    ok = st == 0
    ws = WarmStart(g.n_bus, g.pv, g.pq, g.slack, np.abs(g.flat_start()), n_hidden=256, seed=0)
    _say("warm: training")
    hist = ws.train(torch.as_tensor(S[ok], device="cuda"), torch.as_tensor(V[ok], device="cuda"), epochs=40,
                    batch=256, lr=1e-3)
    return g, ws, hist


def test_training_converges(trained):
    _, _, hist = trained
    assert hist[-1] < 0.05 * hist[0], hist


def test_warm_start_kernel_matches_oracle_forward(trained):
    g, ws, _ = trained
    rows = 100
    S = _sbus(g, rows, seed=7)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, rows)
    B.set_sbus(S)
    B.set_v(g.flat_start())  # fixed buses keep this start V
    B.warm_start(ws.mlp)
    Vw = B.get_v()
    p = {k: v.astype(np.float64) for k, v in ws.mlp.get_params().items()}
    X = M.features(S, ws.in_mu.astype(np.float64), ws.in_sigma.astype(np.float64))
    Y, _ = M.forward(p, X, bf16_ops=True)
    Vref = M.warm_v(Y, ws.out_mu.astype(np.float64), ws.out_sigma.astype(np.float64), ws.kind, ws.vset)
    # bf16 products: a few 1e-3 of the output spread
    tol = 0.02 * float(np.max(ws.out_sigma)) + 1e-6
    assert np.max(np.abs(Vw - Vref)) < tol
    # fixed magnitudes and the slack angle are exact
    fixed = ws.kind != 1
    np.testing.assert_allclose(np.abs(Vw[:, fixed]), np.broadcast_to(ws.vset[fixed], (rows, fixed.sum())),
                               rtol=1e-14)
    assert np.all(np.angle(Vw[:, g.slack]) == 0.0)
    # the host prediction path agrees with the device one
    Vh = ws.predict_v(S)
    assert np.max(np.abs(Vh - Vw)) < 1e-5


def test_streamed_slots_take_the_device_warm_start(trained):
    """Slots filled in a streamed batch get, from g2g_batch_stream_warm, the
    start voltages the whole-batch warm start gives the same rows; the other
    slots keep their V."""
    g, ws, _ = trained
    rows, slots = 70, 256
    _say("stream warm")
    S = _sbus(g, rows, seed=12)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, rows, precision=_capi.MIXED)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.warm_start(ws.mlp)
    Vw = B.get_v()
    Bs = _capi.Batch(plan, slots, precision=_capi.MIXED, per_row_ybus=True)
    Bs.stream_begin()
    sl = list(range(3, 3 + 3 * rows, 3))
    other = [5, 200]
    Bs.stream_fill(other, S[:2], g.flat_start()[None], edits=syn.outage_edits(g, [[], []]))
    Bs.stream_fill(sl, S, g.flat_start()[None], edits=syn.outage_edits(g, [[] for _ in range(rows)]))
    Bs.stream_warm(ws.mlp, sl)
    V = Bs.stream_take(sl)
    np.testing.assert_allclose(V, Vw, rtol=0, atol=1e-12)
    Vo = Bs.stream_take(other)
    np.testing.assert_array_equal(Vo, np.broadcast_to(g.flat_start(), Vo.shape))


@pytest.mark.parametrize("prec", [_capi.FP64, _capi.MIXED])
def test_warm_solve_matches_flat_start_oracle(trained, prec):
    g, ws, _ = trained
    rows = 1000
    _say(f"warm solve prec={prec}")
    S = _sbus(g, rows, seed=1)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, rows, precision=prec)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    stats = B.solve_warm(ws.mlp, k_warm=4, growth=2.0, max_iter=10, tol=1e-8)
    st, it, res = B.results()
    V = B.get_v()
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8, n_threads=16)
    np.testing.assert_array_equal(st, sto)
    conv = st == 0
    assert conv.all()
    # every converged row satisfies the test, on the oracle's own mismatch
    F = max(np.max(np.abs(orc.mismatch_np(g.Ybus, V[r], S[r], g.pv, g.pq))) for r in range(rows))
    assert F < 1e-8
    assert np.max(res) < 1e-8
    # same solution as the flat start (to the conditioning of the test)
    assert np.max(np.abs(V - Vo)) < 1e-6
    # fewer Newton updates than from the flat start
    assert it.mean() < ito.mean(), (it.mean(), ito.mean())
    assert stats["n_converged"] == rows


@pytest.mark.parametrize("start,compact", [("flat", None), ("flat", "1"), ("per_row", None), ("per_row", "1"),
                                           ("tiled", None)])
def test_untrained_mlp_falls_back_to_flat_start(start, compact, monkeypatch):
    """Rows whose learned start fails restart from the batch's V: one start
    for every row, or a start per row; with compaction forced, the restarts
    happen in compacted batches too (compacted again as rows finish)."""
    if compact is not None:
        monkeypatch.setenv("G2G_COMPACT", compact)
    g = syn.grid_118()
    rows = 300
    _say(f"warm fallback (untrained) start={start} compact={compact}")
    S = _sbus(g, rows, seed=3)
    ws = WarmStart(g.n_bus, g.pv, g.pq, g.slack, np.abs(g.flat_start()), n_hidden=64, seed=5)
    n2 = 2 * g.n_bus
    # normalisation that turns the random network's outputs into a poor start:
    # angles spread by one radian
    ws.in_mu = np.zeros(n2, np.float32)
    ws.in_sigma = np.ones(n2, np.float32)
    ws.out_mu = np.tile(np.array([0.0, 1.0], np.float32), g.n_bus)
    ws.out_sigma = np.tile(np.array([1.0, 0.05], np.float32), g.n_bus)
    ws.mlp.set_normalization(ws.in_mu, ws.in_sigma, ws.out_mu, ws.out_sigma)
    V0 = g.flat_start()  # one row: the batch keeps it as its fallback start
    if start == "tiled":  # the same start written row by row: found uniform on the device
        V0 = np.tile(V0, (rows, 1))
    if start == "per_row":  # angles perturbed row by row
        rng = np.random.default_rng(6)
        V0 = V0[None, :] * np.exp(1j * rng.normal(0.0, 0.01, (rows, g.n_bus)))
        V0[:, g.slack] = g.flat_start()[g.slack]
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, rows)
    B.set_sbus(S)
    B.set_v(V0)
    stats = B.solve_warm(ws.mlp, k_warm=2, growth=2.0, max_iter=10, tol=1e-8)
    st, it, _ = B.results()
    V = B.get_v()
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, V0, S, g.pv, g.pq, 10, 1e-8, n_threads=16)
    np.testing.assert_array_equal(st, sto)
    assert stats["n_requeued"] > 0
    # re-queued rows restart from the batch's V: their second pass is the
    # oracle's run from that start, iteration for iteration, on top of the
    # warm pass
    Bf = _capi.Batch(plan, rows)
    Bf.set_sbus(S)
    Bf.set_v(V0)
    Bf.solve(max_iter=10, tol=1e-8)
    stf, itf, _ = Bf.results()
    Vf = Bf.get_v()
    np.testing.assert_array_equal(itf, ito)
    re = it > 2
    assert re.any()
    np.testing.assert_array_equal(it[re] - itf[re] <= 2, True)
    assert np.max(np.abs(V[re] - Vf[re])) < 1e-12


def test_contraction_rule_restarts_rows_at_their_second_update():
    """With the quadratic-contraction rejection at its tightest (q tiny), the
    growth rule and the update budget out of the way, every row that has not
    converged after two updates on its learned start restarts right then:
    its count is 2 + the flat-start run's, and it ends where the flat start
    ends.  Off (q = 0), no row restarts before its budget."""
    g = syn.grid_118()
    rows = 300
    _say("warm contraction")
    S = _sbus(g, rows, seed=3)
    ws = WarmStart(g.n_bus, g.pv, g.pq, g.slack, np.abs(g.flat_start()), n_hidden=64, seed=5)
    n2 = 2 * g.n_bus
    ws.in_mu = np.zeros(n2, np.float32)
    ws.in_sigma = np.ones(n2, np.float32)
    ws.out_mu = np.tile(np.array([0.0, 1.0], np.float32), g.n_bus)
    ws.out_sigma = np.tile(np.array([0.02, 0.002], np.float32), g.n_bus)  # a fair start
    ws.mlp.set_normalization(ws.in_mu, ws.in_sigma, ws.out_mu, ws.out_sigma)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    Bf = _capi.Batch(plan, rows)
    Bf.set_sbus(S)
    Bf.set_v(g.flat_start())
    Bf.solve(max_iter=10, tol=1e-8)
    stf, itf, _ = Bf.results()
    Vf = Bf.get_v()
    assert (stf == 0).all()
    out = {}
    for q in (0.0, 1e-300):
        B = _capi.Batch(plan, rows)
        B.set_sbus(S)
        B.set_v(g.flat_start())
        B.set_warm_contraction(q)
        stats = B.solve_warm(ws.mlp, k_warm=8, growth=1e300, max_iter=10, tol=1e-8)
        st, it, _ = B.results()
        out[q] = (stats, st, it, B.get_v())
    stats, st, it, V = out[1e-300]
    assert (st == 0).all()
    re = it > 2
    assert re.sum() == stats["n_requeued"] > 0
    np.testing.assert_array_equal(it[re], 2 + itf[re])
    assert np.max(np.abs(V[re] - Vf[re])) < 1e-12
    stats0, st0, it0, _ = out[0.0]
    assert (st0 == 0).all()
    # off: the rows that converge on their learned start within the budget
    # never restart, so fewer rows restart than with the rule at its tightest
    assert stats0["n_requeued"] < stats["n_requeued"]
    assert (it0[~re] == it[~re]).all()


def test_warm_start_with_outages_gains_over_flat_start():
    """The learned-warm-start workload in small: the 2,000-bus grid, load
    scaling plus 0-2 random non-islanding outages per row, the outage mask
    among the inputs.  Every row ends as the flat-start oracle's does, and the
    trained network saves Newton updates over the flat start."""
    import torch
    g = syn.grid_c1()
    n_train, rows = 8192, 512
    _say("warm outages: solving the training set")
    S = _sbus(g, n_train, seed=2)
    outs = syn.random_outages(g, n_train, seed=4)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(plan, n_train, precision=_capi.MIXED, per_row_ybus=True)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.set_ybus_edits(*syn.outage_edits(g, outs))
    B.solve(max_iter=10, tol=1e-8)
    st, _, _ = B.results()
    V = B.get_v()
    del B
    ok = st == 0
    mask = syn.outage_mask(g, outs)
    ws = WarmStart(g.n_bus, g.pv, g.pq, g.slack, np.abs(g.flat_start()), n_hidden=512, n_mask=g.n_branch, seed=1)
    _say("warm outages: training")
    hist = ws.train(torch.as_tensor(S[ok], device="cuda"), torch.as_tensor(V[ok], device="cuda"),
                    mask=torch.as_tensor(mask[ok], device="cuda"), epochs=20, batch=256, lr=5e-4)
    assert hist[-1] < 0.5 * hist[0], hist
    _say("warm outages: solving the test rows")
    St = _sbus(g, rows, seed=1)
    outs_t = syn.random_outages(g, rows, seed=3)
    Bt = _capi.Batch(plan, rows, precision=_capi.MIXED, per_row_ybus=True)
    Bt.set_sbus(St)
    Bt.set_v(g.flat_start())
    Bt.set_ybus_edits(*syn.outage_edits(g, outs_t))
    stats = Bt.solve_warm(ws.mlp, k_warm=4, growth=2.0, max_iter=10, tol=1e-8, mask=syn.outage_mask(g, outs_t))
    st, it, res = Bt.results()
    Vw = Bt.get_v()
    Yrows = np.stack([orc.ybus_values_on_pattern(g.Ybus, orc.outage_ybus(g, o)) for o in outs_t])
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), St, g.pv, g.pq, 10, 1e-8, n_threads=16, Yx_rows=Yrows)
    np.testing.assert_array_equal(st, sto)
    conv = sto == 0
    assert conv.mean() > 0.95 and (res[conv] < 1e-8).all()
    assert np.abs(Vw[conv] - Vo[conv]).max() < 1e-6
    _say(f"warm outages: mean updates {it[conv].mean():.2f} vs flat {ito[conv].mean():.2f}, "
         f"re-queued {stats['n_requeued']}")
    assert it[conv].mean() <= ito[conv].mean() - 0.75
    # the same solve with the mask resident on the device
    _say("warm outages: resident mask")
    Bt.set_v(g.flat_start())
    Bt.set_warm_mask(syn.outage_mask(g, outs_t))
    stats2 = Bt.solve_warm(ws.mlp, k_warm=4, growth=2.0, max_iter=10, tol=1e-8)
    st2, it2, _ = Bt.results()
    np.testing.assert_array_equal(st2, st)
    np.testing.assert_array_equal(it2, it)
    assert np.abs(Bt.get_v() - Vw).max() <= 1e-12
    assert stats2["n_requeued"] == stats["n_requeued"]
    Bt.set_warm_mask(None)
    with pytest.raises(ValueError):  # the network needs its mask again
        Bt.solve_warm(ws.mlp, k_warm=4, growth=2.0, max_iter=10, tol=1e-8)
