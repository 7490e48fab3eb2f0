"""Streamed rounds replayed from a captured HIP graph against the same rounds
launched one by one: every round's statuses, iteration counts and residuals,
and every taken row's V, bit for bit, with fills, takes and host reads
between the replays and a change of round shape midway."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec,refine", [(_capi.FP64, False), (_capi.MIXED, False), (_capi.MIXED, True)])
def test_graph_rounds_equal_eager_rounds(prec, refine):
    """refine: one FP64-residual refinement round per update, whose
    multiplier-storing factorization and forward-only walk are captured too."""
    g = syn.grid_118()
    rows, slots = 500, 128
    _say(f"stream graph prec={prec} refine={refine}")
    p, q = syn.injection_scenarios(g, rows, seed=21)
    S = (p + 1j * q) / g.sn_mva
    outs = syn.random_outages(g, rows, seed=22)
    plan = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    Bg = _capi.Batch(plan, slots, precision=prec, per_row_ybus=True)
    Be = _capi.Batch(plan, slots, precision=prec, per_row_ybus=True)
    Be.set_stream_graph(False)
    for B in (Bg, Be):
        if refine:
            B.set_step_control(ir_steps=1)
            B.set_refinement_residual(True)
        B.stream_begin()
    free = list(range(slots))
    busy = {}
    nxt = 0
    got = 0
    rounds = 0
    while got < rows:
        k = min(len(free), rows - nxt, 40)  # partial refills: fills between replays
        if k:
            sl = [free.pop() for _ in range(k)]
            ids = list(range(nxt, nxt + k))
            ed = syn.outage_edits(g, [outs[i] for i in ids])
            for B in (Bg, Be):
                B.stream_fill(sl, S[ids], g.flat_start()[None], edits=ed)
            busy.update(zip(sl, ids))
            nxt += k
        kk = 2 if 10 <= rounds < 14 else 1  # another round shape midway: a second capture
        for B in (Bg, Be):
            B.stream_round(k=kk, max_iter=10, tol=1e-8)
        rounds += 1
        assert rounds < 300
        sg, ig, rg = Bg.results()
        se, ie, re_ = Be.results()
        np.testing.assert_array_equal(sg, se)
        np.testing.assert_array_equal(ig, ie)
        sim = sg != _capi.NOTSIMULATED
        np.testing.assert_array_equal(rg[sim], re_[sim])
        done = [s for s in busy if sg[s] not in (-1, _capi.NOTSIMULATED)]
        if done:
            np.testing.assert_array_equal(Bg.stream_take(done), Be.stream_take(done))
            for s in done:
                busy.pop(s)
                free.append(s)
            got += len(done)
    on, replays = Bg.stream_graph_info()
    assert on and replays >= rounds - 4
    on_e, replays_e = Be.stream_graph_info()
    assert not on_e and replays_e == 0
