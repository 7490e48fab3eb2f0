"""The seeded synthetic grids and scenario sets (CPU)."""
import numpy as np
import pytest

from gpusim2grid_amd import synthetic as syn


@pytest.mark.parametrize("n,n_branch,kw", [(14, 20, dict(n_pv=4)), (118, 180, dict(pv_frac=0.15)),
                                           (2000, 3000, dict(pv_frac=0.15))])
def test_workload_constructors_start_at_the_accepted_draw(n, n_branch, kw):
    """The acceptance search of make_grid(seed=0) ends on ACCEPTED_DRAW[n]:
    every earlier draw fails the flat-start test and that draw passes it."""
    first = syn.ACCEPTED_DRAW[n]
    for draw in range(first):
        g = syn.make_grid(n, n_branch, seed=0, accept=False, first_draw=draw, **kw)
        assert not syn.flat_start_converges(g)
    g = syn.make_grid(n, n_branch, seed=0, accept=False, first_draw=first, **kw)
    assert syn.flat_start_converges(g)


def test_scenarios_scale_loads_and_redispatch():
    """Per bus, P and Q loads share one scale factor; per row, generation is the
    base generation times the row's total-load ratio, so net injection over
    total load is the same constant in every row."""
    g = syn.grid_118()
    p, q = syn.injection_scenarios(g, 64, seed=1)
    assert p.shape == (64, g.n_bus) and q.shape == (64, g.n_bus)
    has_q = g.q_load > 0
    np.testing.assert_allclose(q[:, ~has_q], 0.0)
    pl = (-q[:, has_q] * (g.p_load[has_q] / g.q_load[has_q])).sum(1) * g.sn_mva / g.sn_mva
    ratio = p.sum(1) / pl
    np.testing.assert_allclose(ratio, g.p_gen.sum() / g.p_load[has_q].sum() - 1.0, rtol=1e-9)
    assert np.std(pl) > 0.01 * np.mean(pl)  # the rows do differ

def test_cut_labels_find_bridges_and_two_edge_cuts():
    """random_outages' 2-edge-cut test against brute-force connectivity: a
    label of 0 marks exactly the bridges, equal labels exactly the pairs of
    branches whose removal disconnects the grid."""
    import collections
    for g in (syn.grid_118(), syn.grid_c0()):
        lab = syn._cut_labels(g)
        br = syn.bridges(g.n_bus, g.branch_from, g.branch_to)
        np.testing.assert_array_equal(lab == 0, br)
        cand = np.flatnonzero(~br)
        for i in range(len(cand)):
            for j in range(i + 1, len(cand)):
                a, b = int(cand[i]), int(cand[j])
                assert (lab[a] == lab[b]) == (not syn._connected_without(g, [a, b])), (a, b)


def test_outage_edits_rebuild_the_tripped_ybus():
    from oracle import oracle as orc
    g = syn.grid_118()
    outs = syn.random_outages(g, 6, seed=9)
    row_ptr, ent, delta = syn.outage_edits(g, outs)
    base = g.Ybus.data.copy()
    for r, brs in enumerate(outs):
        vals = base.copy()
        np.subtract.at(vals, ent[row_ptr[r]:row_ptr[r + 1]], delta[row_ptr[r]:row_ptr[r + 1]])
        want = orc.ybus_values_on_pattern(g.Ybus, orc.outage_ybus(g, brs))
        np.testing.assert_allclose(vals, want, rtol=0, atol=1e-12)
    m = syn.outage_mask(g, outs)
    assert m.shape == (6, g.n_branch) and [int(x) for x in m.sum(1)] == [len(o) for o in outs]
