"""The extension-row oracle pinned by finite differences and by special
cases that reduce to the bare system solved by the C oracle."""
import numpy as np
import pytest

from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn
from gpusim2grid_amd.extensions import Extensions, HvdcLine, VoltageGroup
from oracle import ext_oracle as xo
from oracle import oracle as orc
from ext_cases import features


def _layout(g, ext, pv=None, pq=None):
    return xo.Layout(g.n_bus, g.pv if pv is None else pv, g.pq if pq is None else pq, g.slack, ext)


def _point(g, L, seed=0):
    rng = np.random.default_rng(seed)
    V = g.flat_start() * np.exp(1j * rng.normal(0, 0.05, g.n_bus)) * (1 + rng.normal(0, 0.02, g.n_bus))
    xs = rng.normal(0, 0.05, L.n_xs)
    return V, xs


def _unknowns(V, xs, L):
    th, vm = np.angle(V), np.abs(V)
    x = np.zeros(L.d)
    m = L.col_th >= 0
    x[L.col_th[m]] = th[m]
    x[L.col_vm[L.pq]] = vm[L.pq]
    if L.ds:
        x[L.sa] = xs[0]
    for c, qc in enumerate(L.q_col):
        x[qc] = xs[int(L.ds) + c]
    return x


def _from_unknowns(x, V, L):
    th, vm = np.angle(V), np.abs(V)
    m = L.col_th >= 0
    th[m] = x[L.col_th[m]]
    vm[L.pq] = x[L.col_vm[L.pq]]
    xs = np.zeros(L.n_xs)
    if L.ds:
        xs[0] = x[L.sa]
    for c, qc in enumerate(L.q_col):
        xs[int(L.ds) + c] = x[qc]
    return vm * np.exp(1j * th), xs


@pytest.mark.parametrize("which", [("ds",), ("hvdc",), ("vc",), ("ds", "hvdc", "vc")])
def test_jacobian_equals_finite_differences(which):
    g = syn.grid_c0() if which != ("ds", "hvdc", "vc") else syn.grid_118()
    if "hvdc" in which and g.n_bus < 30:
        g = syn.grid_118()
    ext = features(g, which)
    L = _layout(g, ext)
    V, xs = _point(g, L)
    J = xo.j_ext(g.Ybus, V, xs, L)
    x0 = _unknowns(V, xs, L)
    h = 1e-6
    Jfd = np.zeros_like(J)
    for k in range(L.d):
        e = np.zeros(L.d)
        e[k] = h
        Vp, xp = _from_unknowns(x0 + e, V, L)
        Vm_, xm = _from_unknowns(x0 - e, V, L)
        # J is d(mismatch)/dx and F = -mismatch
        Jfd[:, k] = -(xo.f_ext(g.Ybus, Vp, g.Sbus, xp, L) - xo.f_ext(g.Ybus, Vm_, g.Sbus, xm, L)) / (2 * h)
    np.testing.assert_allclose(J, Jfd, rtol=0, atol=2e-7 * max(1.0, np.abs(J).max()))


def test_bare_layout_is_the_c_oracles():
    g = syn.grid_118()
    L = _layout(g, None)
    V, xs = _point(g, L)
    F, Jo, _ = orc.f_and_j(g.Ybus, V, g.Sbus, g.pv, g.pq)
    np.testing.assert_allclose(xo.f_ext(g.Ybus, V, g.Sbus, xs, L), F, atol=1e-12)
    np.testing.assert_allclose(xo.j_ext(g.Ybus, V, xs, L), Jo.toarray(), atol=1e-11)


def test_slack_alone_is_the_plain_slack():
    """Weight 1 on the slack: the bare solution, with sa the slack's P balance."""
    g = syn.grid_118()
    w = np.zeros(g.n_bus)
    s = int(np.atleast_1d(g.slack)[0])
    w[s] = 1.0
    L = _layout(g, Extensions(slack_weights=w))
    V, xs, st, it, _ = xo.nr_ext(g.Ybus, g.flat_start(), g.Sbus, L)
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), g.Sbus, g.pv, g.pq, 10, 1e-8)
    assert st == 0 and sto[0] == 0
    assert np.abs(V - Vo[0]).max() < 1e-9
    P = (Vo[0] * np.conj(g.Ybus @ Vo[0])).real
    assert xs[0] == pytest.approx(g.Sbus[s].real - P[s], abs=1e-8)


def test_shared_slack_moves_every_participant():
    """Weights over the slack and PV buses: each participant absorbs its share."""
    g = syn.grid_118()
    ext = features(g, ("ds",))
    L = _layout(g, ext)
    V, xs, st, it, _ = xo.nr_ext(g.Ybus, g.flat_start(), g.Sbus, L)
    assert st == 0
    P = (V * np.conj(g.Ybus @ V)).real
    w = ext.slack_weights
    part = np.nonzero(w)[0]
    np.testing.assert_allclose(P[part], g.Sbus.real[part] - xs[0] * w[part], atol=1e-8)
    other = np.setdiff1d(g.pvpq if hasattr(g, "pvpq") else np.concatenate([g.pv, g.pq]), part)
    np.testing.assert_allclose(P[other], g.Sbus.real[other], atol=1e-8)


def test_own_bus_controller_is_a_pv_bus():
    """One controller regulating its own bus with slope 0 is a PV bus at v_set."""
    g = syn.grid_118()
    b = int(g.pq[5])
    vset = 1.002
    ext = Extensions(voltage=[VoltageGroup(reg_bus=b, v_set=vset, buses=[b])])
    L = _layout(g, ext)
    V, xs, st, it, _ = xo.nr_ext(g.Ybus, g.flat_start(), g.Sbus, L)
    assert st == 0
    pv = np.sort(np.append(g.pv, b))
    pq = np.setdiff1d(g.pq, [b])
    V0 = g.flat_start()
    V0[b] = vset
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, V0, g.Sbus, pv, pq, 10, 1e-8)
    assert sto[0] == 0
    assert np.abs(V - Vo[0]).max() < 1e-9
    Q = (Vo[0] * np.conj(g.Ybus @ Vo[0])).imag
    assert xs[0] == pytest.approx(Q[b] - g.Sbus[b].imag, abs=1e-8)


def test_constant_droop_line_is_two_injections():
    """k = 0 and no losses: p0 leaves bus1 and arrives at bus2."""
    g = syn.grid_118()
    a, b = int(g.pq[3]), int(g.pq[40])
    ext = Extensions(hvdc=[HvdcLine(a, b, p0=0.05, k=0.0)])
    L = _layout(g, ext)
    V, xs, st, it, _ = xo.nr_ext(g.Ybus, g.flat_start(), g.Sbus, L)
    S = g.Sbus.copy()
    S[a] -= 0.05
    S[b] += 0.05
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8)
    assert st == 0 and sto[0] == 0 and it == ito[0]
    assert np.abs(V - Vo[0]).max() < 1e-10


def test_all_features_converge_and_hold_their_equations():
    g = syn.grid_118()
    ext = features(g)
    L = _layout(g, ext)
    V, xs, st, it, r = xo.nr_ext(g.Ybus, g.flat_start(), g.Sbus, L, max_iter=15)
    assert st == 0, (it, r)
    k0 = 1
    for gi, grp in enumerate(ext.voltage):
        Q = xs[k0:k0 + len(grp.buses)]
        assert abs(V[grp.reg_bus]) + np.dot(grp.slopes, Q) == pytest.approx(grp.v_set, abs=1e-8)
        np.testing.assert_allclose(Q / np.asarray(grp.weights), Q[0] / grp.weights[0], atol=1e-8)
        k0 += len(grp.buses)


def test_plan_pattern_covers_the_oracle_jacobian():
    """The C++ planner's extension layout is the oracle's: same dimension and
    every nonzero of the oracle's J is in the plan's pattern."""
    g = syn.grid_118()
    ext = features(g)
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack, ext=ext)
    inf = P.info()
    L = _layout(g, ext)
    assert inf["d"] == L.d and inf["n_xs"] == L.n_xs and inf["sa_col"] == L.sa
    V, xs = _point(g, L)
    J = xo.j_ext(g.Ybus, V, xs, L)
    jp, ji = P.array("jp"), P.array("ji")
    pat = np.zeros_like(J, dtype=bool)
    for r in range(L.d):
        pat[r, ji[jp[r]:jp[r + 1]]] = True
    assert not (np.abs(J) > 0)[~pat].any()


def test_plan_rejects_bad_extensions():
    g = syn.grid_118()
    with pytest.raises(ValueError):  # a controller at a pv bus
        _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack,
                             ext=Extensions(voltage=[VoltageGroup(int(g.pq[0]), 1.0, [int(g.pv[0])])]))
    with pytest.raises(ValueError):  # an hvdc line looping on one bus
        _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack, ext=Extensions(hvdc=[HvdcLine(3, 3, 0.1, 0.1)]))
    with pytest.raises(ValueError):  # weights of the wrong length
        _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack, ext=Extensions(slack_weights=np.ones(3)))
