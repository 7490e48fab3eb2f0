"""The reference's facades and ``_gpusim2grid`` sessions (tuple mode), on the
HIP path, against the CPU oracle -- modelled on the reference's
tests/python/test_injection_batch.py and test_contingency_batch.py."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tuple(g, V=None):
    return (g.Ybus, g.flat_start() if V is None else V, g.Sbus, np.array([g.slack]), np.array([1.0]), g.pv, g.pq)


# ---- CPU: argument handling ---------------------------------------------------
def test_precision_and_device_names():
    from gpusim2grid_amd import facades as F
    from gpusim2grid_amd import _gpusim2grid as cpp
    assert F._precision("fp64") == cpp.Precision.Fp64
    assert F._precision(None) == cpp.Precision.Fp64
    assert F._precision("mixed") == cpp.Precision.Mixed
    with pytest.raises(ValueError):
        F._precision("fp32")
    with pytest.raises(ValueError):
        F._precision("bf16")
    assert F._device(None) == -1 and F._device("cuda") == -1
    assert F._device("cuda:3") == 3 and F._device(2) == 2
    with pytest.raises(ValueError):
        F._device("cpu")


def test_tuple_mode_only():
    from gpusim2grid_amd import facades as F
    with pytest.raises(TypeError):
        F._tuple(object())
    g = syn.grid_c0()
    t = F._tuple(_tuple(g))
    assert t[1].dtype == complex and t[5].dtype == np.int32
    with pytest.raises(ValueError):
        F._check_common(0, None, None, None, True)
    # ir_n_steps validation (contingency_analysis/__init__.py:157-167)
    with pytest.raises(ValueError):
        F._check_common(-1, None, None, None, None)
    with pytest.raises(TypeError):
        F._check_common(True, None, None, None, None)
    with pytest.raises(TypeError):
        F._check_common(1.5, None, None, None, None)
    assert F._check_common(2, None, None, None, None) == dict(ir_n_steps=2, scaling_max_voltage_change=False,
                                                            max_dVa=0.5, max_dVm=0.1)
    assert F._check_common(0, True, 0.2, None, None)["max_dVa"] == 0.2


def test_module_surface():
    from gpusim2grid_amd import _gpusim2grid as cpp
    for name in ("AcPfNrSession", "InjectionSweepSession", "ContingencyAnalysisSession", "ScenarioSweepSession",
                 "acpf_nr_gpu", "acpf_nr_gpu_injection", "warmup", "RowStatus", "BatchScheduling",
                 "LinearSolver", "ReorderingAlg", "MatchingAlg", "PivotEpsilonAlg", "ContingencySolverType",
                 "AcPfTimings", "BatchTimings", "ContingencyTimings", "TimingEntry"):
        assert hasattr(cpp, name), name
    assert [int(cpp.RowStatus.Converged), int(cpp.RowStatus.MaxIter), int(cpp.RowStatus.Diverged),
            int(cpp.RowStatus.NotSimulated)] == [0, 1, 2, 3]
    assert cpp.is_fp32 is False


# ---- GPU ------------------------------------------------------------------------
@pytest.mark.gpu
def test_acpf_gpu_flat_start_and_resolve():
    from gpusim2grid_amd import AcPfGPU
    g = syn.grid_118()
    _say("AcPfGPU 118")
    ac = AcPfGPU(_tuple(g), init_from_n_powerflow=False)
    V = ac.solve()
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), g.Sbus, g.pv, g.pq, 10, 1e-8)
    assert ac.timings.converged and ac.timings.nb_iter == ito[0]
    assert np.max(np.abs(V - Vo[0])) < 1e-9
    # a converged start passes the presolved check; a flat start does not
    AcPfGPU(_tuple(g, V), init_from_n_powerflow=True)
    with pytest.raises(RuntimeError):
        AcPfGPU(_tuple(g), init_from_n_powerflow=True)
    # resolve with new injections
    S2 = g.Sbus * 1.05
    t = ac.resolve(g.flat_start(), S2)
    Vo2, _, ito2, _ = orc.nr_batch(g.Ybus, g.flat_start(), S2, g.pv, g.pq, 10, 1e-8)
    assert t.converged and t.nb_iter == ito2[0]
    assert np.max(np.abs(ac.solve() - Vo2[0])) < 1e-9
    # ac_pf keeps the set-points of the Vm-fixed buses
    Vac = ac.ac_pf(g.flat_start() * 0.98, 10, 1e-8)
    fixed = np.asarray(ac.session.vm_col_of_bus) < 0
    np.testing.assert_allclose(np.abs(Vac[fixed]), np.abs(g.flat_start()[fixed]), rtol=1e-14)
    # F and J at the solution
    F = ac.get_F()
    assert np.max(np.abs(F)) < 1e-8
    J = ac.get_J()
    Fo, Jo, _ = orc.f_and_j(g.Ybus, ac.solve(), g.Sbus, g.pv, g.pq)
    assert np.max(np.abs(J.toarray() - Jo.toarray())) < 1e-9 * np.max(np.abs(Jo.toarray()))


@pytest.mark.gpu
def test_acpf_nr_gpu_function():
    from gpusim2grid_amd import _gpusim2grid as cpp
    g = syn.grid_c0()
    out = np.zeros(g.n_bus, complex)
    t = cpp.acpf_nr_gpu(out, g.Ybus, g.flat_start(), g.Sbus, np.array([g.slack], np.int32), np.ones(1),
                        np.asarray(g.pv, np.int32), np.asarray(g.pq, np.int32), 10, 1e-8, 2)
    Vo, _, _, _ = orc.nr_batch(g.Ybus, g.flat_start(), g.Sbus, g.pv, g.pq, 10, 1e-8)
    assert t.converged
    assert np.max(np.abs(out - Vo[0])) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("scheduling", ["chunked", "continuous"])
def test_injection_sweep_against_oracle(scheduling):
    from gpusim2grid_amd import InjectionSweepGPU
    g = syn.grid_118()
    rows = 200
    _say(f"InjectionSweepGPU 118 {scheduling}")
    p, q = syn.injection_scenarios(g, rows, seed=4)
    sw = InjectionSweepGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=6, scheduling=scheduling)
    sw.set_injections(p, q, g.sn_mva)
    V = sw.compute(batch_size=64, as_numpy=True)
    Vb = sw.solver.get_v_base()
    S = (p + 1j * q) / g.sn_mva
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, Vb, S, g.pv, g.pq, 6, 1e-8, n_threads=8)
    np.testing.assert_array_equal(sw.get_row_status(), sto)
    if scheduling == "continuous":
        np.testing.assert_array_equal(sw.get_row_iterations(), ito)
    else:
        # chunked: every row runs exactly nb_iter updates; the oracle does the
        # same when its test can never pass
        Vo, _, ito, _ = orc.nr_batch(g.Ybus, Vb, S, g.pv, g.pq, 6, 1e-300, n_threads=8)
        assert (ito == 6).all()
        np.testing.assert_array_equal(sw.get_row_iterations(), ito)
    assert np.max(np.abs(V - Vo)) < 1e-9
    assert (sw.last_residuals() < 1e-8).all()
    # batch-size invariance (test_contingency_batch.py:174-222's pattern)
    V2 = sw.compute(batch_size=1000, as_numpy=True)
    assert np.array_equal(V, V2)
    # the device capsule path
    import torch
    cap = sw.compute(batch_size=1000)
    Vt = torch.from_dlpack(cap).cpu().numpy()
    assert np.array_equal(Vt, V2)


@pytest.mark.gpu
def test_injection_sweep_function():
    from gpusim2grid_amd import _gpusim2grid as cpp
    g = syn.grid_118()
    p, q = syn.injection_scenarios(g, 50, seed=6)
    V, res, t = cpp.acpf_nr_gpu_injection(g.Ybus, g.flat_start(), g.Sbus, np.array([g.slack], np.int32),
                                          np.ones(1), np.asarray(g.pv, np.int32), np.asarray(g.pq, np.int32),
                                          p, q, g.sn_mva, 32, 6)
    assert V.shape == (50 * g.n_bus,)
    assert (res < 1e-8).all() and t.n_chunks == 2


@pytest.mark.gpu
def test_contingency_analysis_against_oracle():
    from gpusim2grid_amd import ContingencyAnalysisGPU, RowStatus
    g = syn.grid_c1()
    _say("ContingencyAnalysisGPU c1")
    br = syn.bridges(g.n_bus, g.branch_from, g.branch_to)
    ok = np.flatnonzero(~br)[:40]
    bad = np.flatnonzero(br)[:2]
    ctg = [[int(l)] for l in ok] + [[int(l)] for l in bad] + [[int(ok[0]), int(ok[1])]]
    ca = ContingencyAnalysisGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=10, scheduling="continuous")
    ca.set_branch_data(g.branch_from, g.branch_to, g.yff, g.yft, g.ytf, g.ytt)
    ca.add_contingencies_by_branch_id(ctg)
    V = ca.compute(batch_size=16, as_numpy=True)
    st = ca.get_row_status()
    it = ca.get_row_iterations()
    Vb = ca.solver.get_v_base()
    Y = g.Ybus.tocsr()
    Y.sort_indices()
    for r, c in enumerate(ctg):
        if r in (40, 41):  # islanding: not simulated, NaN results
            assert st[r] == int(RowStatus.NotSimulated)
            assert np.isnan(ca.last_residuals()[r]) and np.isnan(V[r]).all()
            continue
        yx = orc.ybus_values_on_pattern(Y, orc.outage_ybus(g, c))
        Vo, sto, ito, _ = orc.nr_batch(Y, Vb, g.Sbus, g.pv, g.pq, 10, 1e-8, Yx_rows=yx[None, :])
        assert st[r] == sto[0] and it[r] == ito[0], r
        assert np.max(np.abs(V[r] - Vo[0])) < 1e-9, r


@pytest.mark.gpu
def test_scenario_sweep_injections_and_topology():
    from gpusim2grid_amd import ScenarioSweepGPU
    g = syn.grid_118()
    rows = 40
    _say("ScenarioSweepGPU 118")
    p, q = syn.injection_scenarios(g, rows, seed=8)
    br = syn.bridges(g.n_bus, g.branch_from, g.branch_to)
    nb = np.flatnonzero(~br)
    rng = np.random.default_rng(0)
    topo = [sorted(rng.choice(nb, size=r % 3, replace=False).tolist()) for r in range(rows)]
    ss = ScenarioSweepGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=10, scheduling="continuous")
    ss.set_branch_data(g.branch_from, g.branch_to, g.yff, g.yft, g.ytf, g.ytt)
    ss.set_injections(p, q, g.sn_mva)
    ss.set_topology(topo)
    ss.set_v_init_from(g.flat_start())
    V = ss.compute(batch_size=16, as_numpy=True)
    st, it = ss.get_row_status(), ss.get_row_iterations()
    disc = ss.get_disconnected()
    Y = g.Ybus.tocsr()
    Y.sort_indices()
    S = (p + 1j * q) / g.sn_mva
    for r in range(rows):
        if disc[r]:
            assert st[r] == 3
            continue
        yx = orc.ybus_values_on_pattern(Y, orc.outage_ybus(g, topo[r]))
        Vo, sto, ito, _ = orc.nr_batch(Y, g.flat_start(), S[r], g.pv, g.pq, 10, 1e-8, Yx_rows=yx[None, :])
        assert st[r] == sto[0] and it[r] == ito[0], r
        assert np.max(np.abs(V[r] - Vo[0])) < 1e-9, r


@pytest.mark.gpu
def test_warmup_and_mixed_precision_session():
    import gpusim2grid_amd as G
    G.warmup()
    g = syn.grid_118()
    p, q = syn.injection_scenarios(g, 64, seed=9)
    sw = G.InjectionSweepGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=10, scheduling="continuous",
                             precision="mixed")
    sw.set_injections(p, q, g.sn_mva)
    V = sw.compute(as_numpy=True)
    Vo, sto, _, _ = orc.nr_batch(g.Ybus, sw.solver.get_v_base(), (p + 1j * q) / g.sn_mva, g.pv, g.pq, 10, 1e-8,
                                 n_threads=8)
    np.testing.assert_array_equal(sw.get_row_status(), sto)
    assert np.max(np.abs(V - Vo)) < 1e-7


@pytest.mark.gpu
def test_branch_flows_against_closed_form():
    from gpusim2grid_amd import ContingencyAnalysisGPU
    g = syn.grid_118()
    _say("branch flows 118")
    br = syn.bridges(g.n_bus, g.branch_from, g.branch_to)
    ok = np.flatnonzero(~br)[:20]
    ctg = [[int(l)] for l in ok] + [[int(np.flatnonzero(br)[0])]]
    vn = np.random.default_rng(0).choice([63.0, 225.0, 400.0], size=g.n_bus)
    ca = ContingencyAnalysisGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=10, scheduling="continuous")
    ca.set_branch_data(g.branch_from, g.branch_to, g.yff, g.yft, g.ytf, g.ytt, bus_vn_kv=vn, sn_mva=g.sn_mva)
    ca.add_contingencies_by_branch_id(ctg)
    V = ca.compute(batch_size=8, as_numpy=True)
    ca.compute_flows()
    o, e = ca.or_amps, ca.ex_amps
    assert o.shape == (len(ctg), len(g.branch_from))
    f, t = g.branch_from, g.branch_to
    bo = g.sn_mva * 1e6 / (np.sqrt(3) * vn[f] * 1e3)
    be = g.sn_mva * 1e6 / (np.sqrt(3) * vn[t] * 1e3)
    for r, c in enumerate(ctg[:-1]):
        Io = np.abs(g.yff * V[r, f] + g.yft * V[r, t]) * bo
        Ie = np.abs(g.ytf * V[r, f] + g.ytt * V[r, t]) * be
        Io[c] = 0.0
        Ie[c] = 0.0
        np.testing.assert_allclose(o[r], Io, rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(e[r], Ie, rtol=1e-12, atol=1e-9)
    assert np.isnan(o[-1]).all()  # the islanding row is not simulated
