"""Device-resident sessions: DLPack capsules alias the sessions' device
results (the reference's compute() contract, injection_sweep/gpu_facade.py:
515-524, and dlpack_export.hpp:48-58), and the extension rows through the
facades."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import synthetic as syn
from oracle import ext_oracle as xo
from oracle import oracle as orc
from ext_cases import features

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tuple(g, V=None):
    return (g.Ybus, g.flat_start() if V is None else V, g.Sbus, np.array([g.slack]), np.array([1.0]), g.pv, g.pq)


def test_compute_capsule_aliases_the_session_results():
    import torch
    from gpusim2grid_amd import InjectionSweepGPU
    g = syn.grid_118()
    rows = 150
    _say("capsule aliasing injection sweep")
    sw = InjectionSweepGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=10, scheduling="continuous")
    p, q = syn.injection_scenarios(g, rows, seed=21)
    sw.set_injections(p, q, g.sn_mva)
    t1 = torch.from_dlpack(sw.compute(batch_size=64))
    assert t1.is_cuda and t1.dtype == torch.complex128 and tuple(t1.shape) == (rows, g.n_bus)
    first = t1.cpu().numpy().copy()
    np.testing.assert_array_equal(first, sw.V_results.to_numpy())
    ptr = t1.data_ptr()
    # new injections: the next run writes the same device memory
    sw.set_injections(1.2 * p, 1.2 * q, g.sn_mva)
    t2 = torch.from_dlpack(sw.compute(batch_size=64))
    assert t2.data_ptr() == ptr
    now = t1.cpu().numpy()  # the tensor taken before the run sees the new results
    assert not np.array_equal(now, first)
    np.testing.assert_array_equal(now, sw.V_results.to_numpy())
    S = 1.2 * (p + 1j * q) / g.sn_mva
    Vo, sto, _, _ = orc.nr_batch(g.Ybus, sw.solver.get_v_base(), S, g.pv, g.pq, 10, 1e-8, n_threads=8)
    np.testing.assert_array_equal(sw.get_row_status(), sto)
    assert np.abs(now - Vo).max() < 1e-9


def test_v_dlpack_aliases_the_single_solve():
    import torch
    from gpusim2grid_amd import AcPfGPU
    g = syn.grid_118()
    _say("capsule aliasing acpf")
    ac = AcPfGPU(_tuple(g), init_from_n_powerflow=False)
    t = torch.from_dlpack(ac.v_dlpack())
    v1 = t.cpu().numpy().copy()
    np.testing.assert_array_equal(v1, ac.solve())
    ac.resolve(g.flat_start(), 1.1 * g.Sbus)
    t2 = torch.from_dlpack(ac.v_dlpack())
    assert t2.data_ptr() == t.data_ptr()
    v2 = t.cpu().numpy()
    assert not np.array_equal(v1, v2)
    Vo, sto, _, _ = orc.nr_batch(g.Ybus, g.flat_start(), 1.1 * g.Sbus, g.pv, g.pq, 10, 1e-8)
    assert sto[0] == 0 and np.abs(v2 - Vo[0]).max() < 1e-9


def test_contingency_rows_not_simulated_are_nan_on_the_device():
    import torch
    from gpusim2grid_amd import ContingencyAnalysisGPU
    g = syn.grid_118()
    _say("contingency skipped rows nan")
    ca = ContingencyAnalysisGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=10, scheduling="continuous")
    ca.set_branch_data(g.branch_from, g.branch_to, g.yff, g.yft, g.ytf, g.ytt)
    br = syn.bridges(g.n_bus, g.branch_from, g.branch_to)
    ok = np.flatnonzero(~br)[:20]
    ca.add_contingencies_by_branch_id([[int(l)] for l in ok] + [[int(np.flatnonzero(br)[0])]])
    V = torch.from_dlpack(ca.compute(batch_size=16)).cpu().numpy()
    st = ca.get_row_status()
    assert st[-1] == 3 and np.isnan(V[-1]).all()
    assert (st[:-1] == 0).all() and np.isfinite(V[:-1]).all()


def test_facades_take_extension_rows():
    from gpusim2grid_amd import AcPfGPU, InjectionSweepGPU
    g = syn.grid_118()
    ext = features(g)
    L = xo.Layout(g.n_bus, g.pv, g.pq, g.slack, ext)
    _say("facades with extension rows")
    ac = AcPfGPU(_tuple(g), init_from_n_powerflow=False, max_iter=15, extensions=ext)
    Vo, xso, sto, ito, _ = xo.nr_ext(g.Ybus, g.flat_start(), g.Sbus, L, max_iter=15)
    assert ac.timings.converged and sto == 0 and ac.timings.nb_iter == ito
    assert np.abs(ac.solve() - Vo).max() < 1e-9
    np.testing.assert_allclose(ac.ext_state(), xso, rtol=0, atol=1e-8)
    rows = 40
    p, q = syn.injection_scenarios(g, rows, seed=9)
    sw = InjectionSweepGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=15, max_iter_base=15,
                           scheduling="continuous", extensions=ext)
    sw.set_injections(p, q, g.sn_mva)
    V = sw.compute(batch_size=32, as_numpy=True)
    xs = sw.ext_state()
    Vb = sw.solver.get_v_base()
    S = (p + 1j * q) / g.sn_mva
    for k in range(rows):
        Vk, xk, sk, ik, _ = xo.nr_ext(g.Ybus, Vb, S[k], L, max_iter=15)
        assert sw.get_row_status()[k] == sk == 0 and sw.get_row_iterations()[k] == ik
        assert np.abs(V[k] - Vk).max() < 1e-9
        np.testing.assert_allclose(xs[k], xk, rtol=0, atol=1e-8)
