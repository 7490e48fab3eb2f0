"""The rest of the `_gpusim2grid` surface: the raw batched-LU entry points
(solve_cudss_raw, benchmark_cudss_batch_raw; raw_cudss_solve.hpp:31-121),
per-row generator set-points (set_gen_v, python_bindings.cpp:2232-2243) and
the DLPack exports (v_dlpack, v_results_dlpack), against scipy and the oracle."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tuple(g, V=None):
    return (g.Ybus, g.flat_start() if V is None else V, g.Sbus, np.array([g.slack]), np.array([1.0]), g.pv, g.pq)


def _jacobian(g):
    _, J, _ = orc.f_and_j(g.Ybus, g.flat_start(), g.Sbus, g.pv, g.pq)
    J = sp.csr_matrix(J)
    J.sort_indices()
    return J


def test_solve_cudss_raw_matches_scipy():
    from gpusim2grid_amd import _gpusim2grid as m
    g = syn.grid_118()
    _say("solve_cudss_raw 118")
    J = _jacobian(g)
    rhs = np.random.default_rng(1).normal(size=J.shape[0])
    x = np.asarray(m.solve_cudss_raw(J.shape[0], J.indptr.tolist(), J.indices.tolist(), J.data.tolist(),
                                     rhs.tolist()))
    np.testing.assert_allclose(x, spla.spsolve(J.tocsc(), rhs), rtol=1e-9, atol=1e-11)
    with pytest.raises(ValueError):
        m.solve_cudss_raw(J.shape[0], J.indptr.tolist(), J.indices.tolist(), J.data.tolist(), rhs[:-1].tolist())


def test_benchmark_cudss_batch_raw_solves_every_slot():
    from gpusim2grid_amd import _gpusim2grid as m
    g = syn.grid_118()
    _say("benchmark_cudss_batch_raw 118")
    J = _jacobian(g)
    r = m.benchmark_cudss_batch_raw(J.shape[0], J.indptr.tolist(), J.indices.tolist(), J.data.tolist(), 100,
                                    n_refactorize=3, n_active=37)
    assert (r.dim, r.nnz, r.batch_size, r.n_refactorize, r.n_active) == (J.shape[0], J.nnz, 100, 3, 37)
    assert r.max_rel_residual < 1e-10 and r.n_nonfinite_slots == 0
    assert r.lu_nnz >= J.nnz and r.analysis_ms > 0 and r.factorize_ms > 0 and r.refactorize_ms > 0
    assert r.active_max_rel_residual < 1e-10 and r.n_inactive_touched == 0
    with pytest.raises(ValueError):
        m.benchmark_cudss_batch_raw(J.shape[0], J.indptr.tolist(), J.indices.tolist(), J.data.tolist(), 4,
                                    n_groups=2)


def test_set_gen_v_reseeds_the_fixed_magnitudes():
    from gpusim2grid_amd import InjectionSweepGPU
    g = syn.grid_118()
    rows = 24
    _say("set_gen_v 118")
    p, q = syn.injection_scenarios(g, rows, seed=3)
    inj = InjectionSweepGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=10, scheduling="continuous")
    inj.set_injections(p, q, g.sn_mva)
    inj.set_v_init_from(g.flat_start())
    # generators on every PV bus and the slack, plus one on a PQ bus (ignored)
    gen_bus = np.concatenate([g.pv, np.ravel(g.slack), g.pq[:1]]).astype(np.int32)
    rng = np.random.default_rng(4)
    gen_v = rng.uniform(0.97, 1.05, size=(rows, gen_bus.size))
    gen_v[3, 0] = np.nan  # NaN leaves that (row, generator) alone
    assert inj.solver.is_vm_fixed_bus[g.pv].all() and not inj.solver.is_vm_fixed_bus[g.pq].any()
    inj.set_gen_v(gen_v, gen_bus)
    V = inj.compute(batch_size=16, as_numpy=True)
    st = inj.get_row_status()
    it = inj.get_row_iterations()
    S = (p + 1j * q) / g.sn_mva
    V0 = np.tile(g.flat_start(), (rows, 1))
    for r in range(rows):
        for k, b in enumerate(gen_bus[:-1]):
            if not np.isnan(gen_v[r, k]):
                V0[r, b] = gen_v[r, k] * np.exp(1j * np.angle(V0[r, b]))
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, V0, S, g.pv, g.pq, 10, 1e-8)
    np.testing.assert_array_equal(st, sto)
    np.testing.assert_array_equal(it, ito)
    assert np.abs(V - Vo).max() < 1e-9
    np.testing.assert_allclose(np.abs(V[5, g.pv]), gen_v[5, :g.pv.size], rtol=1e-12)
    np.testing.assert_allclose(np.abs(V[3, g.pv[0]]), np.abs(g.flat_start()[g.pv[0]]), rtol=1e-12)


def test_dlpack_exports():
    import torch
    from gpusim2grid_amd import AcPfGPU, InjectionSweepGPU
    g = syn.grid_118()
    _say("dlpack exports")
    pf = AcPfGPU(_tuple(g), init_from_n_powerflow=False)
    pf.solve()
    v = torch.utils.dlpack.from_dlpack(pf.session.v_dlpack())
    assert v.is_cuda and v.shape == (g.n_bus,)
    np.testing.assert_array_equal(v.cpu().numpy(), pf.session.get_v())
    p, q = syn.injection_scenarios(g, 10, seed=5)
    inj = InjectionSweepGPU(_tuple(g), init_from_n_powerflow=False, nb_iter=10)
    inj.set_injections(p, q, g.sn_mva)
    inj.compute(batch_size=8)
    t = torch.utils.dlpack.from_dlpack(inj.solver.v_results_dlpack())
    assert t.is_cuda and tuple(t.shape) == (10, g.n_bus)
    np.testing.assert_array_equal(t.cpu().numpy(), inj.V_results.to_numpy())
