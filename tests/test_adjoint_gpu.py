"""The batched adjoint J^T x = rhs (through the C ABI and the AcPfGPU facade)
against the oracle's Jacobian at the same voltages."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,prec", [("118", _capi.FP64), ("118", _capi.MIXED), ("c1", _capi.FP64)])
def test_adjoint_solves_the_transposed_jacobian(name, prec):
    g = {"118": syn.grid_118, "c1": syn.grid_c1}[name]()
    rows = 70
    _say(f"adjoint {name} prec={prec}")
    p, q = syn.injection_scenarios(g, rows, seed=13)
    S = (p + 1j * q) / g.sn_mva
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(P, rows, precision=prec)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.solve(max_iter=10, tol=1e-8)
    st0, it0, _ = B.results()
    V = B.get_v()
    d = P.info()["d"]
    rhs = np.random.default_rng(3).normal(size=(rows, d))
    x = B.solve_adjoint(rhs)
    tol = 1e-9 if prec == _capi.FP64 else 1e-4
    for r in (0, 33, rows - 1):
        _, J, _ = orc.f_and_j(g.Ybus, V[r], S[r], g.pv, g.pq)
        res = np.abs(J.T @ x[r] - rhs[r]).max() / np.abs(rhs[r]).max()
        assert res < tol, (r, res)
        # and not the forward system's solution
        assert np.abs(J @ x[r] - rhs[r]).max() > 1e-3 * np.abs(rhs[r]).max()
    st1, it1, _ = B.results()
    np.testing.assert_array_equal(st0, st1)
    np.testing.assert_array_equal(it0, it1)


def test_facade_solve_jt_dlpack():
    import torch
    from torch.utils import dlpack
    from gpusim2grid_amd import AcPfGPU
    g = syn.grid_118()
    _say("adjoint facade")
    ac = AcPfGPU((g.Ybus, g.flat_start(), g.Sbus, np.array([g.slack]), np.array([1.0]), g.pv, g.pq),
                 init_from_n_powerflow=False)
    V = ac.solve()
    _, J, _ = orc.f_and_j(g.Ybus, V, g.Sbus, g.pv, g.pq)
    rhs = torch.arange(J.shape[0], dtype=torch.float64)
    x = dlpack.from_dlpack(ac.solve_JT_dlpack(rhs)).cpu().numpy()
    assert np.abs(J.T @ x - rhs.numpy()).max() < 1e-9 * np.abs(rhs.numpy()).max()
