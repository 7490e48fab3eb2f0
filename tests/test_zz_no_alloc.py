"""The allocation-free raw solve (g2g_lu_solve_batch_noalloc) against the
allocating one (g2g_lu_solve_batch) through the C ABI: bit-identical
solutions and clamp counts, on two calls in a row, the second smaller."""
import sys

import numpy as np
import pytest

from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _jacobians(g, rows, seed):
    rng = np.random.default_rng(seed)
    vals, rhs = [], []
    J0 = None
    for _ in range(rows):
        V = g.flat_start() * np.exp(1j * rng.normal(0, 0.05, g.n_bus)) * (1 + rng.normal(0, 0.02, g.n_bus))
        F, J, _ = orc.f_and_j(g.Ybus, V, g.Sbus, g.pv, g.pq)
        J0 = J if J0 is None else J0
        vals.append(J.data)
        rhs.append(F)
    return J0, np.stack(vals), np.stack(rhs)


def test_noalloc_is_exported():
    assert hasattr(_capi.lib(), "g2g_lu_solve_batch_noalloc")


@pytest.mark.gpu
@pytest.mark.parametrize("prec,ir", [(_capi.FP64, 0), (_capi.MIXED, 0), (_capi.MIXED, 1)])
def test_gpu_noalloc_matches_the_allocating_solve(prec, ir):
    import torch
    g = syn.grid_118()
    J0, vals, rhs = _jacobians(g, 150, seed=4)
    P = _capi.Plan.from_matrix(J0)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    for rows in (150, 37):  # the second call is smaller and must not see the first's leftovers
        _say(f"noalloc prec={prec} ir={ir} rows={rows}")
        v, b = vals[:rows], rhs[:rows]
        x_ref, cl_ref = P.lu_solve(v, b, precision=prec, ir_steps=ir)  # also makes the plan resident
        dv = torch.from_numpy(np.ascontiguousarray(v)).to(dev)
        db = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
        dx = torch.full((rows, b.shape[1]), float("nan"), dtype=torch.float64, device=dev)
        dc = torch.full((rows,), -7, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        _capi.lu_solve_noalloc(P, rows, prec, dv.data_ptr(), db.data_ptr(), dx.data_ptr(), ir_steps=ir,
                               d_clamped=dc.data_ptr(), stream=stream)
        x = dx.cpu().numpy()
        np.testing.assert_array_equal(x, x_ref)
        np.testing.assert_array_equal(dc.cpu().numpy(), cl_ref)
        r = max(np.abs(sp_mat(J0, v[k]) @ x[k] - b[k]).max() / np.abs(b[k]).max() for k in range(rows))
        assert r < (1e-10 if prec == _capi.FP64 else 1e-4)


def sp_mat(J0, data):
    import scipy.sparse as sp
    return sp.csr_matrix((data, J0.indices, J0.indptr), shape=J0.shape)


@pytest.mark.gpu
def test_gpu_noalloc_rejects_rows_beyond_its_scratch():
    import torch
    g = syn.grid_118()
    J0, vals, rhs = _jacobians(g, 1, seed=5)
    P = _capi.Plan.from_matrix(J0)
    P.lu_solve(vals, rhs)
    _say("noalloc oversize rows")
    dev = torch.device("cuda:0")
    rows = 1 << 16  # the 118-bus streams take ~20 KB per row in FP64: over a GB, far beyond 64 MiB
    dv = torch.zeros((1, vals.shape[1]), dtype=torch.float64, device=dev)
    with pytest.raises(ValueError, match="rows"):
        _capi.lu_solve_noalloc(P, rows, _capi.FP64, dv.data_ptr(), dv.data_ptr(), dv.data_ptr())


@pytest.mark.gpu
def test_gpu_noalloc_needs_a_resident_plan():
    import torch
    g = syn.grid_c0()
    J0, vals, rhs = _jacobians(g, 1, seed=6)
    P = _capi.Plan.from_matrix(J0)
    dv = torch.zeros((1, vals.shape[1]), dtype=torch.float64, device="cuda:0")
    _say("noalloc plan not resident")
    with pytest.raises(ValueError, match="resident"):
        _capi.lu_solve_noalloc(P, 1, _capi.FP64, dv.data_ptr(), dv.data_ptr(), dv.data_ptr())
