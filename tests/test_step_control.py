"""Step control of the Newton loop: per-row step scaling (acpf_nr_kernels.cu:677-723,
contingency/driver.cuh:225-235) and iterative refinement of the linear solve
(own_lu/batched_lu.cu:239-260,391-411).

CPU tests pin the oracle's step scaling by invariants that do not restate its
formula, and pin the refinement method (the residual as a new right-hand side
of a refactorization of the same values) on the program interpreter against
the reference's saddle-matrix known answer (tests/python/test_own_lu.py:128-145).
GPU tests compare the HIP path, through the C ABI, with the oracle."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import program_interp as pi
from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc

SADDLE = sp.csr_matrix((np.array([0., 1., 2., 1., 1., 0., 2., 0., 1.]), np.array([0, 1, 2] * 3),
                        np.array([0, 3, 6, 9])), shape=(3, 3))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _unknown_step(g, V0, V1):
    """(dtheta over sorted pvpq, d|V| over pq) between two voltage vectors."""
    pvpq = np.sort(np.concatenate([g.pv, g.pq]))
    return np.angle(V1[pvpq] / V0[pvpq]), np.abs(V1[g.pq]) - np.abs(V0[g.pq])


# ---------------------------------------------------------------------------
# oracle: step scaling
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("bounds", [(0.02, 0.5), (0.5, 0.002), (0.01, 0.001)])
def test_oracle_step_scaling_invariants(bounds):
    """One scaled update is the unscaled one times a single alpha < 1 over
    every unknown, no bound is exceeded and the binding one is met exactly."""
    g = syn.grid_118()
    V0 = g.flat_start()
    mva, mvm = bounds
    V1, *_ = orc.nr_batch(g.Ybus, V0, g.Sbus, g.pv, g.pq, max_iter=1, tol=1e-300)
    V1s, *_ = orc.nr_batch(g.Ybus, V0, g.Sbus, g.pv, g.pq, max_iter=1, tol=1e-300, max_dva=mva, max_dvm=mvm)
    t, m = _unknown_step(g, V0, V1[0])
    ts, ms = _unknown_step(g, V0, V1s[0])
    big_t, big_m = np.abs(t) > 1e-6, np.abs(m) > 1e-6
    ratios = np.concatenate([ts[big_t] / t[big_t], ms[big_m] / m[big_m]])
    alpha = ratios.mean()
    assert 0 < alpha < 1
    np.testing.assert_allclose(ratios, alpha, rtol=1e-8)
    assert np.abs(ts).max() <= mva * (1 + 1e-9) and np.abs(ms).max() <= mvm * (1 + 1e-9)
    assert (abs(np.abs(ts).max() - mva) < 1e-9 * mva) or (abs(np.abs(ms).max() - mvm) < 1e-9 * mvm)


def test_oracle_step_scaling_loose_bounds_change_nothing_and_same_solution():
    g = syn.grid_118()
    V0 = g.flat_start()
    V, st, it, _ = orc.nr_batch(g.Ybus, V0, g.Sbus, g.pv, g.pq)
    Vl, stl, itl, _ = orc.nr_batch(g.Ybus, V0, g.Sbus, g.pv, g.pq, max_dva=10.0, max_dvm=10.0)
    np.testing.assert_array_equal(V, Vl)
    assert itl[0] == it[0]
    Vs, sts, its, _ = orc.nr_batch(g.Ybus, V0, g.Sbus, g.pv, g.pq, max_iter=40, max_dva=0.02, max_dvm=0.005)
    assert sts[0] == 0 and its[0] > it[0]
    np.testing.assert_allclose(Vs, V, rtol=0, atol=1e-8)


# ---------------------------------------------------------------------------
# interpreter: the refinement method on the saddle matrix
# ---------------------------------------------------------------------------

def _interp_refined(A, b, ir_steps, eps):
    P = _capi.Plan.from_matrix(A, order=[0, 1, 2])
    inf = P.info()
    apos, fpos = P.array("apos"), P.array("fpos")
    scale = np.array([np.abs(A.data).max()])
    Ast = pi.astream_from_csr(inf, apos, fpos, A.data[None], b[None])
    U, cl = pi.factor_forward(P, Ast, pivot_eps=eps, scale=scale)
    x = pi.backward(P, U)[0]
    for _ in range(ir_steps):
        r = b - A @ x
        U2, _ = pi.factor_forward(P, pi.astream_from_csr(inf, apos, fpos, A.data[None], r[None]),
                                  pivot_eps=eps, scale=scale)
        # the same factors: the U streams differ only in the y entries
        keep = np.ones(U.shape[1], bool)
        keep[P.array("uoff")[:-1]] = False
        np.testing.assert_array_equal(U2[:, keep], U[:, keep])
        x = x + pi.backward(P, U2)[0]
    return x, cl[0]


def test_refinement_recovers_a_clamped_pivot_interp():
    """own_lu's known answer: the clamp is visible without refinement, and each
    refinement step cuts the relative residual by orders of magnitude."""
    b = np.array([1.0, 2.0, 3.0])
    rel = []
    for k in (0, 1, 2):
        x, cl = _interp_refined(SADDLE, b, k, 1e-12)
        assert cl == 1
        rel.append(np.abs(SADDLE @ x - b).max() / np.abs(b).max())
    r0, r1, r2 = rel
    assert r0 > 1e-8
    assert r1 < 1e-2 * r0 and r2 < 1e-2 * max(r1, 1e-13), rel


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def _sbus_rows(g, rows, seed=1):
    p, q = syn.injection_scenarios(g, rows, seed=seed)
    return (p + 1j * q) / g.sn_mva


def _gpu(g, S, rows, precision, ir_steps=0, scaling=False, max_dva=0.5, max_dvm=0.1, max_iter=10):
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    B = _capi.Batch(P, rows, precision=precision)
    B.set_sbus(S)
    B.set_v(g.flat_start())
    B.set_step_control(ir_steps=ir_steps, scaling=scaling, max_dva=max_dva, max_dvm=max_dvm)
    B.solve(max_iter=max_iter, tol=1e-8)
    st, it, res = B.results()
    return B.get_v(), st, it, res


@pytest.mark.gpu
@pytest.mark.parametrize("name,bounds", [("118", (0.02, 0.005)), ("118", (0.01, 0.5)), ("118", (0.5, 0.002)),
                                         ("c1", (0.05, 0.01))])
def test_gpu_step_scaling_matches_oracle(name, bounds):
    g = {"118": syn.grid_118, "c1": syn.grid_c1}[name]()
    rows = 100
    _say(f"step scaling {name} bounds={bounds} rows={rows}")
    S = _sbus_rows(g, rows)
    V, st, it, res = _gpu(g, S, rows, _capi.FP64, scaling=True, max_dva=bounds[0], max_dvm=bounds[1], max_iter=40)
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 40, 1e-8, n_threads=16,
                                   max_dva=bounds[0], max_dvm=bounds[1])
    np.testing.assert_array_equal(st, sto)
    np.testing.assert_array_equal(it, ito)
    assert np.abs(V - Vo).max() < 1e-8
    # the bounds bind: every row needs more updates than undamped
    _, _, itu, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 40, 1e-8, n_threads=16)
    assert (ito > itu).all()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [_capi.FP64, _capi.MIXED])
def test_gpu_refined_newton_matches_oracle(prec):
    g = syn.grid_118()
    rows = 100
    _say(f"refined newton 118 prec={prec} rows={rows}")
    S = _sbus_rows(g, rows)
    V, st, it, _ = _gpu(g, S, rows, prec, ir_steps=2)
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8, n_threads=16)
    np.testing.assert_array_equal(st, sto)
    np.testing.assert_array_equal(it, ito)
    assert np.abs(V - Vo).max() < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("prec,eps", [(_capi.FP64, 1e-12), (_capi.MIXED, 1e-5)])
def test_gpu_raw_refinement_recovers_a_clamped_pivot(prec, eps):
    """tests/python/test_own_lu.py:128-145 on the raw batched solve."""
    _say(f"raw refinement saddle prec={prec}")
    P = _capi.Plan.from_matrix(SADDLE, order=[0, 1, 2])
    b = np.array([1.0, 2.0, 3.0])
    rows = 4
    vals = np.tile(SADDLE.data, (rows, 1))
    rhs = np.tile(b, (rows, 1))
    rel = []
    for k in (0, 1, 2):
        x, cl = P.lu_solve(vals, rhs, precision=prec, pivot_eps=eps, ir_steps=k)
        assert (cl == 1).all() and np.isfinite(x).all()
        rel.append(max(np.abs(SADDLE @ x[r] - b).max() for r in range(rows)) / np.abs(b).max())
    r0, r1, r2 = rel
    assert r0 > 1e-8
    if prec == _capi.FP64:
        assert r1 < 1e-2 * r0 and r2 < 1e-2 * max(r1, 1e-13), rel
    else:  # residuals in FP32: the contraction stops at float resolution
        assert r1 < max(1e-2 * r0, 1e-6) and r2 < max(1e-2 * r1, 1e-6), rel


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [_capi.FP64, _capi.MIXED])
def test_gpu_raw_refinement_does_not_hurt_on_the_jacobian(prec):
    """test_own_lu.py:102-125: one refinement step on a Jacobian batch leaves
    the residual no larger (to rounding)."""
    g = syn.grid_118()
    rows = 20
    _say(f"raw refinement jacobian prec={prec}")
    rng = np.random.default_rng(5)
    Vs = [g.flat_start() * np.exp(1j * rng.normal(0, 0.05, g.n_bus)) for _ in range(rows)]
    Js, Fs = [], []
    for V in Vs:
        F, J, _ = orc.f_and_j(g.Ybus, V, g.Sbus, g.pv, g.pq)
        Js.append(J)
        Fs.append(F)
    P = _capi.Plan.from_matrix(Js[0])
    vals = np.stack([J.data for J in Js])
    rhs = np.stack(Fs)

    def worst(x):
        return max(np.abs(Js[r] @ x[r] - rhs[r]).max() / np.abs(rhs[r]).max() for r in range(rows))

    x0, _ = P.lu_solve(vals, rhs, precision=prec, ir_steps=0)
    x1, _ = P.lu_solve(vals, rhs, precision=prec, ir_steps=1)
    r0, r1 = worst(x0), worst(x1)
    floor = 1e-13 if prec == _capi.FP64 else 1e-6
    assert r1 <= max(r0, floor), (r0, r1)


@pytest.mark.gpu
def test_gpu_facades_take_the_step_options():
    """AcPfGPU / InjectionSweepGPU with scaling_max_voltage_change and
    ir_n_steps: the damped iteration count and solution of the oracle."""
    from gpusim2grid_amd import AcPfGPU, InjectionSweepGPU
    g = syn.grid_118()
    _say("facades step options 118")
    grid = (g.Ybus, g.flat_start(), g.Sbus, np.array([g.slack]), np.array([1.0]), g.pv, g.pq)
    ac = AcPfGPU(grid, init_from_n_powerflow=False, max_iter=40, scaling_max_voltage_change=True, max_dVa=0.02,
                 max_dVm=0.005, ir_n_steps=1)
    Vo, sto, ito, _ = orc.nr_batch(g.Ybus, g.flat_start(), g.Sbus, g.pv, g.pq, 40, 1e-8, max_dva=0.02, max_dvm=0.005)
    assert ac.timings.converged and ac.timings.nb_iter == ito[0]
    assert np.abs(ac.solve() - Vo[0]).max() < 1e-8
    rows = 24
    p, q = syn.injection_scenarios(g, rows, seed=3)
    sw = InjectionSweepGPU(grid, init_from_n_powerflow=False, nb_iter=40, max_iter_base=40, scheduling="continuous",
                           scaling_max_voltage_change=True, max_dVa=0.05, max_dVm=0.01)
    sw.set_injections(p, q, g.sn_mva)
    V = sw.compute(batch_size=rows, as_numpy=True)
    # rows start from the base case, solved with the same damping
    Vb, *_ = orc.nr_batch(g.Ybus, g.flat_start(), g.Sbus, g.pv, g.pq, 40, 1e-8, max_dva=0.05, max_dvm=0.01)
    So = (p + 1j * q) / g.sn_mva
    Vr, str_, itr, _ = orc.nr_batch(g.Ybus, Vb[0], So, g.pv, g.pq, 40, 1e-8, max_dva=0.05, max_dvm=0.01)
    np.testing.assert_array_equal(sw.get_row_status(), str_)
    np.testing.assert_array_equal(sw.get_row_iterations(), itr)
    assert np.abs(V - Vr).max() < 1e-8


def _saddle_blocks(nb):
    """A block-diagonal matrix of nb saddle blocks and its block-wise identity
    order (every block's (0, 0) pivot is 0 unless a row changes it)."""
    # built by hand: a block_diag would drop SADDLE's explicit zero at (0, 0)
    indptr = np.arange(3 * nb + 1) * 3
    indices = (np.repeat(np.arange(nb) * 3, 9) + np.tile(SADDLE.indices, nb)).astype(np.int32)
    data = np.tile(SADDLE.data, nb)
    blocks = sp.csr_matrix((data, indices, indptr), shape=(3 * nb, 3 * nb))
    return blocks, np.arange(3 * nb)


def test_saddle_blocks_plan_has_two_walkers():
    """The per-row clamp test below runs on a plan whose walk is split over
    two walkers (independent blocks fall into both subtree sets)."""
    A, order = _saddle_blocks(700)
    assert A.nnz == 9 * 700 and A[3, 3] == 0 and A[3, 4] == 1  # (0, 0) of every block stays in the pattern
    inf = _capi.Plan.from_matrix(A, order=order).info()
    assert inf["walkers"] == 2 and inf["d"] == 2100


@pytest.mark.gpu
@pytest.mark.parametrize("prec,eps", [(_capi.FP64, 1e-12), (_capi.MIXED, 1e-5)])
def test_gpu_raw_clamp_counts_are_per_row_on_two_walkers(prec, eps):
    """Rows clamp different numbers of pivots, in blocks walked by both
    walkers: each row's count is its own (own_lu/batched_lu.cu:45-53)."""
    nb = 700
    A, order = _saddle_blocks(nb)
    P = _capi.Plan.from_matrix(A, order=order)
    _say(f"raw clamp counts two walkers prec={prec} info={P.info()['walkers']}")
    rows = 130
    rng = np.random.default_rng(11)
    # entry (0, 0) of each block in CSR order: the first entry of its first row
    first = A.indptr[0:3 * nb:3]
    vals = np.tile(A.data, (rows, 1))
    zero = rng.random((rows, nb)) < rng.random((rows, 1))  # rows differ in how many blocks clamp
    vals[:, first] = np.where(zero, 0.0, 3.0)  # 0: the first pivot clamps; 3: no pivot does
    rhs = rng.normal(size=(rows, 3 * nb))
    _, cl = P.lu_solve(vals, rhs, precision=prec, pivot_eps=eps)
    np.testing.assert_array_equal(cl, zero.sum(axis=1))
    assert len(np.unique(cl)) > 10
