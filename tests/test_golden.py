"""The committed golden fixtures (tests/golden/make_golden.py): the seeded
generator and the oracle reproduce them on the CPU, and the HIP path matches
them on the GPU."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {"c0_single": syn.grid_c0, "g118_sweep": syn.grid_118, "c1_outages": syn.grid_c1}


def _load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def _ybus(z):
    n = int(z["n"])
    return sp.csr_matrix((z["Yx"], z["Yi"], z["Yp"]), shape=(n, n))


def _yx_rows(z):
    Y = _ybus(z)
    rp, ent, dl = z["row_ptr"], z["entry"], z["delta"]
    out = np.tile(Y.data, (rp.size - 1, 1))
    for r in range(rp.size - 1):
        out[r, ent[rp[r]:rp[r + 1]]] -= dl[rp[r]:rp[r + 1]]
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_generator_reproduces_fixture_grid(name):
    z = _load(name)
    g = CASES[name]()
    Y = g.Ybus.tocsr()
    Y.sort_indices()
    np.testing.assert_array_equal(Y.indptr, z["Yp"])
    np.testing.assert_array_equal(Y.indices, z["Yi"])
    np.testing.assert_array_equal(Y.data, z["Yx"])
    np.testing.assert_array_equal(g.pv, z["pv"])
    np.testing.assert_array_equal(g.pq, z["pq"])
    np.testing.assert_array_equal(g.flat_start(), z["V0"])


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_fixture(name):
    z = _load(name)
    yx = _yx_rows(z) if "row_ptr" in z else None
    V, st, it, res = orc.nr_batch(_ybus(z), z["V0"], z["S"], z["pv"], z["pq"], 10, 1e-8, n_threads=4, Yx_rows=yx)
    np.testing.assert_array_equal(st, z["status"])
    np.testing.assert_array_equal(it, z["iters"])
    assert np.max(np.abs(V - np.atleast_2d(z["V"]))) < 1e-12
    # every stored solution satisfies the power-flow equations to the tolerance
    F = orc.mismatch_np(_ybus(z), np.atleast_2d(z["V"])[0], np.atleast_2d(z["S"])[0], z["pv"], z["pq"]) \
        if yx is None else None
    if F is not None:
        assert np.max(np.abs(F)) < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("prec", [0, 1])
def test_gpu_matches_fixture(name, prec):
    from gpusim2grid_amd import _capi
    print(f"golden {name} precision {prec}", flush=True)
    z = _load(name)
    Y = _ybus(z)
    S = np.atleast_2d(z["S"])
    rows = np.atleast_2d(z["V"]).shape[0]
    plan = _capi.Plan.from_grid(Y, z["pv"], z["pq"], int(np.asarray(z["slack"]).reshape(-1)[0]))
    edits = "row_ptr" in z
    B = _capi.Batch(plan, rows, precision=prec, per_row_ybus=edits)
    B.set_sbus(S[0] if S.shape[0] == 1 else S)
    B.set_v(z["V0"])
    if edits:
        B.set_ybus_edits(z["row_ptr"], z["entry"], z["delta"])
    B.solve(max_iter=10, tol=1e-8)
    st, it, res = B.results()
    V = B.get_v()
    np.testing.assert_array_equal(st, z["status"])
    if prec == 0:
        np.testing.assert_array_equal(it, z["iters"])
    assert np.max(np.abs(V - np.atleast_2d(z["V"]))) < (1e-8 if prec == 0 else 1e-6)
