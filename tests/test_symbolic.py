"""Host symbolic layer (CPU): J layout, ordering, programs, C ABI loading."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import program_interp as pi
from gpusim2grid_amd import _capi
from gpusim2grid_amd import synthetic as syn
from oracle import oracle as orc


def test_library_exports_every_declared_symbol():
    L = _capi.lib()
    names = _capi.exported_symbols()
    assert len(names) >= 19
    for name in names:
        assert hasattr(L, name), name
    assert L.g2g_version() >= 100


@pytest.fixture(scope="module", params=["c0", "118", "c1"])
def grid(request):
    return {"c0": syn.grid_c0, "118": syn.grid_118, "c1": syn.grid_c1}[request.param]()


@pytest.fixture(params=[1, 2])
def walkers(request, monkeypatch):
    """Plans built with one walker, or with the walk split over two."""
    monkeypatch.setenv("G2G_WALKERS", str(request.param))
    return request.param


def test_jacobian_layout_matches_oracle(grid):
    P = _capi.Plan.from_grid(grid.Ybus, grid.pv, grid.pq, grid.slack)
    Jp, Ji, maps = orc.jpattern(grid.Ybus, grid.pv, grid.pq)
    np.testing.assert_array_equal(P.array("jp"), Jp)
    np.testing.assert_array_equal(P.array("ji"), Ji)
    np.testing.assert_array_equal(P.array("maps").reshape(4, -1), maps)
    np.testing.assert_array_equal(P.array("pvpq"), np.sort(np.concatenate([grid.pv, grid.pq])))
    inf = P.info()
    assert inf["d"] == grid.pv.size + 2 * grid.pq.size
    # strictly increasing columns per row (acpf_nr.cu:226-235)
    for r in range(inf["d"]):
        assert np.all(np.diff(Ji[Jp[r]:Jp[r + 1]]) > 0)


def _dense_symbolic_fill(A_pattern, perm):
    """Brute-force symbolic elimination of the symmetrised permuted pattern."""
    B = (A_pattern + A_pattern.T)[np.ix_(perm, perm)] != 0
    d = B.shape[0]
    B = B.copy()
    for k in range(d):
        nz = np.flatnonzero(B[k + 1:, k]) + k + 1
        B[np.ix_(nz, nz)] = True
    return int(np.tril(B, -1).sum()), B


def test_symbolic_fill_equals_brute_force():
    g = syn.grid_118()
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    inf = P.info()
    Jp, Ji = P.array("jp"), P.array("ji")
    A = sp.csr_matrix((np.ones(Ji.size), Ji, Jp), shape=(inf["d"], inf["d"])).toarray()
    perm = P.array("perm")
    nL, B = _dense_symbolic_fill(A, perm)
    assert inf["nLU"] == inf["d"] + 2 * nL
    colp, coli = P.array("colp"), P.array("coli")
    for k in range(inf["d"]):
        np.testing.assert_array_equal(coli[colp[k]:colp[k + 1]], np.flatnonzero(B[k + 1:, k]) + k + 1)
    # minimum degree reduces the fill against the natural order
    nL_nat, _ = _dense_symbolic_fill(A, np.arange(inf["d"]))
    assert nL <= nL_nat


def _perturbed(g, seed=0):
    rng = np.random.default_rng(seed)
    return g.flat_start() * np.exp(1j * rng.normal(0, 0.05, g.n_bus)) * (1 + rng.normal(0, 0.02, g.n_bus))


def test_programs_factor_like_superlu_natural(grid, walkers):
    """The static-pivot LU under the plan's order equals SuperLU's factors of
    P A P^T with no pivoting (natural order, diag_pivot_thresh=0)."""
    if grid.n_bus > 200:
        pytest.skip("dense comparison on small grids only")
    P = _capi.Plan.from_grid(grid.Ybus, grid.pv, grid.pq, grid.slack)
    inf = P.info()
    assert inf["walkers"] == walkers
    F, J, _ = orc.f_and_j(grid.Ybus, _perturbed(grid), grid.Sbus, grid.pv, grid.pq)
    A = pi.astream_from_csr(inf, P.array("apos"), P.array("fpos"), J.data[None], F[None])
    U, cl = pi.factor_forward(P, A)
    assert cl[0] == 0
    perm = P.array("perm")
    Jperm = J.toarray()[np.ix_(perm, perm)]
    lu = spla.splu(sp.csc_matrix(Jperm), permc_spec="NATURAL", diag_pivot_thresh=0.0,
                   options=dict(SymmetricMode=True))
    np.testing.assert_array_equal(lu.perm_r, np.arange(inf["d"]))
    Uref = lu.U.toarray()
    uoff = P.array("uoff")
    colp, coli = P.array("colp"), P.array("coli")
    for k in range(inf["d"]):
        rec = U[0, uoff[k]:uoff[k] + 2 + colp[k + 1] - colp[k]]
        assert rec[1] == pytest.approx(1.0 / Uref[k, k], rel=1e-12)  # stored as the reciprocal
        np.testing.assert_allclose(rec[2:], Uref[k, coli[colp[k]:colp[k + 1]]], rtol=1e-11, atol=1e-12)


def test_raw_program_solves(grid, walkers):
    F, J, _ = orc.f_and_j(grid.Ybus, _perturbed(grid, 1), grid.Sbus, grid.pv, grid.pq)
    P = _capi.Plan.from_matrix(J)
    inf = P.info()
    rows = 3
    vals = np.stack([J.data * (1 + 0.01 * k) for k in range(rows)])
    rhs = np.stack([F * (k + 1) for k in range(rows)])
    A = pi.astream_from_csr(inf, P.array("apos"), P.array("fpos"), vals, rhs)
    U, _ = pi.factor_forward(P, A)
    x = pi.backward(P, U)
    for k in range(rows):
        Jk = sp.csr_matrix((vals[k], J.indices, J.indptr), shape=J.shape)
        xs = spla.spsolve(Jk.tocsc(), rhs[k])
        np.testing.assert_allclose(x[k], xs, rtol=1e-10, atol=1e-10 * np.abs(xs).max())


def test_nr_program_applies_the_polar_update(grid, walkers):
    V = _perturbed(grid, 2)
    P = _capi.Plan.from_grid(grid.Ybus, grid.pv, grid.pq, grid.slack)
    inf = P.info()
    F, J, _ = orc.f_and_j(grid.Ybus, V, grid.Sbus, grid.pv, grid.pq)
    A = pi.astream_from_csr(inf, P.array("apos"), P.array("fpos"), J.data[None], F[None])
    U, _ = pi.factor_forward(P, A)
    x = pi.backward(P, U)[0]
    Vn = pi.polar_update(V, x, P.array("pvpq"), grid.pq)
    dx = spla.spsolve(J.tocsc(), F)
    np.testing.assert_allclose(Vn, orc.polar_update_np(V, dx, grid.pv, grid.pq), rtol=0, atol=1e-10)


def test_saddle_matrix_clamp_is_counted():
    """gpusim2grid tests/python/test_own_lu.py:128-145: a 3x3 saddle whose first
    pivot is 0 in the identity order is clamped (and counted)."""
    A = sp.csr_matrix((np.array([0., 1., 2., 1., 1., 0., 2., 0., 1.]), np.array([0, 1, 2] * 3),
                       np.array([0, 3, 6, 9])), shape=(3, 3))
    P = _capi.Plan.from_matrix(A, order=[0, 1, 2])
    inf = P.info()
    np.testing.assert_array_equal(P.array("perm"), [0, 1, 2])
    b = np.array([1.0, 2.0, 3.0])
    Ast = pi.astream_from_csr(inf, P.array("apos"), P.array("fpos"), A.data[None], b[None])
    U, cl = pi.factor_forward(P, Ast, pivot_eps=1e-12, scale=np.array([2.0]))
    assert cl[0] == 1
    x = pi.backward(P, U)[0]
    r = np.abs(A @ x - b).max() / np.abs(b).max()
    assert r > 1e-8  # the clamp is visible


def test_invalid_inputs_raise_value_error():
    g = syn.grid_c0()
    with pytest.raises(ValueError):
        _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, [g.pv[0]])  # slack also PV
    # a bus in neither pv nor pq is fixed, as in the reference's array mode
    inf = _capi.Plan.from_grid(g.Ybus, g.pv[:-1], g.pq, g.slack).info()
    assert inf["npvpq"] == len(g.pv) - 1 + len(g.pq)
    Y = g.Ybus.tolil()
    Y[g.pq[0], g.pq[0]] = 0
    Y = sp.csr_matrix(Y)
    Y.eliminate_zeros()
    with pytest.raises(ValueError):
        _capi.Plan.from_grid(Y, g.pv, g.pq, g.slack)  # no diagonal entry


LDS_BYTES = 160 * 1024


def ring_bytes(W, real_size):
    """program_format.hpp walk_ring_bytes: per walker a control area, a value
    ring (four walkers stream half-size chunks) and a program ring with its
    128-word mirror, then the cross-walker flags."""
    chunk = (16 if real_size == 8 else 32) // (2 if W >= 4 else 1)
    ring_slots = (8 if real_size == 8 else 4) if W == 1 else 4
    return W * (256 + chunk * ring_slots * 64 * real_size + (512 * 4 + 128) * 4) + 1280


def walk_lds(inf, real_size):
    """The factorization kernel's LDS: rings, then the slots, the zero and
    trash slots and the long-column scratch rows."""
    W = inf["walkers"]
    scratch = 2 * inf["max_c"] * W if inf["max_c"] > 7 else 0
    return ring_bytes(W, real_size) + (inf["slots_f"] + 2 + scratch) * 64 * real_size


def backward_lds(inf, real_size):
    return ring_bytes(inf["bwalkers"], real_size) + (inf["slots_b"] + 2) * 64 * real_size


def test_slots_stay_small_on_the_large_grids(monkeypatch):
    """The postorder keeps the live update set small: one walker over one
    backward set fits the LDS in FP64 (the plan every grid falls back to), the
    6,500-bus grid's two-walker split fits it in FP32, and its four backward
    walkers fit in either precision."""
    monkeypatch.setenv("G2G_BWALKERS", "1")
    for g in (syn.grid_c1(), syn.grid_c2()):
        monkeypatch.setenv("G2G_WALKERS", "1")
        inf = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack).info()
        assert inf["max_c"] <= 16 and inf["bwalkers"] == 1
        assert walk_lds(inf, 8) <= LDS_BYTES and backward_lds(inf, 8) <= LDS_BYTES
        assert inf["nLU"] < 1.2 * inf["nJ"]
    monkeypatch.setenv("G2G_WALKERS", "2")
    monkeypatch.delenv("G2G_BWALKERS")
    g = syn.grid_c2()
    inf = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack).info()
    assert inf["bwalkers"] == 4
    assert walk_lds(inf, 4) <= LDS_BYTES
    assert backward_lds(inf, 4) <= LDS_BYTES and backward_lds(inf, 8) <= LDS_BYTES


def test_two_walkers_split_the_tree_evenly(monkeypatch):
    """The 6,500-bus grid's walk splits into two halves within 5 % of each
    other and a top part of a few pivots; segments tile the streams."""
    monkeypatch.setenv("G2G_WALKERS", "2")
    g = syn.grid_c2()
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    inf = P.info()
    seg = P.array("segs").reshape(-1, 9)
    assert inf["walkers"] == 2 and seg.shape[0] == 2
    d = inf["d"]
    assert seg[:, 8].sum() == d and 0 < inf["n_top"] < 0.01 * d
    first = seg[:, 8] - np.array([inf["n_top"], 0])
    assert abs(first[0] - first[1]) < 0.05 * d / 2
    # walker 1's A and U segments come first, walker 0's start at aligned offsets
    assert seg[1, 4] == 0 and seg[0, 4] >= seg[1, 5] and seg[0, 4] % 32 == 0
    assert seg[1, 6] == 0 and seg[0, 6] >= seg[1, 7] and seg[0, 6] % 32 == 0
    assert seg[0, 4] + seg[0, 5] == inf["a_len"] and seg[0, 6] + seg[0, 7] == inf["u_len"]
    # program segments are whole prefetch chunks
    assert (seg[:, [0, 1, 2, 3]] % 512 == 0).all()


def test_four_backward_walkers_split_the_walk_evenly(monkeypatch):
    """The 6,500-bus grid's backward substitution splits into four sets within
    5 % of each other; their U segments tile the U stream from the last set
    down, and their program segments are whole prefetch chunks."""
    monkeypatch.setenv("G2G_WALKERS", "2")
    g = syn.grid_c2()
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    inf = P.info()
    seg = P.array("bsegs").reshape(-1, 9)
    assert inf["bwalkers"] == 4 and seg.shape[0] == 4
    d = inf["d"]
    assert seg[:, 8].sum() == d
    assert seg[:, 8].max() - seg[:, 8].min() < 0.05 * d / 4
    order = np.argsort(seg[:, 6])
    starts, lens = seg[order, 6], seg[order, 7]
    assert starts[0] == 0 and (starts[1:] >= starts[:-1] + lens[:-1]).all()
    assert starts[-1] + lens[-1] == inf["u_len"]
    assert (seg[:, [2, 3]] % 512 == 0).all()


@pytest.mark.parametrize("wf,wb", [(1, 1), (1, 2), (1, 4), (2, 2), (2, 4)])
def test_u_stride_covers_every_whole_chunk_read(monkeypatch, wf, wb):
    """Each walker's U producer streams its segment in whole chunks from the
    segment's start, which for a backward set inside one factorization walker
    is not chunk-aligned: a batch's U stride reaches past every such read, so
    the last tile never reads beyond its buffer."""
    monkeypatch.setenv("G2G_WALKERS", str(wf))
    monkeypatch.setenv("G2G_BWALKERS", str(wb))
    for g in (syn.grid_c1(), syn.grid_c2()):
        P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
        inf = P.info()
        assert (inf["walkers"], inf["bwalkers"]) == (wf, wb)
        for real, chunk, stride in ((8, 16, inf["u_stride_f64"]), (4, 32, inf["u_stride_f32"])):
            assert stride % chunk == 0 and stride >= inf["u_len"]
            for name, W in (("segs", wf), ("bsegs", wb)):
                kch = chunk // (2 if W >= 4 else 1)
                for s in P.array(name).reshape(-1, 9):
                    assert s[6] + -(-s[7] // kch) * kch <= stride


@pytest.mark.parametrize("wf", [1, 2])
def test_a_stride_covers_every_whole_chunk_read(monkeypatch, wf):
    """The factorization's A producers stream each walker's segment in whole
    chunks from its start (full chunks, or half chunks in the half-ring
    geometry two FP32 walkers may take), the program producers stream whole
    program chunks, and the split walk's producers do the same over its own
    segments: a batch's A stride, and each program's length, reach past every
    such read."""
    monkeypatch.setenv("G2G_WALKERS", str(wf))
    for g in (syn.grid_c1(), syn.grid_c2()):
        P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
        inf = P.info()
        assert inf["walkers"] == wf
        for chunk, stride in ((16, inf["a_stride_f64"]), (32, inf["a_stride_f32"])):
            assert stride % chunk == 0 and stride >= max(inf["a_len"], inf["split_a_len"])
            for kch in (chunk, chunk // 2):  # own geometry, half rings
                for s in P.array("segs").reshape(-1, 9):
                    assert s[4] + -(-s[5] // kch) * kch <= stride
            if inf["n_split"] > 0:  # the split walk's factorization rings (two-walker geometry)
                for s in P.array("split_segs").reshape(-1, 9):
                    assert s[4] + -(-s[5] // chunk) * chunk <= stride
        for name, prog in (("segs", "prog_f"), ("bsegs", "prog_b"), ("split_segs", "prog_fs"),
                           ("split_bsegs", "prog_bs")):
            if name.startswith("split") and inf["n_split"] == 0:
                continue
            n_words = P.array(prog).size
            for s in P.array(name).reshape(-1, 9):
                off, ln = (s[0], s[1]) if "bsegs" not in name else (s[2], s[3])
                assert ln % 512 == 0 and off + ln <= n_words


def _split_records(prog, start, length, n_piv):
    """(kind, items) of every record of one split-walk factorization program:
    kind 'gather' with its (dst ref, pend index) pairs, 'dump' with its refs,
    'step' otherwise (no items)."""
    r4 = lambda v: (v + 3) // 4 * 4
    p, end, k, out = start, start + length, 0, []
    h = int(prog[p]) & 0xFFFF
    while p < end and (k < n_piv or h in (0x7FFC, 0x7FFD)):
        c = h & 0x7FFF
        if c == 0x7FFC:
            m = int(prog[p + 1])
            out.append(("gather", [(int(prog[p + 2 + 2 * i]), int(prog[p + 3 + 2 * i])) for i in range(m)]))
            nw = r4(2 + 2 * m)
        elif c == 0x7FFD:
            m = int(prog[p + 1])
            out.append(("dump", [int(prog[p + 3 + i]) for i in range(m)]))
            nw = r4(3 + m)
        elif c == 0x7FFE:
            nw = r4(2 + 2 * int(prog[p + 1]))
            out.append(("merge", []))
        elif h >> 15:
            nw, k = r4(7 + 4 * c + (c + 1) * c), k + 2
            out.append(("step", []))
        else:
            nw, k = r4(3 + 2 * c + (c + 1) * c), k + 1
            out.append(("step", []))
        h = (int(prog[p]) >> 16) & 0xFFFF
        p += nw
    return out


def test_split_walk_gathers_keep_each_target_together():
    """The factorization kernel loads a gather record's pending values and
    target slots in one batch and writes each target once, after summing the
    run of items that share it: that needs every target's items consecutive
    within a record and no slot shared by two runs.  Dump records name
    distinct slots."""
    for g in (syn.grid_c1(), syn.grid_c2()):
        P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
        inf = P.info()
        assert inf["n_split"] > 0
        prog = P.array("prog_fs").astype(np.int64) & 0xFFFFFFFF
        n_gather = 0
        for s in P.array("split_segs").reshape(-1, 9):
            for kind, items in _split_records(prog, s[0], s[1], s[8]):
                if kind == "gather":
                    n_gather += 1
                    assert 0 < len(items) <= 30
                    runs = [d for i, (d, _) in enumerate(items) if i == 0 or items[i - 1][0] != d]
                    assert len(runs) == len(set(runs))
                elif kind == "dump":
                    assert 0 < len(items) <= 30 and len(items) == len(set(items))
        assert n_gather > 10


def test_split_walk_programs_solve_like_the_main_walk():
    """The split walk (fine sets, middle parts, top part, each on slots of its
    own, pending targets through the dump / gather records, top x through the
    preload records) solves the 2,000-bus Jacobian like scipy, in its own
    stream layout; its slots fit a workgroup's LDS in FP32 and FP64."""
    g = syn.grid_c1()
    P = _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    inf = P.info()
    assert inf["n_split"] > 8 and inf["pend_len"] > 0
    F, J, _ = orc.f_and_j(g.Ybus, _perturbed(g, 3), g.Sbus, g.pv, g.pq)
    inf2 = dict(inf, a_len=inf["split_a_len"])
    A = pi.astream_from_csr(inf2, P.array("split_apos"), P.array("split_fpos"), J.data[None], F[None])
    U, cl = pi.factor_forward(P, A, split=True)
    assert cl[0] == 0
    x = pi.backward(P, U, split=True)[0]
    xs = spla.spsolve(J.tocsc(), F)
    assert np.abs(x - xs).max() < 1e-9 * max(1.0, np.abs(xs).max())
    # every equation and entry has a position in the split layout, all distinct
    sa, sf = P.array("split_apos"), P.array("split_fpos")
    pos = np.concatenate([sa[sa >= 0], sf])
    assert (sf >= 0).all() and np.unique(pos).size == pos.size and pos.max() < inf["split_a_len"]
    for rs in (4, 8):
        chunk = 32 if rs == 4 else 16  # full chunks (factorization), half chunks (backward)
        ring_f = 256 + chunk * 4 * 64 * rs + (512 * 4 + 128) * 4 + 1280
        ring_b = 256 + chunk // 2 * 4 * 64 * rs + (512 * 4 + 128) * 4 + 1280
        assert ring_f + (inf["slots_fs"] + 2) * 64 * rs <= LDS_BYTES
        assert ring_b + (inf["slots_bs"] + 2) * 64 * rs <= LDS_BYTES


def test_small_plans_have_no_split_walk():
    g = syn.grid_118()
    assert _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack).info()["n_split"] == 0

