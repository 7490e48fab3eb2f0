"""Writes the golden power-flow fixtures (run from the repository root:
``python tests/golden/make_golden.py``).

The reference's own test grids come from pandapower, which is not available
here, so the fixtures are the CPU oracle's FP64 flat-start solutions on the
seeded synthetic grids:

* ``c0_single.npz``: the 14-bus grid, one solve (Ybus CSR, Sbus, V0, V, status,
  iterations, residual);
* ``g118_sweep.npz``: the 118-bus grid, 64 injection scenarios (seed 1);
* ``c1_outages.npz``: the 2,000-bus grid, 16 single-branch outages (the
  first non-bridge branches), as value edits on the shared Ybus.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gpusim2grid_amd import synthetic as syn  # noqa: E402
from oracle import oracle as orc  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def _grid_arrays(g):
    Y = g.Ybus.tocsr()
    Y.sort_indices()
    return dict(n=g.n_bus, Yp=Y.indptr.astype(np.int32), Yi=Y.indices.astype(np.int32), Yx=Y.data,
                pv=np.asarray(g.pv, np.int32), pq=np.asarray(g.pq, np.int32), slack=np.int32(g.slack),
                V0=g.flat_start())


def main():
    g = syn.grid_c0()
    S = g.Sbus
    V, st, it, res = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8)
    np.savez(os.path.join(HERE, "c0_single.npz"), S=S, V=V[0], status=st, iters=it, res=res, **_grid_arrays(g))

    g = syn.grid_118()
    p, q = syn.injection_scenarios(g, 64, seed=1)
    S = (p + 1j * q) / g.sn_mva
    V, st, it, res = orc.nr_batch(g.Ybus, g.flat_start(), S, g.pv, g.pq, 10, 1e-8, n_threads=8)
    np.savez(os.path.join(HERE, "g118_sweep.npz"), S=S, V=V, status=st, iters=it, res=res, **_grid_arrays(g))

    g = syn.grid_c1()
    Y = g.Ybus.tocsr()
    Y.sort_indices()
    br = np.flatnonzero(~syn.bridges(g.n_bus, g.branch_from, g.branch_to))[:16]
    rows_Y, row_ptr, ent, delta = [], [0], [], []
    for l in br:
        Yo = orc.outage_ybus(g, [l])
        Yx = orc.ybus_values_on_pattern(Y, Yo)
        rows_Y.append(Yx)
        d = Y.data - Yx
        nz = np.flatnonzero(d != 0)
        ent.extend(nz.tolist())
        delta.extend(d[nz].tolist())
        row_ptr.append(len(ent))
    Yx_rows = np.array(rows_Y)
    V, st, it, res = orc.nr_batch(g.Ybus, g.flat_start(), g.Sbus, g.pv, g.pq, 10, 1e-8, n_threads=8,
                                  Yx_rows=Yx_rows)
    np.savez(os.path.join(HERE, "c1_outages.npz"), S=g.Sbus, branches=br.astype(np.int32),
             row_ptr=np.asarray(row_ptr, np.int32), entry=np.asarray(ent, np.int32),
             delta=np.asarray(delta, np.complex128), V=V, status=st, iters=it, res=res, **_grid_arrays(g))


if __name__ == "__main__":
    main()
