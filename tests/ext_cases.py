"""Extension-row cases on the synthetic grids (shared by the oracle and GPU tests)."""
import numpy as np

from gpusim2grid_amd.extensions import Extensions, HvdcLine, VoltageGroup


def _far_pq_pairs(g, k, seed):
    """k pairs of pq buses far apart in the bus graph."""
    import scipy.sparse.csgraph as cg
    rng = np.random.default_rng(seed)
    D = cg.shortest_path(abs(g.Ybus) > 0, unweighted=True)
    pq = np.asarray(g.pq)
    out = []
    while len(out) < k:
        a, b = rng.choice(pq, 2, replace=False)
        if D[a, b] >= 4:
            out.append((int(a), int(b)))
    return out


def features(g, which=("ds", "hvdc", "vc"), seed=0):
    """Extensions on grid g: a distributed slack over the slack and three PV
    buses, two droop lines (one linear with losses, one saturated) and two
    voltage-control groups (an SVC on its own bus; two generators holding a
    neighbouring bus)."""
    ext = Extensions()
    slack = int(np.atleast_1d(g.slack)[0])
    pv = np.asarray(g.pv)
    if "ds" in which:
        w = np.zeros(g.n_bus)
        w[slack] = 0.4
        w[pv[:3]] = 0.2
        ext.slack_weights = w
    if "hvdc" in which:
        (a1, b1), (a2, b2) = _far_pq_pairs(g, 2, seed)
        ext.hvdc = [HvdcLine(a1, b1, p0=0.04, k=0.6, lf1=0.01, lf2=0.015, r=0.02),
                    HvdcLine(a2, b2, p0=0.02, k=0.3, lf1=0.01, lf2=0.01, r=0.01, pmax12=0.03, status=1)]
    if "vc" in which:
        rng = np.random.default_rng(seed + 1)
        used = {h.bus1 for h in ext.hvdc} | {h.bus2 for h in ext.hvdc}
        pq = [int(b) for b in g.pq if int(b) not in used]
        Y = g.Ybus.tocsr()
        svc = pq[rng.integers(len(pq))]
        # a pq bus with two pq neighbours for the remote group
        for reg in rng.permutation(pq):
            nb = [int(j) for j in Y.indices[Y.indptr[reg]:Y.indptr[reg + 1]] if j != reg and int(j) in pq and j != svc]
            if len(nb) >= 2 and reg != svc:
                break
        ext.voltage = [VoltageGroup(reg_bus=svc, v_set=1.0, buses=[svc], slopes=[0.02]),
                       VoltageGroup(reg_bus=int(reg), v_set=1.01, buses=nb[:2], weights=[1.0, 2.0])]
    return ext
