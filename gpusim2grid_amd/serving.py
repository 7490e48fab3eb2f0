"""Serving loop: power-flow requests streamed through resident batches.

Requests (one grid's injection row plus 0..k tripped branches) arrive over
time; each grid keeps its plan and a batch of slots resident on the device
(the batch sizes are chosen so that the batches fit an HBM budget), and a
request waits in its grid's queue until a slot frees.  The loop runs rounds of
Newton iterations per grid (continuous batching: a slot is refilled as soon as
its row finishes, g2g_batch_stream_*), optionally starting every row from the
learned warm start, and records each request's latency from arrival to
result.  This is the reference's continuous scheduling (slot refill,
batch_pf_driver.cu:2720-3100) extended with arrival times, admission by
memory budget and latency accounting, which the reference does not have.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from . import _capi
from . import synthetic as syn


@dataclasses.dataclass
class Request:
    rid: int
    grid: str
    S: np.ndarray            # complex [n] per-unit injections
    outages: list            # tripped branch ids
    t_arrival: float
    t_done: float = 0.0
    status: int = -1
    iterations: int = 0
    V: np.ndarray | None = None
    # where the payload sits: row `row` of `payload` ([rows, n] complex, the
    # receive buffer S is a view of); None for a standalone S
    payload: np.ndarray | None = None
    row: int = -1


def batch_bytes(plan_info, rows, precision, per_row_ybus):
    """Device bytes of a batch of `rows` slots (the arrays g2g_batch_create
    allocates, plus the streaming and compaction buffers)."""
    rp = (rows + 63) // 64 * 64
    rs = 8 if precision == _capi.FP64 else 4
    n, nY, d = plan_info["n"], plan_info["nY"], plan_info["d"]
    b = rp * n * 16 * 2                      # V, Sbus
    b += rp * (plan_info["a_len"] + plan_info["u_len"] + d) * rs
    b += rp * 8 * 8                          # per-row scalars
    if per_row_ybus:  # each row's outage edits (16 to start) and the tiles' edited-bus flags
        b += rp * 16 * 20 + rp // 64 * n * 4
    return int(b * 1.1)


class GridState:
    """One topology's resident state: plan, batch of slots, queue."""

    def __init__(self, name, grid, slots, precision, warm=None, max_iter=10, tol=1e-8, stage_rows=4096):
        import torch
        self.name = name
        self.g = grid
        self.plan = _capi.Plan.from_grid(grid.Ybus, grid.pv, grid.pq, grid.slack)
        self.batch = _capi.Batch(self.plan, slots, precision=precision, per_row_ybus=True)
        self.batch.stream_begin()
        self.free = list(range(slots))[::-1]
        self.busy = {}                      # slot -> Request
        self.queue = collections.deque()
        self.warm = warm
        self.max_iter = max_iter
        self.tol = tol
        self.flat = grid.flat_start()
        # pinned host staging of the rows going in and out (direct DMA)
        self.stage_rows = max(1, min(int(stage_rows), slots))
        self._pin = [torch.empty((self.stage_rows, grid.n_bus), dtype=torch.complex128, pin_memory=True)
                     for _ in range(2)]
        self.S_pin, self.V_pin = (t.numpy() for t in self._pin)
        self.t = collections.Counter()  # host seconds per phase
        self.results = {}  # id(payload buffer) -> (payload, response buffer)

    def admit(self):
        """Fills free slots from the queue; returns how many."""
        k = min(len(self.free), len(self.queue), self.stage_rows)
        if k == 0:
            return 0
        t0 = time.perf_counter()
        reqs = [self.queue.popleft() for _ in range(k)]
        slots = [self.free.pop() for _ in range(k)]
        p0 = reqs[0].payload
        if p0 is not None and all(r.payload is p0 for r in reqs) and reqs[-1].row - reqs[0].row == k - 1 and \
                all(reqs[i].row + 1 == reqs[i + 1].row for i in range(k - 1)):
            S = p0[reqs[0].row:reqs[0].row + k]  # consecutive rows of the receive buffer: no staging copy
        else:
            S = np.stack([r.S for r in reqs], out=self.S_pin[:k])
        t1 = time.perf_counter()
        outs = [r.outages for r in reqs]
        edits = syn.outage_edits(self.g, outs)
        t2 = time.perf_counter()
        self.batch.stream_fill(slots, S, self.flat[None], edits=edits)
        t3 = time.perf_counter()
        if self.warm is not None:  # start voltages predicted on the device, into the slots
            mask = syn.outage_mask(self.g, outs) if self.warm.n_mask else None
            self.batch.stream_warm(self.warm.mlp, slots, mask=mask)
        for s, r in zip(slots, reqs):
            self.busy[s] = r
        t4 = time.perf_counter()
        self.t["admit_stack"] += t1 - t0
        self.t["admit_edits"] += t2 - t1
        self.t["admit_fill"] += t3 - t2
        self.t["admit_warm"] += t4 - t3
        return k

    def round(self, k=1):
        t0 = time.perf_counter()
        self.batch.stream_round(k=k, max_iter=self.max_iter, tol=self.tol)
        self.t["round"] += time.perf_counter() - t0

    def retire(self, now):
        """Takes the finished slots' results out; returns the finished requests."""
        t0 = time.perf_counter()
        st, it, _ = self.batch.results()
        t1 = time.perf_counter()
        busy = np.fromiter(self.busy.keys(), dtype=np.int64, count=len(self.busy))
        fin = (st[busy] != -1) & (st[busy] != _capi.NOTSIMULATED)
        done = busy[fin].tolist()
        out = []
        self.t["retire_results"] += t1 - t0
        t2 = time.perf_counter()
        for c0 in range(0, len(done), self.stage_rows):
            part = done[c0:c0 + self.stage_rows]
            V = self.batch.stream_take(part, out=self.V_pin)
            reqs = [self.busy.pop(s) for s in part]
            res = self._results(reqs[0].payload)
            if res is not None and all(r.payload is reqs[0].payload for r in reqs):
                # one copy into the response buffer (same rows as the payloads);
                # each request's V is a view of its row
                rows = np.fromiter((r.row for r in reqs), dtype=np.int64, count=len(reqs))
                res[rows] = V[:len(reqs)]
                for r, s in zip(reqs, part):
                    r.status, r.iterations, r.V, r.t_done = int(st[s]), int(it[s]), res[r.row], now
            else:
                for q, (r, s) in enumerate(zip(reqs, part)):
                    r.status, r.iterations, r.V, r.t_done = int(st[s]), int(it[s]), V[q].copy(), now
            self.free.extend(part)
            out.extend(reqs)
        self.t["retire_take"] += time.perf_counter() - t2
        return out

    def _results(self, payload):
        """The response buffer of a payload buffer (same shape), made on first use."""
        if payload is None:
            return None
        key = id(payload)
        if key not in self.results:
            self.results[key] = (payload, np.empty_like(payload))
        return self.results[key][1]


class PowerFlowServer:
    """Requests over several resident grids.

    grids: name -> synthetic Grid; slots: name -> batch size (or one size for
    every grid, reduced to fit hbm_budget bytes); warm: name -> WarmStart or
    None.  ``run(requests)`` replays the requests at their arrival times
    (wall clock from the call) and returns the finished requests and a report
    with throughput and p50 / p99 latency."""

    def __init__(self, grids, slots=4096, precision=_capi.MIXED, warm=None, hbm_budget=64e9, k_round=1,
                 max_iter=10, tol=1e-8):
        self.k_round = k_round
        sizes = dict(slots) if isinstance(slots, dict) else {name: int(slots) for name in grids}
        # admission: the resident batches must fit the memory budget
        info = {name: _capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack).info() for name, g in grids.items()}
        while sum(batch_bytes(info[k], sizes[k], precision, True) for k in grids) > hbm_budget:
            big = max(sizes, key=lambda k: batch_bytes(info[k], sizes[k], precision, True))
            if sizes[big] <= 64:
                raise ValueError("the HBM budget does not hold one tile per grid")
            sizes[big] //= 2
        self.slots = sizes
        self.state = {name: GridState(name, g, sizes[name], precision, (warm or {}).get(name), max_iter, tol)
                      for name, g in grids.items()}

    def run(self, requests, timeout=600.0):
        reqs = sorted(requests, key=lambda r: r.t_arrival)
        t0 = time.perf_counter()
        nxt = 0
        done = []
        while len(done) < len(reqs):
            now = time.perf_counter() - t0
            if now > timeout:
                raise RuntimeError("serving loop timed out")
            while nxt < len(reqs) and reqs[nxt].t_arrival <= now:
                self.state[reqs[nxt].grid].queue.append(reqs[nxt])
                nxt += 1
            worked = False
            for st in self.state.values():
                st.admit()
                if st.busy:
                    st.round(self.k_round)
                    done.extend(st.retire(time.perf_counter() - t0))
                    worked = True
            if not worked and nxt < len(reqs):
                time.sleep(max(0.0, min(1e-3, reqs[nxt].t_arrival - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        lat = np.array([r.t_done - r.t_arrival for r in done])
        conv = sum(r.status == 0 for r in done)
        t_last = reqs[-1].t_arrival
        # while requests still arrive: results per second up to the last arrival
        steady = sum(r.t_done <= t_last for r in done) / t_last if t_last > 0 else None
        report = dict(requests=len(done), converged=int(conv), wall_s=wall,
                      throughput_rps=len(done) / wall, served_rps_while_arriving=steady,
                      offered_rps=len(reqs) / max(t_last, 1e-9),
                      p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                      mean_iterations=float(np.mean([r.iterations for r in done])), slots=self.slots,
                      host_s={f"{name}.{k}": round(v, 4) for name, st in self.state.items() for k, v in st.t.items()})
        for st in self.state.values():
            st.t.clear()
        return done, report


def _receive_buffer(S):
    """The grid's request payloads in one page-locked [rows, n] complex
    buffer, rows in arrival order: what a receiving front end would fill, and
    what the device reads by DMA without staging."""
    try:
        import torch
        if torch.cuda.is_available():
            buf = torch.empty(S.shape, dtype=torch.complex128, pin_memory=True).numpy()
            buf[...] = S
            return buf
    except ImportError:
        pass
    return np.ascontiguousarray(S)


def poisson_requests(grids, rate, n, max_outages=3, seed=5):
    """n requests at Poisson arrival times (rate per second) over the grids,
    chosen uniformly; each carries one scenario row of its grid and 0..
    max_outages random non-islanding branch outages."""
    rng = np.random.default_rng(seed)
    names = sorted(grids)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    pick = rng.integers(0, len(names), size=n)
    reqs = []
    for name_i, name in enumerate(names):
        idx = np.flatnonzero(pick == name_i)
        g = grids[name]
        p, q = syn.injection_scenarios(g, max(1, idx.size), seed=seed + 1 + name_i)
        S = (p + 1j * q) / g.sn_mva
        outs = syn.random_outages(g, max(1, idx.size), max_k=min(max_outages, 2), seed=seed + 11 + name_i)
        if max_outages > 2:  # a third outage on some rows, kept non-islanding
            extra = syn.random_outages(g, max(1, idx.size), max_k=1, seed=seed + 21 + name_i)
            outs = [sorted(set(a) | set(b)) if syn._connected_without(g, sorted(set(a) | set(b))) else a
                    for a, b in zip(outs, extra)]
        buf = _receive_buffer(S)
        for q_i, k in enumerate(idx):
            reqs.append(Request(rid=int(k), grid=name, S=buf[q_i], outages=list(outs[q_i]), t_arrival=float(t[k]),
                                payload=buf, row=q_i))
    reqs.sort(key=lambda r: r.t_arrival)
    return reqs
