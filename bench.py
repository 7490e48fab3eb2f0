#!/usr/bin/env python3
"""Converged power flows per second on one MI355X.

Workload: the 6,500-bus synthetic grid (9,000 branches, 12 % PV, 1 slack) with
16,384 injection scenarios (per-row lognormal(0, 0.1) load scaling times a
per-bus lognormal(0, 0.05), PV redispatched proportionally), Newton-Raphson
from a flat start to ||dS||inf < 1e-8 pu.  One step solves every row; the
injections are resident in HBM before the timed region.  Methods:

* ``fp64``  plain FP64 flat-start NR (the reference method);
* ``mixed`` FP32 Jacobian / LU / triangular solves, FP64 mismatch, test and
  state;
* ``mixed_ir`` mixed precision with one refinement round per Newton step whose
  residual is formed in FP64 from V, against the FP32 factors;
* ``warm``  learned warm start + mixed precision (when the MLP is available);
* ``warm_ir`` ``warm`` with ``mixed_ir``'s FP64-residual refinement round on
  every Newton step;
* ``facade_mixed`` / ``facade_warm``: ``mixed`` / ``warm`` through the
  drop-in Python surface (InjectionSweepGPU.compute), to show what the
  facade path costs (``vs_capi``).

Prints one JSON line.  ``value`` is the converged PF/s of the best method
measured (the headline method: warm start + mixed precision, else mixed
precision); every method's figure is in ``methods``.

``--workload c3`` runs the learned-warm-start workload instead: the 2,000-bus
grid, 16,384 rows of the same load scaling plus 0-2 random non-islanding
branch outages each (per-row Ybus value edits, and a 0/1 outage mask among
the network's inputs, resident in HBM like the injections), the warm start
trained on solver-generated solutions.  A row on its learned start falls back
to the base case's solution when its mismatch grows or after 6 updates.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged PF/s at ‖ΔS‖∞<1e-8 pu: warm-start+mixed vs plain FP64 NR"
HBM_PEAK_GBS = 8000.0
# HBM bytes one k_factor launch moves per factored row, measured with rocprofv3 PMC counters
# (FETCH_SIZE x 2 + WRITE_SIZE, scaled on a 1 GiB copy) on the 16,384-row sweep
# (profiles/hbm_traffic_c2_16384.md): 22.46 GB (FP64) and 11.24 GB (FP32) per full launch.
FACTOR_TRAFFIC_PER_ROW = {"fp64": 22.46e9 / 16384, "f32": 11.24e9 / 16384}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=16384)
    ap.add_argument("--max-iter", type=int, default=10)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--methods", default=None,
                    help="comma list (default: fp64,mixed,mixed_ir,warm plus, on c2, facade_mixed,facade_warm)")
    ap.add_argument("--train-rows", type=int, default=200_000,
                    help="warm start: solved training scenarios (the workload text's 200k solver solutions)")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--hidden", type=int, default=1024)
    ap.add_argument("--train-batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--k-warm", type=int, default=6, help="warm start: updates before the fallback")
    ap.add_argument("--growth", type=float, default=1.0, help="warm start: fall back when ||F|| grows (x this)")
    ap.add_argument("--contraction", type=float, default=0.0,
                    help="warm start: fall back from the second update on when ||F|| > q ||F_prev||^2 (0: off)")
    ap.add_argument("--fallback", choices=("flat", "base"), default="base",
                    help="warm start: the fallback start (flat start, or the base case's solution)")
    ap.add_argument("--clip", type=float, default=0.3, help="warm start: gradient-norm clip of the training")
    ap.add_argument("--cpu-rows", type=int, default=0, help="oracle sample rows (0: sized by --cpu-seconds)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="oracle sample's CPU time (0: skip)")
    ap.add_argument("--workload", choices=("c2", "c3", "c4"), default="c2")
    ap.add_argument("--rate", type=float, default=50000.0, help="c4: offered requests per second")
    ap.add_argument("--requests", type=int, default=100000, help="c4: requests replayed")
    ap.add_argument("--no-warm", action="store_true", help="c4: flat starts instead of the learned warm start")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="first copy this many GiB device to device (a dispatch of known size that scales "
                         "rocprofv3 PMC byte counters)")
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args()
    if args.methods is None:
        args.methods = "fp64,mixed,mixed_ir,warm,warm_ir" + (",facade_mixed,facade_warm" if args.workload == "c2" else "")
    return args


class Workload:
    """Grid, scenario rows (Sbus) and, for outage rows, the per-row Ybus edits
    and the outage mask."""

    def __init__(self, args, syn, rows, seed_s, seed_o):
        self.name = args.workload
        self.g = syn.grid_c2() if self.name == "c2" else syn.grid_c1()
        p, q = syn.injection_scenarios(self.g, rows, seed=seed_s)
        self.S = (p + 1j * q) / self.g.sn_mva
        self.outages = None
        self.edits = None
        self.mask = None
        if self.name == "c3":
            self.outages = syn.random_outages(self.g, rows, max_k=2, seed=seed_o)
            self.edits = syn.outage_edits(self.g, self.outages)
            self.mask = syn.outage_mask(self.g, self.outages)

    def describe(self):
        g = self.g
        if self.name == "c2":
            return ("6500-bus injection sweep, 16384 scenarios, ||dS||inf<1e-8",
                    "synthetic (seeded 6,500-bus grid, 16,384 lognormal injection scenarios)")
        return ("2000-bus sweep with 0-2 branch outages per row, 16384 rows, learned warm start, ||dS||inf<1e-8",
                "synthetic (seeded 2,000-bus grid, 16,384 lognormal injection scenarios, 0-2 random "
                "non-islanding outages per row)")


def log(args, *msg):
    if not args.quiet:
        print(*msg, file=sys.stderr, flush=True)


def train_warm(args, wl, torch, capi, syn):
    """Learned warm start: solve fresh training scenarios (seed 2; outages seed
    4) with the mixed-precision solver, then train the MLP on them -- both
    outside the timed region, like the symbolic analysis."""
    from gpusim2grid_amd.warmstart import WarmStart
    g = wl.g
    t0 = time.perf_counter()
    rows = args.train_rows
    tr = Workload(args, syn, rows, seed_s=2, seed_o=4)
    plan = capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    chunk = min(rows, 32768)
    S_all, V_all, M_all = [], [], []
    for r0 in range(0, rows, chunk):
        r1 = min(rows, r0 + chunk)
        S = torch.from_numpy(np.ascontiguousarray(tr.S[r0:r1])).to("cuda")
        B = capi.Batch(plan, r1 - r0, precision=capi.MIXED, per_row_ybus=tr.edits is not None)
        B.set_sbus_device(S.data_ptr(), S.shape[1], r1 - r0)
        B.set_v(g.flat_start())
        if tr.edits is not None:
            B.set_ybus_edits(*syn.outage_edits(g, tr.outages[r0:r1]))
        B.solve(max_iter=args.max_iter, tol=args.tol, scheduling=capi.CONTINUOUS)
        st, _, _ = B.results()
        V = torch.empty((r1 - r0, g.n_bus), dtype=torch.complex128, device="cuda")
        torch.cuda.synchronize()
        B.get_v_device(V.data_ptr(), g.n_bus)
        del B
        ok = torch.from_numpy(st == 0).to("cuda")
        S_all.append(S[ok])
        V_all.append(V[ok])
        if tr.mask is not None:
            M_all.append(torch.from_numpy(tr.mask[r0:r1]).to("cuda")[ok])
    S = torch.cat(S_all).contiguous()
    V = torch.cat(V_all).contiguous()
    mask = torch.cat(M_all).contiguous() if M_all else None
    del S_all, V_all, M_all, tr
    t_data = time.perf_counter() - t0
    ws = WarmStart(g.n_bus, g.pv, g.pq, g.slack, np.abs(g.flat_start()), n_hidden=args.hidden, seed=2,
                   n_mask=0 if mask is None else g.n_branch)
    t1 = time.perf_counter()
    hist = ws.train(S, V, mask=mask, epochs=args.epochs, batch=args.train_batch, lr=args.lr, clip=args.clip,
                    log=(lambda m: log(args, m)))
    lr = args.lr
    if not hist[-1] < 0.5 * hist[0]:
        log(args, f"warm-start training did not converge (loss {hist[0]:.3f} -> {hist[-1]:.3f})")
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    n_train = int(S.shape[0])
    del S, V
    rise = max([hist[i] / hist[i - 1] for i in range(4, len(hist))], default=None)
    return ws, dict(train_rows=n_train, epochs=args.epochs, hidden=args.hidden, lr=lr, clip=args.clip,
                    loss_first=hist[0], loss_last=hist[-1], max_epoch_rise_after_3=rise, data_s=t_data,
                    train_s=t_train)


def run_method(args, name, wl, torch, capi, ws=None):
    g, S = wl.g, wl.S
    prec = capi.FP64 if name == "fp64" else capi.MIXED
    ir = 1 if name in ("mixed_ir", "warm_ir") else 0  # one refinement round per Newton step, FP64 residual
    t_plan = time.perf_counter()
    plan = capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
    info = plan.info()
    t_plan = time.perf_counter() - t_plan
    B = capi.Batch(plan, args.rows, precision=prec, per_row_ybus=wl.edits is not None)
    if wl.edits is not None:
        B.set_ybus_edits(*wl.edits)
    if ir:
        B.set_step_control(ir_steps=ir)
        B.set_refinement_residual(True)
    # injections resident in HBM: one upload before any timing
    S_dev = torch.from_numpy(np.ascontiguousarray(S)).to("cuda")
    B.set_sbus_device(S_dev.data_ptr(), S.shape[1], args.rows)
    V0h = g.flat_start()
    if ws is not None and args.fallback == "base":
        # the fallback start of the warm method: the base case's solution
        # (the grid's own injections, no outage), solved once beforehand
        Pb = capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
        Bb = capi.Batch(Pb, 1)
        Bb.set_sbus(g.Sbus)
        Bb.set_v(g.flat_start())
        Bb.solve(max_iter=args.max_iter, tol=args.tol)
        V0h = Bb.get_v()[0]
        del Bb, Pb
    V0 = torch.from_numpy(V0h).to("cuda")
    if ws is not None:
        B.set_warm_contraction(args.contraction)
        if wl.mask is not None:
            B.set_warm_mask(wl.mask)  # the outage mask resident in HBM, like the injections
    torch.cuda.synchronize()

    def step():
        B.set_v_device(V0.data_ptr(), 0, args.rows)  # flat start (broadcast), inside the step
        if ws is not None:
            return B.solve_warm(ws.mlp, k_warm=args.k_warm, growth=args.growth, max_iter=args.max_iter,
                                tol=args.tol)
        return B.solve(max_iter=args.max_iter, tol=args.tol, scheduling=capi.CONTINUOUS)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st, it, res = B.results()
    conv = int((st == 0).sum())
    s_r = 8 if prec == capi.FP64 else 4
    alg_bytes = sum((info["nJ"] + info["nLU"]) * s_r * s["rows_factored"] for s in stats)
    launches = sum(s["n_factor_launches"] for s in stats)
    ms_factor = sum(s["ms_factor"] for s in stats)
    # the main-layout walk alone (whole-tile launches of the dominant kernel;
    # the split walk's tail launches are another kernel)
    n_main = sum(s["n_factor_main"] for s in stats)
    ms_main = sum(s["ms_factor_main"] for s in stats)
    rows_main = sum(s["rows_factored_main"] for s in stats)
    main_bytes = (info["nJ"] + info["nLU"]) * s_r * rows_main
    out = dict(
        method=name, pfps=conv * args.steps / dt, ms_per_step=1e3 * dt / args.steps,
        analysis_s=t_plan, pfps_incl_analysis=conv / (dt / args.steps + t_plan),
        converged=conv, rows=args.rows, iterations=stats[-1]["iterations"],
        mean_row_iters=float(it[st == 0].mean()) if conv else None,
        iters_hist=np.bincount(it[st == 0]).tolist() if conv else None,
        max_residual=float(res[st == 0].max()) if conv else None,
        ms_mismatch=np.mean([s["ms_mismatch"] for s in stats]), ms_factor=ms_factor / args.steps,
        ms_backward=np.mean([s["ms_backward"] for s in stats]),
        ms_warm_start=np.mean([s["ms_warm_start"] for s in stats]),
        n_requeued=stats[-1]["n_requeued"],
        factor_launch_ms=ms_factor / max(launches, 1),
        factor_alg_bytes_per_launch=alg_bytes / max(launches, 1),
        factor_achieved_gbs=alg_bytes / (ms_factor * 1e-3) / 1e9 if ms_factor > 0 else None,
        factor_rows_per_launch=sum(s["rows_factored"] for s in stats) / max(launches, 1),
        main_launches=n_main / args.steps,
        main_launch_ms=ms_main / max(n_main, 1),
        main_rows_per_launch=rows_main / max(n_main, 1),
        main_alg_bytes_per_launch=main_bytes / max(n_main, 1),
        main_achieved_gbs=main_bytes / (ms_main * 1e-3) / 1e9 if ms_main > 0 else None,
        real="fp64" if prec == capi.FP64 else "f32",
        info=info,
    )
    V = B.get_v()
    del B, S_dev, V0
    torch.cuda.synchronize()
    return out, V, st, it


def run_facade(args, name, wl, torch, ws=None):
    """The same rows through the drop-in Python surface: InjectionSweepGPU
    (array-tuple grid, set_injections in MW / MVAr, compute() returning the
    aliasing DLPack capsule), mixed precision, continuous scheduling; with the
    learned warm start (falling back to the base case) or from the flat
    start."""
    from gpusim2grid_amd.facades import InjectionSweepGPU
    g, S = wl.g, wl.S
    grid = (g.Ybus, g.flat_start(), g.Sbus, np.atleast_1d(g.slack), np.ones(np.atleast_1d(g.slack).size), g.pv,
            g.pq)
    f = InjectionSweepGPU(grid, init_from_n_powerflow=False, precision="mixed", nb_iter=args.max_iter,
                          scheduling="continuous", tol=args.tol)
    f.set_injections(S.real * g.sn_mva, S.imag * g.sn_mva, g.sn_mva)
    if ws is not None:
        f.set_warm_start(ws, k_warm=args.k_warm, growth=args.growth)
        f.set_v_init_from(None)  # the fallback: the base case
    else:
        f.set_v_init_from(g.flat_start())
    for _ in range(args.warmup):
        f.compute(batch_size=args.rows)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cap = f.compute(batch_size=args.rows)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    Vd = torch.from_dlpack(cap)  # aliases the session's results
    st = np.asarray(f.get_row_status())
    it = np.asarray(f.get_row_iterations())
    V = Vd.cpu().numpy().reshape(-1, g.n_bus)
    conv = int((st == 0).sum())
    out = dict(method=name, pfps=conv * args.steps / dt, ms_per_step=1e3 * dt / args.steps, converged=conv,
               rows=args.rows, mean_row_iters=float(it[st == 0].mean()) if conv else None,
               path="InjectionSweepGPU.compute (DLPack result)")
    del f
    return out, V, st, it


def cpu_baseline(args, wl, keep):
    """The oracle's C NR timed on the host's cores over a bounded sample of
    the workload's rows (about --cpu-seconds of work, sized from a pilot),
    with KLU-style refactorizations after each row's first factorization;
    and every GPU method's answers against it on that sample."""
    from oracle import oracle as orc
    g, S = wl.g, wl.S
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))

    def run(r0, r1):
        Yrows = None if wl.outages is None else orc.outage_ybus_rows(g, wl.outages[r0:r1])
        t0 = time.perf_counter()
        out = orc.nr_batch(g.Ybus, g.flat_start(), S[r0:r1], g.pv, g.pq, max_iter=args.max_iter, tol=args.tol,
                           n_threads=cores, Yx_rows=Yrows, refactor=True)
        return out, time.perf_counter() - t0

    # sized from pilots (a small one carries each thread's set-up, so a
    # second, larger one sets the rate)
    n = args.cpu_rows
    if not n:
        n = min(S.shape[0], 16 * cores)
        for _ in range(2):
            _, dt_p = run(0, n)
            n_next = max(n, min(S.shape[0], int(args.cpu_seconds * n / max(dt_p, 1e-6))))
            if n_next <= 2 * n:
                n = n_next
                break
            n = min(n_next, 8 * n)
    n = max(1, min(S.shape[0], n))
    (Vo, sto, ito, _), dt = run(0, n)
    conv = int((sto == 0).sum())
    parity = dict(rows=n)
    for k, (Vk, stk, itk) in keep.items():
        ok = (sto == 0) & (stk[:n] == 0)
        dv = np.abs(Vk[:n][ok] - Vo[ok]).max(axis=1) if ok.any() else np.zeros(0)
        parity[k] = dict(status_equal=bool(np.array_equal(sto, stk[:n])),
                         iters_equal=float((ito[ok] == itk[:n][ok]).mean()) if ok.any() else None,
                         mean_iters=float(itk[:n][ok].mean()) if ok.any() else None,
                         oracle_mean_iters=float(ito[ok].mean()) if ok.any() else None,
                         dV_p50=float(np.percentile(dv, 50)) if dv.size else None,
                         dV_p99=float(np.percentile(dv, 99)) if dv.size else None,
                         dV_max=float(dv.max()) if dv.size else None)
    cpu = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(value=conv / dt, unit="PF/s", cores=cores, kind="port", per_core=conv / dt / cores,
                label="own CPU NR (KLU-style refactor), not lightsim2grid", host_cpu=cpu,
                host_nproc=os.cpu_count(),
                sample=f"first {n} of the {S.shape[0]} rows, oracle C NR (min-degree order, partial-pivot sparse LU "
                       f"on each row's first iteration, refactorization on its pivots after), {cores} threads, "
                       f"{dt:.1f} s"), parity


def run_serving(args, torch, capi, syn):
    """The serving workload: Poisson arrivals over the 118-, 2,000- and 6,500-bus grids, 0-3
    tripped branches per request, resident per-grid batches (continuous
    batching with slot refill), warm start + mixed precision; throughput and
    p50 / p99 latency from arrival to result."""
    from gpusim2grid_amd import serving
    from gpusim2grid_amd.warmstart import WarmStart
    grids = {"g118": syn.grid_118(), "c1": syn.grid_c1(), "c2": syn.grid_c2()}
    warm = {}
    t0 = time.perf_counter()
    if not args.no_warm:
        for name, g in grids.items():
            rows = min(args.train_rows, 16384)
            p, q = syn.injection_scenarios(g, rows, seed=2)
            S = (p + 1j * q) / g.sn_mva
            outs = syn.random_outages(g, rows, max_k=2, seed=4)
            plan = capi.Plan.from_grid(g.Ybus, g.pv, g.pq, g.slack)
            B = capi.Batch(plan, rows, precision=capi.MIXED, per_row_ybus=True)
            B.set_sbus(S)
            B.set_v(g.flat_start())
            B.set_ybus_edits(*syn.outage_edits(g, outs))
            B.solve(max_iter=args.max_iter, tol=args.tol)
            st, _, _ = B.results()
            V = B.get_v()
            del B
            ok = st == 0
            ws = WarmStart(g.n_bus, g.pv, g.pq, g.slack, np.abs(g.flat_start()), n_hidden=args.hidden, seed=2,
                           n_mask=g.n_branch)
            ws.train(torch.as_tensor(S[ok], device="cuda"), torch.as_tensor(V[ok], device="cuda"),
                     mask=torch.as_tensor(syn.outage_mask(g, outs)[ok], device="cuda"), epochs=args.epochs,
                     batch=args.train_batch, lr=args.lr)
            warm[name] = ws
            log(args, f"c4: warm start of {name} trained")
    t_train = time.perf_counter() - t0
    reqs = serving.poisson_requests(grids, rate=args.rate, n=args.requests, max_outages=3, seed=5)
    srv = serving.PowerFlowServer(grids, slots={"g118": 4096, "c1": 8192, "c2": 16384}, precision=capi.MIXED,
                                  warm=warm or None, hbm_budget=200e9)
    # a short replay first (compiles nothing, but pages every kernel in)
    srv.run(serving.poisson_requests(grids, rate=args.rate, n=2000, max_outages=3, seed=99), timeout=300)
    done, rep = srv.run(reqs, timeout=600)
    # capacity: the same requests, all queued at once
    sat = [dataclasses.replace(r, t_arrival=0.0, t_done=0.0, status=-1, iterations=0, V=None) for r in reqs]
    _, cap = srv.run(sat, timeout=600)
    line = dict(metric="served power flows per second, 118/2000/6500-bus request mix (every request queued at "
                       "once; latency at the offered Poisson rate beside it)",
                value=cap["throughput_rps"], unit="PF/s", n_gpus=1, steps=1, warmup=1,
                ms_per_step=cap["wall_s"] * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None,
                dtype="f32+f64", data="synthetic (seeded grids, Poisson arrivals, 0-3 outages per request)",
                config=dict(workload="serving loop, 118/2000/6500-bus grids", offered_rps=args.rate,
                            requests=args.requests, warm_start=not args.no_warm),
                serving=rep, saturated=cap, train_s=t_train)
    print(json.dumps(line, default=float))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from gpusim2grid_amd import _capi as capi
    from gpusim2grid_amd import synthetic as syn
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(2)
    capi.lib()  # fails loudly when the HIP library is missing
    if args.copy_gib > 0:
        nb = int(args.copy_gib * (1 << 30))
        src = torch.empty(nb, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(args, f"copied {nb} bytes device to device")
        del src, dst
    if args.workload == "c4":
        run_serving(args, torch, capi, syn)
        return
    t0 = time.perf_counter()
    wl = Workload(args, syn, args.rows, seed_s=1, seed_o=3)
    g = wl.g
    log(args, f"grid + scenarios {time.perf_counter() - t0:.1f} s")
    methods = {}
    keep = {}
    warm_info = None
    names = args.methods.split(",")
    ws_trained = None
    for name in names:
        t1 = time.perf_counter()
        ws = None
        if name in ("warm", "warm_ir", "facade_warm"):
            if ws_trained is None:
                ws_trained, warm_info = train_warm(args, wl, torch, capi, syn)
                log(args, "warm-start training", warm_info)
            ws = ws_trained
        if name.startswith("facade_"):
            if args.workload != "c2":
                raise SystemExit("the facade methods run the injection sweep (--workload c2)")
            methods[name], V, st, it = run_facade(args, name, wl, torch, ws)
            base = methods.get(name[len("facade_"):])
            if base:
                methods[name]["vs_capi"] = methods[name]["pfps"] / base["pfps"]
        else:
            methods[name], V, st, it = run_method(args, name, wl, torch, capi, ws)
        if name == "warm":
            methods[name]["training"] = warm_info
        keep[name] = (V, st, it)
        log(args, f"{name}: {methods[name]['pfps']:.0f} PF/s ({time.perf_counter() - t1:.1f} s)",
            json.dumps({k: v for k, v in methods[name].items() if k != 'info'}))
    cpu, parity = (None, None)
    if args.cpu_seconds > 0:
        cpu, parity = cpu_baseline(args, wl, keep)
        log(args, "cpu", cpu, parity)
    # every method's answers against the plain FP64 method's (same rows)
    if "fp64" in keep:
        V64, st64, _ = keep["fp64"]
        for k, (Vk, stk, _) in keep.items():
            if k == "fp64":
                continue
            ok = (stk == 0) & (st64 == 0)
            methods[k]["vs_fp64"] = dict(status_equal=bool(np.array_equal(stk, st64)),
                                         max_dV=float(np.abs(Vk[ok] - V64[ok]).max()) if ok.any() else None)
    head = (methods.get("warm") or methods.get("mixed") or methods.get("fp64")
            or next(v for k, v in methods.items() if not k.startswith("facade_")))
    fp64 = methods.get("fp64")
    dom = head
    # the dominant kernel: the main-layout factorization walk (k_factor over
    # whole tiles); every factorization of the method, the split walk's tails
    # included, beside it
    roof = dict(bound="hbm", kernel="k_factor (main-layout walk)",
                achieved=dom["main_achieved_gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                frac=(dom["main_achieved_gbs"] or 0.0) / HBM_PEAK_GBS,
                traffic=(FACTOR_TRAFFIC_PER_ROW[dom["real"]] * dom["main_rows_per_launch"]
                         if args.workload == "c2" else None),
                traffic_source=("rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, on full k_factor launches "
                                "(profiles/hbm_traffic_c2_16384.md): bytes per factored row x this run's "
                                "rows per launch" if args.workload == "c2" else None),
                alg_bytes_per_launch=dom["main_alg_bytes_per_launch"], launch_ms=dom["main_launch_ms"],
                launches_per_step=dom["main_launches"],
                all_factorizations=dict(achieved=dom["factor_achieved_gbs"],
                                        frac=(dom["factor_achieved_gbs"] or 0.0) / HBM_PEAK_GBS,
                                        launch_ms=dom["factor_launch_ms"],
                                        alg_bytes_per_launch=dom["factor_alg_bytes_per_launch"]))
    line = dict(
        metric=METRIC, value=head["pfps"], unit="PF/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=head["ms_per_step"], higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64" if head["method"] == "fp64" else "f32+f64",
        data=wl.describe()[1],
        config=dict(workload=wl.describe()[0], n_bus=g.n_bus,
                    n_branch=g.n_branch, rows=args.rows, method=head["method"], max_iter=args.max_iter),
        ratio_vs_plain_fp64=(head["pfps"] / fp64["pfps"]) if fp64 else None,
        methods={k: {kk: vv for kk, vv in v.items() if kk != "info"} for k, v in methods.items()},
        roofline=roof, cpu_baseline=cpu, parity_vs_oracle=parity,
    )
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
