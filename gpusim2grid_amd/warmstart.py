"""Learned warm start: training the MLP on solver solutions and solving from
its predictions with a flat-start fallback.

The reference starts every row from the base case's converged V or from a
caller's V (``src/_cpp/python_bindings.cpp:2722-2736``); it has no learned
start.  Here an MLP maps a row's per-bus injections (P, Q per unit, plus an
optional 0/1 outage mask) to its voltage angles and magnitudes.  The device
side is ``g2g_mlp_*`` / ``g2g_batch_solve_warm``; this module prepares the
training data on the device (torch tensors used as device buffers), keeps the
normalisation, and drives the training loop.
"""
from __future__ import annotations

import numpy as np

from . import _capi

PV, PQ, SLACK = 0, 1, 2


def bus_kinds(n, pv, pq, slack=None):
    """PV, PQ, or SLACK for every bus in neither list (a fixed bus: it keeps
    the row's start V; `slack` is implied and only checked)."""
    kind = np.full(n, SLACK, np.int32)
    kind[np.asarray(pv, np.int64)] = PV
    kind[np.asarray(pq, np.int64)] = PQ
    if slack is not None and (kind[np.atleast_1d(np.asarray(slack, np.int64))] != SLACK).any():
        raise ValueError("a slack bus is listed in pv or pq")
    return kind


class WarmStart:
    """An MLP warm start for one grid.

    n: buses; pv, pq, slack: bus indices; vset: |V| set-points (the flat
    start's magnitudes, used at PV buses; fixed buses keep the row's start V);
    n_mask: outage-mask
    width (0: injections only).
    """

    def __init__(self, n, pv, pq, slack, vset, n_hidden=1024, n_mask=0, seed=0, stream=None):
        self.n = int(n)
        self.n_mask = int(n_mask)
        self.n_in = 2 * self.n + self.n_mask
        self.n_out = 2 * self.n
        self.mlp = _capi.Mlp(self.n_in, n_hidden, self.n_out, seed=seed, stream=stream)
        self.kind = bus_kinds(self.n, pv, pq, slack)
        self.vset = np.asarray(vset, np.float64)
        self.mlp.set_buses(self.kind, self.vset)
        self.in_mu = self.in_sigma = self.out_mu = self.out_sigma = None

    # -- data ------------------------------------------------------------
    @staticmethod
    def _interleave(a, b):
        out = a.new_empty((a.shape[0], 2 * a.shape[1]))
        out[:, 0::2] = a
        out[:, 1::2] = b
        return out

    def features(self, S, mask=None):
        """S: complex128 tensor [rows, n] (device) -> raw features [rows, n_in] float64."""
        F = self._interleave(S.real, S.imag)
        if self.n_mask:
            import torch
            F = torch.cat([F, mask.to(F.dtype)], dim=1)
        return F

    def targets(self, V):
        return self._interleave(torch_angle(V), V.abs())

    def fit_normalization(self, S, V, mask=None):
        """Feature / target mean and spread from training rows (device tensors)."""
        F = self.features(S, mask)
        T = self.targets(V)

        def ms(x):
            mu = x.mean(0)
            sd = x.std(0)
            sd = sd.where(sd > 1e-9, sd.new_ones(()))
            return mu.float().cpu().numpy(), sd.float().cpu().numpy()

        self.in_mu, self.in_sigma = ms(F)
        if self.n_mask:  # the mask enters raw (0/1)
            self.in_mu[2 * self.n:] = 0.0
            self.in_sigma[2 * self.n:] = 1.0
        # outputs: one scale for all angles and one for all magnitudes, so the
        # loss weighs every bus by its absolute voltage error (what Newton's
        # convergence depends on), not by its own spread
        self.out_mu, _ = ms(T)
        sd = T.std(0)
        self.out_sigma = np.empty(self.n_out, np.float32)
        self.out_sigma[0::2] = float(sd[0::2].pow(2).mean().sqrt())
        self.out_sigma[1::2] = float(sd[1::2].pow(2).mean().sqrt())
        self.out_sigma[self.out_sigma <= 1e-12] = 1.0
        self.mlp.set_normalization(self.in_mu, self.in_sigma, self.out_mu, self.out_sigma)

    def _normalised(self, S, V, mask=None):
        import torch
        dev = S.device
        X = ((self.features(S, mask) - torch.as_tensor(self.in_mu, device=dev, dtype=torch.float64))
             / torch.as_tensor(self.in_sigma, device=dev, dtype=torch.float64)).float().contiguous()
        T = ((self.targets(V) - torch.as_tensor(self.out_mu, device=dev, dtype=torch.float64))
             / torch.as_tensor(self.out_sigma, device=dev, dtype=torch.float64)).float().contiguous()
        return X, T

    # -- training ----------------------------------------------------------
    def train(self, S, V, mask=None, epochs=20, batch=1024, lr=1e-3, lr_final=1e-5, seed=0, log=None,
              warmup=0.05, clip=0.3):
        """Adam on (features(S), targets(V)): the step size rises linearly over
        the first ``warmup`` of the steps, then decays along a cosine from lr to
        lr_final; every gradient is clipped to norm ``clip`` (0: off).  S, V:
        device complex128 tensors of converged training rows.  Returns the
        per-epoch mean losses."""
        import math
        import torch
        if self.in_mu is None:
            self.fit_normalization(S, V, mask)
        X, T = self._normalised(S, V, mask)
        rows = X.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        steps = max(1, epochs * (rows // batch))
        n_warm = int(warmup * steps)
        self.mlp.set_grad_clip(clip)
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(rows, generator=g).to(X.device)
            Xs, Ts = X[perm].contiguous(), T[perm].contiguous()
            torch.cuda.synchronize()
            tot, cnt = 0.0, 0
            for r0 in range(0, rows - batch + 1, batch):
                xp = Xs.data_ptr() + r0 * self.n_in * 4
                tp = Ts.data_ptr() + r0 * self.n_out * 4
                if k < n_warm:
                    lr_k = lr * (k + 1) / n_warm
                else:
                    t = (k - n_warm) / max(1, steps - n_warm)
                    lr_k = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * t))
                tot += self.mlp.train_step_device(xp, tp, batch, lr_k)
                cnt += 1
                k += 1
            hist.append(tot / max(cnt, 1))
            if log:
                log(f"warm-start epoch {ep}: loss {hist[-1]:.3e}")
        return hist

    def predict_v(self, S, mask=None):
        """Start voltages predicted for S (host complex [rows, n]) -> host complex."""
        import torch
        St = torch.as_tensor(np.atleast_2d(S), device="cuda")
        if mask is not None:
            mask = torch.as_tensor(np.atleast_2d(mask), device="cuda")
        X, _ = self._normalised(St, St.new_ones(St.shape), mask)
        Y = torch.empty((X.shape[0], self.n_out), device="cuda", dtype=torch.float32)
        torch.cuda.synchronize()
        self.mlp.forward_device(X.data_ptr(), X.shape[0], Y.data_ptr())
        Yd = Y.double().cpu().numpy() * self.out_sigma + self.out_mu
        va, vm = Yd[:, 0::2], Yd[:, 1::2]
        vm[:, self.kind != PQ] = self.vset[self.kind != PQ]
        va[:, self.kind == SLACK] = 0.0
        return vm * np.exp(1j * va)


def torch_angle(V):
    import torch
    return torch.atan2(V.imag, V.real)
