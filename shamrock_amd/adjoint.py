"""Differentiable SPH rollout: torch.autograd.Function over the GPU solver and its hand-written
adjoint (shamx_solver_adjoint_step / _density, shamrock_amd/csrc/adjoint.hip).

`rollout_density(cfg, fields, v0, u0, steps)` runs `steps` evolve_once from the state `fields`
with initial velocities v0 (N x 3) and internal energies u0 (N) and returns the per-particle
density rho_h(h) of the final state; gradients flow to v0 and u0.  `rollout_state` returns the
final density and internal energy (for losses on both, e.g. an entropy sum).  Sinks (free
boundaries) are part of the differentiated state: their seeds are carried by the solver.  The time steps are those the
solver chooses in the forward pass (or `dts` when given) and are held fixed when
differentiating.  A step whose corrector reran is differentiated through every pass (pass i
evaluates the force from the velocities and u of pass i - 1; the solver keeps them).
cfg["adjoint_fp32"] = 1 computes the adjoint's pair terms in FP32 (Kahan sums, FP64 r_ab);
cfg["mixed_precision"] = 1 runs the forward steps in mixed precision, whose map is then
differentiated as the FP64 step at the mixed trajectory's states.  Memory: one state every `checkpoint_every` steps plus one segment; the
backward pass replays each segment from its checkpoint (the step is deterministic, so replays
are bit-identical) and runs the adjoint of every step on the replayed intermediate products;
the last step of a segment is run once (its products are the solver's when the replay of the
segment ends), and the last step of the rollout not at all (the forward pass left them).

This extension has no counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

from ._lib import KERNEL_IDS
from .setups import KERNEL_HFACT
from .solver import FIELDS, GpuSolver

_KERNEL_NAMES = {v: k for k, v in KERNEL_IDS.items()}


def _torch():
    import torch

    return torch


def _density(cfg, h):
    kernel = cfg["kernel"]
    hf = KERNEL_HFACT[kernel if isinstance(kernel, str) else _KERNEL_NAMES[int(kernel)]]
    return cfg["gpart_mass"] * (hf / h) ** 3


def _step(solver, dt, passes=None):
    """One step (of prescribed dt when given); returns (dt, corrector passes).  A replay must
    take the passes of the step it repeats."""
    if dt is not None:
        solver.set_time(dt=dt)
    log = solver.evolve_once()
    if passes is not None and log["corrector_iterations"] != passes:
        raise RuntimeError("rollout: a replayed step took another number of corrector passes "
                           "than the forward pass")
    return log["dt"], log["corrector_iterations"]


class _Rollout:
    def __init__(self, cfg, fields, steps, checkpoint_every, dts, store_products=True):
        # the adjoint reads the step's products in the caller's storage order
        self.cfg, self.fields, self.steps = dict(cfg, keep_storage_order=1), fields, int(steps)
        self.k = max(1, int(checkpoint_every))
        self.dts = None if dts is None else [float(d) for d in dts]
        # keep every step's products of a segment (the solver's product store) so that the
        # backward pass runs each step once; without it, each step but the last of a segment
        # runs twice (once to lay down the segment's states, once before its adjoint)
        self.store = bool(store_products)
        self.last_c0 = ((self.steps - 1) // self.k) * self.k if self.steps else 0

    def forward(self, v0, u0):
        torch = _torch()
        base = {k: np.asarray(v, np.float64) for k, v in self.fields.items()}
        s = GpuSolver(self.cfg, base)
        for i, k in enumerate(("vx", "vy", "vz")):
            s.set_device(k, v0[:, i].detach().contiguous())
        s.set_device("u", u0.detach().contiguous())
        self.ckpt, used, passes = {}, [], []
        for n in range(self.steps):
            if n % self.k == 0:
                self.ckpt[n] = s.snapshot()
            dt, p = _step(s, None if self.dts is None else self.dts[n])
            used.append(dt)
            passes.append(p)
            if self.store and n >= self.last_c0:
                s.products_save(n - self.last_c0)
        self.dts_used, self.passes_used = used, passes
        self.solver = s
        rho = _density(self.cfg, s.get_device("h"))
        u = s.get_device("u")
        torch.cuda.synchronize()
        return rho, u

    def backward(self, rho_bar, u_bar=None):
        torch = _torch()
        s, N = self.solver, self.solver.n
        seeds = [torch.zeros(N, dtype=torch.float64, device="cuda") for _ in range(11)]
        if rho_bar is None:
            rho_bar = torch.zeros(N, dtype=torch.float64, device="cuda")
        # starts the backward pass (the sinks' seeds are set to zero)
        s.adjoint_density(rho_bar.contiguous(), seeds[0], seeds[1], seeds[2])
        if u_bar is not None:
            seeds[6] += u_bar
        if self.store:
            for c0 in sorted(self.ckpt, reverse=True):
                seg = list(range(c0, min(c0 + self.k, self.steps)))
                if c0 != self.last_c0:  # the forward pass kept the last segment's products
                    s.restore(self.ckpt[c0])
                    for n in seg:
                        _step(s, self.dts_used[n], self.passes_used[n])
                        s.products_save(n - c0)
                for n in reversed(seg):
                    s.products_load(n - c0)
                    s.adjoint_step(seeds)
            s.products_free()
            torch.cuda.synchronize()
            vbar = torch.stack(seeds[3:6], dim=1)
            return vbar, seeds[6], torch.stack(seeds[0:3], dim=1)
        first = True
        for c0 in sorted(self.ckpt, reverse=True):
            seg = list(range(c0, min(c0 + self.k, self.steps)))
            if first:
                # the forward pass has just run the last step: its products are in place
                s.adjoint_step(seeds)
                seg, first = seg[:-1], False
                if not seg:
                    continue
            s.restore(self.ckpt[c0])
            starts = []
            for n in seg[:-1]:
                starts.append(s.snapshot())
                _step(s, self.dts_used[n], self.passes_used[n])
            # the last step of the segment runs once, and its adjoint reads its products
            _step(s, self.dts_used[seg[-1]], self.passes_used[seg[-1]])
            s.adjoint_step(seeds)
            for n, st in zip(reversed(seg[:-1]), reversed(starts)):
                s.restore(st)
                # replay: the adjoint reads this step's products
                _step(s, self.dts_used[n], self.passes_used[n])
                s.adjoint_step(seeds)
        torch.cuda.synchronize()
        vbar = torch.stack(seeds[3:6], dim=1)
        return vbar, seeds[6], torch.stack(seeds[0:3], dim=1)


def _make_function():
    torch = _torch()

    class RolloutState(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v0, u0, cfg, fields, steps, checkpoint_every, dts, store_products):
            r = _Rollout(cfg, fields, steps, checkpoint_every, dts, store_products)
            rho, u = r.forward(v0, u0)
            ctx.rollout = r
            return rho, u

        @staticmethod
        def backward(ctx, rho_bar, u_bar):
            vbar, ubar, _ = ctx.rollout.backward(rho_bar, u_bar)
            ctx.rollout.solver.close()
            return vbar, ubar, None, None, None, None, None, None

    return RolloutState


_FN = None


def rollout_state(cfg, fields, v0, u0, steps, checkpoint_every=5, dts=None,
                  store_products=True):
    """Final-state density and internal energy after `steps` steps, differentiable w.r.t. v0
    (N x 3) and u0 (N)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(v0, u0, cfg, fields, steps, checkpoint_every, dts, store_products)


def rollout_density(cfg, fields, v0, u0, steps, checkpoint_every=5, dts=None,
                    store_products=True):
    """Final-state density after `steps` steps, differentiable w.r.t. v0 (N x 3) and u0 (N)."""
    return rollout_state(cfg, fields, v0, u0, steps, checkpoint_every, dts, store_products)[0]


def rollout_dts(cfg, fields, v0, u0, steps):
    """The time steps the solver takes from this state (to hold them fixed in comparisons)."""
    r = _Rollout(cfg, fields, steps, steps, None, store_products=False)
    r.forward(v0, u0)
    r.solver.close()
    return r.dts_used


def rollout_passes(cfg, fields, v0, u0, steps, dts=None):
    """The corrector passes (1 + reruns) of every step of the rollout."""
    r = _Rollout(cfg, fields, steps, steps, dts, store_products=False)
    r.forward(v0, u0)
    r.solver.close()
    return r.passes_used


__all__ = ["rollout_density", "rollout_state", "rollout_dts", "rollout_passes", "FIELDS"]
