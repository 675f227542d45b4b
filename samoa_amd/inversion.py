"""Tsunami source inversion from gauge records, many instances in lockstep.

Each instance is a frozen tsunami scenario (uniform mesh, bathymetry b) whose
initial surface is an elliptical Gaussian source with parameters
theta = (amplitude [m], axis ratio, x0 [m], y0 [m]) and a semi-major axis of
100 km, as the Fortran scenario builds it (samoa_amd/fortran/swe_scenarios.f90
tsunami_state):

    eta0 = amplitude * exp(-((x - x0)/a)^2 - ((y - y0)/(a/ratio))^2)  (wet cells)
    eta0 = min(b, 0)                                                    (dry cells)

The gauge series of the instances come from batched forward sweeps and
their gradients from batched adjoint sweeps (PooledGaugeInversion: each
instance has its own checkpoint pool, store-all or sqrt(N) checkpoints,
admitted against an HBM budget; the instances of one mode and admission
round share one launch per step); the loss of instance s is the sum over
steps and its gauges of (eta - eta_target)^2.  Because instances are
independent, the gradient of the summed loss with respect to theta_s is the
gradient of instance s's own loss, and `lbfgs` runs one L-BFGS (two-loop
recursion, Armijo backtracking) per instance, in lockstep: every function
evaluation advances all instances together.
"""
from __future__ import annotations

import numpy as np
import torch

from .adjoint import PooledGaugeInversion

SEMI_MAJOR = 1.0e5                     # metres
SCALE = (1.0, 1.0, 1.0e5, 1.0e5)      # optimisation units of (amp, ratio, x0, y0)


class SourceInversion:
    """`meshes[s]` (uniform tsunami meshes of the 1000 km square), bathymetry
    `b[s]` (numpy), gauges (cell indices per instance), n_steps, dt0[s]."""

    def __init__(self, meshes, b, gauges, n_steps, dt0, *, scaling=1.0e6, dry_tolerance=1e-3,
                 courant=0.45, checkpoint_every="auto", hbm_budget=None, device="cuda:0"):
        self.device = torch.device(device)
        self.n_inst = len(meshes)
        self.scaling = float(scaling)
        self.drytol = float(dry_tolerance)
        offs = np.cumsum([0] + [m.n for m in meshes])
        self.offsets = offs
        params = [dict(scaling=scaling, dry_tolerance=dry_tolerance, courant=courant)
                  for _ in meshes]
        self.gauge_counts = [len(g) for g in gauges]
        self.problem = PooledGaugeInversion(list(meshes), gauges, n_steps, params,
                                            checkpoint_every=checkpoint_every,
                                            hbm_budget=hbm_budget, device=device)
        dev = self.device
        self.b = torch.tensor(np.concatenate(b), device=dev)
        cen = np.concatenate([m.centroids() for m in meshes]) * scaling
        self.x = torch.tensor(cen[:, 0], device=dev)
        self.y = torch.tensor(cen[:, 1], device=dev)
        self.inst = torch.tensor(np.repeat(np.arange(self.n_inst), np.diff(offs)), device=dev)
        self.gauge_inst = torch.tensor(np.repeat(np.arange(self.n_inst), self.gauge_counts),
                                       device=dev)
        self.dt0 = np.asarray(dt0, np.float64)
        self.scale = torch.tensor(SCALE, dtype=torch.float64, device=dev)
        self.target = None

    def eta0(self, theta: torch.Tensor) -> torch.Tensor:
        """Initial surface of every instance (concatenated); theta [n_inst, 4]."""
        p = theta[self.inst]
        amp, ratio, x0, y0 = p[:, 0], p[:, 1], p[:, 2], p[:, 3]
        a = SEMI_MAJOR
        c = a / ratio
        src = amp * torch.exp(-(((self.x - x0) / a) ** 2 + ((self.y - y0) / c) ** 2))
        dry = src.detach() < self.b + self.drytol
        return torch.where(dry, torch.clamp(self.b, max=0.0), src)

    def set_target(self, theta_true: torch.Tensor) -> None:
        """Freezes the dt sequence on the true sources and records their gauge
        series as the observations."""
        with torch.no_grad():
            e0 = self.eta0(theta_true)
            self.problem.record_dt(e0, self.b, self.dt0)
            self.target = self.problem.gauge_series(e0, self.b).detach()

    def loss_and_grad(self, z: torch.Tensor):
        """Per-instance loss [n_inst] and gradient [n_inst, 4] in optimisation
        units (theta = z * SCALE), admission round after admission round."""

        def make_inputs():
            zz = z.detach().requires_grad_()
            return self.eta0(zz * self.scale), self.b, [zz]

        def loss_fn(G, cols, members):
            r2 = ((G - self.target[:, cols]) ** 2).sum(0)    # per gauge
            per = torch.zeros(self.n_inst, dtype=torch.float64, device=self.device)
            per = per.index_add(0, self.gauge_inst[cols], r2)
            return per[members]

        losses, (gz,) = self.problem.forward_backward(make_inputs, loss_fn)
        return losses, gz.detach()

    def close(self):
        self.problem.close()


def lbfgs(fg, z0: torch.Tensor, iters: int = 20, memory: int = 6, c1: float = 1e-4,
          max_halvings: int = 12):
    """Lockstep L-BFGS over a batch of independent problems.

    fg(z [B, P]) -> (f [B], g [B, P]).  Returns (z, f, history of f)."""
    z = z0.clone()
    f, g = fg(z)
    S, Y, R = [], [], []
    hist = [f.clone()]
    for _ in range(iters):
        q = g.clone()
        alphas = []
        for s, y, rho in reversed(list(zip(S, Y, R))):
            a = rho * (s * q).sum(1)
            q = q - a[:, None] * y
            alphas.append(a)
        if S:
            sy = (S[-1] * Y[-1]).sum(1)
            yy = (Y[-1] * Y[-1]).sum(1)
            gamma = torch.where(sy > 0, sy / torch.where(yy > 0, yy, torch.ones_like(yy)),
                                torch.ones_like(sy))
        else:
            gamma = 1.0 / torch.clamp(g.norm(dim=1), min=1e-300)
        r = gamma[:, None] * q
        for (s, y, rho), a in zip(zip(S, Y, R), reversed(alphas)):
            bta = rho * (y * r).sum(1)
            r = r + s * (a - bta)[:, None]
        d = -r
        slope = (g * d).sum(1)
        bad = ~(slope < 0)                       # not a descent direction: steepest descent
        d = torch.where(bad[:, None], -g * gamma[:, None], d)
        slope = (g * d).sum(1)
        # Armijo backtracking in lockstep: an instance keeps its step once it
        # satisfies the condition; one that never does stays where it is.
        step = torch.ones_like(f)
        done = torch.zeros_like(f, dtype=torch.bool)
        for _ in range(max_halvings):
            z_new = z + step[:, None] * d
            f_new, g_new = fg(z_new)
            done = done | (f_new <= f + c1 * step * slope)
            if bool(done.all()):
                break
            step = torch.where(done, step, 0.5 * step)
        if not bool(done.all()):
            step = torch.where(done, step, torch.zeros_like(step))
            z_new = z + step[:, None] * d
            f_new, g_new = fg(z_new)
        s_k, y_k = z_new - z, g_new - g
        sy = (s_k * y_k).sum(1)
        rho = torch.where(sy > 1e-300, 1.0 / torch.where(sy > 1e-300, sy, torch.ones_like(sy)),
                          torch.zeros_like(sy))
        S.append(s_k)
        Y.append(y_k)
        R.append(rho)
        if len(S) > memory:
            S.pop(0), Y.pop(0), R.pop(0)
        z, f, g = z_new, f_new, g_new
        hist.append(f.clone())
    return z, f, hist


def gauges_near(mesh, x0: float, y0: float, radius: float, count: int, seed: int,
                scaling: float = 1.0e6) -> np.ndarray:
    """`count` distinct seeded cells whose centroids lie within `radius`
    metres of (x0, y0): gauges that record a source there from the first
    step on."""
    cen = mesh.centroids() * scaling
    near = np.flatnonzero(np.hypot(cen[:, 0] - x0, cen[:, 1] - y0) <= radius)
    if len(near) < count:
        raise ValueError(f"only {len(near)} cells within {radius} m of the source")
    return np.sort(np.random.default_rng(seed).choice(near, count, replace=False))
