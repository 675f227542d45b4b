"""The differentiable inversions on their own workloads (the tsunami scenario:
1000 km square, GRF bathymetry with a coastal shelf, elliptical source, dry
tolerance 1e-3, augmented Riemann solver, outflow), at sizes the CPU
restatement differentiates in seconds:

- the single gauge inversion: squared eta misfit at 32 gauges near the
  source every step against a target run from a perturbed source, over a
  frozen dt sequence; gradients with respect to eta0 and b against
  torch.autograd of the FP64 restatement (tests/torch_ref.py, bit-identical
  to the oracle) and against central finite differences on seeded random
  directions;
- the batched source inversions with per-instance checkpoint pools admitted
  by an HBM budget small enough to force both tape modes and two admission
  rounds: gradients with respect to the source parameters against autograd
  of each instance's own chain, and equal bit for bit to a batch with every
  instance's states kept.

Tolerances: the adjoint sums the same partial derivatives as autograd in a
different order, so agreement is to rounding, 1e-9 of the largest entry;
finite differences to 1e-6 relative."""
import math

import numpy as np
import pytest
import torch

import pyoracle as po
import torch_ref as tr
from samoa_amd import host
from samoa_amd.adjoint import GaugeInversion, PooledGaugeInversion, tape_bytes
from samoa_amd.inversion import SCALE, SourceInversion, gauges_near

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
S, DRY = 1.0e6, 1e-3


def torch_mesh(m):
    normals = np.stack([po.normals(p) if p else np.zeros((3, 2)) for p in range(-8, 9)])
    return tr.TorchMesh(m.depth, m.plotter, m.nbr, normals, po.leg_sizes())


def torch_series(mesh, eta0, b, dts, gauges):
    """Gauge series of the restatement, forward0 = False, alternating."""
    tm = torch_mesh(mesh)
    z = torch.zeros_like(eta0)
    q, out = (eta0, z, z), []
    for n, dt in enumerate(dts):
        e, u, v, _ = tr.step(tm, *q, b, float(dt), n % 2 == 1, S, po.G, DRY, bc=0, solver=0)
        q = (e, u, v)
        out.append(e[torch.as_tensor(gauges)])
    return torch.stack(out)


def rel_err(x, y):
    x = x.detach().cpu().numpy()
    y = y.detach().cpu().numpy()
    return float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300))


def tsunami_problem(d, seed, n_steps):
    mesh = host.mesh_uniform(d)
    eta, hu, hv, b = host.tsunami(mesh, seed=seed, drytol=DRY)
    p = host.tsunami_draw(seed)
    eta_t = host.tsunami(mesh, seed=seed, drytol=DRY, amplitude=1.25 * p["amplitude"])[0]
    gauges = gauges_near(mesh, p["x0"], p["y0"], 2.5e5, 32, seed=2024)
    dt0 = host.initial_dt(mesh, (eta, hu, hv, b), 0.45)
    return mesh, eta, b, eta_t, gauges, dt0


def test_tsunami_gauge_inversion_gradient_matches_autograd_and_fd():
    d, seed, N = 9, 0, 24
    print("case: tsunami inversion d", d, "steps", N, flush=True)
    mesh, eta, b, eta_t, gauges, dt0 = tsunami_problem(d, seed, N)
    prob = GaugeInversion(mesh, gauges, N, dict(scaling=S, dry_tolerance=DRY, courant=0.45),
                          checkpoint_every=math.ceil(math.sqrt(N)))
    e0 = torch.tensor(eta, device=DEV, requires_grad=True)
    bb = torch.tensor(b, device=DEV, requires_grad=True)
    # the first dt lacks the domain scaling, as the reference's does
    dts = prob.record_dt(e0.detach(), bb.detach(), dt0)[:, 0]
    target = prob.gauge_series(torch.tensor(eta_t, device=DEV), bb.detach()).detach()
    L = ((prob.gauge_series(e0, bb) - target) ** 2).sum()
    L.backward()

    te = torch.tensor(eta, requires_grad=True)
    tb = torch.tensor(b, requires_grad=True)
    TG = torch_series(mesh, te, tb, dts.cpu().numpy(), gauges)
    TL = ((TG - target.cpu()) ** 2).sum()
    TL.backward()
    assert abs(float(L) - float(TL)) <= 1e-12 * float(TL)
    assert rel_err(e0.grad, te.grad) <= 1e-9
    assert rel_err(bb.grad, tb.grad) <= 1e-9

    def loss(e, bt):
        return float(((prob.gauge_series(e, bt) - target) ** 2).sum())

    # seeded random directions on the cells that stay wet by a margin
    # (perturbing the coast would move the wet/dry masks the adjoint
    # freezes), windowed around the source; steps of 1 mm in eta and 0.5 m in
    # b.  The FD error is measured against the size of the terms the
    # directional derivative sums, sum |grad_i dir_i|: a random direction's
    # derivative can cancel to far below them.
    p = host.tsunami_draw(seed)
    cen = mesh.centroids() * S
    win = np.exp(-((cen[:, 0] - p["x0"]) ** 2 + (cen[:, 1] - p["y0"]) ** 2) / 3.0e5 ** 2)
    mask = torch.tensor(((eta - b) > 50.0) * win, device=DEV)
    rng = np.random.default_rng(11)
    for k in range(6):
        field = k % 2
        dirn = torch.tensor(rng.standard_normal(mesh.n), device=DEV) * mask
        dirn = dirn / dirn.abs().max()
        eps = 1e-3 if field == 0 else 0.5
        grad = e0.grad if field == 0 else bb.grad
        base = [e0.detach().clone(), bb.detach().clone()]
        p_, m_ = [x.clone() for x in base], [x.clone() for x in base]
        p_[field] += eps * dirn
        m_[field] -= eps * dirn
        fd = (loss(*p_) - loss(*m_)) / (2 * eps)
        ad = float((grad * dirn).sum())
        terms = float((grad * dirn).abs().sum())
        print(f"direction {k} field {field}: adjoint {ad:.12e} fd {fd:.12e} terms {terms:.3e}",
              flush=True)
        assert abs(ad - fd) <= 1e-6 * terms
    prob.close()


def test_batched_inversions_with_pools_match_autograd_per_instance():
    depths, seeds, N = (7, 8, 7), (1, 2, 3), 40
    meshes, b, gauges, dt0, truth = [], [], [], [], []
    for d, s in zip(depths, seeds):
        m = host.mesh_uniform(d)
        eta, hu, hv, bb = host.tsunami(m, seed=s, drytol=DRY)
        p = host.tsunami_draw(s)
        meshes.append(m)
        b.append(bb)
        gauges.append(gauges_near(m, p["x0"], p["y0"], 3.0e5, 8, seed=s))
        dt0.append(host.initial_dt(m, (eta, hu, hv, bb), 0.45))
        truth.append([p["amplitude"], p["ratio"], p["x0"], p["y0"]])
    K = math.ceil(math.sqrt(N))
    # a budget that fits the largest instance's checkpoints plus a little:
    # the two small instances keep all their states, and the big one is
    # admitted in a round of its own
    big = tape_bytes(meshes[1].n, N, False, K)
    budget = big + 0.5 * tape_bytes(meshes[0].n, N, True)
    print("case: pools", depths, "budget", budget, flush=True)
    inv = SourceInversion(meshes, b, gauges, N, dt0, hbm_budget=budget)
    plan = inv.problem.plan()
    print("plan", plan, flush=True)
    modes = {m: g["store_all"] for r in plan["rounds"] for g in r for m in g["instances"]}
    assert len(plan["rounds"]) >= 2 and set(modes.values()) == {True, False}
    theta = torch.tensor(truth, dtype=torch.float64, device=DEV)
    inv.set_target(theta)
    scale = torch.tensor(SCALE, dtype=torch.float64, device=DEV)
    z = (theta * torch.tensor([0.8, 1.1, 1.0, 1.0], dtype=torch.float64, device=DEV)
         + torch.tensor([0.0, 0.0, 2.0e4, -1.0e4], dtype=torch.float64, device=DEV)) / scale
    f, g = inv.loss_and_grad(z)

    # each instance's own chain in the restatement, through eta0(theta)
    zz = z.detach().clone().requires_grad_()
    e0 = inv.eta0(zz * scale)
    total = 0.0
    goff = np.cumsum([0] + [len(x) for x in gauges])
    for s, m in enumerate(meshes):
        sl = slice(int(inv.problem.offsets[s]), int(inv.problem.offsets[s + 1]))
        dts = inv.problem.dt_of(s).cpu().numpy()
        TG = torch_series(m, e0[sl].cpu(), torch.tensor(b[s]), dts, gauges[s])
        Ls = ((TG - inv.target[:, goff[s]:goff[s + 1]].cpu()) ** 2).sum()
        assert abs(float(Ls) - float(f[s])) <= 1e-12 * abs(float(Ls))
        total = total + Ls
    total.backward()
    assert rel_err(g, zz.grad) <= 1e-9

    # the same batch with every instance's states kept: identical gradients
    inv_all = SourceInversion(meshes, b, gauges, N, dt0, checkpoint_every="all")
    inv_all.set_target(theta)
    f2, g2 = inv_all.loss_and_grad(z)
    assert torch.equal(f, f2) and torch.equal(g, g2)
    inv.close()
    inv_all.close()
