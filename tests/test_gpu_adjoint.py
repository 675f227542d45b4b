"""Discrete adjoint on the GPU against PyTorch autograd and finite differences.

swe_adjoint_step_device is checked, one traversal at a time, against
torch.autograd of the plain FP64 restatement (tests/torch_ref.py, itself
bit-identical to the oracle) on wet/dry states with both solvers, both
boundary conditions and both traversal directions; the checkpointed
gauge-series driver (samoa_amd/adjoint.py) is checked against autograd
through a chain of steps, against central finite differences on an all-wet
flow, across checkpoint intervals and between a batch and its single runs.

Tolerance: the adjoint sums the same partial derivatives as autograd in a
different order, so agreement is to rounding: max |x - y| <= 1e-9 max |y|
per field."""
import numpy as np
import pytest
import torch

import pyoracle as po
import torch_ref as tr
from samoa_amd import swe
from samoa_amd.adjoint import GaugeInversion

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DRY = 1e-3
TOL = 1e-9


def make_mesh(d):
    plotter, co, nbr = po.mesh_uniform(d)
    return swe.Mesh(np.full(plotter.shape[0], d, np.int8), plotter, nbr, co)


def torch_mesh(m):
    normals = np.stack([po.normals(p) if p else np.zeros((3, 2)) for p in range(-8, 9)])
    return tr.TorchMesh(m.depth, m.plotter, m.nbr, normals, po.leg_sizes())


def wetdry_state(mesh, seed, S):
    rng = np.random.default_rng(seed)
    cen = mesh.centroids()
    x, y = cen[:, 0], cen[:, 1]
    b = -1.0 + 1.5 * x + 0.3 * np.sin(7 * y) + 0.05 * rng.standard_normal(mesh.n)
    eta = np.where(np.hypot(x - 0.3, y - 0.5) < 0.15, 0.8, 0.1) + 0.01 * rng.standard_normal(mesh.n)
    dry = b > eta
    eta = np.where(dry, np.minimum(b, 0.0), eta)
    h = np.maximum(eta - b, 0)
    hu = h * rng.normal(0, 0.5, mesh.n)
    hv = h * rng.normal(0, 0.5, mesh.n)
    hu[dry] = 0
    hv[dry] = 0
    return eta * S / 10, hu * S / 10, hv * S / 10, b * S / 10


def wet_state(mesh, S):
    cen = mesh.centroids()
    x, y = cen[:, 0], cen[:, 1]
    b = -1.0 - 0.5 * x + 0.2 * np.cos(3 * y)
    eta = 0.3 * np.exp(-((x - 0.4) ** 2 + (y - 0.5) ** 2) / 0.02)
    return eta * S / 10, b * S / 10


def rel_err(x, y):
    x = x.detach().cpu().numpy() if torch.is_tensor(x) else x
    y = y.detach().cpu().numpy() if torch.is_tensor(y) else y
    return float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300))


def P(t):
    return t.data_ptr()


@pytest.mark.parametrize("forward", [True, False])
@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("d", [1, 3, 7, 10, 12])
def test_adjoint_step_matches_autograd(d, solver, bc, forward):
    S = 100.0
    mesh = make_mesh(d)
    tm = torch_mesh(mesh)
    st = wetdry_state(mesh, 10 * d + solver + 2 * bc, S)
    dt = po.initial_dt(mesh.depth, st, 0.45) * S
    print(f"case d={d} solver={solver} bc={bc} forward={forward}", flush=True)

    # autograd of the restatement (CPU)
    q = [torch.tensor(a, requires_grad=True) for a in st]
    e, u, v, _ = tr.step(tm, *q, dt, forward, S, po.G, DRY, bc=bc, solver=solver)
    rng = np.random.default_rng(d)
    lam = rng.normal(size=(3, mesh.n))
    loss = sum((o * torch.tensor(w)).sum() for o, w in zip((e, u, v), lam))
    grads = torch.autograd.grad(loss, q, allow_unused=True)
    grads = [np.zeros(mesh.n) if g is None else g.numpy() for g in grads]

    # forward (masks) and adjoint on the GPU
    op = swe.EulerTimestep(mesh, scaling=S, dry_tolerance=DRY, boundary=bc, solver=solver)
    stream = torch.cuda.current_stream().cuda_stream
    dq = [torch.tensor(a, device=DEV) for a in st]
    out = [torch.empty_like(dq[0]) for _ in range(3)]
    dtd = torch.tensor([dt], dtype=torch.float64, device=DEV)
    masks = torch.full((mesh.n,), -1, dtype=torch.int8, device=DEV)
    op.forward_device(*map(P, dq), *map(P, out), P(dtd), forward, masks=P(masks), stream=stream)
    lin = [torch.tensor(w, device=DEV) for w in lam]
    lout = [torch.full_like(dq[0], float("nan")) for _ in range(3)]
    gb = torch.zeros_like(dq[0])
    op.adjoint_device(*map(P, dq), P(masks), *map(P, lin), *map(P, lout), P(gb), P(dtd), forward,
                      stream=stream)
    torch.cuda.synchronize()
    op.close()

    for a, b_ in zip(out, (e, u, v)):
        assert np.array_equal(a.cpu().numpy(), b_.detach().numpy())
    assert set(np.unique(masks.cpu().numpy())) <= {0, 1, 2}
    for name, x, y in zip(("eta", "hu", "hv", "b"), (*lout, gb), grads):
        err = rel_err(x, y)
        assert err <= TOL, (name, err)


def _torch_chain(mesh, eta0, b, dts, forward0, S, solver, bc, gauges):
    tm = torch_mesh(mesh)
    z = torch.zeros_like(eta0)
    q = (eta0, z, z)
    series = []
    for n, dt in enumerate(dts):
        fwd = forward0 ^ (n % 2 == 1)
        e, u, v, _ = tr.step(tm, *q, b, float(dt), fwd, S, po.G, DRY, bc=bc, solver=solver)
        q = (e, u, v)
        series.append(e[torch.as_tensor(gauges)])
    return torch.stack(series)


@pytest.mark.parametrize("solver,bc", [(0, 0), (1, 1), (0, 1)])
def test_gauge_series_gradient_matches_autograd(solver, bc):
    d, S, N = 8, 100.0, 12
    mesh = make_mesh(d)
    eta, _, _, b = wetdry_state(mesh, 5 + solver, S)
    gauges = np.random.default_rng(1).choice(mesh.n, 16, replace=False)
    prob = GaugeInversion(mesh, gauges, N, dict(scaling=S, dry_tolerance=DRY, solver=solver,
                                                 boundary=bc))
    e0 = torch.tensor(eta, device=DEV, requires_grad=True)
    bb = torch.tensor(b, device=DEV, requires_grad=True)
    dt0 = po.initial_dt(mesh.depth, (eta, np.zeros(mesh.n), np.zeros(mesh.n), b), 0.45) * S
    dts = prob.record_dt(e0.detach(), bb.detach(), dt0)[:, 0].cpu().numpy()
    W = torch.tensor(np.random.default_rng(2).normal(size=(N, len(gauges))), device=DEV)
    G = prob.gauge_series(e0, bb)
    (G * W).sum().backward()

    te = torch.tensor(eta, requires_grad=True)
    tb = torch.tensor(b, requires_grad=True)
    TG = _torch_chain(mesh, te, tb, dts, False, S, solver, bc, gauges)
    (TG * W.cpu()).sum().backward()
    assert np.array_equal(G.detach().cpu().numpy(), TG.detach().numpy())
    assert rel_err(e0.grad, te.grad) <= TOL
    assert rel_err(bb.grad, tb.grad) <= TOL
    prob.close()


def test_checkpoint_interval_does_not_change_gradients():
    d, S, N = 9, 100.0, 17
    mesh = make_mesh(d)
    eta, _, _, b = wetdry_state(mesh, 3, S)
    gauges = np.arange(0, mesh.n, 37)
    dt0 = po.initial_dt(mesh.depth, (eta, np.zeros(mesh.n), np.zeros(mesh.n), b), 0.45) * S
    res = []
    for K in (1, 3, N, "all", "auto"):
        prob = GaugeInversion(mesh, gauges, N, dict(scaling=S, dry_tolerance=DRY),
                              checkpoint_every=K)
        e0 = torch.tensor(eta, device=DEV, requires_grad=True)
        bb = torch.tensor(b, device=DEV, requires_grad=True)
        prob.record_dt(e0.detach(), bb.detach(), dt0)
        (prob.gauge_series(e0, bb) ** 2).sum().backward()
        res.append((e0.grad.cpu().numpy(), bb.grad.cpu().numpy()))
        prob.close()
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0]) and np.array_equal(r[1], res[0][1])


@pytest.mark.parametrize("solver", [0, 1])
def test_gauge_gradient_matches_finite_differences(solver):
    d, S, N = 8, 1000.0, 30
    mesh = make_mesh(d)
    eta, b = wet_state(mesh, S)
    gauges = np.random.default_rng(4).choice(mesh.n, 12, replace=False)
    prob = GaugeInversion(mesh, gauges, N, dict(scaling=S, dry_tolerance=DRY, solver=solver,
                                                 boundary=1))
    z = np.zeros(mesh.n)
    dt0 = po.initial_dt(mesh.depth, (eta, z, z, b), 0.45) * S
    e0 = torch.tensor(eta, device=DEV, requires_grad=True)
    bb = torch.tensor(b, device=DEV, requires_grad=True)
    prob.record_dt(e0.detach(), bb.detach(), dt0)
    target = prob.gauge_series(e0.detach() * 0.7, bb.detach()).detach()

    def loss(e, bt):
        return float(((prob.gauge_series(e, bt) - target) ** 2).sum())

    L = ((prob.gauge_series(e0, bb) - target) ** 2).sum()
    L.backward()
    cen = mesh.centroids()
    for k, (field, grad) in enumerate(((0, e0.grad), (1, bb.grad))):
        delta = np.sin(3 * cen[:, 0] + 1.3 * k) * np.cos(2 * cen[:, 1]) * S / 10
        dd = torch.tensor(delta, device=DEV)
        eps = 1e-6
        args_p = [e0.detach().clone(), bb.detach().clone()]
        args_m = [e0.detach().clone(), bb.detach().clone()]
        args_p[field] += eps * dd
        args_m[field] -= eps * dd
        fd = (loss(*args_p) - loss(*args_m)) / (2 * eps)
        ad = float((grad * dd).sum())
        print(f"field {field}: adjoint {ad:.12e} fd {fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-30)
    prob.close()


def test_batch_adjoint_equals_single_runs():
    S = (100.0, 250.0)
    meshes = [make_mesh(7), make_mesh(9)]
    params = [dict(scaling=S[0], dry_tolerance=DRY, solver=0, boundary=1),
              dict(scaling=S[1], dry_tolerance=DRY, solver=0, boundary=1)]
    N = 9
    states = [wetdry_state(m, 20 + i, s) for i, (m, s) in enumerate(zip(meshes, S))]
    dt0 = [po.initial_dt(m.depth, (st[0], np.zeros(m.n), np.zeros(m.n), st[3]), 0.45) * s
           for m, st, s in zip(meshes, states, S)]
    gl = [np.arange(0, m.n, 7) for m in meshes]
    off = [0, meshes[0].n]
    batch = GaugeInversion(meshes, np.concatenate([g + o for g, o in zip(gl, off)]), N, params)
    e0 = torch.tensor(np.concatenate([st[0] for st in states]), device=DEV, requires_grad=True)
    bb = torch.tensor(np.concatenate([st[3] for st in states]), device=DEV, requires_grad=True)
    batch.record_dt(e0.detach(), bb.detach(), dt0)
    W = torch.tensor(np.random.default_rng(0).normal(size=(N, len(batch.gauges))), device=DEV)
    GB = batch.gauge_series(e0, bb)
    (GB * W).sum().backward()
    col = 0
    for s, m in enumerate(meshes):
        one = GaugeInversion(m, gl[s], N, params[s])
        a = torch.tensor(states[s][0], device=DEV, requires_grad=True)
        c = torch.tensor(states[s][3], device=DEV, requires_grad=True)
        one.record_dt(a.detach(), c.detach(), dt0[s])
        assert torch.equal(one.dt, batch.dt[:, s:s + 1])
        G1 = one.gauge_series(a, c)
        (G1 * W[:, col:col + len(gl[s])]).sum().backward()
        assert torch.equal(G1, GB[:, col:col + len(gl[s])])
        sl = slice(off[s], off[s] + m.n)
        assert torch.equal(a.grad, e0.grad[sl]) and torch.equal(c.grad, bb.grad[sl])
        col += len(gl[s])
        one.close()
    batch.close()


def test_adjoint_step_independent_of_launch_shape():
    """Cells whose neighbour lies in another chunk get its seed formed exactly
    as that chunk forms it, so the chunk size does not change a bit."""
    d, S = 11, 100.0
    mesh = make_mesh(d)
    st = wetdry_state(mesh, 77, S)
    dt = po.initial_dt(mesh.depth, st, 0.45) * S
    op = swe.EulerTimestep(mesh, scaling=S, dry_tolerance=DRY, boundary=1)
    stream = torch.cuda.current_stream().cuda_stream
    dq = [torch.tensor(a, device=DEV) for a in st]
    out = [torch.empty_like(dq[0]) for _ in range(3)]
    dtd = torch.tensor([dt], dtype=torch.float64, device=DEV)
    masks = torch.empty(mesh.n, dtype=torch.int8, device=DEV)
    rng = np.random.default_rng(5)
    lin = [torch.tensor(rng.normal(size=mesh.n), device=DEV) for _ in range(3)]
    results = []
    for chunk, grid in ((256, 0), (128, 0), (512, 0), (256, 16), (128, 2048)):
        print("case: launch shape", chunk, grid, flush=True)
        op.set_tuning(chunk, grid)
        op.forward_device(*map(P, dq), *map(P, out), P(dtd), True, masks=P(masks), stream=stream)
        lout = [torch.empty_like(dq[0]) for _ in range(3)]
        gb = torch.zeros_like(dq[0])
        op.adjoint_device(*map(P, dq), P(masks), *map(P, lin), *map(P, lout), P(gb), P(dtd), True,
                          stream=stream)
        torch.cuda.synchronize()
        results.append(torch.stack(lout + [gb]).cpu())
    op.close()
    for r in results[1:]:
        assert torch.equal(r, results[0])
