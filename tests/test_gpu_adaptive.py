"""The HIP library on the reference's adaptive meshes (cells of six to nine
different depths; tests/golden/adapt_*.npz): bit-identical to the
reference's outputs through the step, the device time loop and a batch, and
the adjoint of a traversal against autograd on the same mesh."""
import numpy as np
import pytest
import torch

import pyoracle as po
import torch_ref as tr
from samoa_amd import swe
from samoa_amd.vtk_mesh import mesh_from_corners
from test_adaptive_mesh import adaptive_runs

pytestmark = pytest.mark.gpu

RUNS = {r[0]: r for r in adaptive_runs()}


def setup(name):
    _, z, meta = RUNS[name]
    mesh = mesh_from_corners(z["corners"], z["depth"], meta["scaling"], meta["offset"])
    params = dict(scaling=meta["scaling"], dry_tolerance=meta["drytol"], courant=meta["courant"],
                  min_depth=meta["dmin"], max_depth=meta["dmax"], boundary=meta["bc"])
    return z, meta, mesh, params


@pytest.mark.parametrize("name", sorted(RUNS))
def test_steps_bit_identical(name):
    z, meta, mesh, params = setup(name)
    print("case:", name, mesh.n, "cells", flush=True)
    op = swe.EulerTimestep(mesh, **params)
    op.upload(*z["states"][0])
    dt = po.initial_dt(mesh.depth, z["states"][0], meta["courant"])
    t = 0.0
    for k in range(1, meta["steps"] + 1):
        res = op.traverse(dt, True, flags=True)
        t += dt
        dt = meta["courant"] * res.dt_cfl
        eta, hu, hv, b = op.download()
        assert z["times"][k] == t
        assert np.array_equal(np.stack([eta, hu, hv, b]), z["states"][k]), f"step {k}"
        assert np.array_equal(res.refine, z["refine"][k])
        assert res.n_refine == int((z["refine"][k] == 1).sum())
    op.close()


@pytest.mark.parametrize("name", sorted(RUNS))
def test_device_loop_and_launch_shapes(name):
    """swe_run's device dt chain (all steps forward here: one step at a time
    keeps the direction) for every launch shape."""
    z, meta, mesh, params = setup(name)
    dt0 = po.initial_dt(mesh.depth, z["states"][0], meta["courant"])
    for chunk, grid in ((256, 0), (128, 0), (512, 8)):
        print("case:", name, chunk, grid, flush=True)
        op = swe.EulerTimestep(mesh, **params)
        op.set_tuning(chunk, grid)
        op.upload(*z["states"][0])
        t, dt = 0.0, dt0
        for k in range(1, meta["steps"] + 1):
            t, dt = op.run(1, dt, forward=True, time=t)
            assert z["times"][k] == t
        eta, hu, hv, b = op.download()
        assert np.array_equal(np.stack([eta, hu, hv, b]), z["states"][meta["steps"]])
        op.close()


@pytest.mark.parametrize("name", sorted(RUNS))
def test_batch_of_adaptive_meshes(name):
    """Two scenarios on one adaptive mesh, started from the reference's
    states of step 0 and step 2 (a batch shares solver, boundary and g, and
    the two runs differ in their boundary)."""
    z, meta, mesh, params = setup(name)
    starts = (0, 2)
    b = swe.Batch([mesh] * len(starts), [params] * len(starts))
    for i, s in enumerate(starts):
        b.upload(i, *z["states"][s])
    # the dt of step s+1 is courant times the CFL bound of step s (oracle)
    dt, dt_at = po.initial_dt(mesh.depth, z["states"][0], meta["courant"]), []
    for k in range(max(starts) + 1):
        dt_at.append(dt)
        cand = po.step(mesh.depth, mesh.plotter, mesh.nbr, z["states"][k], dt,
                       scaling=meta["scaling"], drytol=meta["drytol"], dmin=meta["dmin"],
                       dmax=meta["dmax"], bc=meta["bc"], forward=True)[4]
        dt = meta["courant"] * cand
    dts = np.array([dt_at[s] for s in starts])
    print("case:", name, "starts", starts, flush=True)
    for k in range(1, meta["steps"] - max(starts) + 1):
        dtc, nref = b.step(dts, True)
        dts = meta["courant"] * dtc
        for i, s in enumerate(starts):
            assert np.array_equal(np.stack(b.download(i)), z["states"][s + k]), (name, s, k)
    b.close()


def test_adjoint_on_adaptive_mesh():
    name = sorted(RUNS)[0]
    z, meta, mesh, params = setup(name)
    S = meta["scaling"]
    normals = np.stack([po.normals(p) if p else np.zeros((3, 2)) for p in range(-8, 9)])
    tm = tr.TorchMesh(mesh.depth, mesh.plotter, mesh.nbr, normals, po.leg_sizes())
    st = z["states"][3]
    dt = 1e-3
    q = [torch.tensor(a, requires_grad=True) for a in st]
    e, u, v, _ = tr.step(tm, *q, dt, True, S, po.G, meta["drytol"], bc=meta["bc"])
    lam = np.random.default_rng(0).normal(size=(3, mesh.n))
    loss = sum((o * torch.tensor(w)).sum() for o, w in zip((e, u, v), lam))
    grads = [g.numpy() for g in torch.autograd.grad(loss, q)]
    op = swe.EulerTimestep(mesh, **params)
    P = lambda t: t.data_ptr()  # noqa: E731
    dq = [torch.tensor(a, device="cuda:0") for a in st]
    out = [torch.empty_like(dq[0]) for _ in range(3)]
    dtd = torch.tensor([dt], dtype=torch.float64, device="cuda:0")
    masks = torch.empty(mesh.n, dtype=torch.int8, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    op.forward_device(*map(P, dq), *map(P, out), P(dtd), True, masks=P(masks), stream=stream)
    lin = [torch.tensor(w, device="cuda:0") for w in lam]
    lout = [torch.empty_like(dq[0]) for _ in range(3)]
    gb = torch.zeros_like(dq[0])
    op.adjoint_device(*map(P, dq), P(masks), *map(P, lin), *map(P, lout), P(gb), P(dtd), True,
                      stream=stream)
    torch.cuda.synchronize()
    op.close()
    for a, b_ in zip(out, (e, u, v)):
        assert np.array_equal(a.cpu().numpy(), b_.detach().numpy())
    for x, y in zip((*lout, gb), grads):
        x = x.cpu().numpy()
        assert np.abs(x - y).max() <= 1e-9 * max(np.abs(y).max(), 1e-300)
