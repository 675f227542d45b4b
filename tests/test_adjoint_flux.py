"""The hand-written transpose flux (samoa_amd/csrc/swe_adjoint.hpp) against
PyTorch autograd of the FP64 restatement, on the reference's KAT states
(every branch of the solver) and random states.  CPU only."""
import numpy as np
import pytest
import torch

import golden_io
import native_shim as ns
import pyoracle as po
import torch_ref as tr


def kat_states(n_max=1500):
    z = golden_io.load("kat_geoclaw.npz")
    inp = z["inputs"]
    rng = np.random.default_rng(0)
    idx = rng.permutation(len(inp))[:n_max]
    out = []
    for r in inp[idx]:
        hL, hR, huL, huR, hvL, hvR, bL, bR, dry, g = r
        # edge frame with n = (1, 0): eta = h + b
        out.append((np.array([hL + bL, huL, hvL, bL]), np.array([hR + bR, huR, hvR, bR]), dry))
    return out


def torch_vjp(solver, n, QL, QR, dry, gfL, gfR):
    tl = [torch.tensor(x, requires_grad=True) for x in QL]
    trr = [torch.tensor(x, requires_grad=True) for x in QR]
    nx, ny = torch.tensor([n[0]]), torch.tensor([n[1]])
    fL, fR = tr.edge_flux(nx, ny, [t.reshape(1) for t in tl], [t.reshape(1) for t in trr], dry,
                          po.G, solver)
    loss = sum(fL[i].sum() * gfL[i] + fR[i].sum() * gfR[i] for i in range(3))
    grads = torch.autograd.grad(loss, tl + trr, allow_unused=True)
    g = [0.0 if x is None else float(x) for x in grads]
    return np.array(g[:4]), np.array(g[4:])


def test_host_build_of_solver_matches_oracle():
    rng = np.random.default_rng(1)
    for QL, QR, dry in kat_states(7000):
        for solver in (0, 1):
            n = [(1.0, 0.0), (0.0, -1.0), (np.sqrt(0.5), np.sqrt(0.5))][rng.integers(3)]
            a = ns.edge_flux(solver, n, QL, QR, dry, po.G)
            b = po.flux(np.array(n), QL, QR, dry, solver=solver)
            for x, y in zip(a, b):
                assert np.array_equal(x, y) or (np.isnan(x) == np.isnan(y)).all()


@pytest.mark.parametrize("solver", [0, 1])
def test_transpose_flux_matches_autograd(solver):
    rng = np.random.default_rng(2 + solver)
    worst = 0.0
    checked = 0
    for QL, QR, dry in kat_states():
        fL, fR = ns.edge_flux(solver, (1.0, 0.0), QL, QR, dry, po.G)
        if not (np.isfinite(fL).all() and np.isfinite(fR).all()):
            continue
        gfL, gfR = rng.normal(size=3), rng.normal(size=3)
        a = ns.edge_flux_adj(solver, (1.0, 0.0), QL, QR, dry, po.G, gfL, gfR)
        b = torch_vjp(solver, (1.0, 0.0), QL, QR, dry, gfL, gfR)
        for x, y in zip(a, b):
            if not np.isfinite(y).all():
                continue
            scale = max(1.0, np.abs(y).max())
            err = np.abs(x - y).max() / scale
            worst = max(worst, err)
            assert err < 1e-9, (QL, QR, x, y)
        checked += 1
    assert checked > 1000
    print("worst relative error", worst)


def test_transpose_flux_rotated_frames():
    rng = np.random.default_rng(7)
    for _ in range(300):
        QL = np.array([rng.uniform(0.2, 3), *rng.normal(0, 1, 2), rng.uniform(-1, 0)])
        QR = np.array([rng.uniform(0.2, 3), *rng.normal(0, 1, 2), rng.uniform(-1, 0)])
        th = rng.uniform(0, 2 * np.pi)
        n = (np.cos(th), np.sin(th))
        gfL, gfR = rng.normal(size=3), rng.normal(size=3)
        a = ns.edge_flux_adj(0, n, QL, QR, 1e-3, po.G, gfL, gfR)
        b = torch_vjp(0, n, QL, QR, 1e-3, gfL, gfR)
        for x, y in zip(a, b):
            assert np.allclose(x, y, rtol=1e-10, atol=1e-10)
