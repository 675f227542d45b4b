"""The PyTorch FP64 restatement (tests/torch_ref.py) is the same discrete
scheme as the oracle: bit-identical steps on wet/dry states."""
import numpy as np
import pytest
import torch

import pyoracle as po
import torch_ref as tr
from samoa_amd import host


def torch_mesh(m):
    normals = np.stack([po.normals(p) if p else np.zeros((3, 2)) for p in range(-8, 9)])
    return tr.TorchMesh(m.depth, m.plotter, m.nbr, normals, po.leg_sizes())


def wetdry_state(m, seed, S):
    rng = np.random.default_rng(seed)
    cen = m.centroids()
    x, y = cen[:, 0], cen[:, 1]
    b = (-1.0 + 1.5 * x + 0.3 * np.sin(7 * y)) * S / 10
    eta = np.where(np.hypot(x - 0.3, y - 0.5) < 0.15, 0.8, 0.1) * S / 10
    dry = b > eta
    eta = np.where(dry, np.minimum(b, 0.0), eta)
    h = np.maximum(eta - b, 0)
    hu = h * rng.normal(0, 0.5, m.n)
    hv = h * rng.normal(0, 0.5, m.n)
    hu[dry] = 0
    hv[dry] = 0
    return eta, hu, hv, b


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("bc", [0, 1])
def test_torch_step_matches_oracle(solver, bc):
    d, S = 7, 100.0
    m = host.mesh_uniform(d)
    tm = torch_mesh(m)
    st = wetdry_state(m, 3 + solver + 2 * bc, S)
    dt = po.initial_dt(m.depth, st, 0.45) * S
    ref = list(st)
    cur = [torch.tensor(a) for a in st]
    for fwd in (True, False, True, False):
        e, u, v, _, cand, _ = po.step(m.depth, m.plotter, m.nbr, ref, dt, scaling=S, drytol=1e-3,
                                      bc=bc, solver=solver, forward=fwd)
        te, tu, tv, tc = tr.step(tm, *cur, dt, fwd, S, po.G, 1e-3, bc=bc, solver=solver)
        assert float(tc.min()) == cand
        for a, b_ in zip((te, tu, tv), (e, u, v)):
            assert np.array_equal(a.numpy(), b_)
        ref = [e, u, v, ref[3]]
        cur = [te, tu, tv, cur[3]]
        dt = 0.45 * cand
