"""Batched tsunami-source inversion (samoa_amd/inversion.py) on the GPU:
two instances of uneven size recover their sources from noise-free gauge
records in lockstep."""
import numpy as np
import pytest
import torch

from samoa_amd import host
from samoa_amd.inversion import SCALE, SourceInversion, gauges_near, lbfgs

pytestmark = pytest.mark.gpu


def test_batched_source_inversion_recovers_sources():
    depths, seeds, N = (10, 11), (3, 8), 30
    meshes = [host.mesh_uniform(d) for d in depths]
    truth, b, gauges, dt0 = [], [], [], []
    for m, s in zip(meshes, seeds):
        p = host.tsunami_draw(s)
        eta, hu, hv, bb = host.tsunami(m, seed=s)
        truth.append([p["amplitude"], p["ratio"], p["x0"], p["y0"]])
        b.append(bb)
        gauges.append(gauges_near(m, p["x0"], p["y0"], 1.5e5, 16, seed=s))
        dt0.append(host.initial_dt(m, (eta, hu, hv, bb), 0.45))
    print("case: depths", depths, "seeds", seeds, flush=True)
    inv = SourceInversion(meshes, b, gauges, N, dt0)
    theta_true = torch.tensor(truth, dtype=torch.float64, device="cuda:0")
    inv.set_target(theta_true)
    scale = torch.tensor(SCALE, dtype=torch.float64, device="cuda:0")
    guess = theta_true * torch.tensor([0.6, 1.0, 1.0, 1.0], dtype=torch.float64, device="cuda:0")
    guess = guess + torch.tensor([0.0, 0.4, 3.0e4, -2.0e4], dtype=torch.float64, device="cuda:0")
    z, f, hist = lbfgs(inv.loss_and_grad, guess / scale, iters=25)
    theta = z * scale
    rel = ((theta - theta_true).abs() / theta_true.abs()).cpu().numpy()
    print("loss history", [h.cpu().numpy().tolist() for h in hist[::5]], flush=True)
    print("relative parameter errors", rel, flush=True)
    assert bool((hist[-1] < 1e-6 * hist[0]).all())
    assert rel.max() < 1e-3
    inv.close()
