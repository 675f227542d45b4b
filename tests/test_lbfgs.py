"""Lockstep L-BFGS of samoa_amd/inversion.py on independent host problems
with known minima: every instance converges as an L-BFGS run of its own
would, whatever the others do."""
import torch

from samoa_amd.inversion import lbfgs


def batched_problems(z):
    """Instance 0: ill-conditioned quadratic; 1: Rosenbrock; 2: shifted
    isotropic quadratic (each [4] parameters, minima known)."""
    f = torch.zeros(3, dtype=torch.float64)
    g = torch.zeros_like(z)
    D = torch.tensor([1.0, 10.0, 100.0, 1000.0], dtype=torch.float64)
    f[0] = 0.5 * (D * z[0] ** 2).sum()
    g[0] = D * z[0]
    x, y = z[1, 0], z[1, 1]
    f[1] = (1 - x) ** 2 + 100 * (y - x * x) ** 2 + (z[1, 2:] ** 2).sum()
    g[1, 0] = -2 * (1 - x) - 400 * x * (y - x * x)
    g[1, 1] = 200 * (y - x * x)
    g[1, 2:] = 2 * z[1, 2:]
    c = torch.tensor([3.0, -2.0, 0.5, 7.0], dtype=torch.float64)
    f[2] = ((z[2] - c) ** 2).sum()
    g[2] = 2 * (z[2] - c)
    return f, g


def test_lockstep_lbfgs_converges_per_instance():
    z0 = torch.tensor([[1.0, 1.0, 1.0, 1.0], [-1.2, 1.0, 0.3, -0.2], [0.0, 0.0, 0.0, 0.0]],
                      dtype=torch.float64)
    z, f, hist = lbfgs(batched_problems, z0, iters=60)
    assert torch.allclose(z[0], torch.zeros(4, dtype=torch.float64), atol=1e-6)
    assert torch.allclose(z[1], torch.tensor([1.0, 1.0, 0.0, 0.0], dtype=torch.float64), atol=1e-4)
    assert torch.allclose(z[2], torch.tensor([3.0, -2.0, 0.5, 7.0], dtype=torch.float64), atol=1e-8)
    # monotone per instance (Armijo)
    H = torch.stack(hist)
    assert bool((H[1:] <= H[:-1] + 1e-12).all())


def test_instances_do_not_interact():
    z0 = torch.tensor([[1.0, 1.0, 1.0, 1.0], [-1.2, 1.0, 0.3, -0.2], [0.0, 0.0, 0.0, 0.0]],
                      dtype=torch.float64)
    za, _, _ = lbfgs(batched_problems, z0, iters=15)

    def only_third(z):
        f, g = batched_problems(torch.cat([z0[:2], z]))
        return f[2:], g[2:]

    zb, _, _ = lbfgs(only_third, z0[2:], iters=15)
    assert torch.equal(za[2], zb[0])
