"""Gridded bathymetry (samoa_amd/bathymetry.py): netCDF classic grids read
back exactly, the reference's domain rule (SWE.f90:133-137) on a worked
example, and bilinear sampling that reproduces bilinear functions."""
import os
import tempfile

import numpy as np
import pytest

from samoa_amd import bathymetry as bt
from samoa_amd import host


def write_grid(path, x, y, z, name="z"):
    from scipy.io import netcdf_file

    with netcdf_file(path, "w") as nc:
        nc.createDimension("x", len(x))
        nc.createDimension("y", len(y))
        nc.createVariable("x", "f8", ("x",))[:] = x
        nc.createVariable("y", "f8", ("y",))[:] = y
        nc.createVariable(name, "f4", ("y", "x"))[:] = z


def test_read_netcdf_grid_round_trip():
    x = -1000.0 + 250.0 * np.arange(9)
    y = 500.0 + 125.0 * np.arange(5)
    z = np.random.default_rng(0).normal(size=(5, 9)).astype(np.float32)
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "bath.nc")
        write_grid(f, x, y, z)
        r = bt.read_netcdf_grid(f)
    assert (r.x0, r.y0, r.dx, r.dy) == (-1000.0, 500.0, 250.0, 125.0)
    assert np.array_equal(r.values, z.astype(np.float64))
    assert r.extent == ((-1000.0, 1000.0), (500.0, 1000.0))


def test_read_netcdf_rejects_uneven_axes():
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "bad.nc")
        write_grid(f, np.array([0.0, 1.0, 3.0]), np.array([0.0, 1.0]), np.zeros((2, 3)))
        with pytest.raises(ValueError):
            bt.read_netcdf_grid(f)


def test_asagi_domain_worked_example():
    # bathymetry over [0, 100] x [0, 50], displacement over [70, 90] x [15, 35]:
    # scaling 100; the square centred on (80, 25) starts at (30, -25), which
    # the clamp moves to x = min(max(30, 0), 100 - 100) = 0 and
    # y = min(max(-25, 0), 50 - 100) = -50
    bath = bt.Raster(np.zeros((11, 21)), 0.0, 0.0, 5.0, 5.0)
    displ = bt.Raster(np.zeros((5, 5)), 70.0, 15.0, 5.0, 5.0)
    assert bt.asagi_domain(bath, displ) == (100.0, (0.0, -50.0))
    # without a displacement grid the square sits on the bathymetry
    assert bt.asagi_domain(bath) == (100.0, (0.0, -50.0))


def test_sample_reproduces_bilinear_functions():
    r = bt.Raster(np.zeros((7, 13)), -3.0, 2.0, 0.5, 0.25)
    ny, nx = r.values.shape
    X, Y = np.meshgrid(r.x0 + r.dx * np.arange(nx), r.y0 + r.dy * np.arange(ny))
    f = lambda x, y: 1.5 - 0.25 * x + 2.0 * y + 0.125 * x * y  # noqa: E731
    r.values = f(X, Y)
    rng = np.random.default_rng(1)
    x = rng.uniform(-3.0, -3.0 + 6.0, 1000)
    y = rng.uniform(2.0, 2.0 + 1.5, 1000)
    assert np.allclose(bt.sample(r, x, y), f(x, y), rtol=0, atol=1e-12)
    assert np.array_equal(bt.sample(r, X.ravel(), Y.ravel()), r.values.ravel())
    # outside the nodes' extent: the reference's constant sea floor
    # (SWE_initialize.f90:164-171); the extent's own edges are inside
    out = bt.sample(r, np.array([-4.0, -3.0 + 6.0 + 1e-9, -2.0, 0.0]), np.array([2.5, 3.0, 1.9, 3.5]))
    assert np.array_equal(out[:3], np.full(3, bt.OUTSIDE))
    assert out[3] == f(0.0, 3.5)


def test_cell_bathymetry_uses_centroids():
    mesh = host.mesh_uniform(4)
    r = bt.Raster(np.add.outer(np.arange(6.0), 10.0 * np.arange(6.0)), 0.0, 0.0, 2.5, 2.0)
    b = bt.cell_bathymetry(r, mesh.coords, 10.0, (0.0, 0.0))
    c = 10.0 * mesh.coords.mean(axis=1)
    assert np.allclose(b, 10.0 * c[:, 0] / 2.5 + c[:, 1] / 2.0, atol=1e-12)


def test_integrate_restatement_matches_reference_adaptive_run():
    """The oracle's refine_2D_recursive (SWE_initialize.f90:106-147) with the
    bathymetry function of the reference's patched scenario reproduces, bit
    for bit, the b of every cell of the reference's adaptive run (depths
    3-11, built with _ADAPT_INTEGRATE): 256 samples per depth-3 cell down to
    the centroid at depth 11."""
    import math

    import golden_io
    import pyoracle as po

    z = golden_io.load("adapt_bath_d3_11.npz")
    import json

    meta = json.loads(str(z["meta"]))
    S, off, dmax = meta["scaling"], meta["offset"], meta["dmax"]
    depth = z["depth"]
    b = z["states"][0, 3]
    mesh = host.mesh_from_depths(depth)

    def f(x, y):  # SWE_initialize.f90:185 as the reference's bath variant patches it
        return -0.5 + 0.25 * (S * x + off) + 0.1 * math.cos(2.0 * (S * y + off))

    got = np.array([po.integrate_bathymetry(f, *(tuple(c) for c in mesh.coords[i]),
                                            po.integrate_levels(int(depth[i]), dmax))
                    for i in range(mesh.n)])
    assert (depth < dmax).sum() > 100
    assert np.array_equal(got, b)
    # the centroid alone misses the coarse cells: the recursion is what is pinned
    cen = np.array([f(*(mesh.coords[i].sum(axis=0) / 3.0)) for i in range(mesh.n)])
    assert not np.array_equal(cen[depth < dmax], b[depth < dmax])


def test_raster_integration_of_the_host_paths():
    """The tsunami scenario (Fortran) and cell_bathymetry (numpy) average the
    raster as the oracle's recursion does, with the gridded-data levels
    min(16, min(dmax + 3, data_depth + 3) - d)."""
    import pyoracle as po

    seed, S = 4, 1.0e6
    r = host.tsunami_raster(seed)
    rast = bt.Raster(r["values"], r["x0"], r["y0"], r["dx"], r["dy"])
    dd = bt.data_depth(rast, S)
    assert dd == po.data_depth(S, r["dx"], r["dy"]) == 18
    f = po.bilinear(r["values"], r["x0"], r["y0"], r["dx"], r["dy"])
    for d, dmax in ((12, 12), (15, 17), (6, 6)):
        mesh = host.mesh_uniform(d)
        idx = np.random.default_rng(d).choice(mesh.n, size=min(mesh.n, 40), replace=False)
        levels = po.integrate_levels(d, dmax, dd)
        assert levels == min(16, min(dmax + 3, 21) - d)
        want = np.array([po.integrate_bathymetry(lambda x, y: f(S * x, S * y),
                                                 *(tuple(c) for c in mesh.coords[i]), levels)
                         for i in idx])
        b = host.tsunami(mesh, seed=seed, dmax=dmax)[3]
        assert np.array_equal(b[idx], want), d
        cb = bt.cell_bathymetry(rast, mesh.coords[idx], S, 0.0, depth=mesh.depth[idx], dmax=dmax)
        assert np.array_equal(cb, want), d


def test_data_depth_closed_forms():
    """nint(log2(S^2 / (dx dy))) (SWE_initialize.f90:129): a grid of 2^k x 2^k
    samples over the square is depth 2k; rounding is to the nearest level,
    both above and below the square's own scale."""
    import pyoracle as po

    for k in range(0, 13):
        assert po.data_depth(1.0e6, 1.0e6 / 2 ** k, 1.0e6 / 2 ** k) == 2 * k
    assert po.data_depth(1.0, 0.6, 0.6) == 1      # log2(2.78) = 1.47
    assert po.data_depth(1.0, 0.55, 0.55) == 2    # log2(3.31) = 1.73
    assert po.data_depth(1.0, 3.0, 3.0) == -3     # log2(1/9) = -3.17
    assert po.data_depth(1.0, 4.0, 5.0) == -4     # log2(1/20) = -4.32
    r = bt.Raster(np.zeros((3, 5)), 0.0, 0.0, 1.0e6 / 64, 1.0e6 / 64)
    assert bt.data_depth(r, 1.0e6) == 12
