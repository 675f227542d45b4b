"""Adaptive meshes: the reference's own refined meshes, rebuilt from the
corners it writes (samoa_amd/vtk_mesh.py), stepped by the oracle and
compared bit for bit with the reference's later outputs (states, times and
refinement flags; tests/golden/adapt_*.npz).  Host only."""
import glob
import json
import os

import numpy as np
import pytest

import golden_io
import pyoracle as po
from samoa_amd import host
from samoa_amd.vtk_mesh import mesh_from_corners


def adaptive_runs():
    out = []
    for f in sorted(glob.glob(os.path.join(golden_io.GOLDEN, "adapt_*.npz"))):
        z = np.load(f, allow_pickle=False)
        out.append((os.path.basename(f)[6:-4], z, json.loads(str(z["meta"]))))
    return out


@pytest.mark.parametrize("d", [1, 2, 5, 8])
def test_rebuild_uniform_mesh_from_corners(d):
    """The corners as the reference writes them give back the generator's
    plotter types and neighbours exactly."""
    m = host.mesh_uniform(d)
    S, off = 10.0, -5.0
    world = S * m.coords + off
    ref_order = np.where((m.plotter > 0)[:, None, None], world, world[:, ::-1])
    r = mesh_from_corners(ref_order, m.depth, S, off)
    assert np.array_equal(r.plotter, m.plotter)
    assert np.array_equal(r.nbr, m.nbr.reshape(-1, 3))


@pytest.mark.parametrize("name", [r[0] for r in adaptive_runs()])
def test_oracle_on_reference_adaptive_mesh(name):
    _, z, meta = {r[0]: r for r in adaptive_runs()}[name]
    mesh = mesh_from_corners(z["corners"], z["depth"], meta["scaling"], meta["offset"])
    assert len(np.unique(mesh.depth)) > 3  # really adaptive
    # every pair of neighbours differs by at most one level (conforming)
    nb = mesh.nbr
    for k in range(3):
        o = nb[:, k]
        ok = o >= 0
        assert np.abs(mesh.depth[ok].astype(int) - mesh.depth[o[ok]].astype(int)).max() <= 1
    st = z["states"][0].copy()
    dt = po.initial_dt(mesh.depth, st, meta["courant"])
    t = 0.0
    for k in range(1, meta["steps"] + 1):
        # each step is followed by an output traversal that runs backward in
        # the stored order, so every Euler step runs forward
        e, u, v, rf, cand, _ = po.step(mesh.depth, mesh.plotter, mesh.nbr, st, dt,
                                       scaling=meta["scaling"], drytol=meta["drytol"],
                                       dmin=meta["dmin"], dmax=meta["dmax"], bc=meta["bc"],
                                       forward=True)
        t += dt
        st = np.stack([e, u, v, st[3]])
        dt = meta["courant"] * cand
        assert z["times"][k] == t, f"time after step {k}"
        assert np.array_equal(st, z["states"][k]), f"state after step {k}"
        assert np.array_equal(rf, z["refine"][k]), f"refinement flags after step {k}"
