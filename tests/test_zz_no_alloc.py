"""swe_step_device_noalloc (the step whose scratch is the library's fixed
device array) against swe_step_device on the same inputs, through the C ABI:
two calls in a row, the second on a smaller mesh, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import golden_io
from samoa_amd import _native, host, swe

pytestmark = pytest.mark.gpu


def wet_dry_state(mesh, seed, S):
    rng = np.random.default_rng(seed)
    cen = mesh.centroids()
    x, y = cen[:, 0], cen[:, 1]
    b = (-1.0 + 1.5 * x + 0.3 * np.sin(7 * y)) * S / 10
    eta = np.where(np.hypot(x - 0.3, y - 0.5) < 0.15, 0.8, 0.1) * S / 10
    dry = b > eta
    eta = np.where(dry, np.minimum(b, 0.0), eta)
    h = np.maximum(eta - b, 0.0)
    hu = np.where(dry, 0.0, h * rng.normal(0, 0.5, mesh.n))
    hv = np.where(dry, 0.0, h * rng.normal(0, 0.5, mesh.n))
    return eta, hu, hv, b


def step_both(torch, mesh, st, S, dt, forward, dmin, dmax, bc, solver):
    dev = torch.device("cuda:0")
    lib = _native.load()
    q = [torch.tensor(a, dtype=torch.float64, device=dev) for a in st]
    stream = torch.cuda.current_stream().cuda_stream
    outs = []
    # the handle path
    op = swe.EulerTimestep(mesh, scaling=S, dry_tolerance=1e-3, min_depth=dmin, max_depth=dmax,
                           boundary=bc, solver=solver)
    o = [torch.full_like(q[0], np.nan) for _ in range(3)]
    dtc = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
    nref = torch.full((1,), -1, dtype=torch.int64, device=dev)
    rf = torch.full((mesh.n,), 7, dtype=torch.int8, device=dev)
    op.step_device(*[t.data_ptr() for t in q], *[t.data_ptr() for t in o], dt, forward,
                   dtc.data_ptr(), nref.data_ptr(), rf.data_ptr(), stream=stream)
    torch.cuda.synchronize()
    outs.append((o, dtc, nref, rf))
    op.close()
    # the no-allocation path: the mesh in device memory
    d_depth = torch.tensor(np.ascontiguousarray(mesh.depth, np.int8), device=dev)
    d_plot = torch.tensor(np.ascontiguousarray(mesh.plotter, np.int8), device=dev)
    d_nbr = torch.tensor(np.ascontiguousarray(mesh.nbr, np.int32).reshape(-1), device=dev)
    dm = _native.DeviceMesh(n_cells=mesh.n, depth=d_depth.data_ptr(), plotter=d_plot.data_ptr(),
                            nbr=d_nbr.data_ptr(), scaling=S, g=swe.G, dry_tolerance=1e-3,
                            min_depth=dmin, max_depth=dmax, solver=solver, boundary=bc)
    o = [torch.full_like(q[0], np.nan) for _ in range(3)]
    dtc = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
    nref = torch.full((1,), -1, dtype=torch.int64, device=dev)
    rf = torch.full((mesh.n,), 7, dtype=torch.int8, device=dev)
    _native.check("swe_step_device_noalloc", lib.swe_step_device_noalloc(
        C.byref(dm), *[t.data_ptr() for t in q], *[t.data_ptr() for t in o], float(dt),
        int(bool(forward)), dtc.data_ptr(), nref.data_ptr(), rf.data_ptr(), stream))
    torch.cuda.synchronize()
    outs.append((o, dtc, nref, rf))
    return outs


def assert_same(a, b):
    (oa, da, na, ra), (ob, db, nb, rb) = a, b
    for x, y in zip(oa, ob):
        assert torch_equal_bits(x, y)
    assert torch_equal_bits(da, db)
    assert int(na.item()) == int(nb.item())
    assert bool((ra == rb).all())


def torch_equal_bits(x, y):
    return np.array_equal(x.cpu().numpy().view(np.uint64), y.cpu().numpy().view(np.uint64))


def test_noalloc_step_matches_handle_step_twice():
    torch = pytest.importorskip("torch")
    S = 50.0
    # first call: a uniform depth-12 mesh (8192 cells), walls, forward
    mesh = host.mesh_uniform(12)
    st = wet_dry_state(mesh, 3, S)
    dt = host.initial_dt(mesh, st, 0.45) * S
    a, b = step_both(torch, mesh, st, S, dt, True, 10, 14, swe.BC_WALL, swe.SOLVER_AUG_RIEMANN)
    assert_same(a, b)
    assert np.isfinite(a[0][0].cpu().numpy()).all()
    # second call, smaller: the reference's adaptive mesh of depths 3..11
    # (fewer cells, several depths, a last chunk that is not full), backward
    z = golden_io.load("adapt_bath_d3_11.npz")
    mesh2 = host.mesh_from_depths(z["depth"])
    assert mesh2.n < mesh.n and mesh2.n % 256 != 0
    st2 = wet_dry_state(mesh2, 4, S)
    dt2 = 0.5 * dt
    a, b = step_both(torch, mesh2, st2, S, dt2, False, 3, 11, swe.BC_OUTFLOW, swe.SOLVER_FWAVE)
    assert_same(a, b)


def test_noalloc_rejects_meshes_beyond_its_scratch():
    pytest.importorskip("torch")
    lib = _native.load()
    dm = _native.DeviceMesh(n_cells=(1 << 24) + 1, depth=8, plotter=8, nbr=8, scaling=1.0,
                            g=swe.G, dry_tolerance=1e-3, min_depth=0, max_depth=24, solver=0,
                            boundary=0)
    rc = lib.swe_step_device_noalloc(C.byref(dm), 8, 16, 24, 32, 40, 48, 56, 1.0, 1, 64, None,
                                     None, None)
    assert rc == _native.SWE_ERR_ARG
    assert "out of range" in _native.last_error()
