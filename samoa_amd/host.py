"""ctypes binding of the Fortran host library (samoa_amd/lib/libswe_host.so).

The host side of the drop-in is Fortran, like the reference: uniform
Sierpinski meshes (fortran/swe_sfc_mesh.f90), scenario initial conditions
(fortran/swe_scenarios.f90) and the swe_run time loop
(fortran/swe_driver.f90).  This module exposes them to the Python harness.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native
from .swe import G, Mesh

HOST_LIB = os.path.join(_native.LIB_DIR, "libswe_host.so")
CLI = os.path.join(os.path.dirname(_native.LIB_DIR), "bin", "swe_mi355x")

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def load():
    global _lib
    if _lib is None:
        _native.load()  # one HIP runtime per process, shared with torch
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built; run `python samoa_amd/build.py`")
        L = C.CDLL(HOST_LIB)
        L.swe_host_mesh_uniform.restype = C.c_int64
        L.swe_host_mesh_uniform.argtypes = [C.c_int, _i8p, _i32p, _dp]
        L.swe_host_initial_dt.restype = C.c_double
        L.swe_host_initial_dt.argtypes = [C.c_int64, _i8p, _dp, _dp, _dp, _dp, C.c_double, C.c_double]
        for name in ("swe_host_reference_dam_break", "swe_host_radial_dam_break"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp]
        L.swe_host_tsunami_draw.restype = None
        L.swe_host_tsunami_draw.argtypes = [C.c_int64, C.c_int, C.c_int] + [C.POINTER(C.c_double)] * 4 + [
            C.POINTER(C.c_int)]
        L.swe_host_tsunami.restype = None
        L.swe_host_tsunami.argtypes = [C.c_int64, _dp, _i8p, C.c_int, C.c_int64] + \
            [C.c_double] * 5 + [_dp] * 4
        L.swe_host_time_loop.restype = C.c_int
        L.swe_host_time_loop.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int64,
                                         C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.swe_host_write_vtu.restype = C.c_int
        L.swe_host_write_vtu.argtypes = [C.c_char_p, C.c_int64, _dp, _i8p, _i8p, _dp, _dp, _dp, _dp,
                                         _i8p, C.c_double, C.c_double, C.c_double, C.c_int]
        L.swe_host_write_pvtu.restype = C.c_int
        L.swe_host_write_pvtu.argtypes = [C.c_char_p, C.c_char_p]
        L.swe_host_gauges_sample.restype = None
        L.swe_host_gauges_sample.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp,
                                             C.c_double, C.c_double, C.c_double, _dp]
        L.swe_host_gauges_write.restype = C.c_int
        L.swe_host_gauges_write.argtypes = [C.c_char_p, C.c_int, _dp, _dp, _dp, C.c_int]
        L.swe_host_tsunami_raster.restype = C.c_int
        L.swe_host_tsunami_raster.argtypes = [C.c_int64, _dp, C.POINTER(C.c_double)]
        L.swe_host_tsunami_adaptive_setup.restype = C.c_int
        L.swe_host_tsunami_adaptive_setup.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.c_int,
                                                      C.c_double, C.c_double,
                                                      C.POINTER(C.c_int64),
                                                      C.POINTER(C.c_double)]
        L.swe_host_adaptive_loop.restype = C.c_int
        L.swe_host_adaptive_loop.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_double,
                                             C.c_double, C.c_int64, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.POINTER(C.c_int),
                                             C.POINTER(C.c_int64)]
        L.swe_host_mesh_from_depths.restype = C.c_int
        L.swe_host_mesh_from_depths.argtypes = [C.c_int64, _i8p, _i8p, _dp]
        _lib = L
    return _lib


def mesh_uniform(depth: int) -> Mesh:
    n = 1 << (depth + 1)
    plotter = np.empty(n, np.int8)
    nbr = np.empty(3 * n, np.int32)
    coords = np.empty(6 * n)
    got = load().swe_host_mesh_uniform(depth, plotter, nbr, coords)
    if got != n:
        raise RuntimeError(f"swe_host_mesh_uniform({depth}) failed: {got}")
    return Mesh(np.full(n, depth, np.int8), plotter, nbr.reshape(n, 3), coords.reshape(n, 3, 2))


def neighbours_from_corners(coords: np.ndarray) -> np.ndarray:
    """nbr[n, 3] of a conforming mesh from its corners[n, 3, 2]: two cells are
    neighbours across edge k (opposite corner k) when the edges share their
    midpoint, a dyadic point (exact in the integer key below for depth <= 60)."""
    c = np.asarray(coords, np.float64)
    n = c.shape[0]
    a, b = c[:, [1, 2, 0]], c[:, [2, 0, 1]]          # the ends of edges 0, 1, 2
    mid = np.rint(0.5 * (a + b) * 2.0 ** 33).astype(np.int64).reshape(3 * n, 2)
    order = np.lexsort((mid[:, 1], mid[:, 0]))
    sm = mid[order]
    same = np.all(sm[1:] == sm[:-1], axis=1)
    i = np.nonzero(same)[0]
    nbr = np.full(3 * n, -1, np.int32)
    s0, s1 = order[i], order[i + 1]
    nbr[s0] = s1 // 3
    nbr[s1] = s0 // 3
    return nbr.reshape(n, 3)


def coords_from_depths(depth) -> tuple:
    """(plotter[n], corners[n, 3, 2]) of the mesh with these depths in
    forward curve order (fortran/swe_sfc_mesh.f90 sfc_mesh_from_depths)."""
    depth = np.ascontiguousarray(depth, np.int8)
    n = depth.shape[0]
    plotter = np.empty(n, np.int8)
    coords = np.empty(6 * n)
    if load().swe_host_mesh_from_depths(n, depth, plotter, coords) != 0:
        raise ValueError("depth sequence is not the leaf sequence of the two root triangles")
    return plotter, coords.reshape(n, 3, 2)


def mesh_from_depths(depth) -> Mesh:
    """The mesh whose cells, in forward curve order, have these depths: the
    leaves of the two root triangles (fortran/swe_sfc_mesh.f90
    sfc_mesh_from_depths), with plotter types, corners and neighbours."""
    depth = np.ascontiguousarray(depth, np.int8)
    n = depth.shape[0]
    plotter = np.empty(n, np.int8)
    coords = np.empty(6 * n)
    if load().swe_host_mesh_from_depths(n, depth, plotter, coords) != 0:
        raise ValueError("depth sequence is not the leaf sequence of the two root triangles")
    coords = coords.reshape(n, 3, 2)
    return Mesh(depth.copy(), plotter, neighbours_from_corners(coords), coords)


def _state(mesh, fn):
    out = [np.empty(mesh.n) for _ in range(4)]
    fn(mesh.n, np.ascontiguousarray(mesh.coords.reshape(-1)), *out)
    return tuple(out)


def reference_dam_break(mesh: Mesh):
    """The reference's built-in scenario (SWE_initialize.f90:349-370), domain
    [-5,5]^2: use scaling 10."""
    return _state(mesh, load().swe_host_reference_dam_break)


def radial_dam_break(mesh: Mesh):
    return _state(mesh, load().swe_host_radial_dam_break)


def tsunami_draw(seed: int, dlo: int = 14, dhi: int = 20) -> dict:
    a, r, x, y = (C.c_double() for _ in range(4))
    d = C.c_int()
    load().swe_host_tsunami_draw(seed, dlo, dhi, C.byref(a), C.byref(r), C.byref(x), C.byref(y),
                                 C.byref(d))
    return dict(seed=seed, amplitude=a.value, ratio=r.value, x0=x.value, y0=y.value, depth=d.value)


def tsunami(mesh: Mesh, seed: int, drytol: float = 1e-3, amplitude=None, ratio=None, x0=None,
            y0=None, dmax=None):
    """Tsunami scenario on the 1000 km square (use scaling 1e6).  b is the
    raster averaged over each cell as the reference's _ADAPT_INTEGRATE build
    does for gridded data, with dmax (default: the mesh's deepest cell) the
    run's maximum depth."""
    p = tsunami_draw(seed)
    amp = p["amplitude"] if amplitude is None else amplitude
    rat = p["ratio"] if ratio is None else ratio
    xx = p["x0"] if x0 is None else x0
    yy = p["y0"] if y0 is None else y0
    depth = np.ascontiguousarray(mesh.depth, np.int8)
    dm = int(depth.max()) if dmax is None else int(dmax)
    out = [np.empty(mesh.n) for _ in range(4)]
    load().swe_host_tsunami(mesh.n, np.ascontiguousarray(mesh.coords.reshape(-1)), depth, dm, seed,
                            amp, rat, xx, yy, drytol, *out)
    return tuple(out)


TSUNAMI_RASTER_NODES = 513


def tsunami_raster(seed: int) -> dict:
    """The tsunami scenario's bathymetry raster (fortran/swe_scenarios.f90
    tsunami_raster): values[j, i] at world (i dx, j dx), for
    EulerTimestep.set_bathymetry_raster, which samples it exactly as
    tsunami() does for the cells that remeshing splits."""
    g = TSUNAMI_RASTER_NODES
    values = np.empty(g * g)
    dx = C.c_double()
    got = load().swe_host_tsunami_raster(int(seed), values, C.byref(dx))
    assert got == g
    return dict(values=values.reshape(g, g), x0=0.0, y0=0.0, dx=dx.value, dy=dx.value, offset=0.0)


def tsunami_adaptive_setup(op, seed: int, tol: float, dmax: int, drytol: float,
                           courant: float):
    """The Fortran host's adaptive tsunami set-up (swe_driver.f90
    tsunami_adaptive_setup) on an EulerTimestep holding a uniform mesh:
    initial refinement where the source varies by more than `tol` across a
    cell, scenario state, bathymetry raster.  Returns (cells, dt0)."""
    from . import _native

    n, dt0 = C.c_int64(), C.c_double()
    rc = load().swe_host_tsunami_adaptive_setup(op._h, int(seed), float(tol), int(dmax),
                                                float(drytol), float(courant), C.byref(n),
                                                C.byref(dt0))
    _native.check("tsunami_adaptive_setup", rc)
    op.n = n.value
    op.mesh = None
    return n.value, dt0.value


def adaptive_loop(op, nsteps: int, nadapt: int, refine_tol: float, coarsen_tol: float, k0: int,
                  dt: float, time: float, forward: bool):
    """The Fortran host's time loop with remeshing (swe_driver.f90
    adaptive_loop).  Returns (dt of the next step, time, forward, cells
    stepped in total)."""
    from . import _native

    d, t, f, cells = C.c_double(dt), C.c_double(time), C.c_int(int(bool(forward))), C.c_int64()
    rc = load().swe_host_adaptive_loop(op._h, int(nsteps), int(nadapt), float(refine_tol),
                                       float(coarsen_tol), int(k0), C.byref(d), C.byref(t),
                                       C.byref(f), C.byref(cells))
    _native.check("adaptive_loop", rc)
    op.n = int(op._lib.swe_cells(op._h))
    op.mesh = None
    return d.value, t.value, bool(f.value), cells.value


def initial_dt(mesh: Mesh, state, courant: float, g: float = G) -> float:
    """The init traversal's dt estimate (SWE_initialize.f90:327-335, :240)."""
    eta, hu, hv, b = (np.ascontiguousarray(a, dtype=np.float64) for a in state)
    return load().swe_host_initial_dt(mesh.n, np.ascontiguousarray(mesh.depth, np.int8), eta, hu,
                                      hv, b, g, courant)


def time_loop(op, *, dt: float, time: float = 0.0, forward: bool = False, nmax: int = -1,
              tmax: float = -1.0, tout: float = -1.0, host_steps: bool = False):
    """The Fortran swe_run loop (SWE.f90:380-429) on an EulerTimestep handle.
    Returns (dt, time, forward, steps)."""
    d, t = C.c_double(dt), C.c_double(time)
    fw = C.c_int(int(bool(forward)))
    steps = C.c_int64()
    rc = load().swe_host_time_loop(op._h, op.courant, tmax, tout, nmax, int(bool(host_steps)),
                                   C.byref(d), C.byref(t), C.byref(fw), C.byref(steps))
    _native.check("swe_host_time_loop", rc)
    return d.value, t.value, bool(fw.value), steps.value


def write_vtu(path: str, mesh: Mesh, state, *, time: float, scaling: float, offset: float,
              forward: bool, refine=None) -> None:
    """The reference's VTK layout (fortran/swe_vtu.f90): cells in the order of
    an output traversal running forward or backward along the curve."""
    eta, hu, hv, b = (np.ascontiguousarray(a, dtype=np.float64) for a in state)
    rf = np.zeros(mesh.n, np.int8) if refine is None else np.ascontiguousarray(refine, np.int8)
    rc = load().swe_host_write_vtu(os.fsencode(path), mesh.n,
                                   np.ascontiguousarray(mesh.coords.reshape(-1), np.float64),
                                   np.ascontiguousarray(mesh.plotter, np.int8),
                                   np.ascontiguousarray(mesh.depth, np.int8), eta, b, hu, hv, rf,
                                   float(time), float(scaling), float(offset), int(bool(forward)))
    if rc != 0:
        raise OSError(f"swe_host_write_vtu({path}) failed with I/O status {rc}")


def write_pvtu(path: str, piece: str) -> None:
    rc = load().swe_host_write_pvtu(os.fsencode(path), piece.encode())
    if rc != 0:
        raise OSError(f"swe_host_write_pvtu({path}) failed with I/O status {rc}")


def gauges_sample(mesh: Mesh, state, points, *, time: float, scaling: float,
                  offset: float) -> np.ndarray:
    """Probe values (fortran/swe_gauges.f90) at world points [(x, y), ...]:
    rows x_unit, y_unit, p(1), p(2), h, b, centroid distance, time in days."""
    eta, hu, hv, b = (np.ascontiguousarray(a, dtype=np.float64) for a in state)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((len(pts), 8))
    load().swe_host_gauges_sample(mesh.n, np.ascontiguousarray(mesh.coords.reshape(-1), np.float64),
                                  eta, hu, hv, b, len(pts), np.ascontiguousarray(pts[:, 0]),
                                  np.ascontiguousarray(pts[:, 1]), float(scaling), float(offset),
                                  float(time), out)
    return out


def gauges_write(path: str, points, values: np.ndarray, full: bool = False) -> None:
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    rc = load().swe_host_gauges_write(os.fsencode(path), len(pts), np.ascontiguousarray(pts[:, 0]),
                                      np.ascontiguousarray(pts[:, 1]),
                                      np.ascontiguousarray(values, np.float64), int(bool(full)))
    if rc != 0:
        raise OSError(f"swe_host_gauges_write({path}) failed with I/O status {rc}")
