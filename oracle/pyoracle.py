"""ctypes wrapper around the oracle C restatement (oracle/swe_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (samoa_amd/) never
imports it.

Parity pins of the functions written here in Python: integrate_levels and
integrate_bathymetry are pinned bit for bit by the reference's adaptive run
built with _ADAPT_INTEGRATE (tests/test_bathymetry.py).  bilinear and
data_depth are parity unpinned: the reference samples its grids through
ASAGI, a third-party library absent from it, so they are checked only
against closed forms (bilinear functions, a worked example).

Also holds the readers for the reference's VTU dumps (SWE_xml_output.f90:220-241,
LIB_VTK_IO raw appended big-endian data) used by tests/golden/make_golden.py.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libswe_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

BC_OUTFLOW = 0
BC_WALL = 1
SOLVER_AUG = 0
SOLVER_FWAVE = 1
G = 9.80665

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.swo_leg_sizes.argtypes = [_dp]
        L.swo_riemanntype.argtypes = [C.c_double] * 6 + [C.POINTER(C.c_double)] * 3 + [
            C.POINTER(C.c_int)] * 2
        L.swo_geoclaw.argtypes = [C.c_int] + [C.c_double] * 10 + [_dp, _dp, C.POINTER(C.c_double)]
        L.swo_flux.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
        L.swo_normals.argtypes = [C.c_int, _dp]
        L.swo_mesh_uniform.argtypes = [C.c_int, _i8p, _dp, _i32p]
        L.swo_mesh_uniform.restype = C.c_int64
        step_args = [C.c_int64, _i8p, _i8p, _i32p, C.c_double, C.c_double, C.c_double,
                     C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                     _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i8p, C.POINTER(C.c_int64)]
        L.swo_step.argtypes = step_args
        L.swo_step.restype = C.c_double
        L.swo_step_omp.argtypes = step_args
        L.swo_step_omp.restype = C.c_double
        L.swo_initial_dt.argtypes = [C.c_int64, _i8p, _dp, _dp, _dp, _dp, C.c_double, C.c_double]
        L.swo_initial_dt.restype = C.c_double
        L.swo_adapt.argtypes = [C.c_int64, _i8p, _i8p, _i32p, C.c_void_p, _i8p, C.c_int, C.c_int,
                                _dp, _dp, _dp, _dp, C.c_int64, _i8p, _i8p, C.c_void_p,
                                C.c_void_p, _dp, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.swo_adapt.restype = C.c_int64
        _lib = L
    return _lib


def leg_sizes() -> np.ndarray:
    """l(k) for k = -1..62 at index k+1."""
    out = np.zeros(64)
    lib().swo_leg_sizes(out)
    return out


def riemanntype(hL, hR, uL, uR, drytol, g=G):
    hm, s1, s2 = C.c_double(), C.c_double(), C.c_double()
    r1, r2 = C.c_int(), C.c_int()
    lib().swo_riemanntype(hL, hR, uL, uR, drytol, g, C.byref(hm), C.byref(s1), C.byref(s2),
                          C.byref(r1), C.byref(r2))
    return hm.value, s1.value, s2.value, r1.value, r2.value


def geoclaw(hL, hR, huL, huR, hvL, hvR, bL, bR, drytol, g=G, solver=SOLVER_AUG):
    nl, nr, s = np.zeros(3), np.zeros(3), C.c_double()
    lib().swo_geoclaw(solver, hL, hR, huL, huR, hvL, hvR, bL, bR, drytol, g, nl, nr, C.byref(s))
    return nl, nr, s.value


def flux(n, QL, QR, drytol, g=G, solver=SOLVER_AUG):
    fl, fr = np.zeros(4), np.zeros(4)
    lib().swo_flux(solver, np.ascontiguousarray(n, dtype=np.float64),
                   np.ascontiguousarray(QL, dtype=np.float64),
                   np.ascontiguousarray(QR, dtype=np.float64), drytol, g, fl, fr)
    return fl, fr


def normals(p: int) -> np.ndarray:
    out = np.zeros(6)
    lib().swo_normals(p, out)
    return out.reshape(3, 2)


def mesh_uniform(depth: int):
    """Returns (plotter[N] int8, coords[N,3,2] float64, nbr[N,3] int32) in forward SFC order."""
    n = 1 << (depth + 1)
    plotter = np.zeros(n, np.int8)
    coords = np.zeros(6 * n)
    nbr = np.zeros(3 * n, np.int32)
    got = lib().swo_mesh_uniform(depth, plotter, coords, nbr)
    assert got == n
    return plotter, coords.reshape(n, 3, 2), nbr.reshape(n, 3)


def step(depth, plotter, nbr, state, dt, *, scaling=1.0, drytol=0.01, dmin=None, dmax=None,
         bc=BC_OUTFLOW, solver=SOLVER_AUG, forward=True, g=G, threads=False):
    """One reference Euler traversal on (eta, hu, hv, b).  Returns
    (eta', hu', hv', refine, dt_candidate_min, n_refine)."""
    eta, hu, hv, b = (np.ascontiguousarray(a, dtype=np.float64) for a in state)
    n = eta.shape[0]
    depth = np.ascontiguousarray(depth, dtype=np.int8)
    d = int(depth.max()) if n else 0
    dmin = d if dmin is None else dmin
    dmax = d if dmax is None else dmax
    eo, ho, vo = np.empty_like(eta), np.empty_like(eta), np.empty_like(eta)
    rf = np.zeros(n, np.int8)
    nref = C.c_int64()
    fn = lib().swo_step_omp if threads else lib().swo_step
    cand = fn(n, depth, np.ascontiguousarray(plotter, dtype=np.int8),
              np.ascontiguousarray(nbr, dtype=np.int32).reshape(-1), scaling, g, drytol,
              dmin, dmax, bc, solver, int(bool(forward)), dt, eta, hu, hv, b, eo, ho, vo, rf,
              C.byref(nref))
    return eo, ho, vo, rf, cand, nref.value


def adapt(depth, plotter, nbr, flags, state, *, dmin, dmax, coords=None):
    """One remeshing step (swe_oracle_adapt.c).  Returns (depth', plotter',
    state'[4, n'], nbr'[n', 3] or None, coords'[n', 3, 2] or None, sweeps);
    the neighbour table and corners only when coords are given."""
    depth = np.ascontiguousarray(depth, dtype=np.int8)
    n = depth.shape[0]
    cap = 4 * n
    st = [np.ascontiguousarray(a, dtype=np.float64) for a in state]
    d_o, p_o = np.zeros(cap, np.int8), np.zeros(cap, np.int8)
    s_o = [np.zeros(cap) for _ in range(4)]
    nb_o = c_o = None
    cptr = None
    if coords is not None:
        cin = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
        nb_o = np.zeros(3 * cap, np.int32)
        c_o = np.zeros(6 * cap)
        cptr = cin.ctypes.data
    sweeps = C.c_int()
    m = lib().swo_adapt(n, depth, np.ascontiguousarray(plotter, dtype=np.int8),
                        np.ascontiguousarray(nbr, dtype=np.int32).reshape(-1), cptr,
                        np.ascontiguousarray(flags, dtype=np.int8), dmin, dmax, *st, cap, d_o, p_o,
                        None if nb_o is None else nb_o.ctypes.data,
                        None if c_o is None else c_o.ctypes.data, *s_o, C.byref(sweeps))
    if m < 0:
        raise RuntimeError(f"swo_adapt failed: {m}")
    return (d_o[:m].copy(), p_o[:m].copy(), np.stack([a[:m] for a in s_o]),
            None if nb_o is None else nb_o[:3 * m].reshape(m, 3).copy(),
            None if c_o is None else c_o[:6 * m].reshape(m, 3, 2).copy(), sweeps.value)


def initial_dt(depth, state, courant=0.45, g=G):
    eta, hu, hv, b = (np.ascontiguousarray(a, dtype=np.float64) for a in state)
    return lib().swo_initial_dt(eta.shape[0], np.ascontiguousarray(depth, dtype=np.int8),
                                eta, hu, hv, b, g, courant)


# --------------------------------------------------------------------------
# Reference VTU reader (SWE_xml_output.f90:220-241; LIB_VTK_IO 'binary' =
# raw appended, big-endian, each block prefixed by an Int32 byte count).

_VTK_TYPES = {"Float64": ">f8", "Int32": ">i4", "Int8": ">i1"}


def integrate_levels(depth: int, dmax: int, data_depth=None) -> int:
    """Recursion depth of get_bathymetry_at_element under _ADAPT_INTEGRATE
    (SWE_initialize.f90:127-133): min(16, dmax - d) for a bathymetry
    function, min(16, min(dmax + 3, data_depth + 3) - d) for gridded data
    (the ASAGI branch); <= 0 means one sample at the centroid."""
    if data_depth is None:
        return min(16, dmax - depth)
    return min(16, min(dmax + 3, data_depth + 3) - depth)


def data_depth(scaling: float, dx: float, dy: float) -> int:
    """nint(log(scaling**2 / (dx dy)) / log(2)) of SWE_initialize.f90:129;
    nint rounds half away from zero."""
    import math

    v = math.log(scaling ** 2 / (dx * dy)) / math.log(2.0)
    return int(math.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def integrate_bathymetry(f, x1, x2, x3, levels: int) -> float:
    """refine_2D_recursive (SWE_initialize.f90:106-116): over `levels`
    halvings of the triangle (x1, x2, x3) (x2 the right angle, unit-square
    points), 0.5 (value of (x1, m, x2) + value of (x2, m, x3)) with
    m = 0.5 (x1 + x3); at the bottom f at the centroid (x1 + x2 + x3) / 3.
    Pure-Python recursion for small cases."""
    if levels > 0:
        m = (0.5 * (x1[0] + x3[0]), 0.5 * (x1[1] + x3[1]))
        return 0.5 * (integrate_bathymetry(f, x1, m, x2, levels - 1) +
                      integrate_bathymetry(f, x2, m, x3, levels - 1))
    return f((x1[0] + x2[0] + x3[0]) / 3.0, (x1[1] + x2[1] + x3[1]) / 3.0)


def bilinear(values, x0, y0, dx, dy):
    """The bilinear raster sample of the library's remesher (clamped to the
    last cell, -5000 m outside the nodes' extent as get_bathymetry_at_position
    gives outside its grid) as a function of one world point, for
    integrate_bathymetry."""
    ny, nx = values.shape

    def f(X, Y):
        if not (x0 <= X and y0 <= Y and X <= x0 + float(nx - 1) * dx and
                Y <= y0 + float(ny - 1) * dy):
            return -5000.0  # SWE_initialize.f90:164-171
        fx, fy = (X - x0) / dx, (Y - y0) / dy
        ix = min(max(int(fx), 0), nx - 2)
        iy = min(max(int(fy), 0), ny - 2)
        wx, wy = fx - ix, fy - iy
        return ((1.0 - wx) * (1.0 - wy) * values[iy, ix] + wx * (1.0 - wy) * values[iy, ix + 1] +
                (1.0 - wx) * wy * values[iy + 1, ix] + wx * wy * values[iy + 1, ix + 1])

    return f


def read_vtu(path: str) -> dict:
    raw = open(path, "rb").read()
    head_end = raw.index(b"<AppendedData")
    head = raw[:head_end].decode("latin-1")
    start = raw.index(b"_", head_end) + 1
    out = {}
    for m in re.finditer(r'<DataArray type="(\w+)"([^>]*)>', head):
        typ, attrs = m.group(1), m.group(2)
        name = re.search(r'Name="([^"]+)"', attrs).group(1)
        off = int(re.search(r'offset="(\d+)"', attrs).group(1))
        nb = int(np.frombuffer(raw[start + off:start + off + 4], ">i4")[0])
        arr = np.frombuffer(raw[start + off + 4:start + off + 4 + nb], _VTK_TYPES[typ])
        out[name] = arr.astype(arr.dtype.newbyteorder("="))
    return out
