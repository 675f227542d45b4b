"""Argument and mesh validation of the C ABI (samoa_amd/csrc/swe_api.hip).

Every check below fails before the library touches a device, so these run
without a GPU: a bad call returns a negative status with a message and never
aborts."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from samoa_amd import _native

OK, ERR_ARG, ERR_HIP = 0, -1, -2


@pytest.fixture(scope="module")
def lib():
    return _native.load()


def mesh(d=4):
    plotter, _, nbr = po.mesh_uniform(d)
    n = plotter.shape[0]
    return (np.full(n, d, np.int8), np.ascontiguousarray(plotter, np.int8),
            np.ascontiguousarray(nbr, np.int32).reshape(-1))


def desc(depth, plotter, nbr, **kw):
    p = dict(scaling=1.0, g=9.80665, dry_tolerance=0.01, courant=0.45, min_depth=0,
             max_depth=min(int(depth.max()), 60), solver=0, boundary=0)
    p.update(kw)
    return _native.MeshDesc(n_cells=kw.pop("n_cells", len(depth)), depth=depth.ctypes.data,
                            plotter=plotter.ctypes.data, nbr=nbr.ctypes.data,
                            **{k: v for k, v in p.items() if k != "n_cells"})


def create(lib, d):
    h = C.c_void_p()
    rc = lib.swe_create(C.byref(d), C.byref(h))
    if rc == OK:
        lib.swe_destroy(h)
    return rc


def last_error(lib):
    buf = C.create_string_buffer(512)
    lib.swe_last_error(buf, 512)
    return buf.value.decode()


def test_valid_mesh_passes_validation(lib):
    dep, plo, nbr = mesh()
    # on a machine without a GPU the first device call fails after validation
    assert create(lib, desc(dep, plo, nbr)) in (OK, ERR_HIP)


@pytest.mark.parametrize("field,value,msg", [
    ("n_cells", 0, "n_cells"),
    ("solver", 7, "unknown solver"),
    ("boundary", 3, "unknown boundary"),
    ("dry_tolerance", 0.0, "dry_tolerance"),
    ("dry_tolerance", float("nan"), "dry_tolerance"),
    ("scaling", 0.0, "scaling"),
    ("g", -1.0, "g must be"),
    ("max_depth", 61, "max_depth"),
])
def test_bad_constants(lib, field, value, msg):
    dep, plo, nbr = mesh()
    assert create(lib, desc(dep, plo, nbr, **{field: value})) == ERR_ARG
    assert msg in last_error(lib)


def test_bad_topology(lib):
    dep, plo, nbr = mesh()
    cases = []
    p = plo.copy(); p[3] = 0; cases.append((dep, p, nbr, "plotter[3]"))
    d = dep.copy(); d[5] = 61; cases.append((d, plo, nbr, "depth[5]"))
    b = nbr.copy(); b[7] = -2; cases.append((dep, plo, b, "nbr[7] = -2 invalid"))
    b = nbr.copy(); b[7] = len(dep); cases.append((dep, plo, b, "invalid"))
    b = nbr.copy(); b[3 * 2] = 2; cases.append((dep, plo, b, "invalid"))  # self neighbour
    # break the symmetry of one interior edge
    c = int(np.flatnonzero(nbr >= 0)[0] // 3)
    k = int(np.flatnonzero(nbr[3 * c:3 * c + 3] >= 0)[0])
    o = int(nbr[3 * c + k])
    b = nbr.copy(); b[3 * o:3 * o + 3] = np.where(b[3 * o:3 * o + 3] == c, -1, b[3 * o:3 * o + 3])
    cases.append((dep, plo, b, "not symmetric"))
    # flip one plotter type: the shared edges' normals no longer cancel
    p = plo.copy(); p[c] = -p[c]; cases.append((dep, p, nbr, "disagree on the normal"))
    # swap two cells' order without touching adjacency: no longer a curve
    perm = np.arange(len(dep))
    perm[[1, 5]] = perm[[5, 1]]
    inv = np.argsort(perm)
    nb2 = nbr.reshape(-1, 3)[perm]
    nb2 = np.where(nb2 >= 0, inv[np.maximum(nb2, 0)], -1).astype(np.int32).reshape(-1)
    cases.append((dep[perm].copy(), plo[perm].copy(), nb2, "not in Sierpinski order"))
    for d_, p_, n_, msg in cases:
        assert create(lib, desc(d_, p_, n_)) == ERR_ARG, msg
        assert msg in last_error(lib), (msg, last_error(lib))


def test_batch_constants_must_match(lib):
    dep, plo, nbr = mesh()
    descs = (_native.MeshDesc * 2)(desc(dep, plo, nbr), desc(dep, plo, nbr, solver=1))
    h = C.c_void_p()
    assert lib.swe_batch_create(2, C.cast(descs, C.c_void_p), C.byref(h)) == ERR_ARG
    assert "must match scenario 0" in last_error(lib)
    assert lib.swe_batch_create(0, C.cast(descs, C.c_void_p), C.byref(h)) == ERR_ARG


def test_null_handles(lib):
    d = C.c_double()
    n = C.c_int64()
    assert lib.swe_step(None, 1.0, 1, C.byref(d), C.byref(n), None) == ERR_ARG
    assert lib.swe_upload(None, None, None, None, None) == ERR_ARG
    assert lib.swe_destroy(None) == OK
    assert lib.swe_batch_destroy(None) == OK
    assert lib.swe_batch_cells(None) == -1
    assert lib.swe_batch_offset(None, 0) == -1
    assert lib.swe_create(None, None) == ERR_ARG
