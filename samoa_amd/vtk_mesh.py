"""Meshes recovered from the reference's VTK output.

The reference writes every cell's three corners in world coordinates
(SWE_xml_output.f90:333-342), in the order of its output traversal, with
the corners of a cell of plotter type p in the order (x1, x2, x3) of the
positive type |p|: x2 is the right-angle corner, x1 and x3 the ends of the
hypotenuse.  That is enough to rebuild the mesh this library takes:

- the curve enters a cell of type p > 0 at x1 and leaves it at x3 (a child
  pair splits (x1, x2, x3) into (x1, m, x2) and (x2, m, x3),
  SFC_node_traversal.f90:138-143), and enters a cell of type -|p| at x3; two
  consecutive cells meet at the vertex where the first leaves and the second
  enters, which fixes the sign of every plotter type;
- |p| is the orientation of (x1 - x2, x3 - x2), read from a table built from
  the uniform meshes of fortran/swe_sfc_mesh.f90;
- neighbours are the cells that share an edge (edge k opposite corner k).

Adaptive meshes the reference refined (and froze) can thus be stepped here
and compared with its later outputs.
"""
from __future__ import annotations

import numpy as np

from .swe import Mesh

_ORIENT = None


def _unit(v):
    return np.sign(np.round(v / np.abs(v).max(axis=-1, keepdims=True), 6))


def _orientation_table():
    """(direction of x1 - x2, direction of x3 - x2) -> |p|."""
    global _ORIENT
    if _ORIENT is None:
        from . import host

        table = {}
        for d in (1, 2, 3, 4):
            m = host.mesh_uniform(d)
            for p, x in zip(m.plotter, m.coords):
                x1, x2, x3 = (x[0], x[1], x[2]) if p > 0 else (x[2], x[1], x[0])
                key = tuple(_unit(x1 - x2)) + tuple(_unit(x3 - x2))
                if table.setdefault(key, abs(int(p))) != abs(int(p)):
                    raise RuntimeError("plotter orientation table is not unique")
        _ORIENT = table
    return _ORIENT


def mesh_from_corners(corners, depth, scaling: float, offset: float) -> Mesh:
    """corners[n, 3, 2]: world coordinates of each cell's corners as the
    reference writes them, cells in forward curve order; depth[n]."""
    c = np.asarray(corners, dtype=np.float64)
    n = c.shape[0]
    table = _orientation_table()
    a, b, e = c[:, 0], c[:, 1], c[:, 2]
    absp = np.empty(n, np.int8)
    for i in range(n):
        key = tuple(_unit(a[i] - b[i])) + tuple(_unit(e[i] - b[i]))
        absp[i] = table[key]

    def same(u, v):
        return bool(np.array_equal(u, v))

    sign = np.zeros(n, np.int8)
    for i in range(n):
        s = 0
        if i + 1 < n:  # the vertex where the curve leaves cell i
            na, ne = a[i + 1], e[i + 1]
            ex_e = same(e[i], na) or same(e[i], ne)
            ex_a = same(a[i], na) or same(a[i], ne)
            if ex_e != ex_a:
                s = 1 if ex_e else -1
        if s == 0 and i > 0:  # the vertex where it enters
            pa, pe = a[i - 1], e[i - 1]
            en_a = same(a[i], pa) or same(a[i], pe)
            en_e = same(e[i], pa) or same(e[i], pe)
            if en_a != en_e:
                s = 1 if en_a else -1
        if s == 0:
            raise ValueError(f"cell {i}: cannot tell the direction of the curve")
        sign[i] = s
    plotter = (sign * absp).astype(np.int8)
    # corners in this library's order (x1, x2, x3) of the signed type
    ours = np.where((plotter > 0)[:, None, None], c, c[:, ::-1])
    nbr = np.full((n, 3), -1, np.int32)
    seen = {}
    for i in range(n):
        for k in range(3):
            p, q = ours[i, (k + 1) % 3], ours[i, (k + 2) % 3]
            key = tuple(sorted([tuple(p), tuple(q)]))
            if key in seen:
                j, kj = seen.pop(key)
                nbr[i, k], nbr[j, kj] = j, i
            else:
                seen[key] = (i, k)
    coords = (ours - offset) / scaling
    return Mesh(np.asarray(depth, np.int8).copy(), plotter, nbr, coords)
