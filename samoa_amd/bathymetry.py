"""Gridded bathymetry for the SWE step: the data side of the reference's
ASAGI path (SWE.f90:95-167, SWE_initialize.f90:149-187).

The reference opens a bathymetry and a displacement grid with ASAGI, takes
the domain from their extents (SWE.f90:133-137) and samples them wherever a
cell needs b.  ASAGI and netCDF-4 are not part of this build; grids in the
netCDF classic format (CDF-1/CDF-2, which ASAGI also reads) are read with
scipy's reader.  A grid is node-registered: values at the points of its
coordinate variables, the extent from the first to the last of them.

Sampling is bilinear in the grid cell that holds the point, clamped to the
last cell, and -5000 m outside the nodes' extent (the reference's assumption
outside its ASAGI data), with the arithmetic of the device's remesher
(samoa_amd/csrc/swe_adapt.hip raster_sample), so that b sampled here for an
initial state and b the device samples for a cell it splits agree bit for
bit.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Raster:
    """values[ny, nx] at the world points (x0 + i dx, y0 + j dy)."""

    values: np.ndarray
    x0: float
    y0: float
    dx: float
    dy: float

    @property
    def extent(self):
        """((xmin, xmax), (ymin, ymax)) of the nodes."""
        ny, nx = self.values.shape
        return ((self.x0, self.x0 + (nx - 1) * self.dx), (self.y0, self.y0 + (ny - 1) * self.dy))


def _uniform_axis(c: np.ndarray, name: str):
    c = np.asarray(c, np.float64)
    if c.ndim != 1 or c.size < 2:
        raise ValueError(f"coordinate {name}: need a 1-D axis of at least 2 points")
    d = np.diff(c)
    if not (np.all(d > 0) and np.allclose(d, d[0], rtol=1e-9, atol=0.0)):
        raise ValueError(f"coordinate {name}: spacing is not uniform and increasing")
    return float(c[0]), float((c[-1] - c[0]) / (c.size - 1))


def read_netcdf_grid(path: str, var: str | None = None) -> Raster:
    """A 2-D grid variable of a netCDF classic file.  `var` defaults to the
    only 2-D variable; its dimensions (y, x) must have coordinate variables
    of the same names with uniform spacing."""
    from scipy.io import netcdf_file

    with netcdf_file(path, "r", mmap=False) as nc:
        names = [k for k, v in nc.variables.items() if len(v.dimensions) == 2]
        if var is None:
            if len(names) != 1:
                raise ValueError(f"{path}: choose one of the 2-D variables {names}")
            var = names[0]
        v = nc.variables[var]
        if len(v.dimensions) != 2:
            raise ValueError(f"{path}: {var} is not 2-D")
        ydim, xdim = v.dimensions
        if xdim not in nc.variables or ydim not in nc.variables:
            raise ValueError(f"{path}: {var} needs coordinate variables {ydim}, {xdim}")
        x0, dx = _uniform_axis(nc.variables[xdim][:], xdim)
        y0, dy = _uniform_axis(nc.variables[ydim][:], ydim)
        values = np.array(v[:], dtype=np.float64)
    return Raster(np.ascontiguousarray(values), x0, y0, dx, dy)


def asagi_domain(bath: Raster, displ: Raster | None = None):
    """(scaling, (offset_x, offset_y)) of the domain as SWE.f90:133-137 sets
    cfg%scaling and cfg%offset from the grids' extents: the square of the
    bathymetry's larger extent, centred on the displacement and clamped into
    the bathymetry."""
    (bx0, bx1), (by0, by1) = bath.extent
    (dx0, dx1), (dy0, dy1) = (displ or bath).extent
    scaling = max(bx1 - bx0, by1 - by0)
    off = np.array([0.5 * (dx0 + dx1), 0.5 * (dy0 + dy1)]) - 0.5 * scaling
    off = np.minimum(np.maximum(off, [bx0, by0]), np.array([bx1, by1]) - scaling)
    return scaling, (float(off[0]), float(off[1]))


OUTSIDE = -5000.0  # the sea floor the reference assumes outside its data


def sample(r: Raster, x, y) -> np.ndarray:
    """Bilinear samples at world points (x, y) (arrays), clamped to the last
    cell; the expression of the device's raster_sample, term for term.  A
    point outside the nodes' extent gets OUTSIDE, as get_bathymetry_at_position
    gives outside the ASAGI grid (SWE_initialize.f90:164-171)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    ny, nx = r.values.shape
    inside = ((r.x0 <= x) & (r.y0 <= y) & (x <= r.x0 + float(nx - 1) * r.dx) &
              (y <= r.y0 + float(ny - 1) * r.dy))
    fx = (x - r.x0) / r.dx
    fy = (y - r.y0) / r.dy
    ix = np.clip(np.trunc(fx).astype(np.int64), 0, nx - 2)
    iy = np.clip(np.trunc(fy).astype(np.int64), 0, ny - 2)
    wx = fx - ix.astype(np.float64)
    wy = fy - iy.astype(np.float64)
    v = r.values
    b = ((1.0 - wx) * (1.0 - wy) * v[iy, ix] + wx * (1.0 - wy) * v[iy, ix + 1] +
         (1.0 - wx) * wy * v[iy + 1, ix] + wx * wy * v[iy + 1, ix + 1])
    return np.where(inside, b, OUTSIDE)


def data_depth(r: Raster, scaling: float) -> int:
    """The refinement depth at which a cell is as large as a grid cell,
    nint(log(scaling^2 / (dx dy)) / log(2)) (SWE_initialize.f90:129; nint
    rounds half away from zero)."""
    v = np.log(scaling ** 2 / (r.dx * r.dy)) / np.log(2.0)
    return int(np.sign(v) * np.floor(abs(v) + 0.5))


def integrate_levels(depth: np.ndarray, dmax: int, ddepth: int) -> np.ndarray:
    """Halvings over which get_bathymetry_at_element averages gridded data
    under _ADAPT_INTEGRATE (SWE_initialize.f90:127-133, the ASAGI branch):
    min(16, min(dmax + 3, ddepth + 3) - d); <= 0 means the centroid."""
    return np.minimum(16, min(dmax + 3, ddepth + 3) - depth.astype(np.int64))


def cell_bathymetry(r: Raster, coords: np.ndarray, scaling: float, offset, depth=None,
                    dmax=None) -> np.ndarray:
    """b of each cell of unit-square corners coords[n, 3, 2] (entry, right
    angle, exit), world point = scaling * x + offset.

    Without depth: the raster at the centroid (x1 + x2 + x3) / 3, as the
    reference's _ADAPT_SAMPLE build evaluates it.  With depth and dmax: the
    _ADAPT_INTEGRATE average of get_bathymetry_at_element
    (SWE_initialize.f90:106-147), which the reference is built with and the
    device's remesher applies to split cells: over L = integrate_levels(...)
    halvings of the cell, 0.5 (value of the first half + value of the
    second) with the halves (x1, m, x2), (x2, m, x3), m = 0.5 (x1 + x3), and
    the raster at the centroid at the bottom."""
    ox, oy = (offset, offset) if np.isscalar(offset) else offset
    if depth is None:
        c = (coords[:, 0] + coords[:, 1] + coords[:, 2]) / 3.0
        return sample(r, scaling * c[:, 0] + ox, scaling * c[:, 1] + oy)
    levels = np.maximum(integrate_levels(np.asarray(depth), int(dmax), data_depth(r, scaling)), 0)
    out = np.empty(coords.shape[0])
    for L in np.unique(levels):
        sel = np.nonzero(levels == L)[0]
        out[sel] = _integrate(r, coords[sel], int(L), scaling, ox, oy)
    return out


def _integrate(r: Raster, co: np.ndarray, L: int, scaling: float, ox: float, oy: float):
    """Leaves of L halvings in depth-first order, each level's pairs combined
    as the recursion does (0.5 (first + second))."""
    if L == 0:
        c = (co[:, 0] + co[:, 1] + co[:, 2]) / 3.0
        return sample(r, scaling * c[:, 0] + ox, scaling * c[:, 1] + oy)
    m = 0.5 * (co[:, 0] + co[:, 2])
    first = np.stack([co[:, 0], m, co[:, 1]], axis=1)
    second = np.stack([co[:, 1], m, co[:, 2]], axis=1)
    return 0.5 * (_integrate(r, first, L - 1, scaling, ox, oy) +
                  _integrate(r, second, L - 1, scaling, ox, oy))
