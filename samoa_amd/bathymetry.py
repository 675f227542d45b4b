"""Gridded bathymetry for the SWE step: the data side of the reference's
ASAGI path (SWE.f90:95-167, SWE_initialize.f90:149-187).

The reference opens a bathymetry and a displacement grid with ASAGI, takes
the domain from their extents (SWE.f90:133-137) and samples them wherever a
cell needs b.  ASAGI and netCDF-4 are not part of this build; grids in the
netCDF classic format (CDF-1/CDF-2, which ASAGI also reads) are read with
scipy's reader.  A grid is node-registered: values at the points of its
coordinate variables, the extent from the first to the last of them.

Sampling is bilinear in the grid cell that holds the point, clamped to the
last cell, with the arithmetic of the device's remesher
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


def sample(r: Raster, x, y) -> np.ndarray:
    """Bilinear samples at world points (x, y) (arrays), clamped to the last
    cell; the expression of the device's raster_sample, term for term."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    ny, nx = r.values.shape
    fx = (x - r.x0) / r.dx
    fy = (y - r.y0) / r.dy
    ix = np.clip(np.trunc(fx).astype(np.int64), 0, nx - 2)
    iy = np.clip(np.trunc(fy).astype(np.int64), 0, ny - 2)
    wx = fx - ix.astype(np.float64)
    wy = fy - iy.astype(np.float64)
    v = r.values
    return ((1.0 - wx) * (1.0 - wy) * v[iy, ix] + wx * (1.0 - wy) * v[iy, ix + 1] +
            (1.0 - wx) * wy * v[iy + 1, ix] + wx * wy * v[iy + 1, ix + 1])


def cell_bathymetry(r: Raster, coords: np.ndarray, scaling: float, offset) -> np.ndarray:
    """b of each cell: the raster at the centroid scaling * (x1 + x2 + x3) / 3
    + offset of its unit-square corners coords[n, 3, 2]."""
    c = (coords[:, 0] + coords[:, 1] + coords[:, 2]) / 3.0
    ox, oy = (offset, offset) if np.isscalar(offset) else offset
    return sample(r, scaling * c[:, 0] + ox, scaling * c[:, 1] + oy)
