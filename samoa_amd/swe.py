"""Python mirror of the SWE Euler-step operator behind the C ABI.

`EulerTimestep` plays the role of the reference's
`t_swe_euler_timestep_traversal` (SWE_euler_timestep.f90:9-54): it is created
once per frozen mesh and `traverse()` replaces `swe%euler%traverse(grid)`
(SWE.f90:393).  `Grid` keeps the scenario bookkeeping the reference keeps in
`t_grid` (r_time, r_dt, traversal direction) and applies
pre/post_traversal_grid_op (SWE_euler_timestep.f90:89-131) around each
traversal exactly as the reference does.  Every call goes through the HIP
library; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _native

SOLVER_AUG_RIEMANN = 0
SOLVER_FWAVE = 1
# compute_lf_flux variants (SWE_euler_timestep.f90:359-446); forward steps only
SOLVER_LF = 2
SOLVER_LLF = 3
SOLVER_LF_BATH = 4
SOLVER_LLF_BATH = 5
BC_OUTFLOW = 0
BC_WALL = 1
ARITH_EXACT = 0        # every operation rounds as the reference's
ARITH_TOLERANCED = 1   # shared reciprocals, refined estimates, within 1e-12 per step
G = 9.80665  # SWE_data_types.f90:34


@dataclass
class Mesh:
    """A frozen triangle mesh in forward Sierpinski order.

    depth[n] int8, plotter[n] int8 (forward plotter type), nbr[n,3] int32
    (neighbour across reference edge k+1, -1 on the boundary), coords[n,3,2]
    node positions in unit-square coordinates (optional)."""

    depth: np.ndarray
    plotter: np.ndarray
    nbr: np.ndarray
    coords: Optional[np.ndarray] = None

    @property
    def n(self) -> int:
        return int(self.depth.shape[0])

    def centroids(self) -> np.ndarray:
        return self.coords.mean(axis=1)


@dataclass
class StepResult:
    dt_cfl: float          # min over cells of vol / sum(len * s)
    n_refine: int          # cells flagged for refinement
    refine: Optional[np.ndarray] = None


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class EulerTimestep:
    """One frozen mesh on the GPU; `traverse` performs one Euler traversal."""

    def __init__(self, mesh: Mesh, *, scaling: float = 1.0, g: float = G,
                 dry_tolerance: float = 0.01, courant: float = 0.45,
                 min_depth: Optional[int] = None, max_depth: Optional[int] = None,
                 solver: int = SOLVER_AUG_RIEMANN, boundary: int = BC_OUTFLOW):
        self._lib = _native.load()
        self.mesh = mesh
        self._depth = np.ascontiguousarray(mesh.depth, dtype=np.int8)
        self._plotter = np.ascontiguousarray(mesh.plotter, dtype=np.int8)
        self._nbr = np.ascontiguousarray(mesh.nbr, dtype=np.int32).reshape(-1)
        dmax = int(self._depth.max()) if max_depth is None else int(max_depth)
        dmin = int(self._depth.min()) if min_depth is None else int(min_depth)
        self.desc = _native.MeshDesc(
            n_cells=mesh.n, depth=_ptr(self._depth), plotter=_ptr(self._plotter),
            nbr=_ptr(self._nbr), scaling=float(scaling), g=float(g),
            dry_tolerance=float(dry_tolerance), courant=float(courant), min_depth=dmin,
            max_depth=dmax, solver=int(solver), boundary=int(boundary))
        h = C.c_void_p()
        _native.check("swe_create", self._lib.swe_create(C.byref(self.desc), C.byref(h)))
        self._h = h
        self.n = mesh.n
        self.courant = float(courant)
        self.solver = int(solver)
        self.stream = 0  # the caller's HIP stream (0: the library's own)
        self.arithmetic = ARITH_EXACT

    # -- state -----------------------------------------------------------
    def upload(self, eta, hu, hv, b) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (eta, hu, hv, b)]
        for a in arrs:
            if a.shape != (self.n,):
                raise ValueError(f"state arrays must have shape ({self.n},), got {a.shape}")
        _native.check("swe_upload", self._lib.swe_upload(self._h, *[_ptr(a) for a in arrs]))

    def download(self):
        out = [np.empty(self.n) for _ in range(4)]
        _native.check("swe_download", self._lib.swe_download(self._h, *[_ptr(a) for a in out]))
        return tuple(out)

    # -- traversals --------------------------------------------------------
    def traverse(self, dt: float, forward: bool, flags: bool = False) -> StepResult:
        dtc = C.c_double()
        nref = C.c_int64()
        rf = np.empty(self.n, np.int8) if flags else None
        _native.check("swe_step", self._lib.swe_step(
            self._h, float(dt), int(bool(forward)), C.byref(dtc), C.byref(nref),
            _ptr(rf) if flags else None))
        return StepResult(dtc.value, nref.value, rf)

    # -- remeshing --------------------------------------------------------
    def keep_refinement_flags(self, on: bool = True) -> None:
        """Every step leaves its refinement flags on the device for adapt()."""
        _native.check("swe_keep_refinement_flags",
                      self._lib.swe_keep_refinement_flags(self._h, int(bool(on))))

    def keep_eta_jump_flags(self, refine_tol: float, coarsen_tol: float) -> None:
        """Every step leaves flags from the surface-elevation jumps to wet
        neighbours of the state it reads (the criterion of flag_eta_jumps)
        on the device for adapt() (swe_keep_eta_jump_flags)."""
        _native.check("swe_keep_eta_jump_flags",
                      self._lib.swe_keep_eta_jump_flags(self._h, float(refine_tol),
                                                        float(coarsen_tol)))

    def flag_eta_jumps(self, refine_tol: float, coarsen_tol: float) -> None:
        """Device refinement flags from surface-elevation jumps to wet
        neighbours (swe_flag_eta_jumps), for adapt() with refine=None."""
        _native.check("swe_flag_eta_jumps",
                      self._lib.swe_flag_eta_jumps(self._h, float(refine_tol), float(coarsen_tol)))

    def adapt(self, refine=None):
        """One remeshing step (swe_adapt): conformity of the requests (the
        last step's flags on the device when `refine` is None), split and
        merge, state transfer.  Returns (new cell count, conformity sweeps).
        The reference's traversal direction flips once per adaption."""
        n_new, sweeps = C.c_int64(), C.c_int()
        rf = None
        if refine is not None:
            rf = np.ascontiguousarray(refine, dtype=np.int8)
            if rf.shape != (self.n,):
                raise ValueError(f"refine must have shape ({self.n},), got {rf.shape}")
        _native.check("swe_adapt", self._lib.swe_adapt(
            self._h, None if rf is None else _ptr(rf), C.byref(n_new), C.byref(sweeps)))
        self.n = n_new.value
        self.mesh = None
        return self.n, sweeps.value

    def set_bathymetry_raster(self, values, x0: float, y0: float, dx: float, dy: float,
                              offset) -> None:
        """Cells that adapt() splits take b from this raster (values[ny, nx]
        over the world points (x0 + i dx, y0 + j dy)), bilinear at their
        centroid scaling * c + offset (offset: a number or (x, y)); values
        None: they copy the parent's b."""
        if values is None:
            _native.check("swe_set_bathymetry_raster", self._lib.swe_set_bathymetry_raster(
                self._h, 0, 0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, None))
            return
        v = np.ascontiguousarray(values, dtype=np.float64)
        ny, nx = v.shape
        ox, oy = (offset, offset) if np.isscalar(offset) else offset
        _native.check("swe_set_bathymetry_raster", self._lib.swe_set_bathymetry_raster(
            self._h, nx, ny, float(x0), float(y0), float(dx), float(dy), float(ox), float(oy),
            _ptr(v)))

    def download_mesh(self):
        """(depth[n], plotter[n], nbr[n, 3]) of the current mesh."""
        depth, plotter = np.empty(self.n, np.int8), np.empty(self.n, np.int8)
        nbr = np.empty(3 * self.n, np.int32)
        _native.check("swe_download_mesh", self._lib.swe_download_mesh(
            self._h, _ptr(depth), _ptr(plotter), _ptr(nbr)))
        return depth, plotter, nbr.reshape(self.n, 3)

    def run(self, n_steps: int, dt0, forward: bool, dt_clamp: float = -1.0,
            time: float = 0.0, wait: bool = True):
        """n_steps traversals with the dt chain on the device, starting at
        simulated time `time`; returns (time after the steps, dt of the next
        step).  dt0=None continues the device chain of the last run;
        wait=False returns None at once, leaving time and dt on the device."""
        t = C.c_double(time)
        dtn = C.c_double()
        _native.check("swe_run", self._lib.swe_run(
            self._h, int(n_steps), -1.0 if dt0 is None else float(dt0), int(bool(forward)),
            float(dt_clamp), C.byref(t) if wait else None, C.byref(dtn) if wait else None))
        return (t.value, dtn.value) if wait else None

    def step_device(self, eta, hu, hv, b, eta_out, hu_out, hv_out, dt, forward, dt_cfl,
                    n_refine=0, refine=0, stream=0) -> None:
        """Device-pointer traversal (integers are device addresses)."""
        _native.check("swe_step_device", self._lib.swe_step_device(
            self._h, eta, hu, hv, b, eta_out, hu_out, hv_out, float(dt), int(bool(forward)),
            dt_cfl, n_refine or None, refine or None, stream or None))

    def forward_device(self, eta, hu, hv, b, eta_out, hu_out, hv_out, dt_dev, forward,
                       dt_cfl=0, masks=0, stream=0) -> None:
        """Traversal with dt read from device memory (dt_dev), optionally
        recording the limiter masks the adjoint needs.  Integers are device
        addresses."""
        _native.check("swe_forward_device", self._lib.swe_forward_device(
            self._h, eta, hu, hv, b, eta_out, hu_out, hv_out, dt_dev, int(bool(forward)),
            dt_cfl or None, masks or None, stream or None))

    def adjoint_device(self, eta, hu, hv, b, masks, leta, lhu, lhv, leta_out, lhu_out, lhv_out,
                       gb, dt_dev, forward, stream=0) -> None:
        """Discrete adjoint of forward_device: (leta_out, lhu_out, lhv_out) =
        (dQ'/dQ)^T (leta, lhu, lhv) and gb += (dQ'/db)^T (leta, lhu, lhv), with
        dt and the masks frozen."""
        _native.check("swe_adjoint_step_device", self._lib.swe_adjoint_step_device(
            self._h, eta, hu, hv, b, masks, leta, lhu, lhv, leta_out, lhu_out, lhv_out, gb,
            dt_dev, int(bool(forward)), stream or None))

    def set_stream(self, stream: int) -> None:
        """Launch on a caller's HIP stream (0 = the library's own)."""
        _native.check("swe_set_stream", self._lib.swe_set_stream(self._h, stream or None))
        self.stream = stream or 0

    def set_arithmetic(self, mode: int) -> None:
        """ARITH_EXACT (bit-identical to the reference) or ARITH_TOLERANCED."""
        _native.check("swe_set_arithmetic", self._lib.swe_set_arithmetic(self._h, int(mode)))
        self.arithmetic = int(mode)

    def set_tuning(self, chunk: int, grid: int = 0) -> None:
        _native.check("swe_set_tuning", self._lib.swe_set_tuning(self._h, chunk, grid))

    def get_tuning(self):
        c, g = C.c_int(), C.c_int64()
        _native.check("swe_get_tuning", self._lib.swe_get_tuning(self._h, C.byref(c), C.byref(g)))
        return c.value, g.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.swe_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many frozen meshes advanced by one launch per step (swe_batch_* C ABI).

    `meshes[s]` with constants `params[s]` (keyword arguments of
    EulerTimestep).  Pass the same Mesh object for scenarios of equal meshes:
    they then share one device copy of the topology."""

    def __init__(self, meshes, params):
        self._lib = _native.load()
        if len(meshes) != len(params) or not meshes:
            raise ValueError("one parameter set per mesh required")
        self.meshes = list(meshes)
        self._keep = []
        descs = (_native.MeshDesc * len(meshes))()
        arrays = {}
        for i, (m, p) in enumerate(zip(meshes, params)):
            if id(m) not in arrays:
                arrays[id(m)] = (np.ascontiguousarray(m.depth, dtype=np.int8),
                                 np.ascontiguousarray(m.plotter, dtype=np.int8),
                                 np.ascontiguousarray(m.nbr, dtype=np.int32).reshape(-1))
            dep, plo, nbr = arrays[id(m)]
            descs[i] = _native.MeshDesc(
                n_cells=m.n, depth=_ptr(dep), plotter=_ptr(plo), nbr=_ptr(nbr),
                scaling=float(p.get("scaling", 1.0)), g=float(p.get("g", G)),
                dry_tolerance=float(p.get("dry_tolerance", 0.01)),
                courant=float(p.get("courant", 0.45)),
                min_depth=int(p.get("min_depth", int(dep.min()))),
                max_depth=int(p.get("max_depth", int(dep.max()))),
                solver=int(p.get("solver", SOLVER_AUG_RIEMANN)),
                boundary=int(p.get("boundary", BC_OUTFLOW)))
        self._keep.append(arrays)
        h = C.c_void_p()
        _native.check("swe_batch_create",
                      self._lib.swe_batch_create(len(meshes), C.cast(descs, C.c_void_p), C.byref(h)))
        self._h = h
        self.n_scen = len(meshes)
        self.courant = np.array([float(p.get("courant", 0.45)) for p in params])
        self.cells = int(self._lib.swe_batch_cells(h))
        self.solver = int(params[0].get("solver", SOLVER_AUG_RIEMANN))
        self.stream = 0  # the caller's HIP stream (0: the library's own)
        self.arithmetic = ARITH_EXACT

    def upload(self, s: int, eta, hu, hv, b) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (eta, hu, hv, b)]
        _native.check("swe_batch_upload", self._lib.swe_batch_upload(self._h, s, *[_ptr(a) for a in arrs]))

    def download(self, s: int):
        n = self.meshes[s].n
        out = [np.empty(n) for _ in range(4)]
        _native.check("swe_batch_download",
                      self._lib.swe_batch_download(self._h, s, *[_ptr(a) for a in out]))
        return tuple(out)

    def step(self, dt, forward: bool):
        """One traversal of every scenario; returns (dt_cfl[s], n_refine[s])."""
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        dtc = np.empty(self.n_scen)
        nref = np.empty(self.n_scen, np.int64)
        _native.check("swe_batch_step", self._lib.swe_batch_step(
            self._h, _ptr(dt), int(bool(forward)), _ptr(dtc), _ptr(nref)))
        return dtc, nref

    def run(self, n_steps: int, dt0, forward: bool, dt_clamp: float = -1.0, time=None):
        """n_steps traversals with per-scenario device dt chains; returns
        (time[s], dt_next[s])."""
        dt0 = np.ascontiguousarray(dt0, dtype=np.float64)
        t = np.zeros(self.n_scen) if time is None else np.array(time, dtype=np.float64)
        dtn = np.empty(self.n_scen)
        _native.check("swe_batch_run", self._lib.swe_batch_run(
            self._h, int(n_steps), _ptr(dt0), int(bool(forward)), float(dt_clamp), _ptr(t),
            _ptr(dtn)))
        return t, dtn

    def offset(self, s: int) -> int:
        """First cell of scenario s in the batch's concatenated state arrays."""
        return int(self._lib.swe_batch_offset(self._h, s))

    def forward_device(self, eta, hu, hv, b, eta_out, hu_out, hv_out, dt_dev, forward,
                       dt_cfl=0, masks=0, stream=0) -> None:
        """One traversal of every scenario over caller-owned concatenated
        arrays; dt_dev holds one dt per scenario (device addresses)."""
        _native.check("swe_batch_forward_device", self._lib.swe_batch_forward_device(
            self._h, eta, hu, hv, b, eta_out, hu_out, hv_out, dt_dev, int(bool(forward)),
            dt_cfl or None, masks or None, stream or None))

    def adjoint_device(self, eta, hu, hv, b, masks, leta, lhu, lhv, leta_out, lhu_out, lhv_out,
                       gb, dt_dev, forward, stream=0) -> None:
        _native.check("swe_batch_adjoint_device", self._lib.swe_batch_adjoint_device(
            self._h, eta, hu, hv, b, masks, leta, lhu, lhv, leta_out, lhu_out, lhv_out, gb,
            dt_dev, int(bool(forward)), stream or None))

    def set_arithmetic(self, mode: int) -> None:
        """ARITH_EXACT or ARITH_TOLERANCED for every scenario."""
        _native.check("swe_batch_set_arithmetic",
                      self._lib.swe_batch_set_arithmetic(self._h, int(mode)))
        self.arithmetic = int(mode)

    def set_tuning(self, chunk: int, grid: int = 0) -> None:
        _native.check("swe_batch_set_tuning", self._lib.swe_batch_set_tuning(self._h, chunk, grid))

    def get_tuning(self):
        c, g = C.c_int(), C.c_int64()
        _native.check("swe_batch_get_tuning",
                      self._lib.swe_batch_get_tuning(self._h, C.byref(c), C.byref(g)))
        return c.value, g.value

    def set_stream(self, stream: int) -> None:
        """Launch on a caller's HIP stream (0 = the library's own)."""
        _native.check("swe_batch_set_stream", self._lib.swe_batch_set_stream(self._h, stream or None))
        self.stream = stream or 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.swe_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Grid:
    """Scenario time bookkeeping of t_grid (r_time, r_dt) and the traversal
    direction, driven exactly as swe_run (SWE.f90:380-429) drives it."""

    op: EulerTimestep
    dt: float
    time: float = 0.0
    forward: bool = False
    max_time: float = -1.0          # cfg%r_max_time (-tmax)
    output_time_step: float = -1.0  # cfg%r_output_time_step (-tout)
    history: list = field(default_factory=list)

    def clamp(self) -> float:
        """pre_traversal_grid_op, SWE_euler_timestep.f90:93-99."""
        dt = self.dt
        if self.max_time > 0.0:
            dt = min(self.max_time, dt)
        if self.output_time_step > 0.0:
            dt = min(self.output_time_step, dt)
        return dt

    def euler_step(self, flags: bool = False) -> StepResult:
        self.dt = self.clamp()
        res = self.op.traverse(self.dt, self.forward, flags=flags)
        # post_traversal_grid_op, :115-128
        self.time = self.time + self.dt
        self.dt = self.op.courant * res.dt_cfl
        self.forward = not self.forward  # grid%reverse(), ringbuffer :420
        self.history.append(res)
        return res

    def other_traversal(self) -> None:
        """An output or remeshing traversal: only flips the direction."""
        self.forward = not self.forward
