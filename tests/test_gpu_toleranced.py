"""The toleranced arithmetic mode (SWE_ARITH_TOLERANCED) against the oracle,
whose results are bit-identical to the reference's:

- one step from the same state: every field within 1e-12 of its largest
  value, the CFL minimum within 1e-12 relative, refinement-flag mismatches
  counted (only cells whose criterion sits within rounding of a threshold
  may differ);
- long runs: the radial dam break in a closed basin (depth 10, 1000 steps)
  and the tsunami (depth 10, 1000 steps), each with its own dt chain, within
  1e-9 of the largest field value after the run, and the one-step bound
  re-checked from the oracle's state every 100 steps;
- config 1 at its full size (depth 14, 100 steps);
- the batched ensemble in both modes on the same meshes."""
import numpy as np
import pytest

import pyoracle as po
from samoa_amd import host, swe

pytestmark = pytest.mark.gpu

COURANT = float(np.float32(0.45))
DRYTOL = float(np.float32(0.01))


def field_err(got, want):
    """max |got - want| / max |want| per field."""
    out = []
    for x, y in zip(got, want):
        out.append(float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)))
    return out


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
    return np.stack([eta, hu, hv, b])


def one_step_both(mesh, st, dt, fwd, S, drytol, bc, solver, dmin, dmax):
    op = swe.EulerTimestep(mesh, scaling=S, dry_tolerance=drytol, min_depth=dmin, max_depth=dmax,
                           boundary=bc, solver=solver)
    op.set_arithmetic(swe.ARITH_TOLERANCED)
    op.upload(*st)
    res = op.traverse(dt, fwd, flags=True)
    got = np.stack(op.download()[:3])
    op.close()
    e, u, v, rf, cand, _ = po.step(mesh.depth, mesh.plotter, mesh.nbr, st, dt, scaling=S,
                                   drytol=drytol, bc=bc, solver=solver, forward=fwd, dmin=dmin,
                                   dmax=dmax)
    want = np.stack([e, u, v])
    errs = field_err(got, want)
    return errs, abs(res.dt_cfl - cand) / cand, int((res.refine != rf).sum())


@pytest.mark.parametrize("solver", [swe.SOLVER_AUG_RIEMANN, swe.SOLVER_FWAVE])
@pytest.mark.parametrize("bc", [swe.BC_OUTFLOW, swe.BC_WALL])
@pytest.mark.parametrize("fwd", [True, False])
def test_one_step_within_tolerance(solver, bc, fwd):
    S = 100.0
    for d, seed in ((3, 1), (9, 2), (12, 3)):
        mesh = host.mesh_uniform(d)
        st = wet_dry_state(mesh, seed, S)
        dt = host.initial_dt(mesh, st, 0.45) * S
        errs, dterr, flips = one_step_both(mesh, st, dt, fwd, S, 1e-3, bc, solver, d - 2, d + 2)
        print(f"case d={d} solver={solver} bc={bc} fwd={fwd}: field errors {errs} dt {dterr:.2e} "
              f"flag mismatches {flips}", flush=True)
        assert max(errs) <= 1e-12 and dterr <= 1e-12
        assert flips <= max(1, mesh.n // 1000)


def test_one_step_tsunami_within_tolerance():
    d = 14
    mesh = host.mesh_uniform(d)
    st = np.stack(host.tsunami(mesh, seed=4, drytol=1e-3))
    dt = host.initial_dt(mesh, st, 0.45) * 1.0e6
    errs, dterr, flips = one_step_both(mesh, st, dt, True, 1.0e6, 1e-3, swe.BC_OUTFLOW,
                                       swe.SOLVER_AUG_RIEMANN, 12, 16)
    print("tsunami d14:", errs, dterr, flips, flush=True)
    assert max(errs) <= 1e-12 and dterr <= 1e-12 and flips <= mesh.n // 1000


def run_pair(mesh, st, dt0, steps, every, S, drytol, bc, fwd0):
    """The toleranced device run and the oracle's, each with its own dt
    chain; at every `every` steps the accumulated error and one toleranced
    step from the oracle's state."""
    op = swe.EulerTimestep(mesh, scaling=S, dry_tolerance=drytol, courant=COURANT, boundary=bc)
    op.set_arithmetic(swe.ARITH_TOLERANCED)
    op.upload(*st)
    ref, dt, fwd, worst_step = st.copy(), dt0, fwd0, 0.0
    t_dev, dt_dev = op.run(every, dt0, forward=fwd0, time=0.0)
    errs = None
    for k in range(1, steps + 1):
        e, u, v, _, cand, _ = po.step(mesh.depth, mesh.plotter, mesh.nbr, ref, dt, scaling=S,
                                      drytol=drytol, bc=bc, forward=fwd, threads=True)
        ref = np.stack([e, u, v, ref[3]])
        dt = COURANT * cand
        fwd = not fwd
        if k % every == 0:
            got = np.stack(op.download())
            errs = field_err(got[:3], ref[:3])
            print(f"case: step {k} accumulated errors {errs} dt {abs(dt_dev - dt) / dt:.2e}",
                  flush=True)
            one, _, _ = one_step_both(mesh, ref, dt, fwd, S, drytol, bc, swe.SOLVER_AUG_RIEMANN,
                                      int(mesh.depth.min()), int(mesh.depth.max()))
            worst_step = max(worst_step, max(one))
            if k < steps:
                t_dev, dt_dev = op.run(every, None, forward=fwd)
    op.close()
    return errs, worst_step


def test_thousand_steps_radial_d10():
    mesh = host.mesh_uniform(10)
    st = np.stack(host.radial_dam_break(mesh))
    dt0 = host.initial_dt(mesh, st, COURANT)
    errs, worst = run_pair(mesh, st, dt0, 1000, 100, 1.0, DRYTOL, swe.BC_WALL, True)
    assert max(errs) <= 1e-9 and worst <= 1e-12


def test_thousand_steps_tsunami_d10():
    mesh = host.mesh_uniform(10)
    st = np.stack(host.tsunami(mesh, seed=1, drytol=1e-3))
    dt0 = host.initial_dt(mesh, st, 0.45) * 1.0e6
    errs, worst = run_pair(mesh, st, dt0, 1000, 100, 1.0e6, 1e-3, swe.BC_OUTFLOW, False)
    assert max(errs) <= 1e-9 and worst <= 1e-12


def test_config1_full_size():
    mesh = host.mesh_uniform(14)
    st = np.stack(host.radial_dam_break(mesh))
    dt0 = host.initial_dt(mesh, st, COURANT)
    errs, worst = run_pair(mesh, st, dt0, 100, 10, 1.0, DRYTOL, swe.BC_WALL, True)
    assert max(errs) <= 1e-9 and worst <= 1e-12


def test_batch_modes_on_the_same_meshes():
    """A batch in toleranced mode against the same batch exact: every
    scenario within the one-step bound after one step, and exact mode
    unchanged by having run toleranced."""
    meshes = [host.mesh_uniform(d) for d in (8, 11, 9)]
    params = [dict(scaling=1.0e6, dry_tolerance=1e-3, courant=0.45) for _ in meshes]
    states = [np.stack(host.tsunami(m, seed=s, drytol=1e-3)) for s, m in enumerate(meshes)]
    dts = np.array([host.initial_dt(m, st, 0.45) * 1.0e6 for m, st in zip(meshes, states)])
    out = {}
    for mode in (swe.ARITH_EXACT, swe.ARITH_TOLERANCED, swe.ARITH_EXACT):
        b = swe.Batch(meshes, params)
        b.set_arithmetic(mode)
        for s, st in enumerate(states):
            b.upload(s, *st)
        b.step(dts, True)
        out.setdefault(mode, []).append([np.stack(b.download(s)[:3]) for s in range(len(meshes))])
        b.close()
    for x, y in zip(out[swe.ARITH_EXACT][0], out[swe.ARITH_EXACT][1]):
        assert np.array_equal(x, y)
    for x, y in zip(out[swe.ARITH_TOLERANCED][0], out[swe.ARITH_EXACT][0]):
        errs = field_err(x, y)
        assert max(errs) <= 1e-12, errs


def test_elevation_jump_flags_both_modes():
    """The elevation-jump refinement flags a step leaves (swe_keep_eta_jump_
    flags), toleranced against exact on the same state: the toleranced
    kernel takes each wet edge's jump as |(h + b) - (h' + b')|, so only
    cells whose largest jump sits within rounding of a tolerance may differ;
    the exact kernel's flags equal the numpy restatement of the criterion."""
    from test_gpu_remesh import eta_jump_flags
    d = 12
    mesh = host.mesh_uniform(d)
    st = np.stack(host.tsunami(mesh, seed=5, drytol=1e-3))
    st[0] *= 50.0
    jumps = np.abs(st[0][:, None] - st[0][np.maximum(mesh.nbr, 0)])
    rtol = float(np.quantile(jumps[jumps > 0], 0.9))
    dt = host.initial_dt(mesh, st, 0.45) * 1.0e6
    flags = {}
    for mode in (swe.ARITH_EXACT, swe.ARITH_TOLERANCED):
        op = swe.EulerTimestep(mesh, scaling=1.0e6, dry_tolerance=1e-3, min_depth=d - 2,
                               max_depth=d + 2)
        op.set_arithmetic(mode)
        op.upload(*st)
        op.keep_eta_jump_flags(rtol, rtol / 20)
        flags[mode] = op.traverse(dt, True, flags=True).refine
        op.close()
    want = eta_jump_flags(mesh, st, 1e-3, rtol, rtol / 20, d - 2, d + 2)
    flips = int((flags[swe.ARITH_TOLERANCED] != flags[swe.ARITH_EXACT]).sum())
    print(f"case: jump flags d={d}: refine {int((want == 1).sum())} coarsen "
          f"{int((want == -1).sum())} toleranced mismatches {flips}", flush=True)
    assert np.array_equal(flags[swe.ARITH_EXACT], want)
    assert (want == 1).sum() > 0 and (want == -1).sum() > 0
    assert flips <= mesh.n // 1000
