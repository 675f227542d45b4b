"""Long horizons and full sizes, HIP path against the oracle and the
reference's own runs, bit for bit:

- the radial dam break in a closed basin at its full size (depth 14, 32,768
  cells, flat bed, reflecting walls, the augmented Riemann solver, courant
  0.45, 100 steps), through swe_run (device dt chain) and through swe_step
  (host dt chain), against the oracle every 10 steps and against the
  reference's own depth-14 run (SHA-256 digests of its final state);
- 1000 steps of the same scenario at depth 10 against the oracle;
- 300 steps that remesh before every step from the steps' own refinement
  flags (SWE.f90:385-393 with -nadapt 1), against the oracle's step and
  remeshing after every step;
- one step each way of the tsunami on the depth-21 mesh (4,194,304 cells)
  against the oracle's OpenMP step."""
import hashlib
import json
import os

import numpy as np
import pytest

import golden_io
import pyoracle as po
from samoa_amd import host, swe

pytestmark = pytest.mark.gpu

COURANT = float(np.float32(0.45))   # Config.f90:243 parses a default REAL
DRYTOL = float(np.float32(0.01))    # Config.f90:284


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()


def radial(d):
    mesh = host.mesh_uniform(d)
    return mesh, np.stack(host.radial_dam_break(mesh))


def oracle_run(mesh, st, dt, steps, forward, every, **kw):
    """The oracle's states (and times, next dts) after every `every` steps."""
    out, t, fwd = [], 0.0, forward
    for k in range(1, steps + 1):
        e, u, v, rf, cand, _ = po.step(mesh.depth, mesh.plotter, mesh.nbr, st, dt, threads=True,
                                       forward=fwd, **kw)
        st = np.stack([e, u, v, st[3]])
        t += dt
        dt = COURANT * cand
        fwd = not fwd
        if k % every == 0:
            out.append((k, st, t, dt, rf))
    return out


@pytest.fixture(scope="module")
def config1():
    mesh, st = radial(14)
    assert mesh.n == 32768
    dt0 = host.initial_dt(mesh, st, COURANT)
    kw = dict(scaling=1.0, drytol=DRYTOL, bc=po.BC_WALL, dmin=14, dmax=14)
    print("case: config 1 oracle, 100 steps", flush=True)
    return mesh, st, dt0, oracle_run(mesh, st, dt0, 100, True, 10, **kw)


def test_config1_initial_state_is_the_references():
    fx = json.load(open(os.path.join(golden_io.GOLDEN, "radial_d14_n100.json")))
    mesh, st = radial(14)
    for k, a in zip(("eta", "hu", "hv", "b"), st):
        assert digest(a) == fx["initial"]["sha256"][k], k


def test_config1_run_matches_oracle_and_reference(config1):
    mesh, st, dt0, ref = config1
    fx = json.load(open(os.path.join(golden_io.GOLDEN, "radial_d14_n100.json")))
    op = swe.EulerTimestep(mesh, scaling=1.0, dry_tolerance=DRYTOL, courant=COURANT,
                           min_depth=14, max_depth=14, boundary=swe.BC_WALL)
    op.upload(*st)
    # the reference's traversals: forward at odd steps (golden_io.direction_schedule)
    assert golden_io.direction_schedule(100, 100)[:2] == [True, False]
    t, dt = op.run(10, dt0, forward=True, time=0.0)
    for i, (k, want, tw, dtw, _) in enumerate(ref):
        print("case: config 1 swe_run, step", k, flush=True)
        if i > 0:
            t, dt = op.run(10, None, forward=True)
        got = np.stack(op.download())
        assert np.array_equal(got, want), f"step {k}"
        assert t == tw and dt == dtw, f"step {k}"
    assert t == fx["final"]["time"]
    for name, a in zip(("eta", "hu", "hv", "b"), got):
        assert digest(a) == fx["final"]["sha256"][name], name
    op.close()


def test_config1_host_steps_match_oracle_and_reference(config1):
    mesh, st, dt0, ref = config1
    fx = json.load(open(os.path.join(golden_io.GOLDEN, "radial_d14_n100.json")))
    op = swe.EulerTimestep(mesh, scaling=1.0, dry_tolerance=DRYTOL, courant=COURANT,
                           min_depth=14, max_depth=14, boundary=swe.BC_WALL)
    op.upload(*st)
    t, dt, fwd, i = 0.0, dt0, True, 0
    for k in range(1, 101):
        res = op.traverse(dt, fwd, flags=(k % 10 == 0))
        t += dt
        dt = COURANT * res.dt_cfl
        fwd = not fwd
        if k % 10 == 0:
            kk, want, tw, dtw, rf = ref[i]
            i += 1
            assert kk == k
            assert np.array_equal(np.stack(op.download()), want), f"step {k}"
            assert t == tw and dt == dtw and np.array_equal(res.refine, rf), f"step {k}"
    assert hashlib.sha256(res.refine.tobytes()).hexdigest() == fx["final"]["refine_sha256"]
    for name, a in zip(("eta", "hu", "hv"), op.download()[:3]):
        assert digest(a) == fx["final"]["sha256"][name], name
    op.close()


def test_thousand_steps_radial_d10():
    mesh, st = radial(10)
    dt0 = host.initial_dt(mesh, st, COURANT)
    kw = dict(scaling=1.0, drytol=DRYTOL, bc=po.BC_WALL, dmin=10, dmax=10)
    ref = oracle_run(mesh, st, dt0, 1000, False, 100, **kw)
    op = swe.EulerTimestep(mesh, scaling=1.0, dry_tolerance=DRYTOL, courant=COURANT,
                           min_depth=10, max_depth=10, boundary=swe.BC_WALL)
    op.upload(*st)
    t, dt = op.run(100, dt0, forward=False, time=0.0)
    for i, (k, want, tw, dtw, _) in enumerate(ref):
        print("case: radial d10, step", k, flush=True)
        if i > 0:
            t, dt = op.run(100, None, forward=False)
        assert np.array_equal(np.stack(op.download()), want), f"step {k}"
        assert t == tw and dt == dtw, f"step {k}"
    # the walls keep the water in: mass is conserved to round-off
    vol = 0.5 ** 11
    m0, m1 = (st[0] - st[3]).sum() * vol, (want[0] - want[3]).sum() * vol
    assert abs(m1 - m0) <= 1e-12 * m0
    op.close()


def test_three_hundred_steps_remeshing_every_step():
    """The radial dam break on a mesh of depths 3..12 that remeshes before
    every step from the flags of the step before (volume_op's |dQ_h|
    criterion, keep_refinement_flags), against the oracle's step and
    remeshing: mesh, state, flags, dt and time after every step."""
    dmin, dmax = 3, 12
    mesh, st = radial(6)
    op = swe.EulerTimestep(mesh, scaling=1.0, dry_tolerance=DRYTOL, courant=COURANT,
                           min_depth=dmin, max_depth=dmax, boundary=swe.BC_WALL)
    op.upload(*st)
    op.keep_refinement_flags(True)
    kw = dict(scaling=1.0, drytol=DRYTOL, bc=po.BC_WALL, dmin=dmin, dmax=dmax)
    cur, fwd, t = mesh, True, 0.0
    dt = host.initial_dt(mesh, st, COURANT)
    flags, sizes = None, []
    for k in range(300):
        if k % 50 == 0:
            print("case: remeshing loop, step", k, "cells", cur.n, flush=True)
        if k > 0:
            d2, p2, s2, nb2, c2, _ = po.adapt(cur.depth, cur.plotter, cur.nbr, flags, st,
                                              dmin=dmin, dmax=dmax, coords=cur.coords)
            n, _ = op.adapt()
            assert n == len(d2), k
            d, p, nb = op.download_mesh()
            assert np.array_equal(d, d2) and np.array_equal(p, p2) and np.array_equal(nb, nb2), k
            assert np.array_equal(np.stack(op.download()), s2), k
            cur = swe.Mesh(d2, p2, nb2, c2)
            st = s2
            fwd = not fwd  # the adaption traversal flips the direction
        e, u, v, flags, cand, nref = po.step(cur.depth, cur.plotter, cur.nbr, st, dt, forward=fwd,
                                             **kw)
        res = op.traverse(dt, fwd, flags=True)
        st = np.stack([e, u, v, st[3]])
        assert res.dt_cfl == cand and res.n_refine == nref, k
        assert np.array_equal(res.refine, flags), k
        assert np.array_equal(np.stack(op.download()), st), k
        t += dt
        dt = COURANT * cand
        fwd = not fwd
        sizes.append(cur.n)
    # the mesh followed the front: it grew and changed over the run
    assert max(sizes) > 8 * mesh.n and len(set(sizes)) > 30
    op.close()


@pytest.mark.parametrize("forward", [True, False])
def test_tsunami_d21_step_matches_oracle(forward):
    d = 21
    print("case: tsunami d21 forward", forward, flush=True)
    mesh = host.mesh_uniform(d)
    st = np.stack(host.tsunami(mesh, seed=0, drytol=1e-3))
    dt = host.initial_dt(mesh, st, 0.45) * 1.0e6
    op = swe.EulerTimestep(mesh, scaling=1.0e6, dry_tolerance=1e-3, courant=0.45)
    op.upload(*st)
    res = op.traverse(dt, forward, flags=True)
    e, u, v, rf, cand, nref = po.step(mesh.depth, mesh.plotter, mesh.nbr, st, dt, scaling=1.0e6,
                                      drytol=1e-3, forward=forward, threads=True)
    assert res.dt_cfl == cand and res.n_refine == nref
    assert np.array_equal(res.refine, rf)
    got = op.download()
    for name, a, b in zip(("eta", "hu", "hv"), got, (e, u, v)):
        assert np.array_equal(a, b), name
    op.close()
