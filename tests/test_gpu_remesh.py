"""The device remesher (swe_adapt) against the reference's runs that remesh
before every step and against the oracle's remeshing at larger sizes: the
new mesh (depths, plotter types, neighbour table) and the transferred state
bit for bit, through the flags the steps leave on the device and through
flags passed from the host, across buffer growth and reuse."""
import numpy as np
import pytest

import pyoracle as po
from samoa_amd import _native, host, swe
from test_remesh import RUNS, random_case

pytestmark = pytest.mark.gpu


def params(meta):
    return dict(scaling=meta["scaling"], dry_tolerance=meta["drytol"], courant=meta["courant"],
                min_depth=meta["dmin"], max_depth=meta["dmax"], boundary=meta["bc"])


def check_mesh(op, depth):
    want = host.mesh_from_depths(depth)
    d, p, nb = op.download_mesh()
    assert np.array_equal(d, want.depth)
    assert np.array_equal(p, want.plotter)
    assert np.array_equal(nb, want.nbr)


@pytest.mark.parametrize("name,outs,meta", RUNS, ids=[r[0] for r in RUNS])
@pytest.mark.parametrize("flags_from", ["device", "host"])
def test_follows_reference_remeshing(name, outs, meta, flags_from):
    print("case:", name, flags_from, flush=True)
    mesh = host.mesh_from_depths(outs[0]["depth"])
    op = swe.EulerTimestep(mesh, **params(meta))
    op.keep_refinement_flags(flags_from == "device")
    op.upload(*outs[0]["state"])
    dt = po.initial_dt(mesh.depth, outs[0]["state"], meta["courant"])
    t = 0.0
    for k in range(1, len(outs)):
        a, b = outs[k - 1], outs[k]
        if k == 1 or flags_from == "host":
            flags = a["refine"] if k > 1 else np.zeros(len(a["depth"]), np.int8)
            n, sweeps = op.adapt(flags)
        else:
            n, sweeps = op.adapt()
        assert n == len(b["depth"]) and sweeps >= 1
        check_mesh(op, b["depth"])
        res = op.traverse(dt, not b["forward"], flags=True)
        t += dt
        assert t == b["time"]
        eta, hu, hv, bb = op.download()
        assert np.array_equal(np.stack([eta, hu, hv, bb]), b["state"]), f"step {k}"
        assert np.array_equal(res.refine, b["refine"]), f"step {k}: flags"
        dt = meta["courant"] * res.dt_cfl
    op.close()


@pytest.mark.parametrize("seed", range(6))
def test_random_requests_match_oracle(seed):
    """Rounds of random requests on meshes of 10^4..10^5 cells: mesh and state
    equal the oracle's after every round; a step on the result equals the
    oracle's step."""
    mesh, rng = random_case(seed)
    target = int(rng.integers(10_000, 100_000))
    while mesh.n < target:  # grow it by random refinement
        flags = rng.choice(np.array([0, 1], np.int8), size=mesh.n, p=[0.4, 0.6])
        mesh = host.mesh_from_depths(po.adapt(mesh.depth, mesh.plotter, mesh.nbr, flags,
                                              np.zeros((4, mesh.n)), dmin=1, dmax=16)[0])
    print("case: seed", seed, "cells", mesh.n, flush=True)
    st = np.stack([5.0 + rng.random(mesh.n), rng.normal(size=mesh.n), rng.normal(size=mesh.n),
                   -rng.random(mesh.n)])
    op = swe.EulerTimestep(mesh, scaling=1.0, dry_tolerance=1e-3, min_depth=1, max_depth=16)
    op.upload(*st)
    cur = mesh
    for r in range(4):
        flags = rng.choice(np.array([-1, 0, 1], np.int8), size=cur.n, p=[0.45, 0.35, 0.2])
        d2, p2, s2, nb2, _, sweeps = po.adapt(cur.depth, cur.plotter, cur.nbr, flags, st, dmin=1,
                                              dmax=16, coords=cur.coords)
        n, sw = op.adapt(flags)
        assert n == len(d2)
        d, p, nb = op.download_mesh()
        assert np.array_equal(d, d2) and np.array_equal(p, p2) and np.array_equal(nb, nb2), r
        assert np.array_equal(np.stack(op.download()), s2), r
        cur = host.mesh_from_depths(d2)
        st = s2
    eo, ho, vo, rf, cand, nref = po.step(cur.depth, cur.plotter, cur.nbr, st, 1e-4, scaling=1.0,
                                         drytol=1e-3, dmin=1, dmax=16, forward=False)
    res = op.traverse(1e-4, False, flags=True)
    assert np.array_equal(np.stack(op.download()[:3]), np.stack([eo, ho, vo]))
    assert np.array_equal(res.refine, rf) and res.dt_cfl == cand
    op.close()


def test_merging_pairs_across_count_tiles():
    """The count pass checks a merging cell's sibling in LDS when both lie in
    its tile of 2048 cells, and in memory when the pair straddles two tiles.
    A uniform depth-12 mesh with one cell refined shifts the sibling pairs to
    odd first indices, so that pairs straddle the tile boundaries; then every
    cell requests coarsening.  Mesh and state equal the oracle's."""
    tile = 2048  # cells per workgroup of the count pass
    base = host.mesh_uniform(12)
    flags = np.zeros(base.n, np.int8)
    flags[0] = 1
    cur = host.mesh_from_depths(po.adapt(base.depth, base.plotter, base.nbr, flags,
                                         np.zeros((4, base.n)), dmin=1, dmax=13)[0])
    dep = cur.depth.astype(np.int64)
    off = np.concatenate([[0], np.cumsum(np.int64(1) << (60 - dep))[:-1]])
    first = off % (np.int64(1) << (61 - dep)) == 0
    straddle = [t - 1 for t in range(tile, cur.n, tile)
                if first[t - 1] and dep[t - 1] == dep[t]]
    assert len(straddle) >= 2, straddle  # the case this test is for
    rng = np.random.default_rng(5)
    st = np.stack([5.0 + rng.random(cur.n), rng.normal(size=cur.n), rng.normal(size=cur.n),
                   -rng.random(cur.n)])
    op = swe.EulerTimestep(cur, scaling=1.0, dry_tolerance=1e-3, min_depth=1, max_depth=13)
    op.upload(*st)
    for r in range(2):
        print("case: round", r, "cells", cur.n, "straddling pairs at", straddle, flush=True)
        flags = np.full(cur.n, -1, np.int8)
        d2, p2, s2, nb2, _, _ = po.adapt(cur.depth, cur.plotter, cur.nbr, flags, st, dmin=1,
                                         dmax=13, coords=cur.coords)
        n, _ = op.adapt(flags)
        assert n == len(d2) and n < cur.n, r
        d, p, nb = op.download_mesh()
        assert np.array_equal(d, d2) and np.array_equal(p, p2) and np.array_equal(nb, nb2), r
        assert np.array_equal(np.stack(op.download()), s2), r
        cur = host.mesh_from_depths(d2)
        st = s2
        dep = cur.depth.astype(np.int64)
        off = np.concatenate([[0], np.cumsum(np.int64(1) << (60 - dep))[:-1]])
        first = off % (np.int64(1) << (61 - dep)) == 0
        straddle = [t - 1 for t in range(tile, cur.n, tile)
                    if first[t - 1] and dep[t - 1] == dep[t]]
    op.close()


@pytest.mark.parametrize("pattern", ["all", "random"])
def test_tiny_meshes_down_to_the_two_root_cells(pattern):
    """Meshes smaller than one tile, from the two-cell mesh: four rounds that
    refine (every cell, or a random half), then rounds that coarsen every
    cell back to the two root cells, which cannot coarsen further.  Mesh and
    state equal the oracle's after every round."""
    rng = np.random.default_rng(11)
    cur = host.mesh_uniform(0)
    st = np.stack([5.0 + rng.random(2), rng.normal(size=2), rng.normal(size=2), -rng.random(2)])
    op = swe.EulerTimestep(cur, scaling=1.0, dry_tolerance=1e-3, min_depth=0, max_depth=6)
    op.upload(*st)
    for r, f in enumerate([1] * 4 + [-1] * 7):
        flags = np.full(cur.n, f, np.int8)
        if pattern == "random" and f == 1:
            flags = rng.choice(np.array([0, 1], np.int8), size=cur.n)
        print("case:", pattern, "round", r, "cells", cur.n, flush=True)
        d2, p2, s2, nb2, _, _ = po.adapt(cur.depth, cur.plotter, cur.nbr, flags, st, dmin=0,
                                         dmax=6, coords=cur.coords)
        n, _ = op.adapt(flags)
        assert n == len(d2), r
        d, p, nb = op.download_mesh()
        assert np.array_equal(d, d2) and np.array_equal(p, p2) and np.array_equal(nb, nb2), r
        assert np.array_equal(np.stack(op.download()), s2), r
        cur = host.mesh_from_depths(d2)
        st = s2
    assert cur.n == 2
    op.close()


def test_short_sweep_batch_is_completed():
    """A call without requests changes no word, so the next call starts with
    a single sweep; requests that need more make the count pass find words
    off the fixpoint, and further sweeps and a recount follow (more than one
    sweep reported).  Mesh and state still equal the oracle's, in every case
    and in at least one that takes that path."""
    taken = 0
    for seed in range(10, 16):
        mesh, rng = random_case(seed)
        while mesh.n < 4000:  # tens of sweep tiles
            grow = rng.choice(np.array([0, 1], np.int8), size=mesh.n, p=[0.5, 0.5])
            mesh = host.mesh_from_depths(po.adapt(mesh.depth, mesh.plotter, mesh.nbr, grow,
                                                  np.zeros((4, mesh.n)), dmin=1, dmax=16)[0])
        st = rng.normal(size=(4, mesh.n))
        op = swe.EulerTimestep(mesh, scaling=1.0, dry_tolerance=1e-3, min_depth=1, max_depth=16)
        op.upload(*st)
        n, _ = op.adapt(np.zeros(mesh.n, np.int8))
        assert n == mesh.n
        flags = rng.choice(np.array([-1, 0, 1], np.int8), size=mesh.n, p=[0.3, 0.3, 0.4])
        d2, p2, s2, nb2, _, _ = po.adapt(mesh.depth, mesh.plotter, mesh.nbr, flags, st, dmin=1,
                                         dmax=16, coords=mesh.coords)
        n, sweeps = op.adapt(flags)
        print("case: seed", seed, "cells", mesh.n, "sweeps", sweeps, flush=True)
        taken += sweeps > 1
        assert n == len(d2)
        d, p, nb = op.download_mesh()
        assert np.array_equal(d, d2) and np.array_equal(p, p2) and np.array_equal(nb, nb2)
        assert np.array_equal(np.stack(op.download()), s2)
        op.close()
    assert taken > 0, "no case needed more than the first sweep"


def test_growth_and_reuse_of_buffers():
    """Refine everything from depth 4 to 12 (buffers grow), coarsen back
    (spare buffers reused), against the uniform meshes and the oracle."""
    mesh = host.mesh_uniform(4)
    rng = np.random.default_rng(7)
    st = rng.normal(size=(4, mesh.n))
    op = swe.EulerTimestep(mesh, min_depth=2, max_depth=12)
    op.upload(*st)
    for d in range(5, 13):
        n, _ = op.adapt(np.ones(op.n, np.int8))
        assert n == 2 ** (d + 1)
        st = np.repeat(st, 2, axis=1)
    check_mesh(op, np.full(op.n, 12, np.int8))
    assert np.array_equal(np.stack(op.download()), st)
    for d in range(11, 1, -1):
        n, _ = op.adapt(np.full(op.n, -1, np.int8))
        assert n == 2 ** (d + 1)
        st = 0.5 * (st[:, 0::2] + st[:, 1::2])
    check_mesh(op, np.full(op.n, 2, np.int8))
    assert np.array_equal(np.stack(op.download()), st)
    # a request beyond min_depth is dropped
    n, _ = op.adapt(np.full(op.n, -1, np.int8))
    assert n == op.n == 8
    op.close()


def test_adapt_errors():
    mesh = host.mesh_uniform(3)
    op = swe.EulerTimestep(mesh)
    with pytest.raises(_native.SweError) as e:
        op.adapt()
    assert e.value.code == _native.SWE_ERR_STATE  # no state yet
    op.upload(*np.zeros((4, mesh.n)))
    with pytest.raises(_native.SweError) as e:
        op.adapt()  # the flags of a step were not kept
    assert e.value.code == _native.SWE_ERR_STATE
    op.close()
    # half of the unit square: not a whole mesh
    half = host.mesh_uniform(3)
    n = half.n // 2
    nbr = np.where(half.nbr[:n] >= n, -1, half.nbr[:n])
    part = swe.Mesh(half.depth[:n].copy(), half.plotter[:n].copy(), nbr, half.coords[:n].copy())
    op = swe.EulerTimestep(part)
    op.upload(*np.zeros((4, n)))
    with pytest.raises(_native.SweError) as e:
        op.adapt(np.zeros(n, np.int8))
    assert e.value.code == _native.SWE_ERR_ARG and "leaf sequence" in str(e.value)
    op.close()


def test_split_cells_sample_the_bathymetry_raster():
    """With the tsunami scenario's raster set, the cells a refinement creates
    take b from it averaged over the cell as the reference's _ADAPT_INTEGRATE
    build does (8 samples per cell here), bit-identical to the host's
    scenario code on the finer mesh, while a merge still averages its
    pair."""
    d = 10
    mesh = host.mesh_uniform(d)
    st = host.tsunami(mesh, seed=3, drytol=1e-3)
    op = swe.EulerTimestep(mesh, scaling=1e6, dry_tolerance=1e-3, min_depth=d - 1,
                           max_depth=d + 1)
    op.upload(*st)
    r = host.tsunami_raster(3)
    op.set_bathymetry_raster(r["values"], r["x0"], r["y0"], r["dx"], r["dy"], r["offset"])
    op.adapt(np.ones(mesh.n, np.int8))
    fine = host.tsunami(host.mesh_uniform(d + 1), seed=3, drytol=1e-3)
    b = op.download()[3]
    assert np.array_equal(b, fine[3])
    op.adapt(np.full(op.n, -1, np.int8))
    assert np.array_equal(op.download()[3], 0.5 * (fine[3][0::2] + fine[3][1::2]))
    op.close()


def eta_jump_flags(mesh, st, drytol, refine_tol, coarsen_tol, dmin, dmax):
    """The elevation-jump criterion restated in numpy: j = the largest
    |eta - eta_n| over the wet neighbours n of a wet cell (0 otherwise)."""
    eta, b = st[0], st[3]
    wet = eta - b > drytol
    j = np.zeros(mesh.n)
    for k in range(3):
        o = mesh.nbr[:, k]
        ok = (o >= 0) & wet & wet[np.maximum(o, 0)]
        j = np.where(ok, np.maximum(j, np.abs(eta - eta[np.maximum(o, 0)])), j)
    return np.where((mesh.depth < dmax) & (j > refine_tol), 1,
                    np.where((mesh.depth > dmin) & (j < coarsen_tol), -1, 0)).astype(np.int8)


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("coarsen", [True, False])
def test_steps_leave_eta_jump_flags(seed, coarsen):
    """With swe_keep_eta_jump_flags every step's flags follow the
    elevation-jump criterion on the state it reads (numpy restatement), on
    adaptive meshes with dry land, in both directions; the remeshing they
    drive equals the oracle's with those flags."""
    mesh, rng = random_case(seed, n_rounds=4)
    print("case: seed", seed, "coarsen", coarsen, "cells", mesh.n, flush=True)
    st = np.stack(host.tsunami(mesh, seed=seed, drytol=1e-3))
    # a shelf of dry land along one side
    cen = mesh.centroids()
    st[3] = np.where(cen[:, 0] > 0.85, 50.0, st[3])
    st[0] = np.maximum(st[0], np.minimum(st[3], 0.0))
    dmin, dmax = 1, 14
    j = eta_jump_flags(mesh, st, 1e-3, 0.0, np.inf, 0, 99)  # any wet jump refines
    assert (j == 1).any()
    eta, b = st[0], st[3]
    jumps = np.abs(eta[:, None] - eta[np.maximum(mesh.nbr, 0)])
    rtol = float(np.quantile(jumps[jumps > 0], 0.7))
    ctol = rtol / 8 if coarsen else 0.0
    op = swe.EulerTimestep(mesh, scaling=1e6, dry_tolerance=1e-3, min_depth=dmin, max_depth=dmax)
    op.upload(*st)
    op.keep_eta_jump_flags(rtol, ctol)
    dt, fwd = 5.0, bool(seed % 2)
    for step in range(3):
        cur = np.stack(op.download())
        res = op.traverse(dt, fwd, flags=True)
        want = eta_jump_flags(mesh, cur, 1e-3, rtol, ctol, dmin, dmax)
        assert np.array_equal(res.refine, want), step
        assert res.n_refine == int((want == 1).sum())
        fwd = not fwd
    after = np.stack(op.download())
    d2, p2, s2, nb2, _, _ = po.adapt(mesh.depth, mesh.plotter, mesh.nbr, want, after, dmin=dmin,
                                     dmax=dmax, coords=mesh.coords)
    n, _ = op.adapt()
    assert n == len(d2)
    d, p, nb = op.download_mesh()
    assert np.array_equal(d, d2) and np.array_equal(p, p2) and np.array_equal(nb, nb2)
    assert np.array_equal(np.stack(op.download()), s2)
    op.close()


def test_adaptive_loop_with_step_flags_matches_oracle():
    """Fifteen time steps of the adaptive loop, remeshing before every step
    from the elevation-jump flags the step before left: the oracle's step and
    remeshing with the numpy restatement of the flags, against the device's
    loop, mesh and state bit for bit after every remeshing and step."""
    mesh, rng = random_case(21)
    while mesh.n < 3000:
        grow = rng.choice(np.array([0, 1], np.int8), size=mesh.n, p=[0.4, 0.6])
        mesh = host.mesh_from_depths(po.adapt(mesh.depth, mesh.plotter, mesh.nbr, grow,
                                              np.zeros((4, mesh.n)), dmin=1, dmax=14)[0])
    st = np.stack(host.tsunami(mesh, seed=7, drytol=1e-3))
    st[0] *= 50.0  # a source strong enough to move the mesh within a few steps
    cen = mesh.centroids()
    st[3] = np.where(cen[:, 0] > 0.85, 50.0, st[3])
    st[0] = np.maximum(st[0], np.minimum(st[3], 0.0))
    dmin, dmax, drytol = 2, 14, 1e-3
    jumps = np.abs(st[0][:, None] - st[0][np.maximum(mesh.nbr, 0)])
    rtol = float(np.quantile(jumps[jumps > 0], 0.9))
    ctol = rtol / 20
    op = swe.EulerTimestep(mesh, scaling=1e6, dry_tolerance=drytol, min_depth=dmin, max_depth=dmax)
    op.upload(*st)
    op.keep_eta_jump_flags(rtol, ctol)
    cur, fwd, dt = mesh, False, 20.0
    flags = None
    changed = 0
    for k in range(15):
        print("case: step", k, "cells", cur.n, flush=True)
        if k > 0:
            d2, p2, s2, nb2, _, _ = po.adapt(cur.depth, cur.plotter, cur.nbr, flags, st, dmin=dmin,
                                             dmax=dmax, coords=cur.coords)
            n, _ = op.adapt()
            assert n == len(d2), k
            d, p, nb = op.download_mesh()
            assert np.array_equal(d, d2) and np.array_equal(p, p2) and np.array_equal(nb, nb2), k
            assert np.array_equal(np.stack(op.download()), s2), k
            changed += n != cur.n or not np.array_equal(d2, cur.depth)
            cur, st = host.mesh_from_depths(d2), s2
            fwd = not fwd
        flags = eta_jump_flags(cur, st, drytol, rtol, ctol, dmin, dmax)
        eo, ho, vo, _, cand, _ = po.step(cur.depth, cur.plotter, cur.nbr, st, dt, scaling=1e6,
                                         drytol=drytol, dmin=dmin, dmax=dmax, forward=fwd)
        res = op.traverse(dt, fwd, flags=True)
        assert np.array_equal(res.refine, flags), k
        assert res.dt_cfl == cand, k
        st = np.stack([eo, ho, vo, st[3]])
        assert np.array_equal(np.stack(op.download()), st), k
        fwd = not fwd
    assert changed >= 5
    op.close()


def test_eta_jump_flags_drive_adapt():
    """swe_flag_eta_jumps against its definition restated in numpy, through
    the remeshing it drives (compared with the oracle's remeshing)."""
    mesh = host.mesh_uniform(9)
    st = host.tsunami(mesh, seed=1, drytol=1e-3)
    eta, b = st[0], st[3]
    wet = eta - b > 1e-3
    j = np.zeros(mesh.n)
    for k in range(3):
        o = mesh.nbr[:, k]
        ok = (o >= 0) & wet & wet[np.maximum(o, 0)]
        j = np.where(ok, np.maximum(j, np.abs(eta - eta[np.maximum(o, 0)])), j)
    tol = np.quantile(j[j > 0], 0.8)
    flags = np.where((mesh.depth < 12) & (j > tol), 1,
                     np.where((mesh.depth > 5) & (j < tol / 10), -1, 0)).astype(np.int8)
    assert (flags == 1).any() and (flags == -1).any()
    d2, p2, s2, nb2, _, _ = po.adapt(mesh.depth, mesh.plotter, mesh.nbr, flags, st, dmin=5,
                                     dmax=12, coords=mesh.coords)
    op = swe.EulerTimestep(mesh, scaling=1e6, dry_tolerance=1e-3, min_depth=5, max_depth=12)
    op.upload(*st)
    op.flag_eta_jumps(tol, tol / 10)
    n, _ = op.adapt()
    assert n == len(d2)
    d, p, nb = op.download_mesh()
    assert np.array_equal(d, d2) and np.array_equal(nb, nb2)
    assert np.array_equal(np.stack(op.download()), s2)
    op.close()


def test_run_continues_the_device_chain():
    """swe_run with dt0 <= 0 continues the device dt chain, and without time
    and dt_next it does not wait: 7 steps in pieces equal 7 steps at once."""
    mesh = host.mesh_uniform(8)
    st = host.tsunami(mesh, seed=2, drytol=1e-3)
    dt0 = host.initial_dt(mesh, st, 0.45)
    ops = [swe.EulerTimestep(mesh, scaling=1e6, dry_tolerance=1e-3) for _ in range(2)]
    for op in ops:
        op.upload(*st)
    t_all, dt_all = ops[0].run(7, dt0, False)
    ops[1].run(2, dt0, False)  # backward, forward
    ops[1].run(1, None, False, wait=False)
    ops[1].run(4, None, True, wait=False)
    t, dt = ops[1].run(0, None, True)
    assert (t, dt) == (t_all, dt_all)
    assert np.array_equal(np.stack(ops[0].download()), np.stack(ops[1].download()))
    for op in ops:
        op.close()


@pytest.mark.parametrize("d,dmax", [(9, 10), (4, 16)])
def test_split_cells_sample_a_netcdf_grid(d, dmax):
    """A netCDF bathymetry grid placed with the reference's domain rule:
    the cells a refinement creates take b from it averaged as the
    reference's _ADAPT_INTEGRATE build does for gridded data, bit-identical
    to the host's average on the finer mesh and, for some cells, to the
    oracle's recursion; with a coarse mesh and a deep maximum depth each new
    cell averages 2^9 samples."""
    import os
    import tempfile

    from samoa_amd import bathymetry as bt
    from test_bathymetry import write_grid

    rng = np.random.default_rng(5)
    x = 1000.0 * np.arange(41)
    y = 2000.0 + 1000.0 * np.arange(31)
    z = (-3000.0 + 500.0 * rng.normal(size=(31, 41))).astype(np.float32)
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "bath.nc")
        write_grid(f, x, y, z)
        r = bt.read_netcdf_grid(f)
    S, off = bt.asagi_domain(r)
    mesh = host.mesh_uniform(d)
    print("case: netCDF grid, depth", d, "dmax", dmax, flush=True)
    b = bt.cell_bathymetry(r, mesh.coords, S, off, depth=mesh.depth, dmax=dmax)
    eta = np.zeros(mesh.n)
    op = swe.EulerTimestep(mesh, scaling=S, dry_tolerance=1e-3, min_depth=d, max_depth=dmax)
    op.upload(eta, eta, eta, b)
    op.set_bathymetry_raster(r.values, r.x0, r.y0, r.dx, r.dy, off)
    op.adapt(np.ones(mesh.n, np.int8))
    fine = host.mesh_uniform(d + 1)
    got = op.download()[3]
    assert np.array_equal(got, bt.cell_bathymetry(r, fine.coords, S, off, depth=fine.depth,
                                                  dmax=dmax))
    # the domain's square reaches past the grid in y: cells there take the
    # reference's constant sea floor (SWE_initialize.f90:164-171)
    assert (got == bt.OUTSIDE).any() and (got != bt.OUTSIDE).any()
    levels = po.integrate_levels(d + 1, dmax, bt.data_depth(r, S))
    assert levels == (3 if d == 9 else 9)
    f = po.bilinear(r.values, r.x0, r.y0, r.dx, r.dy)
    for i in (0, fine.n // 3, fine.n - 1):
        want = po.integrate_bathymetry(lambda x, y: f(S * x + off[0], S * y + off[1]),
                                       *(tuple(c) for c in fine.coords[i]), levels)
        assert got[i] == want
    op.close()
