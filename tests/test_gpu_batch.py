"""Batched ensembles: one launch per step over meshes of uneven size must give
each scenario exactly what a single-mesh run gives, for every launch shape."""
import numpy as np
import pytest

import pyoracle as po
from samoa_amd import host, swe

pytestmark = pytest.mark.gpu

DEPTHS = [6, 9, 12, 9, 4, 13, 0]  # 0: the two-cell mesh


def build():
    meshes = {d: host.mesh_uniform(d) for d in set(DEPTHS)}
    ms = [meshes[d] for d in DEPTHS]
    params, states = [], []
    for s, m in enumerate(ms):
        st = host.tsunami(m, seed=s + 1, drytol=1e-3)
        states.append(st)
        params.append(dict(scaling=1.0e6 * (1 + 0.1 * s), dry_tolerance=1e-3 * (1 + s),
                           courant=0.45 - 0.02 * s))
    return ms, params, states


def test_batch_run_matches_single_runs():
    ms, params, states = build()
    b = swe.Batch(ms, params)
    for s, st in enumerate(states):
        b.upload(s, *st)
    dt0 = np.array([host.initial_dt(m, st, p["courant"]) * p["scaling"] for m, st, p in
                    zip(ms, states, params)])
    t, dtn = b.run(9, dt0, forward=False)
    for s, (m, st, p) in enumerate(zip(ms, states, params)):
        op = swe.EulerTimestep(m, **p)
        op.upload(*st)
        ts, dts = op.run(9, dt0[s], forward=False)
        assert ts == t[s] and dts == dtn[s], f"scenario {s}"
        assert all(np.array_equal(x, y) for x, y in zip(op.download(), b.download(s))), s
        op.close()
    b.close()


def test_batch_step_matches_oracle():
    ms, params, states = build()
    b = swe.Batch(ms, params)
    for s, st in enumerate(states):
        b.upload(s, *st)
    ref = [list(st) for st in states]
    dt = np.array([host.initial_dt(m, st, p["courant"]) * p["scaling"] for m, st, p in
                   zip(ms, states, params)])
    for fwd in (True, False, True):
        dtc, _ = b.step(dt, fwd)
        for s, (m, p) in enumerate(zip(ms, params)):
            e, u, v, _, cand, _ = po.step(m.depth, m.plotter, m.nbr, ref[s], dt[s],
                                          scaling=p["scaling"], drytol=p["dry_tolerance"],
                                          forward=fwd)
            ref[s] = [e, u, v, ref[s][3]]
            assert dtc[s] == cand
        dt = dtc * np.array([p["courant"] for p in params])
    for s in range(len(ms)):
        assert all(np.array_equal(x, y) for x, y in zip(b.download(s), ref[s])), s
    b.close()


@pytest.mark.parametrize("chunk,grid", [(128, 0), (512, 0), (256, 64), (512, 16), (128, 2048)])
def test_launch_shape_does_not_change_results(chunk, grid):
    ms, params, states = build()
    dt0 = np.array([host.initial_dt(m, st, p["courant"]) * p["scaling"] for m, st, p in
                    zip(ms, states, params)])
    results = []
    for shape in [(256, 0), (chunk, grid)]:
        b = swe.Batch(ms, params)
        b.set_tuning(*shape)
        assert b.get_tuning() == shape
        for s, st in enumerate(states):
            b.upload(s, *st)
        t, dtn = b.run(7, dt0, forward=True)
        results.append((t, dtn, [b.download(s) for s in range(len(ms))]))
        b.close()
    (t0, d0, s0), (t1, d1, s1) = results
    assert np.array_equal(t0, t1) and np.array_equal(d0, d1)
    for a, c in zip(s0, s1):
        assert all(np.array_equal(x, y) for x, y in zip(a, c))


def test_batch_rejects_bad_tuning():
    ms, params, _ = build()
    b = swe.Batch(ms[:2], params[:2])
    with pytest.raises(RuntimeError):
        b.set_tuning(100, 0)
    with pytest.raises(RuntimeError):
        b.set_tuning(256, 12)
    b.close()
