"""The launch-shape autotuner (samoa_amd/autotune.py): its cache key and
candidate shapes on the CPU; on the GPU, that tuning leaves the state, the
caller's stream and a valid shape behind."""
import numpy as np
import pytest

from samoa_amd import autotune


class FakeBatch:
    def __init__(self, sizes, solver, arithmetic):
        self.meshes = [type("Mesh", (), {"n": n})() for n in sizes]
        self.n_scen = len(sizes)
        self.cells = sum(sizes)
        self.solver = solver
        self.arithmetic = arithmetic


def test_signature_separates_meshes_solver_and_arithmetic():
    a = autotune.signature(FakeBatch([100, 200], 0, 0))
    assert a == autotune.signature(FakeBatch([200, 100], 0, 0))  # order does not matter
    assert a != autotune.signature(FakeBatch([100, 300], 0, 0))
    assert a != autotune.signature(FakeBatch([100, 200], 1, 0))
    assert a != autotune.signature(FakeBatch([100, 200], 0, 1))


def test_candidates():
    cands = autotune.candidates(256)
    assert len(cands) == len(set(cands)) == 4 * len(autotune.CHUNKS)
    for chunk, grid in cands:
        assert chunk in autotune.CHUNKS and grid % 8 == 0 and grid >= 0


@pytest.mark.gpu
def test_tune_restores_state_and_stream():
    import torch

    from samoa_amd import host, swe

    mesh = host.mesh_uniform(12)
    st = np.stack(host.tsunami(mesh, seed=3, drytol=1e-3))
    dt0 = host.initial_dt(mesh, st, 0.45) * 1.0e6
    op = swe.EulerTimestep(mesh, scaling=1.0e6, dry_tolerance=1e-3)
    op.upload(*st)
    op.set_arithmetic(swe.ARITH_TOLERANCED)
    s = torch.cuda.Stream()
    op.set_stream(s.cuda_stream)
    res = autotune.tune(op, dt0, steps=4, repeats=1, use_cache=False)
    print("case: tuned", res["chunk"], res["grid"], flush=True)
    props = torch.cuda.get_device_properties(torch.cuda.current_device())
    assert (res["chunk"], res["grid"]) in autotune.candidates(props.multi_processor_count)
    assert op.get_tuning() == (res["chunk"], res["grid"])
    assert op.stream == s.cuda_stream and op.arithmetic == swe.ARITH_TOLERANCED
    assert np.array_equal(np.stack(op.download()), st)
    op.close()


def test_persistent_grid_needs_a_margin(monkeypatch):
    """The selection rule on made-up timings: a persistent grid within 2 % of
    one workgroup per chunk loses to it; one more than 2 % faster wins."""
    import types

    def run_with(times):
        calls = {}

        class Batch(FakeBatch):
            def download(self, s):
                return (np.zeros(1),) * 4

            def upload(self, s, *st):
                pass

            def set_stream(self, stream):
                pass

            def set_tuning(self, chunk, grid):
                calls["shape"] = (chunk, grid)

            def run(self, n, dt0, forward):
                pass

        class Event:
            def __init__(self, enable_timing=True):
                pass

            def record(self, stream=None):
                pass

            def elapsed_time(self, other):
                return times[calls["shape"]] * 20

        fake_torch = types.SimpleNamespace(
            cuda=types.SimpleNamespace(
                current_device=lambda: 0,
                get_device_properties=lambda d: types.SimpleNamespace(name="x",
                                                                        multi_processor_count=256),
                Stream=lambda device=None: types.SimpleNamespace(cuda_stream=1),
                Event=Event, synchronize=lambda: None))
        monkeypatch.setitem(__import__("sys").modules, "torch", fake_torch)
        b = Batch([10], 0, 0)
        b.n_scen = 1
        return autotune.tune(b, 1.0, steps=20, repeats=1, use_cache=False)

    base = {shape: 1.0 for shape in autotune.candidates(256)}
    near = dict(base)
    near[(256, 0)] = 0.500
    near[(256, 2048)] = 0.495
    r = run_with(near)
    assert (r["chunk"], r["grid"]) == (256, 0)
    far = dict(base)
    far[(256, 0)] = 0.500
    far[(256, 2048)] = 0.480
    r = run_with(far)
    assert (r["chunk"], r["grid"]) == (256, 2048)
