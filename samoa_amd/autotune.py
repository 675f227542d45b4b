"""Launch-shape autotuner for single meshes and batched ensembles.

Tries every compiled chunk size (cells per workgroup, i.e. the LDS tile) and a
few grid sizes (one workgroup per chunk, or persistent grids of k workgroups
per CU) on the operator's own meshes and state, timing graph-replayed steps
of each with HIP events on the stream the kernels run on (the median of a
few repetitions).  The state is saved before the trials and restored before
every candidate, and every candidate's runs start the dt chain at the same
dt0, so each launch shape is timed on the same waves; after the trials the
state is restored once more, so tuning does not advance the simulation.
A persistent grid is chosen only if it beats one workgroup per chunk of the
same size by more than 2 %.  Results are cached per device, mesh sizes,
solver and arithmetic in
~/.cache/samoa_amd/autotune.json; the caller's stream is restored.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

CHUNKS = (128, 256, 512)
# relative margin by which a persistent grid must beat one workgroup per chunk
PERSISTENT_MARGIN = 0.02
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "samoa_amd", "autotune.json")


class _Single:
    """An EulerTimestep seen through the Batch interface (one scenario)."""

    def __init__(self, op):
        self.op = op
        self.meshes = [op.mesh]
        self.n_scen = 1
        self.cells = op.n

    def download(self, s):
        return self.op.download()

    def upload(self, s, *st):
        self.op.upload(*st)

    @property
    def stream(self):
        return self.op.stream

    @property
    def solver(self):
        return self.op.solver

    @property
    def arithmetic(self):
        return self.op.arithmetic

    def set_stream(self, stream):
        self.op.set_stream(stream)

    def set_tuning(self, chunk, grid):
        self.op.set_tuning(chunk, grid)

    def run(self, n, dt0, forward):
        return self.op.run(n, float(np.asarray(dt0).reshape(-1)[0]), forward)


def signature(batch) -> str:
    """The meshes' sizes, the solver and the arithmetic: the best launch
    shape depends on all three."""
    sizes = sorted(m.n for m in batch.meshes)
    h = hashlib.sha1(json.dumps(sizes).encode()).hexdigest()[:16]
    return (f"n{batch.n_scen}_c{batch.cells}_{h}_s{getattr(batch, 'solver', 0)}"
            f"_a{getattr(batch, 'arithmetic', 0)}")


def candidates(n_cus: int):
    out = []
    for chunk in CHUNKS:
        out.append((chunk, 0))
        for k in (2, 4, 8):
            out.append((chunk, 8 * ((k * n_cus + 7) // 8)))
    return out


def tune(batch, dt0, steps: int = 20, repeats: int = 3, use_cache: bool = True,
         verbose: bool = False) -> dict:
    """batch: a swe.Batch or a swe.EulerTimestep; dt0 per scenario."""
    import torch

    from .swe import EulerTimestep

    if isinstance(batch, EulerTimestep):
        batch = _Single(batch)

    dev = torch.cuda.current_device()
    props = torch.cuda.get_device_properties(dev)
    key = f"{props.name}|{signature(batch)}"
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            cache = json.load(open(CACHE))
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            best = cache[key]
            batch.set_tuning(best["chunk"], best["grid"])
            return best

    saved = [batch.download(s) for s in range(batch.n_scen)]
    caller_stream = getattr(batch, "stream", 0)
    stream = torch.cuda.Stream(device=dev)
    batch.set_stream(stream.cuda_stream)
    timings = []
    try:
        for chunk, grid in candidates(props.multi_processor_count):
            # every candidate starts from the saved state
            for s, st in enumerate(saved):
                batch.upload(s, *st)
            batch.set_tuning(chunk, grid)
            batch.run(2, dt0, forward=False)  # capture + warm
            reps = []
            for _ in range(repeats):
                start = torch.cuda.Event(enable_timing=True)
                end = torch.cuda.Event(enable_timing=True)
                torch.cuda.synchronize()
                start.record(stream)
                batch.run(steps, dt0, forward=False)
                end.record(stream)
                torch.cuda.synchronize()
                reps.append(start.elapsed_time(end) / steps)
            ms = float(np.median(reps))
            timings.append({"chunk": chunk, "grid": grid, "ms_per_step": ms})
            if verbose:
                print(f"autotune: chunk {chunk} grid {grid}: {ms:.4f} ms/step", flush=True)
    finally:
        for s, st in enumerate(saved):
            batch.upload(s, *st)
        batch.set_stream(caller_stream)
    best = min(timings, key=lambda r: r["ms_per_step"])
    if best["grid"] != 0:
        # a persistent grid must beat one workgroup per chunk (the hardware's
        # dispatch, which balances uneven chunks over the run) by more than
        # the trials' noise
        one = next(r for r in timings if r["chunk"] == best["chunk"] and r["grid"] == 0)
        if one["ms_per_step"] <= best["ms_per_step"] * (1.0 + PERSISTENT_MARGIN):
            best = one
    result = {"chunk": best["chunk"], "grid": best["grid"], "ms_per_step": best["ms_per_step"],
              "trials": timings}
    batch.set_tuning(result["chunk"], result["grid"])
    if use_cache:
        cache[key] = result
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1)
    return result
