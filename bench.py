"""Benchmark of the MI355X SWE Euler step.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload all|tsunami|frozen|ensemble|adjoint|inversion]
                    [--nadapt k] [--adapt-tol m] [--no-cpu] [--calibrate-copy]

`value` is the adaptive single tsunami: a 1000 km square, b = -4000 m +
Gaussian random field (std 800 m, l = 50 km) + a 150 km shelf rising to
+200 m (a bilinear raster), an elliptical Gaussian source, dry tolerance
1e-3 m, courant 0.45, the augmented Riemann solver, FP64.  The mesh is refined
from depth 16 to depth 26 where the source's surface elevation varies by more
than 1.5 mm across a cell (about 3.2M cells) and remeshed on the device before
every step (the reference's -nadapt 1) from the elevation jumps between wet
neighbours; split cells resample b, merged cells average the state.  A step is
one remeshing (flags, conformity, split/merge, state transfer) plus one Euler
traversal (flux on every edge, cell update with the wet/dry limiter, CFL
reduction, device-side dt); `value` counts Euler cell updates per second of
the whole loop, as the reference's "Cell update throughput" (SWE.f90:474)
counts them per second of its time-step phase, adaptions included.  The
roofline is the Euler step kernel's, timed with HIP events around every step
of that run.

The same line carries, with --workload all (the default): the tsunami on a
frozen uniform depth-21 mesh (4,194,304 cells, K steps from a captured HIP
graph) under "frozen_uniform"; the batched ensemble (64 scenarios from seeds
0-63 on uneven frozen depths 14-20, one launch per step, autotuned launch
shape, 500 steps) under "ensemble"; the differentiable inversion (depth-20
mesh, 200 steps, 32 gauges, sqrt(200) checkpoints) with its adjoint/forward
cost ratio under "adjoint"; and 16 batched source inversions (20 lockstep
L-BFGS iterations) with their recovery error under "inversion".  Inputs are
resident in HBM when the timed region starts.  "frozen_uniform" also carries
"host_buffers": ten of those steps through the host-buffer boundary
(swe_upload, swe_step, swe_download each step), which PCIe transfers bound.
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import gc
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s/GPU plain vs batched ensemble; adjoint/forward cost ratio"
HBM_PEAK = 8.0e12           # MI355X HBM3E, bytes/s
BYTES_PER_CELL = 56         # read eta, hu, hv, b; write eta, hu, hv (FP64)
# how b is sampled: the tsunami raster (513 x 513 nodes) read bilinearly, as
# the reference reads its ASAGI grid, -5000 m outside the nodes; each cell's
# b is that raster averaged over min(16, raster_cap - depth) halvings of the
# cell, as the reference's _ADAPT_INTEGRATE build does for gridded data, at
# set-up and for every cell a remeshing splits
BATHYMETRY = ("tsunami raster 513x513, bilinear (ASAGI-style, -5000 m outside), "
              "cell-averaged as _ADAPT_INTEGRATE")
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", default="all",
                   choices=["all", "tsunami", "frozen", "ensemble", "adjoint", "inversion",
                            "config1"])
    p.add_argument("--ens-steps", type=int, default=500)
    p.add_argument("--ens-scenarios", type=int, default=64)
    p.add_argument("--depth", type=int, default=21)
    p.add_argument("--nadapt", type=int, default=1,
                   help="remesh before every k-th step of the adaptive run")
    p.add_argument("--adapt-tol", type=float, default=1.5e-3,
                   help="surface-elevation variation (m) across a cell that asks for refinement")
    p.add_argument("--adj-depth", type=int, default=20)
    p.add_argument("--adj-steps", type=int, default=200)
    p.add_argument("--fd-dirs", type=int, default=20,
                   help="finite-difference directions checked against the adjoint gradient")
    p.add_argument("--inv-instances", type=int, default=16)
    p.add_argument("--inv-iters", type=int, default=20)
    p.add_argument("--inv-hbm-gib", type=float, default=24.0,
                   help="HBM budget (GiB) the inversions' checkpoint pools are admitted against")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--tuning", default=None,
                   help="CHUNK,GRID launch shape instead of autotuning (e.g. for profiling)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run a 1 GiB device copy (counter calibration under rocprofv3)")
    return p.parse_args()


def fixed_tuning(args, op):
    """--tuning CHUNK,GRID: set the launch shape instead of autotuning."""
    if not args.tuning:
        return None
    chunk, grid = (int(x) for x in args.tuning.split(","))
    op.set_tuning(chunk, grid)
    return {"chunk": chunk, "grid": grid, "trials": []}


def cpu_baseline(mesh, state, dt, seconds_target=15.0):
    """The oracle's OpenMP restatement of the same step on this host's cores,
    on the same mesh and state, for a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    po.build()
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    cores = int(os.environ["OMP_NUM_THREADS"])
    st = [a.copy() for a in state]
    steps, t_used, fwd = 0, 0.0, False
    while t_used < seconds_target and steps < 50:
        t0 = time.perf_counter()
        e, u, v, _, cand, _ = po.step(mesh.depth, mesh.plotter, mesh.nbr, st, dt, scaling=1.0e6,
                                      drytol=1e-3, forward=fwd, threads=True)
        t_used += time.perf_counter() - t0
        st = [e, u, v, st[3]]
        dt = 0.45 * cand
        fwd = not fwd
        steps += 1
    return {"value": mesh.n * steps / t_used, "unit": "cell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"{steps} steps of the same {mesh.n}-cell mesh and state "
                      f"(oracle/swe_oracle.c swo_step_omp, {t_used:.1f} s)"}


def tsunami_adaptive(args, torch, arith=0):
    """The adaptive single tsunami: the tsunami scenario (1000 km square,
    b = -4000 m + Gaussian random field + coastal shelf, as a bilinear raster,
    elliptical source of seed 0) on a mesh refined from depth 16 to depth 26
    where the source's surface elevation varies by more than --adapt-tol
    between a cell's corners (|grad eta| h), then remeshed before every
    --nadapt-th step (the reference's default 1) from the elevation jumps to
    wet neighbours, which span about half a corner range (refine above half
    of --adapt-tol, coarsen below a sixteenth); each step evaluates them on
    the state it reads and leaves them for the next remeshing, as the
    reference's Euler traversal leaves its flags.
    Split cells resample b from the raster, merged cells average it
    (SWE_adapt.f90:95-169).  Throughput counts the Euler cell updates over the
    wall time of the whole phase, remeshing included, as the reference's
    "Cell update throughput" does (SWE.f90:474)."""
    from samoa_amd import host, swe

    S, courant, tol = 1.0e6, 0.45, args.adapt_tol
    dmin, dmax = 16, 26
    stream = torch.cuda.Stream(device=torch.device("cuda:0"))

    def phase(per_step_events):
        """Set-up, warm-up and the timed steps from the same initial state.
        The throughput pass has nothing in the loop but the steps and the
        remeshing; the second pass records events around every step launch
        to time the step kernels (each event is a packet in the queue that
        adds a few microseconds between the kernels)."""
        mesh0 = host.mesh_uniform(dmin)
        op = swe.EulerTimestep(mesh0, scaling=S, dry_tolerance=1e-3, courant=courant,
                               min_depth=dmin, max_depth=dmax, solver=swe.SOLVER_AUG_RIEMANN,
                               boundary=swe.BC_OUTFLOW)
        op.set_arithmetic(arith)
        fixed_tuning(args, op)  # --tuning: a launch shape other than the default
        z = np.zeros(mesh0.n)
        op.upload(z, z, z, z)
        # the Fortran host's set-up: the initial phase of SWE.f90:230-262 with
        # a tolerance on the source's variation across a cell, the scenario
        # state, the bathymetry raster for later splits (swe_driver.f90)
        t0 = time.perf_counter()
        n0, dt0 = host.tsunami_adaptive_setup(op, 0, tol, dmax, 1e-3, courant)
        t_init = time.perf_counter() - t0
        depth = op.download_mesh()[0]
        op.set_stream(stream.cuda_stream)
        state0 = op.download()
        # every step leaves flags from the elevation jumps of the state it
        # reads (a jump to a neighbour's centroid spans about half a corner
        # range)
        op.keep_eta_jump_flags(tol / 2.0, tol / 16.0)
        # the per-step events exist before the timed loop: creating them
        # there would put the harness's own allocations on the host's
        # critical path
        events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(args.steps if per_step_events else 0)]
        step_ev = []
        t, dt, fwd = 0.0, dt0, False

        def one_step(k, timed):
            nonlocal t, dt, fwd
            if (k + 1) % args.nadapt == 0:  # time step k + 1 of SWE.f90:385-390
                if k > 0:  # before the first step the flags are all 0: nothing changes
                    op.adapt()  # the flags the last step left
                fwd = not fwd  # the adaption traversal flips the direction
            n = op.n
            if timed and per_step_events:
                ev = events[len(step_ev)]
                ev[0].record(stream)
            # after the first step the dt chain stays on the device and
            # nothing waits for a step
            if k == 0:
                t, dt = op.run(1, dt, fwd, time=t)
            else:
                op.run(1, None, fwd, wait=False)
            if timed and per_step_events:
                ev[1].record(stream)
                step_ev.append((ev, n))
            fwd = not fwd
            return n

        for k in range(args.warmup):
            one_step(k, False)
        torch.cuda.synchronize()
        cells = []
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        # Python's cycle collector would pause the host loop at random steps;
        # the Fortran host that drives production runs has no such pauses
        gc_was = gc.isenabled()
        gc.disable()
        w0 = time.perf_counter()
        start.record(stream)
        for k in range(args.warmup, args.warmup + args.steps):
            cells.append(one_step(k, True))
        end.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        if gc_was:
            gc.enable()
        t, dt = op.run(0, None, fwd)  # read the device chain
        ms = start.elapsed_time(end)
        eta, hu, hv, b = op.download()
        if not (np.isfinite(eta).all() and np.isfinite(hu).all()):
            raise RuntimeError("non-finite state after the adaptive steps")
        d_end = op.download_mesh()[0]
        op.close()
        return dict(n0=n0, dt0=dt0, t_init=t_init, depth=depth, state0=state0, cells=cells,
                    ms=ms, wall=wall, t=t, d_end=d_end, step_ev=step_ev,
                    digest=hashlib.sha256(eta.tobytes() + hu.tobytes()).hexdigest()[:16])

    run = phase(False)
    prof = phase(True)
    if prof["cells"] != run["cells"] or prof["digest"] != run["digest"]:
        raise RuntimeError("the two passes of the adaptive run differ")
    step_ms = np.array([a.elapsed_time(b) for (a, b), _ in prof["step_ev"]])
    step_cells = np.array([n for _, n in prof["step_ev"]], dtype=np.float64)
    cells, d_end = run["cells"], run["d_end"]
    updates = float(np.sum(cells))
    elapsed = max(run["ms"] * 1e-3, run["wall"])
    elapsed_prof = max(prof["ms"] * 1e-3, prof["wall"])
    return {
        "value": updates / elapsed,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "cell_updates": updates,
        "step_kernel_ms_total": float(step_ms.sum()),
        "step_kernel_bytes_per_s": float(BYTES_PER_CELL * step_cells.sum() / (step_ms.sum() * 1e-3)),
        "remesh_and_host_ms_per_step": float((elapsed * 1e3 - step_ms.sum()) / args.steps),
        "elapsed_gpu_s": run["ms"] * 1e-3, "elapsed_wall_s": run["wall"],
        "ms_per_step_with_step_events": 1e3 * elapsed_prof / args.steps,
        "timing": "throughput from a pass with nothing but the steps and the remeshing in the "
                  "loop; step kernel times from a second pass of the same steps with events "
                  "around every step launch (same cells and final state)",
        "cells_initial": run["n0"], "depths_initial_16_to_26":
            np.bincount(run["depth"], minlength=dmax + 1)[dmin:].tolist(),
        "cells_final": int(len(d_end)),
        "depths_final_16_to_26": np.bincount(d_end, minlength=dmax + 1)[dmin:].tolist(),
        "cells_min": int(min(cells)), "cells_max": int(max(cells)),
        "initial_refinement_s": run["t_init"],
        "nadapt": args.nadapt, "adapt_tol_m": tol, "simulated_time_s": run["t"],
        "depth0": run["depth"], "state0": run["state0"], "dt0": run["dt0"],
    }


def tsunami_frozen(args, torch):
    """The same tsunami on a frozen uniform mesh of depth --depth (21: 4,194,304
    cells): K steps from a captured HIP graph, the device dt chain, no
    remeshing.  Returns (result, mesh, state, dt0)."""
    from samoa_amd import autotune, host, swe

    d = args.depth
    mesh = host.mesh_uniform(d)
    state = host.tsunami(mesh, seed=0, drytol=1e-3)
    S, courant = 1.0e6, 0.45
    dt0 = host.initial_dt(mesh, state, courant)  # unit-square units (reference quirk)

    dev = torch.device("cuda:0")
    torch.cuda.init()
    stream = torch.cuda.Stream(device=dev)
    op = swe.EulerTimestep(mesh, scaling=S, dry_tolerance=1e-3, courant=courant,
                           solver=swe.SOLVER_AUG_RIEMANN, boundary=swe.BC_OUTFLOW)
    op.upload(*state)
    tuned = fixed_tuning(args, op) or autotune.tune(op, dt0, use_cache=False)
    op.set_stream(stream.cuda_stream)

    # warmup: first steps of the run (also captures the graph)
    t_warm, dt = op.run(args.warmup, dt0, forward=False)
    fwd = args.warmup % 2 == 1

    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        start.record(stream)
        t_end, dt_next = op.run(args.steps, dt, forward=fwd, time=t_warm)
        end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end)
    ms_per_step = ms / args.steps
    cells = mesh.n
    value = cells * args.steps / (ms * 1e-3)

    eta, hu, hv, b = op.download()
    if not (np.isfinite(eta).all() and np.isfinite(hu).all()):
        raise RuntimeError("non-finite state after the timed steps")
    exact_end = (eta.copy(), hu.copy(), hv.copy())

    achieved = BYTES_PER_CELL * cells / (ms_per_step * 1e-3)
    traffic, valu = None, None
    if os.path.exists(PROFILE_TRAFFIC):
        prof = json.load(open(PROFILE_TRAFFIC))
        key = f"euler_step_tsunami_d{d}"
        if key in prof:
            traffic = prof[key]["hbm_bytes_per_launch"]
            valu = prof[key].get("valu_issue_frac")

    # The same steps through the host-buffer boundary (swe_upload, swe_step,
    # swe_download): every step takes the state from host memory and hands
    # it back, so PCIe transfers bound it.  Reported beside the resident
    # rate, never as the value.
    r_host = 10
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(r_host):
        op.upload(eta, hu, hv, b)
        op.traverse(dt_next, forward=(fwd + args.steps + k) % 2 == 1)
        eta, hu, hv, b = op.download()
    s_host = time.perf_counter() - t0
    host_buffers = {
        "value": cells * r_host / s_host, "unit": "cell-updates/s", "steps": r_host,
        "ms_per_step": 1e3 * s_host / r_host,
        "per_step": "swe_upload of eta, hu, hv, b from pageable host memory, swe_step, "
                    "swe_download of the four arrays",
    }

    # the toleranced arithmetic on the same mesh, from the same state, with
    # its own launch shape: throughput, and its distance from the exact run
    # after the same steps
    op.set_arithmetic(swe.ARITH_TOLERANCED)
    op.upload(*state)
    tuned_tol = fixed_tuning(args, op) or autotune.tune(op, dt0, use_cache=False)
    op.set_stream(stream.cuda_stream)
    op.upload(*state)
    tw_tol, dt_tol = op.run(args.warmup, dt0, forward=False)
    start2, end2 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start2.record(stream)
    t_tol, _ = op.run(args.steps, dt_tol, forward=fwd, time=tw_tol)
    end2.record(stream)
    torch.cuda.synchronize()
    ms_tol = start2.elapsed_time(end2)
    e2, u2, v2, _ = op.download()
    op.set_arithmetic(swe.ARITH_EXACT)
    ach_tol = BYTES_PER_CELL * cells / (ms_tol / args.steps * 1e-3)
    toleranced = {
        "value": cells * args.steps / (ms_tol * 1e-3), "unit": "cell-updates/s",
        "ms_per_step": ms_tol / args.steps,
        "roofline": {"bound": "hbm", "achieved": ach_tol / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": ach_tol / HBM_PEAK},
        "tuning": {"chunk": tuned_tol["chunk"], "grid": tuned_tol["grid"]},
        "max_rel_diff_vs_exact": {
            k: float(np.abs(x - y).max() / np.abs(y).max())
            for k, x, y in zip(("eta", "hu", "hv"), (e2, u2, v2), exact_end)},
        "simulated_time_rel_diff": abs(t_tol - t_end) / t_end,
    }

    copy_rate = None
    if args.calibrate_copy:
        # three 1 GiB copies for the counters, then twenty timed ones: the
        # rate a plain device copy reaches (bytes read plus bytes written)
        a = torch.empty(1 << 27, dtype=torch.float64, device=dev)
        bb = torch.empty_like(a)
        for _ in range(3):
            bb.copy_(a)
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record()
        for _ in range(20):
            bb.copy_(a)
        c1.record()
        torch.cuda.synchronize()
        copy_rate = 2 * a.numel() * 8 * 20 / (c0.elapsed_time(c1) * 1e-3)
        del a, bb

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"tsunami_frozen_uniform_d{d}",
            "cells": cells, "depth": d, "scaling_m": S, "solver": "aug_riemann",
            "boundary": "outflow", "dry_tolerance": 1e-3, "courant": courant,
            "bathymetry": BATHYMETRY, "simulated_time_s": t_end,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic,
            "kernel": "swe::euler_step_kernel<0, 256, false> + swe::finalize_kernel (one step)",
            "bytes_per_cell_update": BYTES_PER_CELL,
            "valu_issue_frac_profiled": valu,
        },
        "tuning": {"chunk": tuned["chunk"], "grid": tuned["grid"], "trials": tuned["trials"]},
        "host_buffers": host_buffers,
        "toleranced": toleranced,
    }
    if copy_rate is not None:
        out["copy_1gib_bytes_per_s"] = copy_rate
    out["roofline"]["achieved"] = achieved / 1e9
    out["roofline"]["peak"] = HBM_PEAK / 1e9
    op.close()
    return out, mesh, state, dt0


def ensemble(args, torch, arith=0):
    """Batched ensemble: seeds 0..N-1, uneven uniform depths from 14-20."""
    from samoa_amd import autotune, host, swe

    n_scen = args.ens_scenarios
    draws = [host.tsunami_draw(s, 14, 20) for s in range(n_scen)]
    meshes = {d: host.mesh_uniform(d) for d in sorted({p["depth"] for p in draws})}
    ms = [meshes[p["depth"]] for p in draws]
    params = [dict(scaling=1.0e6, dry_tolerance=1e-3, courant=0.45) for _ in draws]
    b = swe.Batch(ms, params)
    b.set_arithmetic(arith)
    dt0 = []
    for s, m in enumerate(ms):
        st = host.tsunami(m, seed=s, drytol=1e-3)
        b.upload(s, *st)
        dt0.append(host.initial_dt(m, st, 0.45))
    dt0 = np.array(dt0)
    tuned = fixed_tuning(args, b) or autotune.tune(b, dt0, use_cache=False)
    stream = torch.cuda.Stream()
    b.set_stream(stream.cuda_stream)
    t, dt = b.run(args.warmup, dt0, forward=False)
    fwd = args.warmup % 2 == 1
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    t, dt = b.run(args.ens_steps, dt, forward=fwd, time=t)
    end.record(stream)
    torch.cuda.synchronize()
    ms_total = start.elapsed_time(end)
    cells = b.cells
    ms_step = ms_total / args.ens_steps
    achieved = BYTES_PER_CELL * cells / (ms_step * 1e-3)
    depths = [p["depth"] for p in draws]
    b.close()
    return {
        "value": cells * args.ens_steps / (ms_total * 1e-3), "unit": "cell-updates/s",
        "ms_per_step": ms_step, "steps": args.ens_steps, "scenarios": n_scen, "cells": cells,
        "depths": {str(d): depths.count(d) for d in sorted(set(depths))},
        "tuning": {"chunk": tuned["chunk"], "grid": tuned["grid"], "trials": tuned["trials"]},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK},
        "workload": "tsunami_ensemble_64_uniform_d14-20",
    }


def config1(args, torch):
    """The radial dam break in a closed basin at its full size: the unit
    square, uniform depth 14 (32,768 cells), flat bed, eta 10 m within 0.1 of
    the centre and 5 m outside, reflecting walls, the augmented Riemann
    solver, courant 0.45 and dry tolerance 0.01 as the reference parses them
    (float32), 100 steps from a captured HIP graph with the device dt chain,
    in both arithmetic modes.  Its cpu_baseline is the reference itself
    (oracle/_ref/samoa_swe_radial_O3: the reference's sources with this
    scenario's one-line patches, built with -O3 in the development container)
    on this host's cores, read from its "Cell update throughput" line
    (SWE.f90:474)."""
    from samoa_amd import host, swe

    mesh = host.mesh_uniform(14)
    st = host.radial_dam_break(mesh)
    courant, drytol = float(np.float32(0.45)), float(np.float32(0.01))
    dt0 = host.initial_dt(mesh, st, courant)
    out = {"workload": "radial_dam_break_walls_uniform_d14_100steps", "cells": mesh.n,
           "steps": 100}
    stream = torch.cuda.Stream()
    for name, mode in (("exact", swe.ARITH_EXACT), ("toleranced", swe.ARITH_TOLERANCED)):
        op = swe.EulerTimestep(mesh, scaling=1.0, dry_tolerance=drytol, courant=courant,
                               min_depth=14, max_depth=14, boundary=swe.BC_WALL)
        op.set_arithmetic(mode)
        op.set_stream(stream.cuda_stream)
        op.upload(*st)
        op.run(100, dt0, forward=True, time=0.0)  # warm: graph capture
        op.upload(*st)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        t_end, _ = op.run(100, dt0, forward=True, time=0.0)
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b)
        res = op.download()
        op.close()
        out[name] = {"value": mesh.n * 100 / (ms * 1e-3), "unit": "cell-updates/s",
                     "ms_per_step": ms / 100, "simulated_time_s": t_end}
        if name == "exact":
            ref_state = res
        else:
            out[name]["max_rel_diff_vs_exact"] = {
                k: float(np.abs(x - y).max() / np.abs(y).max())
                for k, x, y in zip(("eta", "hu", "hv"), res[:3], ref_state[:3])}
    if not args.no_cpu:
        out["cpu_baseline"] = reference_cpu(14, 100)
    return out


def reference_cpu(depth, steps):
    """The reference binary (radial dam break, -O3) on this host's cores."""
    import re
    import subprocess
    import tempfile

    binary = os.path.join(ROOT, "oracle", "_ref", "samoa_swe_radial_O3")
    if not os.path.exists(binary):
        return {"skipped": "oracle/_ref/samoa_swe_radial_O3 is not built (make -C oracle ref-timing)"}
    cores = len(os.sched_getaffinity(0))
    best = None
    for threads in sorted({1, cores}):
        with tempfile.TemporaryDirectory() as od:
            cmd = [binary, "-threads", str(threads), "-sections", "1", "-dstart", str(depth),
                   "-dmin", str(depth), "-dmax", str(depth), "-nadapt", "-1", "-tmax", "-1",
                   "-nmax", str(steps), "-xmloutput", ".false.", "-nout", str(steps + 1),
                   "-output_dir", od]
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=300,
                               env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
        if r.returncode != 0:
            return {"skipped": f"the reference exited with {r.returncode}"}
        rate = float(re.findall(r"Cell update throughput:\s+([0-9.]+)\s+M/s", r.stdout)[-1]) * 1e6
        if best is None or rate > best["value"]:
            best = {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "reference",
                    "sample": f"{steps} steps of the same {2 ** (depth + 1)}-cell mesh and state; "
                              "the faster of 1 and all host threads; the reference's own "
                              "'Cell update throughput' (SWE.f90:474)"}
    return best


def adjoint(args, torch):
    """Differentiable single inversion: the tsunami scenario frozen at a
    uniform depth-20 mesh, 200 steps with the forward dt sequence, squared
    eta misfit at 32 seeded gauges (cells within 150 km of the source) every
    step against a target run from a
    perturbed source; gradients with respect to eta0 and b through the
    discrete adjoint with ceil(sqrt(200)) checkpoints.

    ratio = (backward sweep: recompute + adjoint) / (forward sweep), both over
    the same steps, with ceil(sqrt(N)) checkpoints; sweeps["all"] gives the
    same with every state kept in HBM (no recompute); kernel_ratio = one
    adjoint launch / one forward launch."""
    from samoa_amd import host
    from samoa_amd.adjoint import GaugeInversion
    from samoa_amd.inversion import gauges_near

    d, N = args.adj_depth, args.adj_steps
    mesh = host.mesh_uniform(d)
    eta, hu, hv, b = host.tsunami(mesh, seed=0, drytol=1e-3)
    eta_t = host.tsunami(mesh, seed=0, drytol=1e-3,
                         amplitude=1.25 * host.tsunami_draw(0)["amplitude"])[0]
    dt0 = host.initial_dt(mesh, (eta, hu, hv, b), 0.45)
    src = host.tsunami_draw(0)
    gauges = gauges_near(mesh, src["x0"], src["y0"], 1.5e5, 32, seed=2024)
    dev = torch.device("cuda:0")
    e0 = torch.tensor(eta, device=dev)
    bb = torch.tensor(b, device=dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    sweeps = {}
    for mode in ("sqrt", "all"):
        prob = GaugeInversion(mesh, gauges, N, dict(scaling=1.0e6, dry_tolerance=1e-3, courant=0.45),
                              checkpoint_every=(int(np.ceil(np.sqrt(N))) if mode == "sqrt" else "all"))
        prob.record_dt(e0, bb, dt0)
        target = prob.gauge_series(torch.tensor(eta_t, device=dev), bb).detach()
        for rep in range(2):  # the first repetition warms up
            e0g = e0.clone().requires_grad_()
            bg = bb.clone().requires_grad_()
            torch.cuda.synchronize()
            ev[0].record()
            loss = ((prob.gauge_series(e0g, bg) - target) ** 2).sum()
            ev[1].record()
            loss.backward()
            ev[2].record()
            torch.cuda.synchronize()
        t_fwd = ev[0].elapsed_time(ev[1])
        t_bwd = ev[1].elapsed_time(ev[2])
        finite = bool(torch.isfinite(e0g.grad).all()) and bool(torch.isfinite(bg.grad).all())
        sweeps[mode] = {"K": prob.K, "store_all": prob.store_all, "ms_forward_sweep": t_fwd,
                        "ms_backward_sweep": t_bwd, "ratio": t_bwd / t_fwd, "loss": float(loss.detach()),
                        "gradients_finite": finite}
        if mode == "sqrt":
            g_sqrt = (e0g.grad.clone(), bg.grad.clone())
            fd_check = finite_differences(prob, e0, bb, eta, b, target, g_sqrt, mesh, src,
                                          n_dirs=args.fd_dirs)
        else:
            sweeps["gradients_equal_across_modes"] = bool(
                torch.equal(g_sqrt[0], e0g.grad) and torch.equal(g_sqrt[1], bg.grad))
        if mode == "all":
            break
        prob.close()

    # single-launch costs: one forward traversal with masks vs one adjoint traversal
    z = [torch.zeros_like(e0) for _ in range(9)]
    q, out, lam = z[0:3], z[3:6], z[6:9]
    q[0].copy_(e0)
    masks = torch.empty(mesh.n, dtype=torch.int8, device=dev)
    gb = torch.zeros_like(e0)
    P = lambda t: t.data_ptr()  # noqa: E731
    stream = torch.cuda.current_stream().cuda_stream
    reps = 50

    def fwd_launch():
        prob.op.forward_device(P(q[0]), P(q[1]), P(q[2]), P(bb), P(out[0]), P(out[1]), P(out[2]),
                               P(prob.dt), False, masks=P(masks), stream=stream)

    def adj_launch():
        prob.op.adjoint_device(P(q[0]), P(q[1]), P(q[2]), P(bb), P(masks), P(lam[0]), P(lam[1]),
                               P(lam[2]), P(out[0]), P(out[1]), P(out[2]), P(gb), P(prob.dt),
                               False, stream=stream)

    times = {}
    for name, fn in (("fwd", fwd_launch), ("adj", adj_launch)):
        fn()
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        times[name] = ev[0].elapsed_time(ev[1]) / reps
    prob.close()
    adj_bytes = 96 * mesh.n / (times["adj"] * 1e-3)
    return {
        "workload": f"tsunami_inversion_uniform_d{d}_{N}steps_32gauges",
        "cells": mesh.n, "steps": N,
        "ratio": sweeps["sqrt"]["ratio"],
        "sweeps": sweeps,
        "finite_differences": fd_check,
        "ms_forward_launch": times["fwd"], "ms_adjoint_launch": times["adj"],
        "kernel_ratio": times["adj"] / times["fwd"],
        "roofline_adjoint": {"bound": "hbm", "achieved": adj_bytes / 1e9, "peak": HBM_PEAK / 1e9,
                             "unit": "GB/s", "frac": adj_bytes / HBM_PEAK,
                             "bytes_per_cell_step": 96},
    }


def finite_differences(prob, e0, bb, eta, b, target, grads, mesh, src, n_dirs=20):
    """Central finite differences of the gauge misfit along seeded random
    directions (alternately in eta0 and b) against the adjoint gradient.
    Directions: standard normal values on the cells more than 50 m deep,
    windowed by a 300 km Gaussian around the source and scaled to a largest
    entry of 1; steps of 1 mm (eta0) and 0.5 m (b).  rel_terms divides the
    discrepancy by sum |grad_i dir_i|, the size of the terms the directional
    derivative sums (a random direction's derivative can cancel far below
    them); rel_value divides by |fd|."""
    import torch

    cen = mesh.centroids() * 1.0e6
    win = np.exp(-((cen[:, 0] - src["x0"]) ** 2 + (cen[:, 1] - src["y0"]) ** 2) / 3.0e5 ** 2)
    mask = torch.tensor(((eta - b) > 50.0) * win, device=e0.device)
    rng = np.random.default_rng(20261203)

    def loss(e, bt):
        with torch.no_grad():
            return float(((prob.gauge_series(e, bt) - target) ** 2).sum())

    rows = []
    for k in range(n_dirs):
        field = k % 2
        dirn = torch.tensor(rng.standard_normal(mesh.n), device=e0.device) * mask
        dirn = dirn / dirn.abs().max()
        eps = 1e-3 if field == 0 else 0.5
        base = [e0.clone(), bb.clone()]
        p_, m_ = [x.clone() for x in base], [x.clone() for x in base]
        p_[field] += eps * dirn
        m_[field] -= eps * dirn
        fd = (loss(*p_) - loss(*m_)) / (2 * eps)
        ad = float((grads[field] * dirn).sum())
        terms = float((grads[field] * dirn).abs().sum())
        rows.append({"field": "eta0" if field == 0 else "b", "adjoint": ad, "fd": fd,
                     "rel_terms": abs(ad - fd) / terms, "rel_value": abs(ad - fd) / abs(fd)})
    return {"directions": n_dirs, "steps": {"eta0_m": 1e-3, "b_m": 0.5},
            "max_rel_terms": max(r["rel_terms"] for r in rows),
            "max_rel_value": max(r["rel_value"] for r in rows),
            "median_rel_value": float(np.median([r["rel_value"] for r in rows])),
            "per_direction": rows}


def inversion(args, torch):
    """Batched adjoint inversions: 16 tsunami sources (seeds 0-15, uniform
    depths drawn from 16-20) recovered from 32 gauges each over 200-step
    windows, all instances' forward and adjoint sweeps batched into one
    launch per step, 20 lockstep L-BFGS iterations; reports the final source
    recovery error."""
    from samoa_amd import host
    from samoa_amd.inversion import SCALE, SourceInversion, gauges_near, lbfgs

    n_inst, N = args.inv_instances, args.adj_steps
    dev = torch.device("cuda:0")
    draws = [host.tsunami_draw(s, 16, 20) for s in range(n_inst)]
    meshes = {d: host.mesh_uniform(d) for d in sorted({p["depth"] for p in draws})}
    ms, truth, b, gauges, dt0 = [], [], [], [], []
    for s, p in enumerate(draws):
        m = meshes[p["depth"]]
        eta, hu, hv, bb = host.tsunami(m, seed=s)
        ms.append(m)
        truth.append([p["amplitude"], p["ratio"], p["x0"], p["y0"]])
        b.append(bb)
        gauges.append(gauges_near(m, p["x0"], p["y0"], 1.5e5, 32, seed=s))
        dt0.append(host.initial_dt(m, (eta, hu, hv, bb), 0.45))
    inv = SourceInversion(ms, b, gauges, N, dt0, hbm_budget=args.inv_hbm_gib * 2.0 ** 30)
    theta_true = torch.tensor(truth, dtype=torch.float64, device=dev)
    inv.set_target(theta_true)
    scale = torch.tensor(SCALE, dtype=torch.float64, device=dev)
    guess = theta_true * torch.tensor([0.6, 1.0, 1.0, 1.0], dtype=torch.float64, device=dev)
    guess = guess + torch.tensor([0.0, 0.4, 3.0e4, -2.0e4], dtype=torch.float64, device=dev)
    evals = [0]

    def fg(z):
        evals[0] += 1
        return inv.loss_and_grad(z)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    z, f, hist = lbfgs(fg, guess / scale, iters=args.inv_iters)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    theta = z * scale
    rel = ((theta - theta_true).abs() / theta_true.abs()).cpu().numpy()
    with torch.no_grad():
        e_rec, e_true = inv.eta0(theta), inv.eta0(theta_true)
        src_err = float((e_rec - e_true).norm() / e_true.norm())
    cells = int(inv.problem.offsets[-1])
    plan = inv.problem.plan()
    inv.close()
    return {
        "workload": f"tsunami_source_inversions_{n_inst}_uniform_d16-20_{N}steps_32gauges",
        "instances": n_inst, "cells": cells, "steps": N, "lbfgs_iterations": args.inv_iters,
        "depths": {str(d): [p["depth"] for p in draws].count(d) for d in sorted(meshes)},
        "function_evaluations": evals[0], "seconds": wall,
        "checkpoint_pools": {
            "hbm_budget_gib": args.inv_hbm_gib,
            "rounds": [[{"store_all": g["store_all"], "K": g["K"], "instances": g["instances"],
                         "tape_gib": g["tape_bytes"] / 2.0 ** 30} for g in r]
                       for r in plan["rounds"]]},
        "loss_initial_max": float(hist[0].max()), "loss_final_max": float(f.max()),
        "param_rel_error_max": float(rel.max()), "param_rel_error_median": float(np.median(rel)),
        "source_rel_l2_error": src_err,
    }


import numpy as np  # noqa: E402


_T0 = time.perf_counter()


def progress(what):
    """One line on stderr per finished block, so that a long run shows progress."""
    print(f"bench.py: {what} done at {time.perf_counter() - _T0:.0f} s", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not run "
              "(sharding across GPUs is designed but not built)", file=sys.stderr)
        sys.exit(1)

    import numpy as np
    import torch

    from samoa_amd import autotune, host, swe

    if args.workload == "ensemble":
        print(json.dumps({"metric": METRIC, "ensemble": ensemble(args, torch)}), flush=True)
        return
    torch.cuda.init()
    if args.workload == "adjoint":
        print(json.dumps({"metric": METRIC, "adjoint": adjoint(args, torch)}), flush=True)
        return
    if args.workload == "inversion":
        print(json.dumps({"metric": METRIC, "inversion": inversion(args, torch)}), flush=True)
        return
    if args.workload == "config1":
        print(json.dumps({"metric": METRIC, "config1": config1(args, torch)}), flush=True)
        return

    if args.workload == "frozen":
        out, mesh, state, dt0 = tsunami_frozen(args, torch)
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline(mesh, state, dt0 * 1.0e6)
        print(json.dumps(out), flush=True)
        return

    ad = tsunami_adaptive(args, torch)
    progress("adaptive tsunami")
    achieved = ad["step_kernel_bytes_per_s"]
    traffic, valu = None, None
    if os.path.exists(PROFILE_TRAFFIC):
        prof = json.load(open(PROFILE_TRAFFIC)).get("euler_step_tsunami_adaptive_d16_26")
        if prof:  # bytes per cell of the profiled launches, at this run's mean mesh size
            traffic = prof["hbm_bytes_per_cell"] * ad["cell_updates"] / args.steps
            valu = prof.get("valu_issue_frac")
    out = {
        "metric": METRIC,
        "value": ad["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ad["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "tsunami_adaptive_d16-26_remesh_every_step",
            "cells_initial": ad["cells_initial"], "cells_min": ad["cells_min"],
            "cells_max": ad["cells_max"], "depths": [16, 26], "scaling_m": 1.0e6,
            "solver": "aug_riemann", "boundary": "outflow", "dry_tolerance": 1e-3,
            "courant": 0.45, "nadapt": args.nadapt, "adapt_tol_m": args.adapt_tol,
            "bathymetry": BATHYMETRY,
            "simulated_time_s": ad["simulated_time_s"],
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic,
            "kernel": "swe::euler_step_kernel<0, 256, true> + swe::finalize_kernel (one step)",
            "valu_issue_frac_profiled": valu,
            "bytes_per_cell_update": BYTES_PER_CELL,
            "timing": "HIP events around every step launch of a second pass of the same steps",
            "traffic_source": "HBM bytes per cell of the rocprofv3 FETCH_SIZE and WRITE_SIZE "
                              "passes in profiles/traffic.json (not counted in this run) "
                              "times this run's mean cells per step",
        },
        "adaptive": ad,
    }
    if args.workload == "all":
        frozen, mesh, state, dt0 = tsunami_frozen(args, torch)
        out["frozen_uniform"] = frozen
        progress("frozen uniform tsunami")
        out["ensemble"] = ensemble(args, torch)
        progress("ensemble")
        ad_tol = tsunami_adaptive(args, torch, swe.ARITH_TOLERANCED)
        progress("adaptive tsunami, toleranced")
        for k in ("depth0", "state0", "dt0"):
            ad_tol.pop(k)
        ens_tol = ensemble(args, torch, swe.ARITH_TOLERANCED)
        progress("ensemble, toleranced")
        out["toleranced"] = {
            "note": "SWE_ARITH_TOLERANCED: within 1e-12 of the exact path per step "
                    "(tests/test_gpu_toleranced.py); the exact path is `value`",
            "adaptive": {k: ad_tol[k] for k in ("value", "ms_per_step", "step_kernel_bytes_per_s",
                                                "remesh_and_host_ms_per_step", "cells_min",
                                                "cells_max")},
            "frozen_uniform": frozen["toleranced"],
            "ensemble": {k: ens_tol[k] for k in ("value", "ms_per_step", "roofline", "tuning")},
        }
        out["config1"] = config1(args, torch)
        progress("config 1")
        out["adjoint"] = adjoint(args, torch)
        progress("adjoint")
        out["inversion"] = inversion(args, torch)
        progress("inversions")
    if not args.no_cpu:
        from samoa_amd import host

        mesh0 = host.mesh_from_depths(ad["depth0"])
        out["cpu_baseline"] = cpu_baseline(mesh0, ad["state0"], ad["dt0"] * 1.0e6)
        out["cpu_baseline"]["sample"] += "; the adaptive run's initial mesh, Euler steps only"
    ad.pop("depth0"), ad.pop("state0"), ad.pop("dt0")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
