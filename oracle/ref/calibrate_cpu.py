"""Times the reference CPU path beside the oracle's OpenMP restatement (the
`cpu_baseline` port of bench.py) on the same workloads, in this container
(TEST INFRASTRUCTURE ONLY; the reference never travels to the GPU box).

    make -C oracle ref ref-timing
    python oracle/ref/calibrate_cpu.py [--threads 8] [--out FILE]

Workloads: the radial dam break in a closed basin (VARIANT=radial builds) at
depth 14 (32,768 cells, config 1's size) over 100 steps and at depth 21
(4,194,304 cells, the size of the tsunami's frozen mesh) over 10 steps.  The
reference runs with -threads T -sections T, and its "Cell update throughput"
line (SWE.f90:474, cells over the time-step phase) is read back.  The port is
oracle/swe_oracle.c swo_step_omp on the same mesh and state with T threads,
timed around the steps only.  Prints one JSON object."""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.dirname(HERE)
ROOT = os.path.dirname(ORACLE)
sys.path.insert(0, ORACLE)
sys.path.insert(0, ROOT)

BINARIES = {"O2 -ffp-contract=off (parity build)": "samoa_swe_radial",
            "O3": "samoa_swe_radial_O3",
            "Ofast -march=native": "samoa_swe_radial_Ofast"}


def run_reference(binary: str, depth: int, steps: int, threads: int) -> dict:
    with tempfile.TemporaryDirectory() as od:
        cmd = [os.path.join(ORACLE, "_ref", binary), "-threads", str(threads), "-sections",
               str(threads), "-dstart", str(depth), "-dmin", str(depth), "-dmax", str(depth),
               "-nadapt", "-1", "-tmax", "-1", "-nmax", str(steps), "-xmloutput", ".false.",
               "-nout", str(steps + 1), "-output_dir", od]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True, check=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
        wall = time.perf_counter() - t0
    # the last phase printed is the time-step phase (the first, the initial one)
    m = re.findall(r"Cell update throughput:\s+([0-9.]+)\s+M/s", r.stdout)[-1]
    p = re.findall(r"Phase time:\s+([0-9.]+)\s+s", r.stdout)[-1]
    return {"cell_updates_per_s": float(m) * 1e6, "phase_s": float(p),
            "wall_s": wall, "command": " ".join(cmd[1:]).replace(od, "<dir>")}


def run_port(depth: int, steps: int, threads: int) -> dict:
    import pyoracle as po
    from samoa_amd import host

    os.environ["OMP_NUM_THREADS"] = str(threads)
    mesh = host.mesh_uniform(depth)
    st = list(host.radial_dam_break(mesh))
    dt = host.initial_dt(mesh, st, float(np.float32(0.45)))
    drytol = float(np.float32(0.01))
    t_used, fwd = 0.0, True
    for _ in range(steps):
        t0 = time.perf_counter()
        e, u, v, _, cand, _ = po.step(mesh.depth, mesh.plotter, mesh.nbr, st, dt, scaling=1.0,
                                      drytol=drytol, bc=po.BC_WALL, forward=fwd, threads=True)
        t_used += time.perf_counter() - t0
        st = [e, u, v, st[3]]
        dt = float(np.float32(0.45)) * cand
        fwd = not fwd
    return {"cell_updates_per_s": mesh.n * steps / t_used, "steps_s": t_used}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=len(os.sched_getaffinity(0)))
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    out = {"threads": args.threads, "cpu": open("/proc/cpuinfo").read().split("model name")[1]
           .split("\n")[0].strip(" :\t"), "workloads": {}}
    for depth, steps in ((14, 100), (21, 10)):
        w = {"cells": 2 ** (depth + 1), "steps": steps, "reference": {}}
        for flags, binary in BINARIES.items():
            if os.path.exists(os.path.join(ORACLE, "_ref", binary)):
                for t in sorted({1, args.threads}):
                    w["reference"][f"{flags}, {t} threads"] = run_reference(binary, depth, steps, t)
        w["port"] = run_port(depth, steps, args.threads)
        for flags, r in w["reference"].items():
            r["port_over_reference"] = w["port"]["cell_updates_per_s"] / r["cell_updates_per_s"]
        out["workloads"][f"radial_d{depth}"] = w
        print(json.dumps({f"radial_d{depth}": w}), flush=True)
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
