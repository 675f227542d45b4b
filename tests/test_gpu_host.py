"""The Fortran host over the C ABI on the GPU: the swe_mi355x command-line
driver against the reference's golden runs, and the Fortran time loop
against the oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import golden_io
import pyoracle as po
from samoa_amd import host, swe

pytestmark = pytest.mark.gpu


def read_dump(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:8], "<i8")[0])
    t = float(np.frombuffer(raw[8:16], "<f8")[0])
    arr = np.frombuffer(raw[16:], "<f8").reshape(4, n)
    return t, arr


@pytest.mark.parametrize("hoststeps", [False, True])
@pytest.mark.parametrize("name", ["plain_d6_n1", "plain_d8_n3", "plain_d10_n5"])
def test_cli_matches_reference_run(name, hoststeps):
    _, z, meta = {r[0]: r for r in golden_io.runs()}[name]
    with tempfile.TemporaryDirectory() as td:
        cmd = [host.CLI, "-scenario", "dambreak", "-dmax", str(meta["depth"]), "-nmax",
               str(meta["steps"]), "-nout", str(meta["nout"]), "-tmax", "-1", "-courant", "0.45",
               "-drytolerance", "0.01", "-dump", os.path.join(td, "s")]
        if hoststeps:
            cmd.append("-hoststeps")
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        for i in range(len(z["times"])):
            t, arr = read_dump(os.path.join(td, f"s_{i}.bin"))
            assert t == z["times"][i], f"output {i}"
            assert np.array_equal(arr, z["states"][i]), f"output {i}"


@pytest.mark.parametrize("name,solver,dam", [("llfbath_plain_d7_n2", "llfbath", None),
                                              ("llf_dam_d8_n3", "llf", (10, 5)),
                                              ("lf_dam_d7_n1", "lf", (10, 5))])
def test_cli_lax_friedrichs_runs_match_reference(name, solver, dam):
    """-solver with the reference's flux_solver names: the reference's runs
    of those builds, output for output."""
    _, z, meta = {r[0]: r for r in golden_io.runs()}[name]
    with tempfile.TemporaryDirectory() as td:
        cmd = [host.CLI, "-scenario", "dambreak", "-dmax", str(meta["depth"]), "-nmax",
               str(meta["steps"]), "-nout", str(meta["nout"]), "-tmax", "-1", "-courant", "0.45",
               "-drytolerance", "0.01", "-solver", solver, "-dump", os.path.join(td, "s")]
        if dam:
            cmd += ["-dam_left", str(dam[0]), "-dam_right", str(dam[1])]
        print("case:", " ".join(cmd[1:]), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        for i in range(len(z["times"])):
            t, arr = read_dump(os.path.join(td, f"s_{i}.bin"))
            assert t == z["times"][i], f"output {i}"
            assert np.array_equal(arr, z["states"][i]), f"output {i}"


@pytest.mark.parametrize("d", [3, 4])
def test_cli_vtk_output_identical_to_reference_files(d):
    """The drop-in end to end: the CLI on the GPU writes the same VTK files,
    byte for byte, as the reference binary did for the same run (dam break,
    4 steps, output every 2 steps, tests/golden/vtu)."""
    vdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vtu")
    with tempfile.TemporaryDirectory() as td:
        cmd = [host.CLI, "-scenario", "dambreak", "-dmax", str(d), "-nmax", "4", "-nout", "2",
               "-tmax", "-1", "-xmloutput", ".true.", "-output_dir", td]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        for i in range(3):
            mine = open(os.path.join(td, f"swe_{i}_r0_s1.vtu"), "rb").read()
            ref = open(os.path.join(vdir, f"d{d}_{i}.vtu"), "rb").read()
            assert mine == ref, f"output {i}"
            assert os.path.exists(os.path.join(td, f"swe_{i}.pvtu"))


@pytest.mark.parametrize("d", [3, 4])
def test_cli_probe_output_matches_reference_files(d):
    """With probes the reference runs one more output traversal per output,
    which flips the direction once more: the VTK files and the probe files
    of the CLI match the reference run with -stestpoints (the centroid
    distance to its last bit, the rest exactly)."""
    vdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vtu")
    suffix = open(os.path.join(vdir, "points_name.txt")).read().strip()
    probes = "0.3 0.4,0.71 0.2,-0.5 0.25,-4.9 4.95,2.5 -1.25"
    with tempfile.TemporaryDirectory() as td:
        cmd = [host.CLI, "-scenario", "dambreak", "-dmax", str(d), "-nmax", "4", "-nout", "2",
               "-tmax", "-1", "-xmloutput", ".true.", "-stestpoints", probes, "-output_dir", td]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        for i in range(3):
            mine = open(os.path.join(td, f"swe_{i}_r0_s1.vtu"), "rb").read()
            assert mine == open(os.path.join(vdir, f"p{d}_{i}.vtu"), "rb").read(), f"output {i}"
            name = "swe" + suffix.replace("_d3_", f"_d{d}_").replace("_pointoutput_0", f"_pointoutput_{i}")
            got = open(os.path.join(td, name)).read().splitlines()
            ref = open(os.path.join(vdir, f"p{d}_{i}.points.txt")).read().splitlines()
            assert got[0] == ref[0] and len(got) == len(ref)
            for a, b in zip(got[1:], ref[1:]):
                fa, fb = a.split(","), b.split(",")
                assert fa[:7] == fb[:7] and fa[8] == fb[8], (a, b)
                assert abs(float(fa[7]) - float(fb[7])) <= 2e-15, (a, b)


def test_fortran_time_loop_vs_oracle():
    d, S = 12, 1.0e6
    mesh = host.mesh_uniform(d)
    st = host.tsunami(mesh, seed=9, drytol=1e-3)
    courant = 0.45
    op = swe.EulerTimestep(mesh, scaling=S, dry_tolerance=1e-3, courant=courant)
    op.upload(*st)
    dt0 = host.initial_dt(mesh, st, courant)
    # host-stepped loop with a time limit (pre_traversal_grid_op clamps to tmax)
    dt, t, fwd, steps = host.time_loop(op, dt=dt0, forward=False, nmax=9, tmax=-1.0,
                                       host_steps=True)
    ref = list(st)
    rdt, rt, rfwd = dt0, 0.0, False
    for _ in range(9):
        e, u, v, _, cand, _ = po.step(mesh.depth, mesh.plotter, mesh.nbr, ref, rdt, scaling=S,
                                      drytol=1e-3, forward=rfwd)
        ref = [e, u, v, ref[3]]
        rt += rdt
        rdt = courant * cand
        rfwd = not rfwd
    assert steps == 9 and dt == rdt and t == rt and fwd == rfwd
    assert all(np.array_equal(a, b) for a, b in zip(op.download(), ref))
    # device chain continues the same run
    dt2, t2, fwd2, steps2 = host.time_loop(op, dt=dt, time=t, forward=fwd, nmax=6)
    for _ in range(6):
        e, u, v, _, cand, _ = po.step(mesh.depth, mesh.plotter, mesh.nbr, ref, rdt, scaling=S,
                                      drytol=1e-3, forward=rfwd)
        ref = [e, u, v, ref[3]]
        rt += rdt
        rdt = courant * cand
        rfwd = not rfwd
    assert steps2 == 6 and dt2 == rdt and fwd2 == rfwd and t2 == rt
    assert all(np.array_equal(a, b) for a, b in zip(op.download(), ref))
    op.close()


def test_tmax_stops_and_clamps():
    d = 8
    mesh = host.mesh_uniform(d)
    st = host.reference_dam_break(mesh)
    op = swe.EulerTimestep(mesh, scaling=10.0, dry_tolerance=0.01, courant=0.45)
    op.upload(*st)
    dt0 = host.initial_dt(mesh, st, 0.45)
    dt, t, fwd, steps = host.time_loop(op, dt=dt0, forward=False, tmax=0.05, host_steps=True)
    assert t >= 0.05 and steps > 1
    op.close()


@pytest.mark.parametrize("name", ["plain_d2_9", "dam_d3_12"])
def test_cli_remeshing_run_identical_to_reference_files(name):
    """The whole program with its initial refinement (-dstart below -dmax)
    and remeshing before every step (-nadapt 1) on the device: every VTK file
    equals the reference's byte for byte (tests/golden/vtu/remesh_digests.json;
    dam_d3_12 is the reference's dam break raised to 10 m | 5 m)."""
    import hashlib
    import json

    vdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vtu")
    ref = json.load(open(os.path.join(vdir, "remesh_digests.json")))[name]
    extra = ["-dam_left", "10", "-dam_right", "5"] if ref["binary"] == "samoa_swe_dam" else []
    with tempfile.TemporaryDirectory() as td:
        cmd = [host.CLI, "-scenario", "dambreak", *ref["args"], *extra, "-nadapt", "1", "-tmax",
               "-1", "-xmloutput", ".true.", "-nout", "1", "-output_dir", td]
        print("case:", " ".join(cmd[1:]), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        for i in sorted(ref["files"], key=int):
            path = os.path.join(td, f"swe_{i}_r0_s1.vtu")
            mine = hashlib.sha256(open(path, "rb").read()).hexdigest()
            if mine != ref["files"][i]:
                v = po.read_vtu(path)
                pytest.fail(f"output {i} differs: {len(v['depth'])} cells, time {v['time'][0]!r}, "
                            f"depths {np.bincount(v['depth'])}")


def test_cli_adaptive_tsunami_equals_fortran_loop_from_python():
    """The adaptive tsunami through the CLI (initial refinement by the source's
    variation, remeshing before every step from surface-elevation jumps)
    equals, bit for bit, the same run of the Fortran host's set-up and loop
    driven from Python (swe_driver.f90 tsunami_adaptive_setup, adaptive_loop)."""
    tol, d0, d1, steps = 2e-2, 10, 14, 12
    f32 = lambda x: float(np.float32(x))  # noqa: E731  (the CLI reads REAL)
    with tempfile.TemporaryDirectory() as td:
        cmd = [host.CLI, "-scenario", "tsunami", "-dstart", str(d0), "-dmin", str(d0), "-dmax",
               str(d1), "-refine_tol", str(tol), "-nadapt", "1", "-nmax", str(steps), "-nout",
               str(steps), "-tmax", "-1", "-dump", os.path.join(td, "s")]
        print("case:", " ".join(cmd[1:]), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        t_cli, s_cli = read_dump(os.path.join(td, "s_1.bin"))
        _, s0 = read_dump(os.path.join(td, "s_0.bin"))
    mesh = host.mesh_uniform(d0)
    op = swe.EulerTimestep(mesh, scaling=1e6, dry_tolerance=f32(0.01), courant=f32(0.45),
                           min_depth=d0, max_depth=d1)
    z = np.zeros(mesh.n)
    op.upload(z, z, z, z)
    n, dt0 = host.tsunami_adaptive_setup(op, 0, tol, d1, f32(0.01), f32(0.45))
    assert n == s0.shape[1] > mesh.n
    assert np.array_equal(np.stack(op.download()), s0)
    # the CLI's initial output flipped the direction once
    dt, t, fwd, cells = host.adaptive_loop(op, steps, 1, tol / 2, tol / 16, 0, dt0, 0.0, True)
    assert t == t_cli
    assert op.n == s_cli.shape[1]
    assert np.array_equal(np.stack(op.download()), s_cli)
    op.close()
