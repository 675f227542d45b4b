"""Writes the golden fixtures under tests/golden/ by executing the reference.

Runs only in a container that holds the reference tree and the binaries that
`make -C oracle ref` compiles from it (oracle/_ref/).  The fixtures are data
(inputs and the reference's outputs); the tests compare the oracle and the HIP
library against them, so neither the reference nor its binaries are needed at
test time.

    python tests/golden/make_golden.py        (all fixtures)
    python tests/golden/make_golden.py vtu    (the VTK output files only)
    python tests/golden/make_golden.py adapt  (the adaptive-mesh runs only)
    python tests/golden/make_golden.py remesh (the runs that remesh every step)
    python tests/golden/make_golden.py radial (the radial dam break in a closed basin)

Fixtures:
  kat_geoclaw.npz      inputs (n,10) + reference outputs of c_bind_geoclaw_solver
                       (c_bind_riemannsolvers.f90:16-176) for the augmented
                       (GEOCLAW_AUG_RIEMANN) and f-wave (GEOCLAW_FWAVE) solvers
  kat_riemanntype.npz  inputs (n,6) + outputs of riemanntype (riemannsolvers.f:507-618)
  tables.npz           bit patterns of r_leg_sizes / r_volumes (SFC_data_types.f90:524,557)
  run_*.npz            full SWE runs of the reference binary: mesh point order,
                       per-output state (forward SFC order), time, refinement flags
  vtu/                 the reference's own VTK output files of two small runs
  adapt_*.npz          runs on the reference's adaptively refined (then frozen)
                       meshes: cell corners, depths, states, times, flags
  run_radial_*.npz     the radial dam break in a closed basin (unit square, walls)
  radial_d14_n100.json the same at depth 14 over 100 steps: digests of the
                       initial and final states (the 32,768-cell run itself)
  remesh_*.npz         runs that remesh before every step (-nadapt 1) with the
                       dam-break variant of the reference (10 m | 5 m): per
                       output the depth sequence of the mesh (which fixes it),
                       state, flags, time and output direction
"""
from __future__ import annotations

import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)
import pyoracle as po  # noqa: E402

G = 9.80665
F32_COURANT = float(np.float32(0.45))  # Config.f90:243 rget() is default REAL
F32_DRYTOL = float(np.float32(0.01))   # Config.f90:284


def kat_inputs(seed: int = 20261202) -> np.ndarray:
    """Designed Riemann states: wet/wet shocks, rarefactions, transcritical and
    sonic states, wet/dry with and without a wall, dry/dry, lake at rest,
    identical states, the drytol boundary itself, and random states."""
    rng = np.random.default_rng(seed)
    dry = F32_DRYTOL
    rows = []

    def add(hL, hR, huL, huR, hvL, hvR, bL, bR):
        rows.append([hL, hR, huL, huR, hvL, hvR, bL, bR, dry, G])

    for _ in range(3000):  # random wet/wet
        hL, hR = rng.uniform(0.05, 20.0, 2)
        uL, uR, vL, vR = rng.uniform(-8, 8, 4)
        bL, bR = rng.uniform(-3, 3, 2)
        add(hL, hR, hL * uL, hR * uR, hL * vL, hR * vR, bL, bR)
    for _ in range(800):  # strong shocks / rarefactions, flat bed
        hL, hR = 10 ** rng.uniform(-1.5, 1.5, 2)
        s = rng.choice([-1.0, 1.0])
        uL, uR = s * rng.uniform(0, 12), -s * rng.uniform(0, 12)
        add(hL, hR, hL * uL, hR * uR, 0.0, 0.0, 0.0, 0.0)
    for _ in range(800):  # transcritical / sonic: |u| close to sqrt(g h)
        hL, hR = rng.uniform(0.1, 5.0, 2)
        cL, cR = np.sqrt(G * hL), np.sqrt(G * hR)
        uL = rng.choice([-1, 1]) * cL * (1 + rng.normal(0, 1e-3))
        uR = rng.choice([-1, 1]) * cR * (1 + rng.normal(0, 1e-3))
        bL, bR = rng.uniform(-0.5, 0.5, 2)
        add(hL, hR, hL * uL, hR * uR, hL * rng.normal(), hR * rng.normal(), bL, bR)
    for _ in range(800):  # wet/dry, either side, with and without a wall
        hW = rng.uniform(0.02, 5.0)
        hD = rng.choice([0.0, dry * rng.uniform(0, 1), dry, -rng.uniform(0, 1)])
        uW, vW = rng.uniform(-5, 5, 2)
        bW = rng.uniform(-3, 0)
        bD = bW + rng.choice([rng.uniform(0, 2 * hW), rng.uniform(2 * hW, 10), -rng.uniform(0, 2)])
        huD = rng.normal() * abs(hD)
        if rng.random() < 0.5:
            add(hW, hD, hW * uW, huD, hW * vW, 0.3 * huD, bW, bD)
        else:
            add(hD, hW, huD, hW * uW, 0.3 * huD, hW * vW, bD, bW)
    for _ in range(200):  # dry/dry
        hL, hR = rng.uniform(-1, dry, 2)
        add(hL, hR, rng.normal(), rng.normal(), rng.normal(), rng.normal(), rng.normal(),
            rng.normal())
    for _ in range(600):  # lake at rest: eta equal, p = 0, varying b
        eta = rng.uniform(-1, 1)
        bL, bR = eta - rng.uniform(0.02, 10, 2)
        add(eta - bL, eta - bR, 0.0, 0.0, 0.0, 0.0, bL, bR)
    for _ in range(400):  # identical states (outflow ghost)
        h = rng.uniform(0.0, 10.0)
        u, v, b = rng.normal(0, 3, 3)
        add(h, h, h * u, h * u, h * v, h * v, b, b)
    for _ in range(200):  # mirrored states (wall ghost)
        h = rng.uniform(0.02, 10.0)
        u, v, b = rng.normal(0, 3, 3)
        add(h, h, h * u, -h * u, h * v, h * v, b, b)
    for _ in range(200):  # zero velocities, flat bed, depth jump
        hL, hR = rng.uniform(0.0, 10.0, 2)
        add(hL, hR, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    return np.asarray(rows, dtype=np.float64)


def rt_inputs(seed: int = 7) -> np.ndarray:
    rng = np.random.default_rng(seed)
    dry = F32_DRYTOL
    rows = []
    for _ in range(3000):
        hL, hR = 10 ** rng.uniform(-3, 1.5, 2)
        uL, uR = rng.uniform(-15, 15, 2)
        rows.append([hL, hR, uL, uR, dry, G])
    for _ in range(300):
        h = rng.uniform(0, 3 * dry)
        rows.append([h, rng.uniform(0, 5), rng.normal(0, 3), rng.normal(0, 3), dry, G])
        rows.append([rng.uniform(0, 5), h, rng.normal(0, 3), rng.normal(0, 3), dry, G])
    for _ in range(300):  # the wall test call of c_bind: (hL, hL, uL, -uL)
        h = rng.uniform(dry, 10)
        u = rng.normal(0, 5)
        rows.append([h, h, u, -u, dry, G])
    return np.asarray(rows, dtype=np.float64)


def run_kat(mode: str, inputs: np.ndarray, *extra: str, nout: int) -> np.ndarray:
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        inputs.astype("<f8").tofile(fin)
        subprocess.run([os.path.join(REF, "ref_kat"), mode, *extra, fin, fout], check=True)
        return np.fromfile(fout, "<f8").reshape(-1, nout)


def run_reference(binary: str, depth: int, steps: int, nout: int) -> list[dict]:
    od = tempfile.mkdtemp(prefix="swe_golden_")
    try:
        cmd = [os.path.join(REF, binary), "-threads", "1", "-sections", "1",
               "-dstart", str(depth), "-dmin", str(depth), "-dmax", str(depth),
               "-nadapt", "-1", "-tmax", "-1", "-nmax", str(steps), "-xmloutput", ".true.",
               "-nout", str(nout), "-output_dir", od]
        subprocess.run(cmd, check=True, capture_output=True)
        files = sorted(glob.glob(os.path.join(od, "*_r0_s1.vtu")),
                       key=lambda f: int(f.split("_")[-3]))
        return [po.read_vtu(f) for f in files]
    finally:
        shutil.rmtree(od, ignore_errors=True)


def write_run(name: str, binary: str, depth: int, steps: int, nout: int, bc: int,
              solver: int = 0, scaling: float = 10.0, offset: float = -5.0) -> None:
    """scaling, offset: the domain of the binary's scenario (SWE.f90:169-170;
    [-5,5]^2 unless a variant patches it)."""
    outs = run_reference(binary, depth, steps, nout)
    S, off = scaling, offset
    n = 1 << (depth + 1)
    pts0 = outs[0]["Point"].reshape(n, 3, 3)[:, :, :2]
    # map reference output order -> forward SFC index via the oracle mesh
    _, co, _ = po.mesh_uniform(depth)
    cen = (S * co + off).mean(axis=1)
    index = {tuple(np.round(c, 9)): i for i, c in enumerate(cen)}
    states, times, refine, order_fwd = [], [], [], []
    for v in outs:
        pts = v["Point"].reshape(n, 3, 3)[:, :, :2]
        perm = np.array([index[tuple(np.round(c, 9))] for c in pts.mean(axis=1)])
        st = np.zeros((4, n))
        st[0, perm] = v["water height"]
        st[3, perm] = v["bathymetry"]
        m = v["momentum"].reshape(n, 3)
        st[1, perm] = m[:, 0]
        st[2, perm] = m[:, 1]
        rf = np.zeros(n, np.int8)
        rf[perm] = v["refinement flag"]
        states.append(st)
        times.append(v["time"][0])
        refine.append(rf)
        order_fwd.append(bool(np.array_equal(perm, np.arange(n))))
        assert order_fwd[-1] or np.array_equal(perm, np.arange(n)[::-1])
    assert all(np.isfinite(st).all() for st in states), f"{name}: non-finite output"
    meta = dict(binary=binary, depth=depth, steps=steps, nout=nout, bc=bc, scaling=S,
                offset=off, courant=F32_COURANT, drytol=F32_DRYTOL, solver=solver,
                cmd="-threads 1 -sections 1 -dstart d -dmin d -dmax d -nadapt -1 -tmax -1 "
                    f"-nmax {steps} -xmloutput .true. -nout {nout}")
    np.savez_compressed(os.path.join(HERE, f"run_{name}.npz"), states=np.asarray(states),
                        times=np.asarray(times), refine=np.asarray(refine),
                        order_fwd=np.asarray(order_fwd), points0=pts0,
                        meta=json.dumps(meta))
    print(f"run_{name}: {len(outs)} outputs, n={n}")


PROBES = "0.3 0.4,0.71 0.2,-0.5 0.25,-4.9 4.95,2.5 -1.25"


def write_vtu_fixtures() -> None:
    """The reference's own VTK and probe output files (SWE_xml_output.f90,
    SWE_point_output.f90), unmodified: dam break at depths 3 and 4, 4 steps,
    an output every 2 steps.  Run A writes VTK files only, so the output
    traversals alternate in direction (vtu/d<d>_<i>.vtu, one .pvtu); run B
    adds the probes PROBES, whose traversal flips the direction once more per
    output (vtu/p<d>_<i>.vtu and vtu/p<d>_<i>.points.txt)."""
    vdir = os.path.join(HERE, "vtu")
    os.makedirs(vdir, exist_ok=True)
    for d in (3, 4):
        for tag, extra in (("d", []), ("p", ["-stestpoints", PROBES])):
            od = tempfile.mkdtemp(prefix="swe_vtu_")
            try:
                subprocess.run([os.path.join(REF, "samoa_swe"), "-threads", "1", "-sections", "1",
                                "-dstart", str(d), "-dmin", str(d), "-dmax", str(d), "-nadapt",
                                "-1", "-tmax", "-1", "-nmax", "4", "-xmloutput", ".true.", "-nout",
                                "2", *extra, "-output_dir", od], check=True, capture_output=True)
                for f in glob.glob(os.path.join(od, "*_r0_s1.vtu")):
                    i = int(f.split("_")[-3])
                    shutil.copy(f, os.path.join(vdir, f"{tag}{d}_{i}.vtu"))
                for f in glob.glob(os.path.join(od, "*_pointoutput_*.txt")):
                    i = int(f.rsplit("_", 1)[1].split(".")[0])
                    shutil.copy(f, os.path.join(vdir, f"{tag}{d}_{i}.points.txt"))
                if d == 3 and tag == "d":
                    p0 = sorted(glob.glob(os.path.join(od, "*_0.pvtu")))[0]
                    shutil.copy(p0, os.path.join(vdir, "d3_0.pvtu"))
                if d == 3 and tag == "p":
                    name = os.path.basename(glob.glob(os.path.join(od, "*_pointoutput_0.txt"))[0])
                    with open(os.path.join(vdir, "points_name.txt"), "w") as fh:
                        fh.write(name[name.index("_d"):] + "\n")
            finally:
                shutil.rmtree(od, ignore_errors=True)
    for f in os.listdir(vdir):
        os.chmod(os.path.join(vdir, f), 0o644)
    print("vtu fixtures:", sorted(os.listdir(vdir)))


def write_adaptive_run(name: str, binary: str, dstart: int, dmin: int, dmax: int, steps: int,
                       bc: int) -> None:
    """A run on an adaptive mesh: the reference's initial refinement from
    depth dstart up to dmax (SWE_initialize.f90), then frozen (-nadapt -1),
    an output after every step.  The cells are stored in forward curve order
    (the reference writes them in the order of its output traversal, which
    runs backward here); the mesh is rebuilt from the corners by the test."""
    od = tempfile.mkdtemp(prefix="swe_golden_")
    try:
        cmd = [os.path.join(REF, binary), "-threads", "1", "-sections", "1", "-dstart", str(dstart),
               "-dmin", str(dmin), "-dmax", str(dmax), "-nadapt", "-1", "-tmax", "-1", "-nmax",
               str(steps), "-xmloutput", ".true.", "-nout", "1", "-output_dir", od]
        subprocess.run(cmd, check=True, capture_output=True)
        files = sorted(glob.glob(os.path.join(od, "*_r0_s1.vtu")),
                       key=lambda f: int(f.split("_")[-3]))
        outs = [po.read_vtu(f) for f in files]
    finally:
        shutil.rmtree(od, ignore_errors=True)
    n = len(outs[0]["depth"])
    pts = outs[0]["Point"].reshape(n, 3, 3)[:, :, :2]
    for v in outs[1:]:  # frozen mesh, same output order
        assert np.array_equal(v["Point"].reshape(n, 3, 3)[:, :, :2], pts)
    rev = np.arange(n)[::-1]
    states, times, refine = [], [], []
    for v in outs:
        m = v["momentum"].reshape(n, 3)
        states.append(np.stack([v["water height"], m[:, 0], m[:, 1], v["bathymetry"]])[:, rev])
        times.append(v["time"][0])
        refine.append(v["refinement flag"][rev])
    meta = dict(binary=binary, dstart=dstart, dmin=dmin, dmax=dmax, steps=steps, nout=1, bc=bc,
                scaling=10.0, offset=-5.0, courant=F32_COURANT, drytol=F32_DRYTOL,
                cmd=" ".join(cmd[1:]).replace(od, "<dir>"))
    np.savez_compressed(os.path.join(HERE, f"adapt_{name}.npz"), corners=pts[rev],
                        depth=outs[0]["depth"][rev], states=np.array(states), times=np.array(times),
                        refine=np.array(refine), meta=json.dumps(meta))
    print(f"adapt_{name}: {len(outs)} outputs, n={n}, depths {np.bincount(outs[0]['depth'])}")


# The reference built with its Lax-Friedrichs flux solvers (oracle/Makefile:
# samoa_swe_*_lf, _llf, _lfbath, _llfbath): solver ids of the library.
LF_RUNS = [
    # the global variant's fixed alpha = 100 outruns the CFL step: the
    # reference's own run turns non-finite after three steps
    ("lf_dam_d7_n1", "samoa_swe_dam_lf", 7, 3, 1, 0, 2),
    ("llf_dam_d8_n3", "samoa_swe_dam_llf", 8, 24, 3, 0, 3),
    ("lfbath_bath_d7_n2", "samoa_swe_bath_lfbath", 7, 12, 2, 1, 4),
    ("llfbath_bath_d8_n3", "samoa_swe_bath_llfbath", 8, 24, 3, 1, 5),
    ("llfbath_plain_d7_n2", "samoa_swe_llfbath", 7, 16, 2, 0, 5),
]


def write_lf_runs() -> None:
    for args in LF_RUNS:
        write_run(*args)


def main() -> None:
    if not os.path.exists(os.path.join(REF, "ref_kat")):
        sys.exit("build the reference first: make -C oracle ref")
    if sys.argv[1:] == ["vtu"]:
        write_vtu_fixtures()
        return
    if sys.argv[1:] == ["adapt"]:
        write_adaptive_fixtures()
        return
    if sys.argv[1:] == ["lf"]:
        write_lf_runs()
        return
    if sys.argv[1:] == ["remesh"]:
        write_remesh_fixtures()
        return
    if sys.argv[1:] == ["radial"]:
        write_radial_fixtures()
        return
    kin = kat_inputs()
    aug = run_kat("geoclaw", kin, "3", nout=7)
    fwave = run_kat("geoclaw", kin, "1", nout=7)
    np.savez_compressed(os.path.join(HERE, "kat_geoclaw.npz"), inputs=kin, aug=aug, fwave=fwave)
    rin = rt_inputs()
    rt = run_kat("riemanntype", rin, nout=5)
    np.savez_compressed(os.path.join(HERE, "kat_riemanntype.npz"), inputs=rin, outputs=rt)
    print(f"kat: {len(kin)} geoclaw records, {len(rin)} riemanntype records")

    lines = subprocess.run([os.path.join(REF, "ref_tables")], check=True, capture_output=True,
                           text=True).stdout.split("\n")
    tab = {"leg": {}, "vol": {}, "scl": {}}
    for line in lines:
        if line.strip():
            k, i, h = line.split()
            tab[k][int(i)] = int(h, 16)
    np.savez_compressed(os.path.join(HERE, "tables.npz"),
                        leg=np.array([tab["leg"][i] for i in range(-1, 61)], np.uint64),
                        vol=np.array([tab["vol"][i] for i in range(0, 61)], np.uint64))

    write_run("plain_d6_n1", "samoa_swe", 6, 20, 1, 0)
    write_run("plain_d8_n3", "samoa_swe", 8, 30, 3, 0)
    write_run("plain_d10_n5", "samoa_swe", 10, 25, 5, 0)
    write_run("bath_d6_n1", "samoa_swe_bath", 6, 30, 1, 1)
    write_run("bath_d7_n2", "samoa_swe_bath", 7, 40, 2, 1)
    write_run("bath_d9_n7", "samoa_swe_bath", 9, 35, 7, 1)
    write_vtu_fixtures()
    write_adaptive_fixtures()
    write_radial_fixtures()


def write_radial_fixtures() -> None:
    """The radial dam break in a closed basin (the reference's VARIANT=radial
    build: unit square, eta 10 m inside r < 0.1 around the centre, 5 m
    outside, flat bed, reflecting walls), 100 steps at depths 8 and 10, and
    the same run at depth 14 (32,768 cells) whose final state is stored as
    SHA-256 digests of its bytes beside a few sums."""
    write_run("radial_d8_n10", "samoa_swe_radial", 8, 100, 10, 1, scaling=1.0, offset=0.0)
    write_run("radial_d10_n50", "samoa_swe_radial", 10, 100, 50, 1, scaling=1.0, offset=0.0)
    import hashlib

    d, steps = 14, 100
    outs = run_reference("samoa_swe_radial", d, steps, steps)
    n = 1 << (d + 1)
    _, co, _ = po.mesh_uniform(d)
    index = {tuple(np.round(c, 9)): i for i, c in enumerate(co.mean(axis=1))}
    res = {}
    for tag, v in (("initial", outs[0]), ("final", outs[-1])):
        pts = v["Point"].reshape(n, 3, 3)[:, :, :2]
        perm = np.array([index[tuple(np.round(c, 9))] for c in pts.mean(axis=1)])
        st = np.zeros((4, n))
        st[0, perm] = v["water height"]
        m = v["momentum"].reshape(n, 3)
        st[1, perm] = m[:, 0]
        st[2, perm] = m[:, 1]
        st[3, perm] = v["bathymetry"]
        rf = np.zeros(n, np.int8)
        rf[perm] = v["refinement flag"]
        res[tag] = dict(
            time=float(v["time"][0]),
            sha256={k: hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()
                    for k, a in zip(("eta", "hu", "hv", "b"), st)},
            refine_sha256=hashlib.sha256(rf.tobytes()).hexdigest(),
            sums={k: float(a.sum()) for k, a in zip(("eta", "hu", "hv"), st[:3])},
            forward=bool(np.array_equal(perm, np.arange(n))))
    meta = dict(binary="samoa_swe_radial", depth=d, steps=steps, nout=steps, bc=1, scaling=1.0,
                offset=0.0, courant=F32_COURANT, drytol=F32_DRYTOL, solver=0,
                cmd="-threads 1 -sections 1 -dstart 14 -dmin 14 -dmax 14 -nadapt -1 -tmax -1 "
                    f"-nmax {steps} -xmloutput .true. -nout {steps}")
    with open(os.path.join(HERE, "radial_d14_n100.json"), "w") as fh:
        json.dump(dict(meta=meta, **res), fh, indent=1, sort_keys=True)
    print(f"radial_d14_n100: final time {res['final']['time']!r}")


def write_remesh_run(name: str, binary: str, dstart: int, dmin: int, dmax: int, steps: int,
                     bc: int) -> None:
    """A run that remeshes before every time step (SWE.f90:387-390 with
    -nadapt 1), an output after every step.  The output traversal alternates
    its direction; each output is stored in forward curve order (a forward
    output starts in the root triangle below the diagonal).  The mesh of an
    output is stored as its depth sequence; the corners the reference wrote
    are checked here against the mesh that sequence fixes."""
    from samoa_amd import host

    od = tempfile.mkdtemp(prefix="swe_golden_")
    try:
        cmd = [os.path.join(REF, binary), "-threads", "1", "-sections", "1", "-dstart", str(dstart),
               "-dmin", str(dmin), "-dmax", str(dmax), "-nadapt", "1", "-tmax", "-1", "-nmax",
               str(steps), "-xmloutput", ".true.", "-nout", "1", "-output_dir", od]
        subprocess.run(cmd, check=True, capture_output=True)
        files = sorted(glob.glob(os.path.join(od, "*_r0_s1.vtu")),
                       key=lambda f: int(f.split("_")[-3]))
        outs = [po.read_vtu(f) for f in files]
    finally:
        shutil.rmtree(od, ignore_errors=True)
    depth, states, refine, times, fwd, offsets = [], [], [], [], [], [0]
    for v in outs:
        n = len(v["depth"])
        pts = v["Point"].reshape(n, 3, 3)[:, :, :2]
        cen = pts.mean(axis=1)
        forward = bool(cen[0, 1] < cen[0, 0])
        idx = np.arange(n) if forward else np.arange(n)[::-1]
        mesh = host.mesh_from_depths(v["depth"][idx])
        ours = (10.0 * mesh.coords - 5.0)
        ref = np.where((mesh.plotter > 0)[:, None, None], pts[idx], pts[idx][:, ::-1])
        assert np.array_equal(ours, ref), "corners differ from the depth sequence's mesh"
        m = v["momentum"].reshape(n, 3)
        depth.append(v["depth"][idx])
        states.append(np.stack([v["water height"], m[:, 0], m[:, 1], v["bathymetry"]])[:, idx])
        refine.append(v["refinement flag"][idx])
        times.append(v["time"][0])
        fwd.append(forward)
        offsets.append(offsets[-1] + n)
    meta = dict(binary=binary, dstart=dstart, dmin=dmin, dmax=dmax, steps=steps, nout=1, bc=bc,
                scaling=10.0, offset=-5.0, courant=F32_COURANT, drytol=F32_DRYTOL,
                cmd=" ".join(cmd[1:]).replace(od, "<dir>"))
    np.savez_compressed(os.path.join(HERE, f"remesh_{name}.npz"), depth=np.concatenate(depth),
                        offsets=np.array(offsets), states=np.concatenate(states, axis=1),
                        refine=np.concatenate(refine), times=np.array(times),
                        forward=np.array(fwd), meta=json.dumps(meta))
    print(f"remesh_{name}: {len(outs)} outputs, cells {[offsets[i + 1] - offsets[i] for i in range(len(outs))]}")


def write_remesh_fixtures() -> None:
    write_remesh_run("dam_d3_12", "samoa_swe_dam", 3, 3, 12, 30, 0)
    write_remesh_run("dam_d2_9", "samoa_swe_dam", 2, 2, 9, 30, 0)
    write_remesh_vtu_digests()


# runs of the whole reference program with the initial refinement (-dstart <
# -dmax) and remeshing before every step, VTK output after every step
REMESH_CLI_RUNS = {
    "plain_d2_9": ("samoa_swe", ["-dstart", "2", "-dmin", "2", "-dmax", "9", "-nmax", "12"]),
    "dam_d3_12": ("samoa_swe_dam", ["-dstart", "3", "-dmin", "3", "-dmax", "12", "-nmax", "20"]),
}


def write_remesh_vtu_digests() -> None:
    """SHA-256 of every VTK file the reference writes in REMESH_CLI_RUNS
    (-nadapt 1 -nout 1), so that the CLI's files can be compared byte for
    byte without storing them (vtu/remesh_digests.json)."""
    import hashlib

    out = {}
    for name, (binary, args) in REMESH_CLI_RUNS.items():
        od = tempfile.mkdtemp(prefix="swe_vtu_")
        try:
            cmd = [os.path.join(REF, binary), "-threads", "1", "-sections", "1", *args, "-nadapt",
                   "1", "-tmax", "-1", "-xmloutput", ".true.", "-nout", "1", "-output_dir", od]
            subprocess.run(cmd, check=True, capture_output=True)
            files = {}
            for f in glob.glob(os.path.join(od, "*_r0_s1.vtu")):
                i = int(f.split("_")[-3])
                files[str(i)] = hashlib.sha256(open(f, "rb").read()).hexdigest()
        finally:
            shutil.rmtree(od, ignore_errors=True)
        out[name] = dict(binary=binary, args=args, files=files)
        print(f"{name}: {len(files)} VTK files")
    with open(os.path.join(HERE, "vtu", "remesh_digests.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)


def write_adaptive_fixtures() -> None:
    write_adaptive_run("plain_d2_9", "samoa_swe", 2, 2, 9, 12, 0)
    write_adaptive_run("bath_d3_11", "samoa_swe_bath", 3, 3, 11, 15, 1)


if __name__ == "__main__":
    main()
