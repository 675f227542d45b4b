"""VTK output (samoa_amd/fortran/swe_vtu.f90) against the reference's own
output files (tests/golden/vtu, written by SWE_xml_output.f90 through
LIB_VTK_IO): the same cell state, time, traversal direction and refinement
flags give the same file, byte for byte."""
import os
import re

import numpy as np
import pytest

import pyoracle as po
from samoa_amd import host

VTU = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vtu")
S, OFF = 10.0, -5.0  # the reference scenario's domain [-5, 5]^2 (SWE.f90:169-170)


def fixtures(tag="d"):
    return sorted(f for f in os.listdir(VTU) if f.endswith(".vtu") and f.startswith(tag))


def unpack(name, mesh):
    """State (forward SFC order), time, flags and direction of a reference file."""
    v = po.read_vtu(os.path.join(VTU, name))
    n = mesh.n
    pts = v["Point"].reshape(n, 3, 3)[:, :, :2].mean(axis=1)
    cen = (S * mesh.coords + OFF).mean(axis=1)
    order = np.array([np.argmin(np.abs(cen - p).sum(axis=1)) for p in pts])
    forward = bool(np.array_equal(order, np.arange(n)))
    assert forward or np.array_equal(order, np.arange(n)[::-1])
    st = np.zeros((4, n))
    st[0, order] = v["water height"]
    st[3, order] = v["bathymetry"]
    m = v["momentum"].reshape(n, 3)
    st[1, order], st[2, order] = m[:, 0], m[:, 1]
    rf = np.zeros(n, np.int8)
    rf[order] = v["refinement flag"]
    return st, float(v["time"][0]), rf, forward


@pytest.mark.parametrize("name", fixtures("d") + fixtures("p"))
def test_writer_reproduces_reference_file(name, tmp_path):
    d = int(re.match(r"[dp](\d+)_", name).group(1))
    mesh = host.mesh_uniform(d)
    st, t, rf, forward = unpack(name, mesh)
    out = tmp_path / "out.vtu"
    host.write_vtu(str(out), mesh, st, time=t, scaling=S, offset=OFF, forward=forward, refine=rf)
    ref = open(os.path.join(VTU, name), "rb").read()
    mine = out.read_bytes()
    assert len(mine) == len(ref)
    assert mine == ref


def test_both_directions_covered():
    dirs = set()
    for name in fixtures():
        d = int(re.match(r"[dp](\d+)_", name).group(1))
        dirs.add(unpack(name, host.mesh_uniform(d))[3])
    assert dirs == {True, False}


def test_pvtu_matches_reference(tmp_path):
    ref = open(os.path.join(VTU, "d3_0.pvtu"), "rb").read().decode()
    piece = re.search(r'Source="([^"]+)"', ref).group(1)
    out = tmp_path / "x.pvtu"
    host.write_pvtu(str(out), piece)
    assert out.read_text() == ref


PROBES = [(0.3, 0.4), (0.71, 0.2), (-0.5, 0.25), (-4.9, 4.95), (2.5, -1.25)]


@pytest.mark.parametrize("name", fixtures("p"))
def test_probe_output_matches_reference(name, tmp_path):
    """SWE_point_output.f90's probe files for the same states: the same text,
    except that the centroid distance may differ in its last bit (the
    reference evaluates the centroid through its element transform)."""
    d = int(re.match(r"[dp](\d+)_", name).group(1))
    mesh = host.mesh_uniform(d)
    st, t, _, _ = unpack(name, mesh)
    vals = host.gauges_sample(mesh, st, PROBES, time=t, scaling=S, offset=OFF)
    out = tmp_path / "p.txt"
    host.gauges_write(str(out), PROBES, vals)
    ref = open(os.path.join(VTU, name.replace(".vtu", ".points.txt"))).read().splitlines()
    mine = out.read_text().splitlines()
    assert mine[0] == ref[0] and len(mine) == len(ref)
    for a, b in zip(mine[1:], ref[1:]):
        fa, fb = a.split(","), b.split(",")
        assert fa[:7] == fb[:7] and fa[8] == fb[8], (a, b)
        assert abs(float(fa[7]) - float(fb[7])) <= 2e-15, (a, b)
    full = tmp_path / "f.txt"
    host.gauges_write(str(full), PROBES, vals, full=True)
    back = np.array([[float(x) for x in line.split(",")] for line in full.read_text().splitlines()[1:]])
    assert np.array_equal(back[:, 3:], vals[:, 2:])
