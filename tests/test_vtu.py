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


def fixtures():
    return sorted(f for f in os.listdir(VTU) if f.endswith(".vtu"))


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


@pytest.mark.parametrize("name", fixtures())
def test_writer_reproduces_reference_file(name, tmp_path):
    d = int(re.match(r"d(\d+)_", name).group(1))
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
        d = int(re.match(r"d(\d+)_", name).group(1))
        dirs.add(unpack(name, host.mesh_uniform(d))[3])
    assert dirs == {True, False}


def test_pvtu_matches_reference(tmp_path):
    ref = open(os.path.join(VTU, "d3_0.pvtu"), "rb").read().decode()
    piece = re.search(r'Source="([^"]+)"', ref).group(1)
    out = tmp_path / "x.pvtu"
    host.write_pvtu(str(out), piece)
    assert out.read_text() == ref
