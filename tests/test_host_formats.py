"""Host-side formats of the case driver, without a GPU: the reference's control files
(tests/golden/ctrl, copied from the reference's testcases and doc), the SU2 and Gmsh readers on
the reference's laminar NACA0012 mesh given in both formats, and the VTU writer.

The C++ headers are exercised through tests/host_harness.cpp, compiled here with g++."""
import math
import os
import shutil
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from conftest import ROOT, golden
from fvens_amd import gmsh, su2


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("host") / "host_harness")
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "fvens_amd", "host"),
                    os.path.join(ROOT, "tests", "host_harness.cpp"), "-o", exe], check=True)
    return exe


def _parse(harness, name):
    out = subprocess.run([harness, "ctrl", golden(os.path.join("ctrl", name))], check=True,
                         capture_output=True, text=True).stdout
    d = {}
    for line in out.splitlines():
        k, _, v = line.partition(" ")
        d[k] = v
    return d


def test_viscous_example_control_file(harness):
    d = _parse(harness, "example-control-file.ctrl")
    assert d["flowtype"] == "NAVIERSTOKES" and d["viscsim"] == "1" and d["lognres"] == "1"
    assert math.isclose(float(d["alpha"]), math.radians(2.0), rel_tol=1e-15)
    assert float(d["Reinf"]) == 5000.0 and float(d["Tinf"]) == 290.0 and float(d["Pr"]) == 0.72
    # bc0 farfield 4; bc1 subsonic inflow 3 with two values; bc2 adiabatic wall 5; bc3 isothermal 7
    assert d["bc0"].split() == ["1", "4"]
    assert d["bc1"].split() == ["3", "3", "2", "373"]
    assert d["bc2"].split() == ["7", "5", "20"]
    assert d["bc3"].split() == ["6", "7", "10", "290"]
    assert d["lwalls"].split() == ["5", "7"] and d["lothers"].split() == ["4"]
    assert d["invflux"] == "ROE" and d["invfluxjac"] == "ROE"       # "consistent"
    # limiter_parameter 20.0 is in the file; the reference does not parse it (zero)
    assert d["limiter"] == "WENO" and float(d["limparam"]) == 0.0
    assert d["pseudotime"] == "IMPLICIT"
    assert [float(x) for x in d["main"].split()] == [100.0, 2000.0, 1e-5, 500]
    assert [float(x) for x in d["starter"].split()] == [1, 50.0, 500.0, 0.1, 50]
    assert d["nlupdate"].split() == ["ROBUST_FLOW", "0.29999999999999999"]


def test_inviscid_and_explicit_control_files(harness):
    d = _parse(harness, "naca0012-transonic-implicit.ctrl")
    assert d["meshfile"] == "testcases/naca0012/grids/NACA0012_inv.su2"
    assert d["flowtype"] == "EULER" and d["viscsim"] == "0" and d["Reinf"] == "inf"
    assert math.isclose(float(d["Minf"]), 0.8) and d["limiter"] == "WENO"
    assert d["bc0"].split() == ["0", "2"] and d["bc1"].split() == ["1", "4"]
    assert d["nlupdate"].split()[0] == "FULL"
    e = _parse(harness, "2dcylinder-explicit.ctrl")
    assert e["pseudotime"] == "EXPLICIT" and e["invflux"] == "HLL" and e["invfluxjac"] == ""
    assert [float(x) for x in e["main"].split()] == [0.2, 0.2, 1e-5, 500000]
    v = _parse(harness, "visc-naca0012-laminar-implicit.ctrl")
    assert v["bc0"].split() == ["2", "4"] and v["bc1"].split() == ["7", "2", "0"]
    assert v["gradient"] == "LEASTSQUARES" and v["limiter"] == "NONE" and v["order2"] == "1"
    assert [float(x) for x in v["main"].split()] == [500.0, 5000.0, 1e-6, 80]
    assert [float(x) for x in v["starter"].split()] == [1, 200.0, 1000.0, 0.1, 50]


def test_su2_and_gmsh_readers_agree(harness):
    # the reference ships its laminar NACA0012 mesh in both formats
    a = gmsh.read_gmsh2(golden("NACA0012_lam_hybrid_1.msh"))
    b = su2.read_su2(golden("NACA0012_lam_hybrid_1.su2"))
    assert (a.npoin, a.nelem, a.nbface) == (b.npoin, b.nelem, b.nbface) == (8741, 13156, 250)
    assert np.array_equal(a.coords, b.coords)
    assert np.array_equal(a.inpoel, b.inpoel) and np.array_equal(a.nnode, b.nnode)
    key = lambda m: sorted(map(tuple, np.c_[np.sort(m.bface, 1), m.btag]))  # noqa: E731
    assert key(a) == key(b)
    outs = [subprocess.run([harness, "mesh", golden(f)], check=True, capture_output=True,
                           text=True).stdout for f in ("NACA0012_lam_hybrid_1.msh",
                                                        "NACA0012_lam_hybrid_1.su2")]
    assert outs[0] == outs[1] and outs[0].split()[:3] == ["8741", "13156", "250"]
    assert su2.read_mesh(golden("NACA0012_lam_hybrid_1.su2")).nelem == 13156


def test_vtu_point_data(harness, tmp_path):
    mesh = golden("squareunsquad0.msh")
    out = str(tmp_path / "s.vtu")
    subprocess.run([harness, "vtu", mesh, out], check=True)
    raw = gmsh.read_gmsh2(mesh)
    root = ET.parse(out).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == raw.npoin
    assert int(piece.get("NumberOfCells")) == raw.nelem
    arrays = {a.get("Name"): np.array(a.text.split(), float)
              for a in piece.iter("DataArray") if a.get("Name")}
    # the harness's state: rho = 1 + 0.001 (e mod 17), averaged to points with area weights
    x, y = raw.coords[:, 0], raw.coords[:, 1]
    area = np.zeros(raw.nelem)
    for e in range(raw.nelem):
        n = raw.nnode[e]
        p = raw.inpoel[e, :n]
        area[e] = 0.5 * np.sum(x[p] * y[np.roll(p, -1)] - x[np.roll(p, -1)] * y[p])
    e_ = np.arange(raw.nelem)
    rho = 1 + 0.001 * (e_ % 17)
    vx, vy, pr = 0.5, 0.01 * (e_ % 5), 1.0 / (1.4 * 0.25)
    u = np.stack([rho, rho * vx, rho * vy, pr / 0.4 + 0.5 * rho * (vx * vx + vy * vy)], 1)
    num, den = np.zeros((raw.npoin, 4)), np.zeros(raw.npoin)
    for e in range(raw.nelem):
        for p in raw.inpoel[e, :raw.nnode[e]]:
            num[p] += u[e] * area[e]
            den[p] += area[e]
    q = num / den[:, None]
    np.testing.assert_allclose(arrays["density"], q[:, 0], rtol=1e-9)
    wx, wy = q[:, 1] / q[:, 0], q[:, 2] / q[:, 0]
    pp = 0.4 * (q[:, 3] - 0.5 * q[:, 0] * (wx * wx + wy * wy))
    np.testing.assert_allclose(arrays["pressure"], pp, rtol=1e-9)
    np.testing.assert_allclose(arrays["mach-number"], np.hypot(wx, wy) / np.sqrt(1.4 * pp / q[:, 0]),
                               rtol=1e-9)
    np.testing.assert_allclose(arrays["temperature"], 1.4 * 0.25 * pp / q[:, 0], rtol=1e-9)
    offsets = arrays["offsets"].astype(int)
    assert np.array_equal(offsets, np.cumsum(raw.nnode))
    assert set(arrays["types"].astype(int)) <= {5, 9}


def test_other_marker_surface_file(harness, tmp_path):
    # exportSurfaceData's file for a non-wall marker (src/spatial/aoutput.cpp:243-296): face
    # centre and the host cell's velocity per boundary face with that marker, in boundary order
    mesh = golden("NACA0012_lam_hybrid_1.msh")
    raw = gmsh.read_gmsh2(mesh)
    prefix = str(tmp_path / "case")
    out = subprocess.run([harness, "surf", mesh, prefix, "4"], check=True, capture_output=True,
                         text=True).stdout
    host = np.array(out.split(), int)
    # the host cell of each boundary face holds both of its points
    for b in range(raw.nbface):
        pts = set(raw.inpoel[host[b], :raw.nnode[host[b]]])
        assert set(raw.bface[b]) <= pts
    lines = open(prefix + "-surf_o4.out").read().splitlines()
    assert lines[0].split() == ["#", "x", "y", "u", "v"]
    rows = np.array([ln.split() for ln in lines[1:]], float)
    far = np.flatnonzero(raw.btag == 4)
    assert len(rows) == len(far) == 50
    mid = 0.5 * (raw.coords[raw.bface[far, 0]] + raw.coords[raw.bface[far, 1]])
    np.testing.assert_allclose(rows[:, :2], mid, rtol=1e-5, atol=1e-9)
    e = host[far]
    np.testing.assert_allclose(rows[:, 2], 0.5, rtol=1e-12)
    np.testing.assert_allclose(rows[:, 3], 0.01 * (e % 5), rtol=1e-5, atol=1e-12)


def test_gmsh_writer_round_trip(harness, tmp_path):
    # UMesh::writeGmsh2 layout (src/mesh/mesh.cpp:204-281), used by the reference's format
    # converter: the SU2 mesh written as Gmsh 2.2 reads back as the same mesh, with the
    # reference's element numbering and tags
    out = str(tmp_path / "naca.msh")
    subprocess.run([harness, "msh", golden("NACA0012_lam_hybrid_1.su2"), out], check=True)
    a = su2.read_su2(golden("NACA0012_lam_hybrid_1.su2"))
    b = gmsh.read_gmsh2(out)
    assert np.array_equal(a.coords, b.coords)
    assert np.array_equal(a.inpoel, b.inpoel) and np.array_equal(a.nnode, b.nnode)
    assert np.array_equal(a.bface, b.bface) and np.array_equal(a.btag, b.btag)
    lines = open(out).read().splitlines()
    first = lines.index("$Elements") + 2
    assert lines[first].split()[:5] == ["1", "1", "2", str(a.btag[0]), str(1 + a.btag[0])]
    cell = lines[first + a.nbface].split()
    assert cell[0] == str(a.nbface + 1) and cell[2:5] == ["2", "1", "1"]
