"""Line and reverse Cuthill-McKee orderings (host code in libfvx, no GPU) against the
reference's mesh-ordering fixtures: tests/common-input/testanisotropic.msh with the lines of
findLines (testanisotropic-lines.txt, tests/mesh/testlineordering.cpp) and the hybrid
line-point orderings in natural and RCM order (testanisotropic-lineordering.txt,
testanisotropic-line_rcm_ordering.txt, tests/mesh/testhybridlineordering.cpp), all at
anisotropy threshold 10. The fixtures number cells as Gmsh elements: index + nbface + 1."""
import numpy as np
import pytest

from conftest import golden
from fvens_amd import fvx, gmsh, meshgen


@pytest.fixture(scope="module")
def aniso():
    return gmsh.read_gmsh2(golden("testanisotropic.msh"))


def _ids(raw, cells):
    return [int(c) + raw.nbface + 1 for c in cells]


def test_find_lines_fixture(aniso):
    want = [[int(t) for t in ln.split()]
            for ln in open(golden("testanisotropic-lines.txt")).read().splitlines() if ln.strip()]
    got = [_ids(aniso, ln) for ln in fvx.find_lines(aniso, 10.0)]
    assert got == want


@pytest.mark.parametrize("ordering,fixture", [
    ("line", "testanisotropic-lineordering.txt"),
    ("line_natural", "testanisotropic-lineordering.txt"),
    ("line_rcm", "testanisotropic-line_rcm_ordering.txt")])
def test_line_orderings_fixture(aniso, ordering, fixture):
    want = [int(t) for t in open(golden(fixture)).read().split()]
    assert _ids(aniso, fvx.line_permutation(aniso, 10.0, ordering)) == want


def test_no_lines_below_the_threshold(aniso):
    # an isotropic threshold above every cell's anisotropy finds no line; the line ordering is
    # then the identity
    assert fvx.find_lines(aniso, 1e6) == []
    assert np.array_equal(fvx.line_permutation(aniso, 1e6, "line"), np.arange(aniso.nelem))


def test_cell_rcm_on_a_strip_is_a_band_of_width_one():
    # a 1 x n strip of quads, numbered randomly: reverse Cuthill-McKee recovers the strip order
    # (or its reverse), so every interior face joins consecutive cells
    n = 40
    x = np.linspace(0, 1, n + 1)
    coords = np.array([[xi, y] for y in (0.0, 0.1) for xi in x])
    quads = np.array([[i, i + 1, n + 2 + i, n + 1 + i] for i in range(n)])
    shuffle = np.random.default_rng(3).permutation(n)
    inpoel = quads[shuffle]
    bface = np.array([[i, i + 1] for i in range(n)] + [[n + 2 + i, n + 1 + i] for i in range(n)]
                     + [[n + 1, 0], [n, 2 * n + 1]])
    raw = meshgen.RawMesh(coords=coords, inpoel=inpoel.astype(np.int32),
                          nnode=np.full(n, 4, np.int32), bface=bface.astype(np.int32),
                          btag=np.full(len(bface), 2, np.int32))
    perm = fvx.rcm_permutation(raw)
    assert sorted(perm) == list(range(n))
    pos = shuffle[perm]            # strip position of each new cell
    assert np.all(np.abs(np.diff(pos)) == 1)


def test_line_ordering_rejects_bad_arguments(aniso):
    import ctypes as C
    d, keep = fvx._desc(aniso)
    perm = np.zeros(aniso.nelem, np.int32)
    rc = fvx.lib().fvx_mesh_line_order(C.byref(d), C.c_double(10.0), 7,
                                       perm.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc != 0 and b"line-graph ordering" in fvx.lib().fvx_last_error()
    n = C.c_int()
    rc = fvx.lib().fvx_mesh_lines(C.byref(d), C.c_double(10.0), C.byref(n), None, None)
    assert rc != 0
    with pytest.raises(KeyError):
        fvx.line_permutation(aniso, 10.0, "line_amd")
