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


# ------------------------------------------------------------------ Hilbert order -----------

def _grid(n):
    """n x n unit squares, cells numbered row by row (a scrambled start for the ordering)."""
    x = np.arange(n + 1, dtype=float)
    coords = np.array([[xi, yi] for yi in x for xi in x])
    nid = lambda i, j: j * (n + 1) + i  # noqa: E731
    quads = np.array([[nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)]
                      for j in range(n) for i in range(n)], np.int32)
    bface = []
    for i in range(n):
        bface += [[nid(i, 0), nid(i + 1, 0)], [nid(i + 1, n), nid(i, n)],
                  [nid(0, i + 1), nid(0, i)], [nid(n, i), nid(n, i + 1)]]
    return meshgen.RawMesh(coords=coords, inpoel=quads, nnode=np.full(n * n, 4, np.int32),
                           bface=np.array(bface, np.int32),
                           btag=np.full(len(bface), 2, np.int32))


@pytest.mark.parametrize("n", [4, 16])
def test_hilbert_order_walks_a_grid_one_neighbour_at_a_time(n):
    # on a 2^k x 2^k grid the Hilbert curve visits every cell once and each step moves to an
    # edge neighbour; it starts in a corner cell
    raw = _grid(n)
    perm = fvx.hilbert_permutation(raw)
    assert sorted(perm) == list(range(n * n))
    ij = np.stack([perm % n, perm // n], 1)
    steps = np.abs(np.diff(ij, axis=0)).sum(1)
    assert np.all(steps == 1)
    assert tuple(ij[0]) in {(0, 0), (0, n - 1), (n - 1, 0), (n - 1, n - 1)}


def test_hilbert_order_makes_compact_blocks():
    # on the 98,304-cell airfoil O-mesh most interior faces join two cells of the same 64-cell
    # block of consecutive cells (the Hilbert curve's compact patches: 76%), against 14% in RCM
    # order, whose blocks are thin strips across the rings
    import pyoracle as po
    raw = meshgen.mesh_k2()
    frac = {}
    for name, perm in (("hilbert", fvx.hilbert_permutation(raw)), ("rcm", fvx.rcm_permutation(raw))):
        m = po.Mesh(raw.permuted(perm))
        f = m.intfac[raw.nbface:, :2]
        frac[name] = np.mean(f[:, 0] // 64 == f[:, 1] // 64)
    assert frac["hilbert"] >= 0.7 and frac["rcm"] < 0.3, frac
