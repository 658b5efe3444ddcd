"""SU2 ASCII mesh reader for linear triangle / quad meshes.

Follows the reference reader (src/mesh/meshreaders.cpp:269-397): `NDIME=`, then `NELEM=` cells
as VTK type (5 triangle, 9 quad), 0-based node numbers and a trailing index; `NPOIN=` nodes
(x, y, index); `NMARK=` markers, each `MARKER_TAG=` (an integer physical marker) and
`MARKER_ELEMS=` line elements (type 3) with two node numbers. Boundary faces are kept in
marker order.
"""
from __future__ import annotations

import numpy as np

from .meshgen import RawMesh


def _value(line: str, key: str) -> str:
    k, _, v = line.partition("=")
    if k.strip() != key:
        raise ValueError(f"expected {key}= but read {line!r}")
    return v.strip()


def read_su2(path: str) -> RawMesh:
    with open(path, "r") as fh:
        lines = [ln for ln in fh.read().split("\n")]
    pos = 0

    def nxt():
        nonlocal pos
        while pos < len(lines) and (not lines[pos].strip() or lines[pos].lstrip().startswith("%")):
            pos += 1
        if pos >= len(lines):
            raise ValueError("unexpected end of SU2 file")
        ln = lines[pos]
        pos += 1
        return ln

    if int(_value(nxt(), "NDIME")) != 2:
        raise ValueError("only two-dimensional SU2 meshes are supported")
    nelem = int(_value(nxt(), "NELEM"))
    inpoel = np.full((nelem, 4), -1, np.int32)
    nnode = np.zeros(nelem, np.int32)
    for i in range(nelem):
        tok = nxt().split()
        t = int(tok[0])
        if t == 5:
            nnode[i] = 3
        elif t == 9:
            nnode[i] = 4
        else:
            raise ValueError(f"unsupported SU2 element type {t}")
        inpoel[i, :nnode[i]] = [int(v) for v in tok[1:1 + nnode[i]]]
    npoin = int(_value(nxt(), "NPOIN").split()[0])
    coords = np.array([nxt().split()[:2] for _ in range(npoin)], dtype=np.float64)
    nmark = int(_value(nxt(), "NMARK"))
    bf, bt = [], []
    for _ in range(nmark):
        tag = int(_value(nxt(), "MARKER_TAG"))
        nf = int(_value(nxt(), "MARKER_ELEMS"))
        for _ in range(nf):
            tok = nxt().split()
            if int(tok[0]) != 3:
                raise ValueError("SU2 boundary elements must be lines (type 3)")
            bf.append([int(tok[1]), int(tok[2])])
            bt.append(tag)
    return RawMesh(coords, inpoel, nnode, np.array(bf, np.int32).reshape(-1, 2),
                   np.array(bt, np.int32), path)


def read_mesh(path: str) -> RawMesh:
    """SU2 when the file name ends in ".su2", Gmsh 2.2 otherwise, as the reference's readMesh
    (src/mesh/meshreaders.cpp:35-61)."""
    from .gmsh import read_gmsh2
    if path.rsplit(".", 1)[-1] == "su2":
        return read_su2(path)
    return read_gmsh2(path)
