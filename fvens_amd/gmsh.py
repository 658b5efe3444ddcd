"""Gmsh 2.2 ASCII mesh reader and writer for linear triangle/quad meshes.

Follows the reference reader (src/mesh/meshreaders.cpp:66-266): node coordinates (z dropped),
linear edges (type 1) are boundary faces whose first tag is the physical marker, triangles
(type 2) and quads (type 3) are cells; node numbers in the file are 1-based.
"""
from __future__ import annotations

import numpy as np

from .meshgen import RawMesh


def read_gmsh2(path: str) -> RawMesh:
    with open(path, "r") as fh:
        lines = fh.read().split("\n")
    it = iter(range(len(lines)))
    i = 0
    coords = None
    bf, bt, cells, nn = [], [], [], []
    while i < len(lines):
        s = lines[i].strip()
        if s == "$Nodes":
            npoin = int(lines[i + 1])
            arr = np.array([lines[i + 2 + k].split() for k in range(npoin)], dtype=np.float64)
            ids = arr[:, 0].astype(np.int64)
            if not np.array_equal(ids, np.arange(1, npoin + 1)):
                raise ValueError("node numbering must be contiguous from 1")
            coords = arr[:, 1:3].copy()
            i += 2 + npoin
            continue
        if s == "$Elements":
            nelm = int(lines[i + 1])
            for k in range(nelm):
                tok = lines[i + 2 + k].split()
                etype = int(tok[1])
                ntags = int(tok[2])
                tags = [int(t) for t in tok[3:3 + ntags]]
                nodes = [int(t) - 1 for t in tok[3 + ntags:]]
                if etype == 1:
                    bf.append(nodes[:2])
                    bt.append(tags[0] if tags else 0)
                elif etype == 2:
                    cells.append(nodes[:3] + [-1]); nn.append(3)
                elif etype == 3:
                    cells.append(nodes[:4]); nn.append(4)
                elif etype == 15:
                    continue            # points carry no cells
                else:
                    raise ValueError(f"unsupported Gmsh element type {etype}")
            i += 2 + nelm
            continue
        i += 1
    if coords is None:
        raise ValueError("no $Nodes section")
    return RawMesh(coords, np.array(cells, np.int32).reshape(-1, 4), np.array(nn, np.int32),
                   np.array(bf, np.int32).reshape(-1, 2), np.array(bt, np.int32), path)


def write_gmsh2(m: RawMesh, path: str, domain_tag: int = 1) -> None:
    with open(path, "w") as fh:
        fh.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n")
        fh.write(f"{m.npoin}\n")
        for k, (x, y) in enumerate(m.coords):
            fh.write(f"{k + 1} {x:.17g} {y:.17g} 0\n")
        fh.write("$EndNodes\n$Elements\n")
        fh.write(f"{m.nbface + m.nelem}\n")
        e = 1
        for (a, b), t in zip(m.bface, m.btag):
            fh.write(f"{e} 1 2 {t} {t} {a + 1} {b + 1}\n"); e += 1
        for row, n in zip(m.inpoel, m.nnode):
            typ = 2 if n == 3 else 3
            nodes = " ".join(str(v + 1) for v in row[:n])
            fh.write(f"{e} {typ} 2 {domain_tag} {domain_tag} {nodes}\n"); e += 1
        fh.write("$EndElements\n")
