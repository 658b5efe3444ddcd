"""Deterministic generators for the benchmark and parity meshes.

Each generator returns a :class:`RawMesh` -- node coordinates, CCW cell connectivity
(triangles padded with -1 to 4 columns) and tagged boundary edges -- which is exactly the
information the reference's Gmsh 2.2 reader produces (src/mesh/meshreaders.cpp:66-266).
Markers follow the reference control files:

* Gaussian-bump channel: walls 2, inlet 3, outlet 4 (tests/inv-gaussianbump/base.ctrl,
  tests/inv-gaussianbump/gaussian_channel.geo).
* Airfoil O-meshes: wall 2, far field 4 (testcases/naca0012/transonic-sanity-test-muscl.ctrl,
  testcases/visc-naca0012/laminar-implicit.ctrl).
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class RawMesh:
    coords: np.ndarray      # (npoin, 2) float64
    inpoel: np.ndarray      # (nelem, 4) int32, -1 padded for triangles, CCW
    nnode: np.ndarray       # (nelem,) int32, 3 or 4
    bface: np.ndarray       # (nbface, 2) int32 node pairs
    btag: np.ndarray        # (nbface,) int32 physical markers
    name: str = ""

    @property
    def nelem(self) -> int:
        return int(self.inpoel.shape[0])

    @property
    def npoin(self) -> int:
        return int(self.coords.shape[0])

    @property
    def nbface(self) -> int:
        return int(self.bface.shape[0])

    def signed_areas(self) -> np.ndarray:
        x = self.coords[:, 0]
        y = self.coords[:, 1]
        a = np.zeros(self.nelem)
        for k in range(1, 3):
            i0 = self.inpoel[:, 0]
            i1 = self.inpoel[:, k]
            i2 = self.inpoel[:, k + 1]
            valid = (i2 >= 0) if k == 2 else np.ones(self.nelem, bool)
            i2c = np.where(i2 >= 0, i2, i0)
            t = 0.5 * ((x[i1] - x[i0]) * (y[i2c] - y[i0]) - (x[i2c] - x[i0]) * (y[i1] - y[i0]))
            a += np.where(valid, t, 0.0)
        return a

    def permuted(self, perm: np.ndarray) -> "RawMesh":
        """Cells renumbered so that new cell i is old cell perm[i]."""
        perm = np.asarray(perm, dtype=np.int64)
        return RawMesh(self.coords, self.inpoel[perm].copy(), self.nnode[perm].copy(),
                       self.bface, self.btag, self.name)


def _cells_from_grid(nid: np.ndarray, split_from_row: int, split_rows_below: bool):
    """Quads of a structured node grid nid[i, k] (i periodic or not handled by caller).

    Rows k with (k < split_from_row if split_rows_below else k >= split_from_row) are split
    into two triangles along the (i,k)-(i+1,k+1) diagonal.
    """
    ni, nk = nid.shape[0] - 1, nid.shape[1] - 1
    quads = []
    tris = []
    for k in range(nk):
        split = (k < split_from_row) if split_rows_below else (k >= split_from_row)
        a = nid[:-1, k]
        b = nid[1:, k]
        c = nid[1:, k + 1]
        d = nid[:-1, k + 1]
        if split:
            t = np.empty((2 * ni, 4), np.int32)
            t[0::2] = np.stack([a, b, c, -np.ones_like(a)], axis=1)
            t[1::2] = np.stack([a, c, d, -np.ones_like(a)], axis=1)
            tris.append((k, t))
        else:
            quads.append((k, np.stack([a, b, c, d], axis=1).astype(np.int32)))
    rows = sorted(quads + tris, key=lambda kv: kv[0])
    inpoel = np.concatenate([r[1] for r in rows], axis=0).astype(np.int32)
    nnode = np.where(inpoel[:, 3] >= 0, 4, 3).astype(np.int32)
    return inpoel, nnode


def _orient_ccw(m: RawMesh) -> RawMesh:
    a = m.signed_areas()
    neg = a < 0
    if np.any(neg):
        ip = m.inpoel.copy()
        q = neg & (m.nnode == 4)
        t = neg & (m.nnode == 3)
        ip[q] = ip[q][:, ::-1]
        ip[t, :3] = ip[t, :3][:, ::-1]
        m = RawMesh(m.coords, ip, m.nnode, m.bface, m.btag, m.name)
    if np.any(m.signed_areas() <= 0):
        raise ValueError("generated mesh has degenerate cells")
    return m


def gaussian_bump(nx: int = 96, ny: int = 32, nsplit: int = 16) -> RawMesh:
    """Channel x in [-1.5,1.5] between y = 0.0625 exp(-25 x^2) and y = 0.8.

    nx*ny quads; the lowest ``nsplit`` rows are split into two triangles each.
    """
    xs = np.linspace(-1.5, 1.5, nx + 1)
    ylow = 0.0625 * np.exp(-25.0 * xs ** 2)
    t = np.arange(ny + 1) / ny
    X = np.repeat(xs[:, None], ny + 1, axis=1)
    Y = ylow[:, None] + (0.8 - ylow[:, None]) * t[None, :]
    nid = np.arange((nx + 1) * (ny + 1), dtype=np.int32).reshape(nx + 1, ny + 1)
    coords = np.stack([X.ravel(), Y.ravel()], axis=1)
    inpoel, nnode = _cells_from_grid(nid, nsplit, True)
    bf, bt = [], []
    for i in range(nx):          # lower wall
        bf.append((nid[i, 0], nid[i + 1, 0])); bt.append(2)
    for i in range(nx):          # upper wall
        bf.append((nid[i + 1, ny], nid[i, ny])); bt.append(2)
    for k in range(ny):          # inlet (x = -1.5)
        bf.append((nid[0, k + 1], nid[0, k])); bt.append(3)
    for k in range(ny):          # outlet (x = +1.5)
        bf.append((nid[nx, k], nid[nx, k + 1])); bt.append(4)
    m = RawMesh(coords, inpoel, nnode, np.array(bf, np.int32), np.array(bt, np.int32),
                f"bump{nx}x{ny}")
    return _orient_ccw(m)


def gaussian_channel(nx: int, ny: int, xend: float = 1.0, height: float = 0.75, a: float = 0.02,
                     b: float = 100.0) -> RawMesh:
    """All-triangle channel x in [-xend, xend] between y = a exp(-b x^2) and y = height: the
    geometry and markers of the reference's entropy-convergence case
    (tests/inv-gaussianbump/gaussian_channel.geo: walls 2, inlet 3, outlet 4), meshed here as
    nx*ny quads each split into two triangles instead of Gmsh's unstructured triangles."""
    xs = np.linspace(-xend, xend, nx + 1)
    ylow = a * np.exp(-b * xs ** 2)
    t = np.arange(ny + 1) / ny
    X = np.repeat(xs[:, None], ny + 1, axis=1)
    Y = ylow[:, None] + (height - ylow[:, None]) * t[None, :]
    nid = np.arange((nx + 1) * (ny + 1), dtype=np.int32).reshape(nx + 1, ny + 1)
    coords = np.stack([X.ravel(), Y.ravel()], axis=1)
    inpoel, nnode = _cells_from_grid(nid, ny, True)
    bf, bt = [], []
    for i in range(nx):
        bf.append((nid[i, 0], nid[i + 1, 0])); bt.append(2)
    for i in range(nx):
        bf.append((nid[i + 1, ny], nid[i, ny])); bt.append(2)
    for k in range(ny):
        bf.append((nid[0, k + 1], nid[0, k])); bt.append(3)
    for k in range(ny):
        bf.append((nid[nx, k], nid[nx, k + 1])); bt.append(4)
    m = RawMesh(coords, inpoel, nnode, np.array(bf, np.int32), np.array(bt, np.int32),
                f"channel{nx}x{ny}")
    return _orient_ccw(m)


def naca4_thickness(x: np.ndarray, t: float = 0.12) -> np.ndarray:
    return 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3
                      - 0.1036 * x ** 4)


def _geometric_ratio(first: float, total: float, n: int) -> float:
    lo, hi = 1.0 + 1e-12, 2.0
    f = lambda q: first * (q ** n - 1.0) / (q - 1.0) - total
    while f(hi) < 0:
        hi *= 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def airfoil_omesh(ni: int = 512, nk: int = 128, nsplit_outer: int = 64,
                  first_spacing: float = 1e-3, rfar: float = 50.0,
                  thickness: float = 0.12) -> RawMesh:
    """O-mesh around a symmetric NACA 4-digit airfoil of unit chord.

    ``ni`` cells around, ``nk`` cells in the wall-normal direction with geometric stretching
    from ``first_spacing`` to a far-field circle of radius ``rfar`` centred at mid-chord; the
    outermost ``nsplit_outer`` rings are split into triangles.
    """
    th = 2.0 * np.pi * np.arange(ni) / ni
    xw = 0.5 * (1.0 + np.cos(th))
    yt = naca4_thickness(xw, thickness)
    yw = np.where(th <= np.pi, yt, -yt)
    yw[0] = 0.0
    W = np.stack([xw, yw], axis=1)
    # outward wall normals from the centred tangent (wall runs counter-clockwise)
    tang = np.roll(W, -1, axis=0) - np.roll(W, 1, axis=0)
    nrm = np.stack([tang[:, 1], -tang[:, 0]], axis=1)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    F = np.stack([0.5 + rfar * np.cos(th), rfar * np.sin(th)], axis=1)
    E = F - W
    dist = np.linalg.norm(E, axis=1)
    E /= dist[:, None]
    D = rfar - 0.5
    q = _geometric_ratio(first_spacing, D, nk)
    s = first_spacing * (q ** np.arange(nk + 1) - 1.0) / (q - 1.0)
    s[-1] = D
    frac = s / D
    blend = np.clip(frac / 0.02, 0.0, 1.0) ** 0.5
    P = np.empty((ni, nk + 1, 2))
    for k in range(nk + 1):
        dvec = (1.0 - blend[k]) * nrm + blend[k] * E
        dvec /= np.linalg.norm(dvec, axis=1)[:, None]
        P[:, k, :] = W + (frac[k] * dist)[:, None] * dvec
    P[:, nk, :] = F
    coords = P.reshape(-1, 2)
    nid = np.arange(ni * (nk + 1), dtype=np.int32).reshape(ni, nk + 1)
    nidp = np.concatenate([nid, nid[:1]], axis=0)      # periodic in i
    inpoel, nnode = _cells_from_grid(nidp, nk - nsplit_outer, False)
    bf, bt = [], []
    for i in range(ni):
        bf.append((nidp[i + 1, 0], nidp[i, 0])); bt.append(2)
    for i in range(ni):
        bf.append((nidp[i, nk], nidp[i + 1, nk])); bt.append(4)
    m = RawMesh(coords, inpoel, nnode, np.array(bf, np.int32), np.array(bt, np.int32),
                f"naca_o{ni}x{nk}s{nsplit_outer}")
    return _orient_ccw(m)


# The benchmark / parity configurations.
def mesh_k1() -> RawMesh:
    """Gaussian bump, 96x32 quads with the lower 16 rows split: 4,608 cells."""
    return gaussian_bump(96, 32, 16)


def mesh_k2() -> RawMesh:
    """Inviscid airfoil, 512x128 O-mesh, outer 64 rings split: 98,304 cells."""
    return airfoil_omesh(512, 128, 64, 1e-3, 50.0)


def mesh_k2l() -> RawMesh:
    """Inviscid airfoil refined 4x per direction, 2048x512, outer 256 rings split: 1,572,864 cells."""
    return airfoil_omesh(2048, 512, 256, 2.5e-4, 50.0)


def mesh_k3() -> RawMesh:
    """Viscous airfoil, 1024x256 all-quad O-mesh, wall spacing 2e-5: 262,144 cells."""
    return airfoil_omesh(1024, 256, 0, 2e-5, 50.0)
