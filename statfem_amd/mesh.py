"""Meshes, DOF numbering, Dirichlet sets and sensor projection rows.

Host-side data preparation for the GPU path. Conventions follow the reference:

* reference cells are [0,1]^d with the node orders of corlib/Shapefun.ipp
  (LINE2 :39-55, TRI3 :181-197, QUAD4 :460-476, HEX8 :926-975);
* nodes keep file/generation order and DOFs follow node order
  (del2/fem/NodePotential.hpp:65); vector DOFs are interleaved 3n+c
  (solid/fem/NodeStatic.ipp:15-20) and constrained DOFs stay in the system;
* boundary detection uses corlib::fuzzyEqual with tol 1e-8
  (corlib/fuzzyEqual.hpp:27-31), as in the laplaceRHSinfer driver
  (del2/apps/reference/laplaceRHSinfer/fem.hpp:36-44).
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

LINE2, TRI3, QUAD4, HEX8 = 1, 2, 3, 4
_DIM = {LINE2: 1, TRI3: 2, QUAD4: 2, HEX8: 3}
_NEN = {LINE2: 2, TRI3: 3, QUAD4: 4, HEX8: 8}


@dataclass
class Mesh:
    kind: int
    coords: np.ndarray            # (nnodes, dim) float64
    conn: np.ndarray              # (nel, nen) int32
    grid: Optional[tuple] = None  # per-axis node coordinates of a tensor grid (x fastest)
    meta: dict = field(default_factory=dict)

    @property
    def dim(self):
        return _DIM[self.kind]

    @property
    def nen(self):
        return _NEN[self.kind]

    @property
    def nnodes(self):
        return self.coords.shape[0]

    @property
    def nel(self):
        return self.conn.shape[0]

    def element_dofs(self, dofpn=1):
        """Element DOF lists in local node order, node-major / component-minor
        (del2/fem/ElementPotential.ipp:25-42, solid/fem/ElementStatic.ipp:20-36)."""
        if dofpn == 1:
            return np.ascontiguousarray(self.conn, dtype=np.int32)
        e = (self.conn[:, :, None] * dofpn + np.arange(dofpn)[None, None, :]).reshape(self.nel, -1)
        return np.ascontiguousarray(e, dtype=np.int32)


def line1d(n, length=1.0):
    x = np.linspace(0.0, length, n + 1)
    conn = np.stack([np.arange(n), np.arange(1, n + 1)], axis=1).astype(np.int32)
    return Mesh(LINE2, x[:, None].copy(), conn, grid=(x,))


def grid2d(nx, ny=None, kind=QUAD4):
    """Unit square, (nx+1)(ny+1) nodes numbered x-fastest. QUAD4 cells are CCW
    (0,0),(1,0),(1,1),(0,1); TRI3 splits each cell along its (0,0)-(1,1) diagonal
    into (00,10,11) and (00,11,01), both counter-clockwise."""
    ny = nx if ny is None else ny
    x = np.linspace(0.0, 1.0, nx + 1)
    y = np.linspace(0.0, 1.0, ny + 1)
    X, Y = np.meshgrid(x, y)   # (ny+1, nx+1), x fastest
    coords = np.stack([X.ravel(), Y.ravel()], axis=1)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n00 = (j * (nx + 1) + i).ravel()
    n10, n11, n01 = n00 + 1, n00 + nx + 2, n00 + nx + 1
    if kind == QUAD4:
        conn = np.stack([n00, n10, n11, n01], axis=1)
    elif kind == TRI3:
        t1 = np.stack([n00, n10, n11], axis=1)
        t2 = np.stack([n00, n11, n01], axis=1)
        conn = np.stack([t1, t2], axis=1).reshape(-1, 3)
    else:
        raise ValueError("grid2d supports QUAD4 and TRI3")
    return Mesh(kind, coords, conn.astype(np.int32), grid=(x, y))


def grid3d(nx, ny=None, nz=None):
    """Unit cube of HEX8 cells, nodes numbered x fastest then y then z."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    x = np.linspace(0.0, 1.0, nx + 1)
    y = np.linspace(0.0, 1.0, ny + 1)
    z = np.linspace(0.0, 1.0, nz + 1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    sx, sy = 1, nx + 1
    sz = (nx + 1) * (ny + 1)
    n0 = (k * sz + j * sy + i).ravel()
    conn = np.stack([n0, n0 + sx, n0 + sx + sy, n0 + sy,
                     n0 + sz, n0 + sz + sx, n0 + sz + sx + sy, n0 + sz + sy], axis=1)
    return Mesh(HEX8, coords, conn.astype(np.int32), grid=(x, y, z))


def read_smf(path):
    """SMF reader (corlib/Mesh.ipp:62-102, corlib/SmfHead.ipp): '!' header lines
    give the element shape, then 'nNodes nElements', node lines with three
    coordinates, then connectivity lines. Nodes keep file order."""
    shape = None
    with open(path) as fh:
        lines = [ln.strip() for ln in fh if ln.strip()]
    body = []
    for ln in lines:
        if ln.startswith("!"):
            parts = ln[1:].split()
            if len(parts) >= 2 and parts[0] == "elementShape":
                shape = parts[1]
        else:
            body.append(ln)
    nn, ne = (int(t) for t in body[0].split()[:2])
    xyz = np.array([[float(t) for t in body[1 + i].split()[:3]] for i in range(nn)])
    conn = np.array([[int(t) for t in body[1 + nn + e].split()] for e in range(ne)], dtype=np.int32)
    kind = {"line": LINE2, "triangle": TRI3, "quadrilateral": QUAD4, "hexahedron": HEX8}[shape]
    d = _DIM[kind]
    return Mesh(kind, np.ascontiguousarray(xyz[:, :d]), conn)


def fuzzy_boundary(coords, axes, lo=0.0, hi=1.0, tol=1e-8, sides=("lo", "hi")):
    """Node ids with any listed coordinate fuzzy-equal to lo or hi."""
    m = np.zeros(coords.shape[0], dtype=bool)
    for a in axes:
        if "lo" in sides:
            m |= np.abs(coords[:, a] - lo) < tol
        if "hi" in sides:
            m |= np.abs(coords[:, a] - hi) < tol
    return np.nonzero(m)[0].astype(np.int32)


def dirichlet_all_boundary(mesh):
    return fuzzy_boundary(mesh.coords, range(mesh.dim))


def sensor_rows(mesh, pts):
    """Interpolation rows of the observation operator P (ns x N): each sensor
    gets the shape-function values of the element that contains it (barycentric
    for TRI3, bilinear for QUAD4, trilinear for HEX8, linear for LINE2).
    Tensor-grid meshes only. Returns CSR-like (cols, vals) of shape (ns, nen)."""
    assert mesh.grid is not None, "sensor_rows needs a structured mesh"
    pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
    ns, d = pts.shape
    cells, loc = [], []
    for a in range(d):
        g = mesh.grid[a]
        c = np.clip(np.searchsorted(g, pts[:, a], side="right") - 1, 0, len(g) - 2)
        cells.append(c)
        loc.append((pts[:, a] - g[c]) / (g[c + 1] - g[c]))
    if mesh.kind == LINE2:
        n0 = cells[0]
        cols = np.stack([n0, n0 + 1], 1)
        vals = np.stack([1 - loc[0], loc[0]], 1)
    elif mesh.kind in (QUAD4, TRI3):
        nx = len(mesh.grid[0]) - 1
        n00 = cells[1] * (nx + 1) + cells[0]
        a, b = loc
        if mesh.kind == QUAD4:
            cols = np.stack([n00, n00 + 1, n00 + nx + 2, n00 + nx + 1], 1)
            vals = np.stack([(1 - a) * (1 - b), a * (1 - b), a * b, (1 - a) * b], 1)
        else:
            lower = a >= b   # triangle (00,10,11) holds points below the diagonal
            cols = np.where(lower[:, None], np.stack([n00, n00 + 1, n00 + nx + 2], 1),
                            np.stack([n00, n00 + nx + 2, n00 + nx + 1], 1))
            vals = np.where(lower[:, None], np.stack([1 - a, a - b, b], 1),
                            np.stack([1 - b, a, b - a], 1))
    else:
        nx, ny = len(mesh.grid[0]) - 1, len(mesh.grid[1]) - 1
        sy, sz = nx + 1, (nx + 1) * (ny + 1)
        n0 = cells[2] * sz + cells[1] * sy + cells[0]
        a, b, c = loc
        cols = np.stack([n0, n0 + 1, n0 + 1 + sy, n0 + sy,
                         n0 + sz, n0 + sz + 1, n0 + sz + 1 + sy, n0 + sz + sy], 1)
        vals = np.stack([(1 - a) * (1 - b) * (1 - c), a * (1 - b) * (1 - c), a * b * (1 - c),
                         (1 - a) * b * (1 - c), (1 - a) * (1 - b) * c, a * (1 - b) * c,
                         a * b * c, (1 - a) * b * c], 1)
    return cols.astype(np.int32), vals.astype(np.float64)


def dense_projection(ncols, cols, vals, dofpn=1, comp=0):
    """Dense ns x (ncols*dofpn) P from sensor rows (component `comp` for vector DOFs)."""
    ns = cols.shape[0]
    P = np.zeros((ns, ncols * dofpn))
    for s in range(ns):
        np.add.at(P[s], cols[s] * dofpn + comp, vals[s])
    return P


def union(meshes):
    """Disjoint union of meshes of one element kind: nodes and elements of mesh q
    follow those of meshes 0..q-1, so the union's operator is block diagonal.
    Returns the union mesh and the node offsets (len(meshes) + 1)."""
    kinds = {m.kind for m in meshes}
    assert len(kinds) == 1, "union needs one element kind"
    off = np.zeros(len(meshes) + 1, dtype=np.int64)
    off[1:] = np.cumsum([m.nnodes for m in meshes])
    coords = np.concatenate([m.coords for m in meshes], axis=0)
    conn = np.concatenate([m.conn + off[q] for q, m in enumerate(meshes)], axis=0).astype(np.int32)
    return Mesh(kinds.pop(), coords, conn), off
