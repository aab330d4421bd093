"""Structured mesh generators for the workloads on the path.

Meshes are plain arrays in the form OGS-5 reads from ``#FEM_MSH``
(MSH/msh_mesh.cpp:454 $NODES, :475 $ELEMENTS; MSH/msh_elem.cpp:578): node
coordinates and an element -> node connectivity table whose node order is the
element's local order (quad: counter-clockwise; hex: bottom face
counter-clockwise then the top face; tet: the four vertices).  The equation
index of a node is its node index (MSH/msh_mesh.cpp:862-870), so the
natural x-fastest numbering below is also the row order of the CRS matrix.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

LINE2, QUAD4, HEX8, TET4 = 1, 2, 3, 5  # MshElemType::type numbering
NPE = {LINE2: 2, QUAD4: 4, HEX8: 8, TET4: 4}
EDIM = {LINE2: 1, QUAD4: 2, HEX8: 3, TET4: 3}


@dataclass
class Mesh:
    etype: int
    dim: int
    x: np.ndarray  # float64 [n_nodes]
    y: np.ndarray
    z: np.ndarray
    conn: np.ndarray  # int64 [n_elem, npe]
    shape: tuple  # cells per axis
    h: tuple  # cell size per axis
    cell_of_elem: np.ndarray  # int64 [n_elem] -> structured cell index

    @property
    def n_nodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_elem(self) -> int:
        return int(self.conn.shape[0])

    @property
    def npe(self) -> int:
        return NPE[self.etype]

    def node_index(self, *ijk) -> int:
        """Node number of structured node (i, j[, k])."""
        if self.dim == 1:
            return int(ijk[0])
        if self.dim == 2:
            i, j = ijk
            return int(j * (self.shape[0] + 1) + i)
        i, j, k = ijk
        nx, ny = self.shape[0] + 1, self.shape[1] + 1
        return int((k * ny + j) * nx + i)

    def nodes_where(self, axis: int, value: float, tol: float = 1e-9) -> np.ndarray:
        c = (self.x, self.y, self.z)[axis]
        return np.nonzero(np.abs(c - value) <= tol)[0].astype(np.int64)


def line_mesh(n: int, length: float = 1.0) -> Mesh:
    xs = np.linspace(0.0, length, n + 1)
    conn = np.stack([np.arange(n), np.arange(1, n + 1)], axis=1).astype(np.int64)
    zeros = np.zeros_like(xs)
    return Mesh(LINE2, 1, xs, zeros.copy(), zeros.copy(), conn, (n,), (length / n,),
                np.arange(n, dtype=np.int64))


def quad_mesh(nx: int, ny: int, lx: float = 1.0, ly: float = 1.0) -> Mesh:
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")  # [ny+1, nx+1], x fastest
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (j * (nx + 1) + i).ravel()
    conn = np.stack([n0, n0 + 1, n0 + 1 + (nx + 1), n0 + (nx + 1)], axis=1)
    return Mesh(QUAD4, 2, X.ravel().copy(), Y.ravel().copy(), np.zeros(X.size),
                conn.astype(np.int64), (nx, ny), (lx / nx, ly / ny),
                np.arange(nx * ny, dtype=np.int64))


def _grid_nodes(nx, ny, nz, lx, ly, lz):
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    zs = np.linspace(0.0, lz, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")  # x fastest
    return X.ravel().copy(), Y.ravel().copy(), Z.ravel().copy()


def hex_mesh(nx: int, ny: int, nz: int, lx: float = None, ly: float = None,
             lz: float = None) -> Mesh:
    lx = float(nx) if lx is None else lx
    ly = float(ny) if ly is None else ly
    lz = float(nz) if lz is None else lz
    X, Y, Z = _grid_nodes(nx, ny, nz, lx, ly, lz)
    sx, sy = 1, nx + 1
    sz = (nx + 1) * (ny + 1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (k * sz + j * sy + i).ravel()
    conn = np.stack([n0, n0 + sx, n0 + sx + sy, n0 + sy,
                     n0 + sz, n0 + sx + sz, n0 + sx + sy + sz, n0 + sy + sz], axis=1)
    return Mesh(HEX8, 3, X, Y, Z, conn.astype(np.int64), (nx, ny, nz),
                (lx / nx, ly / ny, lz / nz), np.arange(nx * ny * nz, dtype=np.int64))


# Kuhn/Freudenthal split of a cube into 6 tetrahedra along its main diagonal:
# one tet per axis permutation (a, b, c): v000 -> +e_a -> +e_a+e_b -> v111.
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def tet_mesh(nx: int, ny: int, nz: int, lx: float = None, ly: float = None,
             lz: float = None) -> Mesh:
    lx = float(nx) if lx is None else lx
    ly = float(ny) if ly is None else ly
    lz = float(nz) if lz is None else lz
    X, Y, Z = _grid_nodes(nx, ny, nz, lx, ly, lz)
    strides = np.array([1, nx + 1, (nx + 1) * (ny + 1)], dtype=np.int64)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (k * strides[2] + j * strides[1] + i).ravel()
    tets = []
    for a, b, c in _PERMS:
        v1 = n0 + strides[a]
        v2 = v1 + strides[b]
        v3 = v2 + strides[c]
        tets.append(np.stack([n0, v1, v2, v3], axis=1))
    # element order: cube-major, 6 tets per cube
    conn = np.stack(tets, axis=1).reshape(-1, 4)
    cells = np.repeat(np.arange(nx * ny * nz, dtype=np.int64), 6)
    return Mesh(TET4, 3, X, Y, Z, conn.astype(np.int64), (nx, ny, nz),
                (lx / nx, ly / ny, lz / nz), cells)


def element_centroids(mesh: Mesh) -> np.ndarray:
    c = mesh.conn
    return np.stack([mesh.x[c].mean(1), mesh.y[c].mean(1), mesh.z[c].mean(1)], axis=1)
