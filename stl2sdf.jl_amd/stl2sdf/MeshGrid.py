"""Grid and triangle-mesh containers (host side), mirroring
src/MeshGrid/Grid.jl, Grid_setup.jl and MeshInformations.jl.

Arrays keep the reference's layouts: X is 3 x nnp float64 and IEN 3 x nel
int64 with 1-based indices, both column-major, so they can be handed to the
C ABI without copies.  Grid arithmetic is plain IEEE double in the
reference's operation order.
"""
import math

import numpy as np


class TriangularMesh:
    """TriangularMesh(X, IEN) -- MeshInformations.jl:57-84.

    X: 3 x nnp (or a list of 3-vectors); IEN: 3 x nel, 1-based.  The
    node-to-element lists (INE) are only used by the reference's 3-nearest
    vertex fallback and are built on demand."""

    def __init__(self, X, IEN):
        X = np.asarray(X, dtype=np.float64)
        IEN = np.asarray(IEN, dtype=np.int64)
        if X.ndim == 2 and X.shape[0] != 3 and X.shape[1] == 3:
            X = X.T
        if IEN.ndim == 2 and IEN.shape[0] != 3 and IEN.shape[1] == 3:
            IEN = IEN.T
        if X.ndim != 2 or X.shape[0] != 3 or IEN.ndim != 2 or IEN.shape[0] != 3:
            raise ValueError("X must be 3 x nnp and IEN 3 x nel")
        self.X = np.asfortranarray(X)
        self.IEN = np.asfortranarray(IEN)
        self.nsd = 3
        self.nnp = self.X.shape[1]
        self.nen = 3
        self.nel = self.IEN.shape[1]
        self._INE = None

    @property
    def INE(self):
        if self._INE is None:
            ine = [[] for _ in range(self.nnp)]
            for el in range(self.nel):
                for c in range(3):
                    ine[self.IEN[c, el] - 1].append(el + 1)
            self._INE = ine
        return self._INE


class Grid:
    """Grid(AABB_min, AABB_max, N_max, margineCells=3) -- Grid.jl:9-33."""

    def __init__(self, AABB_min, AABB_max, N_max, margineCells=3, _raw=None):
        if _raw is not None:
            self.AABB_min, self.AABB_max, self.N, self.cell_size, self.ngp = _raw
            return
        lo = [float(v) for v in AABB_min]
        hi = [float(v) for v in AABB_max]
        cell = max(hi[d] - lo[d] for d in range(3)) / N_max
        mh = margineCells * cell
        lo = [lo[d] - mh for d in range(3)]
        hi = [hi[d] + mh for d in range(3)]
        N = [int(math.ceil((hi[d] - lo[d]) / cell)) for d in range(3)]
        hi = [lo[d] + N[d] * cell for d in range(3)]
        self.AABB_min = np.array(lo)
        self.AABB_max = np.array(hi)
        self.N = np.array(N, dtype=np.int64)
        self.cell_size = cell
        self.ngp = int((N[0] + 1) * (N[1] + 1) * (N[2] + 1))

    def __repr__(self):
        return (f"Grid(AABB_min={self.AABB_min.tolist()}, AABB_max={self.AABB_max.tolist()}, "
                f"N={self.N.tolist()}, cell_size={self.cell_size!r}, ngp={self.ngp})")


def getMesh_AABB(X):
    """getMesh_AABB -- Grid.jl:74-78."""
    X = np.asarray(X)
    return X.min(axis=1).astype(np.float64), X.max(axis=1).astype(np.float64)


def generateGridPoints(grid):
    """generateGridPoints -- Grid.jl:82-96: 3 x ngp, i fastest, column-major."""
    nx, ny, nz = (int(n) + 1 for n in grid.N)
    h = grid.cell_size
    xs = grid.AABB_min[0] + h * np.arange(nx, dtype=np.float64)
    ys = grid.AABB_min[1] + h * np.arange(ny, dtype=np.float64)
    zs = grid.AABB_min[2] + h * np.arange(nz, dtype=np.float64)
    P = np.empty((3, nx * ny * nz), dtype=np.float64, order="F")
    P[0] = np.tile(xs, ny * nz)
    P[1] = np.tile(np.repeat(ys, nx), nz)
    P[2] = np.repeat(zs, nx * ny)
    return P


def noninteractive_sdf_grid_setup(mesh, B):
    """noninteractive_sdf_grid_setup -- Grid_setup.jl:104-124 (margin 3)."""
    X_min, X_max = getMesh_AABB(mesh.X)
    N_new = int(math.floor(max(X_max[d] - X_min[d] for d in range(3)) / B))
    return Grid(X_min, X_max, N_new, 3)
