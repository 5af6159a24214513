"""Shared test inputs (seeded synthetic meshes and grids)."""
import os

import numpy as np

from stl2sdf import MeshGrid, synthetic
from stl2sdf.DataImport import import_stl

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def icosphere_case(subdiv=4, n_max=40, margin=2):
    """Config 1: icosphere r=1 subdivided 4x, float32 vertices, cell 0.05."""
    X, IEN = synthetic.icosphere(1.0, subdiv)
    X = synthetic.to_float32(X)
    mesh = MeshGrid.TriangularMesh(X, IEN)
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    return mesh, MeshGrid.Grid(xmin, xmax, n_max, margin)


def artifacts_case(cell=0.5):
    X, IEN = import_stl(os.path.join(GOLDEN, "artifacts.stl"))
    mesh = MeshGrid.TriangularMesh(X, IEN)
    return mesh, MeshGrid.noninteractive_sdf_grid_setup(mesh, cell)


def star_case(n=60, n_max=45, seed=0):
    X, IEN, _ = synthetic.star_surface(n, n, seed)
    X = synthetic.to_float32(X)
    mesh = MeshGrid.TriangularMesh(X, IEN)
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    return mesh, MeshGrid.Grid(xmin, xmax, n_max, 3)


def open_case():
    """Icosphere with a band of triangles removed: rays disagree, the
    adaptive vote and the winding fallback both run."""
    X, IEN = synthetic.icosphere(1.0, 3)
    X = synthetic.to_float32(X)
    c = X[:, IEN - 1].mean(axis=1)          # 3 x nel centroids
    keep = ~((np.abs(c[2]) < 0.15) & (c[0] > 0))
    mesh = MeshGrid.TriangularMesh(X, np.asfortranarray(IEN[:, keep]))
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    return mesh, MeshGrid.Grid(xmin, xmax, 24, 2)


def diag(grid_or_mesh):
    X = grid_or_mesh.X
    return float(np.linalg.norm(X.max(1) - X.min(1)))
