"""Closest points (the xp matrix of evalDistancesOnTriMesh) on the GPU.

xp is the exact FP64 closest point on the mesh, the smallest (squared
distance, triangle index) over all triangles, so it must equal the oracle's
brute-force loop bit for bit, at every distance; dist is unchanged by asking
for xp.
"""
import numpy as np
import pytest

from oracle import oracle
from stl2sdf import MeshGrid, SignedDistances, synthetic
from stl2sdf.DataExport import SelectProjectedNodes
from tests import _cases

pytestmark = pytest.mark.gpu


def _check(mesh, grid, points=None):
    P = MeshGrid.generateGridPoints(grid) if points is None else points
    print(f"case: nel={mesh.nel} npts={P.shape[1]} lattice={points is None}", flush=True)
    d, xp = SignedDistances.evalDistancesOnTriMesh(mesh, grid, points)
    d0, none = SignedDistances.evalDistancesOnTriMesh(mesh, grid, points, want_xp=False)
    assert none is None and np.array_equal(d, d0)
    _, xp_ref = oracle.dist_exact(mesh.X, mesh.IEN, P, grid.cell_size, want_xp=True)
    bad = np.flatnonzero(np.any(xp != xp_ref, axis=0))
    assert bad.size == 0, f"{bad.size} xp mismatches, first {bad[:5]}"
    return P, d, xp


def test_icosphere_lattice():
    mesh, grid = _cases.icosphere_case(subdiv=3, n_max=24)
    _check(mesh, grid)


def test_star_lattice_and_far_points():
    mesh, grid = _cases.star_case(n=40, n_max=20, seed=1)
    _check(mesh, grid)
    rng = np.random.default_rng(5)
    lo, hi = mesh.X.min(1), mesh.X.max(1)
    far = rng.uniform(lo - 3.0, hi + 3.0, size=(3000, 3)).T
    _check(mesh, grid, np.asfortranarray(far))


def test_artifacts_mesh_lattice():
    mesh, grid = _cases.artifacts_case(0.5)
    _check(mesh, grid)


def test_degenerate_and_duplicate_triangles():
    """Zero-area triangles (collinear and repeated vertices) take the edge
    fallback; duplicated triangles tie and the lower index wins."""
    X, IEN = synthetic.icosphere(1.0, 2)
    X = np.asfortranarray(X)
    nnp = X.shape[1]
    extra_X = np.array([[0.0, 0, 0], [0.5, 0.5, 0.5], [1.0, 1.0, 1.0]]).T   # collinear
    X2 = np.asfortranarray(np.hstack([X, extra_X]))
    col = np.array([[nnp + 1], [nnp + 2], [nnp + 3]])
    rep = np.array([[1], [1], [2]])                     # repeated vertex
    dup = IEN[:, :7]                                    # duplicates of 7 faces
    IEN2 = np.asfortranarray(np.hstack([dup, IEN, col, rep]))
    mesh = MeshGrid.TriangularMesh(X2, IEN2)
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    grid = MeshGrid.Grid(xmin, xmax, 20, 3)
    _check(mesh, grid)


def test_empty_mesh_gives_zero_xp():
    X = np.zeros((3, 3), order="F")
    mesh = MeshGrid.TriangularMesh(X, np.zeros((3, 0), dtype=np.int64, order="F"))
    grid = MeshGrid.Grid(np.zeros(3), np.ones(3), 4, 0)
    d, xp = SignedDistances.evalDistancesOnTriMesh(mesh, grid)
    assert xp.shape == (3, grid.ngp) and not np.any(xp)


def test_large_star_samples_and_properties():
    """A 200k-triangle surface on a 96-point grid: a sample of points equals
    the brute-force oracle bit for bit; every point's |p - xp| is >= its
    clamped distance and equals it (to the distance tolerance) below h; and
    SelectProjectedNodes keeps every point (xp is never the zero vector
    here) with max projection distance = max |p - xp|."""
    X, IEN, _ = synthetic.star_surface(316, 316, 0)
    X = synthetic.to_float32(X)
    mesh = MeshGrid.TriangularMesh(X, IEN)
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    grid = MeshGrid.Grid(xmin, xmax, 93, 3)
    P = MeshGrid.generateGridPoints(grid)
    print(f"case: nel={mesh.nel} ngp={grid.ngp}", flush=True)
    d, xp = SignedDistances.evalDistancesOnTriMesh(mesh, grid)
    r = np.linalg.norm(P - xp, axis=0)
    tol = 1e-6 * _cases.diag(mesh)
    assert np.all(r >= d - tol)
    band = d < grid.cell_size
    assert np.all(np.abs(r[band] - d[band]) <= tol)
    rng = np.random.default_rng(2)
    idx = np.sort(rng.choice(P.shape[1], 384, replace=False))
    _, xp_ref = oracle.dist_exact(mesh.X, mesh.IEN, np.asfortranarray(P[:, idx]), 1e9,
                                  want_xp=True)
    assert np.array_equal(xp[:, idx], xp_ref)
    Xs, Xp, mean_pd, max_pd = SelectProjectedNodes(mesh, grid, xp, P)
    assert Xs.shape[1] == grid.ngp and abs(max_pd - float(r.max())) <= 1e-12
