"""Batched signed distance: every mesh of a batch equals its single-mesh
result, bit for bit, across runs, after the meshes change in place, and
with the autotuner's choices."""
import numpy as np
import pytest

from stl2sdf import Batch, MeshGrid, SignedDistances, synthetic
from tests import _cases

pytestmark = pytest.mark.gpu


def _family(count, seed, tri_range, grid_points):
    meshes, grids = [], []
    for X, IEN, g, _ in synthetic.batch_family(count, seed, tri_range, grid_points):
        mesh = MeshGrid.TriangularMesh(X, IEN)
        xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
        meshes.append(mesh)
        grids.append(MeshGrid.Grid(xmin, xmax, g - 7, 3))
    return meshes, grids


def _check(b, meshes, grids):
    for i, (m, g, got) in enumerate(zip(meshes, grids, b.results())):
        ref = SignedDistances.signed_distance(m, g)
        assert np.array_equal(got, ref), f"mesh {i}: {np.count_nonzero(got != ref)} differ"


def test_batch_matches_single_mesh():
    meshes, grids = _family(6, 7, (500, 20000), (24, 32, 48))
    ico, gi = _cases.icosphere_case(3, 20, 2)
    meshes.append(ico)
    grids.append(gi)
    print("batch:", [(m.nel, int(g.ngp)) for m, g in zip(meshes, grids)], flush=True)
    b = Batch.SignedDistanceBatch(meshes, grids, autotune=False)
    b.run()
    _check(b, meshes, grids)
    b.run()   # a second run
    _check(b, meshes, grids)
    print("info:", b.info())
    b.close()


def test_batch_replay_after_in_place_change():
    """Replays recompute from the current device arrays; lists that outgrow
    their capacities are re-planned."""
    meshes, grids = _family(4, 8, (1000, 8000), (24, 40))
    b = Batch.SignedDistanceBatch(meshes, grids, autotune=False)
    b.run()
    _check(b, meshes, grids)
    # refine nothing, but move every vertex: the same topology, new geometry
    for k, (m, g) in enumerate(zip(meshes, grids)):
        Xn = np.asfortranarray(m.X * 0.93 + 0.01 * (k + 1))
        meshes[k] = MeshGrid.TriangularMesh(Xn, m.IEN)
        b.set_vertices(k, Xn)
    b.run()
    _check(b, meshes, grids)
    b.close()


def test_batch_autotune():
    meshes, grids = _family(5, 9, (2000, 30000), (32, 48))
    b = Batch.SignedDistanceBatch(meshes, grids, autotune=True)
    info = b.info()
    print("autotune:", {k: v for k, v in info.items() if k != "tuning"})
    print("candidates:", info["tuning"])
    assert info["candidates"] >= 10
    assert info["chunk_len"] in (16, 32, 64) and info["ray_stack"] in (1024, 2048)
    b.run()
    _check(b, meshes, grids)
    b.close()


def test_batch_config4_family():
    """The batched configuration at full size: 64 seeded star surfaces with
    triangle counts log-uniform in [2k, 500k] and grids of 64, 96, 128 or 192
    points on the longest axis; every mesh equals its single-mesh result bit
    for bit."""
    meshes, grids = [], []
    for X, IEN, g, _ in synthetic.batch_family(64, 2):
        mesh = MeshGrid.TriangularMesh(X, IEN)
        xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
        meshes.append(mesh)
        grids.append(MeshGrid.Grid(xmin, xmax, g - 7, 3))
    print("config 4:", sum(m.nel for m in meshes), "triangles,",
          sum(int(g.ngp) for g in grids), "points", flush=True)
    b = Batch.SignedDistanceBatch(meshes, grids, autotune=False)
    b.run()
    print("info:", b.info(), flush=True)
    _check(b, meshes, grids)
    b.close()
