"""Parity at the benchmark configurations (GPU tests).

The full grid runs on the GPU; the oracle (FP64, the reference's ray
arithmetic, exact distances through its bounding-volume hierarchy) checks
* a seeded uniform sample of 32768 lattice points,
* a seeded sample of 32768 points of the surface band (|sdf| < h),
* every point the sign kernel handed to the adaptive vote,
* one full z-plane through the middle of the body.
Signs and confidences must be bit-identical, distances within
1e-6 x bbox diagonal.  Over the whole grid: the fused signed distance and
the separate sign entry point agree bit for bit, every |sdf| <= h, and
|sdf| is 1-Lipschitz along each axis (|d(p) - d(q)| <= h + 2 tol between
neighbours), a size-independent check of the points no sample visits.

Workloads: the star surface r = 1 + 0.3 * (8 seeded spherical harmonics)
on a 450 x 450 lat-long mesh (404,100 triangles) with Grid(X_min, X_max,
249, 3), 256 points on the longest axis; and the benchmark's 708 x 708 mesh
(1,001,112 triangles) with Grid(X_min, X_max, 505, 3), 504 x 512 x 320
points.  The reference paths are RaycastingSignDetection.jl:337-404 and
dfOnTriangularMesh.jl:19-209.
"""
import numpy as np
import pytest

from oracle import oracle
from stl2sdf import MeshGrid, SdfSmoothing, SignedDistances, _lib, synthetic

pytestmark = pytest.mark.gpu

DIST_TOL = 1e-6  # x bbox diagonal


def _star(n, n_max):
    X, IEN, info = synthetic.star_surface(n, n, 0)
    mesh = MeshGrid.TriangularMesh(synthetic.to_float32(X), IEN)
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    return mesh, MeshGrid.Grid(xmin, xmax, n_max, 3)


def _lattice_points(grid, idx):
    """generateGridPoints (Grid.jl:82-96) at the given linear indices."""
    N = np.asarray(grid.N, np.int64)
    nx, ny = N[0] + 1, N[1] + 1
    i, j, k = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    P = np.empty((3, idx.size), order="F")
    for a, q in enumerate((i, j, k)):
        P[a] = grid.AABB_min[a] + grid.cell_size * q.astype(np.float64)
    return P


def _check_scale(n, n_max, expect_nel, expect_vote):
    mesh, grid = _star(n, n_max)
    assert mesh.nel == expect_nel
    h = grid.cell_size
    N = np.asarray(grid.N, np.int64)
    print(f"case: star {n}x{n} nel={mesh.nel} N={list(N)} ngp={grid.ngp}", flush=True)
    sdf = SignedDistances.signed_distance(mesh, grid)
    vote = _lib.vote_points(_lib.ctx(0))
    assert vote.size == _lib.counters(_lib.ctx(0))[0]
    s, c = SignedDistances.raycast_sign_detection(mesh, grid)
    vote2 = _lib.vote_points(_lib.ctx(0))
    print(f"  vote-stage points: {vote.size} (sign entry point: {vote2.size})", flush=True)
    assert np.array_equal(vote, vote2) and (vote.size > 0) == expect_vote

    # whole grid: both entry points agree; clamp; 1-Lipschitz along each axis
    assert np.array_equal(np.signbit(sdf), s < 0)
    d = np.abs(sdf)
    tol = DIST_TOL * float(np.linalg.norm(mesh.X.max(1) - mesh.X.min(1)))
    assert d.max() <= h
    D = d.reshape(N[2] + 1, N[1] + 1, N[0] + 1)
    for ax in range(3):
        assert np.abs(np.diff(D, axis=ax)).max() <= h * (1 + 1e-12) + 2 * tol

    rng = np.random.default_rng(0)
    uni = rng.choice(grid.ngp, 32768, replace=False)
    band_all = np.flatnonzero(d < h)
    band = rng.choice(band_all, min(32768, band_all.size), replace=False)
    nxy = (N[0] + 1) * (N[1] + 1)
    kz = (N[2] + 1) // 2
    slab = np.arange(kz * nxy, (kz + 1) * nxy)
    idx = np.unique(np.concatenate([uni, band, vote, slab]))
    print(f"  checked points: {idx.size} (band {band.size}, slab {slab.size})", flush=True)
    P = _lattice_points(grid, idx)
    s_ref, c_ref = oracle.signs(mesh.X, mesh.IEN, P)
    d_ref = oracle.dist_exact_bvh(mesh.X, mesh.IEN, P, h)

    mism = np.flatnonzero(s[idx] != s_ref)
    assert mism.size == 0, f"{mism.size} sign mismatches, first at {idx[mism[:5]]}"
    assert np.array_equal(c[idx], c_ref), "confidence mismatch"
    assert np.array_equal(np.signbit(sdf[idx]), s_ref < 0)
    err = np.abs(d[idx] - d_ref)
    assert err.max() <= tol, f"max distance error {err.max():.3e} > {tol:.3e}"
    on_vote = np.isin(idx, vote)
    assert np.array_equal(s[idx][on_vote], s_ref[on_vote])
    print(f"  max distance error {err.max():.3e} (tol {tol:.3e}); "
          f"{int((c_ref != 1).sum())} checked points with confidence < 1", flush=True)


def test_star_450_grid_256_class():
    _check_scale(450, 249, 404100, False)


def test_star_708_grid_512_class_benchmark_workload():
    _check_scale(708, 505, 1001112, True)


def test_config3_chain_shells_artifacts_rbf_s2():
    """The stl_to_sdf chain of the artifact + smoothing configuration at a
    reduced grid (96 points on the longest axis instead of 192): the 450 x
    450 star body plus 40 seeded icosphere shells (radius 0.01-0.05, 80
    triangles each) -> signed distance -> remove_sdf_artifacts! (ratio 0.01)
    -> RBF interpolation at grid_refinement 2 (stl_to_sdf.jl:101-146).
    Signs bit-exact and distances within tolerance on every grid point; the
    flip set equal to the oracle's on the same field; the fine field within
    1e-4 x max|f| of the oracle's FP64 solve when both are converged tightly,
    and the default-tolerance residual within sqrt(eps(Float32))."""
    Xb, Ib, _ = synthetic.star_surface(450, 450, 0)
    Xs, Is = synthetic.shells(40, 0.01, 0.05, Xb.min(1), Xb.max(1), seed=1)
    X, IEN = synthetic.merge((synthetic.to_float32(Xb), Ib), (synthetic.to_float32(Xs), Is))
    mesh = MeshGrid.TriangularMesh(X, IEN)
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    grid = MeshGrid.Grid(xmin, xmax, 89, 3)
    N = np.asarray(grid.N, np.int64)
    h = grid.cell_size
    print(f"case: config-3 chain nel={mesh.nel} N={list(N)} ngp={grid.ngp}", flush=True)
    assert mesh.nel == 404100 + 40 * 80 and max(N) + 1 in (96, 97)

    sdf = SignedDistances.signed_distance(mesh, grid)
    P = MeshGrid.generateGridPoints(grid)
    s_ref, _ = oracle.signs(mesh.X, mesh.IEN, P)
    d_ref = oracle.dist_exact_bvh(mesh.X, mesh.IEN, P, h)
    assert np.array_equal(np.signbit(sdf), s_ref < 0)
    tol = DIST_TOL * float(np.linalg.norm(xmax - xmin))
    assert np.abs(np.abs(sdf) - d_ref).max() <= tol

    out = sdf.copy()
    n = SignedDistances.remove_sdf_artifacts_(out, grid, min_component_ratio=0.01)
    ref, nref = oracle.remove_artifacts(sdf, N, 0.0, 0.01)
    print(f"  flipped {n} nodes (oracle {nref})", flush=True)
    assert n == nref and n > 0
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))

    info = {}
    fine, fgrid = SdfSmoothing.RBFs_smoothing(out, grid, True, 2, reltol=1e-7, info=info)
    w_ref, it, res = oracle.rbf_weights(out.astype(np.float32), N, h, reltol=1e-10)
    f_ref = oracle.rbf_fine(w_ref, grid.AABB_min, grid.AABB_max, N, h, 2)
    assert fine.shape == f_ref.shape == (int(np.prod(2 * N + 1)),)
    err = np.max(np.abs(fine - f_ref))
    print(f"  RBF: {info['cg_iterations']} CG iterations, fine error {err:.3e} "
          f"(max |f| {np.abs(f_ref).max():.3e}; oracle {it} iterations)", flush=True)
    assert err <= 1e-4 * np.abs(f_ref).max()
    info2 = {}
    SdfSmoothing.RBFs_smoothing(out, grid, True, 2, info=info2)
    assert info2["rel_residual"] <= SdfSmoothing.DEFAULT_RELTOL
