"""Parity of the HIP path against the CPU oracle (GPU tests).

Signs must be bit-identical to the oracle (which follows the reference's
Moller-Trumbore operation order); distances must lie within
1e-6 x bbox diagonal of the exact FP64 min(d, h).
"""
import numpy as np
import pytest

from oracle import oracle
from stl2sdf import MeshGrid, SignedDistances, _lib
from tests import _cases

pytestmark = pytest.mark.gpu

DIST_TOL = 1e-6  # x bbox diagonal


def _check(mesh, grid, points=None):
    P = MeshGrid.generateGridPoints(grid) if points is None else points
    print(f"case: nel={mesh.nel} ngp={P.shape[1]} N={list(grid.N)}", flush=True)
    d_gpu, _ = SignedDistances.evalDistancesOnTriMesh(mesh, grid, points)
    s_gpu, c_gpu = SignedDistances.raycast_sign_detection(mesh, grid, points)
    d_ref = oracle.dist_exact(mesh.X, mesh.IEN, P, grid.cell_size)
    s_ref, c_ref = oracle.signs(mesh.X, mesh.IEN, P)
    tol = DIST_TOL * _cases.diag(mesh)
    err = np.abs(d_gpu - d_ref)
    assert err.max() <= tol, f"max distance error {err.max():.3e} > {tol:.3e}"
    mism = np.flatnonzero(s_gpu != s_ref)
    assert mism.size == 0, f"{mism.size} sign mismatches, first at {mism[:5]}"
    assert np.array_equal(c_gpu, c_ref), "confidence mismatch"
    return d_gpu, s_gpu


def test_icosphere_config1():
    mesh, grid = _cases.icosphere_case()
    assert grid.ngp == 91125
    _check(mesh, grid)


def test_icosphere_margin3_lattice_points_passed():
    # stl_to_sdf's margin-3 grid, lattice handed in explicitly
    mesh, grid = _cases.icosphere_case(margin=3)
    P = MeshGrid.generateGridPoints(grid)
    _check(mesh, grid, P)


def test_artifacts_stl():
    mesh, grid = _cases.artifacts_case(0.5)
    assert list(grid.N) == [126, 46, 14] and grid.ngp == 89535
    _check(mesh, grid)


def test_star_surface():
    mesh, grid = _cases.star_case()
    _check(mesh, grid)


def test_open_mesh_adaptive_and_winding():
    mesh, grid = _cases.open_case()
    P = MeshGrid.generateGridPoints(grid)
    _, _, rays, wind = oracle.signs(mesh.X, mesh.IEN, P, diagnostics=True)
    assert (rays > 6).sum() > 100 and wind.sum() > 0  # the case exercises both
    _check(mesh, grid)


def test_general_points():
    mesh, grid = _cases.icosphere_case(subdiv=3, n_max=20)
    rng = np.random.default_rng(5)
    P = np.asfortranarray(rng.uniform(-1.3, 1.3, size=(3, 3000)))
    _check(mesh, grid, P)


def test_min_rays_other_than_six():
    mesh, grid = _cases.open_case()
    P = MeshGrid.generateGridPoints(grid)
    s, c = SignedDistances.raycast_sign_detection(mesh, grid, None, min_rays=4, max_rays=20,
                                                  confidence_threshold=0.7)
    s_ref, c_ref = oracle.signs(mesh.X, mesh.IEN, P, min_rays=4, max_rays=20, conf_thr=0.7)
    assert np.array_equal(s, s_ref) and np.array_equal(c, c_ref)


def test_no_winding_fallback():
    mesh, grid = _cases.open_case()
    P = MeshGrid.generateGridPoints(grid)
    s, c = SignedDistances.raycast_sign_detection(mesh, grid, None, use_winding_fallback=False)
    s_ref, c_ref = oracle.signs(mesh.X, mesh.IEN, P, use_winding=False)
    assert np.array_equal(s, s_ref) and np.array_equal(c, c_ref)
    assert (c < 0.6).any()


def test_fused_equals_product():
    mesh, grid = _cases.star_case(n=40, n_max=30)
    d, _ = SignedDistances.evalDistancesOnTriMesh(mesh, grid)
    s, _ = SignedDistances.raycast_sign_detection(mesh, grid)
    sdf = SignedDistances.signed_distance(mesh, grid)
    assert np.array_equal(sdf, d * s)


def test_bad_index_raises():
    mesh, grid = _cases.icosphere_case(subdiv=1, n_max=8)
    bad = MeshGrid.TriangularMesh(mesh.X, mesh.IEN.copy())
    bad.IEN[1, 3] = mesh.nnp + 1
    with pytest.raises(_lib.SdfError):
        SignedDistances.evalDistancesOnTriMesh(bad, grid)


def test_empty_mesh():
    mesh, grid = _cases.icosphere_case(subdiv=1, n_max=8)
    empty = MeshGrid.TriangularMesh(mesh.X, np.zeros((3, 0), np.int64))
    d, _ = SignedDistances.evalDistancesOnTriMesh(empty, grid)
    s, c = SignedDistances.raycast_sign_detection(empty, grid)
    assert np.all(d == grid.cell_size)
    assert np.all(s == -1.0) and np.all(c == 1.0)


def test_repeat_calls_stable():
    mesh, grid = _cases.star_case(n=40, n_max=30)
    a = SignedDistances.signed_distance(mesh, grid)
    b = SignedDistances.signed_distance(mesh, grid)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name,case", [("c1_icosphere.npz", "icosphere"),
                                       ("artifacts_05.npz", "artifacts")])
def test_against_golden_fixtures(name, case):
    import os
    g = np.load(os.path.join(_cases.GOLDEN, name))
    mesh, grid = _cases.icosphere_case() if case == "icosphere" else _cases.artifacts_case(0.5)
    s, c = SignedDistances.raycast_sign_detection(mesh, grid)
    assert np.array_equal(np.packbits(s > 0), g["signs"])
    idx = np.flatnonzero(c != 1.0)
    assert np.array_equal(idx, g["conf_idx"]) and np.array_equal(c[idx], g["conf_val"])
    d, _ = SignedDistances.evalDistancesOnTriMesh(mesh, grid)
    assert np.max(np.abs(d - g["dist"])) <= DIST_TOL * _cases.diag(mesh) + 1e-7 * grid.cell_size


def test_general_points_many_pending_star():
    # off-lattice points near the surface of a 7k-triangle star: LBVH rays
    mesh, grid = _cases.star_case()
    rng = np.random.default_rng(11)
    P = MeshGrid.generateGridPoints(grid)
    sel = rng.choice(P.shape[1], 4000, replace=False)
    Q = np.asfortranarray(P[:, sel] + rng.normal(scale=grid.cell_size * 0.3, size=(3, 4000)))
    _check(mesh, grid, Q)


def test_large_triangles_cube_fine_grid():
    # 12 triangles spanning thousands of tiles and lines (the per-workgroup
    # binning path); lattice planes pass within 1e-17 of the faces, so many
    # ray decisions sit at the eps boundaries and take the exact path
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]])
    mesh = MeshGrid.TriangularMesh(np.asfortranarray(v.T), np.asfortranarray((f + 1).T))
    grid = MeshGrid.Grid(np.zeros(3), np.ones(3), 98, 3)
    _check(mesh, grid)


def test_list_capacity_growth_on_one_context():
    """A context sized on a small mesh, then given a mesh with many more
    (tile, triangle) pairs and axis-line hits: the first pass overflows its
    lists, reads only what was stored, and the call repeats with larger
    capacities; the results equal the oracle's."""
    import ctypes as C
    L = _lib.lib()
    h = C.c_void_p()
    _lib.check(L.sdf_ctx_create(0, C.byref(h)))
    try:
        for n, n_max in ((12, 10), (120, 30)):
            mesh, grid = _cases.star_case(n=n, n_max=n_max)
            X = np.asfortranarray(mesh.X, np.float64)
            IEN = np.asfortranarray(mesh.IEN, np.int64)
            amin, N = _lib.f64_3(grid.AABB_min), _lib.i64_3(grid.N)
            sdf = np.empty(int(grid.ngp))
            print(f"case: nel={mesh.nel} ngp={grid.ngp}", flush=True)
            _lib.check(L.sdf_signed_distance(h, _lib.dptr(X), X.shape[1], _lib.i64ptr(IEN),
                                             IEN.shape[1], _lib.dptr(amin),
                                             float(grid.cell_size), _lib.i64ptr(N),
                                             _lib.dptr(sdf)))
            P = MeshGrid.generateGridPoints(grid)
            d_ref = oracle.dist_exact(mesh.X, mesh.IEN, P, grid.cell_size)
            s_ref, _ = oracle.signs(mesh.X, mesh.IEN, P)
            assert np.array_equal(np.signbit(sdf), s_ref < 0)   # d * s keeps the sign at 0
            assert np.max(np.abs(np.abs(sdf) - d_ref)) <= DIST_TOL * _cases.diag(mesh)
    finally:
        _lib.check(L.sdf_ctx_destroy(h))


def test_bad_index_on_device_mesh_reported_by_next_call():
    """sdf_mesh_set_device does not wait for the device; an IEN index outside
    1..nnp is reported by every call on that mesh, and a valid mesh clears it.
    Device buffers come from the HIP runtime the library itself uses."""
    import ctypes as C
    mesh, grid = _cases.icosphere_case(subdiv=1, n_max=8)
    L = _lib.lib()
    c = _lib.ctx(0)
    hip = C.CDLL("libamdhip64.so.7")   # the runtime already loaded for the library
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    bufs = []

    def dev(a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
        bufs.append(p)
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p

    try:
        X = dev(np.asfortranarray(mesh.X, np.float64).T.copy())    # 3 x nnp column-major
        bad = mesh.IEN.astype(np.int64).copy()
        bad[1, 3] = mesh.nnp + 1
        Ib = dev(np.asfortranarray(bad).T.copy())
        Ig = dev(np.asfortranarray(mesh.IEN, np.int64).T.copy())
        out = dev(np.zeros(int(grid.ngp)))
        amin, N = _lib.f64_3(grid.AABB_min), _lib.i64_3(grid.N)
        call = lambda: L.sdf_signed_distance_device(c, _lib.dptr(amin), float(grid.cell_size),
                                                    _lib.i64ptr(N), out)
        _lib.check(L.sdf_mesh_set_device(c, X, mesh.nnp, Ib, mesh.nel))
        for _ in range(2):
            with pytest.raises(_lib.SdfError, match="IEN"):
                _lib.check(call())
        _lib.check(L.sdf_mesh_set_device(c, X, mesh.nnp, Ig, mesh.nel))
        _lib.check(call())
        got = np.empty(int(grid.ngp))
        assert hip.hipMemcpy(got.ctypes.data, out, got.nbytes, 2) == 0   # device to host
        ref = SignedDistances.signed_distance(mesh, grid)
        assert np.array_equal(got, ref)
    finally:
        for p in bufs:
            hip.hipFree(p)


def test_grid_covering_part_of_the_mesh():
    """A grid over part of the mesh: triangles outside the grid still count
    for the rays (the LBVH holds the whole mesh), so signs and distances
    equal the oracle's, as for the full grid."""
    mesh, full = _cases.star_case(n=40, n_max=24, seed=2)
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    mid = 0.5 * (np.asarray(xmin) + np.asarray(xmax))
    part = MeshGrid.Grid(np.asarray(xmin), mid, 20, 2)      # lower octant and a margin
    for grid in (full, part):
        P = MeshGrid.generateGridPoints(grid)
        print(f"case: nel={mesh.nel} ngp={grid.ngp} N={list(grid.N)}", flush=True)
        sdf = SignedDistances.signed_distance(mesh, grid)
        s_ref, _ = oracle.signs(mesh.X, mesh.IEN, P)
        d_ref = oracle.dist_exact(mesh.X, mesh.IEN, P, grid.cell_size)
        assert np.array_equal(np.signbit(sdf), s_ref < 0)
        assert np.max(np.abs(np.abs(sdf) - d_ref)) <= DIST_TOL * _cases.diag(mesh)


@pytest.mark.parametrize("min_rays,max_rays", [(6, 40), (6, 36), (5, 37)])
def test_max_rays_beyond_the_direction_table(min_rays, max_rays):
    """max_rays above the 36 directions: the reference skips rays past the
    36th (RaycastingSignDetection.jl:265-271), so the vote ends at 36 rays
    with the same counts; the lattice path and the general-point path both
    equal the oracle, which runs the reference's loop unclamped."""
    mesh, grid = _cases.open_case()
    P = MeshGrid.generateGridPoints(grid)
    print(f"case: open mesh, min_rays={min_rays} max_rays={max_rays}", flush=True)
    s_ref, c_ref = oracle.signs(mesh.X, mesh.IEN, P, min_rays=min_rays, max_rays=max_rays)
    s, c = SignedDistances.raycast_sign_detection(mesh, grid, None, min_rays=min_rays,
                                                  max_rays=max_rays)
    assert np.array_equal(s, s_ref) and np.array_equal(c, c_ref)
    Q = np.asfortranarray(P[:, ::7] + 1e-3 * grid.cell_size)
    s2, c2 = SignedDistances.raycast_sign_detection(mesh, grid, Q, min_rays=min_rays,
                                                    max_rays=max_rays)
    s2_ref, c2_ref = oracle.signs(mesh.X, mesh.IEN, Q, min_rays=min_rays, max_rays=max_rays)
    assert np.array_equal(s2, s2_ref) and np.array_equal(c2, c2_ref)
