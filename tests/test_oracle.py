"""Pins the CPU oracle (no GPU): each check is chosen so that a plausible
mistake in the restatement (dropped term, wrong sign, swapped operand,
wrong tie rule) fails it.  The reference holds no golden outputs for this
path, so the pins are closed forms, brute-force alternatives, scipy, and the
grid sizes the reference's formulas give for its own meshes."""
import json
import math
import os

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle
from stl2sdf import MeshGrid, synthetic
from stl2sdf.DataImport import import_stl
from tests import _cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------- grid (Grid.jl:9-33, Grid_setup.jl:104-124) ----------------

def test_grid_sizes_of_reference_meshes():
    mesh, grid = _cases.artifacts_case(0.5)
    assert list(grid.N) == [126, 46, 14] and grid.ngp == 89535
    info = json.load(open(os.path.join(GOLDEN, "voronoi_rabbit_aabb.json")))
    xmin, xmax = np.array(info["xmin"]), np.array(info["xmax"])
    n_new = int(math.floor(max(xmax - xmin) / 0.5))
    amin, amax, N, h, ngp = oracle.grid(xmin, xmax, n_new, 3)
    assert list(N) == [76, 127, 130] and ngp == 1291136
    # icosphere config: margin 2 -> 45^3, stl_to_sdf margin 3 -> 47^3
    assert _cases.icosphere_case(margin=2)[1].ngp == 91125
    assert _cases.icosphere_case(margin=3)[1].ngp == 103823


def test_grid_covers_mesh_and_is_uniform():
    rng = np.random.default_rng(1)
    for _ in range(20):
        xmin = rng.uniform(-5, 5, 3)
        xmax = xmin + rng.uniform(0.1, 4, 3)
        nmax = int(rng.integers(3, 90))
        m = int(rng.integers(0, 4))
        amin, amax, N, h, ngp = oracle.grid(xmin, xmax, nmax, m)
        assert h == max(xmax - xmin) / nmax
        assert np.all(amin <= xmin - m * h * (1 - 1e-12))
        assert np.all(amax >= xmax + m * h * (1 - 1e-9))
        assert ngp == np.prod(N + 1)
        # ceil of a rounded quotient: one extra cell when it lands just above
        assert max(N) in (nmax + 2 * m, nmax + 2 * m + 1)


# ---------------- ray directions (RaycastingSignDetection.jl:162-227) -------

def test_ray_directions():
    d = oracle.ray_directions()
    assert d.shape == (36, 3)
    assert np.array_equal(d[:6], np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1],
                                           [-1, 0, 0], [0, -1, 0], [0, 0, -1]], float))
    r2, r3 = 1.0 / math.sqrt(2.0), 1.0 / math.sqrt(3.0)
    assert set(np.abs(d[6:18]).ravel()) == {0.0, r2}
    assert set(np.abs(d[18:26]).ravel()) == {r3}
    # the 26 lattice directions are all distinct unit vectors
    assert np.allclose(np.linalg.norm(d[:26], axis=1), 1.0, atol=2e-16)
    assert len({tuple(x) for x in d[:26]}) == 26
    # golden-angle directions: z = 1 - 2 i/10, on the unit sphere, azimuth 2pi frac(i phi)
    phi = (1 + 5 ** 0.5) / 2
    for i in range(1, 11):
        x, y, z = d[25 + i]
        assert z == 1.0 - 2.0 * (i / 10.0)
        assert abs(x * x + y * y + z * z - 1.0) < 1e-15
        if i != 5 and i != 10 and abs(z) < 1:
            az = math.atan2(y, x) % (2 * math.pi)
            assert abs(az - 2 * math.pi * ((i * phi) % 1.0)) < 1e-12


# ---------------- Moller-Trumbore (:64-106) ----------------

T0 = ([0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0])


def test_mt_known_hits_and_misses():
    down = [0, 0, -1.0]
    assert oracle.ray_triangle([0.25, 0.25, 1.0], down, *T0) == 1.0
    assert oracle.ray_triangle([0.25, 0.25, 0.5], down, *T0) == 0.5
    assert oracle.ray_triangle([0.25, 0.25, -1.0], down, *T0) is None      # behind
    assert oracle.ray_triangle([0.25, 0.25, 1.0], [1.0, 0, 0], *T0) is None  # parallel
    assert oracle.ray_triangle([0.5, 0.5, 1.0], down, *T0) == 1.0          # on edge u+v=1
    assert oracle.ray_triangle([0.0, 0.0, 1.0], down, *T0) == 1.0          # at a vertex
    assert oracle.ray_triangle([0.5 + 1e-9, 0.5, 1.0], down, *T0) is None  # just outside
    assert oracle.ray_triangle([-1e-12, 0.2, 1.0], down, *T0) is None      # u < 0 side
    assert oracle.ray_triangle([0.25, 0.25, 5e-9], down, *T0) is None      # t <= eps
    assert oracle.ray_triangle([0.25, 0.25, 2e-8], down, *T0) is not None  # t > eps
    # oblique ray: t is the distance along a unit direction
    d = np.array([1.0, 1.0, -2.0]) / math.sqrt(6.0)
    t = oracle.ray_triangle([0.1, 0.1, 0.4], d, *T0)
    assert abs(t - 0.2 * math.sqrt(6.0)) < 1e-15


def _cube():
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]])
    return np.asfortranarray(v.T), np.asfortranarray((f + 1).T)


def test_ray_parity_dedup_on_shared_edges():
    X, IEN = _cube()
    c = [0.5, 0.5, 0.5]
    # through a face interior, through a cube edge (two triangles, same t),
    # through a face diagonal (shared edge of the two face triangles), and
    # through a cube corner (several triangles meet): all count once -> inside
    for d in ([1, 0, 0], [1, 1, 0], [1, 1, 1], [0, 1, 1]):
        d = np.array(d, float) / np.linalg.norm(d)
        assert oracle.ray_inside(X, IEN, c, d)
    # outside point, ray through two faces -> even
    assert not oracle.ray_inside(X, IEN, [-0.5, 0.5, 0.5], [1.0, 0, 0])
    # outside point whose ray crosses the bottom and the top face exactly on
    # their diagonals (two triangles each at the same t, counted once) and
    # runs parallel to the four side faces (|a| < eps): two crossings
    assert oracle.ray_inside(X, IEN, [0.3, 0.3, -1.0], [0, 0, 1.0]) is False
    # the same ray from the cube's inside crosses only the top face
    assert oracle.ray_inside(X, IEN, [0.3, 0.3, 0.5], [0, 0, 1.0]) is True


def test_signs_closed_form_and_bvh_equals_brute():
    mesh, grid = _cases.icosphere_case(subdiv=3, n_max=20)
    P = MeshGrid.generateGridPoints(grid)
    s1, c1 = oracle.signs(mesh.X, mesh.IEN, P, use_bvh=True)
    s0, c0 = oracle.signs(mesh.X, mesh.IEN, P, use_bvh=False)
    assert np.array_equal(s0, s1) and np.array_equal(c0, c1)
    r = np.linalg.norm(P, axis=0)
    r_in = 0.98  # inscribed radius of the subdivided icosphere is above this
    assert np.all(s1[r < r_in] == 1) and np.all(s1[r > 1.0 + 1e-9] == -1)
    mesh, grid = _cases.open_case()
    P = MeshGrid.generateGridPoints(grid)[:, ::5]
    s1, c1 = oracle.signs(mesh.X, mesh.IEN, P, use_bvh=True)
    s0, c0 = oracle.signs(mesh.X, mesh.IEN, P, use_bvh=False)
    assert np.array_equal(s0, s1) and np.array_equal(c0, c1)


def test_adaptive_vote_counts():
    """Rays used follow 6 -> +4 batches (<= 30) and the 0.85 / 0.95 rules."""
    mesh, grid = _cases.open_case()
    P = MeshGrid.generateGridPoints(grid)
    s, c, rays, wind = oracle.signs(mesh.X, mesh.IEN, P, diagnostics=True)
    assert set(np.unique(rays)) <= {6, 10, 14, 18, 22, 26, 30}
    assert np.all(c[rays == 6] == 1.0)                        # 6/6 or 0/6 only
    assert np.all((c[wind == 1] == 0.8))
    assert np.all(c[(wind == 0)] >= 0.6)
    assert np.all(rays[wind == 1] == 30)                     # fallback only after 30 rays


def test_winding_sign_closed_mesh():
    X, IEN = synthetic.icosphere(1.0, 2)
    for p, want in (([0, 0, 0], 1.0), ([0.3, -0.2, 0.5], 1.0), ([1.5, 0, 0], -1.0),
                    ([0, 3, 0], -1.0)):
        assert oracle.winding_sign(X, IEN, p) == want


# ---------------- exact distance ----------------

def _dist_np(p, a, b, c):
    """Distance point-triangle by a different route: projection onto the
    plane when it falls inside, else the nearest of the three segments."""
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    # barycentric of q by areas
    def area(x, y, z):
        return np.dot(np.cross(y - x, z - x), n)
    A = area(a, b, c)
    l1, l2, l3 = area(q, b, c) / A, area(a, q, c) / A, area(a, b, q) / A
    if min(l1, l2, l3) >= 0:
        return abs(np.dot(p - a, n))
    best = np.inf
    for s0, s1 in ((a, b), (b, c), (c, a)):
        e = s1 - s0
        t = np.clip(np.dot(p - s0, e) / np.dot(e, e), 0, 1)
        best = min(best, np.linalg.norm(p - (s0 + t * e)))
    return best


def test_exact_distance_against_independent_formula():
    rng = np.random.default_rng(7)
    for _ in range(30):
        V = rng.normal(size=(3, 3))
        X = np.asfortranarray(V.T)
        IEN = np.array([[1], [2], [3]])
        P = np.asfortranarray(rng.normal(scale=2.0, size=(3, 50)))
        d = oracle.dist_exact(X, IEN, P, 1e9)
        for k in range(P.shape[1]):
            ref = _dist_np(P[:, k], V[0], V[1], V[2])
            assert abs(d[k] - ref) <= 1e-12 * (1 + ref)


def test_exact_distance_closed_forms_and_clamp():
    X = np.asfortranarray(np.array(T0).T)
    IEN = np.array([[1], [2], [3]])
    P = np.asfortranarray(np.array([[0.2, 0.2, 0.7], [-0.3, -0.4, 0.0], [0.5, -2.0, 0.0],
                                    [2.0, 2.0, 0.0], [0.25, 0.25, 100.0]]).T)
    d = oracle.dist_exact(X, IEN, P, 10.0)
    assert d[0] == 0.7                               # face region
    assert abs(d[1] - 0.5) < 1e-15                   # vertex region (3-4-5)
    assert d[2] == 2.0                               # edge region (y = 0 edge)
    assert abs(d[3] - 1.5 * math.sqrt(2)) < 1e-15    # hypotenuse edge
    assert d[4] == 10.0                              # clamped at h
    mesh, grid = _cases.icosphere_case(subdiv=3, n_max=16)
    P = MeshGrid.generateGridPoints(grid)
    d = oracle.dist_exact(mesh.X, mesh.IEN, P, 1e9)
    r = np.linalg.norm(P, axis=0)
    out = r > 1.0
    # vertices on the unit sphere, faces inside it: |p| - 1 <= d outside
    assert np.all(d[out] >= r[out] - 1.0 - 1e-7)
    assert np.all(d <= np.abs(r - 1.0) + 0.02)


def test_exact_closest_point_closed_forms_and_tie_rule():
    """xp of the exact distance: the projection in the face region, the
    vertex and edge points elsewhere, and on an exact tie the triangle with
    the lower index."""
    X = np.asfortranarray(np.array(T0).T)
    IEN = np.array([[1], [2], [3]])
    P = np.asfortranarray(np.array([[0.2, 0.2, 0.7], [-0.3, -0.4, 0.0], [0.5, -2.0, 0.0],
                                    [2.0, 2.0, 0.0], [0.25, 0.25, 100.0]]).T)
    _, xp = oracle.dist_exact(X, IEN, P, 10.0, want_xp=True)
    assert np.allclose(xp[:, 0], [0.2, 0.2, 0.0], rtol=0, atol=1e-15)  # face: foot of the normal
    assert np.array_equal(xp[:, 1], [0.0, 0.0, 0.0])     # vertex a
    assert np.array_equal(xp[:, 2], [0.5, 0.0, 0.0])     # edge y = 0
    assert np.allclose(xp[:, 3], [0.5, 0.5, 0.0], atol=1e-15)  # hypotenuse midpoint
    assert np.allclose(xp[:, 4], [0.25, 0.25, 0.0], rtol=0, atol=1e-15)  # beyond h: not clamped
    # two mirror triangles at z = +-1; the origin is 1 from both
    up = [[0.0, 0, 1], [1.0, 0, 1], [0.0, 1, 1]]
    dn = [[0.0, 0, -1], [1.0, 0, -1], [0.0, 1, -1]]
    X2 = np.asfortranarray(np.array(up + dn).T)
    P0 = np.asfortranarray(np.array([[0.1, 0.1, 0.0]]).T)
    for first, z in ((np.array([[1, 4], [2, 5], [3, 6]]), 1.0),
                     (np.array([[4, 1], [5, 2], [6, 3]]), -1.0)):
        _, xp = oracle.dist_exact(X2, first, P0, 10.0, want_xp=True)
        assert np.allclose(xp[:, 0], [0.1, 0.1, z], rtol=0, atol=1e-15)


def test_exact_closest_point_lies_on_mesh_and_realises_distance():
    """Independent of the closest-point formula: |p - xp| equals the exact
    distance (pinned above against an area-coordinate formula) and xp is on
    the mesh (its own distance to the mesh is zero)."""
    mesh, grid = _cases.star_case(n=24, n_max=16, seed=3)
    rng = np.random.default_rng(11)
    lo, hi = mesh.X.min(1), mesh.X.max(1)
    P = np.asfortranarray(rng.uniform(lo - 1.0, hi + 1.0, size=(400, 3)).T)
    d, xp = oracle.dist_exact(mesh.X, mesh.IEN, P, 1e9, want_xp=True)
    r = np.linalg.norm(P - xp, axis=0)
    assert np.all(np.abs(r - d) <= 1e-12 * (1 + d))
    d_on = oracle.dist_exact(mesh.X, mesh.IEN, np.asfortranarray(xp), 1e9)
    assert d_on.max() <= 1e-12
    for k in range(0, 400, 40):   # brute force with the independent formula
        V = mesh.X[:, mesh.IEN - 1]
        best = min(_dist_np(P[:, k], V[:, 0, e], V[:, 1, e], V[:, 2, e])
                   for e in range(mesh.nel))
        assert abs(best - d[k]) <= 1e-12 * (1 + best)


def test_reference_faithful_distance_bounds():
    mesh, grid = _cases.icosphere_case(subdiv=3, n_max=16)
    P = MeshGrid.generateGridPoints(grid)
    h = grid.cell_size
    de = oracle.dist_exact(mesh.X, mesh.IEN, P, h)
    dr = oracle.dist_reference(mesh.X, mesh.IEN, P, h)
    # every value the reference logic produces is a distance to a point on
    # the mesh, clamped at h: never below the exact one
    assert np.all(dr >= de - 1e-12) and np.all(dr <= h)
    assert np.mean(np.abs(dr - de) < 1e-12) > 0.95


# ---------------- components / artifacts (SdfArtifactRemoval.jl) ----------

def test_components_match_scipy_label():
    rng = np.random.default_rng(3)
    for shape in ((7, 9, 11), (20, 13, 9), (1, 5, 30)):
        nz, ny, nx = shape
        sdf = rng.normal(size=shape)
        sdf = ndimage.uniform_filter(sdf, 3)
        N = np.array([nx - 1, ny - 1, nz - 1])
        lab, nc = oracle.components(sdf.ravel(), N)
        sl, n2 = ndimage.label(sdf >= 0.0, structure=ndimage.generate_binary_structure(3, 1))
        assert nc == n2
        flat = sl.ravel()
        for comp in range(1, n2 + 1):
            members = np.flatnonzero(flat == comp)
            assert np.all(lab[members] == members.min())
        assert np.all(lab[flat == 0] == -1)


def test_remove_artifacts_round_half_even_and_negative_zero():
    # components along x of sizes 50, 3, 2, 1 separated by exterior nodes
    sizes = [50, 3, 2, 1]
    line = []
    for s in sizes:
        line += [1.0] * s + [-1.0]
    sdf = np.array(line)
    N = np.array([len(line) - 1, 0, 0])
    # ratio * 50 = 2.5 -> round-half-even 2: sizes 1 flipped only
    out, n = oracle.remove_artifacts(sdf, N, 0.0, 0.05)
    assert n == 1
    # ratio * 50 = 3.5 -> 4: sizes 3, 2, 1 flipped
    out, n = oracle.remove_artifacts(sdf, N, 0.0, 0.07)
    assert n == 6 and np.all(out[:50] == 1.0) and np.all(out[50:] <= 0)
    # -0.0 counts as interior (sdf >= 0)
    z = np.array([1.0] * 10 + [-1.0, -0.0, -1.0])
    lab, nc = oracle.components(z, np.array([12, 0, 0]))
    assert nc == 2 and lab[11] == 11
    out, n = oracle.remove_artifacts(z, np.array([12, 0, 0]), 0.0, 0.2)
    assert n == 1 and np.signbit(out[11])
    # no interior at all -> 0
    out, n = oracle.remove_artifacts(-np.ones(8), np.array([7, 0, 0]), 0.0, 0.01)
    assert n == 0


# ---------------- RBF (RBFs4Smoothing.jl) ----------------

def _dense_K(N, threshold=1e-3):
    nx, ny, nz = (int(v) + 1 for v in N)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    C = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(float)
    D2 = ((C[:, None, :] - C[None, :, :]) ** 2).sum(-1)
    K = np.exp(-D2)
    K[K <= threshold] = 0.0
    return K.astype(np.float32).astype(np.float64)


def test_rbf_weights_solve_dense_system():
    N = np.array([5, 4, 6])
    rng = np.random.default_rng(2)
    b = rng.normal(size=int(np.prod(N + 1))).astype(np.float32)
    K = _dense_K(N)
    assert int((K[K.shape[0] // 2] != 0).sum()) <= 81
    w_ref = np.linalg.solve(K, b.astype(np.float64))
    w, it, res = oracle.rbf_weights(b, N, 1.0, reltol=1e-10)
    assert np.max(np.abs(w - w_ref)) <= 1e-5 * np.max(np.abs(w_ref))
    # default stopping rule of the reference's cg: ||r|| <= sqrt(eps(Float32)) ||b||
    w2, it2, res2 = oracle.rbf_weights(b, N, 1.0)
    assert res2 <= 3.46e-4 and it2 <= it


def test_rbf_interior_row_has_81_taps():
    K = _dense_K(np.array([8, 8, 8]))
    centre = (4 * 9 + 4) * 9 + 4
    assert int((K[centre] != 0).sum()) == 81


def test_rbf_fine_reproduces_coarse_at_s1():
    N = np.array([6, 5, 4])
    rng = np.random.default_rng(4)
    w = rng.normal(size=int(np.prod(N + 1))).astype(np.float32)
    f = oracle.rbf_fine(w, np.zeros(3), N.astype(float), N, 1.0, 1)
    K = _dense_K(N)
    # fine evaluation at s = 1 is K w (same support, exp(-d^2) weights)
    assert np.max(np.abs(f - K @ w.astype(np.float64))) < 1e-5 * np.abs(w).sum()


# ---------------- STL import ----------------

def test_stl_readers_agree():
    path = os.path.join(GOLDEN, "artifacts.stl")
    X1, I1 = oracle.read_stl_binary(path)
    X2, I2 = import_stl(path)
    assert np.array_equal(X1, X2) and np.array_equal(I1, I2)
    assert I1.shape[1] == 3712


# ---------------- golden fixtures ----------------

@pytest.mark.parametrize("name,case", [("c1_icosphere.npz", "icosphere"),
                                       ("artifacts_05.npz", "artifacts")])
def test_oracle_reproduces_golden(name, case):
    g = np.load(os.path.join(GOLDEN, name))
    mesh, grid = _cases.icosphere_case() if case == "icosphere" else _cases.artifacts_case(0.5)
    assert np.array_equal(g["N"], grid.N) and float(g["h"]) == grid.cell_size
    P = MeshGrid.generateGridPoints(grid)
    s, c = oracle.signs(mesh.X, mesh.IEN, P)
    assert np.array_equal(np.packbits(s > 0), g["signs"])
    idx = np.flatnonzero(c != 1.0)
    assert np.array_equal(idx, g["conf_idx"]) and np.array_equal(c[idx], g["conf_val"])
    d = oracle.dist_exact(mesh.X, mesh.IEN, P, grid.cell_size)
    assert np.max(np.abs(d - g["dist"])) <= 1e-7 * grid.cell_size
