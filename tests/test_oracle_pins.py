"""Pins of oracle functions that the restated formulas alone leave open.

* dist_exact_bvh: bit-identical to the brute-force dist_exact it replaces
  on meshes too large for brute force.
* rbf_fine at smooth = 2: a dense evaluation at the reference's own Float32
  fine-grid coordinates (RBFs4Smoothing.jl:38-46, 60-74, 219-253).
* dist_reference: hand-built cases whose reference result differs from the
  exact distance -- the 2.5h triangle-centre candidate rule
  (dfOnTriangularMesh.jl:48, 79), the 3-nearest-vertex fallback (:81-94) and
  the vertex test suppressed by an in-range edge projection (:141-175).
* the adaptive vote's tie rule: ratio 0.5 -> sign -1
  (RaycastingSignDetection.jl:285).
"""
import math

import numpy as np
import pytest

from oracle import oracle
from stl2sdf import MeshGrid, synthetic
from tests import _cases


def _mesh(v, f):
    return (np.asfortranarray(np.asarray(v, float).T),
            np.asfortranarray(np.asarray(f, np.int64).T + 1))


# ---------------- exact distance through the hierarchy ----------------

@pytest.mark.parametrize("case", ["icosphere", "star", "cube", "far_points"])
def test_dist_exact_bvh_bit_identical_to_brute_force(case):
    rng = np.random.default_rng(3)
    if case == "icosphere":
        mesh, grid = _cases.icosphere_case(subdiv=3, n_max=24)
        P, h = MeshGrid.generateGridPoints(grid), grid.cell_size
        X, IEN = mesh.X, mesh.IEN
    elif case == "star":
        X, IEN, _ = synthetic.star_surface(40, 40, 1)
        X = synthetic.to_float32(X)
        P = np.asfortranarray(rng.uniform(-1.6, 1.6, size=(3, 6000)))
        h = 0.07
    elif case == "cube":
        # 12 large triangles: every node box spans most of the mesh
        v = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1],
             [0, 1, 1]]
        f = [[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [2, 3, 7],
             [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]]
        X, IEN = _mesh(v, f)
        P = np.asfortranarray(rng.uniform(-0.3, 1.3, size=(3, 4000)))
        h = 0.2
    else:
        # h far above every distance: the true minimum over all triangles
        X, IEN, _ = synthetic.star_surface(30, 30, 2)
        P = np.asfortranarray(rng.uniform(-4.0, 4.0, size=(3, 3000)))
        h = 1e9
    a = oracle.dist_exact(X, IEN, P, h)
    b = oracle.dist_exact_bvh(X, IEN, P, h)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert (a < h).any() and (case == "far_points" or (a == h).any())


def test_dist_exact_bvh_no_triangles():
    X = np.zeros((3, 3), order="F")
    IEN = np.zeros((3, 0), np.int64, order="F")
    P = np.asfortranarray(np.eye(3))
    assert np.all(oracle.dist_exact_bvh(X, IEN, P, 0.5) == 0.5)


# ---------------- RBF fine evaluation at smooth = 2 ----------------

def _dense_fine_reference_coordinates(w, amin, amax, N, h, s, threshold=1e-3, fine_shift=0.0):
    """rbf_interpolation_kdtree (RBFs4Smoothing.jl:219-253) evaluated densely:
    coarse points range(Float32(xmin), Float32(xmax), length = N + 1)
    (:38-46), fine points xmin + (i - 1) dx in Float32 with
    dx = (xmax - xmin) / (N_x s) from the x extent for every axis (:60-74),
    every coarse point within Float32(sqrt(-log(threshold) sigma^2)) taken
    with weight exp(-(d / sigma)^2), sigma = h.  fine_shift moves the fine
    points only (the negative control)."""
    lo = np.asarray(amin, np.float32)
    hi = np.asarray(amax, np.float32)
    n = np.asarray(N) + 1
    axes_c = [(lo[a].astype(np.float64) + np.arange(n[a]) *
               ((hi[a].astype(np.float64) - lo[a]) / (n[a] - 1))).astype(np.float32)
              for a in range(3)]
    nf = np.asarray(N) * s + 1
    dx = np.float32((hi[0] - lo[0]) / np.float32(nf[0] - 1))
    axes_f = [(lo[a] + np.arange(nf[a]).astype(np.float32) * dx).astype(np.float32)
              for a in range(3)]
    C = np.stack(np.meshgrid(*axes_c, indexing="ij"), -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    F = np.stack(np.meshgrid(*axes_f, indexing="ij"), -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    C = C.astype(np.float64)
    F = F.astype(np.float64) + fine_shift
    rmax = np.float32(math.sqrt(-math.log(threshold) * h * h))
    out = np.zeros(F.shape[0])
    for i0 in range(0, F.shape[0], 512):
        d = np.sqrt(((F[i0:i0 + 512, None, :] - C[None, :, :]) ** 2).sum(-1))
        k = np.where(d <= rmax, np.exp(-(d / h) ** 2), 0.0)
        out[i0:i0 + 512] = k @ w.astype(np.float64)
    return out


def test_rbf_fine_s2_against_dense_evaluation_at_float32_coordinates():
    N = np.array([7, 6, 5])
    h = 0.1
    amin = np.array([0.3, -0.7, 1.1])
    amax = amin + N * h
    rng = np.random.default_rng(12)
    w = rng.normal(size=int(np.prod(N + 1))).astype(np.float32)
    f = oracle.rbf_fine(w, amin, amax, N, h, 2)
    ref = _dense_fine_reference_coordinates(w, amin, amax, N, h, 2)
    assert f.shape == ref.shape == (int(np.prod(N * 2 + 1)),)
    # Float32 coordinates of magnitude <= 1.6 move offsets by <= 1e-6 cells
    assert np.max(np.abs(f - ref)) <= 2e-5 * np.abs(ref).max()
    # fine points half a fine cell off (an index mapped to the wrong node)
    # differ by O(1): the comparison can see such a mistake
    shifted = _dense_fine_reference_coordinates(w, amin, amax, N, h, 2, fine_shift=0.25 * h)
    assert np.max(np.abs(f - shifted)) > 0.05 * np.abs(ref).max()


# ---------------- reference-faithful distance, closed forms ----------------

def test_reference_distance_misses_a_near_triangle_with_a_far_centre():
    # big triangle in z = 0.5 under the origin, centroid (0, 6, 0.5), more
    # than 2.5h away; a tiny triangle at z = -2 has its centre within 2.5h,
    # so the candidate list is not empty and the big one is never tested
    big = [[-3, -1, 0.5], [3, -1, 0.5], [0, 20, 0.5]]
    tiny = [[-0.1, 0, -2], [0.1, 0, -2], [0, 0.1, -2]]
    X, IEN = _mesh(big + tiny, [[0, 1, 2], [3, 4, 5]])
    p = np.zeros((3, 1), order="F")
    h = 1.0
    assert oracle.dist_exact(X, IEN, p, h)[0] == 0.5
    assert oracle.dist_reference(X, IEN, p, h)[0] == h
    # larger h: the big triangle's centre (distance 6.02) is a candidate
    assert oracle.dist_reference(X, IEN, p, 2.5)[0] == 0.5


def test_reference_distance_three_nearest_vertex_fallback():
    # no triangle centre within 2.5h: the triangles around the three nearest
    # vertices are tested (here the big triangle, through its vertex A)
    big = [[-3, -1, 0.5], [3, -1, 0.5], [0, 20, 0.5]]
    X, IEN = _mesh(big, [[0, 1, 2]])
    p = np.zeros((3, 1), order="F")
    assert oracle.dist_reference(X, IEN, p, 1.0)[0] == 0.5


def test_reference_distance_vertex_test_suppressed_by_edge_projection():
    # p = (2, 3, 0) lies in the Voronoi region of v3 = (2, 1, 0): the exact
    # distance is 2.  Its projection onto edge v1 -> v2 is in range and
    # improves min_dist to 3, which marks the triangle face-or-edge, so the
    # vertex test is skipped (dfOnTriangularMesh.jl:141-175): 3.
    tri = [[0, 0, 0], [4, 0, 0], [2, 1, 0]]
    p = np.asfortranarray([[2.0], [3.0], [0.0]])
    X, IEN = _mesh(tri, [[0, 1, 2]])
    assert oracle.dist_exact(X, IEN, p, 4.0)[0] == 2.0
    assert oracle.dist_reference(X, IEN, p, 4.0)[0] == 3.0
    assert oracle.dist_reference(X, IEN, p, 2.5)[0] == 2.5   # then clamped at h
    # An earlier candidate at distance 2.5 keeps min_dist below the edge
    # projection's 3, the edge does not count, and the vertex test finds 2;
    # with the big triangle first, its 3 stands and the other gives 2.5.
    other = [[1.5, 5.5, 0], [2.5, 5.5, 0], [2, 6.5, 0]]   # nearest point (2, 5.5, 0)
    X1, I1 = _mesh(other + tri, [[0, 1, 2], [3, 4, 5]])
    X2, I2 = _mesh(tri + other, [[0, 1, 2], [3, 4, 5]])
    assert oracle.dist_exact(X1, I1, p, 4.0)[0] == 2.0
    assert oracle.dist_reference(X1, I1, p, 4.0)[0] == 2.0
    assert oracle.dist_reference(X2, I2, p, 4.0)[0] == 2.5


# ---------------- vote tie rule ----------------

def test_vote_tie_counts_as_outside():
    # one triangle in x = 1 around (1, 0, 0): the +x ray from the origin
    # crosses it once (inside), the +y ray crosses nothing (outside)
    X, IEN = _mesh([[1, -1, -1], [1, 1, -1], [1, 0, 1]], [[0, 1, 2]])
    p = np.zeros((3, 1), order="F")
    s, c = oracle.signs(X, IEN, p, min_rays=2, max_rays=2, use_winding=False)
    assert s[0] == -1.0 and c[0] == 0.5
    # one more ray (+z, misses): 1 of 3 inside, still outside, confidence
    # max(r, 1 - r) with r = 1/3, rounded as the reference computes it
    s, c = oracle.signs(X, IEN, p, min_rays=3, max_rays=3, use_winding=False)
    assert s[0] == -1.0 and c[0] == 1.0 - 1.0 / 3.0
    # the -x ray alone crosses nothing; the +x ray alone is inside
    s, c = oracle.signs(X, IEN, p, min_rays=1, max_rays=1, use_winding=False)
    assert s[0] == 1.0 and c[0] == 1.0
