"""GPU parity for artifact removal, component analysis and RBF smoothing,
and the stl_to_sdf pipeline end to end.

Components: canonical labels (smallest index per component) and the flip
set must equal the oracle's exactly.  RBF: fine fields within 1e-4 x max|f|
of the oracle's FP64 solve when both are converged tightly; with the
reference's default tolerance the GPU residual must meet
||b - K w|| <= sqrt(eps(Float32)) ||b||."""
import os

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle
from stl2sdf import MeshGrid, SdfSmoothing, SignedDistances, synthetic
from stl2sdf.DataExport import read_vti
from stl2sdf.stl_to_sdf import Options, stl_to_sdf
from tests import _cases

pytestmark = pytest.mark.gpu


def _grid_for(N):
    N = np.asarray(N, np.int64)
    return MeshGrid.Grid(None, None, None, _raw=(np.zeros(3), N.astype(float), N, 1.0,
                                                 int(np.prod(N + 1))))


def _random_field(shape, seed, smooth=2):
    rng = np.random.default_rng(seed)
    f = ndimage.uniform_filter(rng.normal(size=shape), smooth)
    return np.ascontiguousarray(f.ravel())


@pytest.mark.parametrize("shape,seed", [((9, 11, 13), 0), ((40, 33, 27), 1), ((1, 1, 200), 2),
                                        ((64, 64, 64), 3)])
def test_components_labels_exact(shape, seed):
    sdf = _random_field(shape, seed)
    N = np.array([shape[2] - 1, shape[1] - 1, shape[0] - 1])
    grid = _grid_for(N)
    print(f"case: components shape={shape} seed={seed}", flush=True)
    d, lab = SignedDistances.analyze_sdf_components(sdf, grid, labels=True)
    ref, nc = oracle.components(sdf, N)
    assert len(d) == nc
    assert np.array_equal(np.where(ref >= 0, ref + 1, 0), lab)
    sizes = np.bincount(ref[ref >= 0])
    for root, size in d.items():
        assert sizes[root - 1] == size


@pytest.mark.parametrize("ratio", [0.01, 0.05, 0.3])
def test_remove_artifacts_matches_oracle(ratio):
    sdf = _random_field((48, 40, 36), 7, smooth=3)
    N = np.array([35, 39, 47])
    grid = _grid_for(N)
    ref, nref = oracle.remove_artifacts(sdf, N, 0.0, ratio)
    out = sdf.copy()
    n = SignedDistances.remove_sdf_artifacts_(out, grid, min_component_ratio=ratio)
    assert n == nref and n > 0
    assert np.array_equal(out, ref)


def test_remove_artifacts_edge_rules():
    line = []
    for s in [50, 3, 2, 1]:
        line += [1.0] * s + [-1.0]
    base = np.array(line)
    grid = _grid_for([len(line) - 1, 0, 0])
    for ratio, want in ((0.05, 1), (0.07, 6)):   # round-half-even: 2.5 -> 2, 3.5 -> 4
        out = base.copy()
        assert SignedDistances.remove_sdf_artifacts_(out, grid, min_component_ratio=ratio) == want
        assert np.array_equal(out, oracle.remove_artifacts(base, grid.N, 0.0, ratio)[0])
    z = np.array([1.0] * 10 + [-1.0, -0.0, -1.0])
    g2 = _grid_for([12, 0, 0])
    out = z.copy()
    assert SignedDistances.remove_sdf_artifacts_(out, g2, min_component_ratio=0.2) == 1
    assert np.signbit(out[11])
    neg = -np.ones(8)
    assert SignedDistances.remove_sdf_artifacts_(neg, _grid_for([7, 0, 0])) == 0
    with pytest.raises(ValueError):
        SignedDistances.remove_sdf_artifacts_(np.ones(5), _grid_for([7, 0, 0]))


def test_artifacts_on_signed_field_with_shells():
    # config-3 family at small size: star body plus 40 small shells
    Xb, Ib, _ = synthetic.star_surface(60, 60, 0)
    Xs, Is = synthetic.shells(40, 0.03, 0.08, Xb.min(1), Xb.max(1), seed=1)
    X, IEN = synthetic.merge((synthetic.to_float32(Xb), Ib), (synthetic.to_float32(Xs), Is))
    mesh = MeshGrid.TriangularMesh(X, IEN)
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    grid = MeshGrid.Grid(xmin, xmax, 64, 3)
    sdf = SignedDistances.signed_distance(mesh, grid)
    P = MeshGrid.generateGridPoints(grid)
    s_ref, _ = oracle.signs(mesh.X, mesh.IEN, P)
    assert np.array_equal(np.sign(sdf)[sdf != 0], s_ref[sdf != 0])
    ref, nref = oracle.remove_artifacts(sdf, grid.N, 0.0, 0.01)
    out = sdf.copy()
    n = SignedDistances.remove_sdf_artifacts_(out, grid, min_component_ratio=0.01)
    assert n == nref and np.array_equal(out, ref)


@pytest.mark.parametrize("N,s", [((10, 8, 12), 1), ((10, 8, 12), 2), ((24, 20, 18), 2)])
def test_rbf_interpolation_matches_oracle(N, s):
    N = np.array(N)
    grid = _grid_for(N)
    vals = _random_field(tuple((N + 1)[::-1]), 5)
    info = {}
    fine, fgrid = SdfSmoothing.RBFs_smoothing(vals, grid, True, s, reltol=1e-7, info=info)
    w_ref, it, res = oracle.rbf_weights(vals.astype(np.float32), N, 1.0, reltol=1e-10)
    f_ref = oracle.rbf_fine(w_ref, np.zeros(3), N.astype(float), N, 1.0, s)
    assert fine.shape == f_ref.shape
    scale = np.abs(f_ref).max()
    assert np.max(np.abs(fine - f_ref)) <= 1e-4 * scale
    assert info["rel_residual"] <= 1e-7
    if s == 1:  # interpolation reproduces the coarse values
        assert np.max(np.abs(fine - vals)) <= 1e-4 * np.abs(vals).max()


def test_rbf_default_tolerance_residual():
    N = np.array([30, 26, 22])
    grid = _grid_for(N)
    vals = _random_field(tuple((N + 1)[::-1]), 9)
    info = {}
    SdfSmoothing.RBFs_smoothing(vals, grid, True, 1, info=info)
    assert info["rel_residual"] <= SdfSmoothing.DEFAULT_RELTOL
    assert info["cg_iterations"] > 0


def test_rbf_approximation_direct_sum():
    N = np.array([12, 9, 7])
    grid = _grid_for(N)
    vals = _random_field(tuple((N + 1)[::-1]), 4)
    for s in (1, 2):
        fine, _ = SdfSmoothing.RBFs_smoothing(vals, grid, False, s)
        f_ref = oracle.rbf_fine(vals.astype(np.float32), np.zeros(3), N.astype(float), N, 1.0, s)
        assert np.max(np.abs(fine - f_ref)) <= 1e-5 * np.abs(f_ref).max()


def test_stl_to_sdf_reference_test_options(tmp_path):
    # the reference's own test run (test/runtests.jl:39-49)
    stl = os.path.join(_cases.GOLDEN, "artifacts.stl")
    opts = Options(smoothing_method="interpolation", grid_refinement=1, cell_size=0.5,
                   remove_artifacts=True, artifact_ratio=0.01)
    assert stl_to_sdf(stl, options=opts, out_dir=str(tmp_path)) == ()
    v, dims, origin, spacing, name = read_vti(str(tmp_path / "artifacts_sdf.vti"))
    mesh, grid = _cases.artifacts_case(0.5)
    assert dims == (127, 47, 15) and name == "distance" and spacing[0] == grid.cell_size
    P = MeshGrid.generateGridPoints(grid)
    s_ref, _ = oracle.signs(mesh.X, mesh.IEN, P)
    d_ref = oracle.dist_exact(mesh.X, mesh.IEN, P, grid.cell_size)
    expect, _ = oracle.remove_artifacts(d_ref * s_ref, grid.N, 0.0, 0.01)
    assert np.array_equal(np.sign(v), np.sign(expect))
    assert np.max(np.abs(v - expect)) <= 1e-6 * _cases.diag(mesh)
    fv, fdims, _, _, _ = read_vti(str(tmp_path / "artifacts_fine_sdf.vti"))
    assert fdims == dims and fv.dtype == np.float32
    assert np.max(np.abs(fv - v)) <= 2e-3 * np.abs(v).max()  # interpolation at s = 1


def _shell_mesh(n=60, n_max=64):
    Xb, Ib, _ = synthetic.star_surface(n, n, 0)
    Xs, Is = synthetic.shells(40, 0.03, 0.08, Xb.min(1), Xb.max(1), seed=1)
    X, IEN = synthetic.merge((synthetic.to_float32(Xb), Ib), (synthetic.to_float32(Xs), Is))
    mesh = MeshGrid.TriangularMesh(X, IEN)
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    return mesh, MeshGrid.Grid(xmin, xmax, n_max, 3)


@pytest.mark.parametrize("remove", [False, True])
def test_signed_fused_equals_the_separate_steps(remove):
    """sdf_signed_fused (distance, sign, product and artifact removal on the
    device) gives the separate entry points' field bit for bit."""
    mesh, grid = _shell_mesh()
    buf, n = SignedDistances.signed_fused_device(mesh, grid, remove, 0.01)
    fused = buf.to_host(np.float64, grid.ngp)
    ref = SignedDistances.signed_distance(mesh, grid)
    nref = SignedDistances.remove_sdf_artifacts_(ref, grid, min_component_ratio=0.01) \
        if remove else 0
    print(f"flipped {n} (separate steps {nref})", flush=True)
    assert n == nref and (n > 0) == remove
    assert np.array_equal(fused.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("interp,s", [(True, 1), (True, 2), (False, 2)])
def test_rbf_device_equals_host_entry_point(interp, s):
    N = np.array([24, 20, 18])
    grid = _grid_for(N)
    vals = _random_field(tuple((N + 1)[::-1]), 7)
    info_h, info_d = {}, {}
    fine, fg = SdfSmoothing.RBFs_smoothing(vals, grid, interp, s, info=info_h)
    from stl2sdf import _lib
    d_in = _lib.DeviceBuffer(vals.nbytes).from_host(vals)
    d_fine, fg2 = SdfSmoothing.RBFs_smoothing_device(d_in, grid, interp, s, info=info_d)
    fine_d = d_fine.to_host(np.float32, fine.size)
    assert np.array_equal(fine.view(np.uint32), fine_d.view(np.uint32))
    assert info_h == info_d and np.array_equal(fg.N, fg2.N)


def test_stl_to_sdf_fine_grid_mismatch_fails_before_the_solve(tmp_path):
    # cell size 0.25: N = [245, 85, 21], the rebuilt fine grid at s = 2 is
    # [490, 171, 42]; the coarse export is written, then the fine one's
    # length error is raised (ExportToVTI.jl:43-47) without the solve
    stl = os.path.join(_cases.GOLDEN, "artifacts.stl")
    opts = Options(smoothing_method="interpolation", grid_refinement=2, cell_size=0.25)
    with pytest.raises(ValueError, match="doesn't match grid dimensions"):
        stl_to_sdf(stl, options=opts, out_dir=str(tmp_path))
    assert os.path.exists(tmp_path / "artifacts_sdf.vti")
    assert not os.path.exists(tmp_path / "artifacts_fine_sdf.vti")
