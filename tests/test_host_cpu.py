"""Host-side pieces that need no GPU: Options validation, VTI round trip,
STL reading, fine-grid construction."""
import os

import numpy as np
import pytest

from stl2sdf import MeshGrid, synthetic
from stl2sdf.DataExport import exportSdfToVTI, read_vti
from stl2sdf.DataImport import import_stl
from stl2sdf.stl_to_sdf import Options, stl_to_sdf
from tests import _cases


def test_options_asserts_match_reference(tmp_path):
    stl = os.path.join(_cases.GOLDEN, "artifacts.stl")
    for bad in (Options(smoothing_method="cubic", cell_size=0.5),
                Options(grid_refinement=3, cell_size=0.5),
                Options(artifact_ratio=1.0, cell_size=0.5),
                Options(artifact_ratio=0.0, cell_size=0.5)):
        with pytest.raises(AssertionError):
            stl_to_sdf(stl, options=bad, out_dir=str(tmp_path))


def test_vti_roundtrip(tmp_path):
    g = MeshGrid.Grid(np.array([-1.0, 0.0, 2.0]), np.array([1.0, 0.5, 2.5]), 17, 3)
    v = np.random.default_rng(0).normal(size=g.ngp)
    f = exportSdfToVTI(str(tmp_path / "a"), g, v, "distance")
    r, dims, origin, spacing, name = read_vti(f)
    assert np.array_equal(r, v) and dims == tuple(int(n) + 1 for n in g.N)
    assert origin == list(g.AABB_min) and spacing == [g.cell_size] * 3 and name == "distance"
    fine = np.arange(int(np.prod(g.N * 2 + 1)), dtype=np.float32)
    f2 = exportSdfToVTI(str(tmp_path / "b.vti"), g, fine, "distance", smooth=2)
    r2, dims2, _, sp2, _ = read_vti(f2)
    assert np.array_equal(r2, fine) and sp2 == [g.cell_size / 2] * 3
    with pytest.raises(ValueError):
        exportSdfToVTI(str(tmp_path / "c"), g, v[:-1], "distance")


def test_stl_write_read_roundtrip(tmp_path):
    X, IEN = synthetic.icosphere(1.0, 2)
    p = str(tmp_path / "ico.stl")
    synthetic.write_stl_binary(p, X, IEN)
    X2, I2 = import_stl(p)
    assert X2.shape[1] == X.shape[1] and I2.shape == IEN.shape
    # same triangles after float32 rounding (vertex numbering by first use)
    t1 = np.sort(synthetic.to_float32(X)[:, IEN - 1].reshape(9, -1), axis=0)
    t2 = np.sort(X2[:, I2 - 1].reshape(9, -1), axis=0)
    assert np.array_equal(np.sort(t1, axis=1), np.sort(t2, axis=1))


def test_fine_grid_of_rbf_smoothing():
    g = MeshGrid.Grid(np.zeros(3), np.array([1.0, 0.7, 0.3]), 20, 3)
    fg = MeshGrid.Grid(g.AABB_min, g.AABB_max, int(max(g.N)) * 2, 0)
    assert abs(fg.cell_size - g.cell_size / 2) < 1e-15
    assert all(abs(int(a) - 2 * int(b)) <= 1 for a, b in zip(fg.N, g.N))


def test_fine_grid_guard_matches_the_export_error(tmp_path):
    """RBFs_smoothing returns Grid(AABB_min, AABB_max, maximum(N) * s, 0)
    (RBFs4Smoothing.jl:316), rebuilt with ceil rounding, beside
    prod(N * s + 1) values; where the two disagree the reference's fine
    export fails (ExportToVTI.jl:43-47).  The guard raises that same error
    before the solve, and only there."""
    from stl2sdf.SdfSmoothing import check_fine_grid, fine_grid
    mesh, g = _cases.artifacts_case(0.25)
    assert g.N.tolist() == [245, 85, 21]
    for s, fine_n in ((1, [245, 86, 21]), (2, [490, 171, 42])):
        fg = fine_grid(g, s)
        assert fg.N.tolist() == fine_n
        values = np.zeros(int(np.prod(g.N * s + 1)), np.float32)
        with pytest.raises(ValueError) as export_err:
            exportSdfToVTI(str(tmp_path / "f"), fg, values, "distance")
        with pytest.raises(ValueError) as guard_err:
            check_fine_grid(g, s)
        assert str(guard_err.value) == str(export_err.value)
    # the reference's own test grid (cell size 0.5) exports at s = 1 and 2
    _, g5 = _cases.artifacts_case(0.5)
    for s in (1, 2):
        check_fine_grid(g5, s)
        assert np.array_equal(fine_grid(g5, s).N, g5.N * s)


def test_select_projected_nodes():
    """SelectProjectedNodes (DataPostProcess.jl:21-58): keeps the points with
    sum |xp| > 1e-10 in order, mean and max of |X - Xp|, 0.0 / 0.0 when none."""
    from stl2sdf import MeshGrid
    from stl2sdf.DataExport import SelectProjectedNodes
    X = np.asfortranarray(np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0]]).T)
    mesh = MeshGrid.TriangularMesh(X, np.array([[1], [2], [3]]))
    grid = MeshGrid.Grid(np.zeros(3), np.ones(3), 1, 0)     # 2 x 2 x 2 points
    pts = MeshGrid.generateGridPoints(grid)
    xp = np.zeros((3, 8))
    xp[:, 1] = [3.0, 0.0, 0.0]          # |X - Xp| = 2
    xp[:, 5] = [1.0, 0.0, 5.0]          # |X - Xp| = 4
    xp[:, 6] = [1e-11, 0.0, 0.0]        # below the threshold: dropped
    Xs, Xp, mean_pd, max_pd = SelectProjectedNodes(mesh, grid, xp, pts)
    assert Xs.shape == (3, 2) and np.array_equal(Xs, pts[:, [1, 5]])
    assert np.array_equal(Xp, xp[:, [1, 5]])
    assert mean_pd == 3.0 and max_pd == 4.0
    Xs, Xp, mean_pd, max_pd = SelectProjectedNodes(mesh, grid, np.zeros((3, 8)), pts)
    assert Xs.shape == (3, 0) and mean_pd == 0.0 and max_pd == 0.0
    with pytest.raises(ValueError):
        SelectProjectedNodes(mesh, grid, np.zeros((3, 7)), pts)
