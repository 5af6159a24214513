"""stl_to_sdf and Options, mirroring src/stl_to_sdf.jl:21-147.

The orchestration stays on the host as in the reference; distances, signs,
the combine step and artifact removal run in one libsdf_mi355x.so call and
the RBF smoothing reads that call's device buffer, so the coarse field
crosses PCIe once (for its export) and the fine one once (for its own)."""
import os

import numpy as np

from . import SignedDistances
from .DataExport import exportSdfToVTI
from .DataImport import import_stl
from .MeshGrid import TriangularMesh, noninteractive_sdf_grid_setup
from .SdfSmoothing import RBFs_smoothing_device, check_fine_grid


class Options:
    """Options(; smoothing_method=None, grid_refinement=1, cell_size=None,
    remove_artifacts=False, artifact_ratio=0.01) -- stl_to_sdf.jl:21-43.
    smoothing_method is None, "interpolation" or "approximation" (Julia
    symbols :interpolation / :approximation)."""

    def __init__(self, smoothing_method=None, grid_refinement=1, cell_size=None,
                 remove_artifacts=False, artifact_ratio=0.01):
        self.smoothing_method = smoothing_method
        self.grid_refinement = grid_refinement
        self.cell_size = cell_size
        self.remove_artifacts = remove_artifacts
        self.artifact_ratio = artifact_ratio


def stl_to_sdf(stl_filename, options=None, out_dir=".", device=0):
    """stl_to_sdf(stl_filename; options) -- stl_to_sdf.jl:75-147.

    Writes <base>_sdf.vti (and <base>_fine_sdf.vti when smoothing) into
    out_dir and returns ().  The interactive grid setup of the reference
    (cell_size = nothing, which reads stdin) is not offered: cell_size must be
    given."""
    o = options if options is not None else Options()
    # input validation (:77-81)
    if o.smoothing_method is not None:
        assert o.smoothing_method in ("interpolation", "approximation"), \
            "smoothing_method must be :interpolation, :approximation, or nothing"
    assert o.grid_refinement in (1, 2), "grid_refinement must be 1 or 2"
    assert 0.0 < o.artifact_ratio < 1.0, "artifact_ratio must be in range (0, 1)"
    if o.cell_size is None:
        raise ValueError("cell_size is required (the interactive grid setup reads stdin)")
    base = os.path.splitext(os.path.basename(stl_filename))[0]
    X, IEN = import_stl(stl_filename)
    mesh = TriangularMesh(X, IEN)
    grid = noninteractive_sdf_grid_setup(mesh, o.cell_size)
    # distances -> signs -> product -> artifact removal in one device call;
    # the field crosses PCIe once, for the export
    d_sdf, _ = SignedDistances.signed_fused_device(
        mesh, grid, o.remove_artifacts, o.artifact_ratio, device=device)
    sdf = d_sdf.to_host(np.float64, grid.ngp)
    exportSdfToVTI(os.path.join(out_dir, f"{base}_sdf.vti"), grid, sdf, "distance")
    if o.smoothing_method is not None:
        # the reference fails in the fine export when the rebuilt fine grid
        # does not hold prod(N * s + 1) points: fail before the solve
        check_fine_grid(grid, o.grid_refinement)
        d_fine, fgrid = RBFs_smoothing_device(d_sdf, grid, o.smoothing_method == "interpolation",
                                              o.grid_refinement, device=device)
        fine = d_fine.to_host(np.float32, int(np.prod(np.asarray(grid.N) * o.grid_refinement + 1)))
        exportSdfToVTI(os.path.join(out_dir, f"{base}_fine_sdf.vti"), fgrid, fine, "distance")
    return ()
