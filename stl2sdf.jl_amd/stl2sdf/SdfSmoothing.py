"""SdfSmoothing entry point on the MI355X library
(src/SdfSmoothing/SdfSmoothing.jl:3, RBFs4Smoothing.jl:275-320)."""
import ctypes as C

import numpy as np

from . import _lib
from .MeshGrid import Grid

# sqrt(eps(Float32)): IterativeSolvers' default relative tolerance for cg
DEFAULT_RELTOL = 3.4526698300124393e-4


def RBFs_smoothing(dist, my_grid, Is_interpolation, smooth, threshold=1e-3,
                   reltol=DEFAULT_RELTOL, device=0, info=None):
    """RBFs_smoothing(dist, grid, is_interpolation, smooth, threshold=1e-3)
    -> (fine::float32 vector, Grid).

    Gaussian exp(-(r/h)^2) kernel truncated where it falls to ``threshold``;
    interpolation solves K w = Float32.(dist) to ||b - K w|| <= reltol ||b||
    (FP32 CG with FP64 residual refinement); approximation uses the values as
    weights.  The fine field has prod(N*smooth + 1) points, i fastest.  The
    returned grid is Grid(AABB_min, AABB_max, maximum(N)*smooth, 0) as in the
    reference.  ``info`` (a dict) receives the CG iteration count and the
    final relative residual."""
    d = np.ascontiguousarray(dist, dtype=np.float64)
    if d.ndim != 1 or d.size != my_grid.ngp:
        raise ValueError("dist must hold one value per grid point")
    N = _lib.i64_3(my_grid.N)
    nf = int(np.prod(N * int(smooth) + 1))
    fine = np.empty(nf, dtype=np.float32)
    it = C.c_int32(0)
    rel = C.c_double(0.0)
    _lib.check(_lib.lib().sdf_rbf_smoothing(
        _lib.ctx(device), _lib.dptr(d), _lib.i64ptr(N), float(my_grid.cell_size),
        1 if Is_interpolation else 0, int(smooth), float(threshold), float(reltol),
        fine.ctypes.data_as(C.POINTER(C.c_float)), C.byref(it), C.byref(rel)))
    if info is not None:
        info["cg_iterations"] = int(it.value)
        info["rel_residual"] = float(rel.value)
    return fine, fine_grid(my_grid, smooth)


def fine_grid(my_grid, smooth):
    """The grid RBFs_smoothing returns: Grid(AABB_min, AABB_max,
    maximum(N) * smooth, 0) (RBFs4Smoothing.jl:316)."""
    return Grid(my_grid.AABB_min, my_grid.AABB_max, int(max(my_grid.N)) * int(smooth), 0)


def check_fine_grid(my_grid, smooth):
    """The fine field has prod(N * smooth + 1) values, but the grid returned
    with it is rebuilt from the bounding box with ceil rounding, whose N + 1
    can differ; exportSdfToVTI then fails on the length (ExportToVTI.jl:43-47).
    Raises that error up front, so the smoothing is not computed for a
    result that cannot be exported."""
    nf = int(np.prod(np.asarray(my_grid.N, np.int64) * int(smooth) + 1))
    dims = int(np.prod(np.asarray(fine_grid(my_grid, smooth).N, np.int64) + 1))
    if nf != dims:
        raise ValueError(f"Values vector length ({nf}) doesn't match grid dimensions ({dims}).")


def RBFs_smoothing_device(d_dist, my_grid, Is_interpolation, smooth, threshold=1e-3,
                          reltol=DEFAULT_RELTOL, device=0, info=None):
    """RBFs_smoothing on a device buffer of grid.ngp doubles (for example the
    one SignedDistances.signed_fused_device returns).  Returns (buffer, Grid):
    the fine field, prod(N*smooth + 1) float32 values, stays on the device."""
    N = _lib.i64_3(my_grid.N)
    nf = int(np.prod(N * int(smooth) + 1))
    out = _lib.DeviceBuffer(4 * nf, device)
    it = C.c_int32(0)
    rel = C.c_double(0.0)
    _lib.check(_lib.lib().sdf_rbf_smoothing_device(
        _lib.ctx(device), d_dist.ptr, _lib.i64ptr(N), float(my_grid.cell_size),
        1 if Is_interpolation else 0, int(smooth), float(threshold), float(reltol), out.ptr,
        C.byref(it), C.byref(rel)))
    if info is not None:
        info["cg_iterations"] = int(it.value)
        info["rel_residual"] = float(rel.value)
    return out, fine_grid(my_grid, smooth)
