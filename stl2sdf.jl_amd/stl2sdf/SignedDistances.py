"""SignedDistances entry points on the MI355X library
(src/SignedDistances/SignedDistances.jl:3-4).

Same names, arguments and results as the reference; the work runs in
libsdf_mi355x.so.  Julia's ``remove_sdf_artifacts!`` is spelled
``remove_sdf_artifacts_`` here and mutates its argument the same way.
"""
import ctypes as C

import numpy as np

from . import _lib
from .MeshGrid import generateGridPoints  # noqa: F401  (re-exported for callers)


def _mesh_args(mesh):
    X = np.asfortranarray(mesh.X, dtype=np.float64)
    IEN = np.asfortranarray(mesh.IEN, dtype=np.int64)
    return X, IEN


def _grid_args(grid):
    return _lib.f64_3(grid.AABB_min), float(grid.cell_size), _lib.i64_3(grid.N)


def evalDistancesOnTriMesh(mesh, grid, points=None, device=0, want_xp=True):
    """evalDistancesOnTriMesh(mesh, grid, points) -- dfOnTriangularMesh.jl:19-209.

    Returns (dist, xp).  dist[i] = min(exact distance, grid.cell_size);
    xp[:, i] is the exact closest point on the mesh (3 x npts, FP64) at any
    distance (the reference's is the closest among its candidate triangles,
    which equals it within ~2.5 cells).  want_xp=False skips xp and returns
    None for it.  points=None means the grid's lattice (the usual
    generateGridPoints(grid) input, which is also recognised when passed
    explicitly)."""
    X, IEN = _mesh_args(mesh)
    amin, h, N = _grid_args(grid)
    if points is None:
        npts = int(grid.ngp)
        pp = None
    else:
        P = _lib.as_points(points)
        npts = P.shape[1]
        pp = _lib.dptr(P)
    dist = np.empty(npts, dtype=np.float64)
    L = _lib.lib()
    if not want_xp:
        _lib.check(L.sdf_eval_distances(
            _lib.ctx(device), _lib.dptr(X), X.shape[1], _lib.i64ptr(IEN), IEN.shape[1],
            _lib.dptr(amin), h, _lib.i64ptr(N), pp, npts, _lib.dptr(dist)))
        return dist, None
    xp = np.empty((3, npts), dtype=np.float64, order="F")
    _lib.check(L.sdf_eval_distances_xp(
        _lib.ctx(device), _lib.dptr(X), X.shape[1], _lib.i64ptr(IEN), IEN.shape[1],
        _lib.dptr(amin), h, _lib.i64ptr(N), pp, npts, _lib.dptr(dist), _lib.dptr(xp)))
    return dist, xp


def raycast_sign_detection(mesh, grid, points=None, min_rays=6, max_rays=30,
                           confidence_threshold=0.6, use_winding_fallback=True,
                           device=0):
    """raycast_sign_detection -- RaycastingSignDetection.jl:337-404.

    Returns (signs, confidences); inside = +1."""
    X, IEN = _mesh_args(mesh)
    amin, h, N = _grid_args(grid)
    if points is None:
        npts = int(grid.ngp)
        pp = None
    else:
        P = _lib.as_points(points)
        npts = P.shape[1]
        pp = _lib.dptr(P)
    signs = np.empty(npts, dtype=np.float64)
    conf = np.empty(npts, dtype=np.float64)
    L = _lib.lib()
    _lib.check(L.sdf_raycast_signs(
        _lib.ctx(device), _lib.dptr(X), X.shape[1], _lib.i64ptr(IEN), IEN.shape[1],
        _lib.dptr(amin), h, _lib.i64ptr(N), pp, npts, int(min_rays), int(max_rays),
        float(confidence_threshold), 1 if use_winding_fallback else 0,
        _lib.dptr(signs), _lib.dptr(conf)))
    return signs, conf


def signed_distance(mesh, grid, device=0):
    """dists .* signs over the grid lattice in one library call
    (stl_to_sdf.jl:101-119 with the default ray options)."""
    X, IEN = _mesh_args(mesh)
    amin, h, N = _grid_args(grid)
    out = np.empty(int(grid.ngp), dtype=np.float64)
    L = _lib.lib()
    _lib.check(L.sdf_signed_distance(
        _lib.ctx(device), _lib.dptr(X), X.shape[1], _lib.i64ptr(IEN), IEN.shape[1],
        _lib.dptr(amin), h, _lib.i64ptr(N), _lib.dptr(out)))
    return out


def signed_fused_device(mesh, grid, remove_artifacts=False, min_component_ratio=0.01,
                        device=0):
    """The device part of stl_to_sdf (stl_to_sdf.jl:101-126): distances,
    signs (default ray options), their product and, with remove_artifacts,
    remove_sdf_artifacts!(sdf, grid; min_component_ratio) in one library
    call.  Returns (buffer, nflipped): the field stays in the device buffer
    (``_lib.DeviceBuffer``, grid.ngp doubles); only the mesh crosses PCIe."""
    X, IEN = _mesh_args(mesh)
    amin, h, N = _grid_args(grid)
    L = _lib.lib()
    c = _lib.ctx(device)
    _lib.check(L.sdf_mesh_set(c, _lib.dptr(X), X.shape[1], _lib.i64ptr(IEN), IEN.shape[1]))
    buf = _lib.DeviceBuffer(8 * int(grid.ngp), device)
    n = C.c_int64(0)
    _lib.check(L.sdf_signed_fused(c, _lib.dptr(amin), h, _lib.i64ptr(N),
                                  1 if remove_artifacts else 0, float(min_component_ratio),
                                  buf.ptr, C.byref(n)))
    return buf, int(n.value)


def _check_len(sdf_values, grid):
    v = np.asarray(sdf_values)
    if v.ndim != 1 or v.shape[0] != grid.ngp:
        # SdfArtifactRemoval.jl:144-146
        raise ValueError(f"SDF values length ({v.size}) doesn't match grid points ({grid.ngp})")


def remove_sdf_artifacts_(sdf_values, grid, threshold=0.0, min_component_ratio=0.01, device=0):
    """remove_sdf_artifacts!(sdf_values, grid; threshold, min_component_ratio)
    -- SdfArtifactRemoval.jl:134-248.  Modifies ``sdf_values`` (float64,
    contiguous) in place and returns the number of flipped nodes."""
    _check_len(sdf_values, grid)
    if not (isinstance(sdf_values, np.ndarray) and sdf_values.dtype == np.float64
            and sdf_values.flags.c_contiguous):
        raise TypeError("sdf_values must be a contiguous float64 numpy array (modified in place)")
    n = C.c_int64(0)
    _lib.check(_lib.lib().sdf_remove_artifacts(
        _lib.ctx(device), _lib.dptr(sdf_values), _lib.i64ptr(_lib.i64_3(grid.N)),
        float(threshold), float(min_component_ratio), C.byref(n)))
    return int(n.value)


def analyze_sdf_components(sdf_values, grid, threshold=0.0, device=0, labels=False):
    """analyze_sdf_components(sdf_values, grid; threshold) -> {root: size}
    -- SdfArtifactRemoval.jl:259-318.  Roots are canonical: the smallest
    1-based linear index of each component (the reference's roots depend on
    its union order).  labels=True also returns the per-node 1-based root
    (0 for exterior nodes)."""
    _check_len(sdf_values, grid)
    v = np.ascontiguousarray(sdf_values, dtype=np.float64)
    N = _lib.i64_3(grid.N)
    L = _lib.lib()
    c = _lib.ctx(device)
    ncomp = C.c_int64(0)
    lab = np.empty(v.size, np.int64) if labels else None
    _lib.check(L.sdf_analyze_components(c, _lib.dptr(v), _lib.i64ptr(N), float(threshold),
                                        _lib.i64ptr(lab) if labels else None, None, None, 0,
                                        C.byref(ncomp)))
    k = int(ncomp.value)
    roots = np.empty(max(k, 1), np.int64)
    sizes = np.empty(max(k, 1), np.int64)
    _lib.check(L.sdf_analyze_components(c, _lib.dptr(v), _lib.i64ptr(N), float(threshold), None,
                                        _lib.i64ptr(roots), _lib.i64ptr(sizes), k,
                                        C.byref(ncomp)))
    d = {int(r) + 1: int(s) for r, s in zip(roots[:k], sizes[:k])}
    if labels:
        return d, np.where(lab >= 0, lab + 1, 0)
    return d


def ray_directions():
    """The 36 prioritised directions (RaycastingSignDetection.jl:162-227)."""
    d = np.zeros(108, dtype=np.float64)
    _lib.check(_lib.lib().sdf_ray_directions(_lib.dptr(d)))
    return d.reshape(36, 3)
