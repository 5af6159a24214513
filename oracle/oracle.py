"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product library never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)

_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_grid.argtypes = [_dp, _dp, C.c_int64, C.c_int64, _dp, _dp, _i64p, _dp, _i64p]
        L.orc_grid_points.argtypes = [_dp, C.c_double, _i64p, _dp]
        L.orc_ray_directions.argtypes = [_dp]
        L.orc_ray_triangle.argtypes = [_dp, _dp, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_ray_triangle.restype = C.c_int
        L.orc_dist_exact.argtypes = [_dp, C.c_int64, _i64p, C.c_int64, _dp, C.c_int64,
                                     C.c_double, _dp, _dp, C.c_int]
        L.orc_dist_exact_bvh.argtypes = [_dp, C.c_int64, _i64p, C.c_int64, _dp, C.c_int64,
                                         C.c_double, _dp, C.c_int]
        L.orc_dist_reference.argtypes = [_dp, C.c_int64, _i64p, C.c_int64, _dp, C.c_int64,
                                         C.c_double, _dp, C.c_int]
        L.orc_signs.argtypes = [_dp, C.c_int64, _i64p, C.c_int64, _dp, C.c_int64, _dp,
                                C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp,
                                _i32p, _i32p, C.c_int]
        L.orc_ray_inside.argtypes = [_dp, _i64p, C.c_int64, _dp, _dp]
        L.orc_ray_inside.restype = C.c_int
        L.orc_winding_sign.argtypes = [_dp, _i64p, C.c_int64, _dp]
        L.orc_winding_sign.restype = C.c_double
        L.orc_components.argtypes = [_dp, _i64p, C.c_double, _i64p]
        L.orc_components.restype = C.c_int64
        L.orc_remove_artifacts.argtypes = [_dp, _i64p, C.c_double, C.c_double]
        L.orc_remove_artifacts.restype = C.c_int64
        L.orc_rbf_weights.argtypes = [_fp, _i64p, C.c_double, C.c_double, C.c_double, C.c_int,
                                      _fp, _dp]
        L.orc_rbf_weights.restype = C.c_int
        L.orc_rbf_fine.argtypes = [_fp, _dp, _dp, _i64p, C.c_double, C.c_int, C.c_double, _fp,
                                   C.c_int]
        L.orc_read_stl_binary.argtypes = [C.c_char_p, _dp, _i64p, _i64p, _i64p]
        L.orc_process_stl_mesh.argtypes = [_dp, C.c_int64, _dp, _i64p]
        L.orc_process_stl_mesh.restype = C.c_int64
        L.orc_read_stl_binary.restype = C.c_int
        L.orc_bf16.argtypes = [C.c_float]
        L.orc_bf16.restype = C.c_float
        L.orc_mlp_freqs.argtypes = [_fp]
        L.orc_mlp_features.argtypes = [_fp, C.c_int, _dp]
        L.orc_mlp_forward.argtypes = [_dp, _fp, C.c_int64, C.c_int, _dp]
        L.orc_mlp_loss_grad.argtypes = [_dp, _fp, _fp, C.c_int64, C.c_int, _dp]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_mlp_adam.argtypes = [_fp, _fp, _fp, _fp, C.c_int64, C.c_int, C.c_float]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_i64p)


def _mesh(X, IEN):
    return (np.asfortranarray(X, dtype=np.float64), np.asfortranarray(IEN, dtype=np.int64))


def grid(xmin, xmax, n_max, margin=3):
    amin, amax, N = np.zeros(3), np.zeros(3), np.zeros(3, np.int64)
    h, ngp = C.c_double(), C.c_int64()
    lib().orc_grid(_d(np.asarray(xmin, np.float64)), _d(np.asarray(xmax, np.float64)),
                   int(n_max), int(margin), _d(amin), _d(amax), _i(N), C.byref(h), C.byref(ngp))
    return amin, amax, N, h.value, ngp.value


def grid_points(amin, h, N):
    N = np.asarray(N, np.int64)
    P = np.empty((3, int(np.prod(N + 1))), order="F")
    lib().orc_grid_points(_d(np.asarray(amin, np.float64)), float(h), _i(N), _d(P))
    return P


def ray_directions():
    d = np.zeros(108)
    lib().orc_ray_directions(_d(d))
    return d.reshape(36, 3)


def ray_triangle(o, d, v1, v2, v3, eps=1e-8):
    t = C.c_double()
    args = [np.ascontiguousarray(a, np.float64) for a in (o, d, v1, v2, v3)]
    hit = lib().orc_ray_triangle(*[_d(a) for a in args], eps, C.byref(t))
    return t.value if hit else None


def dist_exact(X, IEN, pts, h, threads=0, want_xp=False):
    X, IEN = _mesh(X, IEN)
    P = np.asfortranarray(pts, np.float64)
    out = np.empty(P.shape[1])
    xp = np.empty((3, P.shape[1]), order="F") if want_xp else None
    lib().orc_dist_exact(_d(X), X.shape[1], _i(IEN), IEN.shape[1], _d(P), P.shape[1], float(h),
                         _d(out), _d(xp) if want_xp else None, int(threads))
    return (out, xp) if want_xp else out


def dist_exact_bvh(X, IEN, pts, h, threads=0):
    """dist_exact's min(d, h) through a bounding-volume hierarchy (same bits)."""
    X, IEN = _mesh(X, IEN)
    P = np.asfortranarray(pts, np.float64)
    out = np.empty(P.shape[1])
    lib().orc_dist_exact_bvh(_d(X), X.shape[1], _i(IEN), IEN.shape[1], _d(P), P.shape[1],
                             float(h), _d(out), int(threads))
    return out


def dist_reference(X, IEN, pts, h, threads=0):
    X, IEN = _mesh(X, IEN)
    P = np.asfortranarray(pts, np.float64)
    out = np.empty(P.shape[1])
    lib().orc_dist_reference(_d(X), X.shape[1], _i(IEN), IEN.shape[1], _d(P), P.shape[1],
                             float(h), _d(out), int(threads))
    return out


def signs(X, IEN, pts, min_rays=6, max_rays=30, conf_thr=0.6, use_winding=True,
          use_bvh=True, threads=0, diagnostics=False):
    X, IEN = _mesh(X, IEN)
    P = np.asfortranarray(pts, np.float64)
    n = P.shape[1]
    s, c = np.empty(n), np.empty(n)
    ru, wu = np.empty(n, np.int32), np.empty(n, np.int32)
    d = ray_directions()
    lib().orc_signs(_d(X), X.shape[1], _i(IEN), IEN.shape[1], _d(P), n, _d(np.ascontiguousarray(d)),
                    int(min_rays), int(max_rays), float(conf_thr), int(bool(use_winding)),
                    int(bool(use_bvh)), _d(s), _d(c), ru.ctypes.data_as(_i32p),
                    wu.ctypes.data_as(_i32p), int(threads))
    if diagnostics:
        return s, c, ru, wu
    return s, c


def ray_inside(X, IEN, o, d):
    X, IEN = _mesh(X, IEN)
    return bool(lib().orc_ray_inside(_d(X), _i(IEN), IEN.shape[1],
                                     _d(np.ascontiguousarray(o, np.float64)),
                                     _d(np.ascontiguousarray(d, np.float64))))


def winding_sign(X, IEN, p):
    X, IEN = _mesh(X, IEN)
    return lib().orc_winding_sign(_d(X), _i(IEN), IEN.shape[1],
                                  _d(np.ascontiguousarray(p, np.float64)))


def components(sdf, N, threshold=0.0):
    sdf = np.ascontiguousarray(sdf, np.float64)
    lab = np.empty(sdf.shape[0], np.int64)
    nc = lib().orc_components(_d(sdf), _i(np.asarray(N, np.int64)), float(threshold), _i(lab))
    return lab, nc


def remove_artifacts(sdf, N, threshold=0.0, ratio=0.01):
    s = np.array(sdf, np.float64, copy=True)
    n = lib().orc_remove_artifacts(_d(s), _i(np.asarray(N, np.int64)), float(threshold),
                                   float(ratio))
    return s, n


def rbf_weights(coarse, N, h, threshold=1e-3, reltol=3.4526698e-4, maxiter=0):
    c = np.ascontiguousarray(coarse, np.float32)
    w = np.empty_like(c)
    res = C.c_double()
    it = lib().orc_rbf_weights(c.ctypes.data_as(_fp), _i(np.asarray(N, np.int64)), float(h),
                               float(threshold), float(reltol), int(maxiter),
                               w.ctypes.data_as(_fp), C.byref(res))
    return w, it, res.value


def rbf_fine(weights, amin, amax, N, h, smooth, threshold=1e-3, threads=0):
    N = np.asarray(N, np.int64)
    nf = int(np.prod(N * smooth + 1))
    out = np.empty(nf, np.float32)
    w = np.ascontiguousarray(weights, np.float32)
    lib().orc_rbf_fine(w.ctypes.data_as(_fp), _d(np.asarray(amin, np.float64)),
                       _d(np.asarray(amax, np.float64)), _i(N), float(h), int(smooth),
                       float(threshold), out.ctypes.data_as(_fp), int(threads))
    return out


def read_stl_binary(path):
    nnp, nel = C.c_int64(0), C.c_int64(0)
    rc = lib().orc_read_stl_binary(path.encode(), None, C.byref(nnp), None, C.byref(nel))
    if rc != 0:
        raise IOError(f"orc_read_stl_binary failed: {rc}")
    X = np.empty((3, nnp.value), order="F")
    IEN = np.empty((3, nel.value), np.int64, order="F")
    rc = lib().orc_read_stl_binary(path.encode(), _d(X), C.byref(nnp), _i(IEN), C.byref(nel))
    if rc != 0:
        raise IOError(f"orc_read_stl_binary failed: {rc}")
    return X, IEN


def process_stl_mesh(corners):
    """process_stl_mesh on triangle corners (n*3 x 3): (X 3 x nnp, IEN 3 x nel)."""
    V = np.ascontiguousarray(np.asarray(corners, np.float64).reshape(-1, 3))
    nv = V.shape[0]
    X = np.empty(3 * max(nv, 1))
    IEN = np.empty(max(nv, 1), np.int64)
    nu = lib().orc_process_stl_mesh(_d(V), nv, _d(X), _i(IEN))
    return (np.asfortranarray(X[:3 * nu].reshape(nu, 3).T),
            np.asfortranarray(IEN[:nv].reshape(-1, 3).T))


# ---------------------------------------------------------------- MLP
MLP_NPARAM = 49665
MLP_EXACT, MLP_BF16 = 0, 1
MLP_OFFSETS = {"W1": 0, "b1": 16384, "W2": 16512, "b2": 32896, "W3": 33024,
               "b3": 49408, "w4": 49536, "b4": 49664}


def bf16(x):
    return float(lib().orc_bf16(float(x)))


def mlp_freqs():
    f = np.empty(16, np.float32)
    lib().orc_mlp_freqs(f.ctypes.data_as(_fp))
    return f


def _u(u):
    u = np.ascontiguousarray(np.asarray(u, np.float32).reshape(-1, 3))
    return u, u.shape[0]


def mlp_features(u, mode=MLP_BF16):
    u = np.ascontiguousarray(np.asarray(u, np.float32).reshape(3))
    phi = np.empty(128)
    lib().orc_mlp_features(u.ctypes.data_as(_fp), int(mode), _d(phi))
    return phi


def mlp_forward(params, u, mode=MLP_BF16):
    """Network output for points u (n x 3 normalised coordinates)."""
    p = np.ascontiguousarray(params, np.float64)
    u, n = _u(u)
    y = np.empty(n)
    lib().orc_mlp_forward(_d(p), u.ctypes.data_as(_fp), n, int(mode), _d(y))
    return y


def mlp_loss_grad(params, u, t, mode=MLP_BF16):
    p = np.ascontiguousarray(params, np.float64)
    u, n = _u(u)
    t = np.ascontiguousarray(t, np.float32)
    g = np.empty(MLP_NPARAM)
    loss = lib().orc_mlp_loss_grad(_d(p), u.ctypes.data_as(_fp), t.ctypes.data_as(_fp), n,
                                   int(mode), _d(g))
    return loss, g


def mlp_adam(params, m, v, grad, t, lr):
    """One FP32 Adam step; returns new (params, m, v)."""
    p = np.array(params, np.float32)
    m = np.array(m, np.float32)
    v = np.array(v, np.float32)
    g = np.ascontiguousarray(grad, np.float32)
    lib().orc_mlp_adam(p.ctypes.data_as(_fp), m.ctypes.data_as(_fp), v.ctypes.data_as(_fp),
                       g.ctypes.data_as(_fp), p.size, int(t), float(lr))
    return p, m, v
