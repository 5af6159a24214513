"""Learned search radius on the MI355X library (no counterpart in the reference).

A per-mesh coordinate MLP (16 Fourier frequencies, three hidden layers of
128, ReLU; BF16 weights and activations, FP32 accumulation on the matrix
cores) is fitted to exact distances computed by the plain path on a coarse
lattice, then proposes a search radius for every fine-grid node.  The
distance loop visits only triangles within that radius; a node whose radius
holds no triangle is settled by an exact LBVH query with the same
arithmetic, so
``LearnedRadius.signed_distance`` returns exactly what
``SignedDistances.signed_distance`` returns; ``stats()`` reports how much
work the network saved and how often it fell back.

Low-level network calls (``mlp_*``) expose the forward pass, the loss
gradient and Adam training on caller-given samples.
"""
import ctypes as C

import numpy as np

from . import _lib

NPARAM = 49665
OFFSETS = {"W1": 0, "b1": 16384, "W2": 16512, "b2": 32896, "W3": 33024,
           "b3": 49408, "w4": 49536, "b4": 49664}

_fp = C.POINTER(C.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u(u):
    u = _f32(u).reshape(-1, 3)
    return u, u.shape[0]


def mlp_set_params(params, device=0):
    p = _f32(params)
    if p.size != NPARAM:
        raise ValueError(f"params must hold {NPARAM} floats")
    _lib.check(_lib.lib().sdf_mlp_set_params(_lib.ctx(device), p.ctypes.data_as(_fp)))


def mlp_get_params(device=0):
    p = np.empty(NPARAM, np.float32)
    _lib.check(_lib.lib().sdf_mlp_get_params(_lib.ctx(device), p.ctypes.data_as(_fp)))
    return p


def mlp_forward(u, device=0):
    """Network output for normalised points u (n x 3)."""
    u, n = _u(u)
    y = np.empty(n, np.float32)
    _lib.check(_lib.lib().sdf_mlp_forward(_lib.ctx(device), u.ctypes.data_as(_fp), n,
                                          y.ctypes.data_as(_fp)))
    return y


def mlp_loss_grad(u, t, device=0):
    """(mean squared error, gradient over all parameters)."""
    u, n = _u(u)
    t = _f32(t)
    g = np.empty(NPARAM, np.float32)
    loss = C.c_double()
    _lib.check(_lib.lib().sdf_mlp_loss_grad(_lib.ctx(device), u.ctypes.data_as(_fp),
                                            t.ctypes.data_as(_fp), n, g.ctypes.data_as(_fp),
                                            C.byref(loss)))
    return loss.value, g


def mlp_train(u, t, steps, batch, lr, seed=0, device=0):
    """Adam training on (u, t); returns the per-step batch losses."""
    u, n = _u(u)
    t = _f32(t)
    hist = np.empty(int(steps), np.float32)
    _lib.check(_lib.lib().sdf_mlp_train(_lib.ctx(device), u.ctypes.data_as(_fp),
                                        t.ctypes.data_as(_fp), n, int(steps), int(batch),
                                        float(lr), int(seed), hist.ctypes.data_as(_fp)))
    return hist


FIT_STATS = ("loss", "mae_band", "delta", "samples", "samples_band", "train_ms",
             "coarse_ms", "hc", "steps")
RUN_STATS = ("fallback_nodes", "evals", "tile_pairs", "lane_pairs_radius",
             "lane_pairs_h", "delta", "hc", "evals_plain", "network_skipped")


class LearnedRadius:
    """Learned-radius signed distance for one mesh on one device."""

    def __init__(self, device=0):
        self.device = device
        self.fit_stats = None

    def _set_mesh(self, mesh):
        X = np.asfortranarray(mesh.X, dtype=np.float64)
        IEN = np.asfortranarray(mesh.IEN, dtype=np.int64)
        _lib.check(_lib.lib().sdf_mesh_set(_lib.ctx(self.device), _lib.dptr(X), X.shape[1],
                                           _lib.i64ptr(IEN), IEN.shape[1]))
        return X, IEN

    def fit(self, mesh, grid, n_coarse=128, steps=1500, batch=4096, lr=2e-3, seed=1,
            quantile=0.99):
        """Fit the network to `mesh` for the fine grid `grid`; returns the fit stats."""
        self._set_mesh(mesh)
        st = np.zeros(len(FIT_STATS))
        _lib.check(_lib.lib().sdf_learned_fit(
            _lib.ctx(self.device), _lib.dptr(_lib.f64_3(grid.AABB_min)),
            float(grid.cell_size), _lib.i64ptr(_lib.i64_3(grid.N)), int(n_coarse),
            int(steps), int(batch), float(lr), int(seed), float(quantile), _lib.dptr(st),
            len(FIT_STATS)))
        self.fit_stats = dict(zip(FIT_STATS, st.tolist()))
        return self.fit_stats

    def signed_distance(self, mesh, grid):
        """Same result as SignedDistances.signed_distance(mesh, grid)."""
        X = np.asfortranarray(mesh.X, dtype=np.float64)
        IEN = np.asfortranarray(mesh.IEN, dtype=np.int64)
        out = np.empty(int(grid.ngp), np.float64)
        _lib.check(_lib.lib().sdf_learned_signed_distance(
            _lib.ctx(self.device), _lib.dptr(X), X.shape[1], _lib.i64ptr(IEN), IEN.shape[1],
            _lib.dptr(_lib.f64_3(grid.AABB_min)), float(grid.cell_size),
            _lib.i64ptr(_lib.i64_3(grid.N)), _lib.dptr(out)))
        return out

    def stats(self):
        """Counters of the last learned call."""
        out = np.zeros(len(RUN_STATS))
        n = C.c_int32()
        _lib.check(_lib.lib().sdf_learned_stats(_lib.ctx(self.device), _lib.dptr(out),
                                                len(RUN_STATS), C.byref(n)))
        return dict(zip(RUN_STATS, out.tolist()))
