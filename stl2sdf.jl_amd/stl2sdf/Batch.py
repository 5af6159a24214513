"""Batched signed distance on the MI355X library (no counterpart in the
reference, which computes one mesh per call).

``SignedDistanceBatch(meshes, grids)`` uploads the meshes once and builds a
batch in the library: the meshes are dealt to worker contexts whose passes
overlap on the GPU, each run issues every mesh with no host wait and reads
all statuses once, and ``autotune`` picks the number of worker contexts,
the distance work-item length and the ray-stack depth by timing the batch.
Each result equals ``SignedDistances.signed_distance(mesh, grid)`` bit for
bit.
Device buffers come from the library's own HIP runtime
(``_lib.DeviceBuffer``), not from torch: torch ships its own HIP and HSA
runtimes, and a process that loads this library before importing torch ends
up with both copies, whose allocations the other copy does not know.
"""
import ctypes as C

import numpy as np

from . import _lib

INFO = ("meshes", "grid_points", "last_ms", "chunk_len", "ray_stack", "lanes", "issue_ms",
        "threads", "candidates")


class SignedDistanceBatch:
    def __init__(self, meshes, grids, device=0, autotune=True):
        if len(meshes) != len(grids) or not meshes:
            raise ValueError("need one grid per mesh")
        self.device = device
        self._X, self._I, self.out, self._ngp = [], [], [], []
        for mesh, grid in zip(meshes, grids):
            X = np.ascontiguousarray(np.asarray(mesh.X, np.float64).T)   # nnp x 3 = 3 x nnp col-major
            I = np.ascontiguousarray(np.asarray(mesh.IEN, np.int64).T)
            self._X.append(_lib.DeviceBuffer(X.nbytes, device).from_host(X))
            self._I.append(_lib.DeviceBuffer(I.nbytes, device).from_host(I))
            self.out.append(_lib.DeviceBuffer(8 * int(grid.ngp), device))
            self._ngp.append(int(grid.ngp))
        n = len(meshes)
        ptr = lambda bs: (C.c_void_p * n)(*[b.ptr.value for b in bs])
        nnp = np.array([m.X.shape[1] for m in meshes], np.int64)
        nel = np.array([m.IEN.shape[1] for m in meshes], np.int64)
        amin = np.concatenate([np.asarray(g.AABB_min, np.float64) for g in grids])
        h = np.array([g.cell_size for g in grids], np.float64)
        N = np.concatenate([np.asarray(g.N, np.int64) for g in grids])
        L = _lib.lib()
        c = _lib.ctx(device)
        self._h = C.c_void_p()
        _lib.check(L.sdf_batch_create(c, n, ptr(self._X), _lib.i64ptr(nnp), ptr(self._I),
                                      _lib.i64ptr(nel), _lib.dptr(amin), _lib.dptr(h),
                                      _lib.i64ptr(N), ptr(self.out), 1 if autotune else 0,
                                      C.byref(self._h)))

    def set_vertices(self, k, X):
        """Overwrite mesh k's vertices in place (same count), for replays."""
        X = np.ascontiguousarray(np.asarray(X, np.float64).T)
        self._X[k].from_host(X)

    def run(self):
        _lib.check(_lib.lib().sdf_batch_run(self._h))

    def info(self):
        out = np.zeros(256)
        n = C.c_int32()
        _lib.check(_lib.lib().sdf_batch_info(self._h, _lib.dptr(out), 256, C.byref(n)))
        d = dict(zip(INFO, out[:len(INFO)].tolist()))
        k = int(d["candidates"])
        d["tuning"] = [(int(out[len(INFO) + 2 * i]), float(out[len(INFO) + 2 * i + 1]))
                       for i in range(k)]
        return d

    def results(self):
        return [b.to_host(np.float64, n) for b, n in zip(self.out, self._ngp)]

    def close(self):
        if self._h:
            _lib.check(_lib.lib().sdf_batch_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
