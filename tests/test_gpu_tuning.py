"""Launch tunables (sdf_ctx_tune) change only how the work is launched: every
setting gives the default's signed distances bit for bit."""
import ctypes as C

import numpy as np
import pytest

from stl2sdf import SignedDistances, _lib
from tests import _cases

pytestmark = pytest.mark.gpu


def _run(h, mesh, grid):
    X = np.asfortranarray(mesh.X, np.float64)
    IEN = np.asfortranarray(mesh.IEN, np.int64)
    out = np.empty(int(grid.ngp))
    _lib.check(_lib.lib().sdf_signed_distance(
        h, _lib.dptr(X), X.shape[1], _lib.i64ptr(IEN), IEN.shape[1],
        _lib.dptr(_lib.f64_3(grid.AABB_min)), float(grid.cell_size),
        _lib.i64ptr(_lib.i64_3(grid.N)), _lib.dptr(out)))
    return out


@pytest.mark.parametrize("name,value", [("chunk_len", 16), ("chunk_len", 37),
                                        ("ray_stack", 2048), ("dist_per_cu", 0),
                                        ("dist_per_cu", 1), ("dist_per_cu", 16),
                                        ("fused_per_cu", 4), ("fused_per_cu", 32),
                                        ("lines_early", 1), ("lines_early", 2),
                                        ("line_per_cu", 1), ("lbvh_early", 0),
                                        ("lbvh_early", 1), ("lbvh_early", 2),
                                        ("lbvh_early", 3), ("sign_split", 0), ("wind_ahead", 1),
                                        ("same_per_cu", 4)])
def test_tunable_keeps_results(name, value):
    mesh, grid = _cases.star_case(n=60, n_max=45)
    ref = SignedDistances.signed_distance(mesh, grid)
    L = _lib.lib()
    h = C.c_void_p()
    _lib.check(L.sdf_ctx_create(0, C.byref(h)))
    try:
        _lib.check(L.sdf_ctx_tune(h, name.encode(), value))
        print(f"case: {name}={value} nel={mesh.nel} ngp={grid.ngp}", flush=True)
        got = _run(h, mesh, grid)
        assert np.array_equal(got, ref)
        assert np.array_equal(_run(h, mesh, grid), ref)   # second call: recorded capacities
    finally:
        _lib.check(L.sdf_ctx_destroy(h))


def test_tunable_errors():
    L = _lib.lib()
    c = _lib.ctx(0)
    for name, value in (("chunk_len", 0), ("chunk_len", 65), ("ray_stack", 512),
                        ("dist_per_cu", -1), ("lines_early", 3), ("lbvh_early", 4), ("sign_split", 2), ("wind_ahead", 2), ("same_per_cu", 65),
                        ("no_such_knob", 1)):
        with pytest.raises(_lib.SdfError):
            _lib.check(L.sdf_ctx_tune(c, name.encode(), value))
