"""C ABI checks that need no GPU: the library loads, exports every declared
symbol, and its host-only entry points match the oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle
from stl2sdf import MeshGrid, _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    inc = os.path.join(REPO, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(sdf_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared_symbols()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) == names


def test_abi_version():
    assert _lib.lib().sdf_abi_version() == 100


@pytest.mark.parametrize("xmin,xmax,nmax,margin", [
    ((-1, -1, -1), (1, 1, 1), 40, 2),
    ((-1, -1, -1), (1, 1, 1), 40, 3),
    ((0.1, -2.3, 5.0), (3.7, 0.4, 5.9), 77, 3),
    ((-0.3, 0.0, 1e-3), (0.2, 1e-3, 0.9), 13, 0),
])
def test_grid_setup_matches_oracle_and_host(xmin, xmax, nmax, margin):
    amin, amax, N = np.zeros(3), np.zeros(3), np.zeros(3, np.int64)
    h, ngp = C.c_double(), C.c_int64()
    _lib.check(_lib.lib().sdf_grid_setup(
        _lib.dptr(np.array(xmin, float)), _lib.dptr(np.array(xmax, float)), nmax, margin,
        _lib.dptr(amin), _lib.dptr(amax), _lib.i64ptr(N), C.byref(h), C.byref(ngp)))
    oa, ob, oN, oh, ong = oracle.grid(xmin, xmax, nmax, margin)
    g = MeshGrid.Grid(np.array(xmin, float), np.array(xmax, float), nmax, margin)
    assert np.array_equal(amin, oa) and np.array_equal(amax, ob) and np.array_equal(N, oN)
    assert h.value == oh and ngp.value == ong
    assert np.array_equal(g.AABB_min, oa) and np.array_equal(g.AABB_max, ob)
    assert list(g.N) == list(oN) and g.cell_size == oh and g.ngp == ong


def test_ray_directions_match_oracle_bitwise():
    d = np.zeros(108)
    _lib.check(_lib.lib().sdf_ray_directions(_lib.dptr(d)))
    assert np.array_equal(d.reshape(36, 3), oracle.ray_directions())


def test_lattice_points_match_oracle():
    g = MeshGrid.Grid(np.array([-0.7, 0.1, 2.0]), np.array([0.9, 1.3, 2.4]), 23, 3)
    P = MeshGrid.generateGridPoints(g)
    Q = oracle.grid_points(g.AABB_min, g.cell_size, g.N)
    assert np.array_equal(P, Q)


def test_ctx_create_without_device_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    h = C.c_void_p()
    st = _lib.lib().sdf_ctx_create(0, C.byref(h))
    assert st != 0 and h.value is None
    assert _lib.lib().sdf_last_error()


def test_null_arguments_rejected():
    st = _lib.lib().sdf_grid_setup(None, None, 1, 0, None, None, None, None, None)
    assert st == -1
    assert b"NULL" in _lib.lib().sdf_last_error()
