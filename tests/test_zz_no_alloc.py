"""The allocation-free artifact removal (sdf_remove_artifacts_device_noalloc)
against sdf_remove_artifacts_device, through the C ABI: identical fields
and flip counts on two calls in a row, the second on a smaller grid (the
static scratch keeps the first call's contents), and a grid above the
scratch's maximum rejected before any device work."""
import ctypes as C

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle
from stl2sdf import _lib

pytestmark = pytest.mark.gpu

MAX_NODES = 8388608  # SDF_NOALLOC_MAX_NODES in include/sdf_mi355x.h


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    f = ndimage.uniform_filter(rng.normal(size=shape), 3)
    return np.ascontiguousarray(f.ravel())


def _hip():
    """The HIP runtime the library itself is linked against (already loaded)."""
    _lib.lib()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _run(fn, c, sdf, N, ratio):
    """One call on a device copy of sdf; returns the field and the count."""
    hip = _hip()
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), sdf.nbytes) == 0
    try:
        assert hip.hipMemcpy(p, sdf.ctypes.data, sdf.nbytes, 1) == 0   # host to device
        n = C.c_int64(-1)
        _lib.check(fn(c, p, _lib.i64ptr(N), 0.0, ratio, C.byref(n)))
        out = np.empty_like(sdf)
        assert hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
    finally:
        hip.hipFree(p)
    return out, n.value


def test_noalloc_equals_allocating_entry_point():
    L = _lib.lib()
    c = _lib.ctx(0)
    for shape, seed, ratio in (((72, 64, 56), 21, 0.02), ((30, 26, 20), 22, 0.05)):
        sdf = _field(shape, seed)
        N = np.array([shape[2] - 1, shape[1] - 1, shape[0] - 1], np.int64)
        print(f"case: noalloc shape={shape} seed={seed} ratio={ratio}", flush=True)
        ref, nref = _run(L.sdf_remove_artifacts_device, c, sdf, N, ratio)
        got, ngot = _run(L.sdf_remove_artifacts_device_noalloc, c, sdf, N, ratio)
        assert ngot == nref and nref > 0
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
        orc, norc = oracle.remove_artifacts(sdf, N, 0.0, ratio)
        assert norc == nref and np.array_equal(got.view(np.uint64), orc.view(np.uint64))


def test_noalloc_rejects_grids_above_the_scratch():
    L = _lib.lib()
    c = _lib.ctx(0)
    N = np.array([255, 255, 128], np.int64)   # 256 * 256 * 129 nodes > 8388608
    assert np.prod(N + 1) > MAX_NODES
    n = C.c_int64(-1)
    st = L.sdf_remove_artifacts_device_noalloc(c, None, _lib.i64ptr(N), 0.0, 0.01, C.byref(n))
    assert st == -3 and n.value == -1    # SDF_ERR_LIMIT, nothing computed
    assert b"at most" in L.sdf_last_error()
