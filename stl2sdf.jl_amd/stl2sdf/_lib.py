"""ctypes binding of libsdf_mi355x.so (the C ABI in include/sdf_mi355x.h).

The library is the only compute path: if it is missing or no GPU is
present, calls raise instead of falling back to anything else.
"""
import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libsdf_mi355x.so")

_lib = None
_lock = threading.Lock()
_ctx = {}

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_vp = C.c_void_p

# name -> (restype, argtypes); mirrors include/sdf_mi355x.h
SIGNATURES = {
    "sdf_abi_version": (C.c_int32, []),
    "sdf_last_error": (C.c_char_p, []),
    "sdf_ctx_create": (C.c_int32, [C.c_int32, C.POINTER(_vp)]),
    "sdf_ctx_destroy": (C.c_int32, [_vp]),
    "sdf_ctx_set_stream": (C.c_int32, [_vp, _vp]),
    "sdf_ctx_tune": (C.c_int32, [_vp, C.c_char_p, C.c_int64]),
    "sdf_ctx_synchronize": (C.c_int32, [_vp]),
    "sdf_ctx_enable_timing": (C.c_int32, [_vp, C.c_int32]),
    "sdf_ctx_stage_times": (C.c_int32, [_vp, C.c_char_p, C.c_int32, _dp, C.c_int32, _i32p]),
    "sdf_ctx_counters": (C.c_int32, [_vp, _i64p, C.c_int32, _i32p]),
    "sdf_ctx_vote_points": (C.c_int32, [_vp, _i64p, C.c_int64, _i64p]),
    "sdf_grid_setup": (C.c_int32, [_dp, _dp, C.c_int64, C.c_int64, _dp, _dp, _i64p, _dp, _i64p]),
    "sdf_eval_distances": (C.c_int32, [_vp, _dp, C.c_int64, _i64p, C.c_int64, _dp, C.c_double,
                                       _i64p, _dp, C.c_int64, _dp]),
    "sdf_eval_distances_xp": (C.c_int32, [_vp, _dp, C.c_int64, _i64p, C.c_int64, _dp,
                                          C.c_double, _i64p, _dp, C.c_int64, _dp, _dp]),
    "sdf_raycast_signs": (C.c_int32, [_vp, _dp, C.c_int64, _i64p, C.c_int64, _dp, C.c_double,
                                      _i64p, _dp, C.c_int64, C.c_int32, C.c_int32, C.c_double,
                                      C.c_int32, _dp, _dp]),
    "sdf_signed_distance": (C.c_int32, [_vp, _dp, C.c_int64, _i64p, C.c_int64, _dp, C.c_double,
                                        _i64p, _dp]),
    "sdf_mesh_set_device": (C.c_int32, [_vp, _vp, C.c_int64, _vp, C.c_int64]),
    "sdf_signed_distance_device": (C.c_int32, [_vp, _dp, C.c_double, _i64p, _vp]),
    "sdf_ray_directions": (C.c_int32, [_dp]),
    "sdf_import_stl": (C.c_int32, [C.c_char_p, _dp, _i64p, _i64p, _i64p]),
    "sdf_mesh_set": (C.c_int32, [_vp, _dp, C.c_int64, _i64p, C.c_int64]),
    "sdf_batch_create": (C.c_int32, [_vp, C.c_int32, C.POINTER(_vp), _i64p, C.POINTER(_vp),
                                     _i64p, _dp, _dp, _i64p, C.POINTER(_vp), C.c_int32,
                                     C.POINTER(_vp)]),
    "sdf_batch_run": (C.c_int32, [_vp]),
    "sdf_batch_info": (C.c_int32, [_vp, _dp, C.c_int32, _i32p]),
    "sdf_batch_destroy": (C.c_int32, [_vp]),
    "sdf_mlp_set_params": (C.c_int32, [_vp, _fp]),
    "sdf_mlp_get_params": (C.c_int32, [_vp, _fp]),
    "sdf_mlp_forward": (C.c_int32, [_vp, _fp, C.c_int64, _fp]),
    "sdf_mlp_loss_grad": (C.c_int32, [_vp, _fp, _fp, C.c_int64, _fp, _dp]),
    "sdf_mlp_train": (C.c_int32, [_vp, _fp, _fp, C.c_int64, C.c_int32, C.c_int32, C.c_float,
                                  C.c_uint64, _fp]),
    "sdf_learned_fit": (C.c_int32, [_vp, _dp, C.c_double, _i64p, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_float, C.c_uint64, C.c_double, _dp,
                                    C.c_int32]),
    "sdf_learned_signed_distance": (C.c_int32, [_vp, _dp, C.c_int64, _i64p, C.c_int64, _dp,
                                                C.c_double, _i64p, _dp]),
    "sdf_learned_signed_distance_device": (C.c_int32, [_vp, _dp, C.c_double, _i64p, _vp]),
    "sdf_learned_stats": (C.c_int32, [_vp, _dp, C.c_int32, _i32p]),
    "sdf_remove_artifacts": (C.c_int32, [_vp, _dp, _i64p, C.c_double, C.c_double, _i64p]),
    "sdf_remove_artifacts_device": (C.c_int32, [_vp, _vp, _i64p, C.c_double, C.c_double,
                                                _i64p]),
    "sdf_remove_artifacts_device_noalloc": (C.c_int32, [_vp, _vp, _i64p, C.c_double,
                                                        C.c_double, _i64p]),
    "sdf_analyze_components": (C.c_int32, [_vp, _dp, _i64p, C.c_double, _i64p, _i64p, _i64p,
                                           C.c_int64, _i64p]),
    "sdf_rbf_smoothing": (C.c_int32, [_vp, _dp, _i64p, C.c_double, C.c_int32, C.c_int32,
                                      C.c_double, C.c_double, _fp, _i32p, _dp]),
    "sdf_rbf_smoothing_device": (C.c_int32, [_vp, _vp, _i64p, C.c_double, C.c_int32,
                                             C.c_int32, C.c_double, C.c_double, _vp, _i32p,
                                             _dp]),
    "sdf_signed_fused": (C.c_int32, [_vp, _dp, C.c_double, _i64p, C.c_int32, C.c_double, _vp,
                                     _i64p]),
}


class SdfError(RuntimeError):
    """A non-zero status returned by the C ABI."""

    def __init__(self, status, msg):
        super().__init__(f"sdf status {status}: {msg}")
        self.status = status


def lib():
    """Load the shared library once; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} not found: build it with "
                    "`python stl2sdf.jl_amd/csrc/build.py` (hipcc, gfx950)")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = lib().sdf_last_error()
        raise SdfError(status, msg.decode() if msg else "")


def ctx(device=0):
    """Process-wide context for a device (created on first use)."""
    with _lock:
        c = _ctx.get(device)
    if c is None:
        h = _vp()
        check(lib().sdf_ctx_create(device, C.byref(h)))
        with _lock:
            _ctx[device] = h
        c = h
    return c


_hip = None


def hip():
    """The HIP runtime the library runs on (loaded with it), for device
    buffers handed to the ``_device`` entry points."""
    global _hip
    lib()
    with _lock:
        if _hip is None:
            H = C.CDLL("libamdhip64.so.7")
            H.hipSetDevice.argtypes = [C.c_int]
            H.hipMalloc.argtypes = [C.POINTER(_vp), C.c_size_t]
            H.hipFree.argtypes = [_vp]
            H.hipMemcpy.argtypes = [_vp, _vp, C.c_size_t, C.c_int]
            for f in ("hipSetDevice", "hipMalloc", "hipFree", "hipMemcpy"):
                getattr(H, f).restype = C.c_int
            _hip = H
    return _hip


class DeviceBuffer:
    """A device allocation of ``nbytes`` on ``device`` (hipMalloc), freed
    with the object.  ``ptr`` is the device address for the ``_device``
    entry points; to_host / from_host copy synchronously."""

    def __init__(self, nbytes, device=0):
        H = hip()
        self.device = device
        self.nbytes = int(nbytes)
        self.ptr = _vp()
        if H.hipSetDevice(device) != 0:
            raise SdfError(-1, f"hipSetDevice({device}) failed")
        if H.hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)) != 0:
            raise SdfError(-1, f"hipMalloc of {self.nbytes} bytes failed")

    def from_host(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        if a.nbytes and hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) != 0:
            raise SdfError(-1, "hipMemcpy host to device failed")
        return self

    def to_host(self, dtype, count):
        out = np.empty(int(count), dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("read beyond the buffer")
        if out.nbytes and hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) != 0:
            raise SdfError(-1, "hipMemcpy device to host failed")
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def dptr(a):
    return a.ctypes.data_as(_dp)


def i64ptr(a):
    return a.ctypes.data_as(_i64p)


def f64_3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def i64_3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int64).reshape(3))


def as_points(points):
    """Column-major 3 x n float64 (the layout of a Julia Matrix{Float64})."""
    p = np.asarray(points, dtype=np.float64)
    if p.ndim != 2 or p.shape[0] != 3:
        raise ValueError("points must be a 3 x n array")
    return np.asfortranarray(p)


def stage_times(c):
    names = C.create_string_buffer(1024)
    ms = np.zeros(64, dtype=np.float64)
    n = C.c_int32(0)
    check(lib().sdf_ctx_stage_times(c, names, 1024, dptr(ms), 64, C.byref(n)))
    keys = names.value.decode().split(";") if n.value else []
    return dict(zip(keys, ms[: n.value].tolist()))


def vote_points(c):
    """Linear indices of the points the last lattice call sent to the vote."""
    n = C.c_int64(0)
    check(lib().sdf_ctx_vote_points(c, None, 0, C.byref(n)))
    out = np.zeros(max(n.value, 1), dtype=np.int64)
    check(lib().sdf_ctx_vote_points(c, i64ptr(out), n.value, C.byref(n)))
    return out[: n.value]


def counters(c):
    out = np.zeros(8, dtype=np.int64)
    n = C.c_int32(0)
    check(lib().sdf_ctx_counters(c, i64ptr(out), 8, C.byref(n)))
    return out[: n.value].tolist()
