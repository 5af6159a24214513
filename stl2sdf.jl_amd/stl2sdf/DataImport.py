"""STL import, mirroring src/DataImport/ImportStlData.jl.

import_stl(path) -> (X, IEN): X 3 x nnp float64, IEN 3 x nel int64 1-based,
read by the library's native reader (sdf_import_stl): the reference's own
readers (:102-196; a file starting with "solid" is ASCII with Float64
coordinates, any other binary with float32 ones) and process_stl_mesh's
vertex merge (:198-309: a corner joins the first earlier unique vertex
within 1e-10, searched in its 1e-10 hash cell and then its 26 neighbours in
the reference's order).  The numpy readers below return the raw corners
(3 per triangle) for tests and tools.
"""
import ctypes as C

import numpy as np


def _detect_ascii(path):
    with open(path, "rb") as f:
        return f.read(5) == b"solid"


def read_binary_stl(path):
    with open(path, "rb") as f:
        f.read(80)
        n = int(np.frombuffer(f.read(4), "<u4")[0])
        rec = np.frombuffer(f.read(50 * n), dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)),
                                                   ("a", "<u2")], count=n)
    return rec["v"].astype(np.float64).reshape(-1, 3)


def read_ascii_stl(path):
    """Corners of read_ascii_stl (ImportStlData.jl:163-196): a short "vertex"
    line counts but adds no vertex, and a triangle is the last three vertices
    added whenever a vertex makes the count a multiple of three."""
    verts, corners, count = [], [], 0
    with open(path, "r", errors="replace") as f:
        for line in f:
            s = line.strip()
            if s.startswith("vertex"):
                count += 1
                parts = s.split()
                if len(parts) >= 4:
                    verts.append([float(parts[1]), float(parts[2]), float(parts[3])])
                    if count % 3 == 0:
                        if len(verts) < 3:
                            raise ValueError("triangle before the first vertex")
                        corners += verts[-3:]
    return np.array(corners, dtype=np.float64).reshape(-1, 3)


def import_stl(path):
    """import_stl -- ImportStlData.jl:10-27 (native reader and merge)."""
    from . import _lib
    L = _lib.lib()
    nnp, nel = C.c_int64(0), C.c_int64(0)
    p = str(path).encode()
    _lib.check(L.sdf_import_stl(p, None, C.byref(nnp), None, C.byref(nel)))
    X = np.empty((3, nnp.value), order="F")
    IEN = np.empty((3, nel.value), np.int64, order="F")
    _lib.check(L.sdf_import_stl(p, _lib.dptr(X), C.byref(nnp), _lib.i64ptr(IEN),
                                C.byref(nel)))
    return X, IEN


def read_corners(path):
    """The raw triangle corners (n*3 x 3 float64) as the reference's readers
    produce them, before the merge."""
    return read_ascii_stl(path) if _detect_ascii(path) else read_binary_stl(path)
