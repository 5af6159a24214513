"""STL import (host side), mirroring src/DataImport/ImportStlData.jl.

import_stl(path) -> (X, IEN): X 3 x nnp float64, IEN 3 x nel int64 1-based.
Binary and ASCII files are read like the reference's readers (:127-196).
Vertex de-duplication follows process_stl_mesh (:198-309): a vertex joins
the first earlier unique vertex within 1e-10; STL coordinates come from
float32 (binary) or decimal text (ASCII), so vertices within 1e-10 of each
other are bitwise equal in practice and the first-occurrence order of
``np.unique`` reproduces the reference's numbering.
"""
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
    verts = []
    with open(path, "r", errors="replace") as f:
        for line in f:
            s = line.strip()
            if s.startswith("vertex"):
                parts = s.split()
                if len(parts) >= 4:
                    verts.append([float(parts[1]), float(parts[2]), float(parts[3])])
    V = np.array(verts, dtype=np.float64).reshape(-1, 3)
    return V[: 3 * (len(V) // 3)]


def process_stl_mesh(V):
    """Dedup per-face vertices (n*3 x 3) -> (X 3 x nnp, IEN 3 x nel)."""
    if len(V) == 0:
        return np.zeros((3, 0), order="F"), np.zeros((3, 0), np.int64, order="F")
    _, first, inv = np.unique(V, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")          # first-occurrence order
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    X = V[first[order]]
    IEN = (rank[inv.reshape(-1)] + 1).reshape(-1, 3)
    return np.asfortranarray(X.T), np.asfortranarray(IEN.T.astype(np.int64))


def import_stl(path):
    """import_stl -- ImportStlData.jl:10-27."""
    V = read_ascii_stl(path) if _detect_ascii(path) else read_binary_stl(path)
    return process_stl_mesh(V)
