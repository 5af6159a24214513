"""STL import (sdf_import_stl, no GPU): the reference's readers and its 1e-10
vertex merge (src/DataImport/ImportStlData.jl:102-309).

The merge is pinned by hand-built corner lists whose result follows from
the reference's rules: a corner joins the first earlier *unique* vertex
closer than 1e-10 (so a chain of near-duplicates does not collapse), the
search reaches into neighbouring 1e-10 cells, and vertices are numbered in
order of first appearance.  The oracle's restatement of process_stl_mesh
must give the same answers, and the library's reader must equal the oracle
on binary and ASCII files.
"""
import os

import numpy as np
import pytest

from oracle import oracle
from stl2sdf import _lib, synthetic
from stl2sdf.DataImport import import_stl, read_corners

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _ascii(path, corners, extra_lines=()):
    with open(path, "w") as f:
        f.write("solid test\n")
        for t in range(len(corners) // 3):
            f.write(" facet normal 0 0 0\n  outer loop\n")
            for v in corners[3 * t:3 * t + 3]:
                f.write("   vertex " + " ".join(repr(float(c)) for c in v) + "\n")
            f.write("  endloop\n endfacet\n")
        for line in extra_lines:
            f.write(line + "\n")
        f.write("endsolid test\n")


# corner list with near-duplicates and the result the reference's rules give
A, B, C = (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)
A2 = (5e-11, 0.0, 0.0)                      # 5e-11 from A: joins A
B2 = (1.0 + 1.5e-10, 0.0, 0.0)              # 1.5e-10 from B: new vertex
D = (1.0, 1.0, 0.0)
P1, P2 = (0.7 - 3e-11, 0.2, 0.1), (0.7 + 3e-11, 0.2, 0.1)  # 6e-11 apart, other 1e-10 cells
Q1 = (2.0, 2.0, 2.0)
Q2 = (2.0 + 8e-11, 2.0, 2.0)                # joins Q1
Q3 = (2.0 + 1.6e-10, 2.0, 2.0)              # 8e-11 from Q2 but 1.6e-10 from Q1: new
CORNERS = [A, B, C, A2, B2, D, P1, P2, Q1, Q2, Q3, D]
EXPECT_X = [A, B, C, B2, D, P1, Q1, Q3]
EXPECT_IEN = [[1, 2, 3], [1, 4, 5], [6, 6, 7], [7, 8, 5]]


def test_oracle_merge_known_answers():
    assert np.floor(P1[0] * 1e10) != np.floor(P2[0] * 1e10)   # different hash cells
    X, IEN = oracle.process_stl_mesh(np.array(CORNERS))
    assert np.array_equal(X, np.array(EXPECT_X).T)
    assert np.array_equal(IEN, np.array(EXPECT_IEN).T)


def test_native_ascii_merge_equals_known_answers_and_oracle(tmp_path):
    path = str(tmp_path / "near.stl")
    _ascii(path, CORNERS)
    X, IEN = import_stl(path)
    assert np.array_equal(X, np.array(EXPECT_X).T)
    assert np.array_equal(IEN, np.array(EXPECT_IEN).T)
    Xo, Io = oracle.process_stl_mesh(read_corners(path))
    assert np.array_equal(X, Xo) and np.array_equal(IEN, Io)


def test_native_ascii_short_vertex_line_shifts_the_grouping(tmp_path):
    # read_ascii_stl counts a "vertex" line with fewer than three numbers but
    # adds no vertex; a triangle is the last three vertices added whenever a
    # vertex makes the count a multiple of three (ImportStlData.jl:172-190)
    path = str(tmp_path / "short.stl")
    with open(path, "w") as f:
        f.write("solid s\n")
        for v in ([0, 0, 0], [1, 0, 0], [0, 1, 0]):
            f.write(f"vertex {v[0]} {v[1]} {v[2]}\n")
        f.write("vertex 5 5\n")                 # counted, not added
        for v in ([2, 0, 0], [0, 2, 0], [3, 3, 3]):
            f.write(f"vertex {v[0]} {v[1]} {v[2]}\n")
        f.write("endsolid s\n")
    X, IEN = import_stl(path)
    # count 3 -> (v1 v2 v3); count 6 at (0, 2, 0) -> (v3, (2,0,0), (0,2,0));
    # the last vertex (count 7) closes nothing
    assert np.array_equal(X.T, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [2, 0, 0], [0, 2, 0]])
    assert np.array_equal(IEN.T, [[1, 2, 3], [3, 4, 5]])
    Xo, Io = oracle.process_stl_mesh(read_corners(path))
    assert np.array_equal(X, Xo) and np.array_equal(IEN, Io)


def test_native_binary_equals_oracle_reader():
    path = os.path.join(GOLDEN, "artifacts.stl")
    X, IEN = import_stl(path)
    Xo, Io = oracle.read_stl_binary(path)
    assert np.array_equal(X, Xo) and np.array_equal(IEN, Io)
    assert IEN.shape[1] == 3712


def test_native_binary_round_trip_of_a_generated_mesh(tmp_path):
    X0, I0, _ = synthetic.star_surface(30, 30, 4)
    X0 = synthetic.to_float32(X0)
    path = str(tmp_path / "star.stl")
    synthetic.write_stl_binary(path, X0, I0)
    X, IEN = import_stl(path)
    assert X.shape[1] == X0.shape[1] and IEN.shape == I0.shape   # every vertex shared
    assert np.array_equal(X[:, IEN - 1], X0[:, I0 - 1])           # same triangles
    Xo, Io = oracle.read_stl_binary(path)
    assert np.array_equal(X, Xo) and np.array_equal(IEN, Io)


def test_native_reader_errors(tmp_path):
    bad = str(tmp_path / "trunc.stl")
    with open(bad, "wb") as f:
        f.write(b"\0" * 80 + np.uint32(5).tobytes() + b"\0" * 60)
    with pytest.raises(_lib.SdfError, match="ends before triangle"):
        import_stl(bad)
    with pytest.raises(_lib.SdfError, match="cannot open"):
        import_stl(str(tmp_path / "missing.stl"))
    nan = str(tmp_path / "word.stl")
    with open(nan, "w") as f:
        f.write("solid w\nvertex 1 2 x3\nvertex 0 0 0\nvertex 1 1 1\nendsolid w\n")
    with pytest.raises(_lib.SdfError, match="cannot parse"):
        import_stl(nan)
