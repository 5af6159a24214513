"""Writes the golden fixtures in this folder from the CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py [--rabbit PATH]

Fixtures (expected outputs for seeded inputs the tests rebuild exactly):
  c1_icosphere.npz   config 1: icosphere r=1 subdiv 4 (float32 vertices),
                     Grid(X_min, X_max, 40, 2): 45^3 points.  signs (packed
                     bits, 1 = inside), confidences != 1 (index, value),
                     exact min(d, h) distances as float32.
  artifacts_05.npz   artifacts.stl at cell_size 0.5 (stl_to_sdf's grid,
                     margin 3): same fields.
  voronoi_rabbit_aabb.json
                     AABB of the reference's data/VoronoiRabbit.stl (the mesh
                     itself is not copied), for the grid-size check of its
                     documented cell 0.5 setup.
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "stl2sdf.jl_amd"))
sys.path.insert(0, REPO)


def pack_case(mesh, grid, path):
    from oracle import oracle
    from stl2sdf import MeshGrid
    P = MeshGrid.generateGridPoints(grid)
    s, c = oracle.signs(mesh.X, mesh.IEN, P)
    d = oracle.dist_exact(mesh.X, mesh.IEN, P, grid.cell_size)
    idx = np.flatnonzero(c != 1.0)
    np.savez_compressed(
        path, N=np.asarray(grid.N), h=grid.cell_size, aabb_min=grid.AABB_min,
        signs=np.packbits(s > 0), conf_idx=idx.astype(np.int64), conf_val=c[idx],
        dist=d.astype(np.float32))
    print(path, grid.ngp, "points,", idx.size, "conf != 1")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rabbit", default="/root/reference/data/VoronoiRabbit.stl")
    args = ap.parse_args()
    from tests import _cases
    mesh, grid = _cases.icosphere_case()
    pack_case(mesh, grid, os.path.join(HERE, "c1_icosphere.npz"))
    mesh, grid = _cases.artifacts_case(0.5)
    pack_case(mesh, grid, os.path.join(HERE, "artifacts_05.npz"))
    if os.path.exists(args.rabbit):
        from stl2sdf.DataImport import import_stl
        X, IEN = import_stl(args.rabbit)
        info = {"xmin": X.min(1).tolist(), "xmax": X.max(1).tolist(),
                "nnp": int(X.shape[1]), "nel": int(IEN.shape[1])}
        with open(os.path.join(HERE, "voronoi_rabbit_aabb.json"), "w") as f:
            json.dump(info, f, indent=1)
        print("rabbit", info)


if __name__ == "__main__":
    main()
