"""Benchmark: signed grid points per second on one MI355X.

Workload (default): the seeded star-shaped surface r(theta, phi) = 1 + 0.3 *
sum of 8 random real spherical harmonics (l <= 6, seed 0) on a 708 x 708
latitude-longitude mesh (1,001,112 triangles, float32 vertices) and the grid
Grid(X_min, X_max, 505, 3), i.e. 512 points on the longest axis.  One step is
the whole device-resident signed-distance path: triangle gather from X/IEN,
binning, fused distance + axis-ray kernel, adaptive vote for undecided
points, written as a float64 SDF array in HBM.

Timed steps run without instrumentation; the JSON line also carries the
stage split and the roofline of the dominant (fused) kernel, both from HIP
events on the kernels' stream over 5 further steps with the stage timers on, the learned-radius path on the same
workload (network fitted once on a 128-point coarse lattice, not timed; each
timed step as above with the radius-limited distance loop; its output is
checked bit-for-bit against the plain path's), and the CPU oracle timed on a
bounded random sample of the same grid points on the host cores.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "stl2sdf.jl_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "fused_lattice_star512_hbm_traffic.md")


# the sources of the sign stage's two kernels and of the brick marks that
# split the bricks between them; the traffic summary records their hash, and a
# summary measured on other sources is not reported
KERNEL_SOURCES = ["stl2sdf.jl_amd/csrc/sdf_query.hip", "stl2sdf.jl_amd/csrc/sdf_common.h",
                  "stl2sdf.jl_amd/csrc/sdf_internal.h", "stl2sdf.jl_amd/csrc/sdf_build.hip"]


def kernel_source_hash():
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(REPO, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(workload):
    """HBM bytes per step of the sign stage's two kernels (fused_list_kernel,
    fused_same_kernel) from the committed counter summary (FETCH_SIZE / WRITE_SIZE passes, copy-calibrated), with the
    summary's source hash and commit.  traffic is None when the summary does
    not cover this workload or was measured on other kernel sources."""
    info = {"traffic": None, "traffic_profile": os.path.relpath(TRAFFIC_PROFILE, REPO),
            "traffic_source_hash": None, "traffic_commit": None,
            "kernel_source_hash": kernel_source_hash(), "traffic_stale": None}
    if workload != "star512" or not os.path.exists(TRAFFIC_PROFILE):
        return info
    value = None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("fused_lattice traffic_bytes_per_step:"):
            value = float(line.split(":")[1])
        elif line.startswith("kernel_source_hash:"):
            info["traffic_source_hash"] = line.split(":")[1].strip()
        elif line.startswith("measured_at_commit:"):
            info["traffic_commit"] = line.split(":")[1].strip()
    info["traffic_stale"] = info["traffic_source_hash"] != info["kernel_source_hash"]
    if not info["traffic_stale"]:
        info["traffic"] = value
    return info


def staged(L, ctx, step, torch, n=5):
    """Per-stage times (HIP events on the stream the kernels run on) over n
    extra steps with the library's stage timers on; lists of ms per stage."""
    from stl2sdf import _lib
    _lib.check(L.sdf_ctx_enable_timing(ctx, 1))
    stages = {}
    for _ in range(n):
        step()
        for k, v in _lib.stage_times(ctx).items():
            stages.setdefault(k, []).append(v)
    torch.cuda.synchronize()
    _lib.check(L.sdf_ctx_enable_timing(ctx, 0))
    return stages


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(name):
    from stl2sdf import MeshGrid, synthetic
    if name == "star512":
        X, IEN, info = synthetic.star_surface(708, 708, 0)
        n_max, desc = 505, "star surface 708x708 lat-long (seed 0), 512^3-class grid"
    elif name == "star256":
        X, IEN, info = synthetic.star_surface(450, 450, 0)
        n_max, desc = 249, "star surface 450x450 lat-long (seed 0), 256^3-class grid"
    elif name == "icosphere":
        X, IEN = synthetic.icosphere(1.0, 4)
        info, n_max, desc = {}, 40, "icosphere r=1 subdiv 4, cell 0.05"
    else:
        raise SystemExit(f"unknown workload {name}")
    X = synthetic.to_float32(X)
    mesh = MeshGrid.TriangularMesh(X, IEN)
    xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
    grid = MeshGrid.Grid(xmin, xmax, n_max, 3 if name != "icosphere" else 2)
    return mesh, grid, desc, info


def cpu_baseline(mesh, grid, target_s, threads):
    """Oracle (reference-faithful distance + BVH ray vote) on random samples
    of the grid's lattice points.  Each oracle call rebuilds its trees (a
    fixed cost that a full grid amortises), so two sample sizes are timed
    (three runs each, median) and the per-point cost is their slope; the
    larger size is chosen so that its runs take about target_s in all."""
    from oracle import oracle
    rng = np.random.default_rng(0)
    N = np.asarray(grid.N, np.int64)
    h = grid.cell_size

    def sample(n):
        idx = rng.integers(0, grid.ngp, size=n)
        nx, ny = N[0] + 1, N[1] + 1
        i, j, k = idx % nx, (idx // nx) % ny, idx // (nx * ny)
        P = np.empty((3, n), order="F")
        P[0] = grid.AABB_min[0] + h * i.astype(np.float64)
        P[1] = grid.AABB_min[1] + h * j.astype(np.float64)
        P[2] = grid.AABB_min[2] + h * k.astype(np.float64)
        return P

    def run(n):
        P = sample(n)
        t0 = time.perf_counter()
        oracle.dist_reference(mesh.X, mesh.IEN, P, h, threads)
        oracle.signs(mesh.X, mesh.IEN, P, threads=threads)
        return time.perf_counter() - t0

    def median3(n):
        ts = sorted(run(n) for _ in range(3))
        return ts[1], ts

    n1 = 1 << 14
    t1, ts1 = median3(n1)
    # per-point cost from one probe of 2^18 points, then the size whose runs
    # take about target_s / 3 each
    np_ = 1 << 18
    tp = run(np_)
    per_pt = max((tp - t1) / (np_ - n1), 1e-9)
    n2 = int((max(target_s / 3.0 - t1, 0.5)) / per_pt) + n1
    n2 = min(max(n2, np_), 1 << 24)
    t2, ts2 = median3(n2)
    slope = (t2 - t1) / max(n2 - n1, 1)
    build_s = max(t1 - slope * n1, 0.0)
    full_s = build_s + slope * grid.ngp
    return grid.ngp / full_s, (n1, ts1, n2, ts2, build_s, full_s)


def batch_bench(L, ctx, torch, runs=5):
    """The batched configuration: 64 seeded star surfaces, triangle counts
    log-uniform in [2k, 500k], grids of 64/96/128/192 points on the longest
    axis, as one autotuned batch; timed over `runs` runs (median), beside
    the same meshes one after another through the single-mesh device entry
    point on one context, and checked against it bit for bit."""
    from stl2sdf import Batch, MeshGrid, synthetic
    meshes, grids = [], []
    for X, IEN, g, _ in synthetic.batch_family(64, 2):
        mesh = MeshGrid.TriangularMesh(X, IEN)
        xmin, xmax = MeshGrid.getMesh_AABB(mesh.X)
        meshes.append(mesh)
        grids.append(MeshGrid.Grid(xmin, xmax, g - 7, 3))
    pts = sum(int(g.ngp) for g in grids)
    log(f"[bench] batch: {len(meshes)} meshes, {sum(m.nel for m in meshes)} triangles, "
        f"{pts} points")
    b = Batch.SignedDistanceBatch(meshes, grids, autotune=True)
    b.run()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        b.run()
        ts.append(time.perf_counter() - t0)
    t_batch = float(np.median(ts))
    info = b.info()
    # the same meshes one by one through sdf_signed_distance_device
    from stl2sdf import _lib
    outs = [_lib.DeviceBuffer(8 * int(g.ngp)) for g in grids]
    args_ = [(_lib.f64_3(g.AABB_min), float(g.cell_size), _lib.i64_3(g.N)) for g in grids]

    def seq():
        for i, m in enumerate(meshes):
            _lib.check(L.sdf_mesh_set_device(ctx, b._X[i].ptr, m.nnp, b._I[i].ptr, m.nel))
            amin, h, N = args_[i]
            _lib.check(L.sdf_signed_distance_device(ctx, _lib.dptr(amin), h, _lib.i64ptr(N),
                                                    outs[i].ptr))
        torch.cuda.synchronize()

    seq()
    tq = []
    for _ in range(runs):
        t0 = time.perf_counter()
        seq()
        tq.append(time.perf_counter() - t0)
    t_seq = float(np.median(tq))
    got = b.results()
    identical = all(np.array_equal(got[i], outs[i].to_host(np.float64, int(g.ngp)))
                    for i, g in enumerate(grids))
    b.close()
    log(f"[bench] batch {t_batch * 1e3:.2f} ms, one by one {t_seq * 1e3:.2f} ms, "
        f"identical={identical}")
    return {"value": pts / t_batch, "unit": "signed grid points/s", "ms_per_run": t_batch * 1e3,
            "meshes": len(meshes), "triangles": int(sum(m.nel for m in meshes)),
            "points": pts, "sequential_ms": t_seq * 1e3, "vs_sequential": t_seq / t_batch,
            "identical_to_single_mesh": identical,
            "issue_ms": info["issue_ms"],
            "tune": {k: info[k] for k in ("lanes", "chunk_len", "ray_stack", "threads")},
            "candidates": info["tuning"]}


def batch_child(timeout_s=600):
    """batch_bench in a child process (its own library context and runtime
    state), so that a failure there costs the batched figure only."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--batch-only"]
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, timeout=timeout_s, text=True)
    except subprocess.TimeoutExpired:
        log(f"[bench] batch child timed out after {timeout_s} s")
        return {"error": f"timed out after {timeout_s} s"}
    lines = [s for s in r.stdout.splitlines() if s.startswith("{")]
    if r.returncode != 0 or not lines:
        log(f"[bench] batch child exited with {r.returncode}")
        return {"error": f"child exited with status {r.returncode}"}
    return json.loads(lines[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="star512")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--fit-steps", type=int, default=1500)
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched 64-mesh family (the batched configuration)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched configuration, its result as one JSON line (the "
                         "default run takes it from such a child process)")
    ap.add_argument("--tune", action="append", default=[], metavar="NAME=VALUE",
                    help="launch tunable of the library (sdf_ctx_tune); results do not change")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only copy an MB-sized device buffer 3 times (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.empty(args.copy_probe * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(each GPU would take its own z-slab of the grid with no exchange).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from stl2sdf import _lib

    if args.batch_only:
        L = _lib.lib()
        ctx = _lib.ctx(0)
        stream = torch.cuda.current_stream(torch.device("cuda:0"))
        _lib.check(L.sdf_ctx_set_stream(ctx, C.c_void_p(stream.cuda_stream)))
        print(json.dumps(batch_bench(L, ctx, torch)))
        return

    log(f"[bench] building workload {args.workload}")
    mesh, grid, desc, info = make_workload(args.workload)
    log(f"[bench] nel={mesh.nel} nnp={mesh.nnp} N={list(grid.N)} ngp={grid.ngp}")

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    L = _lib.lib()
    ctx = _lib.ctx(0)
    _lib.check(L.sdf_ctx_set_stream(ctx, C.c_void_p(stream.cuda_stream)))
    for t in args.tune:
        name, value = t.split("=")
        _lib.check(L.sdf_ctx_tune(ctx, name.encode(), int(value)))
    dX = torch.from_numpy(np.ascontiguousarray(mesh.X.T)).to(dev)      # 3 x nnp column-major
    dI = torch.from_numpy(np.ascontiguousarray(mesh.IEN.T)).to(dev)
    out = torch.empty(grid.ngp, dtype=torch.float64, device=dev)
    amin = _lib.f64_3(grid.AABB_min)
    N = _lib.i64_3(grid.N)
    h = float(grid.cell_size)

    def step():
        _lib.check(L.sdf_mesh_set_device(ctx, C.c_void_p(dX.data_ptr()), mesh.nnp,
                                         C.c_void_p(dI.data_ptr()), mesh.nel))
        _lib.check(L.sdf_signed_distance_device(ctx, _lib.dptr(amin), h, _lib.i64ptr(N),
                                                C.c_void_p(out.data_ptr())))

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"[bench] warmup {w + 1}/{args.warmup} done")
    # timed steps without instrumentation, then a few steps with the stage
    # events on (their time is not counted)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log(f"[bench] timed {args.steps} steps in {dt:.3f} s")
    stages = staged(L, ctx, step, torch)
    cnt = _lib.counters(ctx)

    learned = None
    if not args.no_learned:
        import torch as _t
        fst = np.zeros(9)
        step()
        _lib.check(L.sdf_learned_fit(ctx, _lib.dptr(amin), h, _lib.i64ptr(N), 128,
                                     args.fit_steps, 4096, C.c_float(2e-3), 1, 0.99,
                                     _lib.dptr(fst), 9))
        fit = dict(zip(("loss", "mae_band_hc", "delta_hc", "samples", "samples_band",
                        "train_ms", "coarse_ms", "hc", "steps"), fst.tolist()))
        log(f"[bench] learned fit: {fit}")
        out_l = torch.empty_like(out)

        def lstep():
            _lib.check(L.sdf_mesh_set_device(ctx, C.c_void_p(dX.data_ptr()), mesh.nnp,
                                             C.c_void_p(dI.data_ptr()), mesh.nel))
            _lib.check(L.sdf_learned_signed_distance_device(
                ctx, _lib.dptr(amin), h, _lib.i64ptr(N), C.c_void_p(out_l.data_ptr())))

        run_names = ("fallback_nodes", "exact_evals", "tile_pairs", "node_pairs_radius",
                     "node_pairs_h", "delta", "hc", "exact_evals_plain", "network_skipped")

        def learned_run(force):
            # force: evaluate the network even where its radius is h everywhere
            _lib.check(L.sdf_ctx_tune(ctx, b"learned_force", int(force)))
            for w in range(args.warmup):
                lstep()
                torch.cuda.synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                lstep()
            torch.cuda.synchronize()
            ldt = time.perf_counter() - t0
            lstages = staged(L, ctx, lstep, torch)
            lst = np.zeros(len(run_names))
            nst = C.c_int32()
            _lib.check(L.sdf_learned_stats(ctx, _lib.dptr(lst), len(run_names), C.byref(nst)))
            step()
            torch.cuda.synchronize()
            identical = bool(_t.equal(out, out_l))
            lval = grid.ngp / (ldt / args.steps)
            return {
                "value": lval, "ms_per_step": ldt / args.steps * 1e3,
                "vs_plain": lval / (grid.ngp / (dt / args.steps)),
                "identical_to_plain": identical,
                "stages_ms": {k: float(np.mean(v)) for k, v in lstages.items()},
                "run": dict(zip(run_names, lst.tolist())),
            }

        learned = learned_run(False)
        learned["fit"] = fit
        forced = learned_run(True)
        _lib.check(L.sdf_ctx_tune(ctx, b"learned_force", 0))
        learned["network_evaluated"] = forced
        log(f"[bench] learned: {learned['value']:.4g} pts/s, identical="
            f"{learned['identical_to_plain']}, skipped={learned['run']['network_skipped']}; "
            f"with the network {forced['value']:.4g} pts/s")

    batched = None if args.no_batch else batch_child()

    ms_step = dt / args.steps * 1e3
    value = grid.ngp / (dt / args.steps)
    stage_ms = {k: float(np.mean(v)) for k, v in stages.items()}
    # the sign stage: one launch, or (split stage) the listed bricks on the
    # main stream plus the marked bricks on a helper stream, each timed by
    # events on its own stream; their sum is the stage's kernel time
    fused_ms = stage_ms.get("fused", float("nan")) + stage_ms.get("fused_same", 0.0)
    alg_bytes = 8.0 * grid.ngp + 24.0 * mesh.nnp + 24.0 * mesh.nel
    achieved = alg_bytes / (fused_ms * 1e-3) / 1e9
    res = {
        "metric": "signed grid pts/sec at 512^3, 1M tris: learned-radius vs plain GPU vs CPU ref",
        "value": value,
        "unit": "signed grid points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 distance + f64 sign predicates, f64 output",
        "data": "synthetic (seeded star surface, float32 vertices)",
        "config": {"workload": desc, "nel": mesh.nel, "nnp": mesh.nnp,
                   "grid_N": [int(v) for v in grid.N], "ngp": int(grid.ngp),
                   "cell_size": h, "tune": args.tune},
        "stages_ms": stage_ms,
        "counters": {"pending_points": cnt[0], "tile_pairs": cnt[1], "line_hits": cnt[2],
                     "device_bytes": cnt[3], "vote_second_phase": cnt[4],
                     "winding_points": cnt[5], "margin_points": cnt[6]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "kernel": "fused_list_kernel + fused_same_kernel (sign stage)",
                     "alg_bytes_per_step": alg_bytes,
                     "fused_ms": fused_ms, **measured_traffic(args.workload)},
    }
    res["paths"] = {"value": "plain GPU path (the drop-in default)",
                    "learned_radius": "same workload with the learned search radius"}
    if learned is not None:
        res["learned_radius"] = learned
    if batched is not None:
        res["batched"] = batched
    if not args.no_cpu_baseline:
        # every host core this process may use: the affinity mask, bounded by
        # the OpenMP thread count the environment sets (the share of the
        # host's cores given to one GPU where that is set)
        affinity = len(os.sched_getaffinity(0))
        omp = os.environ.get("OMP_NUM_THREADS", "")
        threads = min(affinity, int(omp)) if omp.isdigit() and int(omp) > 0 else affinity
        log(f"[bench] cpu baseline with {threads} threads (nproc {os.cpu_count()}, "
            f"affinity {affinity}, OMP_NUM_THREADS {omp or 'unset'})")
        v, (n1, ts1, n2, ts2, b_s, f_s) = cpu_baseline(mesh, grid, args.cpu_seconds, threads)
        fmt = lambda ts: "/".join(f"{t:.2f}" for t in ts)
        res["cpu_baseline"] = {
            "value": v, "unit": "signed grid points/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "affinity_cores": affinity,
            "omp_num_threads": omp or None,
            "sample": (f"oracle (reference-faithful distance + bounding-sphere BVH adaptive ray "
                       f"vote, FP64, {threads} threads) on {n1} and {n2} uniformly random "
                       f"lattice points of the same grid, 3 runs each ({fmt(ts1)} s; {fmt(ts2)} "
                       f"s), medians' slope = per-point cost; tree builds {b_s:.1f} s; full grid "
                       f"extrapolated to {f_s:.0f} s")}
        res["gpu_over_cpu"] = value / v
    print(json.dumps(res))


if __name__ == "__main__":
    main()
