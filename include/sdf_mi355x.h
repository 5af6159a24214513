/*
 * sdf_mi355x.h -- C ABI of the MI355X signed-distance-field library.
 *
 * Drop-in boundary for the SDF hot path of jezekon/stl2sdf.jl.  The Julia
 * package has no FFI of its own; each entry point below replaces one Julia
 * function of the reference and keeps its argument meaning and array
 * layouts, so a ccall binding (or any other FFI) can sit directly behind the
 * reference's function names.
 *
 * Conventions (identical to the reference's in-memory layouts):
 *   X        3 x nnp doubles, column-major (x,y,z of vertex 1, then vertex 2 ...)
 *            (TriangularMesh.X, src/MeshGrid/MeshInformations.jl:58)
 *   IEN      3 x nel int64, column-major, 1-based vertex indices
 *            (TriangularMesh.IEN, MeshInformations.jl:59)
 *   grid     aabb_min[3], cell size h, N[3] cells per axis; (N+1) points per
 *            axis, linear index i + (N1+1)*(j + (N2+1)*k), 0-based here
 *            (Grid, src/MeshGrid/Grid.jl:1-33, 82-96)
 *   points   3 x npts doubles or NULL; NULL means the grid's own lattice
 *            (generateGridPoints).  A points array equal to the lattice is
 *            detected and takes the lattice path.
 *   signs    inside = +1, outside = -1 (RaycastingSignDetection.jl:285)
 *
 * Every call returns an sdf_status: 0 on success, < 0 on error, with a
 * message from sdf_last_error() (thread-local).  Nothing aborts or throws
 * across the ABI.  Host-pointer entry points are synchronous: outputs are
 * written before they return and no host pointer is retained.  Entry points
 * with a _device suffix take device pointers, are asynchronous on the
 * context's stream, and need sdf_ctx_synchronize() before the caller reads.
 * A context serialises its own calls; distinct contexts may be used from
 * different threads.
 */
#ifndef SDF_MI355X_H
#define SDF_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t sdf_status;
typedef struct sdf_ctx sdf_ctx;

enum {
  SDF_OK = 0,
  SDF_ERR_ARG = -1,      /* invalid argument (sizes, indices, NULL) */
  SDF_ERR_HIP = -2,      /* HIP runtime failure (no device, OOM, ...) */
  SDF_ERR_LIMIT = -3,    /* an internal capacity was exceeded */
  SDF_ERR_INTERNAL = -4
};

/* ABI version (major*100 + minor). */
int32_t sdf_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char *sdf_last_error(void);

/* Context: one HIP device, one stream, cached device buffers.
 * No reference counterpart (the reference has no device state). */
sdf_status sdf_ctx_create(int32_t device, sdf_ctx **out);
sdf_status sdf_ctx_destroy(sdf_ctx *ctx);
/* Run on a caller-owned hipStream_t (NULL: back to the context's own). */
sdf_status sdf_ctx_set_stream(sdf_ctx *ctx, void *hip_stream);
sdf_status sdf_ctx_synchronize(sdf_ctx *ctx);
/* Launch tunables of the lattice path (no counterpart in the reference;
 * results do not depend on them): "chunk_len" triangles per distance work
 * item (1..64, default 64), "ray_stack" LDS ray-stack depth of the vote
 * stage (1024 or 2048), "dist_per_cu" / "fused_per_cu" workgroups per CU
 * of the distance / sign kernels (0: one per work item; defaults 7 / 0),
 * "sign_split" 1: the sign stage's bricks whose lines all agree run on a
 * helper stream beside the vote (default 1), "same_per_cu" workgroups per CU
 * of that pass (0: one per brick, the default), "wind_ahead" 1: the
 * winding sums of the vote's likely winding items run beside its ray phases
 * (0: only after the vote, for the items it leaves; default 0),
 * "lbvh_early" when the LBVH build runs: 0 after the tile binning, issued
 * after the distance and axis-line work, 1 the same, issued before it, 2
 * after the axis-line hits end, 3 beside the tile binning (default 3),
 * "lines_early" 1: the axis-line hits start beside the tile binning
 * instead of after it, 2: only their counting pass does (default 0), "line_per_cu" workgroups per CU of the
 * axis-line kernels (0: one per triangle or line, the default). */
sdf_status sdf_ctx_tune(sdf_ctx *ctx, const char *name, int64_t value);
/* Stage timers (hipEvents) of the last call, when enabled: names are
 * written as one ';'-separated string, milliseconds into ms[0..n). */
sdf_status sdf_ctx_enable_timing(sdf_ctx *ctx, int32_t on);
sdf_status sdf_ctx_stage_times(sdf_ctx *ctx, char *names, int32_t names_cap,
                               double *ms, int32_t cap, int32_t *n);
/* Counters of the last call: [0] points handed to the adaptive vote (left
 * undecided by the six axis rays, or decided inside a rounding margin),
 * [1] triangles listed in distance tiles, [2] axis-line hits,
 * [3] device bytes held by the context, [4] vote points still open after
 * the first vote phase (10 rays), [5] points settled by the winding
 * number, [6] of the points of [0], those decided inside a rounding margin
 * (they vote from ray 1). */
sdf_status sdf_ctx_counters(sdf_ctx *ctx, int64_t *out, int32_t cap,
                            int32_t *n);

/* Diagnostics, no counterpart in the reference: linear indices of the
 * lattice points the last lattice call with signs handed to the adaptive
 * vote (left undecided by the six axis rays, or decided inside a rounding
 * margin), ascending, up to cap of them; *n receives the total. */
sdf_status sdf_ctx_vote_points(sdf_ctx *ctx, int64_t *idx_out, int64_t cap,
                               int64_t *n);

/* Grid(X_min, X_max, N_max, margin) -- src/MeshGrid/Grid.jl:9-33. */
sdf_status sdf_grid_setup(const double xmin[3], const double xmax[3],
                          int64_t n_max, int64_t margin, double aabb_min[3],
                          double aabb_max[3], int64_t N[3], double *cell_size,
                          int64_t *ngp);

/* evalDistancesOnTriMesh(mesh, grid, points) -> dist
 * -- src/SignedDistances/dfOnTriangularMesh.jl:19-209.
 * dist_out[npts]: exact point-triangle distance clamped at h (min(d, h)),
 * FP32 arithmetic on lattice-relative coordinates, FP64 output.  The
 * closest-point matrix xp comes from sdf_eval_distances_xp. */
sdf_status sdf_eval_distances(sdf_ctx *ctx, const double *X, int64_t nnp,
                              const int64_t *IEN, int64_t nel,
                              const double aabb_min[3], double h,
                              const int64_t N[3], const double *points,
                              int64_t npts, double *dist_out);

/* evalDistancesOnTriMesh(mesh, grid, points) -> (dist, xp)
 * -- src/SignedDistances/dfOnTriangularMesh.jl:19-209 (xp: :108-186, merged
 * at :194-199).  dist_out as sdf_eval_distances; xp_out[3 * npts]
 * (column-major 3 x npts): the closest point on the mesh to each point,
 * exact FP64 over all triangles (Voronoi-region closest point, smallest
 * (squared distance, triangle index) on ties) at every distance, also where
 * the reference's candidate search is heuristic (beyond ~2.5h); all zeros
 * for a mesh with no triangle, as the reference's initial xp. */
sdf_status sdf_eval_distances_xp(sdf_ctx *ctx, const double *X, int64_t nnp,
                                 const int64_t *IEN, int64_t nel,
                                 const double aabb_min[3], double h,
                                 const int64_t N[3], const double *points,
                                 int64_t npts, double *dist_out, double *xp_out);

/* raycast_sign_detection(mesh, grid, points; min_rays, max_rays,
 * confidence_threshold, use_winding_fallback) -> (signs, confidences)
 * -- src/SignedDistances/RaycastingSignDetection.jl:337-404.
 * Bit-identical decisions: FP64 Moller-Trumbore in the reference's operation
 * order, eps 1e-8, sorted-t dedup, 6 -> 30 adaptive vote, winding fallback. */
sdf_status sdf_raycast_signs(sdf_ctx *ctx, const double *X, int64_t nnp,
                             const int64_t *IEN, int64_t nel,
                             const double aabb_min[3], double h,
                             const int64_t N[3], const double *points,
                             int64_t npts, int32_t min_rays, int32_t max_rays,
                             double confidence_threshold,
                             int32_t use_winding_fallback, double *signs_out,
                             double *conf_out);

/* The stl_to_sdf distance + sign + combine steps (src/stl_to_sdf.jl:101-119):
 * sdf_out[ngp] = evalDistancesOnTriMesh(...) .* raycast_sign_detection(...)
 * with default ray options, over the grid's lattice. */
sdf_status sdf_signed_distance(sdf_ctx *ctx, const double *X, int64_t nnp,
                               const int64_t *IEN, int64_t nel,
                               const double aabb_min[3], double h,
                               const int64_t N[3], double *sdf_out);

/* Device-resident form of the same steps: the mesh is given as device
 * pointers (same layouts) and is kept by the context until the next
 * sdf_mesh_set_device call, which does not wait for the device (an IEN
 * index outside 1..nnp is reported by every later call on that mesh);
 * sdf_signed_distance_device writes ngp doubles to a device buffer on the
 * context stream and returns after one status read. */
sdf_status sdf_mesh_set_device(sdf_ctx *ctx, const double *d_X, int64_t nnp,
                               const int64_t *d_IEN, int64_t nel);
sdf_status sdf_signed_distance_device(sdf_ctx *ctx, const double aabb_min[3],
                                      double h, const int64_t N[3],
                                      double *d_sdf_out);
/* Host-array form of sdf_mesh_set_device (copies X / IEN, synchronous). */
sdf_status sdf_mesh_set(sdf_ctx *ctx, const double *X, int64_t nnp, const int64_t *IEN,
                        int64_t nel);

/* The device part of stl_to_sdf (src/stl_to_sdf.jl:101-126) on the mesh of
 * the last sdf_mesh_set[_device] call: evalDistancesOnTriMesh,
 * raycast_sign_detection (default ray options), sdf_dists = dists .* signs
 * and, when remove_artifacts is nonzero, remove_sdf_artifacts!(sdf_dists,
 * grid; threshold = 0, min_component_ratio), all on the context stream into
 * the device buffer d_sdf_out (ngp doubles), which never crosses PCIe.
 * *nflipped (nullable) receives the artifact count (0 without removal).
 * Synchronous; two status reads (lattice sizes, flip count). */
sdf_status sdf_signed_fused(sdf_ctx *ctx, const double aabb_min[3], double h,
                            const int64_t N[3], int32_t remove_artifacts,
                            double min_component_ratio, double *d_sdf_out,
                            int64_t *nflipped);

/* remove_sdf_artifacts!(sdf, grid; threshold, min_component_ratio)
 * -- src/SignedDistances/SdfArtifactRemoval.jl:134-248.
 * sdf_inout[(N+1) product] is modified in place: interior nodes
 * (sdf >= threshold) of every 6-connected component other than the largest
 * whose size is below max(1, round(ratio * largest)) (round half to even)
 * become -abs(sdf).  *nflipped receives the count. */
sdf_status sdf_remove_artifacts(sdf_ctx *ctx, double *sdf_inout,
                                const int64_t N[3], double threshold,
                                double min_component_ratio, int64_t *nflipped);
/* Same on a device buffer (synchronous: *nflipped is set on return). */
sdf_status sdf_remove_artifacts_device(sdf_ctx *ctx, double *d_sdf,
                                       const int64_t N[3], double threshold,
                                       double min_component_ratio,
                                       int64_t *nflipped);

/* Allocation-free variant of sdf_remove_artifacts_device.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (serving loops, harnesses that forbid hipMalloc, set-up next to HIP graph
 * capture).  Same arguments and result, bit for bit; it calls no hipMalloc,
 * hipFree or allocating library routine.  Its scratch (component labels
 * and sizes, 8 bytes per node) is a static per-device array of
 * SDF_NOALLOC_MAX_NODES + 1 nodes (64 MiB) that each GPU the library runs on
 * holds until the process ends; grids with more nodes are rejected with
 * SDF_ERR_LIMIT.  Every call on one GPU shares that array, from any context
 * or thread: calls that overlap in time on one GPU race.  Synchronous
 * (*nflipped is set on return). */
#define SDF_NOALLOC_MAX_NODES 8388608
sdf_status sdf_remove_artifacts_device_noalloc(sdf_ctx *ctx, double *d_sdf,
                                               const int64_t N[3], double threshold,
                                               double min_component_ratio,
                                               int64_t *nflipped);

/* analyze_sdf_components(sdf, grid; threshold) -> Dict(root => size)
 * -- SdfArtifactRemoval.jl:259-318.  Roots are canonical: the smallest
 * 0-based linear index of each component.  roots_out/sizes_out receive up to
 * cap entries sorted by root; *ncomp is the total.  labels_out (nullable,
 * one int64 per node) receives each node's root or -1. */
sdf_status sdf_analyze_components(sdf_ctx *ctx, const double *sdf,
                                  const int64_t N[3], double threshold,
                                  int64_t *labels_out, int64_t *roots_out,
                                  int64_t *sizes_out, int64_t cap,
                                  int64_t *ncomp);

/* RBFs_smoothing(dist, grid, is_interpolation, smooth, threshold)
 * -- src/SdfSmoothing/RBFs4Smoothing.jl:275-320.
 * dist: (N+1) product doubles; fine_out: prod(N*smooth + 1) floats (i
 * fastest).  Kernel exp(-(r/h)^2) truncated where it drops to threshold;
 * interpolation solves K w = Float32.(dist) until the FP64 residual satisfies
 * ||b - K w|| <= reltol ||b|| (reltol <= 0: sqrt(eps(Float32)), the
 * reference's cg default); approximation uses w = Float32.(dist).
 * iters_out / rel_residual_out (nullable): CG iterations and final relative
 * residual. */
sdf_status sdf_rbf_smoothing(sdf_ctx *ctx, const double *dist,
                             const int64_t N[3], double h,
                             int32_t is_interpolation, int32_t smooth,
                             double threshold, double reltol, float *fine_out,
                             int32_t *iters_out, double *rel_residual_out);
/* Device-buffer form of sdf_rbf_smoothing: d_dist (ngp doubles) and
 * d_fine_out (prod(N*smooth + 1) floats) are device pointers; the same
 * arithmetic and results.  Synchronous. */
sdf_status sdf_rbf_smoothing_device(sdf_ctx *ctx, const double *d_dist,
                                    const int64_t N[3], double h,
                                    int32_t is_interpolation, int32_t smooth,
                                    double threshold, double reltol,
                                    float *d_fine_out, int32_t *iters_out,
                                    double *rel_residual_out);

/* import_stl(filename) -> (X, IEN) -- src/DataImport/ImportStlData.jl:10-27
 * with the package's readers (:102-196: a file whose first five bytes are
 * "solid" is read as ASCII with Float64 coordinates, any other as binary
 * with float32 ones) and process_stl_mesh's merge (:198-309: each triangle
 * corner joins the first earlier unique vertex within 1e-10 in its 1e-10
 * hash cell, then in the 26 neighbouring cells in the reference's order).
 * Host code, no device work.  Two calls: with X and IEN NULL, *nnp and *nel
 * receive the sizes; then, with *nnp / *nel the buffer capacities, X
 * (3 x nnp doubles) and IEN (3 x nel int64, 1-based) are filled. */
sdf_status sdf_import_stl(const char *path, double *X, int64_t *nnp, int64_t *IEN,
                          int64_t *nel);

/* The 36 prioritised ray directions (generate_ray_directions_prioritized,
 * RaycastingSignDetection.jl:162-227), 3 doubles each. */
sdf_status sdf_ray_directions(double *dirs_out /* 108 */);

/* ------------------------------------------------------------------
 * Learned search radius.  No counterpart in the reference: a per-mesh
 * coordinate MLP (16 Fourier frequencies, three hidden layers of 128, ReLU,
 * BF16 weights and activations with FP32 accumulation) proposes a search
 * radius for each fine-grid node; the distance loop evaluates only triangles
 * within it and re-runs a node's tile over all triangles when none is found,
 * so every output equals the one of sdf_signed_distance (nodes whose radius
 * held no triangle are settled by exact LBVH queries with the same
 * arithmetic).
 *
 * Parameter vector (SDF_MLP_NPARAM floats, weights out x in, row-major):
 *   W1 @0, b1 @16384, W2 @16512, b2 @32896, W3 @33024, b3 @49408,
 *   w4 @49536, b4 @49664.
 * Inputs u are normalised coordinates (3 floats per point); features are
 * u, 0, then sin/cos(pi 2^(0.4k) u_a) for k < 16, zero padded to 128. */
#define SDF_MLP_NPARAM 49665

sdf_status sdf_mlp_set_params(sdf_ctx *ctx, const float *params);
sdf_status sdf_mlp_get_params(sdf_ctx *ctx, float *params_out);
/* y_out[i] = network(u[3i..3i+2]) */
sdf_status sdf_mlp_forward(sdf_ctx *ctx, const float *u, int64_t n, float *y_out);
/* Mean squared error over (u, t) and its gradient (SDF_MLP_NPARAM floats). */
sdf_status sdf_mlp_loss_grad(sdf_ctx *ctx, const float *u, const float *t, int64_t n,
                             float *grad_out, double *loss_out);
/* `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8) on batches of `batch`
 * samples, learning rate decayed by a cosine from lr to lr/20; seed 0 takes
 * the samples in order, otherwise reshuffles every epoch.  loss_hist
 * (nullable) receives each step's batch loss. */
sdf_status sdf_mlp_train(sdf_ctx *ctx, const float *u, const float *t, int64_t n,
                         int32_t steps, int32_t batch, float lr, uint64_t seed,
                         float *loss_hist);
/* Fit the network to the context's mesh for the fine grid (aabb_min, h, N):
 * exact distances on a coarse lattice over the same box with n_coarse points
 * on the longest axis (spacing hc), clamped at hc, are the targets (d/hc);
 * training set = every node with d < hc plus about as many farther nodes.
 * The radius offset delta is the given quantile (<= 0: 0.99) of
 * (target - max(output, 0)) over the nodes with d < hc; a fine node's radius
 * is min(max((max(output, 0) + delta) * hc, h / 100), h).  stats_out (nullable, up to
 * 9): final loss, mean absolute error (units of hc), delta, samples, samples
 * with d < hc, training ms, coarse-distance ms, hc, steps. */
sdf_status sdf_learned_fit(sdf_ctx *ctx, const double aabb_min[3], double h,
                           const int64_t N[3], int32_t n_coarse, int32_t steps,
                           int32_t batch, float lr, uint64_t seed, double quantile,
                           double *stats_out, int32_t stats_cap);
/* sdf_signed_distance / sdf_signed_distance_device with the learned radius
 * (identical outputs).  When delta * hc >= h every radius is h whatever the
 * network outputs: it could prune nothing, and the call runs the plain
 * distance loop without evaluating it. */
sdf_status sdf_learned_signed_distance(sdf_ctx *ctx, const double *X, int64_t nnp,
                                       const int64_t *IEN, int64_t nel,
                                       const double aabb_min[3], double h,
                                       const int64_t N[3], double *sdf_out);
sdf_status sdf_learned_signed_distance_device(sdf_ctx *ctx, const double aabb_min[3],
                                              double h, const int64_t N[3],
                                              double *d_sdf_out);
/* Counters of the last learned call (synchronises), up to 9: nodes whose
 * radius held no triangle (settled by exact LBVH queries), exact
 * evaluations (per wave: its longest candidate list, summed), (tile,
 * triangle) pairs, (node, triangle) pairs within the learned radius and
 * within h (box/plane lower bound), delta, hc, the exact evaluations the
 * plain loop makes on the same tiles, and 1 when the call ran without the
 * network (delta * hc >= h; the counters before it are then 0 but for the
 * tile pairs, delta and hc). */
sdf_status sdf_learned_stats(sdf_ctx *ctx, double *out, int32_t cap, int32_t *n);

/* ------------------------------------------------------------------
 * Batches.  No counterpart in the reference (it handles one mesh per
 * call): many device-resident meshes, each with its own grid, computed as
 * by sdf_signed_distance_device (bit for bit) in one submission.  The
 * meshes are dealt round-robin to worker contexts ("lanes", each with its
 * own streams and buffers); a run issues every mesh's pass with no host
 * wait, so the lanes' passes overlap on the GPU, and reads all status
 * blocks once at the end.  Creation plans every mesh (sizes its lists and
 * the lanes' buffers) and, with autotune != 0, first times candidate lane
 * counts, distance work-item lengths and ray-stack depths on the batch
 * itself and keeps the fastest.
 * Arrays are per mesh: d_X[i] (3 x nnp[i] doubles), d_IEN[i] (3 x nel[i]
 * 1-based int64), aabb_min[3i..], h[i], N[3i..], d_sdf_out[i] (prod(N+1)
 * doubles); all device pointers stay valid and in place for the batch's
 * life.  At most 4096 meshes. */
typedef struct sdf_batch sdf_batch;
sdf_status sdf_batch_create(sdf_ctx *ctx, int32_t nmesh, const double *const *d_X,
                            const int64_t *nnp, const int64_t *const *d_IEN,
                            const int64_t *nel, const double *aabb_min, const double *h,
                            const int64_t *N, double *const *d_sdf_out, int32_t autotune,
                            sdf_batch **out);
/* One run: recomputes every mesh from its current device arrays and
 * synchronises (a mesh whose lists outgrew their capacities is computed
 * again on its own). */
sdf_status sdf_batch_run(sdf_batch *batch);
/* meshes, grid points, last run ms (wall time of issue, passes and status
 * read), then the chosen tunables (work-item triangles, ray-stack depth,
 * lanes), the last run's host issue ms, whether each lane has its own
 * issuing host thread (1) or not (0), the number of timed candidates and
 * their (id, ms) pairs (id: lanes, 100 + work-item triangles, 1000 +
 * ray-stack depth, 10000 + threads). */
sdf_status sdf_batch_info(sdf_batch *batch, double *out, int32_t cap, int32_t *n);
sdf_status sdf_batch_destroy(sdf_batch *batch);

#ifdef __cplusplus
}
#endif
#endif /* SDF_MI355X_H */
