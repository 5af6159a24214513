/*
 * CPU FP64 restatement of the stl2sdf.jl SDF algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the timed CPU baseline.  The product library never links, loads
 * or calls it; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may.
 *
 * Conventions shared with the reference (citations are to the reference
 * repository, jezekon/stl2sdf.jl):
 *   X    : 3 x nnp doubles, column-major (xyz of vertex 0, then vertex 1, ...)
 *   IEN  : 3 x nel int64, 1-based vertex indices
 *   pts  : 3 x npts doubles, column-major
 *   grid linear index: i fastest, then j, then k (src/MeshGrid/Grid.jl:87-94)
 *   sign convention: inside = +1 (RaycastingSignDetection.jl:285)
 */
#ifndef SDF_ORACLE_H
#define SDF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid ctor, src/MeshGrid/Grid.jl:9-33. */
void orc_grid(const double xmin[3], const double xmax[3], int64_t n_max,
              int64_t margin, double aabb_min[3], double aabb_max[3],
              int64_t N[3], double *cell_size, int64_t *ngp);

/* generateGridPoints, src/MeshGrid/Grid.jl:82-96 (3 x ngp, i fastest). */
void orc_grid_points(const double aabb_min[3], double h, const int64_t N[3],
                     double *pts);

/* The 36 prioritised ray directions, RaycastingSignDetection.jl:162-227. */
void orc_ray_directions(double dirs[36 * 3]);

/* Moller-Trumbore with the reference's operation order,
 * RaycastingSignDetection.jl:64-106.  Returns 1 and *t on a hit. */
int orc_ray_triangle(const double o[3], const double d[3], const double v1[3],
                     const double v2[3], const double v3[3], double eps,
                     double *t);

/* Exact point-triangle distance in FP64 (closest point by Voronoi regions),
 * brute force over all triangles, then clamped at h like
 * dfOnTriangularMesh.jl:202-206.  Optional closest points (3 x npts). */
void orc_dist_exact(const double *X, int64_t nnp, const int64_t *IEN,
                    int64_t nel, const double *pts, int64_t npts, double h,
                    double *dist, double *xp, int nthreads);

/* orc_dist_exact's min(d, h) (no xp) through a bounding-volume hierarchy:
 * bit-identical results, for meshes too large for the brute-force loop. */
void orc_dist_exact_bvh(const double *X, int64_t nnp, const int64_t *IEN,
                        int64_t nel, const double *pts, int64_t npts, double h,
                        double *dist, int nthreads);

/* Reference-faithful distance, dfOnTriangularMesh.jl:19-209: triangle-centre
 * candidates within 2.5h (sorted by triangle index), 3-nearest-vertex
 * fallback, barycentric face test, edge projections, vertex test only when no
 * edge improved, clamp at h. */
void orc_dist_reference(const double *X, int64_t nnp, const int64_t *IEN,
                        int64_t nel, const double *pts, int64_t npts, double h,
                        double *dist, int nthreads);

/* raycast_sign_detection, RaycastingSignDetection.jl:337-404.
 * dirs: 36 x 3 (from orc_ray_directions).  use_bvh=0 tests every triangle
 * per ray; use_bvh=1 uses a bounding-sphere BVH (same result, faster).
 * Per-point diagnostics (nullable): rays used and 1 when the winding
 * fallback decided the sign. */
void orc_signs(const double *X, int64_t nnp, const int64_t *IEN, int64_t nel,
               const double *pts, int64_t npts, const double *dirs,
               int min_rays, int max_rays, double conf_thr, int use_winding,
               int use_bvh, double *signs, double *conf, int32_t *rays_used,
               int32_t *winding_used, int nthreads);

/* One ray: RaycastingSignDetection.jl:114-154 over all triangles. */
int orc_ray_inside(const double *X, const int64_t *IEN, int64_t nel,
                   const double o[3], const double d[3]);

/* generalized_winding_number, RaycastingSignDetection.jl:295-328. */
double orc_winding_sign(const double *X, const int64_t *IEN, int64_t nel,
                        const double p[3]);

/* 6-connected components of sdf >= threshold (SdfArtifactRemoval.jl:148-194)
 * with canonical labels: label = smallest 0-based linear index of the
 * component, -1 for exterior nodes.  Returns the number of components. */
int64_t orc_components(const double *sdf, const int64_t N[3], double threshold,
                       int64_t *labels);

/* remove_sdf_artifacts!, SdfArtifactRemoval.jl:134-248.  In place; returns
 * the number of flipped nodes. */
int64_t orc_remove_artifacts(double *sdf, const int64_t N[3], double threshold,
                             double ratio);

/* RBF smoothing, RBFs4Smoothing.jl:275-320.
 * coarse: ngp floats (already Float32 as process_vector makes them).
 * Interpolation solves K w = b with CG (x0 = 0, stop when
 * ||r|| <= reltol * ||b||, maxiter = n), FP64 arithmetic on the FP32 matrix
 * values.  Returns the iteration count; weights (ngp) and fine values
 * (prod(N*s+1)) are written. */
int orc_rbf_weights(const float *coarse, const int64_t N[3], double h,
                    double threshold, double reltol, int maxiter,
                    float *weights, double *resid_out);
void orc_rbf_fine(const float *weights, const double aabb_min[3],
                  const double aabb_max[3], const int64_t N[3], double h,
                  int smooth, double threshold, float *fine, int nthreads);

/* Binary STL reader with the reference's 1e-10 vertex dedup
 * (ImportStlData.jl:127-161, 198-309).  Two-call protocol: with X/IEN NULL
 * the counts are returned; then call again with buffers. Returns 0 on
 * success. */
int orc_read_stl_binary(const char *path, double *X, int64_t *nnp,
                        int64_t *IEN, int64_t *nel);

/* process_stl_mesh (ImportStlData.jl:198-309) on nv triangle corners V
 * (3 doubles each, 3 corners per triangle): X (3 x nnp, capacity 3 nv) and
 * IEN (nv 1-based indices) when not NULL; returns nnp. */
int64_t orc_process_stl_mesh(const double *V, int64_t nv, double *X, int64_t *IEN);

#ifdef __cplusplus
}
#endif
#endif
