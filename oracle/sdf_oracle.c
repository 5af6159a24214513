/*
 * CPU FP64 restatement of the stl2sdf.jl SDF algorithm.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path and the
 * timed CPU baseline.  The product library never links or calls this code.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  FP
 * contraction is off on purpose: the reference evaluates every product and
 * sum with its own rounding (Julia does not fuse a*b+c unless asked), and the
 * sign predicates must reproduce that bit for bit.
 *
 * Function -> reference map (jezekon/stl2sdf.jl):
 *   orc_grid / orc_grid_points    src/MeshGrid/Grid.jl:9-33, 82-96
 *   orc_ray_directions            src/SignedDistances/RaycastingSignDetection.jl:162-227
 *   orc_ray_triangle              RaycastingSignDetection.jl:64-106
 *   ray_inside_* / orc_ray_inside RaycastingSignDetection.jl:114-154
 *   adaptive vote in orc_signs    RaycastingSignDetection.jl:234-288, 337-404
 *   orc_winding_sign              RaycastingSignDetection.jl:295-328
 *   orc_dist_exact                exact min(d, h); clamp of dfOnTriangularMesh.jl:202-206
 *   orc_dist_exact_bvh            the same values through a bounding-volume hierarchy
 *   orc_dist_reference            dfOnTriangularMesh.jl:19-209, 226-251, 329-392
 *   orc_components                SdfArtifactRemoval.jl:148-194 (sequential semantics)
 *   orc_remove_artifacts          SdfArtifactRemoval.jl:134-248
 *   orc_rbf_weights / orc_rbf_fine src/SdfSmoothing/RBFs4Smoothing.jl:86-108, 142-253, 275-320
 *   orc_read_stl_binary           src/DataImport/ImportStlData.jl:127-161, 198-309
 */
#include "sdf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RAY_EPS 1e-8 /* RaycastingSignDetection.jl:118 (passed at :132) */

/* ------------------------------------------------------------------ */
/* small vector helpers, reference operation order                     */
/* ------------------------------------------------------------------ */

/* LinearAlgebra.dot on length-3 Float64 vectors: ((x1*y1 + x2*y2) + x3*y3),
 * every product and sum rounded (no fused multiply-add). */
static inline double dot3(const double *a, const double *b) {
  double s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

/* LinearAlgebra.cross: [a2*b3-a3*b2, a3*b1-a1*b3, a1*b2-a2*b1]. */
static inline void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

static inline void sub3(const double *a, const double *b, double *c) {
  c[0] = a[0] - b[0];
  c[1] = a[1] - b[1];
  c[2] = a[2] - b[2];
}

/* LinearAlgebra.norm for a length-3 vector (generic_norm2: plain sum of
 * squares, then sqrt; no scaling for finite non-zero inputs). */
static inline double norm3(const double *a) {
  double s = a[0] * a[0];
  s = s + a[1] * a[1];
  s = s + a[2] * a[2];
  return sqrt(s);
}

/* LinearAlgebra.normalize: v * inv(norm(v)); zero vector -> NaN. */
static inline void normalize3(const double *a, double *o) {
  double n = norm3(a);
  if (n == 0.0) {
    o[0] = o[1] = o[2] = NAN;
    return;
  }
  double inv = 1.0 / n;
  o[0] = a[0] * inv;
  o[1] = a[1] * inv;
  o[2] = a[2] * inv;
}

static inline const double *vtx(const double *X, const int64_t *IEN, int64_t el,
                                int c) {
  return X + 3 * (IEN[3 * el + c] - 1);
}

/* ------------------------------------------------------------------ */
/* grid                                                                */
/* ------------------------------------------------------------------ */

void orc_grid(const double xmin[3], const double xmax[3], int64_t n_max,
              int64_t margin, double aabb_min[3], double aabb_max[3],
              int64_t N[3], double *cell_size, int64_t *ngp) {
  /* cell_size = maximum(AABB_max .- AABB_min) ./ N_max */
  double ext = xmax[0] - xmin[0];
  for (int d = 1; d < 3; d++) {
    double e = xmax[d] - xmin[d];
    if (e > ext) ext = e;
  }
  double h = ext / (double)n_max;
  double mh = (double)margin * h;
  int64_t prod = 1;
  for (int d = 0; d < 3; d++) {
    double lo = xmin[d] - mh;
    double hi = xmax[d] + mh;
    N[d] = (int64_t)ceil((hi - lo) / h);
    aabb_min[d] = lo;
    aabb_max[d] = lo + (double)N[d] * h;
    prod *= N[d] + 1;
  }
  *cell_size = h;
  *ngp = prod;
}

void orc_grid_points(const double aabb_min[3], double h, const int64_t N[3],
                     double *pts) {
  int64_t a = 0;
  for (int64_t k = 0; k <= N[2]; k++)
    for (int64_t j = 0; j <= N[1]; j++)
      for (int64_t i = 0; i <= N[0]; i++) {
        pts[3 * a + 0] = aabb_min[0] + h * (double)i;
        pts[3 * a + 1] = aabb_min[1] + h * (double)j;
        pts[3 * a + 2] = aabb_min[2] + h * (double)k;
        a++;
      }
}

/* ------------------------------------------------------------------ */
/* ray directions                                                      */
/* ------------------------------------------------------------------ */

void orc_ray_directions(double dirs[36 * 3]) {
  static const double raw[26][3] = {
      {1, 0, 0},   {0, 1, 0},   {0, 0, 1},    {-1, 0, 0},  {0, -1, 0},
      {0, 0, -1},  {1, 1, 0},   {1, -1, 0},   {-1, 1, 0},  {-1, -1, 0},
      {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},   {-1, 0, -1}, {0, 1, 1},
      {0, 1, -1},  {0, -1, 1},  {0, -1, -1},  {1, 1, 1},   {1, 1, -1},
      {1, -1, 1},  {1, -1, -1}, {-1, 1, 1},   {-1, 1, -1}, {-1, -1, 1},
      {-1, -1, -1}};
  for (int i = 0; i < 26; i++) normalize3(raw[i], dirs + 3 * i);
  /* golden-angle directions, not normalised (:218-224) */
  double phi = (1.0 + sqrt(5.0)) / 2.0;
  double twopi = 6.283185307179586; /* Float64(2pi) */
  for (int i = 1; i <= 10; i++) {
    double x = (double)i * phi;
    double fr = fmod(x, 1.0); /* mod(x, 1.0) for x > 0 */
    double theta = twopi * fr;
    double z = 1.0 - 2.0 * ((double)i / 10.0);
    double r = sqrt(1.0 - z * z);
    double *d = dirs + 3 * (25 + i);
    d[0] = r * cos(theta);
    d[1] = r * sin(theta);
    d[2] = z;
  }
}

/* ------------------------------------------------------------------ */
/* Moller-Trumbore, reference order                                    */
/* ------------------------------------------------------------------ */

int orc_ray_triangle(const double o[3], const double d[3], const double v1[3],
                     const double v2[3], const double v3[3], double eps,
                     double *t_out) {
  double e1[3], e2[3], h[3], s[3], q[3];
  sub3(v2, v1, e1);
  sub3(v3, v1, e2);
  cross3(d, e2, h);
  double a = dot3(e1, h);
  if (fabs(a) < eps) return 0;
  double f = 1.0 / a;
  sub3(o, v1, s);
  double u = f * dot3(s, h);
  if (u < 0.0 || u > 1.0) return 0;
  cross3(s, e1, q);
  double v = f * dot3(d, q);
  if (v < 0.0 || u + v > 1.0) return 0;
  double t = f * dot3(e2, q);
  if (!(t > eps)) return 0;
  *t_out = t;
  return 1;
}

/* Parity of the deduplicated hit list (RaycastingSignDetection.jl:138-153). */
static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

static int parity_of_hits(double *ts, int64_t n) {
  if (n == 0) return 0;
  qsort(ts, (size_t)n, sizeof(double), cmp_double);
  int64_t unique = 1;
  for (int64_t i = 1; i < n; i++)
    if (fabs(ts[i] - ts[i - 1]) > RAY_EPS) unique++;
  return (int)(unique % 2 == 1);
}

typedef struct {
  double *buf;
  int64_t n, cap;
} tbuf;

static void tbuf_push(tbuf *b, double t) {
  if (b->n == b->cap) {
    b->cap = b->cap ? 2 * b->cap : 64;
    b->buf = (double *)realloc(b->buf, (size_t)b->cap * sizeof(double));
  }
  b->buf[b->n++] = t;
}

static int ray_inside_brute(const double *X, const int64_t *IEN, int64_t nel,
                            const double o[3], const double d[3], tbuf *tb) {
  tb->n = 0;
  for (int64_t el = 0; el < nel; el++) {
    double t;
    if (orc_ray_triangle(o, d, vtx(X, IEN, el, 0), vtx(X, IEN, el, 1),
                         vtx(X, IEN, el, 2), RAY_EPS, &t))
      tbuf_push(tb, t);
  }
  return parity_of_hits(tb->buf, tb->n);
}

int orc_ray_inside(const double *X, const int64_t *IEN, int64_t nel,
                   const double o[3], const double d[3]) {
  tbuf tb = {0, 0, 0};
  int r = ray_inside_brute(X, IEN, nel, o, d, &tb);
  free(tb.buf);
  return r;
}

/* ------------------------------------------------------------------ */
/* bounding-sphere BVH (RaycastingSignDetection.jl:35-56)              */
/* Pruning only drops triangles whose padded bounding sphere the ray   */
/* misses; such triangles cannot produce a hit with t > eps, so the    */
/* result equals the all-triangle loop.                                */
/* ------------------------------------------------------------------ */

typedef struct {
  double lo[3], hi[3];
  int64_t left, right; /* children, or -1 */
  int64_t first, count; /* leaf range into perm */
} bnode;

typedef struct {
  bnode *nodes;
  int64_t nnodes;
  int64_t *perm;
  double *sph; /* 4 per triangle: centre, radius */
} sbvh;

static double *g_sort_key;
static int cmp_key_idx(const void *a, const void *b) {
  double x = g_sort_key[*(const int64_t *)a], y = g_sort_key[*(const int64_t *)b];
  if (x < y) return -1;
  if (x > y) return 1;
  int64_t i = *(const int64_t *)a, j = *(const int64_t *)b;
  return (i > j) - (i < j);
}

static int64_t sbvh_build_rec(sbvh *b, int64_t first, int64_t count,
                              double *key) {
  int64_t id = b->nnodes++;
  bnode *nd = &b->nodes[id];
  for (int d = 0; d < 3; d++) {
    nd->lo[d] = INFINITY;
    nd->hi[d] = -INFINITY;
  }
  double clo[3] = {INFINITY, INFINITY, INFINITY},
         chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = first; i < first + count; i++) {
    const double *s = b->sph + 4 * b->perm[i];
    for (int d = 0; d < 3; d++) {
      if (s[d] - s[3] < nd->lo[d]) nd->lo[d] = s[d] - s[3];
      if (s[d] + s[3] > nd->hi[d]) nd->hi[d] = s[d] + s[3];
      if (s[d] < clo[d]) clo[d] = s[d];
      if (s[d] > chi[d]) chi[d] = s[d];
    }
  }
  if (count <= 4) {
    nd->left = nd->right = -1;
    nd->first = first;
    nd->count = count;
    return id;
  }
  int ax = 0;
  for (int d = 1; d < 3; d++)
    if (chi[d] - clo[d] > chi[ax] - clo[ax]) ax = d;
  for (int64_t i = first; i < first + count; i++)
    key[b->perm[i]] = b->sph[4 * b->perm[i] + ax];
  g_sort_key = key;
  qsort(b->perm + first, (size_t)count, sizeof(int64_t), cmp_key_idx);
  int64_t half = count / 2;
  int64_t l = sbvh_build_rec(b, first, half, key);
  int64_t r = sbvh_build_rec(b, first + half, count - half, key);
  nd = &b->nodes[id];
  nd->left = l;
  nd->right = r;
  nd->first = nd->count = 0;
  return id;
}

static void sbvh_build(sbvh *b, const double *X, const int64_t *IEN,
                       int64_t nel) {
  b->sph = (double *)malloc((size_t)(4 * nel) * sizeof(double));
  b->perm = (int64_t *)malloc((size_t)nel * sizeof(int64_t));
  b->nodes = (bnode *)malloc((size_t)(2 * nel + 1) * sizeof(bnode));
  b->nnodes = 0;
  for (int64_t el = 0; el < nel; el++) {
    const double *v1 = vtx(X, IEN, el, 0), *v2 = vtx(X, IEN, el, 1),
                 *v3 = vtx(X, IEN, el, 2);
    double c[3], t[3];
    for (int d = 0; d < 3; d++) c[d] = ((v1[d] + v2[d]) + v3[d]) / 3.0;
    sub3(v1, c, t);
    double r = norm3(t);
    sub3(v2, c, t);
    double r2 = norm3(t);
    sub3(v3, c, t);
    double r3 = norm3(t);
    if (r2 > r) r = r2;
    if (r3 > r) r = r3;
    double *s = b->sph + 4 * el;
    s[0] = c[0];
    s[1] = c[1];
    s[2] = c[2];
    /* 1.01 padding of the reference plus a relative guard so that the
     * prune stays conservative under rounding */
    s[3] = r * 1.01 + 1e-12 * (fabs(c[0]) + fabs(c[1]) + fabs(c[2]) + r);
    b->perm[el] = el;
  }
  if (nel > 0) {
    double *key = (double *)malloc((size_t)nel * sizeof(double));
    sbvh_build_rec(b, 0, nel, key);
    free(key);
  }
}

static void sbvh_free(sbvh *b) {
  free(b->sph);
  free(b->perm);
  free(b->nodes);
}

static int ray_box(const double o[3], const double d[3], const double lo[3],
                   const double hi[3]) {
  double tmin = 0.0, tmax = INFINITY;
  for (int a = 0; a < 3; a++) {
    double pad = 1e-12 * (fabs(lo[a]) + fabs(hi[a]) + 1.0);
    double l = lo[a] - pad, h = hi[a] + pad;
    if (d[a] == 0.0) {
      if (o[a] < l || o[a] > h) return 0;
      continue;
    }
    double t1 = (l - o[a]) / d[a], t2 = (h - o[a]) / d[a];
    if (t1 > t2) {
      double tt = t1;
      t1 = t2;
      t2 = tt;
    }
    if (t1 > tmin) tmin = t1;
    if (t2 < tmax) tmax = t2;
    if (tmin > tmax * (1.0 + 1e-12) + 1e-300) return 0;
  }
  return 1;
}

static int ray_sphere(const double o[3], const double d[3], const double *s) {
  double oc[3] = {s[0] - o[0], s[1] - o[1], s[2] - o[2]};
  double dd = dot3(d, d);
  double tca = dot3(oc, d) / dd;
  double r2 = s[3] * s[3];
  double oc2 = dot3(oc, oc);
  if (tca < 0.0) return oc2 <= r2 * (1.0 + 1e-9);
  double d2 = oc2 - tca * tca * dd;
  return d2 <= r2 * (1.0 + 1e-9) + 1e-300;
}

static int ray_inside_bvh(const sbvh *b, const double *X, const int64_t *IEN,
                          const double o[3], const double d[3], tbuf *tb) {
  tb->n = 0;
  if (b->nnodes == 0) return 0;
  int64_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const bnode *nd = &b->nodes[stack[--sp]];
    if (!ray_box(o, d, nd->lo, nd->hi)) continue;
    if (nd->left < 0) {
      for (int64_t i = nd->first; i < nd->first + nd->count; i++) {
        int64_t el = b->perm[i];
        if (!ray_sphere(o, d, b->sph + 4 * el)) continue;
        double t;
        if (orc_ray_triangle(o, d, vtx(X, IEN, el, 0), vtx(X, IEN, el, 1),
                             vtx(X, IEN, el, 2), RAY_EPS, &t))
          tbuf_push(tb, t);
      }
    } else {
      stack[sp++] = nd->left;
      stack[sp++] = nd->right;
    }
  }
  return parity_of_hits(tb->buf, tb->n);
}

/* ------------------------------------------------------------------ */
/* winding number                                                      */
/* ------------------------------------------------------------------ */

double orc_winding_sign(const double *X, const int64_t *IEN, int64_t nel,
                        const double p[3]) {
  double w = 0.0;
  for (int64_t el = 0; el < nel; el++) {
    double t[3], a[3], b[3], c[3], bc[3];
    sub3(vtx(X, IEN, el, 0), p, t);
    normalize3(t, a);
    sub3(vtx(X, IEN, el, 1), p, t);
    normalize3(t, b);
    sub3(vtx(X, IEN, el, 2), p, t);
    normalize3(t, c);
    cross3(b, c, bc);
    double num = dot3(a, bc);
    double den = ((1.0 + dot3(a, b)) + dot3(b, c)) + dot3(c, a);
    if (fabs(den) > 1e-10) {
      double sa = 2.0 * atan2(fabs(num), den);
      if (num < 0) sa = -sa;
      w += sa;
    }
  }
  w = w / 12.566370614359172; /* Float64(4pi) */
  return fabs(w) > 0.5 ? 1.0 : -1.0;
}

/* ------------------------------------------------------------------ */
/* adaptive vote + driver                                              */
/* ------------------------------------------------------------------ */

void orc_signs(const double *X, int64_t nnp, const int64_t *IEN, int64_t nel,
               const double *pts, int64_t npts, const double *dirs,
               int min_rays, int max_rays, double conf_thr, int use_winding,
               int use_bvh, double *signs, double *conf, int32_t *rays_used_out,
               int32_t *winding_out, int nthreads) {
  (void)nnp;
  sbvh b;
  memset(&b, 0, sizeof(b));
  if (use_bvh) sbvh_build(&b, X, IEN, nel);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    tbuf tb = {0, 0, 0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int64_t i = 0; i < npts; i++) {
      const double *p = pts + 3 * i;
      int inside = 0, total = 0;
#define CAST(k)                                                              \
  (use_bvh ? ray_inside_bvh(&b, X, IEN, p, dirs + 3 * (k), &tb)              \
           : ray_inside_brute(X, IEN, nel, p, dirs + 3 * (k), &tb))
      for (int k = 0; k < min_rays; k++) {
        if (CAST(k)) inside++;
        total++;
      }
      double ratio = (double)inside / (double)total;
      double c = ratio > 1.0 - ratio ? ratio : 1.0 - ratio;
      int used = min_rays;
      while (c < 0.85 && used < max_rays) {
        int batch = max_rays - used < 4 ? max_rays - used : 4;
        for (int k = used; k < used + batch; k++) {
          if (k < 36) {
            if (CAST(k)) inside++;
            total++;
          }
        }
        used += batch;
        ratio = (double)inside / (double)total;
        c = ratio > 1.0 - ratio ? ratio : 1.0 - ratio;
        if (c > 0.95) break;
      }
#undef CAST
      double sg = ratio > 0.5 ? 1.0 : -1.0;
      int wused = 0;
      if (use_winding && c < conf_thr) {
        sg = orc_winding_sign(X, IEN, nel, p);
        c = 0.8;
        wused = 1;
      }
      signs[i] = sg;
      conf[i] = c;
      if (rays_used_out) rays_used_out[i] = used;
      if (winding_out) winding_out[i] = wused;
    }
    free(tb.buf);
  }
  if (use_bvh) sbvh_free(&b);
}

/* ------------------------------------------------------------------ */
/* exact point-triangle distance                                       */
/* ------------------------------------------------------------------ */

static double seg_dist2(const double p[3], const double a[3], const double b[3],
                        double cp[3]) {
  double ab[3], ap[3];
  sub3(b, a, ab);
  sub3(p, a, ap);
  double l2 = dot3(ab, ab);
  double t = l2 > 0.0 ? dot3(ap, ab) / l2 : 0.0;
  if (t < 0.0) t = 0.0;
  if (t > 1.0) t = 1.0;
  for (int d = 0; d < 3; d++) cp[d] = a[d] + t * ab[d];
  double r[3];
  sub3(p, cp, r);
  return dot3(r, r);
}

/* Closest point on triangle abc to p by Voronoi regions (Ericson, Real-Time
 * Collision Detection, 5.1.5).  Degenerate triangles use the three edges. */
static double tri_dist2(const double p[3], const double a[3], const double b[3],
                        const double c[3], double cp[3]) {
  double ab[3], ac[3], ap[3], bp[3], cpv[3];
  sub3(b, a, ab);
  sub3(c, a, ac);
  sub3(p, a, ap);
  double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  double r[3];
  if (d1 <= 0.0 && d2 <= 0.0) {
    memcpy(cp, a, 3 * sizeof(double));
    goto done;
  }
  sub3(p, b, bp);
  double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) {
    memcpy(cp, b, 3 * sizeof(double));
    goto done;
  }
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    double v = d1 / (d1 - d3);
    for (int d = 0; d < 3; d++) cp[d] = a[d] + v * ab[d];
    goto done;
  }
  sub3(p, c, cpv);
  double d5 = dot3(ab, cpv), d6 = dot3(ac, cpv);
  if (d6 >= 0.0 && d5 <= d6) {
    memcpy(cp, c, 3 * sizeof(double));
    goto done;
  }
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    double w = d2 / (d2 - d6);
    for (int d = 0; d < 3; d++) cp[d] = a[d] + w * ac[d];
    goto done;
  }
  double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int d = 0; d < 3; d++) cp[d] = b[d] + w * (c[d] - b[d]);
    goto done;
  }
  {
    double den = va + vb + vc;
    if (!(fabs(den) > 0.0) || !isfinite(den)) {
      double c1[3], c2[3], c3[3];
      double s1 = seg_dist2(p, a, b, c1), s2 = seg_dist2(p, b, c, c2),
             s3 = seg_dist2(p, c, a, c3);
      double best = s1;
      memcpy(cp, c1, sizeof c1);
      if (s2 < best) {
        best = s2;
        memcpy(cp, c2, sizeof c2);
      }
      if (s3 < best) {
        best = s3;
        memcpy(cp, c3, sizeof c3);
      }
      return best;
    }
    double inv = 1.0 / den;
    double v = vb * inv, w = vc * inv;
    for (int d = 0; d < 3; d++) cp[d] = a[d] + ab[d] * v + ac[d] * w;
  }
done:
  sub3(p, cp, r);
  return dot3(r, r);
}

void orc_dist_exact(const double *X, int64_t nnp, const int64_t *IEN,
                    int64_t nel, const double *pts, int64_t npts, double h,
                    double *dist, double *xp, int nthreads) {
  (void)nnp;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t i = 0; i < npts; i++) {
    const double *p = pts + 3 * i;
    double best = INFINITY, bcp[3] = {0, 0, 0};
    for (int64_t el = 0; el < nel; el++) {
      double cp[3];
      double d2 = tri_dist2(p, vtx(X, IEN, el, 0), vtx(X, IEN, el, 1),
                            vtx(X, IEN, el, 2), cp);
      if (d2 < best) {
        best = d2;
        memcpy(bcp, cp, sizeof bcp);
      }
    }
    double d = sqrt(best);
    dist[i] = d > h ? h : d;
    if (xp) memcpy(xp + 3 * i, bcp, sizeof bcp);
  }
}

/* Squared distance from p to a box (0 inside). */
static double box_dist2(const double p[3], const double lo[3], const double hi[3]) {
  double s = 0.0;
  for (int d = 0; d < 3; d++) {
    double e = 0.0;
    if (p[d] < lo[d]) e = lo[d] - p[d];
    else if (p[d] > hi[d]) e = p[d] - hi[d];
    s += e * e;
  }
  return s;
}

/* min(d, h) of orc_dist_exact, with the bounding-sphere hierarchy of the
 * sign oracle (each node box holds the spheres, each sphere its triangle).
 * The search bound starts at (h (1 + 1e-6))^2 and shrinks to the smallest
 * squared distance found; a node is skipped only when its box distance
 * exceeds the bound by a relative 1e-9, far above the rounding of either
 * distance, so every triangle that could hold the minimum is evaluated with
 * the same tri_dist2.  The result equals orc_dist_exact's bit for bit: the
 * same minimum when it is below the bound, and h otherwise. */
void orc_dist_exact_bvh(const double *X, int64_t nnp, const int64_t *IEN,
                        int64_t nel, const double *pts, int64_t npts, double h,
                        double *dist, int nthreads) {
  (void)nnp;
  sbvh b;
  memset(&b, 0, sizeof(b));
  if (nel > 0) sbvh_build(&b, X, IEN, nel);
  const double hb = h * (1.0 + 1e-6);
  const double bound0 = hb * hb;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    int64_t cap = 256;
    int64_t *stack = (int64_t *)malloc((size_t)cap * sizeof(int64_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int64_t i = 0; i < npts; i++) {
      const double *p = pts + 3 * i;
      double best = INFINITY, bound = bound0;
      int64_t sp = 0;
      if (nel > 0) stack[sp++] = 0;
      while (sp > 0) {
        const bnode *nd = &b.nodes[stack[--sp]];
        if (box_dist2(p, nd->lo, nd->hi) > bound * (1.0 + 1e-9)) continue;
        if (nd->left < 0) {
          for (int64_t q = nd->first; q < nd->first + nd->count; q++) {
            const int64_t el = b.perm[q];
            double cp[3];
            double d2 = tri_dist2(p, vtx(X, IEN, el, 0), vtx(X, IEN, el, 1),
                                  vtx(X, IEN, el, 2), cp);
            if (d2 < best) best = d2;
            if (d2 < bound) bound = d2;
          }
        } else {
          if (sp + 2 > cap) {
            cap *= 2;
            stack = (int64_t *)realloc(stack, (size_t)cap * sizeof(int64_t));
          }
          stack[sp++] = nd->right;
          stack[sp++] = nd->left;
        }
      }
      const double d = sqrt(best);
      dist[i] = d > h ? h : d;
    }
    free(stack);
  }
  if (nel > 0) sbvh_free(&b);
}

/* ------------------------------------------------------------------ */
/* reference-faithful distance                                         */
/* ------------------------------------------------------------------ */

/* kd-tree over 3-D points (vertices or triangle centres) */
typedef struct {
  const double *P; /* 3 x n */
  int64_t *idx;
  int64_t n;
} kdt;

static const double *g_kd_P;
static int g_kd_ax;
static int cmp_kd(const void *a, const void *b) {
  double x = g_kd_P[3 * *(const int64_t *)a + g_kd_ax];
  double y = g_kd_P[3 * *(const int64_t *)b + g_kd_ax];
  return (x > y) - (x < y);
}

/* implicit tree over idx[lo, hi): median at mid, split axis = depth % 3 */
static void kd_build_rec(kdt *t, int64_t lo, int64_t hi, int depth) {
  if (hi - lo <= 1) return;
  g_kd_P = t->P;
  g_kd_ax = depth % 3;
  qsort(t->idx + lo, (size_t)(hi - lo), sizeof(int64_t), cmp_kd);
  int64_t mid = lo + (hi - lo) / 2;
  kd_build_rec(t, lo, mid, depth + 1);
  kd_build_rec(t, mid + 1, hi, depth + 1);
}

static void kd_build(kdt *t, const double *P, int64_t n) {
  t->P = P;
  t->n = n;
  t->idx = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) t->idx[i] = i;
  kd_build_rec(t, 0, n, 0);
}

static double d2pt(const double *a, const double *b) {
  double r[3];
  sub3(a, b, r);
  return dot3(r, r);
}

typedef struct {
  int64_t *v;
  int64_t n, cap;
} ivec;
static void ivec_push(ivec *a, int64_t x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 64;
    a->v = (int64_t *)realloc(a->v, (size_t)a->cap * sizeof(int64_t));
  }
  a->v[a->n++] = x;
}

static void kd_range_rec(const kdt *t, int64_t lo, int64_t hi, int depth,
                         const double *q, double r2, ivec *out) {
  if (hi <= lo) return;
  int64_t mid = lo + (hi - lo) / 2;
  int64_t id = t->idx[mid];
  const double *p = t->P + 3 * id;
  if (d2pt(p, q) <= r2) ivec_push(out, id);
  if (hi - lo == 1) return;
  int ax = depth % 3;
  double diff = q[ax] - p[ax];
  if (diff <= 0 || diff * diff <= r2)
    kd_range_rec(t, lo, mid, depth + 1, q, r2, out);
  if (diff >= 0 || diff * diff <= r2)
    kd_range_rec(t, mid + 1, hi, depth + 1, q, r2, out);
}

static void kd_knn3_rec(const kdt *t, int64_t lo, int64_t hi, int depth,
                        const double *q, int64_t best[3], double bd[3]) {
  if (hi <= lo) return;
  int64_t mid = lo + (hi - lo) / 2;
  int64_t id = t->idx[mid];
  const double *p = t->P + 3 * id;
  double d = d2pt(p, q);
  if (d < bd[2] || (d == bd[2] && id < best[2])) {
    int k = 2;
    while (k > 0 && (d < bd[k - 1] || (d == bd[k - 1] && id < best[k - 1]))) {
      bd[k] = bd[k - 1];
      best[k] = best[k - 1];
      k--;
    }
    bd[k] = d;
    best[k] = id;
  }
  if (hi - lo == 1) return;
  int ax = depth % 3;
  double diff = q[ax] - p[ax];
  if (diff <= 0) {
    kd_knn3_rec(t, lo, mid, depth + 1, q, best, bd);
    if (diff * diff <= bd[2]) kd_knn3_rec(t, mid + 1, hi, depth + 1, q, best, bd);
  } else {
    kd_knn3_rec(t, mid + 1, hi, depth + 1, q, best, bd);
    if (diff * diff <= bd[2]) kd_knn3_rec(t, lo, mid, depth + 1, q, best, bd);
  }
}

static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* barycentricCoordinates, dfOnTriangularMesh.jl:226-251: the 3x3 system
 * solved by Gaussian elimination with partial pivoting. */
static void bary_ref(const double *x1, const double *x2, const double *x3,
                     const double *n, const double *x, double lam[3]) {
  double A[3][3] = {
      {x1[1] * n[2] - x1[2] * n[1], x2[1] * n[2] - x2[2] * n[1],
       x3[1] * n[2] - x3[2] * n[1]},
      {x1[2] * n[0] - x1[0] * n[2], x2[2] * n[0] - x2[0] * n[2],
       x3[2] * n[0] - x3[0] * n[2]},
      {x1[0] * n[1] - x1[1] * n[0], x2[0] * n[1] - x2[1] * n[0],
       x3[0] * n[1] - x3[1] * n[0]}};
  double b[3] = {x[1] * n[2] - x[2] * n[1], x[2] * n[0] - x[0] * n[2],
                 x[0] * n[1] - x[1] * n[0]};
  int im = 0;
  double nm = fabs(n[0]);
  for (int d = 1; d < 3; d++)
    if (fabs(n[d]) > nm) {
      nm = fabs(n[d]);
      im = d;
    }
  A[im][0] = A[im][1] = A[im][2] = 1.0;
  b[im] = 1.0;
  for (int c = 0; c < 3; c++) {
    int pr = c;
    for (int r = c + 1; r < 3; r++)
      if (fabs(A[r][c]) > fabs(A[pr][c])) pr = r;
    if (pr != c) {
      for (int k = 0; k < 3; k++) {
        double tt = A[c][k];
        A[c][k] = A[pr][k];
        A[pr][k] = tt;
      }
      double tt = b[c];
      b[c] = b[pr];
      b[pr] = tt;
    }
    for (int r = c + 1; r < 3; r++) {
      double m = A[r][c] / A[c][c];
      for (int k = c; k < 3; k++) A[r][k] -= m * A[c][k];
      b[r] -= m * b[c];
    }
  }
  for (int r = 2; r >= 0; r--) {
    double s = b[r];
    for (int k = r + 1; k < 3; k++) s -= A[r][k] * lam[k];
    lam[r] = s / A[r][r];
  }
}

void orc_dist_reference(const double *X, int64_t nnp, const int64_t *IEN,
                        int64_t nel, const double *pts, int64_t npts, double h,
                        double *dist, int nthreads) {
  /* prepare_triangle_data (:329-368) */
  double *cen = (double *)malloc((size_t)(3 * nel + 1) * sizeof(double));
  double *nrm = (double *)malloc((size_t)(3 * nel + 1) * sizeof(double));
  for (int64_t el = 0; el < nel; el++) {
    const double *v1 = vtx(X, IEN, el, 0), *v2 = vtx(X, IEN, el, 1),
                 *v3 = vtx(X, IEN, el, 2);
    double e1[3], e3[3], me3[3], n[3];
    sub3(v2, v1, e1);
    sub3(v1, v3, e3);
    for (int d = 0; d < 3; d++) me3[d] = -e3[d];
    cross3(e1, me3, n);
    double nn = norm3(n);
    for (int d = 0; d < 3; d++) {
      nrm[3 * el + d] = n[d] / nn;
      cen[3 * el + d] = ((v1[d] + v2[d]) + v3[d]) / 3.0;
    }
  }
  /* node-to-element connectivity (MeshInformations.jl:39-54) as CSR */
  int64_t *ine_off = (int64_t *)calloc((size_t)(nnp + 1), sizeof(int64_t));
  int64_t *ine = (int64_t *)malloc((size_t)(3 * nel + 1) * sizeof(int64_t));
  for (int64_t el = 0; el < nel; el++)
    for (int c = 0; c < 3; c++) ine_off[IEN[3 * el + c]]++;
  for (int64_t v = 0; v < nnp; v++) ine_off[v + 1] += ine_off[v];
  {
    int64_t *cur = (int64_t *)malloc((size_t)(nnp + 1) * sizeof(int64_t));
    memcpy(cur, ine_off, (size_t)(nnp + 1) * sizeof(int64_t));
    for (int64_t el = 0; el < nel; el++)
      for (int c = 0; c < 3; c++) ine[cur[IEN[3 * el + c] - 1]++] = el;
    free(cur);
  }
  kdt tc, tv;
  kd_build(&tc, cen, nel);
  kd_build(&tv, X, nnp);
  double r = 2.5 * h;
  double r2 = r * r;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    ivec cand = {0, 0, 0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 128)
#endif
    for (int64_t i = 0; i < npts; i++) {
      const double *p = pts + 3 * i;
      cand.n = 0;
      kd_range_rec(&tc, 0, nel, 0, p, r2, &cand);
      if (cand.n == 0 && nnp > 0) {
        int64_t best[3] = {-1, -1, -1};
        double bd[3] = {INFINITY, INFINITY, INFINITY};
        kd_knn3_rec(&tv, 0, nnp, 0, p, best, bd);
        for (int k = 0; k < 3; k++)
          if (best[k] >= 0)
            for (int64_t q = ine_off[best[k]]; q < ine_off[best[k] + 1]; q++)
              ivec_push(&cand, ine[q]);
      }
      qsort(cand.v, (size_t)cand.n, sizeof(int64_t), cmp_i64);
      double min_dist = 1e10; /* abs(-1e10) */
      int64_t prev = -1;
      for (int64_t q = 0; q < cand.n; q++) {
        int64_t el = cand.v[q];
        if (el == prev) continue; /* Set semantics */
        prev = el;
        const double *v[3] = {vtx(X, IEN, el, 0), vtx(X, IEN, el, 1),
                              vtx(X, IEN, el, 2)};
        int reject = 0;
        for (int d = 0; d < 3; d++) {
          double lo = v[0][d], hi = v[0][d];
          for (int c = 1; c < 3; c++) {
            if (v[c][d] < lo) lo = v[c][d];
            if (v[c][d] > hi) hi = v[c][d];
          }
          if (p[d] < lo - min_dist || p[d] > hi + min_dist) reject = 1;
        }
        if (reject) continue;
        double lam[3];
        bary_ref(v[0], v[1], v[2], nrm + 3 * el, p, lam);
        double mn = lam[0];
        if (lam[1] < mn || isnan(lam[1])) mn = lam[1];
        if (lam[2] < mn || isnan(lam[2])) mn = lam[2];
        if (isnan(lam[0])) mn = NAN;
        int face_or_edge = 0;
        if (mn >= 0.0) {
          double xp[3], t[3];
          for (int d = 0; d < 3; d++)
            xp[d] = (lam[0] * v[0][d] + lam[1] * v[1][d]) + lam[2] * v[2][d];
          sub3(p, xp, t);
          double dt = norm3(t);
          if (dt < min_dist) min_dist = dt;
        } else {
          for (int j = 0; j < 3; j++) {
            const double *xv = v[j];
            const double *xn = v[(j + 1) % 3];
            double e[3], eu[3], t[3];
            sub3(xn, xv, e);
            double L = norm3(e);
            for (int d = 0; d < 3; d++) eu[d] = e[d] / L;
            sub3(p, xv, t);
            double P = dot3(t, eu);
            if (P >= 0 && P <= L) {
              double xp[3];
              for (int d = 0; d < 3; d++) xp[d] = xv[d] + eu[d] * P;
              sub3(p, xp, t);
              double dt = norm3(t);
              if (dt < min_dist) {
                min_dist = dt;
                face_or_edge = 1;
              }
            }
          }
          if (!face_or_edge) {
            double best = INFINITY;
            for (int c = 0; c < 3; c++) {
              double t[3];
              sub3(p, v[c], t);
              double dt = norm3(t);
              if (dt < best) best = dt;
            }
            if (best < min_dist) min_dist = best;
          }
        }
      }
      dist[i] = min_dist > h ? h : min_dist;
    }
    free(cand.v);
  }
  free(cen);
  free(nrm);
  free(ine_off);
  free(ine);
  free(tc.idx);
  free(tv.idx);
}

/* ------------------------------------------------------------------ */
/* connected components / artifact removal                             */
/* ------------------------------------------------------------------ */

static int64_t uf_find(int64_t *par, int64_t x) {
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}

static void uf_union(int64_t *par, int64_t a, int64_t b) {
  a = uf_find(par, a);
  b = uf_find(par, b);
  if (a == b) return;
  if (a < b)
    par[b] = a;
  else
    par[a] = b;
}

int64_t orc_components(const double *sdf, const int64_t N[3], double threshold,
                       int64_t *labels) {
  int64_t nx = N[0] + 1, ny = N[1] + 1, nz = N[2] + 1;
  int64_t n = nx * ny * nz;
  for (int64_t i = 0; i < n; i++) labels[i] = sdf[i] >= threshold ? i : -1;
  for (int64_t k = 0; k < nz; k++)
    for (int64_t j = 0; j < ny; j++)
      for (int64_t i = 0; i < nx; i++) {
        int64_t a = (k * ny + j) * nx + i;
        if (labels[a] < 0) continue;
        if (i + 1 < nx && labels[a + 1] >= 0) uf_union(labels, a, a + 1);
        if (j + 1 < ny && labels[a + nx] >= 0) uf_union(labels, a, a + nx);
        if (k + 1 < nz && labels[a + nx * ny] >= 0)
          uf_union(labels, a, a + nx * ny);
      }
  int64_t nc = 0;
  for (int64_t a = 0; a < n; a++)
    if (labels[a] >= 0) {
      labels[a] = uf_find(labels, a);
      if (labels[a] == a) nc++;
    }
  return nc;
}

int64_t orc_remove_artifacts(double *sdf, const int64_t N[3], double threshold,
                             double ratio) {
  int64_t n = (N[0] + 1) * (N[1] + 1) * (N[2] + 1);
  int64_t *lab = (int64_t *)malloc((size_t)n * sizeof(int64_t));
  int64_t nc = orc_components(sdf, N, threshold, lab);
  if (nc == 0) {
    free(lab);
    return 0;
  }
  int64_t *size = (int64_t *)calloc((size_t)n, sizeof(int64_t));
  for (int64_t a = 0; a < n; a++)
    if (lab[a] >= 0) size[lab[a]]++;
  int64_t largest = -1, lsize = 0;
  for (int64_t a = 0; a < n; a++)
    if (size[a] > lsize) {
      lsize = size[a];
      largest = a;
    }
  /* round(Int, ratio * largest): round half to even */
  double x = ratio * (double)lsize;
  double rmin = nearbyint(x); /* default rounding mode: ties to even */
  int64_t min_size = (int64_t)rmin;
  if (min_size < 1) min_size = 1;
  int64_t flipped = 0;
  for (int64_t a = 0; a < n; a++) {
    if (lab[a] < 0) continue;
    if (lab[a] != largest && size[lab[a]] < min_size) {
      sdf[a] = -fabs(sdf[a]);
      flipped++;
    }
  }
  free(lab);
  free(size);
  return flipped;
}

/* ------------------------------------------------------------------ */
/* RBF smoothing                                                       */
/* ------------------------------------------------------------------ */

/* Lattice offsets within the truncation radius R = h*sqrt(-log(threshold))
 * with value exp(-m) > threshold (m = di^2+dj^2+dk^2): the rows of
 * compute_sparse_kernel_matrix (RBFs4Smoothing.jl:142-176) on a uniform grid. */
static int rbf_offsets(double threshold, int off[][3], float *val) {
  int n = 0;
  double R2 = -log(threshold);
  for (int dk = -3; dk <= 3; dk++)
    for (int dj = -3; dj <= 3; dj++)
      for (int di = -3; di <= 3; di++) {
        int m = di * di + dj * dj + dk * dk;
        if ((double)m > R2) continue;
        double v = exp(-(double)m);
        if (!(v > threshold)) continue;
        off[n][0] = di;
        off[n][1] = dj;
        off[n][2] = dk;
        val[n] = (float)v;
        n++;
      }
  return n;
}

static void rbf_apply(const double *x, double *y, const int64_t N[3], int noff,
                      int off[][3], const float *val) {
  int64_t nx = N[0] + 1, ny = N[1] + 1, nz = N[2] + 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int64_t k = 0; k < nz; k++)
    for (int64_t j = 0; j < ny; j++)
      for (int64_t i = 0; i < nx; i++) {
        double s = 0.0;
        for (int q = 0; q < noff; q++) {
          int64_t ii = i + off[q][0], jj = j + off[q][1], kk = k + off[q][2];
          if (ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz)
            continue;
          s += (double)val[q] * x[(kk * ny + jj) * nx + ii];
        }
        y[(k * ny + j) * nx + i] = s;
      }
}

int orc_rbf_weights(const float *coarse, const int64_t N[3], double h,
                    double threshold, double reltol, int maxiter,
                    float *weights, double *resid_out) {
  (void)h;
  int off[400][3];
  float val[400];
  int noff = rbf_offsets(threshold, off, val);
  int64_t n = (N[0] + 1) * (N[1] + 1) * (N[2] + 1);
  double *x = (double *)calloc((size_t)n, sizeof(double));
  double *r = (double *)malloc((size_t)n * sizeof(double));
  double *p = (double *)malloc((size_t)n * sizeof(double));
  double *Ap = (double *)malloc((size_t)n * sizeof(double));
  double bb = 0.0;
  for (int64_t i = 0; i < n; i++) {
    r[i] = coarse[i];
    p[i] = r[i];
    bb += r[i] * r[i];
  }
  double rr = bb;
  double stop = reltol * sqrt(bb);
  int it = 0;
  if (maxiter <= 0) maxiter = (int)(n < 2000000000 ? n : 2000000000);
  while (sqrt(rr) > stop && it < maxiter) {
    rbf_apply(p, Ap, N, noff, off, val);
    double pAp = 0.0;
    for (int64_t i = 0; i < n; i++) pAp += p[i] * Ap[i];
    double alpha = rr / pAp;
    double rr2 = 0.0;
    for (int64_t i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
      rr2 += r[i] * r[i];
    }
    double beta = rr2 / rr;
    rr = rr2;
    for (int64_t i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
    it++;
  }
  /* true residual of the returned weights */
  for (int64_t i = 0; i < n; i++) weights[i] = (float)x[i];
  for (int64_t i = 0; i < n; i++) p[i] = weights[i];
  rbf_apply(p, Ap, N, noff, off, val);
  double res = 0.0;
  for (int64_t i = 0; i < n; i++) {
    double d = Ap[i] - (double)coarse[i];
    res += d * d;
  }
  if (resid_out) *resid_out = sqrt(res) / sqrt(bb > 0 ? bb : 1.0);
  free(x);
  free(r);
  free(p);
  free(Ap);
  return it;
}

void orc_rbf_fine(const float *weights, const double aabb_min[3],
                  const double aabb_max[3], const int64_t N[3], double h,
                  int smooth, double threshold, float *fine, int nthreads) {
  (void)aabb_min;
  (void)aabb_max;
  int64_t nx = N[0] + 1, ny = N[1] + 1, nz = N[2] + 1;
  int64_t fx = N[0] * smooth + 1, fy = N[1] * smooth + 1,
          fz = N[2] * smooth + 1;
  double R2 = -log(threshold); /* in units of h^2 */
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (int64_t c = 0; c < fz; c++)
    for (int64_t b = 0; b < fy; b++)
      for (int64_t a = 0; a < fx; a++) {
        double s = 0.0;
        int64_t i0 = a / smooth, j0 = b / smooth, k0 = c / smooth;
        for (int64_t kk = k0 - 3; kk <= k0 + 4; kk++) {
          if (kk < 0 || kk >= nz) continue;
          double dz = (double)c / smooth - (double)kk;
          for (int64_t jj = j0 - 3; jj <= j0 + 4; jj++) {
            if (jj < 0 || jj >= ny) continue;
            double dy = (double)b / smooth - (double)jj;
            for (int64_t ii = i0 - 3; ii <= i0 + 4; ii++) {
              if (ii < 0 || ii >= nx) continue;
              double dx = (double)a / smooth - (double)ii;
              double m = dx * dx + dy * dy + dz * dz;
              if (m > R2) continue;
              s += (double)weights[(kk * ny + jj) * nx + ii] * exp(-m);
            }
          }
        }
        fine[(c * fy + b) * fx + a] = (float)s;
      }
  (void)h;
}

/* ------------------------------------------------------------------ */
/* binary STL + dedup                                                  */
/* ------------------------------------------------------------------ */

typedef struct {
  int64_t kx, ky, kz;
  int64_t head; /* first unique vertex in this cell */
  int used;
} hslot;

static uint64_t hkey(int64_t a, int64_t b, int64_t c) {
  uint64_t h = (uint64_t)a * 0x9E3779B97F4A7C15ull;
  h ^= (uint64_t)b * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
  h ^= (uint64_t)c * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
  return h;
}

/* process_stl_mesh (ImportStlData.jl:198-309) over nv triangle corners V
 * (3 doubles each): the unique vertices into U (3 per vertex, capacity nv)
 * and each corner's 1-based vertex index into tri; returns the count. */
static int64_t stl_merge(const double *V, int64_t nv, double *U, int64_t *tri) {
  const double tol = 1e-10;
  const double scale = floor(1.0 / tol);
  int64_t cap = 1;
  while (cap < 2 * nv + 16) cap <<= 1;
  hslot *tab = (hslot *)calloc((size_t)cap, sizeof(hslot));
  int64_t *next = (int64_t *)malloc((size_t)(nv + 1) * sizeof(int64_t));
  int64_t nu = 0;
  for (int64_t q = 0; q < nv; q++) {
    double v[3] = {V[3 * q], V[3 * q + 1], V[3 * q + 2]};
    int64_t g[3];
    double pic[3];
    for (int d = 0; d < 3; d++) {
      g[d] = (int64_t)floor(v[d] * scale);
      pic[d] = v[d] * scale - floor(v[d] * scale);
    }
    int64_t found = -1;
    for (int pass = 0; pass < 27 && found < 0; pass++) {
      int dx, dy, dz;
      if (pass == 0) {
        dx = dy = dz = 0;
      } else {
        int m = pass - 1; /* neighbours in Julia order, dz fastest */
        int lin = m < 13 ? m : m + 1; /* skip the centre */
        dx = lin / 9 - 1;
        dy = (lin / 3) % 3 - 1;
        dz = lin % 3 - 1;
        int ok = (dx == -1 && pic[0] < tol * scale) ||
                 (dx == 1 && pic[0] > 1 - tol * scale) ||
                 (dy == -1 && pic[1] < tol * scale) ||
                 (dy == 1 && pic[1] > 1 - tol * scale) ||
                 (dz == -1 && pic[2] < tol * scale) ||
                 (dz == 1 && pic[2] > 1 - tol * scale);
        if (!ok) continue;
      }
      int64_t kx = g[0] + dx, ky = g[1] + dy, kz = g[2] + dz;
      uint64_t hh = hkey(kx, ky, kz) & (uint64_t)(cap - 1);
      while (tab[hh].used) {
        if (tab[hh].kx == kx && tab[hh].ky == ky && tab[hh].kz == kz) {
          for (int64_t u = tab[hh].head; u >= 0; u = next[u]) {
            double r[3] = {v[0] - U[3 * u], v[1] - U[3 * u + 1],
                           v[2] - U[3 * u + 2]};
            if (norm3(r) < tol) {
              found = u;
              break;
            }
          }
          break;
        }
        hh = (hh + 1) & (uint64_t)(cap - 1);
      }
    }
    if (found < 0) {
      found = nu++;
      U[3 * found] = v[0];
      U[3 * found + 1] = v[1];
      U[3 * found + 2] = v[2];
      uint64_t hh = hkey(g[0], g[1], g[2]) & (uint64_t)(cap - 1);
      while (tab[hh].used &&
             !(tab[hh].kx == g[0] && tab[hh].ky == g[1] && tab[hh].kz == g[2]))
        hh = (hh + 1) & (uint64_t)(cap - 1);
      if (!tab[hh].used) {
        tab[hh].used = 1;
        tab[hh].kx = g[0];
        tab[hh].ky = g[1];
        tab[hh].kz = g[2];
        tab[hh].head = -1;
      }
      /* append at the tail to keep first-found order */
      next[found] = -1;
      if (tab[hh].head < 0)
        tab[hh].head = found;
      else {
        int64_t u = tab[hh].head;
        while (next[u] >= 0) u = next[u];
        next[u] = found;
      }
    }
    tri[q] = found + 1;
  }
  free(tab);
  free(next);
  return nu;
}

int64_t orc_process_stl_mesh(const double *V, int64_t nv, double *X, int64_t *IEN) {
  double *U = (double *)malloc((size_t)(3 * nv + 3) * sizeof(double));
  int64_t *tri = (int64_t *)malloc((size_t)(nv + 1) * sizeof(int64_t));
  const int64_t nu = stl_merge(V, nv, U, tri);
  if (X) memcpy(X, U, (size_t)(3 * nu) * sizeof(double));
  if (IEN) memcpy(IEN, tri, (size_t)nv * sizeof(int64_t));
  free(U);
  free(tri);
  return nu;
}

int orc_read_stl_binary(const char *path, double *X, int64_t *nnp,
                        int64_t *IEN, int64_t *nel) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  unsigned char hdr[80];
  uint32_t nt = 0;
  if (fread(hdr, 1, 80, f) != 80 || fread(&nt, 4, 1, f) != 1) {
    fclose(f);
    return -2;
  }
  float *raw = (float *)malloc((size_t)nt * 9 * sizeof(float) + 1);
  for (uint32_t t = 0; t < nt; t++) {
    float rec[12];
    uint16_t attr;
    if (fread(rec, 4, 12, f) != 12 || fread(&attr, 2, 1, f) != 1) {
      free(raw);
      fclose(f);
      return -3;
    }
    memcpy(raw + 9 * (size_t)t, rec + 3, 9 * sizeof(float));
  }
  fclose(f);
  int64_t nv = 3 * (int64_t)nt;
  double *V = (double *)malloc((size_t)(3 * nv + 3) * sizeof(double));
  for (int64_t q = 0; q < 3 * nv; q++) V[q] = (double)raw[q];
  double *U = (double *)malloc((size_t)(3 * nv + 3) * sizeof(double));
  int64_t *tri = (int64_t *)malloc((size_t)(nv + 1) * sizeof(int64_t));
  int64_t nu = stl_merge(V, nv, U, tri);
  free(V);
  int rc = 0;
  if (X && IEN) {
    if (*nnp < nu || *nel < (int64_t)nt)
      rc = -4;
    else {
      memcpy(X, U, (size_t)(3 * nu) * sizeof(double));
      memcpy(IEN, tri, (size_t)nv * sizeof(int64_t));
    }
  }
  *nnp = nu;
  *nel = nt;
  free(raw);
  free(U);
  free(tri);
  return rc;
}
