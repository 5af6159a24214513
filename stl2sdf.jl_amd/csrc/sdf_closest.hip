// Closest point on the mesh for every query point (the xp matrix of
// evalDistancesOnTriMesh, src/SignedDistances/dfOnTriangularMesh.jl:19-209;
// per-candidate projection :108-186, thread merge :194-199).
//
// The reference keeps, per point, the closest projection among its
// candidate triangles (centres within 2.5h, else the triangles around the 3
// nearest vertices), so beyond a few cells its xp depends on that candidate
// set.  Here xp is the exact closest point over all triangles, in FP64:
// Voronoi-region closest point on each triangle (Ericson, Real-Time
// Collision Detection 5.1.5; degenerate triangles fall back to their three
// edges), and the winner is the smallest (squared distance, triangle index)
// pair.  That winner does not depend on the traversal order, so the result
// is reproducible bit for bit by a brute-force loop with the same
// arithmetic.
//
// One wavefront serves 64 points (an 8 x 4 x 2 lattice tile, or 64
// consecutive caller points) and walks the LBVH depth-first, nearer child
// first by majority vote of its lanes.  A child is entered when any lane's
// bound admits its box; the node index is wave-uniform, so node and
// triangle records come through scalar loads and every lane evaluates every
// visited leaf.
#include "sdf_internal.h"
#include "sdf_lbvh.h"

namespace sdf {

namespace {

constexpr int kClosestStack = 128;  // > 30 Morton bits + log2(triangles) + 1

#pragma clang fp contract(off)

__device__ inline void csub(const double *a, const double *b, double *c) {
  c[0] = a[0] - b[0];
  c[1] = a[1] - b[1];
  c[2] = a[2] - b[2];
}

__device__ inline double cdot(const double *a, const double *b) {
  double s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

// Closest point of segment ab to p; returns the squared distance.
__device__ inline double seg_closest(const double *p, const double *a, const double *b,
                                     double *q) {
  double ab[3], ap[3], r[3];
  csub(b, a, ab);
  csub(p, a, ap);
  const double den = cdot(ab, ab);
  double t = den > 0.0 ? cdot(ap, ab) / den : 0.0;
  t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  for (int d = 0; d < 3; d++) q[d] = a[d] + t * ab[d];
  csub(p, q, r);
  return cdot(r, r);
}

// Closest point of triangle abc to p by Voronoi regions; returns the
// squared distance |p - q|^2.  Every product and sum rounded (no fused
// multiply-add), so the value is the same on any IEEE FP64 unit.
__device__ inline double tri_closest(const double *p, const double *a, const double *b,
                                     const double *c, double *q) {
  double ab[3], ac[3], ap[3], r[3];
  csub(b, a, ab);
  csub(c, a, ac);
  csub(p, a, ap);
  const double d1 = cdot(ab, ap), d2 = cdot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) {
    for (int d = 0; d < 3; d++) q[d] = a[d];
  } else {
    double bp[3];
    csub(p, b, bp);
    const double d3 = cdot(ab, bp), d4 = cdot(ac, bp);
    const double vc = d1 * d4 - d3 * d2;
    if (d3 >= 0.0 && d4 <= d3) {
      for (int d = 0; d < 3; d++) q[d] = b[d];
    } else if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
      const double v = d1 / (d1 - d3);
      for (int d = 0; d < 3; d++) q[d] = a[d] + v * ab[d];
    } else {
      double cpv[3];
      csub(p, c, cpv);
      const double d5 = cdot(ab, cpv), d6 = cdot(ac, cpv);
      const double vb = d5 * d2 - d1 * d6;
      const double va = d3 * d6 - d5 * d4;
      if (d6 >= 0.0 && d5 <= d6) {
        for (int d = 0; d < 3; d++) q[d] = c[d];
      } else if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double w = d2 / (d2 - d6);
        for (int d = 0; d < 3; d++) q[d] = a[d] + w * ac[d];
      } else if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int d = 0; d < 3; d++) q[d] = b[d] + w * (c[d] - b[d]);
      } else {
        const double den = va + vb + vc;
        if (!(fabs(den) > 0.0) || !isfinite(den)) {
          double c1[3], c2[3], c3[3];
          const double s1 = seg_closest(p, a, b, c1), s2 = seg_closest(p, b, c, c2),
                       s3 = seg_closest(p, c, a, c3);
          double best = s1;
          for (int d = 0; d < 3; d++) q[d] = c1[d];
          if (s2 < best) {
            best = s2;
            for (int d = 0; d < 3; d++) q[d] = c2[d];
          }
          if (s3 < best) {
            best = s3;
            for (int d = 0; d < 3; d++) q[d] = c3[d];
          }
          return best;
        }
        const double inv = 1.0 / den;
        const double v = vb * inv, w = vc * inv;
        for (int d = 0; d < 3; d++) q[d] = a[d] + ab[d] * v + ac[d] * w;
      }
    }
  }
  csub(p, q, r);
  return cdot(r, r);
}

#pragma clang fp contract(fast)

// Pruning bound of one lane: a box is admitted when its FP32 distance to
// the rounded point is within lim = sqrt(best2) (1 + 2^-20) + 2^-20 |p|inf.
// The margin covers the rounding of p to FP32 (2^-24 |p| per coordinate)
// and of the FP32 box distance, so the box of any triangle at FP64 squared
// distance <= best2 (and every ancestor box) is always admitted.
__device__ inline float lane_lim2(double best2, float pmax) {
  if (!(best2 < __builtin_huge_val())) return __builtin_huge_valf();
  const float lim = (float)sqrt(best2) * (1.0f + 0x1p-20f) + 0x1p-20f * pmax;
  return lim * lim * (1.0f + 0x1p-20f);
}

template <bool kLattice>
__global__ __launch_bounds__(64) void closest_points_kernel(
    GridP g, TileDims td, const double *__restrict__ pts, int64_t npts, int64_t nwaves,
    const Tri64 *__restrict__ tri, Bvh bvh, double *__restrict__ xp,
    uint32_t *__restrict__ err) {
  __shared__ int32_t stack[kClosestStack];
  const int lane = threadIdx.x;
  // consecutive waves of one round on one XCD (its own L2)
  const int64_t first = xcd_remap(blockIdx.x, gridDim.x);
  for (int64_t w = first; w < nwaves; w += gridDim.x) {
    bool active;
    int64_t idx;
    double p[3];
    if (kLattice) {
      int32_t tx, ty, tz;
      tile_coords((uint32_t)w, td, &tx, &ty, &tz);
      const int32_t i = tx * kTileX + (lane & (kTileX - 1));
      const int32_t j = ty * kTileY + ((lane / kTileX) & (kTileY - 1));
      const int32_t k = tz * kTileZ + lane / (kTileX * kTileY);
      active = i < g.nx && j < g.ny && k < g.nz;
      idx = (int64_t)i + (int64_t)g.nx * ((int64_t)j + (int64_t)g.ny * k);
      const Vec3d v = lattice_point(g, i, j, k);
      p[0] = v.x;
      p[1] = v.y;
      p[2] = v.z;
    } else {
      idx = w * kTilePts + lane;
      active = idx < npts;
      const int64_t r = active ? idx : 0;
      p[0] = pts[3 * r];
      p[1] = pts[3 * r + 1];
      p[2] = pts[3 * r + 2];
    }
    const float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
    const float pmax = fmaxf(fmaxf(fabsf(pf[0]), fabsf(pf[1])), fabsf(pf[2]));
    double best2 = __builtin_huge_val();
    uint32_t bt = 0xFFFFFFFFu;
    double q[3] = {0.0, 0.0, 0.0};
    float lim2 = active ? __builtin_huge_valf() : -1.0f;
    const uint64_t act = __ballot(active);
    auto leaf = [&](uint32_t t) {
      const Tri64 T = tri[t];
      double c[3];
      const double d2 = tri_closest(p, T.v, T.v + 3, T.v + 6, c);
      if (active && (d2 < best2 || (d2 == best2 && t < bt))) {
        best2 = d2;
        bt = t;
        q[0] = c[0];
        q[1] = c[1];
        q[2] = c[2];
        lim2 = lane_lim2(best2, pmax);
      }
    };
    if (bvh.n == 1) {
      leaf(bvh.leaf_tri[0]);
    } else if (bvh.n > 1) {
      int sp = 0;
      int32_t node = 0;
      while (true) {
        const BvhNode nd = bvh.nodes[node];
        const float e0 = box_dist2(pf, nd.box[0]), e1 = box_dist2(pf, nd.box[1]);
        const int votes1 = __popcll(__ballot(active && e1 < e0));
        const bool swap = 2 * votes1 > __popcll(act);
        const int32_t c0 = swap ? nd.child[1] : nd.child[0];
        const int32_t c1 = swap ? nd.child[0] : nd.child[1];
        const float f0 = swap ? e1 : e0, f1 = swap ? e0 : e1;
        int32_t next = -1;
        if (__ballot(f0 <= lim2)) {
          if (c0 < 0)
            leaf((uint32_t)(-(c0 + 1)));
          else
            next = c0;
        }
        if (__ballot(f1 <= lim2)) {  // lim2 may have shrunk in the first leaf
          if (c1 < 0) {
            leaf((uint32_t)(-(c1 + 1)));
          } else if (next < 0) {
            next = c1;
          } else {
            if (sp >= kClosestStack) {
              if (lane == 0) atomicOr(err, kErrStackOverflow);
              break;
            }
            if (lane == 0) stack[sp] = c1;
            sp++;
          }
        }
        if (next >= 0) {
          node = next;
        } else if (sp > 0) {
          sp--;  // LDS accesses of one wave complete in order
          node = __builtin_amdgcn_readfirstlane(stack[sp]);
        } else {
          break;
        }
      }
    }
    if (active) {
      xp[3 * idx] = q[0];
      xp[3 * idx + 1] = q[1];
      xp[3 * idx + 2] = q[2];
    }
  }
}

}  // namespace

void launch_closest_points(hipStream_t s, const GridP &g, const double *pts, int64_t npts,
                           const Tri64 *tri, const Bvh &bvh, double *xp, uint32_t *err) {
  const bool lattice = pts == nullptr;
  const TileDims td = tile_dims(g);
  const int64_t nwaves = lattice ? td.ntiles : (npts + kTilePts - 1) / kTilePts;
  if (nwaves <= 0) return;
  const int64_t grid = nwaves < 65536 * 8 ? nwaves : 65536 * 8;
  if (lattice)
    hipLaunchKernelGGL(closest_points_kernel<true>, dim3((unsigned)grid), dim3(64), 0, s, g, td,
                       pts, npts, nwaves, tri, bvh, xp, err);
  else
    hipLaunchKernelGGL(closest_points_kernel<false>, dim3((unsigned)grid), dim3(64), 0, s, g,
                       td, pts, npts, nwaves, tri, bvh, xp, err);
  SDF_HIP(hipGetLastError());
}

}  // namespace sdf
