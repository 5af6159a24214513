// LBVH layout and device traversal routines (ray parity, nearest triangle).
#pragma once

#include "sdf_internal.h"

namespace sdf {

// Internal node: both children's boxes (lo xyz, hi xyz) so one 64-byte
// fetch tests two children.  child >= 0: internal node; child < 0: the
// leaf of triangle -(child+1).
struct BvhNode {
  float box[2][6];
  int32_t child[2];
  int32_t pad_[2];
};

// Four-wide view of a binary hierarchy: wide node i holds the children of
// binary node i's internal children and its leaf children as they are, so
// one 128-byte fetch tests up to four boxes and a traversal from the root
// visits every other level.  child: >= 0 node index, kNoChild for an empty
// slot, else a leaf reference as in the binary nodes.
constexpr int32_t kNoChild = INT32_MIN;

struct alignas(128) BvhWide {
  float box[4][6];
  int32_t child[4];
  int32_t pad_[4];
};

// Triangles per leaf of the ray hierarchy.
constexpr int kRayLeaf = 8;

struct Bvh {
  int64_t n = 0;               // leaves (triangles)
  BvhNode *nodes = nullptr;    // n-1 internal nodes, root = 0 (null unless built full)
  float *leaf_box = nullptr;   // 6 floats per triangle (by triangle id)
  uint32_t *leaf_tri = nullptr;  // sorted leaf -> triangle id
  // The ray hierarchy: leaves are runs of kRayLeaf consecutive sorted
  // triangles (run c = leaf_tri[c kRayLeaf ..]); nc - 1 internal nodes
  // (root 0) whose leaf children are -(c+1).  nc == 1: no nodes.
  int64_t nc = 0;
  BvhNode *cnodes = nullptr;
  BvhWide *cwide = nullptr;  // four-wide view of cnodes (nc - 1 nodes)
};

// Morton quantisation box: lo corner and extent per axis (extent > 0).
struct QBox {
  float lo[3], ext[3];
};

// qbox null: the box of the triangles' centroids (one reduction pass).
// The ray hierarchy is always built; the per-triangle hierarchy (nodes,
// for nearest-point queries) only when full.
Bvh build_lbvh(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t n,
               const QBox *qbox = nullptr, bool full = true);
// Stable LSD sort of (key, value) pairs by the low key_bits bits; n is a
// capacity when n_dev (the actual count, on the device) is given.  With
// keys_out, the sorted pairs are left in whichever buffers the last pass
// wrote (the inputs or the arena's scratch pair), returned there.
void radix_sort_u32_pairs(hipStream_t s, Arena &ar, uint32_t *keys,
                          uint32_t *vals, int64_t n, int key_bits,
                          const uint64_t *n_dev = nullptr, uint32_t **keys_out = nullptr,
                          uint32_t **vals_out = nullptr);

constexpr int kStackDepth = 64;

// Slab test in FP64 on a box padded by a relative 1e-7 (conservative: a
// triangle the FP64 predicates hit is never pruned).  Only t >= 0 matters.
__device__ inline bool ray_box(const double *o, const double *d,
                               const double *inv, const float *b) {
  double tmin = 0.0, tmax = __builtin_huge_val();
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double lo = (double)b[a], hi = (double)b[3 + a];
    double pad = 1e-7 * (fabs(lo) + fabs(hi)) + 1e-30;
    lo -= pad;
    hi += pad;
    if (d[a] == 0.0) {
      if (o[a] < lo || o[a] > hi) return false;
      continue;
    }
    double t1 = (lo - o[a]) * inv[a], t2 = (hi - o[a]) * inv[a];
    double tn = fmin(t1, t2), tf = fmax(t1, t2);
    tmin = fmax(tmin, tn);
    tmax = fmin(tmax, tf);
  }
  return tmin <= tmax;
}

// Squared distance from p (FP32, absolute) to a box.
__device__ inline float box_dist2(const float *p, const float *b) {
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float e = fmaxf(fmaxf(b[a] - p[a], p[a] - b[3 + a]), 0.f);
    s += e * e;
  }
  return s;
}

// Nearest triangle within radius r of p: returns min(exact distance, r)
// (FP32 arithmetic on FP64 differences).  Sets *found when a triangle lies
// within r.  Box pruning is padded by `slack` (absolute) to stay
// conservative under FP32 rounding of p.
__device__ inline double nearest_bvh(const Bvh &bvh, const Tri64 *__restrict__ tri,
                                     const double *p, double r, float slack,
                                     bool *found, bool *overflow) {
  float pf[3] = {(float)p[0], (float)p[1], (float)p[2]};
  float best = (float)r;  // current search radius
  float best2 = best * best;
  bool any = false;
  auto leaf = [&](uint32_t t) {
    const Tri64 q = tri[t];
    float d2 = tri_dist2<float>(
        0.f, 0.f, 0.f, (float)(q.v[0] - p[0]), (float)(q.v[1] - p[1]),
        (float)(q.v[2] - p[2]), (float)(q.v[3] - p[0]), (float)(q.v[4] - p[1]),
        (float)(q.v[5] - p[2]), (float)(q.v[6] - p[0]), (float)(q.v[7] - p[1]),
        (float)(q.v[8] - p[2]));
    if (d2 < best2) {
      best2 = d2;
      any = true;
    }
  };
  *overflow = false;
  if (bvh.n == 0) {
    *found = false;
    return r;
  }
  if (bvh.n == 1) {
    leaf(bvh.leaf_tri[0]);
  } else {
    int32_t stack[kStackDepth];
    int sp = 0;
    int32_t node = 0;
    while (true) {
      const BvhNode nd = bvh.nodes[node];
      float lim = sqrtf(best2) + slack;
      float lim2 = lim * lim;
      float d0 = box_dist2(pf, nd.box[0]);
      float d1 = box_dist2(pf, nd.box[1]);
      int32_t c0 = nd.child[0], c1 = nd.child[1];
      if (d1 < d0) {  // visit the nearer child first
        float tf = d0;
        d0 = d1;
        d1 = tf;
        int32_t tc = c0;
        c0 = c1;
        c1 = tc;
      }
      int32_t next = -0x7FFFFFFF;
      if (d0 <= lim2) {
        if (c0 < 0)
          leaf((uint32_t)(-(c0 + 1)));
        else
          next = c0;
      }
      if (d1 <= lim2) {
        if (c1 < 0) {
          leaf((uint32_t)(-(c1 + 1)));
        } else if (next == -0x7FFFFFFF) {
          next = c1;
        } else {
          if (sp >= kStackDepth) {
            *overflow = true;
            break;
          }
          stack[sp++] = c1;
        }
      }
      if (next != -0x7FFFFFFF) {
        node = next;
      } else if (sp > 0) {
        node = stack[--sp];
      } else {
        break;
      }
    }
  }
  *found = any;
  if (!any) return r;
  double dd = sqrt((double)best2);
  return dd < r ? dd : r;
}


// Smallest leaf value over the triangles whose boxes lie within
// min(sqrt(best), rcap) + slack of p (absolute FP32 coordinates); leaf(t)
// returns a squared distance computed by the caller's own arithmetic.
// Returns +inf when no leaf was evaluated.
template <typename Leaf>
__device__ inline float nearest_bvh_with(const Bvh &bvh, const float *pf, float rcap,
                                         float slack, Leaf leaf, bool *overflow) {
  float best2 = __builtin_huge_valf();
  *overflow = false;
  if (bvh.n == 0) return best2;
  if (bvh.n == 1) return leaf(bvh.leaf_tri[0]);
  int32_t stack[kStackDepth];
  int sp = 0;
  int32_t node = 0;
  while (true) {
    const BvhNode nd = bvh.nodes[node];
    const float lim = fminf(sqrtf(best2), rcap) + slack;
    const float lim2 = lim * lim;
    float d0 = box_dist2(pf, nd.box[0]);
    float d1 = box_dist2(pf, nd.box[1]);
    int32_t c0 = nd.child[0], c1 = nd.child[1];
    if (d1 < d0) {
      const float tf = d0;
      d0 = d1;
      d1 = tf;
      const int32_t tc = c0;
      c0 = c1;
      c1 = tc;
    }
    int32_t next = -0x7FFFFFFF;
    if (d0 <= lim2) {
      if (c0 < 0)
        best2 = fminf(best2, leaf((uint32_t)(-(c0 + 1))));
      else
        next = c0;
    }
    if (d1 <= lim2) {
      if (c1 < 0) {
        best2 = fminf(best2, leaf((uint32_t)(-(c1 + 1))));
      } else if (next == -0x7FFFFFFF) {
        next = c1;
      } else {
        if (sp >= kStackDepth) {
          *overflow = true;
          break;
        }
        stack[sp++] = c1;
      }
    }
    if (next != -0x7FFFFFFF) {
      node = next;
    } else if (sp > 0) {
      node = stack[--sp];
    } else {
      break;
    }
  }
  return best2;
}

}  // namespace sdf
