// Shared device/host helpers for the MI355X SDF kernels.
//
// Exactness rules used throughout:
//  * The sign predicates (Moller-Trumbore, RaycastingSignDetection.jl:64-106)
//    are evaluated in FP64 with the reference's operation order and with FP
//    contraction switched off, so every product and sum is rounded exactly
//    like the reference's separate operations.
//  * Lattice coordinates are AABB_min[d] + h * i in FP64 without contraction
//    (src/MeshGrid/Grid.jl:90).
//  * Distances are computed in FP32 on coordinates taken relative to a
//    nearby lattice point (the difference itself formed in FP64), which keeps
//    the absolute error far below 1e-6 x bbox diagonal.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdf {

constexpr double kRayEps = 1e-8;  // RaycastingSignDetection.jl:118,132
constexpr int kNumDirs = 36;      // RaycastingSignDetection.jl:162-227

// Lattice description: points per axis (N+1), origin, spacing.
struct GridP {
  double ox, oy, oz;
  double h;
  int32_t nx, ny, nz;
  int32_t pad_;
};

// Distance tiles: 8 x 4 x 2 lattice points = one 64-lane wavefront.
constexpr int kTileX = 8, kTileY = 4, kTileZ = 2;
constexpr int kTilePts = kTileX * kTileY * kTileZ;

struct TileDims {
  int32_t tx, ty, tz;  // tiles per axis
  int64_t ntiles;
};

__host__ __device__ inline TileDims tile_dims(const GridP &g) {
  TileDims t;
  t.tx = (g.nx + kTileX - 1) / kTileX;
  t.ty = (g.ny + kTileY - 1) / kTileY;
  t.tz = (g.nz + kTileZ - 1) / kTileZ;
  t.ntiles = (int64_t)t.tx * t.ty * t.tz;
  return t;
}

// Lattice coordinate AABB_min[d] + h * i (Grid.jl:90), no contraction.
__host__ __device__ inline double lattice_coord(double o, double h, int64_t i) {
#pragma clang fp contract(off)
  double hi = h * (double)i;
  return o + hi;
}

struct Vec3d {
  double x, y, z;
};

__host__ __device__ inline Vec3d lattice_point(const GridP &g, int64_t i,
                                               int64_t j, int64_t k) {
  Vec3d p;
  p.x = lattice_coord(g.ox, g.h, i);
  p.y = lattice_coord(g.oy, g.h, j);
  p.z = lattice_coord(g.oz, g.h, k);
  return p;
}

// ---------------------------------------------------------------------
// Reference-order FP64 vector helpers (no contraction).
// dot(a,b) = ((a1*b1 + a2*b2) + a3*b3); cross as LinearAlgebra.cross.
// ---------------------------------------------------------------------
__host__ __device__ inline double dot3_ref(const double *a, const double *b) {
#pragma clang fp contract(off)
  double s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

__host__ __device__ inline void cross3_ref(const double *a, const double *b,
                                           double *c) {
#pragma clang fp contract(off)
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Moller-Trumbore exactly as RaycastingSignDetection.jl:64-106 with eps=1e-8.
// Returns true and t on a hit with t > eps.
__host__ __device__ inline bool mt_ref(const double *o, const double *d,
                                       const double *v1, const double *v2,
                                       const double *v3, double *t_out) {
#pragma clang fp contract(off)
  double e1[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
  double e2[3] = {v3[0] - v1[0], v3[1] - v1[1], v3[2] - v1[2]};
  double hv[3];
  cross3_ref(d, e2, hv);
  double a = dot3_ref(e1, hv);
  if (fabs(a) < kRayEps) return false;
  double f = 1.0 / a;
  double s[3] = {o[0] - v1[0], o[1] - v1[1], o[2] - v1[2]};
  double u = f * dot3_ref(s, hv);
  if (u < 0.0 || u > 1.0) return false;
  double q[3];
  cross3_ref(s, e1, q);
  double v = f * dot3_ref(d, q);
  if (v < 0.0 || u + v > 1.0) return false;
  double t = f * dot3_ref(e2, q);
  if (!(t > kRayEps)) return false;
  *t_out = t;
  return true;
}

// Axis-line hit record: everything MT needs to evaluate t for any origin on
// the line (v1, e1, e2 and f = 1/a for the +axis direction).
struct LineHit {
  double v1[3];
  double e1[3];
  double e2[3];
  double f;
};

// The record line_hit_test writes for a hit of the triangle on an axis line
// (it does not depend on the line), by the same arithmetic.
__host__ __device__ inline void line_hit_record(int axis, const double *v1, const double *v2,
                                                const double *v3, LineHit *rec) {
#pragma clang fp contract(off)
  double d[3] = {0.0, 0.0, 0.0};
  d[axis] = 1.0;
  double e1[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
  double e2[3] = {v3[0] - v1[0], v3[1] - v1[1], v3[2] - v1[2]};
  double hv[3];
  cross3_ref(d, e2, hv);
  const double a = dot3_ref(e1, hv);
  for (int c = 0; c < 3; c++) {
    rec->v1[c] = v1[c];
    rec->e1[c] = e1[c];
    rec->e2[c] = e2[c];
  }
  rec->f = 1.0 / a;
}

// u/v/|a| part of MT for the +axis direction on an axis-aligned grid line.
// For d = +e_axis those quantities do not depend on the origin's coordinate
// along the axis (the along-axis component only enters through products with
// an exact zero), so one test per (line, triangle) decides the hit for every
// lattice point of the line, and for the -axis ray as well (a, f negate
// exactly; u, v are bitwise equal; t negates exactly).
__host__ __device__ inline bool line_hit_test(const double *o, int axis,
                                              const double *v1,
                                              const double *v2,
                                              const double *v3, LineHit *rec) {
#pragma clang fp contract(off)
  double d[3] = {0.0, 0.0, 0.0};
  d[axis] = 1.0;
  double e1[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
  double e2[3] = {v3[0] - v1[0], v3[1] - v1[1], v3[2] - v1[2]};
  double hv[3];
  cross3_ref(d, e2, hv);
  double a = dot3_ref(e1, hv);
  if (fabs(a) < kRayEps) return false;
  double f = 1.0 / a;
  double s[3] = {o[0] - v1[0], o[1] - v1[1], o[2] - v1[2]};
  double u = f * dot3_ref(s, hv);
  if (u < 0.0 || u > 1.0) return false;
  double q[3];
  cross3_ref(s, e1, q);
  double v = f * dot3_ref(d, q);
  if (v < 0.0 || u + v > 1.0) return false;
  if (rec) {
    for (int c = 0; c < 3; c++) {
      rec->v1[c] = v1[c];
      rec->e1[c] = e1[c];
      rec->e2[c] = e2[c];
    }
    rec->f = f;
  }
  return true;
}

// Certified summary of one hit for the per-point fast path: the crossing
// coordinate c along the line (c - x approximates the reference's t for the
// point at coordinate x), a rigorous bound delta on |t(x) - (c - x)| over the
// whole line, and the hit's cluster index among the line's sorted hits
// (clusters: maximal runs whose consecutive gaps are <= eps).
struct LineFast {
  double c;
  float delta;
  uint16_t cid;
  uint16_t pad_;
};
constexpr uint32_t kLineExact = 0x80000000u;  // line meta flag: no fast path
constexpr int kLineFastMax = 64;             // lines with more hits: exact path

// Sign bricks: 8 x 8 lattice columns x kSignZ points per 64-lane wave.
constexpr int kSignZ = 4;
// Points per sign-brick segment of an x-, y- and z-line.
constexpr int kSegLen[3] = {8, 8, kSignZ};

// Inline copy of a short line's summary (up to 4 hits) indexed directly by
// line id: one 64-byte load instead of offsets + records.  info: bits 0-7
// hit count (255 = more than 254), 8-15 cluster count, bit 30 kLineSeg,
// bit 31 kLineExact.
//
// Sign bricks cut a line into segments of kSegLen[axis] points (segment s
// = points [s L, min(s L + L, n) - 1]).  With kLineSeg, seg[h] = a | b << 8
// says where crossing h stands for every segment: for s < a it lies above
// the segment (c - hi_s > eps + delta), for s >= b below it (c - lo_s <
// -eps - delta), for a <= s < b neither (the segment is not uniform).  The
// rounded differences are monotone in s, so these are the same decisions
// the per-segment comparisons make.
constexpr int kInlineHits = 4;
constexpr uint32_t kLineSeg = 0x40000000u;
struct __attribute__((aligned(16))) LineInline {
  double c[kInlineHits];
  float delta[kInlineHits];
  uint8_t cid[kInlineHits];
  uint32_t info;
  uint16_t seg[kInlineHits];
};

// The integer tail of a LineInline (cid, info, seg: its last 16 bytes), all
// a kLineSeg uniformity test reads, in one 16-byte load.
struct LineTail {
  uint32_t cid4;          // cid[h] in byte h
  uint32_t info;
  uint32_t seg01, seg23;  // seg[0] | seg[1] << 16, seg[2] | seg[3] << 16
};
static_assert(sizeof(LineInline) == 64 && __builtin_offsetof(LineInline, cid) == 48 &&
                  __builtin_offsetof(LineInline, info) == 52 &&
                  __builtin_offsetof(LineInline, seg) == 56,
              "LineTail mirrors the last 16 bytes of LineInline");
__device__ inline LineTail load_line_tail(const LineInline *p) {
  const uint4 v = *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(p) + 48);
  return LineTail{v.x, v.y, v.z, v.w};
}

// Both rays of a line for every point whose coordinate lies in [lo, hi],
// when no crossing is within its rounding margin of that range: each
// crossing is then right of every point (fl(c - hi) > eps + D) or left of
// every point (fl(c - lo) < -eps - D).  Rounded subtraction is monotone, so
// every point's comparisons in inline_parity come out the same as at the
// range's ends, none of them ambiguous; the parities are shared.  Returns
// false when some crossing is near the range (or the summary is not
// usable); the caller then classifies each point.
__host__ __device__ inline bool inline_uniform(const LineInline &r, double lo, double hi, int *pos,
                                      int *neg) {
  // selects, not branches: the per-hit cases would otherwise cost scalar
  // exec-mask work on every call
  const uint32_t n = r.info & 0xFFu;
  const bool usable = !(r.info & kLineExact) && n <= (uint32_t)kInlineHits;
  const int ncl = (int)((r.info >> 8) & 0xFFu);
  int first_pos = -1, last_neg = -1;
  bool uni = true;
#pragma unroll
  for (int h2 = 0; h2 < kInlineHits; h2++) {
    const bool valid = h2 < (int)n;
    const double D = (double)r.delta[h2];
    const int cid = r.cid[h2];
    const bool above = r.c[h2] - hi > kRayEps + D;
    const bool below = !above & (r.c[h2] - lo < -kRayEps - D);
    first_pos = (valid & above & (first_pos < 0)) ? cid : first_pos;
    last_neg = (valid & below) ? cid : last_neg;
    uni = uni & (!valid | above | below);
  }
  *pos = first_pos >= 0 ? ((ncl - first_pos) & 1) : 0;
  *neg = last_neg >= 0 ? ((last_neg + 1) & 1) : 0;
  return usable & uni;
}

// inline_uniform for segment s of the line from the crossings' segment
// ranges (kLineSeg): the same decisions by integer compares.
__host__ __device__ inline bool seg_uniform(const LineTail &r, int s, int *pos, int *neg) {
  const uint32_t n = r.info & 0xFFu;
  const bool usable = !(r.info & kLineExact) && n <= (uint32_t)kInlineHits;
  const int ncl = (int)((r.info >> 8) & 0xFFu);
  int first_pos = -1, last_neg = -1;
  bool uni = true;
#pragma unroll
  for (int h2 = 0; h2 < kInlineHits; h2++) {
    const bool valid = h2 < (int)n;
    const int cid = (int)((r.cid4 >> (8 * h2)) & 0xFFu);
    const uint32_t sg = ((h2 < 2 ? r.seg01 : r.seg23) >> (16 * (h2 & 1))) & 0xFFFFu;
    const int a = (int)(sg & 0xFFu), b = (int)(sg >> 8);
    const bool above = s < a;
    const bool below = !above & (s >= b);
    first_pos = (valid & above & (first_pos < 0)) ? cid : first_pos;
    last_neg = (valid & below) ? cid : last_neg;
    uni = uni & (!valid | above | below);
  }
  *pos = first_pos >= 0 ? ((ncl - first_pos) & 1) : 0;
  *neg = last_neg >= 0 ? ((last_neg + 1) & 1) : 0;
  return usable & uni;
}

// Sign-brick segment bytes: for every axis line and every segment of it
// (kSegLen points), bit 0 is set when the segment's points compare every
// crossing of the line the same way, so they share the line's two ray
// parities, which bits 1 (+axis ray) and 2 (-axis ray) hold; a segment
// that is not uniform has byte 0.  Four consecutive segments of a line share
// a 32-bit word (segment s in byte s % 4 of word s / 4).  Layout, in words:
// x-lines [wx][ny nz], then y-lines [wy][nx nz], then z-lines [wz][nx ny],
// each line index as in the line lists (x: k ny + j, y: k nx + i,
// z: j nx + i), w = ceil(segments / 4).
struct SegLayout {
  int64_t off[3];  // first word of each axis' part
  int64_t nl[3];   // lines per axis (the stride between words of a line)
  int32_t ns[3];   // segments per line
  int32_t pad_;
  int64_t total;   // words
};
__host__ __device__ inline SegLayout seg_layout(const GridP &g) {
  SegLayout sl;
  sl.nl[0] = (int64_t)g.ny * g.nz;
  sl.nl[1] = (int64_t)g.nx * g.nz;
  sl.nl[2] = (int64_t)g.nx * g.ny;
  const int32_t na[3] = {g.nx, g.ny, g.nz};
  int64_t off = 0;
  for (int a = 0; a < 3; a++) {
    sl.ns[a] = (na[a] + kSegLen[a] - 1) / kSegLen[a];
    sl.off[a] = off;
    off += (int64_t)((sl.ns[a] + 3) / 4) * sl.nl[a];
  }
  sl.pad_ = 0;
  sl.total = off;
  return sl;
}

// Bits [0, a) of a 128-bit mask (lo, hi), a in 0..128.
__host__ __device__ inline void low_bits128(int a, uint64_t &lo, uint64_t &hi) {
  lo = a >= 64 ? ~0ull : (a <= 0 ? 0ull : ((1ull << a) - 1ull));
  hi = a >= 128 ? ~0ull : (a <= 64 ? 0ull : ((1ull << (a - 64)) - 1ull));
}

// Bits 0-3 of a nibble to bits 0, 8, 16, 24.
__host__ __device__ inline uint32_t spread_nibble(uint32_t x) {
  return (x & 1u) | ((x & 2u) << 7) | ((x & 4u) << 14) | ((x & 8u) << 21);
}

// The segment words of a line with segment ranges (kLineSeg) and ns <= 128
// segments: seg_uniform for every segment at once, as 128-bit masks.
// Crossing h is above the segments of A_h = [0, a_h) and below those of
// B_h = [b_h, ns) minus A_h; a segment is uniform when every crossing is
// above or below it, its +ray parity comes from the first crossing above it
// (A_h minus the earlier A's), its -ray parity from the last one below it
// (B_h minus the later B's).  put(w, word) receives words 0 .. (ns+3)/4-1.
template <typename Put>
__host__ __device__ inline void seg_words_masks(const LineTail &r, int ns, Put put,
                                                uint64_t *same = nullptr) {
  const uint32_t n = r.info & 0xFFu;
  const bool usable = !(r.info & kLineExact) && n <= (uint32_t)kInlineHits;
  const int ncl = (int)((r.info >> 8) & 0xFFu);
  uint64_t Nlo, Nhi;
  low_bits128(ns, Nlo, Nhi);
  uint64_t ulo = usable ? Nlo : 0ull, uhi = usable ? Nhi : 0ull;
  uint64_t plo = 0ull, phi = 0ull, qlo = 0ull, qhi = 0ull;
  uint64_t salo = 0ull, sahi = 0ull;  // union of the earlier A's
  uint64_t blo[kInlineHits], bhi[kInlineHits];
  int cid[kInlineHits];
#pragma unroll
  for (int h2 = 0; h2 < kInlineHits; h2++) {
    const bool valid = usable && h2 < (int)n;
    cid[h2] = (int)((r.cid4 >> (8 * h2)) & 0xFFu);
    const uint32_t sg = ((h2 < 2 ? r.seg01 : r.seg23) >> (16 * (h2 & 1))) & 0xFFFFu;
    uint64_t alo, ahi, lbl, lbh;
    low_bits128((int)(sg & 0xFFu), alo, ahi);
    low_bits128((int)(sg >> 8), lbl, lbh);
    alo &= Nlo;
    ahi &= Nhi;
    const uint64_t b_lo = Nlo & ~lbl & ~alo, b_hi = Nhi & ~lbh & ~ahi;
    blo[h2] = valid ? b_lo : 0ull;
    bhi[h2] = valid ? b_hi : 0ull;
    if (valid) {
      ulo &= alo | b_lo;
      uhi &= ahi | b_hi;
      if ((ncl - cid[h2]) & 1) {
        plo |= alo & ~salo;
        phi |= ahi & ~sahi;
      }
      salo |= alo;
      sahi |= ahi;
    }
  }
  uint64_t sblo = 0ull, sbhi = 0ull;  // union of the later B's
#pragma unroll
  for (int h2 = kInlineHits - 1; h2 >= 0; h2--) {
    if ((cid[h2] + 1) & 1) {
      qlo |= blo[h2] & ~sblo;
      qhi |= bhi[h2] & ~sbhi;
    }
    sblo |= blo[h2];
    sbhi |= bhi[h2];
  }
  if (same) {  // segments whose byte is 1 (u, neither ray odd) and 7 (u, both odd)
    same[0] = ulo & ~plo & ~qlo;
    same[1] = uhi & ~phi & ~qhi;
    same[2] = ulo & plo & qlo;
    same[3] = uhi & phi & qhi;
  }
  const int nw = (ns + 3) / 4;
  for (int w = 0; w < nw; w++) {
    const int s0 = 4 * w;
    const uint32_t u = (uint32_t)((s0 < 64 ? ulo >> s0 : uhi >> (s0 - 64)) & 0xFull);
    const uint32_t p = (uint32_t)((s0 < 64 ? plo >> s0 : phi >> (s0 - 64)) & 0xFull);
    const uint32_t q = (uint32_t)((s0 < 64 ? qlo >> s0 : qhi >> (s0 - 64)) & 0xFull);
    put(w, spread_nibble(u) | (spread_nibble(u & p) << 1) | (spread_nibble(u & q) << 2));
  }
}

// t for origin o on the +axis ray of a recorded line hit (reference order).
__host__ __device__ inline double line_hit_t(const LineHit &r, const double *o) {
#pragma clang fp contract(off)
  double s[3] = {o[0] - r.v1[0], o[1] - r.v1[1], o[2] - r.v1[2]};
  double q[3];
  cross3_ref(s, r.e1, q);
  return r.f * dot3_ref(r.e2, q);
}

// ---------------------------------------------------------------------
// Point-triangle squared distance (closest point by Voronoi regions,
// Ericson, Real-Time Collision Detection 5.1.5).  T = float or double.
// a, b, c and p share one local origin.
// ---------------------------------------------------------------------
template <typename T>
__host__ __device__ inline T tri_dist2(T px, T py, T pz, T ax, T ay, T az, T bx,
                                       T by, T bz, T cx, T cy, T cz) {
  T abx = bx - ax, aby = by - ay, abz = bz - az;
  T acx = cx - ax, acy = cy - ay, acz = cz - az;
  T apx = px - ax, apy = py - ay, apz = pz - az;
  T d1 = abx * apx + aby * apy + abz * apz;
  T d2 = acx * apx + acy * apy + acz * apz;
  T qx, qy, qz;
  if (d1 <= T(0) && d2 <= T(0)) {
    qx = ax;
    qy = ay;
    qz = az;
  } else {
    T bpx = px - bx, bpy = py - by, bpz = pz - bz;
    T d3 = abx * bpx + aby * bpy + abz * bpz;
    T d4 = acx * bpx + acy * bpy + acz * bpz;
    T cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    T d5 = abx * cpx + aby * cpy + abz * cpz;
    T d6 = acx * cpx + acy * cpy + acz * cpz;
    T vc = d1 * d4 - d3 * d2;
    T vb = d5 * d2 - d1 * d6;
    T va = d3 * d6 - d5 * d4;
    if (d3 >= T(0) && d4 <= d3) {
      qx = bx;
      qy = by;
      qz = bz;
    } else if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
      T v = d1 / (d1 - d3);
      qx = ax + v * abx;
      qy = ay + v * aby;
      qz = az + v * abz;
    } else if (d6 >= T(0) && d5 <= d6) {
      qx = cx;
      qy = cy;
      qz = cz;
    } else if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
      T w = d2 / (d2 - d6);
      qx = ax + w * acx;
      qy = ay + w * acy;
      qz = az + w * acz;
    } else if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
      T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      qx = bx + w * (cx - bx);
      qy = by + w * (cy - by);
      qz = bz + w * (cz - bz);
    } else {
      T den = va + vb + vc;
      if (!(den > T(0) && den - den == T(0))) {
        // degenerate triangle: nearest of the three edges
        T best = T(1e30f);
        const T P[3][3] = {{ax, ay, az}, {bx, by, bz}, {cx, cy, cz}};
        for (int e = 0; e < 3; e++) {
          const T *s0 = P[e], *s1 = P[(e + 1) % 3];
          T ex = s1[0] - s0[0], ey = s1[1] - s0[1], ez = s1[2] - s0[2];
          T wx = px - s0[0], wy = py - s0[1], wz = pz - s0[2];
          T l2 = ex * ex + ey * ey + ez * ez;
          T t = l2 > T(0) ? (wx * ex + wy * ey + wz * ez) / l2 : T(0);
          t = t < T(0) ? T(0) : (t > T(1) ? T(1) : t);
          T rx = wx - t * ex, ry = wy - t * ey, rz = wz - t * ez;
          T dd = rx * rx + ry * ry + rz * rz;
          best = dd < best ? dd : best;
        }
        return best;
      }
      T inv = T(1) / den;
      T v = vb * inv, w = vc * inv;
      qx = ax + abx * v + acx * w;
      qy = ay + aby * v + acy * w;
      qz = az + abz * v + acz * w;
    }
  }
  T rx = px - qx, ry = py - qy, rz = pz - qz;
  return rx * rx + ry * ry + rz * rz;
}

// Triangle vertices in FP64, 9 doubles per triangle (v1, v2, v3).
struct Tri64 {
  double v[9];
};

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Map launch index b of B so that XCD x receives the contiguous range
// [x*B/8, (x+1)*B/8) of work items: neighbouring tiles then share one L2.
constexpr int kNumXcd = 8;
// Tile index -> tile coordinates with 32-bit divisions (tile counts stay
// below 2^31).
__device__ inline void tile_coords(uint32_t tile, const TileDims &td, int32_t *tx,
                                   int32_t *ty, int32_t *tz) {
  const uint32_t q = tile / (uint32_t)td.tx;
  *tx = (int32_t)(tile - q * (uint32_t)td.tx);
  *ty = (int32_t)(q % (uint32_t)td.ty);
  *tz = (int32_t)(q / (uint32_t)td.ty);
}

__device__ inline int64_t xcd_remap(int64_t b, int64_t B) {
  // 32-bit arithmetic: launch sizes stay below 2^31 blocks
  const uint32_t bb = (uint32_t)b, BB = (uint32_t)B;
  const uint32_t per = BB / kNumXcd, rem = BB % kNumXcd;
  const uint32_t x = bb % kNumXcd, local = bb / kNumXcd;
  return (int64_t)(x * per + (x < rem ? x : rem) + local);
}

// Cache-blocked tile order: bricks of 16x16x16 tiles (bricks x-fastest),
// Morton order inside a brick.  Neighbouring launch indices then touch
// neighbouring tiles in all three directions, so the axis lines a brick
// needs (about 1 MB of hit data) stay in one XCD's L2.  Returns false for
// the padding positions of partial bricks.
constexpr int kBrickBits = 4;  // 16 tiles per brick edge
__host__ __device__ inline int64_t brick_launch_size(const TileDims &td) {
  const int64_t e = 1 << kBrickBits;
  return ((td.tx + e - 1) / e) * ((td.ty + e - 1) / e) * ((td.tz + e - 1) / e) *
         (e * e * e);
}
__host__ __device__ inline bool brick_tile(int64_t m, const TileDims &td, int32_t *tx,
                                  int32_t *ty, int32_t *tz) {
  // 32-bit decode (brick counts and positions stay below 2^31)
  const int32_t e = 1 << kBrickBits;
  const uint32_t brick = (uint32_t)(m >> (3 * kBrickBits));
  const uint32_t w = (uint32_t)(m & ((1 << (3 * kBrickBits)) - 1));
  const uint32_t nbx = (uint32_t)((td.tx + e - 1) / e), nby = (uint32_t)((td.ty + e - 1) / e);
  const uint32_t q = brick / nbx;
  const int32_t bx = (int32_t)(brick - q * nbx), by = (int32_t)(q % nby),
                bz = (int32_t)(q / nby);
  int32_t lx = 0, ly = 0, lz = 0;
#pragma unroll
  for (int b = 0; b < kBrickBits; b++) {
    lx |= ((w >> (3 * b + 0)) & 1) << b;
    ly |= ((w >> (3 * b + 1)) & 1) << b;
    lz |= ((w >> (3 * b + 2)) & 1) << b;
  }
  *tx = bx * e + lx;
  *ty = by * e + ly;
  *tz = bz * e + lz;
  return *tx < td.tx && *ty < td.ty && *tz < td.tz;
}

// Launch-order position (brick_tile's m) of the brick with linear index b
// (x fastest).
__host__ __device__ inline int64_t brick_order_of(uint32_t b, const TileDims &td) {
  const int32_t e = 1 << kBrickBits;
  const uint32_t tx = (uint32_t)td.tx, ty = (uint32_t)td.ty;
  const uint32_t x = b % tx, y = (b / tx) % ty, z = b / (tx * ty);
  const uint32_t nbx = (tx + e - 1) / e, nby = (ty + e - 1) / e;
  const uint32_t blk = ((z >> kBrickBits) * nby + (y >> kBrickBits)) * nbx + (x >> kBrickBits);
  uint32_t w = 0;
#pragma unroll
  for (int k = 0; k < kBrickBits; k++)
    w |= (((x >> k) & 1u) << (3 * k)) | (((y >> k) & 1u) << (3 * k + 1)) |
         (((z >> k) & 1u) << (3 * k + 2));
  return ((int64_t)blk << (3 * kBrickBits)) | (int64_t)w;
}

// Per-triangle data for the branch-free point-triangle distance, relative
// to a tile origin (FP32).  Fields, each 3 floats unless noted: a, b, c,
// ab, bc, ca, il (1/|edge|^2, 0 for a zero-length edge), n = ab x (c - a),
// inn (1 float: 1/|n|^2, 0 when degenerate), mab = n x ab, mbc, mca.
constexpr int kTriPreFloats = 34;

__device__ inline void tri_pre_make(const float *a, const float *b,
                                    const float *c, float *o) {
  float ab[3], bc[3], ca[3], ac[3], n[3];
  for (int k = 0; k < 3; k++) {
    ab[k] = b[k] - a[k];
    bc[k] = c[k] - b[k];
    ca[k] = a[k] - c[k];
    ac[k] = c[k] - a[k];
  }
  n[0] = ab[1] * ac[2] - ab[2] * ac[1];
  n[1] = ab[2] * ac[0] - ab[0] * ac[2];
  n[2] = ab[0] * ac[1] - ab[1] * ac[0];
  const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  const float *E[3] = {ab, bc, ca};
  for (int k = 0; k < 3; k++) {
    o[k] = a[k];
    o[3 + k] = b[k];
    o[6 + k] = c[k];
    o[9 + k] = ab[k];
    o[12 + k] = bc[k];
    o[15 + k] = ca[k];
    o[21 + k] = n[k];
  }
  for (int e = 0; e < 3; e++) {
    const float l2 = E[e][0] * E[e][0] + E[e][1] * E[e][1] + E[e][2] * E[e][2];
    o[18 + e] = l2 > 0.f ? 1.0f / l2 : 0.f;
    float m[3] = {n[1] * E[e][2] - n[2] * E[e][1], n[2] * E[e][0] - n[0] * E[e][2],
                  n[0] * E[e][1] - n[1] * E[e][0]};
    for (int k = 0; k < 3; k++) o[25 + 3 * e + k] = m[k];
  }
  o[24] = nn > 0.f ? 1.0f / nn : 0.f;
}

// Squared distance from p to the triangle: the closest point lies in the
// face interior (all three edge functions (p - v) . (n x edge) >= 0, then
// the plane distance) or on one of the three edges (clamped projection);
// the minimum of those four candidates, without branches.  Split so that
// points sharing y and z (a row) share the work that does not depend on x:
// tri_pre_yz forms each dot product's y and z terms, tri_pre_dist2_x adds
// the x term (every caller goes through these, so all kernels round alike).
// T(i) reads field i of the precomputed record.
struct TriYZ {
  float pay, paz, pby, pbz, pcy, pcz;  // p - a, p - b, p - c (y, z)
  float kwc, kwa, kwb, kpn;            // edge functions and plane: y + z terms
  float k0, k1, k2;                    // edge projections: y + z terms
};
struct TriX {  // the record's fields the x part reads
  float ax, bx, cx, wcx, wax, wbx, nx, inn;
  float e0[3], il0, e1[3], il1, e2[3], il2;
};
template <typename Get>
__device__ inline TriX tri_pre_x(Get T) {
  TriX r;
  r.ax = T(0);
  r.bx = T(3);
  r.cx = T(6);
  r.wcx = T(25);
  r.wax = T(28);
  r.wbx = T(31);
  r.nx = T(21);
  r.inn = T(24);
  for (int k = 0; k < 3; k++) {
    r.e0[k] = T(9 + k);
    r.e1[k] = T(12 + k);
    r.e2[k] = T(15 + k);
  }
  r.il0 = T(18);
  r.il1 = T(19);
  r.il2 = T(20);
  return r;
}
// (explicit fused multiply-adds, no contraction left to the compiler: every
// kernel that inlines these rounds exactly alike)
template <typename Get>
__device__ inline TriYZ tri_pre_yz(float py, float pz, Get T) {
#pragma clang fp contract(off)
  TriYZ q;
  q.pay = py - T(1);
  q.paz = pz - T(2);
  q.pby = py - T(4);
  q.pbz = pz - T(5);
  q.pcy = py - T(7);
  q.pcz = pz - T(8);
  q.kwc = __builtin_fmaf(q.pay, T(26), q.paz * T(27));
  q.kwa = __builtin_fmaf(q.pby, T(29), q.pbz * T(30));
  q.kwb = __builtin_fmaf(q.pcy, T(32), q.pcz * T(33));
  q.kpn = __builtin_fmaf(q.pay, T(22), q.paz * T(23));
  q.k0 = __builtin_fmaf(q.pay, T(10), q.paz * T(11));
  q.k1 = __builtin_fmaf(q.pby, T(13), q.pbz * T(14));
  q.k2 = __builtin_fmaf(q.pcy, T(16), q.pcz * T(17));
  return q;
}
// distance to the segment from v (p - v given) along e, 1/|e|^2 = il
__device__ inline float seg_dist2_x(float pvx, float pvy, float pvz, float k, const float *e,
                                    float il) {
#pragma clang fp contract(off)
  const float t = fminf(fmaxf(__builtin_fmaf(pvx, e[0], k) * il, 0.f), 1.f);
  const float rx = __builtin_fmaf(-t, e[0], pvx);
  const float ry = __builtin_fmaf(-t, e[1], pvy);
  const float rz = __builtin_fmaf(-t, e[2], pvz);
  return __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, rz * rz));
}
__device__ inline float tri_pre_dist2_x(float px, const TriX &r, const TriYZ &q) {
#pragma clang fp contract(off)
  const float pax = px - r.ax, pbx = px - r.bx, pcx = px - r.cx;
  const float wc = __builtin_fmaf(pax, r.wcx, q.kwc);
  const float wa = __builtin_fmaf(pbx, r.wax, q.kwa);
  const float wb = __builtin_fmaf(pcx, r.wbx, q.kwb);
  const float pn = __builtin_fmaf(pax, r.nx, q.kpn);
  const bool inside = (wa >= 0.f) & (wb >= 0.f) & (wc >= 0.f) & (r.inn > 0.f);
  const float face = inside ? (pn * pn) * r.inn : __builtin_huge_valf();
  const float d0 = seg_dist2_x(pax, q.pay, q.paz, q.k0, r.e0, r.il0);
  const float d1 = seg_dist2_x(pbx, q.pby, q.pbz, q.k1, r.e1, r.il1);
  const float d2 = seg_dist2_x(pcx, q.pcy, q.pcz, q.k2, r.e2, r.il2);
  return fminf(fminf(face, d0), fminf(d1, d2));
}
template <typename Get>
__device__ inline float tri_pre_dist2(float px, float py, float pz, Get T) {
  return tri_pre_dist2_x(px, tri_pre_x(T), tri_pre_yz(py, pz, T));
}

// Number of work items when the host holds only a capacity and the device
// holds the actual count (null: the capacity is the count).
// Read-only view of device memory that nothing writes while the reading
// kernel runs: wave-uniform loads through it are scalar loads (scalar cache,
// their own wait counter, so waiting for them never waits for the wave's
// vector loads and stores).
template <typename T>
using KConst = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ inline KConst<T> kconst(const T *p) {
  return (KConst<T>)p;
}

__device__ inline int64_t chunk_count(int64_t cap, const uint64_t *dev) {
  if (!dev) return cap;
  const int64_t n = (int64_t)*kconst(dev);
  return n < cap ? n : cap;
}


struct SignDims {
  int32_t sx, sy, sz;
};

// What the sign kernel needs of one position m of the brick launch order,
// worked out before it runs (brick_plan_kernel): the brick's coordinates
// (xy = bx | by << 16, z = bz), its four distance slots (the first work
// item of each of its 1 x 2 x 2 distance tiles, or the spare slot
// chunk_cap for a tile without work) and flags (bit 0: some tile has
// distance work; bit 1: m is a brick of the grid, not padding).  32 bytes,
// one scalar load.
struct BrickPlan {
  uint32_t slot[4];
  uint32_t xy, z, flags, pad_;
};
static_assert(sizeof(BrickPlan) == 32, "one 8-dword scalar load per brick");

__host__ __device__ inline SignDims sign_dims(const GridP &g) {
  SignDims d;
  d.sx = (g.nx + 7) / 8;
  d.sy = (g.ny + 7) / 8;
  d.sz = (g.nz + kSignZ - 1) / kSignZ;
  return d;
}

}  // namespace sdf
