// Query kernels: the fused lattice kernel (distance + six axis rays + combine)
// and the adaptive-vote finisher for points the axis rays leave undecided.
//
// Reference path replaced (jezekon/stl2sdf.jl):
//   evalDistancesOnTriMesh      src/SignedDistances/dfOnTriangularMesh.jl:19-209
//   raycast_sign_detection      src/SignedDistances/RaycastingSignDetection.jl:337-404
//   adaptive_point_in_mesh_raycast / check_ray_inside_bvh  (:234-288, :114-154)
//   generalized_winding_number  (:295-328)
//   sdf_dists = dists .* signs  src/stl_to_sdf.jl:119
#include "sdf_internal.h"
#include "sdf_lbvh.h"

namespace sdf {

namespace {

constexpr int kHitCap = 2048;  // per-ray hit buffer of the finisher (LDS)

// Fast path for one axis line: classify every hit of the line for the point
// at coordinate x from its certified crossing (c - x vs +-eps, margin delta)
// and count clusters.  Returns false when any decision falls inside a margin
// or the line is flagged exact; the caller then evaluates the reference's
// arithmetic.  +dir accepts t > eps (a suffix in crossing order), -dir
// accepts t < -eps (a prefix).
__device__ inline bool axis_pair_parity_fast(const LineFast *__restrict__ fast,
                                             uint64_t hb, uint64_t he, uint32_t meta,
                                             double x, int *pos_par, int *neg_par) {
  if (meta & kLineExact) return false;
  const int ncl = (int)(meta & 0xFFFFu);
  int first_pos = -1, last_neg = -1;
  for (uint64_t q = hb; q < he; q++) {
    const LineFast r = fast[q];
    const double A = r.c - x, D = (double)r.delta;
    if (A > kRayEps + D) {
      if (first_pos < 0) first_pos = r.cid;
    } else if (!(A < kRayEps - D)) {
      return false;
    }
    if (A < -kRayEps - D) {
      last_neg = r.cid;
    } else if (!(A > -kRayEps + D)) {
      return false;
    }
  }
  *pos_par = first_pos >= 0 ? ((ncl - first_pos) & 1) : 0;
  *neg_par = last_neg >= 0 ? ((last_neg + 1) & 1) : 0;
  return true;
}

// Fast path on a line's inline summary for NQ points at coordinates xs
// (same rules as axis_pair_parity_fast).  ok[q] turns false for a point
// whose decisions fall inside a margin; the whole call returns false when
// the summary is not usable (more than kInlineHits hits, or exact line).
template <int NQ>
__device__ inline bool inline_parity(const LineInline &r, const double *xs,
                                     int *pos, int *neg, bool *ok) {
  const uint32_t n = r.info & 0xFFu;
  if ((r.info & kLineExact) || n > (uint32_t)kInlineHits) return false;
  const int ncl = (int)((r.info >> 8) & 0xFFu);
  int first_pos[NQ], last_neg[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    first_pos[q] = -1;
    last_neg[q] = -1;
    ok[q] = true;
  }
#pragma unroll
  for (int h2 = 0; h2 < kInlineHits; h2++) {
    if (h2 < (int)n) {
      const double D = (double)r.delta[h2];
      const int cid = r.cid[h2];
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const double A = r.c[h2] - xs[q];
        if (A > kRayEps + D) {
          if (first_pos[q] < 0) first_pos[q] = cid;
        } else if (!(A < kRayEps - D)) {
          ok[q] = false;
        }
        if (A < -kRayEps - D) {
          last_neg[q] = cid;
        } else if (!(A > -kRayEps + D)) {
          ok[q] = false;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    pos[q] = first_pos[q] >= 0 ? ((ncl - first_pos[q]) & 1) : 0;
    neg[q] = last_neg[q] >= 0 ? ((last_neg[q] + 1) & 1) : 0;
  }
  return true;
}

// LDS of one distance workgroup: the chunk's triangle records, the
// (row, triangle) jobs of phase 2 and the per-point minima.
constexpr int kTileRows = kTileY * kTileZ;
struct DistShared {
  float pre[kTriPreFloats + 6][kTilePts];
  uint16_t job[kTileRows * kTilePts];  // row << 8 | triangle slot
  uint32_t best[kTilePts];             // squared distances as ordered bits
};

// One distance work item (tile, chunk of <= clen triangles) for the 64-lane
// block: the chunk's triangles are staged in LDS as precomputed records, a
// row test keeps each row's candidates, and the exact distances of every
// (row, candidate) job are spread evenly over the wave, 8 lanes (the row's
// points) per job.  Uniform per block (contains barriers).
// bounded: the caller has checked item against the item count.
__device__ inline void dist_item(const DistJob &J, int64_t item, DistShared &S,
                                 bool bounded = false) {
  const GridP &g = J.g;
  const TileDims &td = J.td;
  const Tri64 *__restrict__ tri = J.tri;
  const uint32_t *__restrict__ tile_tris = J.tile_tris;
  float *__restrict__ d2buf = J.d2buf;
  if (!bounded && item >= chunk_count(J.nchunks, J.nchunks_dev)) return;
  // the item's header, wave-uniform and read-only here: one 16-byte scalar
  // load (as two words: tile | first << 32, range)
  static_assert(sizeof(DistItem) == 16, "DistItem is two 64-bit words");
  const auto hp = kconst(reinterpret_cast<const uint64_t *>(J.items)) + 2 * item;
  const uint64_t h0 = hp[0], range = hp[1];
  const int lane = threadIdx.x;
  const uint32_t tile = (uint32_t)h0;
  const uint64_t first = h0 >> 32;
  const bool single = (range >> 55) & 1ull;  // the tile's only work item
  if (J.only && kconst(J.only)[first] == 0) return;  // re-run restricted to flagged tiles
  int32_t tx, ty, tz;
  tile_coords(tile, td, &tx, &ty, &tz);
  const int32_t i0 = tx * kTileX, j0 = ty * kTileY, k0 = tz * kTileZ;
  const Vec3d T0 = lattice_point(g, i0, j0, k0);
  const uint64_t b = range & kItemBeginMask;
  const int n = (int)((range >> 48) & 127ull);
  const float hf = (float)g.h;
  const float H2 = hf * hf * (1.0f + 0x1p-10f);
  // Phase 1, one triangle per lane: which of the tile's 8 rows of 8 points
  // (fixed j, k; row r = (j - j0) + 4 (k - k0)) can hold a point within h
  // of the triangle.  The lower bound is the larger of the squared distance
  // from the row's box to the triangle's box and the squared distance of
  // the row segment to the triangle's plane (0 when its ends straddle the
  // plane); it is below every row point's own bound, and inflated by 2^-10
  // to cover its rounding against the exact evaluation, so the nearest
  // triangle of every point within h is kept.
  uint32_t rowbits = 0;
  if (lane < n) {
    const Tri64 t = tri[tile_tris[b + lane]];
    float v[9];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[3 * k + 0] = (float)(t.v[3 * k + 0] - T0.x);
      v[3 * k + 1] = (float)(t.v[3 * k + 1] - T0.y);
      v[3 * k + 2] = (float)(t.v[3 * k + 2] - T0.z);
    }
    float rec[kTriPreFloats];
    tri_pre_make(v, v + 3, v + 6, rec);
#pragma unroll
    for (int f = 0; f < kTriPreFloats; f++) S.pre[f][lane] = rec[f];
    float bmin[3], bmax[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {  // the triangle's box
      bmin[a] = fminf(fminf(v[a], v[3 + a]), v[6 + a]);
      bmax[a] = fmaxf(fmaxf(v[a], v[3 + a]), v[6 + a]);
      S.pre[kTriPreFloats + a][lane] = bmin[a];
      S.pre[kTriPreFloats + 3 + a][lane] = bmax[a];
    }
    // row coordinates relative to the tile origin, as the lanes form them
    const float x1 = (float)(lattice_coord(g.ox, g.h, i0 + kTileX - 1) - T0.x);
    const float ex = fmaxf(fmaxf(bmin[0] - x1, -bmax[0]), 0.f);  // rows span x in [0, x1]
    const float nx = rec[21], ny = rec[22], nz = rec[23], inn = rec[24];
    const float s0x = (0.f - rec[0]) * nx, s1x = (x1 - rec[0]) * nx;
#pragma unroll
    for (int zz = 0; zz < kTileZ; zz++) {
      const float pz = (float)(lattice_coord(g.oz, g.h, k0 + zz) - T0.z);
      const float ez = fmaxf(fmaxf(bmin[2] - pz, pz - bmax[2]), 0.f);
      const float cz = (pz - rec[2]) * nz;
#pragma unroll
      for (int yy = 0; yy < kTileY; yy++) {
        const float py = (float)(lattice_coord(g.oy, g.h, j0 + yy) - T0.y);
        const float ey = fmaxf(fmaxf(bmin[1] - py, py - bmax[1]), 0.f);
        const float box2 = ex * ex + ey * ey + ez * ez;
        const float c = (py - rec[1]) * ny + cz;
        const float s0 = s0x + c, s1 = s1x + c;
        const float sm = ((s0 > 0.f) == (s1 > 0.f)) ? fminf(fabsf(s0), fabsf(s1)) : 0.f;
        if (fmaxf(box2, sm * sm * inn) <= H2) rowbits |= 1u << (yy + kTileY * zz);
      }
    }
  }
  // job list: row 0's candidates in lane order, then row 1's, ...
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int njobs = 0;
#pragma unroll
  for (int r = 0; r < kTileRows; r++) {
    const unsigned long long m = __ballot((rowbits >> r) & 1u);
    if ((rowbits >> r) & 1u) S.job[njobs + __popcll(m & below)] = (uint16_t)((r << 8) | lane);
    njobs += __popcll(m);
  }
  S.best[lane] = 0x7F800000u;  // +inf
  __syncthreads();
  // this lane's x and the rows' y and z, relative to the tile origin (exact
  // small multiples of h), as the row test formed them
  const float px = (float)(lattice_coord(g.ox, g.h, i0 + (lane & (kTileX - 1))) - T0.x);
  const float py0 = (float)(lattice_coord(g.oy, g.h, j0) - T0.y);
  const float py1 = (float)(lattice_coord(g.oy, g.h, j0 + 1) - T0.y);
  const float py2 = (float)(lattice_coord(g.oy, g.h, j0 + 2) - T0.y);
  const float py3 = (float)(lattice_coord(g.oy, g.h, j0 + 3) - T0.y);
  const float pz0 = (float)(lattice_coord(g.oz, g.h, k0) - T0.z);
  const float pz1 = (float)(lattice_coord(g.oz, g.h, k0 + 1) - T0.z);
  // Phase 2: exact distances, job q = lane / 8 + 8 t for t = 0, 1, ...
  for (int q = lane / kTileX; q < njobs; q += kTilePts / kTileX) {
    const uint32_t e = S.job[q];
    const int r = (int)(e >> 8), c = (int)(e & 63u);
    const int ry = r & (kTileY - 1);
    const float py = ry < 2 ? (ry == 0 ? py0 : py1) : (ry == 2 ? py2 : py3);
    const float pz = r < kTileY ? pz0 : pz1;
    const float d2 = tri_pre_dist2(px, py, pz, [&](int f) { return S.pre[f][c]; });
    // squared distances are >= +0, so their bits order like the values
    atomicMin(&S.best[r * kTileX + (lane & (kTileX - 1))], __float_as_uint(d2));
  }
  __syncthreads();
  const float best = __uint_as_float(S.best[lane]);
  float *slot = d2buf + first * kTilePts + lane;
  if (single)
    *slot = best;
  else
    atomicMin((unsigned int *)slot, __float_as_uint(best));
  __syncthreads();  // S is reused by the block's next item
}

// One work item per workgroup and round; a grid smaller than the item
// count leaves wave slots and LDS free for kernels on other streams.
__global__ __launch_bounds__(kTilePts) void dist_chunks_kernel(DistJob J) {
  __shared__ DistShared S;
  const int64_t n = chunk_count(J.nchunks, J.nchunks_dev);
  for (int64_t it = xcd_remap(blockIdx.x, gridDim.x); it < n; it += gridDim.x)
    dist_item(J, it, S, true);
}

// Pair windows.  The sorted (tile, triangle) pairs are cut into windows of
// kTilePts consecutive pairs, one per workgroup round.  Phase 1 stages every
// pair of the window at once, one per lane (a tile's list averages a few
// triangles, so a wave per tile would leave most lanes idle here): the
// triangle record relative to the pair's own tile origin and dist_item's row
// test, applied to each half row (4 points): a triangle near a tile is
// usually within h of only a few consecutive points of a row, so half-row
// jobs evaluate fewer points beyond h.  The window holds the runs of one or
// more tiles; phase 2 and the minima then take one run at a time, with
// dist_item's arithmetic, so a point's value is the one dist_item gives
// wherever that is below h^2 (a half row the test passes over has no point
// within h of the triangle; at and above h^2 the output is h either way).
// A run inside the window stores its tile's 64 minima; a run cut by a
// window boundary combines them by atomicMin into the slot tile_bins_finish
// set to +inf.
constexpr int kHalfX = kTileX / 2;
constexpr int kHalfRows = 2 * kTileRows;  // half row hr = 2 row + (x >= kHalfX)
struct PairShared {
  float pre[kTriPreFloats][kTilePts];
  uint16_t job[kHalfRows * kTilePts];  // half row << 8 | lane, run by run
  uint32_t best[kTilePts];             // squared distances as ordered bits
  // per run, from its first lane: the points' coordinates relative to the
  // tile origin (x of the 8 columns, y of the 4 rows, z of the 2 layers)
  float pc[kTilePts][kTileX + kTileY + kTileZ];
};

__device__ inline void dist_window(const DistJob &J, int64_t w, int64_t n, PairShared &S) {
  const GridP &g = J.g;
  const TileDims &td = J.td;
  const int lane = threadIdx.x;
  const int64_t q0 = w * kTilePts;
  const int64_t q = q0 + lane;
  const bool valid = q < n;
  const uint32_t tile = valid ? J.keys[q] : 0xFFFFFFFFu;
  const uint32_t tid = valid ? J.tile_tris[q] : 0u;
  // runs cut by the window's ends (wave-uniform, scalar loads: the keys are
  // read-only here)
  const auto kc = kconst(J.keys);
  const bool first_split = w > 0 && kc[q0 - 1] == kc[q0];
  const int64_t qe = q0 + kTilePts;
  const bool last_split = qe < n && kc[qe] == kc[qe - 1];
  const uint32_t prev = __shfl_up(tile, 1, kTilePts);
  const bool head = valid && (lane == 0 || prev != tile);
  const uint64_t hm = __ballot(head);
  const uint64_t vm = __ballot(valid);
  const float hf = (float)g.h;
  const float H2 = hf * hf * (1.0f + 0x1p-10f);
  uint32_t rowbits = 0;
  if (valid) {
    int32_t tx, ty, tz;
    tile_coords(tile, td, &tx, &ty, &tz);
    const int32_t i0 = tx * kTileX, j0 = ty * kTileY, k0 = tz * kTileZ;
    const Vec3d T0 = lattice_point(g, i0, j0, k0);
    const Tri64 t = J.tri[tid];
    float v[9];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[3 * k + 0] = (float)(t.v[3 * k + 0] - T0.x);
      v[3 * k + 1] = (float)(t.v[3 * k + 1] - T0.y);
      v[3 * k + 2] = (float)(t.v[3 * k + 2] - T0.z);
    }
    float rec[kTriPreFloats];
    tri_pre_make(v, v + 3, v + 6, rec);
#pragma unroll
    for (int f = 0; f < kTriPreFloats; f++) S.pre[f][lane] = rec[f];
    float bmin[3], bmax[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      bmin[a] = fminf(fminf(v[a], v[3 + a]), v[6 + a]);
      bmax[a] = fmaxf(fmaxf(v[a], v[3 + a]), v[6 + a]);
    }
    // the half rows span x in [0, xa] and [xb, x1] (relative to the tile
    // origin); per half, the row test's two bounds on that x range
    const float xa = (float)(lattice_coord(g.ox, g.h, i0 + kHalfX - 1) - T0.x);
    const float xb = (float)(lattice_coord(g.ox, g.h, i0 + kHalfX) - T0.x);
    const float x1 = (float)(lattice_coord(g.ox, g.h, i0 + kTileX - 1) - T0.x);
    const float ex0 = fmaxf(fmaxf(bmin[0] - xa, -bmax[0]), 0.f);
    const float ex1 = fmaxf(fmaxf(bmin[0] - x1, xb - bmax[0]), 0.f);
    const float nx = rec[21], ny = rec[22], nz = rec[23], inn = rec[24];
    const float s0x = (0.f - rec[0]) * nx, sax = (xa - rec[0]) * nx;
    const float sbx = (xb - rec[0]) * nx, s1x = (x1 - rec[0]) * nx;
#pragma unroll
    for (int zz = 0; zz < kTileZ; zz++) {
      const float pz = (float)(lattice_coord(g.oz, g.h, k0 + zz) - T0.z);
      const float ez = fmaxf(fmaxf(bmin[2] - pz, pz - bmax[2]), 0.f);
      const float cz = (pz - rec[2]) * nz;
#pragma unroll
      for (int yy = 0; yy < kTileY; yy++) {
        const float py = (float)(lattice_coord(g.oy, g.h, j0 + yy) - T0.y);
        const float ey = fmaxf(fmaxf(bmin[1] - py, py - bmax[1]), 0.f);
        const float eyz = ey * ey + ez * ez;
        const float c = (py - rec[1]) * ny + cz;
        const int r = yy + kTileY * zz;
        {
          const float s0 = s0x + c, s1 = sax + c;
          const float sm = ((s0 > 0.f) == (s1 > 0.f)) ? fminf(fabsf(s0), fabsf(s1)) : 0.f;
          if (fmaxf(ex0 * ex0 + eyz, sm * sm * inn) <= H2) rowbits |= 1u << (2 * r);
        }
        {
          const float s0 = sbx + c, s1 = s1x + c;
          const float sm = ((s0 > 0.f) == (s1 > 0.f)) ? fminf(fabsf(s0), fabsf(s1)) : 0.f;
          if (fmaxf(ex1 * ex1 + eyz, sm * sm * inn) <= H2) rowbits |= 1u << (2 * r + 1);
        }
      }
    }
    if (head) {  // the run's coordinates, formed as the row test formed them
      float *pc = S.pc[lane];
#pragma unroll
      for (int x = 0; x < kTileX; x++)
        pc[x] = (float)(lattice_coord(g.ox, g.h, i0 + x) - T0.x);
#pragma unroll
      for (int y = 0; y < kTileY; y++)
        pc[kTileX + y] = (float)(lattice_coord(g.oy, g.h, j0 + y) - T0.y);
#pragma unroll
      for (int z = 0; z < kTileZ; z++)
        pc[kTileX + kTileY + z] = (float)(lattice_coord(g.oz, g.h, k0 + z) - T0.z);
    }
  }
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  // the window's job list, run by run, each run's row 0 candidates in lane
  // order, then row 1's, ...: job (lane, r) sits after every job of the runs
  // before (all lanes below its run's first lane la), then after its run's
  // jobs of the rows before r, then after its run's row-r jobs of lower lanes
  const uint64_t upto = lane == 63 ? ~0ull : (~0ull >> (63 - lane));
  const int la_own = 63 - __builtin_clzll(hm & upto | 1ull);  // the lane's run's first lane
  const uint64_t before_run = la_own == 0 ? 0ull : (~0ull >> (64 - la_own));
  const uint64_t hm_above = hm & ~upto;
  const uint64_t run_mask = (hm_above ? (hm_above & (0ull - hm_above)) - 1ull : ~0ull) & ~before_run;
  unsigned long long mr[kHalfRows];
#pragma unroll
  for (int r = 0; r < kHalfRows; r++) mr[r] = __ballot((rowbits >> r) & 1u);
  {
    int pos = 0;
#pragma unroll
    for (int r = 0; r < kHalfRows; r++) pos += __popcll(mr[r] & before_run);
#pragma unroll
    for (int r = 0; r < kHalfRows; r++) {
      if ((rowbits >> r) & 1u)
        S.job[pos + __popcll(mr[r] & run_mask & below)] = (uint16_t)((r << 8) | lane);
      pos += __popcll(mr[r] & run_mask);
    }
  }
  // one run (local tile) at a time: lanes [la, lb) of the window
  uint64_t rest = hm;
  int jstart = 0;
  while (rest != 0ull) {
    const int la = __builtin_ctzll(rest);
    rest &= rest - 1ull;
    const int lb = rest ? __builtin_ctzll(rest) : __popcll(vm);
    const uint64_t m_run = (lb == 64 ? ~0ull : ((1ull << lb) - 1ull)) & ~((1ull << la) - 1ull);
    int njobs = 0;
#pragma unroll
    for (int r = 0; r < kHalfRows; r++) njobs += __popcll(mr[r] & m_run);
    const uint32_t rt = __shfl(tile, la, kTilePts);  // the run's tile
    S.best[lane] = 0x7F800000u;  // +inf
    __syncthreads();
    const float *pc = S.pc[la];
    const float py0 = pc[kTileX], py1 = pc[kTileX + 1], py2 = pc[kTileX + 2],
                py3 = pc[kTileX + 3];
    const float pz0 = pc[kTileX + kTileY], pz1 = pc[kTileX + kTileY + 1];
    // one job per lane: the triangle's record and the half row's y and z
    // terms once, then its 4 points
    for (int jq = jstart + lane; jq < jstart + njobs; jq += kTilePts) {
      const uint32_t e = S.job[jq];
      const int hr = (int)(e >> 8), c = (int)(e & 63u);
      const int r = hr >> 1, half = hr & 1;
      const int ry = r & (kTileY - 1);
      const float py = ry < 2 ? (ry == 0 ? py0 : py1) : (ry == 2 ? py2 : py3);
      const float pz = r < kTileY ? pz0 : pz1;
      auto T = [&](int f) { return S.pre[f][c]; };
      const TriX tx = tri_pre_x(T);
      const TriYZ yz = tri_pre_yz(py, pz, T);
#pragma unroll
      for (int xi = 0; xi < kHalfX; xi++) {
        const float px = pc[half * kHalfX + xi];
        const float d2 = tri_pre_dist2_x(px, tx, yz);
        // squared distances are >= +0, so their bits order like the values
        atomicMin(&S.best[r * kTileX + half * kHalfX + xi], __float_as_uint(d2));
      }
    }
    __syncthreads();
    const uint32_t bv = S.best[lane];
    float *slot = J.d2buf + (uint64_t)rt * kTilePts + lane;
    if ((la == 0 && first_split) || (rest == 0ull && last_split))
      atomicMin((unsigned int *)slot, bv);
    else
      *slot = __uint_as_float(bv);
    jstart += njobs;
    __syncthreads();  // S.best is reused by the next run
  }
  __syncthreads();  // S.pre is reused by the next window
}

// Windows of the sorted pairs, grid-stride.
__global__ __launch_bounds__(kTilePts) void dist_pairs_kernel(DistJob J) {
  __shared__ PairShared S;
  int64_t n = J.npairs;
  if (J.npairs_dev && (int64_t)*kconst(J.npairs_dev) < n) n = (int64_t)*kconst(J.npairs_dev);
  const int64_t nw = (n + kTilePts - 1) / kTilePts;
  for (int64_t w = xcd_remap(blockIdx.x, gridDim.x); w < nw; w += gridDim.x)
    dist_window(J, w, n, S);
}

// Learned-radius variant of the distance item.  Lane p carries a squared
// search radius R2(p) <= h^2 proposed by the network and evaluates only the
// triangles whose lower bound (the larger of the squared distances to the
// triangle's box and to its plane) is within R2 inflated by 2^-10 (which
// covers the rounding of the bound against the exact evaluation).  A lane
// that ends with best >= R2 < h^2 has no triangle within its radius and is
// settled afterwards by an exact query; every other lane's best is the value
// the full loop gives.  Counters per block slot (b & 255): [8..] the wave's
// exact evaluations (its longest candidate list), [264..] candidates within
// the learned radius, [520..] candidates within h, [776..] the wave's
// longest list within h (the exact evaluations of the plain loop).
__global__ __launch_bounds__(kTilePts) void dist_chunks_learned_kernel(
    GridP g, TileDims td, const Tri64 *__restrict__ tri,
    const uint64_t *__restrict__ tile_off, const uint64_t *__restrict__ tile_end,
    const uint32_t *__restrict__ tile_tris,
    const uint64_t *__restrict__ chunk_off, const uint32_t *__restrict__ chunk_tile,
    int64_t nchunks, const uint64_t *__restrict__ nchunks_dev, int clen,
    const float *__restrict__ r2buf, float *__restrict__ d2buf,
    unsigned long long *__restrict__ stats) {
  __shared__ float s_pre[kTriPreFloats][kTilePts];
  __shared__ float s_box[6][kTilePts];
  const int64_t item = xcd_remap(blockIdx.x, gridDim.x);
  if (item >= chunk_count(nchunks, nchunks_dev)) return;
  const int lane = threadIdx.x;
  const int64_t tile = chunk_tile[item];
  const uint64_t first = chunk_off[tile];
  const int64_t q = item - (int64_t)first;
  int32_t tx, ty, tz;
  tile_coords((uint32_t)tile, td, &tx, &ty, &tz);
  const int32_t i0 = tx * kTileX, j0 = ty * kTileY, k0 = tz * kTileZ;
  const Vec3d T0 = lattice_point(g, i0, j0, k0);
  const uint64_t b = tile_off[tile] + (uint64_t)q * clen;
  const uint64_t e = min(b + (uint64_t)clen, tile_end[tile]);
  const int n = (int)(e - b);
  if (lane < n) {
    const Tri64 t = tri[tile_tris[b + lane]];
    float v[9];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[3 * k + 0] = (float)(t.v[3 * k + 0] - T0.x);
      v[3 * k + 1] = (float)(t.v[3 * k + 1] - T0.y);
      v[3 * k + 2] = (float)(t.v[3 * k + 2] - T0.z);
    }
    float rec[kTriPreFloats];
    tri_pre_make(v, v + 3, v + 6, rec);
#pragma unroll
    for (int f = 0; f < kTriPreFloats; f++) s_pre[f][lane] = rec[f];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      s_box[a][lane] = fminf(fminf(v[a], v[3 + a]), v[6 + a]);
      s_box[3 + a][lane] = fmaxf(fmaxf(v[a], v[3 + a]), v[6 + a]);
    }
  }
  __syncthreads();
  const int32_t i = i0 + (lane & (kTileX - 1));
  const int32_t j = j0 + ((lane / kTileX) & (kTileY - 1));
  const int32_t k = k0 + lane / (kTileX * kTileY);
  const Vec3d p = lattice_point(g, i, j, k);
  const float px = (float)(p.x - T0.x), py = (float)(p.y - T0.y), pz = (float)(p.z - T0.z);
  const float R2 = r2buf[first * kTilePts + lane] * (1.0f + 0x1p-10f);
  const float hf = (float)g.h;
  const float H2 = hf * hf * (1.0f + 0x1p-10f);
  // phase 1: candidates within this lane's radius (and, for the counters,
  // within h); phase 2: exact distances of the candidates only
  unsigned long long cand = 0ull;
  int lanes_h = 0;
  for (int c = 0; c < n; c++) {
    const float ex = fmaxf(fmaxf(s_box[0][c] - px, px - s_box[3][c]), 0.f);
    const float ey = fmaxf(fmaxf(s_box[1][c] - py, py - s_box[4][c]), 0.f);
    const float ez = fmaxf(fmaxf(s_box[2][c] - pz, pz - s_box[5][c]), 0.f);
    const float box2 = ex * ex + ey * ey + ez * ez;
    const float pn = (px - s_pre[0][c]) * s_pre[21][c] + (py - s_pre[1][c]) * s_pre[22][c] +
                     (pz - s_pre[2][c]) * s_pre[23][c];
    const float lb2 = fmaxf(box2, pn * pn * s_pre[24][c]);
    if (lb2 <= R2) cand |= 1ull << c;
    lanes_h += lb2 <= H2 ? 1 : 0;
  }
  const int lanes_r = __popcll(cand);
  float best = __builtin_huge_valf();
  int visited = 0;
  while (cand) {
    const int c = __builtin_ctzll(cand);
    cand &= cand - 1ull;
    const float d2 = tri_pre_dist2(px, py, pz, [&](int f) { return s_pre[f][c]; });
    best = fminf(best, d2);
    visited++;
  }
  // wave totals for the counters
  int tot_r = lanes_r, tot_h = lanes_h, vmax = visited, hmax = lanes_h;
  for (int off = 32; off > 0; off >>= 1) {
    tot_r += __shfl_xor(tot_r, off);
    tot_h += __shfl_xor(tot_h, off);
    vmax = max(vmax, __shfl_xor(vmax, off));
    hmax = max(hmax, __shfl_xor(hmax, off));
  }
  float *slot = d2buf + first * kTilePts + lane;
  if (chunk_off[tile + 1] - first == 1)
    *slot = best;
  else
    atomicMin((unsigned int *)slot, __float_as_uint(best));
  if (lane == 0 && stats) {
    atomicAdd(stats + 8 + (blockIdx.x & 255), (unsigned long long)vmax);
    atomicAdd(stats + 264 + (blockIdx.x & 255), (unsigned long long)tot_r);
    atomicAdd(stats + 520 + (blockIdx.x & 255), (unsigned long long)tot_h);
    atomicAdd(stats + 776 + (blockIdx.x & 255), (unsigned long long)hmax);
  }
}

// Appends every lane whose learned radius R2 < h^2 enclosed no triangle to
// the fallback list (packed first_item * 64 + lane); counts them in stats[0].
__global__ __launch_bounds__(256) void learned_check_kernel(
    const uint64_t *__restrict__ chunk_off, const uint32_t *__restrict__ chunk_tile,
    int64_t nchunks, const uint64_t *__restrict__ nchunks_dev,
    const float *__restrict__ r2buf, const float *__restrict__ d2buf,
    float h2, uint64_t *__restrict__ list, unsigned long long *__restrict__ stats) {
  const int64_t item = blockIdx.x * 4LL + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (item >= chunk_count(nchunks, nchunks_dev)) return;
  const uint64_t first = chunk_off[chunk_tile[item]];
  if ((uint64_t)item != first) return;
  const float R2 = r2buf[first * kTilePts + lane];
  const float d2 = d2buf[first * kTilePts + lane];
  const bool fb = d2 >= R2 && R2 < h2;
  const unsigned long long m = __ballot(fb);
  if (!m) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(stats + 0, (unsigned long long)__popcll(m));
  base = __shfl(base, 0);
  if (fb) {
    const unsigned long long below = m & ((1ull << lane) - 1ull);
    list[base + __popcll(below)] = first * kTilePts + (uint64_t)lane;
  }
}

// Exact distance of each fallback node over the LBVH with the tile kernel's
// arithmetic (records and point relative to the tile origin), so the value
// equals the full tile loop's: the nearest triangle within h lies in both
// candidate sets and is evaluated the same way.  Overwrites the node's slot.
__global__ __launch_bounds__(256) void learned_fallback_kernel(
    GridP g, TileDims td, const Tri64 *__restrict__ tri, Bvh bvh,
    const uint32_t *__restrict__ chunk_tile, const uint64_t *__restrict__ list,
    const unsigned long long *__restrict__ count, float slack, float *__restrict__ d2buf,
    uint32_t *__restrict__ err) {
  const int64_t n = (int64_t)*count;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t code = list[i];
    const uint64_t first = code / kTilePts;
    const int lane = (int)(code % kTilePts);
    const int64_t tile = chunk_tile[first];
    int32_t tx, ty, tz;
    tile_coords((uint32_t)tile, td, &tx, &ty, &tz);
    const int32_t i0 = tx * kTileX, j0 = ty * kTileY, k0 = tz * kTileZ;
    const Vec3d T0 = lattice_point(g, i0, j0, k0);
    const Vec3d p = lattice_point(g, i0 + (lane & (kTileX - 1)), j0 + ((lane / kTileX) & (kTileY - 1)),
                                  k0 + lane / (kTileX * kTileY));
    const float px = (float)(p.x - T0.x), py = (float)(p.y - T0.y), pz = (float)(p.z - T0.z);
    const float pa[3] = {(float)p.x, (float)p.y, (float)p.z};
    auto leaf = [&](uint32_t t) {
      const Tri64 q = tri[t];
      float v[9];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        v[3 * k + 0] = (float)(q.v[3 * k + 0] - T0.x);
        v[3 * k + 1] = (float)(q.v[3 * k + 1] - T0.y);
        v[3 * k + 2] = (float)(q.v[3 * k + 2] - T0.z);
      }
      float rec[kTriPreFloats];
      tri_pre_make(v, v + 3, v + 6, rec);
      return tri_pre_dist2(px, py, pz, [&](int f) { return rec[f]; });
    };
    bool ovf = false;
    const float best = nearest_bvh_with(bvh, pa, (float)g.h * (1.0f + 0x1p-10f), slack, leaf, &ovf);
    if (ovf) atomicOr(err, kErrStackOverflow);
    d2buf[code] = best;
  }
}

// Sign / combine kernel.  One wavefront covers 8 x 8 x 4 lattice points:
// lane (i, j) takes the 4 points k0..k0+3 of its z-line, so the z-line's hit
// summary is read once and classifies all 4 points; the x- and y-lines of
// the 4 points are independent loads.  Per point: distance from the chunk
// results (h when its distance tile lists no triangle), the six axis-ray
// votes (certified fast path, else the reference's arithmetic), combine,
// outputs; points whose votes are not unanimous go to the pending list.

// LDS staging of one sign brick: its 32 x-line and 32 y-line summaries and
// the distance slots of its 4 distance tiles, fetched by one round of
// independent loads so each lane's per-point reads come from LDS.
// LDS hand-off between the lanes of a one-wave workgroup (64 lanes): a
// wave's LDS operations complete in order, so a wavefront-scope fence, which
// keeps the compiler from moving memory operations across it, is all that is
// needed.  A workgroup barrier's release fence would also wait for every
// global store still in flight.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct FusedShared {
  LineInline lx[kSignZ][8];  // x-line (k0 + q, 8 by + w)
  LineInline ly[kSignZ][8];  // y-line (k0 + q, 8 bx + w)
  // per x- and y-line: 1 | pos << 1 | neg << 2 when both rays are the same
  // for all 8 points of the line in the brick (inline_uniform), else 0
  uint8_t ux[kSignZ][8], uy[kSignZ][8];
};

// Output sets of the sign kernel (bits); an instantiation keeps only the
// pointers of its set in registers.
enum : int { kOutSdf = 1, kOutDist = 2, kOutSigns = 4, kOutConf = 8, kOutAll = 15 };

// One sign brick (8 x 8 x 4 points) of job J for the 64-lane block (one
// wave: its LDS hand-offs are wave_lds_sync); m is the brick's position in
// the brick launch order; C is the LDS copy of the job's cold fields;
// outputs outside kOut are not written.  The brick's coordinates and
// distance slots come from J.plan[m] (BrickPlan, scalar loads).  Uniform
// per block.
template <int kOut>
__device__ inline void fused_brick(const FusedJob &J, int64_t m, FusedShared &S,
                                   const FusedCold &C) {
  const GridP &g = J.g;
  const float *__restrict__ d2buf = J.d2buf;
  const LineInline *__restrict__ inl = J.inl;
  const int64_t base1 = J.base1, base2 = J.base2;
  // stages fixed at compile time for the specialised output sets (the
  // launcher picks them only when the job matches): bit 0 distances (d2buf
  // set), bit 1 signs; 0 = both read from the job
  constexpr int kStages = kOut == kOutSdf                  ? 3
                          : kOut == kOutDist               ? 1
                          : kOut == (kOutSigns | kOutConf) ? 2
                                                           : 0;
  const bool has_d2 = kStages ? (kStages & 1) != 0 : d2buf != nullptr;
  const int do_sign = kStages ? (kStages >> 1) : J.do_sign;
  const FusedOut out = J.out;
  BrickPlan P;
  {
    // adjacent words through the constant view: scalar loads
    const KConst<uint32_t> pw = kconst(reinterpret_cast<const uint32_t *>(J.plan + m));
#pragma unroll
    for (int t = 0; t < 4; t++) P.slot[t] = pw[t];
    P.xy = pw[4];
    P.z = pw[5];
    P.flags = pw[6];
  }
  if (!(P.flags & 2u)) return;  // padding position (uniform)
  const int32_t bx = (int32_t)(P.xy & 0xFFFFu), by = (int32_t)(P.xy >> 16),
                bz = (int32_t)P.z;
  const int lane = threadIdx.x;
  const int32_t i = bx * 8 + (lane & 7);
  const int32_t j = by * 8 + (lane >> 3);
  const int32_t k0 = bz * kSignZ;
  const bool in_grid = i < g.nx && j < g.ny;
  const int nk = min(kSignZ, g.nz - k0);

  // ---- issue the brick's loads: the segment bytes of the lane's own z-line
  // and of one x- or y-line together with the four distance tiles' offsets,
  // then the lane's four squared distances.  Vector loads complete in issue
  // order, so the bytes go first and the uniformity test below runs while
  // the distances are in flight.
  // Lanes 0-31 stage the x-lines (k0 + sq, 8 by + sw), lanes 32-63 the
  // y-lines (k0 + sq, 8 bx + sw): one code path, operands by select.
  const int sq = (lane >> 3) & 3, sw = lane & 7;
  const bool xl = lane < 32;
  const int32_t la = (xl ? by : bx) * 8 + sw;  // the line's y (x-lines) or x (y-lines)
  const int32_t across = xl ? g.ny : g.nx;
  const bool valid = k0 + sq < g.nz && la < across;
  // the staged line (0 when it serves no point) and the lane's own z-line
  const int64_t Lxy = valid ? (xl ? 0 : base1) + (int64_t)(k0 + sq) * across + la : 0;
  const int64_t Lz = in_grid ? base2 + (int64_t)j * g.nx + i : 0;
  // The brick covers 1 x 2 x 2 distance tiles (8 x 4 x 2 points; tile
  // t = 2 tz local + ty local), whose slots the plan holds.
  static_assert(kTileX == 8 && 2 * kTileY == 8 && 2 * kTileZ == kSignZ,
                "a sign brick covers 1 x 2 x 2 distance tiles");
  // segment bytes (SegLayout): segment bx of the staged x-line or by of the
  // staged y-line, segment bz of the lane's z-line.  A line that serves no
  // point reads the sentinel word after the last one (byte 1 at every
  // position: uniform, both parities even), so the loads and the byte
  // extraction are unconditional and nothing waits on them before the
  // distance loads are issued.
  uint32_t bxy = 1u, bzl = 1u;
  if (do_sign) {
    const int32_t sxy = xl ? bx : by;
    const int64_t nl = (int64_t)across * g.nz;
    // a brick whose lines all carry byte 1, or all byte 7, at the brick
    // (plan flag bits 2, 3: brick_same_flags_kernel) reads the matching
    // sentinel word instead (one cached address), so its loads return at
    // once and it takes the shared-byte path below with the same bytes
    const uint32_t same = P.flags & 12u;
    const int64_t sent = J.segtotal + (same == 8u ? 1 : 0);
    const int64_t oxy = valid && !same ? (xl ? 0 : J.segoff1) + (int64_t)(sxy >> 2) * nl +
                                             (int64_t)(k0 + sq) * across + la
                                       : sent;
    const int64_t oz = in_grid && !same
                           ? J.segoff2 + (int64_t)(bz >> 2) * ((int64_t)g.nx * g.ny) +
                                 (int64_t)j * g.nx + i
                           : sent;
    const uint32_t wxy = J.segw[oxy], wz = J.segw[oz];
    bxy = (wxy >> (8 * (sxy & 3))) & 0xFFu;
    bzl = (wz >> (8 * (bz & 3))) & 0xFFu;
  }
  // The lane's squared distances of points (i, j, k0 + q), straight from the
  // four tiles: a tile's slot (its first work item) is wave-uniform; for each
  // q the wave reads two runs of 32 consecutive floats.
  float best[kSignZ];
#pragma unroll
  for (int q = 0; q < kSignZ; q++) best[q] = __builtin_huge_valf();
  bool work = false;  // wave-uniform: some tile of the brick has distance work
  if (has_d2) {
    // a tile without distance work reads the spare tile of +inf after the
    // last work item: the loads are unconditional and need no select
    uint32_t slot[4];
#pragma unroll
    for (int t = 0; t < 4; t++) slot[t] = P.slot[t];
    work = (P.flags & 1u) != 0u;
    // a brick no triangle is near (most of the grid) keeps +inf: distance h
    if (work) {
      const int tyl = lane >> 5;  // the lane's y tile in the brick
#pragma unroll
      for (int q = 0; q < kSignZ; q++) {
        const uint32_t sl = tyl ? slot[(q >> 1) * 2 + 1] : slot[(q >> 1) * 2];
        const int tl = (lane & 7) + kTileX * (((lane >> 3) & 3) + kTileY * (q & 1));
        best[q] = d2buf[(uint64_t)sl * kTilePts + tl];
      }
    }
  }

  // ---- uniformity of the lane's z-line over its nk points and of its staged
  // line over the brick's 8 points on it, from the tails; the full x- and
  // y-line summaries go to LDS only for a brick with a line that is not
  // uniform
  bool zuni = false;  // the lane's z-line is the same for its nk points
  int zp = 0, zn = 0;
  // every line this lane staged or owns is shared by the brick's points
  bool lines_shared = true;
  if (do_sign) {
    // the byte is 1 | pos << 1 | neg << 2 for a uniform segment, else 0
    const bool uni = (bxy & 1u) != 0u;
    // (a line that serves no point holds the sentinel's 1, read by no point)
    *(xl ? &S.ux[sq][sw] : &S.uy[sq][sw]) = (uint8_t)bxy;
    zuni = (bzl & 1u) != 0u;
    zp = (int)((bzl >> 1) & 1u);
    zn = (int)((bzl >> 2) & 1u);
    // a staged line outside the grid (or a layer beyond nz) serves no point
    lines_shared = (!valid | uni) & (!in_grid | zuni);
  }
  // the distances are read here, on every path: a load no path reads would
  // still be in flight at the next brick, whose first write to its register
  // then waits for every load and store before it
#pragma unroll
  for (int q = 0; q < kSignZ; q++) asm volatile("" : "+v"(best[q]));
  // wave-uniform shortcut: every line of the brick is shared by its points
  // (no crossing near the brick), so each point only adds six stored parities
  const bool brick_uni = __all(lines_shared);
  // every line of the brick holds the same uniform byte, with equal parities
  // for its two rays (1: all even, 7: all odd): every point's six votes agree
  // (outside, or inside), with no staged parities to add up.  Lane 0 is
  // always inside the grid.
  const uint32_t b0 = __builtin_amdgcn_readfirstlane(bzl);
  const bool same = do_sign && brick_uni && (b0 == 1u || b0 == 7u) &&
                    __all((!in_grid | (bzl == b0)) & (!valid | (bxy == b0)));
  if (do_sign && !brick_uni) *(xl ? &S.lx[sq][sw] : &S.ly[sq][sw]) = inl[Lxy];
  wave_lds_sync();

  // the point's distance: min(sqrt(best), h), h when no triangle was near.
  // best is an FP32 squared distance: its correctly rounded FP32 root adds at
  // most half an FP32 ulp (below 3e-10 for distances up to h = 0.005) and
  // costs a fraction of the FP64 root's correction sequence
  auto point_dist = [&](int q) -> double {
    double dist = g.h;
    if (best[q] < __builtin_huge_valf()) {
      const double d = (double)__fsqrt_rn(best[q]);
      dist = d < g.h ? d : g.h;
    }
    return dist;
  };
  // pending entry when a decision sat inside its rounding margin (amb: the
  // point restarts the vote from ray 1 with the reference's arithmetic, list
  // 0) or the six votes disagree (list 1 for a 5:1 split, list 2 for 4:2 or
  // 3:3); the vote stage writes such a point's outputs later, after the
  // provisional ones written here
  auto finish = [&](int64_t idx, double dist, int inside, bool amb) {
    if (do_sign && (amb | ((inside != 0) & (inside != 6)))) {
      Pending pr;
      pr.idx = idx;
      pr.dist = dist;
      pr.inside = amb ? 0 : inside;
      pr.used = amb ? 0 : 6;
      pr.bits = 0ull;
      const int li = amb ? 0 : ((inside == 1) | (inside == 5)) ? 1 : 2;
      const uint32_t slot = atomicAdd(C.pend.n[li], 1u);
      if (slot < C.pend.cap[li])
        C.pend.p[li][slot] = pr;
      else
        atomicOr(C.err, kErrPendingOverflow);
    }
    const double sgn = inside == 6 ? 1.0 : -1.0;
    if ((kOut & kOutDist) && out.dist) out.dist[idx] = dist;
    if ((kOut & kOutSdf) && out.sdf) out.sdf[idx] = dist * sgn;
    if ((kOut & kOutSigns) && out.signs) out.signs[idx] = sgn;
    if ((kOut & kOutConf) && out.conf) out.conf[idx] = 1.0;
  };
  if (in_grid && same) {
    const int64_t plane = (int64_t)g.nx * g.ny;
    const int64_t idx0 = ((int64_t)k0 * g.ny + j) * g.nx + i;
    const int inside = b0 == 7u ? 6 : 0;
#pragma unroll
    for (int q = 0; q < kSignZ; q++) {
      if (q >= nk) break;  // uniform per wave
      finish(idx0 + q * plane, work ? point_dist(q) : g.h, inside, false);
    }
  } else if (in_grid && brick_uni) {
    // every line of the brick is shared by its points: the six parities are
    // the stored ones
    const int64_t plane = (int64_t)g.nx * g.ny;
    const int64_t idx0 = ((int64_t)k0 * g.ny + j) * g.nx + i;
#pragma unroll
    for (int q = 0; q < kSignZ; q++) {
      if (q >= nk) break;  // uniform per wave
      int inside = 0;
      if (do_sign) {
        const uint32_t ux = S.ux[q][lane >> 3], uy = S.uy[q][lane & 7];
        inside = zp + zn + (int)(((ux >> 1) & 1u) + ((ux >> 2) & 1u) + ((uy >> 1) & 1u) +
                                 ((uy >> 2) & 1u));
      }
      finish(idx0 + q * plane, point_dist(q), inside, false);
    }
  } else if (in_grid) {
    const double x = lattice_coord(g.ox, g.h, i), y = lattice_coord(g.oy, g.h, j);
    double zq[kSignZ];
#pragma unroll
    for (int q = 0; q < kSignZ; q++) zq[q] = lattice_coord(g.oz, g.h, k0 + q);
    int zpos[kSignZ], zneg[kSignZ];
    bool zok[kSignZ];
#pragma unroll
    for (int q = 0; q < kSignZ; q++) {
      zpos[q] = zp;
      zneg[q] = zn;
      zok[q] = true;
    }
    if (do_sign && !zuni) {
      // z-line (i, j): one summary classifies all 4 points, else the line's
      // crossing list
      if (!inline_parity<kSignZ>(inl[Lz], zq, zpos, zneg, zok)) {
        const uint64_t hb = C.line_off[Lz], he = C.line_off[Lz + 1];
        const uint32_t mz = C.meta[Lz];
#pragma unroll
        for (int q = 0; q < kSignZ; q++)
          zok[q] = q < nk ? axis_pair_parity_fast(C.fast, hb, he, mz, zq[q], &zpos[q], &zneg[q])
                          : true;
      }
    }
#pragma unroll
    for (int q = 0; q < kSignZ; q++) {
      if (q >= nk) break;  // uniform per wave
      const int32_t k = k0 + q;
      const int64_t idx = ((int64_t)k * g.ny + j) * g.nx + i;
      int inside = 0;
      bool amb = false;
      if (do_sign) {
        const uint32_t ux = S.ux[q][lane >> 3], uy = S.uy[q][lane & 7];
        inside = zpos[q] + zneg[q];
        int a, bpar;
        bool okq;
        amb = !zok[q];
        if (ux & 1u) {
          a = (int)((ux >> 1) & 1u);
          bpar = (int)((ux >> 2) & 1u);
        } else {
          const LineInline &rx = S.lx[q][lane >> 3];
          if (!inline_parity<1>(rx, &x, &a, &bpar, &okq) || !okq) {
            const int64_t Lx = (int64_t)k * g.ny + j;
            amb |= !axis_pair_parity_fast(C.fast, C.line_off[Lx], C.line_off[Lx + 1],
                                          C.meta[Lx], x, &a, &bpar);
          }
        }
        inside += a + bpar;
        if (uy & 1u) {
          a = (int)((uy >> 1) & 1u);
          bpar = (int)((uy >> 2) & 1u);
        } else {
          const LineInline &ry = S.ly[q][lane & 7];
          if (!inline_parity<1>(ry, &y, &a, &bpar, &okq) || !okq) {
            const int64_t Ly = base1 + (int64_t)k * g.nx + i;
            amb |= !axis_pair_parity_fast(C.fast, C.line_off[Ly], C.line_off[Ly + 1],
                                          C.meta[Ly], y, &a, &bpar);
          }
        }
        inside += a + bpar;
      }
      finish(idx, point_dist(q), inside, amb);
    }
  }
  wave_lds_sync();  // S is reused by the block's next brick
}

// Eight waves per SIMD (64 registers, no spills to scratch).  A brick waits
// on two load round trips (its plan and segment bytes, then its distances),
// so occupancy pays: on the star surface (82.6 M points) 0.294-0.297 ms at
// eight waves against 0.300-0.303 ms at seven (earlier builds: 0.358 ms at
// six against 0.370 ms at five).
template <int kOut>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8)))
void fused_lattice_kernel(FusedJob J) {
  __shared__ FusedShared S;
  __shared__ FusedCold C;
  if (threadIdx.x == 0) C = J.cold;
  wave_lds_sync();
  // 32-bit brick loop (the launcher keeps nbricks below 2^31): the loop test
  // stays on the scalar unit and takes no vector register that a load in
  // flight may still target
  const uint32_t nb = (uint32_t)J.nbricks;
  for (uint32_t m = (uint32_t)xcd_remap(blockIdx.x, gridDim.x); m < nb; m += gridDim.x)
    fused_brick<kOut>(J, (int64_t)m, S, C);
}

// The bricks launch_brick_same marked (every line byte 1 or every line byte
// 7), one per workgroup; the others return after their plan's flags.  A
// marked brick without distance work (plan flag bit 0 clear: no triangle
// near any of its points) is h from the surface at every point, outside
// (byte 1) or inside (byte 7): it only stores, with no load after its plan.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8)))
void fused_same_kernel(FusedJob J) {
  __shared__ FusedShared S;
  __shared__ FusedCold C;
  if (threadIdx.x == 0) C = J.cold;
  wave_lds_sync();
  const GridP &g = J.g;
  const int lane = threadIdx.x;
  const int64_t plane = (int64_t)g.nx * g.ny;
  const uint32_t nb = (uint32_t)J.nbricks;
  for (uint32_t m = (uint32_t)xcd_remap(blockIdx.x, gridDim.x); m < nb; m += gridDim.x) {
    // BrickPlan words 4-6: xy, z, flags (scalar loads)
    const KConst<uint32_t> pw = kconst(reinterpret_cast<const uint32_t *>(J.plan + m));
    const uint32_t flags = pw[6];
    if (!(flags & 12u)) continue;
    if (flags & 1u) {
      fused_brick<kOutSdf>(J, (int64_t)m, S, C);
      continue;
    }
    const uint32_t xy = pw[4], bz = pw[5];
    const int32_t i = (int32_t)(xy & 0xFFFFu) * 8 + (lane & 7);
    const int32_t j = (int32_t)(xy >> 16) * 8 + (lane >> 3);
    const int32_t k0 = (int32_t)bz * kSignZ;
    const int nk = min(kSignZ, g.nz - k0);
    const double v = (flags & 8u) ? g.h : -g.h;
    if (i < g.nx && j < g.ny) {
      const int64_t idx0 = ((int64_t)k0 * g.ny + j) * g.nx + i;
#pragma unroll
      for (int q = 0; q < kSignZ; q++) {
        if (q >= nk) break;  // uniform per wave
        J.out.sdf[idx0 + q * plane] = v;
      }
    }
  }
}

// The bricks in list[0, *n) (launch positions).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8)))
void fused_list_kernel(FusedJob J, const uint32_t *__restrict__ list, const uint32_t *n) {
  __shared__ FusedShared S;
  __shared__ FusedCold C;
  if (threadIdx.x == 0) C = J.cold;
  wave_lds_sync();
  const uint32_t cnt = *kconst(n);
  for (uint32_t t = blockIdx.x; t < cnt; t += gridDim.x)
    fused_brick<kOutSdf>(J, (int64_t)kconst(list)[t], S, C);
}

// The sign kernel's plan, one thread per launch position: the brick decode
// and the distance-slot lookup the kernel would otherwise repeat per brick
// on its scalar unit, with the same rules (a tile outside the grid or
// without work items, or whose first item lies beyond the capacity, takes
// the spare slot chunk_cap).
__global__ __launch_bounds__(256) void brick_plan_kernel(FusedJob J, BrickPlan *__restrict__ plan) {
  const TileDims &td = J.td;
  const int64_t chunk_cap = J.chunk_cap;
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < J.nbricks;
       m += (int64_t)gridDim.x * blockDim.x) {
    BrickPlan P;
    P.slot[0] = P.slot[1] = P.slot[2] = P.slot[3] = (uint32_t)chunk_cap;
    P.xy = P.z = P.flags = P.pad_ = 0u;
    int32_t bx, by, bz;
    if (brick_tile(m, J.bd, &bx, &by, &bz)) {
      P.xy = (uint32_t)bx | (uint32_t)by << 16;
      P.z = (uint32_t)bz;
      P.flags = 2u;
      if (J.d2buf && (J.chunk_off || J.tile_end)) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int32_t tyy = 2 * by + (t & 1), tzz = 2 * bz + (t >> 1);
          if (tyy < td.ty && tzz < td.tz) {
            const int64_t tile = ((int64_t)tzz * td.ty + tyy) * td.tx + bx;
            if (J.tile_end) {  // distances by tile: a tile with a run has its slot
              if (J.tile_end[tile] != 0u) {
                P.slot[t] = (uint32_t)tile;
                P.flags |= 1u;
              }
            } else {
              const uint64_t c0 = J.chunk_off[tile], c1 = J.chunk_off[tile + 1];
              if (c1 > c0 && (int64_t)c0 < chunk_cap) {
                P.slot[t] = (uint32_t)c0;
                P.flags |= 1u;
              }
            }
          }
        }
      }
    }
    plan[m] = P;
  }
}

// Parity of one ray over all triangles, one workgroup (reference:
// check_ray_inside_bvh with a brute-force contact list).
__device__ int block_ray_parity(const double *o, const double *d,
                                const Tri64 *__restrict__ tri, int64_t nel,
                                double *s_t, int *s_n, int *s_cnt,
                                uint32_t *err) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    *s_n = 0;
    *s_cnt = 0;
  }
  __syncthreads();
  for (int64_t el = tid; el < nel; el += blockDim.x) {
    const Tri64 t = tri[el];
    double tv;
    if (mt_ref(o, d, t.v, t.v + 3, t.v + 6, &tv)) {
      int slot = atomicAdd(s_n, 1);
      if (slot < kHitCap) s_t[slot] = tv;
    }
  }
  __syncthreads();
  int n = *s_n;
  if (n > kHitCap) {
    if (tid == 0) atomicOr(err, kErrHitOverflow);
    n = kHitCap;
  }
  for (int a = tid; a < n; a += blockDim.x) {
    double x = s_t[a];
    bool dup = false, have = false;
    double pred = 0.0;
    for (int b = 0; b < n; b++) {
      double y = s_t[b];
      if (b != a && y == x && b < a) dup = true;
      if (y < x && (!have || y > pred)) {
        pred = y;
        have = true;
      }
    }
    if (!dup && (!have || x - pred > kRayEps)) atomicAdd(s_cnt, 1);
  }
  __syncthreads();
  int c = *s_cnt;
  __syncthreads();
  return c & 1;
}

// Solid angle sum (generalized_winding_number, :295-328), block reduction.
__device__ inline void normalize_ref(const double *a, double *o) {
#pragma clang fp contract(off)
  double s = a[0] * a[0];
  s = s + a[1] * a[1];
  s = s + a[2] * a[2];
  double n = sqrt(s);
  if (n == 0.0) {
    o[0] = o[1] = o[2] = __builtin_nan("");
    return;
  }
  double inv = 1.0 / n;
  o[0] = a[0] * inv;
  o[1] = a[1] * inv;
  o[2] = a[2] * inv;
}

__device__ inline void item_point(const GridP &g, const double *pts, int64_t idx,
                                  double *o) {
  if (pts) {
    o[0] = pts[3 * idx + 0];
    o[1] = pts[3 * idx + 1];
    o[2] = pts[3 * idx + 2];
  } else {
    int64_t nxy = (int64_t)g.nx * g.ny;
    int64_t kk = idx / nxy, rem = idx % nxy;
    Vec3d q = lattice_point(g, rem % g.nx, rem / g.nx, kk);
    o[0] = q.x;
    o[1] = q.y;
    o[2] = q.z;
  }
}

// Live count of list k.
__device__ inline uint32_t list_count(const PendLists &L, int k) {
  uint32_t n = L.cap[k];
  if (L.n_dev[k] && *L.n_dev[k] < n) n = *L.n_dev[k];
  return n;
}

// The list of a slot over the lists (k) and its index in that list.
__device__ inline int slot_list(const PendLists &L, uint64_t slot, uint64_t *index) {
  int k = 0;
  while (k + 1 < L.count && slot >= L.cap[k]) {
    slot -= L.cap[k];
    k++;
  }
  *index = slot;
  return k;
}

// The item in a slot over the lists.
__device__ inline Pending &slot_item(const PendLists &L, uint64_t slot) {
  uint64_t i;
  const int k = slot_list(L, slot, &i);
  return L.items[k][i];
}

// Rays [r0, r1) of list k (none when r1 <= r0).
__device__ inline int list_rays(const PendLists &L, int k) {
  return L.r1[k] > L.r0[k] ? L.r1[k] - L.r0[k] : 0;
}

// Wavefront-cooperative ray parity: one wavefront per (item, ray) over the
// four-wide view of the ray hierarchy (leaves are runs of kRayLeaf sorted
// triangles).  The 64 lanes pop up to 16 wide nodes at a time from an LDS
// stack, four lanes per node, and test one child box each (FP64 slab test
// on boxes rounded outward); internal children are pushed and reached runs
// queued with ballot prefix sums; in the same
// iteration up to 64 / kRayLeaf queued runs are tested, kRayLeaf lanes per
// run, with the reference's FP64 Moller-Trumbore.  Hits go to an LDS list; the unique count follows the
// reference rule (sorted, consecutive gap > eps) computed as: a value counts
// when it is the smallest or its strict predecessor lies more than eps
// below, equal values counting once.  Stack or hit-list overflow sends the
// ray to the brute-force fallback.
// Hits one wave-traversed ray can hold (LDS); a ray with more is redone by
// the brute-force kernel.  Small, so that many rays are in flight per CU.
constexpr int kWaveHits = 256;

constexpr int kRayWaves = 8192;  // persistent grid of the ray stage (waves)
constexpr int kRunQueue = 512;   // queued runs of one ray (LDS); more: brute force

template <int kWaveStack>
__global__ __launch_bounds__(64) void ray_wave_kernel(
    GridP g, const double *__restrict__ pts, const Tri64 *__restrict__ tri, Bvh bvh,
    RayParams rp, PendLists lists, uint64_t *__restrict__ fb, uint32_t *nfb, uint32_t fb_cap,
    uint32_t *err) {
  __shared__ int32_t stack[kWaveStack];
  __shared__ int32_t runq[kRunQueue];  // runs reached, waiting for their triangle tests
  __shared__ double hit_t[kWaveHits];
  __shared__ int s_top, s_nq, s_nh, s_ovf, s_cnt;
  const int lane = threadIdx.x;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  // list k contributes n_k * nr_k (item, ray) pairs, after the lists before it
  int64_t wend[kPendLists];
  int64_t total = 0;
  for (int k = 0; k < kPendLists; k++) {
    if (k < lists.count) total += (int64_t)list_count(lists, k) * list_rays(lists, k);
    wend[k] = total;
  }
  for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
    int k = 0;
    while (w >= wend[k]) k++;
    const int64_t wk = w - (k ? wend[k - 1] : 0);
    const int nr = list_rays(lists, k);
    const uint64_t slot = lists.base(k) + (uint64_t)(wk / nr);
    const int r = lists.r0[k] + (int)(wk % nr);
    Pending &item = slot_item(lists, slot);
    if (r < item.used) continue;  // counted in the item's state (uniform per wave)
    double o[3];
    item_point(g, pts, item.idx, o);
    const double *d = rp.dirs + 3 * r;
    double inv[3];
    for (int a = 0; a < 3; a++) inv[a] = d[a] != 0.0 ? 1.0 / d[a] : 0.0;
    if (lane == 0) {
      s_top = 0;
      s_nq = 0;
      s_nh = 0;
      s_ovf = 0;
      s_cnt = 0;
      if (bvh.nc == 1) {
        runq[0] = 0;
        s_nq = 1;
      } else if (bvh.nc > 1) {
        stack[0] = 0;
        s_top = 1;
      }
    }
    __syncthreads();
    // Each iteration: up to 16 wide nodes of the stack (a lane per child
    // box), then up to 64 / kRayLeaf queued runs (kRayLeaf lanes per run, a
    // triangle per lane).
    while (true) {
      const int top = s_top, nq = s_nq;
      if ((top == 0 && nq == 0) || s_ovf) break;
      const int take = top < 16 ? top : 16;
      const int base = top - take;
      const int32_t node = (lane >> 2) < take ? stack[base + (lane >> 2)] : -1;
      constexpr int kRunsPerStep = 64 / kRayLeaf;
      const int qtake = nq < kRunsPerStep ? nq : kRunsPerStep;
      const int qbase = nq - qtake;
      const int32_t run = lane / kRayLeaf < qtake ? runq[qbase + lane / kRayLeaf] : -1;
      __syncthreads();
      int32_t push = -1, leaf = -1;
      if (node >= 0) {
        const BvhWide &nd = bvh.cwide[node];
        const int32_t ch = nd.child[lane & 3];
        if (ch != kNoChild && ray_box(o, d, inv, nd.box[lane & 3])) {
          if (ch < 0)
            leaf = -(ch + 1);
          else
            push = ch;
        }
      }
      if (run >= 0) {
        const int64_t li = (int64_t)run * kRayLeaf + (lane & (kRayLeaf - 1));
        if (li < bvh.n) {
          const Tri64 q = tri[bvh.leaf_tri[li]];
          double tv;
          if (mt_ref(o, d, q.v, q.v + 3, q.v + 6, &tv)) {
            const int slot = atomicAdd(&s_nh, 1);
            if (slot < kWaveHits)
              hit_t[slot] = tv;
            else
              s_ovf = 1;
          }
        }
      }
      const uint64_t b1 = __ballot(push >= 0), c1 = __ballot(leaf >= 0);
      const int ntop = base + __popcll(b1);
      const int nnq = qbase + __popcll(c1);
      if (ntop > kWaveStack || nnq > kRunQueue) {
        if (lane == 0) s_ovf = 1;
      } else {
        if (push >= 0) stack[base + __popcll(b1 & below)] = push;
        if (leaf >= 0) runq[qbase + __popcll(c1 & below)] = leaf;
      }
      __syncthreads();
      if (lane == 0) {
        const bool ok = ntop <= kWaveStack && nnq <= kRunQueue;
        s_top = ok ? ntop : 0;
        s_nq = ok ? nnq : 0;
      }
      __syncthreads();
    }
    const int n = s_nh;
    const bool ovf = s_ovf != 0;
    if (!ovf) {
      for (int a = lane; a < n; a += 64) {
        const double x = hit_t[a];
        bool dup = false, have = false;
        double pred = 0.0;
        for (int b = 0; b < n; b++) {
          const double y = hit_t[b];
          if (y == x && b < a) dup = true;
          if (y < x && (!have || y > pred)) {
            pred = y;
            have = true;
          }
        }
        if (!dup && (!have || x - pred > kRayEps)) atomicAdd(&s_cnt, 1);
      }
    }
    __syncthreads();
    if (lane == 0) {
      if (ovf) {
        const uint32_t q = atomicAdd(nfb, 1u);
        if (q < fb_cap)
          fb[q] = slot * 64u + (uint64_t)r;
        else
          atomicOr(err, kErrHitOverflow);
      } else if (s_cnt & 1) {
        atomicOr(&item.bits, 1ull << r);
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void ray_bits_brute_kernel(
    GridP g, const double *__restrict__ pts, const Tri64 *__restrict__ tri,
    int64_t nel, RayParams rp, PendLists lists, const uint64_t *__restrict__ fb,
    uint32_t nfb, const uint32_t *__restrict__ nfb_dev, uint32_t *err) {
  __shared__ double s_t[kHitCap];
  __shared__ int s_n, s_cnt;
  if (nfb_dev && *nfb_dev < nfb) nfb = *nfb_dev;
  for (uint32_t q = blockIdx.x; q < nfb; q += gridDim.x) {
    const uint64_t slot = fb[q] / 64u;
    const int r = (int)(fb[q] % 64u);
    Pending &item = slot_item(lists, slot);
    double o[3];
    item_point(g, pts, item.idx, o);
    int par = block_ray_parity(o, rp.dirs + 3 * r, tri, nel, s_t, &s_n, &s_cnt, err);
    if (threadIdx.x == 0 && par) atomicOr(&item.bits, 1ull << r);
    __syncthreads();
  }
}

// Replay of adaptive_point_in_mesh_raycast (:234-288) on the traced ray
// parities.  An item of list k may use rays below lists.avail[k]: one whose
// vote needs a later ray (the first min_rays, or the next batch of 4) is
// appended to `next` (count *nnext, capacity next_cap) with its state and
// cleared ray bits; one whose final confidence is below the threshold goes
// to the winding list.
__global__ void vote_kernel(RayParams rp, PendLists lists, FusedOut out,
                            Pending *__restrict__ next, uint32_t *nnext, uint32_t next_cap,
                            WindItem *__restrict__ wind, uint32_t *nwind, uint32_t wind_cap,
                            uint32_t *err) {
  // live items: list k's first n_k, after the lists before it
  uint64_t tend[kPendLists];
  uint64_t n = 0;
  for (int k = 0; k < kPendLists; k++) {
    if (k < lists.count) n += list_count(lists, k);
    tend[k] = n;
  }
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
       t += (uint64_t)gridDim.x * blockDim.x) {
    int k = 0;
    while (t >= tend[k]) k++;
    const uint64_t index = t - (k ? tend[k - 1] : 0u);
    Pending P = lists.items[k][index];
    const int avail = lists.avail[k];
    const unsigned long long m = P.bits;
    int inside = P.inside, total = P.used, used = P.used;
    bool carry = false;
    if (used < rp.min_rays) {
      if (rp.min_rays > avail) {
        carry = true;
      } else {
        for (int r = used; r < rp.min_rays; r++) {
          if ((m >> r) & 1ull) inside++;
          total++;
        }
        used = rp.min_rays;
      }
    }
    double ratio = total > 0 ? (double)inside / (double)total : 0.0;
    double conf = ratio > 1.0 - ratio ? ratio : 1.0 - ratio;
    while (!carry && conf < 0.85 && used < rp.max_rays) {
      int batch = rp.max_rays - used < 4 ? rp.max_rays - used : 4;
      if (used + batch > avail) {
        carry = true;
        break;
      }
      for (int r = used; r < used + batch; r++) {
        if (r < kNumDirs) {
          if ((m >> r) & 1ull) inside++;
          total++;
        }
      }
      used += batch;
      ratio = (double)inside / (double)total;
      conf = ratio > 1.0 - ratio ? ratio : 1.0 - ratio;
      if (conf > 0.95) break;
    }
    if (carry) {
      P.inside = inside;
      P.used = used;
      P.bits = 0ull;
      const uint32_t q = next ? atomicAdd(nnext, 1u) : next_cap;
      if (q < next_cap)
        next[q] = P;
      else
        atomicOr(err, kErrPendingOverflow);
      continue;
    }
    double sgn = ratio > 0.5 ? 1.0 : -1.0;
    if (rp.use_winding && conf < rp.conf_thr) {
      const uint32_t q = atomicAdd(nwind, 1u);
      if (q < wind_cap) {
        WindItem wi;
        wi.idx = P.idx;
        wi.dist = P.dist;
        wi.row = k == lists.ahead_list && index < lists.ahead_rows
                     ? (int64_t)(lists.ahead_base + index)
                     : (int64_t)-1;
        wind[q] = wi;
      }
      continue;
    }
    if (out.sdf) out.sdf[P.idx] = P.dist * sgn;
    if (out.signs) out.signs[P.idx] = sgn;
    if (out.conf) out.conf[P.idx] = conf;
  }
}

// generalized_winding_number (:295-328) of point o over triangles [e0, e1),
// one block: the block's threads take triangles at stride blockDim.x and a
// fixed reduction tree adds their sums (deterministic).  Every thread
// returns the chunk's sum.
__device__ double winding_chunk_sum(const double *o, const Tri64 *__restrict__ tri,
                                    int64_t e0, int64_t e1, double *s_red) {
#pragma clang fp contract(off)
  double acc = 0.0;
  for (int64_t el = e0 + threadIdx.x; el < e1; el += blockDim.x) {
    const Tri64 t = tri[el];
    double r[3], a[3], b[3], cc[3], bc[3];
    for (int k = 0; k < 3; k++) r[k] = t.v[k] - o[k];
    normalize_ref(r, a);
    for (int k = 0; k < 3; k++) r[k] = t.v[3 + k] - o[k];
    normalize_ref(r, b);
    for (int k = 0; k < 3; k++) r[k] = t.v[6 + k] - o[k];
    normalize_ref(r, cc);
    cross3_ref(b, cc, bc);
    double num = dot3_ref(a, bc);
    double den = ((1.0 + dot3_ref(a, b)) + dot3_ref(b, cc)) + dot3_ref(cc, a);
    if (fabs(den) > 1e-10) {
      double sa = 2.0 * atan2(fabs(num), den);
      if (num < 0) sa = -sa;
      acc += sa;
    }
  }
  s_red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s_red[threadIdx.x] += s_red[threadIdx.x + off];
    __syncthreads();
  }
  const double sum = s_red[0];
  __syncthreads();
  return sum;
}

// Chunk sums ahead of the vote: the first min(*count_dev, rows) items of a
// pending list, item i's chunk c into ahead[(base + i) * chunks + c].
__global__ __launch_bounds__(256) void winding_ahead_kernel(
    GridP g, const double *__restrict__ pts, const Tri64 *__restrict__ tri, int64_t nel,
    const Pending *__restrict__ items, const uint32_t *__restrict__ count_dev, uint32_t rows,
    uint32_t base, int chunks, int64_t chunk, double *__restrict__ ahead) {
  __shared__ double s_red[256];
  const uint32_t n = *count_dev < rows ? *count_dev : rows;
  const int64_t nblk = (int64_t)n * chunks;
  for (int64_t q = blockIdx.x; q < nblk; q += gridDim.x) {
    const uint32_t i = (uint32_t)(q / chunks);
    const int64_t c = q % chunks;
    double o[3];
    item_point(g, pts, items[i].idx, o);
    const int64_t e0 = c * chunk, e1 = min(nel, e0 + chunk);
    const double sum = winding_chunk_sum(o, tri, e0, e1, s_red);
    if (threadIdx.x == 0) ahead[((int64_t)base + i) * chunks + c] = sum;
  }
}

// Chunk sums of the winding items without an ahead row: item w's chunk c
// into partial[w * chunks + c].
__global__ __launch_bounds__(256) void winding_partial_kernel(
    GridP g, const double *__restrict__ pts, const Tri64 *__restrict__ tri, int64_t nel,
    const WindItem *__restrict__ wind, const uint32_t *__restrict__ nwind_dev, uint32_t cap,
    int chunks, int64_t chunk, double *__restrict__ partial) {
  __shared__ double s_red[256];
  const uint32_t n = *nwind_dev < cap ? *nwind_dev : cap;
  const int64_t nblk = (int64_t)n * chunks;
  for (int64_t q = blockIdx.x; q < nblk; q += gridDim.x) {
    const uint32_t w = (uint32_t)(q / chunks);
    const int64_t c = q % chunks;
    const WindItem wi = wind[w];
    if (wi.row >= 0) continue;  // computed ahead (uniform per block)
    double o[3];
    item_point(g, pts, wi.idx, o);
    const int64_t e0 = c * chunk, e1 = min(nel, e0 + chunk);
    const double sum = winding_chunk_sum(o, tri, e0, e1, s_red);
    if (threadIdx.x == 0) partial[q] = sum;
  }
}

// One wave per item: lanes add the chunk sums at stride 64 in chunk order,
// then a fixed butterfly; every lane holds the same total.
__global__ __launch_bounds__(64) void winding_final_kernel(
    const WindItem *__restrict__ wind, const uint32_t *__restrict__ nwind_dev, uint32_t cap,
    int chunks, const double *__restrict__ partial, const double *__restrict__ ahead,
    FusedOut out) {
#pragma clang fp contract(off)
  const uint32_t n = *nwind_dev < cap ? *nwind_dev : cap;
  for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
    const WindItem wi = wind[w];
    const double *row = wi.row >= 0 ? ahead + wi.row * chunks : partial + (int64_t)w * chunks;
    double tot = 0.0;
    for (int c = threadIdx.x; c < chunks; c += 64) tot += row[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
    if (threadIdx.x == 0) {
      tot = tot / 12.566370614359172;  // Float64(4pi)
      const double sgn = fabs(tot) > 0.5 ? 1.0 : -1.0;
      if (out.sdf) out.sdf[wi.idx] = wi.dist * sgn;
      if (out.signs) out.signs[wi.idx] = sgn;
      if (out.conf) out.conf[wi.idx] = 0.8;
    }
  }
}

__global__ void dist_points_bvh_kernel(const double *__restrict__ pts,
                                       int64_t npts, const Tri64 *__restrict__ tri,
                                       Bvh bvh, double h, float slack,
                                       double *__restrict__ dist, uint32_t *err) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npts;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    bool found, ovf;
    double d = nearest_bvh(bvh, tri, p, h, slack, &found, &ovf);
    if (ovf) atomicOr(err, kErrStackOverflow);
    dist[i] = d;
  }
}

__global__ void init_pending_all_kernel(Pending *pend, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    Pending p;
    p.idx = i;
    p.dist = 0.0;
    p.inside = 0;
    p.used = 0;
    p.bits = 0ull;
    pend[i] = p;
  }
}

}  // namespace

DistJob make_dist_job(const GridP &g, const Tri64 *tri, const Bins &b, float *d2buf,
                      const uint32_t *only) {
  DistJob J;
  J.g = g;
  J.td = tile_dims(g);
  J.tri = tri;
  J.tile_off = b.tile_off;
  J.tile_tris = b.tile_tris;
  J.chunk_off = b.chunk_off;
  J.chunk_tile = b.chunk_tile;
  J.items = b.chunk_item;
  J.nchunks = b.nchunks;
  J.nchunks_dev = b.nchunks_dev;
  J.only = only;
  J.keys = b.tile_keys;
  J.npairs = b.ntile_pairs;
  J.npairs_dev = b.npairs_dev;
  J.d2buf = d2buf;
  J.clen = b.chunk_len;
  J.pad_ = 0;
  return J;
}

void launch_dist_chunks(hipStream_t s, const GridP &g, const Tri64 *tri,
                        const Bins &b, float *d2buf, const uint32_t *only, int64_t max_blocks) {
  if (b.nchunks <= 0) return;
  const int64_t grid = max_blocks > 0 && max_blocks < b.nchunks ? max_blocks : b.nchunks;
  hipLaunchKernelGGL(dist_chunks_kernel, dim3((unsigned)grid), dim3(kTilePts), 0, s,
                     make_dist_job(g, tri, b, d2buf, only));
  SDF_HIP(hipGetLastError());
}


void launch_dist_pairs(hipStream_t s, const GridP &g, const Tri64 *tri, const Bins &b,
                       float *d2buf, int per_cu, int ncu) {
  if (b.ntile_pairs <= 0) return;
  DistJob J = make_dist_job(g, tri, b, d2buf, nullptr);
  const int64_t nw = (b.ntile_pairs + kTilePts - 1) / kTilePts;
  // at most the workgroups the LDS lets each CU hold (in a persistent grid,
  // workgroups that wait for a slot would start their share late)
  const int fit = (int)(160 * 1024 / sizeof(PairShared));
  if (per_cu <= 0 || per_cu > fit) per_cu = fit;
  const int64_t cap = (int64_t)per_cu * (ncu > 0 ? ncu : 1);
  const int64_t grid = nw < cap ? nw : cap;
  hipLaunchKernelGGL(dist_pairs_kernel, dim3((unsigned)grid), dim3(kTilePts), 0, s, J);
  SDF_HIP(hipGetLastError());
}

void launch_dist_chunks_learned(hipStream_t s, const GridP &g, const Tri64 *tri,
                                const Bins &b, const float *r2buf, float *d2buf,
                                uint64_t *list, unsigned long long *stats) {
  if (b.nchunks <= 0) return;
  TileDims td = tile_dims(g);
  hipLaunchKernelGGL(dist_chunks_learned_kernel, dim3((unsigned)b.nchunks), dim3(kTilePts), 0,
                     s, g, td, tri, b.tile_off, b.tile_end, b.tile_tris, b.chunk_off,
                     b.chunk_tile,
                     b.nchunks, b.nchunks_dev, (int)b.chunk_len, r2buf, d2buf, stats);
  const float hf = (float)g.h;
  hipLaunchKernelGGL(learned_check_kernel, dim3((unsigned)((b.nchunks + 3) / 4)), dim3(256),
                     0, s, b.chunk_off, b.chunk_tile, b.nchunks, b.nchunks_dev, r2buf,
                     (const float *)d2buf, hf * hf, list, stats);
  SDF_HIP(hipGetLastError());
}

void launch_learned_fallback(hipStream_t s, const GridP &g, const Tri64 *tri, const Bvh &bvh,
                             const Bins &b, const uint64_t *list,
                             const unsigned long long *stats, float slack, float *d2buf,
                             uint32_t *err) {
  if (b.nchunks <= 0) return;
  hipLaunchKernelGGL(learned_fallback_kernel, dim3(256), dim3(256), 0, s, g, tile_dims(g), tri,
                     bvh, b.chunk_tile, list, stats, slack, d2buf, err);
  SDF_HIP(hipGetLastError());
}

FusedJob make_fused_job(const GridP &g, const Bins &b, const float *d2buf, bool do_sign,
                        FusedOut out, const PendOut &pend, uint32_t *err) {
  FusedJob J;
  J.g = g;
  J.td = tile_dims(g);
  const SignDims sd = sign_dims(g);
  J.bd.tx = sd.sx;
  J.bd.ty = sd.sy;
  J.bd.tz = sd.sz;
  J.bd.ntiles = (int64_t)sd.sx * sd.sy * sd.sz;
  J.chunk_off = b.d2_by_tile ? nullptr : b.chunk_off;
  J.tile_end = b.d2_by_tile ? b.tile_end : nullptr;
  J.d2buf = d2buf;
  J.inl = b.line_inline;
  J.segw = b.seg_words;
  J.plan = nullptr;
  {
    const SegLayout sl = seg_layout(g);
    J.segoff1 = sl.off[1];
    J.segoff2 = sl.off[2];
    J.segtotal = sl.total;
  }
  J.base1 = b.line_base[1];
  J.base2 = b.line_base[2];
  J.out = out;
  // the spare slot of +inf: after the last work item, or after the last tile
  J.chunk_cap = b.d2_by_tile ? J.td.ntiles : b.nchunks;
  J.nbricks = J.td.ntiles > 0 ? brick_launch_size(J.bd) : 0;
  J.do_sign = do_sign ? 1 : 0;
  J.pad_ = 0;
  J.cold.line_off = b.line_off;
  J.cold.fast = b.line_fast;
  J.cold.meta = b.line_meta;
  J.cold.pend = pend;
  J.cold.err = err;
  return J;
}

void launch_fused_lattice(hipStream_t s, const GridP &g, const Tri64 *tri,
                          const Bins &b, const float *d2buf, bool do_sign,
                          FusedOut out, const PendOut &pend, uint32_t *err,
                          const BrickPlan *plan, int64_t max_blocks, int select,
                          const uint32_t *list, const uint32_t *nlist) {
  (void)tri;
  FusedJob J = make_fused_job(g, b, d2buf, do_sign, out, pend, err);
  if (J.nbricks <= 0) return;
  if (!plan) throw Error{-1, "the sign kernel needs its brick plan"};
  J.plan = plan;
  if (J.nbricks >= ((int64_t)1 << 31)) throw Error{-1, "too many sign bricks for one launch"};
  const int64_t grid = max_blocks > 0 && max_blocks < J.nbricks ? max_blocks : J.nbricks;
  const int set = (out.sdf ? kOutSdf : 0) | (out.dist ? kOutDist : 0) |
                  (out.signs ? kOutSigns : 0) | (out.conf ? kOutConf : 0);
  // the specialised sets fix their stages (fused_brick): sdf = distances +
  // signs, dist = distances, signs + conf = signs; anything else runs the
  // generic instantiation
  if (select != 0 && !(set == kOutSdf && do_sign && d2buf && b.d2_by_tile))
    throw Error{-1, "brick selection needs the sdf output with distances by tile"};
  if (select == 1) {
    // (max_blocks caps the grid: wave slots left for the vote beside it)
    hipLaunchKernelGGL(fused_same_kernel, dim3((unsigned)grid), dim3(64), 0, s, J);
  } else if (select == 2) {
    if (!list || !nlist) throw Error{-1, "brick list missing"};
    const int64_t g2 = J.nbricks < 16384 ? J.nbricks : 16384;
    hipLaunchKernelGGL(fused_list_kernel, dim3((unsigned)g2), dim3(64), 0, s, J, list, nlist);
  } else if (set == kOutSdf && do_sign && d2buf)
    hipLaunchKernelGGL(fused_lattice_kernel<kOutSdf>, dim3((unsigned)grid), dim3(64), 0, s, J);
  else if (set == kOutDist && !do_sign && d2buf)
    hipLaunchKernelGGL(fused_lattice_kernel<kOutDist>, dim3((unsigned)grid), dim3(64), 0, s, J);
  else if (set == (kOutSigns | kOutConf) && do_sign && !d2buf)
    hipLaunchKernelGGL(fused_lattice_kernel<kOutSigns | kOutConf>, dim3((unsigned)grid),
                       dim3(64), 0, s, J);
  else
    hipLaunchKernelGGL(fused_lattice_kernel<kOutAll>, dim3((unsigned)grid), dim3(64), 0, s, J);
  SDF_HIP(hipGetLastError());
}

int64_t brick_plan_size(const GridP &g) {
  const SignDims sd = sign_dims(g);
  TileDims bd;
  bd.tx = sd.sx;
  bd.ty = sd.sy;
  bd.tz = sd.sz;
  bd.ntiles = (int64_t)sd.sx * sd.sy * sd.sz;
  return tile_dims(g).ntiles > 0 ? brick_launch_size(bd) : 0;
}

void launch_brick_plan(hipStream_t s, const GridP &g, const Bins &b, const float *d2buf,
                       BrickPlan *plan) {
  const FusedJob J = make_fused_job(g, b, d2buf, false, FusedOut(), PendOut{}, nullptr);
  if (J.nbricks <= 0) return;
  if (J.nbricks >= ((int64_t)1 << 31)) throw Error{-1, "too many sign bricks for one launch"};
  if (J.bd.tx >= 65536 || J.bd.ty >= 65536)
    throw Error{-3, "grid too wide for the sign kernel's 16-bit brick coordinates"};
  const int64_t blocks = (J.nbricks + 255) / 256;
  hipLaunchKernelGGL(brick_plan_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)),
                     dim3(256), 0, s, J, plan);
  SDF_HIP(hipGetLastError());
}


static int blocks_for(int64_t n, int t) {
  int64_t b = (n + t - 1) / t;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (int)b;
}

// One phase of the adaptive vote: each list's rays [r0, r1) by LBVH
// traversal, overflowing rays by brute force, then the vote replay.
void resolve_pending(hipStream_t s, Arena &ar, const GridP &g, const double *pts,
                     const Tri64 *tri, int64_t nel, const Bvh &bvh, const RayParams &rp,
                     const PendLists &lists, FusedOut out, Pending *next, uint32_t *nnext,
                     uint32_t next_cap, WindItem *wind, uint32_t *nwind, uint32_t wind_cap,
                     uint32_t *d_counts, uint32_t *err, int stack) {
  const uint64_t slots = lists.slots();
  if (slots == 0) return;
  const uint64_t fb_want = slots * 4u + 1024u;
  const uint32_t fb_cap = fb_want < 0xFFFFFFF0ull ? (uint32_t)fb_want : 0xFFFFFFF0u;
  uint64_t *fb = ar.get<uint64_t>(ar.tmp("res_fb"), fb_cap);
  bool rays = false;
  for (int k = 0; k < lists.count; k++) rays = rays || lists.r1[k] > lists.r0[k];
  if (rays) {
    // LDS traversal stack depth (entries); a ray that overflows it is redone
    // by the brute-force kernel, so the depth only trades occupancy
    if (stack <= 1024)
      hipLaunchKernelGGL(ray_wave_kernel<1024>, dim3(kRayWaves), dim3(64), 0, s, g, pts, tri,
                         bvh, rp, lists, fb, d_counts, fb_cap, err);
    else
      hipLaunchKernelGGL(ray_wave_kernel<2048>, dim3(kRayWaves), dim3(64), 0, s, g, pts, tri,
                         bvh, rp, lists, fb, d_counts, fb_cap, err);
    hipLaunchKernelGGL(ray_bits_brute_kernel, dim3(256), dim3(256), 0, s, g, pts, tri, nel, rp,
                       lists, (const uint64_t *)fb, fb_cap, (const uint32_t *)d_counts, err);
    SDF_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(vote_kernel, dim3(256), dim3(256), 0, s, rp, lists, out, next, nnext,
                     next_cap, wind, nwind, wind_cap, err);
  SDF_HIP(hipGetLastError());
}

void resolve_winding(hipStream_t s, Arena &ar, const GridP &g, const double *pts,
                     const Tri64 *tri, int64_t nel, const WindItem *wind,
                     const uint32_t *nwind, uint32_t wind_cap, const double *ahead,
                     FusedOut out) {
  if (wind_cap == 0) return;
  int64_t chunk = 0;
  const int chunks = winding_chunks(nel, &chunk);
  double *partial = ar.get<double>(ar.tmp("res_wpartial"), (size_t)wind_cap * chunks);
  hipLaunchKernelGGL(winding_partial_kernel, dim3(1024), dim3(256), 0, s, g, pts, tri, nel,
                     wind, nwind, wind_cap, chunks, chunk, partial);
  hipLaunchKernelGGL(winding_final_kernel, dim3(wind_cap < 1024u ? wind_cap : 1024u), dim3(64),
                     0, s, wind, nwind, wind_cap, chunks, (const double *)partial, ahead, out);
  SDF_HIP(hipGetLastError());
}

int winding_chunks(int64_t nel, int64_t *chunk) {
  int64_t ch = 4096;
  int chunks = (int)((nel + ch - 1) / ch);
  if (chunks < 1) chunks = 1;
  if (chunks > 1024) {
    chunks = 1024;
    ch = (nel + chunks - 1) / chunks;
  }
  *chunk = ch;
  return chunks;
}

void launch_winding_ahead(hipStream_t s, const GridP &g, const double *pts, const Tri64 *tri,
                          int64_t nel, const Pending *items, const uint32_t *count_dev,
                          uint32_t rows, uint32_t base, double *ahead) {
  if (rows == 0) return;
  int64_t chunk = 0;
  const int chunks = winding_chunks(nel, &chunk);
  // a quarter of the post-vote pass's grid: it runs beside the vote's rays,
  // which it should not crowd
  hipLaunchKernelGGL(winding_ahead_kernel, dim3(256), dim3(256), 0, s, g, pts, tri, nel, items,
                     count_dev, rows, base, chunks, chunk, ahead);
  SDF_HIP(hipGetLastError());
}

void launch_dist_points_bvh(hipStream_t s, const double *pts, int64_t npts,
                            const Tri64 *tri, const Bvh &bvh, double h,
                            float slack, double *dist, uint32_t *err) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(dist_points_bvh_kernel, dim3(blocks_for(npts, 256)),
                     dim3(256), 0, s, pts, npts, tri, bvh, h, slack, dist, err);
  SDF_HIP(hipGetLastError());
}

void launch_init_pending_all(hipStream_t s, Pending *pend, int64_t n) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(init_pending_all_kernel, dim3((unsigned)blocks), dim3(256),
                     0, s, pend, n);
  SDF_HIP(hipGetLastError());
}

}  // namespace sdf
