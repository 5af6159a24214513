// Build-side kernels: triangle gather, exclusive scan, triangle binning into
// distance tiles, and axis-line hit lists for the sign vote.
//
// Binning replaces the reference's per-point KD-tree candidate search
// (dfOnTriangularMesh.jl:27-35,79-94) for lattice points: every triangle is
// listed in each 8x4x2 point tile that lies within the clamp radius h of it,
// so a tile's list holds every triangle that can change min(d, h) for any of
// its points.  Axis-line hit lists replace the per-point BVH ray traversal
// (RaycastingSignDetection.jl:114-136) for the six axis rays.
#include "sdf_internal.h"
#include "sdf_lbvh.h"

namespace sdf {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanBlock = kScanThreads * kScanItems;

__global__ void status_kernel(StatusSrc src, uint64_t *out) {
  const int i = threadIdx.x;
  if (i >= 16) return;
  uint64_t v = 0;
  if (i < src.n && src.p[i])
    v = src.is64[i] ? *(const uint64_t *)src.p[i] : (uint64_t) * (const uint32_t *)src.p[i];
  out[i] = v;
}

__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_u64_kernel(uint64_t *p, int64_t n, uint64_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_f64_kernel(double *p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

inline int grid_for(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (int)b;
}

// grid_for, at most max_blocks workgroups when max_blocks > 0 (grid-stride
// kernels only).
inline int capped_grid(int64_t n, int threads, int64_t max_blocks) {
  const int b = grid_for(n, threads);
  return max_blocks > 0 && max_blocks < b ? (int)max_blocks : b;
}

// One thread per triangle: 1-based IEN -> 9 doubles.  Out-of-range indices
// raise kErrBadIndex and read vertex 1 instead (never out of bounds).  A
// workgroup gathers its 256 triangles into LDS and writes their records as
// one contiguous run (a thread's own 72-byte record would spread each store
// instruction over 4.6 KB).
constexpr int kGatherThreads = 256;
__global__ __launch_bounds__(kGatherThreads) void gather_tris_kernel(
    const double *__restrict__ X, int64_t nnp, const int64_t *__restrict__ IEN, int64_t nel,
    Tri64 *__restrict__ tri, uint32_t *err) {
  __shared__ double s[kGatherThreads * 9];
  double *out = reinterpret_cast<double *>(tri);
  for (int64_t base = (int64_t)blockIdx.x * kGatherThreads; base < nel;
       base += (int64_t)gridDim.x * kGatherThreads) {
    const int64_t t = base + threadIdx.x;
    if (t < nel) {
      for (int c = 0; c < 3; c++) {
        int64_t id = IEN[3 * t + c];
        if (id < 1 || id > nnp) {
          atomicOr(err, kErrBadIndex);
          id = 1;
        }
        s[9 * threadIdx.x + 3 * c + 0] = X[3 * (id - 1) + 0];
        s[9 * threadIdx.x + 3 * c + 1] = X[3 * (id - 1) + 1];
        s[9 * threadIdx.x + 3 * c + 2] = X[3 * (id - 1) + 2];
      }
    }
    __syncthreads();
    const int64_t nv = 9 * (min((int64_t)kGatherThreads, nel - base));
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int e = k * kGatherThreads + threadIdx.x;
      if (e < nv) out[9 * base + e] = s[e];
    }
    __syncthreads();  // s is reused by the next round
  }
}

// Block-level exclusive scan.  in has n_in entries (missing entries are 0),
// out gets n_out entries; bsum[b] = block total.
template <typename TIn>
__global__ __launch_bounds__(kScanThreads) void scan_block_kernel(
    const TIn *__restrict__ in, const TIn *__restrict__ in2, int64_t n_in,
    uint64_t *__restrict__ out, int64_t n_out, uint64_t *__restrict__ bsum) {
  __shared__ uint64_t buf[2][kScanThreads];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)tid * kScanItems;
  uint64_t v[kScanItems];
  uint64_t tot = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; q++) {
    int64_t i = base + q;
    uint64_t x = (i < n_in) ? (uint64_t)in[i] + (in2 ? (uint64_t)in2[i] : 0ull) : 0ull;
    v[q] = tot;
    tot += x;
  }
  buf[0][tid] = tot;
  __syncthreads();
  int src = 0;
  for (int off = 1; off < kScanThreads; off <<= 1) {
    uint64_t x = buf[src][tid];
    if (tid >= off) x += buf[src][tid - off];
    buf[src ^ 1][tid] = x;
    src ^= 1;
    __syncthreads();
  }
  uint64_t incl = buf[src][tid];
  uint64_t excl = incl - tot;
#pragma unroll
  for (int q = 0; q < kScanItems; q++) {
    int64_t i = base + q;
    if (i < n_out) out[i] = excl + v[q];
  }
  if (tid == kScanThreads - 1) bsum[blockIdx.x] = incl;
}

__global__ void scan_add_kernel(uint64_t *__restrict__ out, int64_t n,
                                const uint64_t *__restrict__ boff) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] += boff[i / kScanBlock];
}

// Second pass for up to kScanFlatBlocks scan blocks: workgroup b sums the
// totals of blocks 0..b-1 itself and adds that to block b's outputs (one
// launch instead of a scan of the totals and an add pass).
constexpr int64_t kScanFlatBlocks = 8192;
__global__ __launch_bounds__(kScanThreads) void scan_add_flat_kernel(
    uint64_t *__restrict__ out, int64_t n_out, const uint64_t *__restrict__ bsum) {
  __shared__ uint64_t red[kScanThreads / 64];
  const int64_t b = blockIdx.x;
  if (b == 0) return;  // uniform: block 0 adds nothing
  uint64_t p = 0;
  for (int64_t j = threadIdx.x; j < b; j += kScanThreads) p += bsum[j];
  for (int off = 32; off > 0; off >>= 1) p += __shfl_down(p, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = p;
  __syncthreads();
  uint64_t pre = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; w++) pre += red[w];
  const int64_t base = b * kScanBlock;
#pragma unroll
  for (int q = 0; q < kScanItems; q++) {
    const int64_t i = base + q * kScanThreads + threadIdx.x;
    if (i < n_out) out[i] += pre;
  }
}

// Exclusive scan of in (+ in2 elementwise when not null) into out.
template <typename TIn>
void scan_rec(hipStream_t s, Arena &ar, const TIn *in, int64_t n_in,
              uint64_t *out, int64_t n_out, int level, const TIn *in2 = nullptr) {
  int64_t nb = (n_out + kScanBlock - 1) / kScanBlock;
  if (nb < 1) nb = 1;
  uint64_t *bsum = ar.get<uint64_t>(ar.tmp("scan_bsum" + std::to_string(level)), nb + 1);
  hipLaunchKernelGGL(scan_block_kernel<TIn>, dim3((unsigned)nb),
                     dim3(kScanThreads), 0, s, in, in2, n_in, out, n_out, bsum);
  SDF_HIP(hipGetLastError());
  if (nb > 1 && nb <= kScanFlatBlocks) {
    hipLaunchKernelGGL(scan_add_flat_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s,
                       out, n_out, (const uint64_t *)bsum);
    SDF_HIP(hipGetLastError());
  } else if (nb > 1) {
    uint64_t *boff = ar.get<uint64_t>(ar.tmp("scan_boff" + std::to_string(level)), nb + 1);
    scan_rec<uint64_t>(s, ar, bsum, nb, boff, nb, level + 1);
    hipLaunchKernelGGL(scan_add_kernel, dim3(grid_for(n_out, 256)), dim3(256),
                       0, s, out, n_out, boff);
    SDF_HIP(hipGetLastError());
  }
}

__device__ inline void tri_minmax(const Tri64 &t, double *mn, double *mx) {
  for (int d = 0; d < 3; d++) {
    double a = t.v[d], b = t.v[3 + d], c = t.v[6 + d];
    mn[d] = fmin(a, fmin(b, c));
    mx[d] = fmax(a, fmax(b, c));
  }
}

// Inclusive lattice index range [lo, hi] of points within [a, b] on one axis.
__device__ inline void point_range(double a, double b, double o, double h,
                                   int32_t n, int32_t *lo, int32_t *hi) {
  double fa = ceil((a - o) / h - 1e-7);
  double fb = floor((b - o) / h + 1e-7);
  if (fa < 0.0) fa = 0.0;
  if (fb > (double)(n - 1)) fb = (double)(n - 1);
  if (fa > fb || !(fa == fa) || !(fb == fb)) {
    *lo = 1;
    *hi = 0;
    return;
  }
  *lo = (int32_t)fa;
  *hi = (int32_t)fb;
}

struct TileRange {
  int32_t lo[3], hi[3];  // tile index range (inclusive)
  int64_t count;
};

__device__ inline TileRange tri_tile_range(const Tri64 &t, const GridP &g,
                                           double R) {
  double mn[3], mx[3];
  tri_minmax(t, mn, mx);
  const double o[3] = {g.ox, g.oy, g.oz};
  const int32_t n[3] = {g.nx, g.ny, g.nz};
  const int32_t tsz[3] = {kTileX, kTileY, kTileZ};
  TileRange r;
  r.count = 1;
  for (int d = 0; d < 3; d++) {
    int32_t plo, phi;
    point_range(mn[d] - R, mx[d] + R, o[d], g.h, n[d], &plo, &phi);
    if (plo > phi) {
      r.count = 0;
      r.lo[d] = 1;
      r.hi[d] = 0;
      continue;
    }
    r.lo[d] = plo / tsz[d];
    r.hi[d] = phi / tsz[d];
    r.count *= (int64_t)(r.hi[d] - r.lo[d] + 1);
  }
  return r;
}

// Does the tile's point box, grown by R on every side, meet the triangle's
// plane?  A point within R of the triangle has its nearest triangle point
// inside that grown box, so rejecting on this test never drops a needed
// triangle.
__device__ inline bool tile_plane_test(const Tri64 &t, const GridP &g,
                                       const TileDims &td, int32_t tx,
                                       int32_t ty, int32_t tz, double R) {
  double e1[3] = {t.v[3] - t.v[0], t.v[4] - t.v[1], t.v[5] - t.v[2]};
  double e2[3] = {t.v[6] - t.v[0], t.v[7] - t.v[1], t.v[8] - t.v[2]};
  double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                 e1[0] * e2[1] - e1[1] * e2[0]};
  double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (!(nn > 0.0)) return true;
  int32_t i0 = tx * kTileX, j0 = ty * kTileY, k0 = tz * kTileZ;
  int32_t i1 = min(i0 + kTileX - 1, g.nx - 1), j1 = min(j0 + kTileY - 1, g.ny - 1),
          k1 = min(k0 + kTileZ - 1, g.nz - 1);
  double lo[3] = {lattice_coord(g.ox, g.h, i0), lattice_coord(g.oy, g.h, j0),
                  lattice_coord(g.oz, g.h, k0)};
  double hi[3] = {lattice_coord(g.ox, g.h, i1), lattice_coord(g.oy, g.h, j1),
                  lattice_coord(g.oz, g.h, k1)};
  double proj = 0.0, dist = 0.0, mag = 0.0;
  for (int d = 0; d < 3; d++) {
    double c = 0.5 * (lo[d] + hi[d]);
    double e = 0.5 * (hi[d] - lo[d]) + R;
    proj += fabs(n[d]) * e;
    dist += n[d] * (c - t.v[d]);
    mag += fabs(n[d]) * (fabs(c) + fabs(t.v[d]));
  }
  return fabs(dist) <= proj * (1.0 + 1e-9) + 1e-12 * mag;
}

constexpr int64_t kSmallMax = 256;  // candidates looped by one thread

// Atomic-free tile binning: per triangle, count the tiles that pass the
// plane test, scan, emit (tile, triangle) pairs at the triangle's offset,
// then sort the pairs by tile (stable LDS radix sort).  Tile lists come out
// in triangle order, independent of scheduling.  Triangles with more than
// kSmallMax candidate tiles are counted and emitted by one workgroup each.
// Also zeroes the nzero words of the tiles' run bounds (a side job that
// saves the chain a launch).
__global__ void tile_count_small_kernel(const Tri64 *__restrict__ tri, int64_t nel,
                                        GridP g, TileDims td, double R,
                                        uint32_t *__restrict__ cnt,
                                        uint32_t *__restrict__ big, uint32_t *nbig,
                                        uint64_t *__restrict__ zero, int64_t nzero) {
  for (int64_t z = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; z < nzero;
       z += (int64_t)gridDim.x * blockDim.x)
    zero[z] = 0ull;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nel;
       t += (int64_t)gridDim.x * blockDim.x) {
    const Tri64 tr = tri[t];
    const TileRange r = tri_tile_range(tr, g, R);
    uint32_t c = 0;
    if (r.count > kSmallMax) {
      big[atomicAdd(nbig, 1u)] = (uint32_t)t;
    } else if (r.count > 0) {
      for (int32_t tz = r.lo[2]; tz <= r.hi[2]; tz++)
        for (int32_t ty = r.lo[1]; ty <= r.hi[1]; ty++)
          for (int32_t tx = r.lo[0]; tx <= r.hi[0]; tx++)
            c += tile_plane_test(tr, g, td, tx, ty, tz, R) ? 1u : 0u;
    }
    cnt[t] = c;
  }
}

__global__ void tile_emit_small_kernel(const Tri64 *__restrict__ tri, int64_t nel,
                                       GridP g, TileDims td, double R,
                                       const uint64_t *__restrict__ toff,
                                       uint32_t *__restrict__ keys,
                                       uint32_t *__restrict__ vals, uint64_t cap) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nel;
       t += (int64_t)gridDim.x * blockDim.x) {
    const Tri64 tr = tri[t];
    const TileRange r = tri_tile_range(tr, g, R);
    if (r.count == 0 || r.count > kSmallMax) continue;
    uint64_t pos = toff[t];
    for (int32_t tz = r.lo[2]; tz <= r.hi[2]; tz++)
      for (int32_t ty = r.lo[1]; ty <= r.hi[1]; ty++)
        for (int32_t tx = r.lo[0]; tx <= r.hi[0]; tx++)
          if (tile_plane_test(tr, g, td, tx, ty, tz, R)) {
            if (pos < cap) {  // beyond: overflow, reported by the sizes
              keys[pos] = (uint32_t)(((int64_t)tz * td.ty + ty) * td.tx + tx);
              vals[pos] = (uint32_t)t;
            }
            pos++;
          }
  }
}

// One workgroup (256 threads) per large triangle; mode 0 counts, mode 1
// emits in candidate order (block prefix sum over the pass flags).
__global__ __launch_bounds__(256) void tile_big_kernel(
    const Tri64 *__restrict__ tri, GridP g, TileDims td, double R, int mode,
    const uint32_t *__restrict__ big, const uint32_t *nbig,
    uint32_t *__restrict__ cnt, const uint64_t *__restrict__ toff,
    uint32_t *__restrict__ keys, uint32_t *__restrict__ vals, uint64_t cap) {
  __shared__ uint32_t s_scan[256];
  __shared__ uint32_t s_base;
  const uint32_t nb = *nbig;
  for (uint32_t q = blockIdx.x; q < nb; q += gridDim.x) {
    const uint32_t t = big[q];
    const Tri64 tr = tri[t];
    const TileRange r = tri_tile_range(tr, g, R);
    const int64_t wx = r.hi[0] - r.lo[0] + 1, wy = r.hi[1] - r.lo[1] + 1;
    if (threadIdx.x == 0) s_base = 0;
    __syncthreads();
    for (int64_t c0 = 0; c0 < r.count; c0 += blockDim.x) {
      const int64_t c = c0 + threadIdx.x;
      int32_t tx = 0, ty = 0, tz = 0;
      uint32_t pass = 0;
      if (c < r.count) {
        tx = r.lo[0] + (int32_t)(c % wx);
        ty = r.lo[1] + (int32_t)((c / wx) % wy);
        tz = r.lo[2] + (int32_t)(c / (wx * wy));
        pass = tile_plane_test(tr, g, td, tx, ty, tz, R) ? 1u : 0u;
      }
      s_scan[threadIdx.x] = pass;
      __syncthreads();
      for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? s_scan[threadIdx.x - off] : 0u;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
      }
      const uint32_t incl = s_scan[threadIdx.x];
      if (mode == 1 && pass) {
        const uint64_t pos = toff[t] + s_base + incl - 1;
        if (pos < cap) {
          keys[pos] = (uint32_t)(((int64_t)tz * td.ty + ty) * td.tx + tx);
          vals[pos] = t;
        }
      }
      __syncthreads();
      if (threadIdx.x == 255) s_base += incl;
      __syncthreads();
    }
    if (mode == 0 && threadIdx.x == 0) cnt[t] = s_base;
    __syncthreads();
  }
}

// Run bounds of every non-empty tile from the sorted tile keys: the first
// pair of a run writes its begin, the last its end (one coalesced pass; the
// bounds of empty tiles stay 0).  With d2 (distances by tile, ntiles + 1
// slots): the spare slot and the slot of every tile whose run reaches past
// the next window boundary (a multiple of kTilePts pairs), which several
// windows combine by atomicMin, are set to +inf.
__global__ void tile_bounds_kernel(const uint32_t *__restrict__ keys, int64_t n,
                                   const uint64_t *__restrict__ n_dev,
                                   uint64_t *__restrict__ begin, uint64_t *__restrict__ end,
                                   float *__restrict__ d2, int64_t ntiles) {
  const float4 inf4 = make_float4(__builtin_huge_valf(), __builtin_huge_valf(),
                                  __builtin_huge_valf(), __builtin_huge_valf());
  if (d2 && blockIdx.x == 0 && threadIdx.x < kTilePts / 4)
    reinterpret_cast<float4 *>(d2 + ntiles * kTilePts)[threadIdx.x] = inf4;
  if (n_dev && (int64_t)*n_dev < n) n = (int64_t)*n_dev;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) {
      begin[k] = (uint64_t)i;
      const int64_t wb = (i | (kTilePts - 1)) + 1;  // the next window boundary
      if (d2 && wb < n && keys[wb] == k) {
        float4 *slot = reinterpret_cast<float4 *>(d2 + (int64_t)k * kTilePts);
#pragma unroll
        for (int q = 0; q < kTilePts / 4; q++) slot[q] = inf4;
      }
    }
    if (i == n - 1 || keys[i + 1] != k) end[k] = (uint64_t)(i + 1);
  }
}

// With the run bounds deferred (tile_bins_bounds on another stream): the
// spare slot of d2 and the slots tile_bounds_kernel would set to +inf, one
// thread per window boundary wb (a run holding pairs wb - 1 and wb).
__global__ void split_slots_kernel(const uint32_t *__restrict__ keys, int64_t n,
                                   const uint64_t *__restrict__ n_dev, float *__restrict__ d2,
                                   int64_t ntiles) {
  const float4 inf4 = make_float4(__builtin_huge_valf(), __builtin_huge_valf(),
                                  __builtin_huge_valf(), __builtin_huge_valf());
  if (blockIdx.x == 0 && threadIdx.x < kTilePts / 4)
    reinterpret_cast<float4 *>(d2 + ntiles * kTilePts)[threadIdx.x] = inf4;
  if (n_dev && (int64_t)*n_dev < n) n = (int64_t)*n_dev;
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x + 1; m * kTilePts < n;
       m += (int64_t)gridDim.x * blockDim.x) {
    const int64_t wb = m * kTilePts;
    const uint32_t k = keys[wb];
    if (keys[wb - 1] == k) {
      float4 *slot = reinterpret_cast<float4 *>(d2 + (int64_t)k * kTilePts);
#pragma unroll
      for (int q = 0; q < kTilePts / 4; q++) slot[q] = inf4;
    }
  }
}

// ---- axis-line hits ----

struct LineRange {
  int32_t lo[3][2], hi[3][2];  // per axis: the two transverse index ranges
  int64_t cnt[3];
};

__device__ inline LineRange tri_line_range(const Tri64 &t, const GridP &g) {
  double mn[3], mx[3];
  tri_minmax(t, mn, mx);
  const double o[3] = {g.ox, g.oy, g.oz};
  const int32_t n[3] = {g.nx, g.ny, g.nz};
  int32_t plo[3], phi[3];
  for (int d = 0; d < 3; d++) {
    // pad by a small fraction of a cell; the exact test decides
    double pad = 1e-6 * g.h;
    point_range(mn[d] - pad, mx[d] + pad, o[d], g.h, n[d], &plo[d], &phi[d]);
  }
  LineRange r;
  for (int ax = 0; ax < 3; ax++) {
    int b = ax == 0 ? 1 : 0;
    int c = ax == 2 ? 1 : 2;
    r.lo[ax][0] = plo[b];
    r.hi[ax][0] = phi[b];
    r.lo[ax][1] = plo[c];
    r.hi[ax][1] = phi[c];
    int64_t w0 = phi[b] - plo[b] + 1, w1 = phi[c] - plo[c] + 1;
    r.cnt[ax] = (w0 > 0 && w1 > 0) ? w0 * w1 : 0;
  }
  return r;
}

// Line (ax, a, b): its id and its origin, the lattice point with index 0
// along the axis.
__device__ inline int64_t line_origin(const GridP &g, int ax, int32_t a, int32_t b,
                                      int64_t base1, int64_t base2, double *o) {
  if (ax == 0) {  // lines over (j, k)
    o[0] = lattice_coord(g.ox, g.h, 0);
    o[1] = lattice_coord(g.oy, g.h, a);
    o[2] = lattice_coord(g.oz, g.h, b);
    return (int64_t)b * g.ny + a;
  } else if (ax == 1) {  // lines over (i, k)
    o[0] = lattice_coord(g.ox, g.h, a);
    o[1] = lattice_coord(g.oy, g.h, 0);
    o[2] = lattice_coord(g.oz, g.h, b);
    return base1 + (int64_t)b * g.nx + a;
  }
  // lines over (i, j)
  o[0] = lattice_coord(g.ox, g.h, a);
  o[1] = lattice_coord(g.oy, g.h, b);
  o[2] = lattice_coord(g.oz, g.h, 0);
  return base2 + (int64_t)b * g.nx + a;
}

__device__ inline void line_visit(const Tri64 &tr, uint32_t tid, const GridP &g, int ax,
                                  int32_t a, int32_t b, int64_t base1,
                                  int64_t base2, int mode,
                                  uint32_t *__restrict__ cnt,
                                  const uint64_t *__restrict__ off,
                                  uint32_t *__restrict__ hits, uint32_t *__restrict__ hline,
                                  uint64_t cap) {
  double o[3];
  const int64_t line = line_origin(g, ax, a, b, base1, base2, o);
  if (!line_hit_test(o, ax, tr.v, tr.v + 3, tr.v + 6, nullptr)) return;
  if (mode == 0) {
    atomicAdd(&cnt[line], 1u);
  } else {
    const uint32_t slot = atomicAdd(&cnt[line], 1u);  // cnt is the cursor here
    const uint64_t pos = off[line] + slot;
    if (pos < cap) {  // beyond: overflow, reported by the sizes
      hits[pos] = tid;
      hline[pos] = (uint32_t)line;
    }
  }
}

// Triangles with at most kSmallMax candidate lines, one thread each.  The
// counting pass runs the exact hit test; a triangle's first kEmitBatch hits
// (in candidate order) keep their line ids and take their slots in `cnt` by
// one batch of returning atomics at the end (hitn[t] = hits, kHitsMany for
// more; lines and slots at hitl / hslot[kEmitBatch t ..]), further hits are
// counted in `cnt2`.  The emit pass then writes the triangle's id at the
// first hits' slots without atomics (line_prep_kernel forms each hit's
// record from its triangle); for a triangle with more hits it tests again,
// skips the first kEmitBatch hits and takes cursor slots in `cnt`, which
// continue after every slot the counting pass gave.
constexpr int kEmitBatch = 4;
constexpr uint8_t kHitsMany = 255;

__global__ void line_count_small_kernel(const Tri64 *__restrict__ tri, int64_t nel, GridP g,
                                        int64_t base1, int64_t base2,
                                        uint32_t *__restrict__ cnt, uint32_t *__restrict__ cnt2,
                                        uint32_t *__restrict__ big, uint32_t *nbig,
                                        uint8_t *__restrict__ hitn,
                                        uint32_t *__restrict__ hitl,
                                        uint32_t *__restrict__ hslot) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nel;
       t += (int64_t)gridDim.x * blockDim.x) {
    const Tri64 tr = tri[t];
    const LineRange r = tri_line_range(tr, g);
    const int64_t total = r.cnt[0] + r.cnt[1] + r.cnt[2];
    int nh = 0;
    uint32_t ln[kEmitBatch] = {0, 0, 0, 0};
    if (total > kSmallMax) {
      big[atomicAdd(nbig, 1u)] = (uint32_t)t;
    } else if (total > 0) {
      for (int ax = 0; ax < 3; ax++)
        for (int32_t b = r.lo[ax][1]; b <= r.hi[ax][1]; b++)
          for (int32_t a = r.lo[ax][0]; a <= r.hi[ax][0]; a++) {
            double o[3];
            const int64_t line = line_origin(g, ax, a, b, base1, base2, o);
            if (!line_hit_test(o, ax, tr.v, tr.v + 3, tr.v + 6, nullptr)) continue;
            if (nh >= kEmitBatch) atomicAdd(&cnt2[line], 1u);
#pragma unroll
            for (int k = 0; k < kEmitBatch; k++) ln[k] = nh == k ? (uint32_t)line : ln[k];
            nh++;
          }
    }
    hitn[t] = nh > kEmitBatch ? kHitsMany : (uint8_t)nh;
    if (nh > 0) {
      uint32_t sl[kEmitBatch];
#pragma unroll
      for (int k = 0; k < kEmitBatch; k++) sl[k] = k < nh ? atomicAdd(&cnt[ln[k]], 1u) : 0u;
#pragma unroll
      for (int k = 0; k < kEmitBatch; k++) {
        hitl[kEmitBatch * t + k] = ln[k];
        hslot[kEmitBatch * t + k] = sl[k];
      }
    }
  }
}

__global__ void line_emit_small_kernel(const Tri64 *__restrict__ tri, int64_t nel, GridP g,
                                       int64_t base1, int64_t base2,
                                       const uint8_t *__restrict__ hitn,
                                       const uint32_t *__restrict__ hitl,
                                       const uint32_t *__restrict__ hslot,
                                       uint32_t *__restrict__ cursor,
                                       const uint64_t *__restrict__ off,
                                       uint32_t *__restrict__ hits,
                                       uint32_t *__restrict__ hline, uint64_t cap) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nel;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int nh = hitn[t];
    if (nh == 0) continue;
    const int nf = nh < kEmitBatch ? nh : kEmitBatch;  // hits with a slot
#pragma unroll
    for (int k = 0; k < kEmitBatch; k++) {
      if (k < nf) {
        const uint32_t ln = hitl[kEmitBatch * t + k];
        const uint64_t pos = off[ln] + hslot[kEmitBatch * t + k];
        if (pos < cap) {  // beyond: overflow, reported by the sizes
          hits[pos] = (uint32_t)t;
          hline[pos] = ln;
        }
      }
    }
    if (nh == kHitsMany) {  // the hits after the first kEmitBatch, in the same order
      const Tri64 tr = tri[t];
      const LineRange r = tri_line_range(tr, g);
      int h = 0;
      for (int ax = 0; ax < 3; ax++)
        for (int32_t b = r.lo[ax][1]; b <= r.hi[ax][1]; b++)
          for (int32_t a = r.lo[ax][0]; a <= r.hi[ax][0]; a++) {
            double o[3];
            const int64_t line = line_origin(g, ax, a, b, base1, base2, o);
            if (!line_hit_test(o, ax, tr.v, tr.v + 3, tr.v + 6, nullptr)) continue;
            if (h++ < kEmitBatch) continue;
            const uint64_t pos = off[line] + atomicAdd(&cursor[line], 1u);
            if (pos < cap) {
              hits[pos] = (uint32_t)t;
              hline[pos] = (uint32_t)line;
            }
          }
    }
  }
}

__global__ void line_bin_big_kernel(const Tri64 *__restrict__ tri, GridP g,
                                    int64_t base1, int64_t base2, int mode,
                                    uint32_t *__restrict__ cnt,
                                    const uint64_t *__restrict__ off,
                                    uint32_t *__restrict__ hits, uint32_t *__restrict__ hline,
                                    uint64_t cap, const uint32_t *__restrict__ big,
                                    const uint32_t *nbig) {
  const uint32_t nb = *nbig;
  for (uint32_t q = blockIdx.x; q < nb; q += gridDim.x) {
    const uint32_t tid = big[q];
    const Tri64 tr = tri[tid];
    const LineRange r = tri_line_range(tr, g);
    const int64_t total = r.cnt[0] + r.cnt[1] + r.cnt[2];
    for (int64_t c = threadIdx.x; c < total; c += blockDim.x) {
      int64_t local = c;
      int ax = 0;
      while (ax < 2 && local >= r.cnt[ax]) {
        local -= r.cnt[ax];
        ax++;
      }
      const int64_t w0 = r.hi[ax][0] - r.lo[ax][0] + 1;
      line_visit(tr, tid, g, ax, r.lo[ax][0] + (int32_t)(local % w0),
                 r.lo[ax][1] + (int32_t)(local / w0), base1, base2, mode, cnt, off,
                 hits, hline, cap);
    }
  }
}

__global__ void tile_chunks_kernel(const uint64_t *__restrict__ tile_off,
                                   const uint64_t *__restrict__ tile_end, int64_t ntiles,
                                   int clen, uint32_t *__restrict__ nchunk) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntiles;
       t += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t c = tile_end[t] - tile_off[t];
    nchunk[t] = (uint32_t)((c + clen - 1) / clen);
  }
}

__global__ void chunk_items_kernel(const uint64_t *__restrict__ chunk_off,
                                   const uint64_t *__restrict__ tile_off,
                                   const uint64_t *__restrict__ tile_end, int64_t ntiles,
                                   int clen, uint32_t *__restrict__ items,
                                   DistItem *__restrict__ recs, uint64_t cap,
                                   float *__restrict__ d2buf) {
  const float4 inf4 = make_float4(__builtin_huge_valf(), __builtin_huge_valf(),
                                  __builtin_huge_valf(), __builtin_huge_valf());
  if (d2buf && blockIdx.x == 0 && threadIdx.x < kTilePts / 4)  // the spare slot
    reinterpret_cast<float4 *>(d2buf + cap * kTilePts)[threadIdx.x] = inf4;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntiles;
       t += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t c0 = chunk_off[t], c1 = chunk_off[t + 1];
    if (c1 == c0) continue;  // empty tile
    if (d2buf && c1 - c0 > 1 && c0 < cap) {  // the slot several work items atomicMin into
      float4 *slot = reinterpret_cast<float4 *>(d2buf + c0 * kTilePts);
#pragma unroll
      for (int k = 0; k < kTilePts / 4; k++) slot[k] = inf4;
    }
    const uint64_t tb = tile_off[t], te = tile_end[t];
    const uint64_t single = c1 - c0 == 1 ? 1ull << 55 : 0ull;
    for (uint64_t q = c0; q < c1 && q < cap; q++) {
      items[q] = (uint32_t)t;
      const uint64_t b = tb + (q - c0) * (uint64_t)clen;
      const uint64_t n = min((uint64_t)clen, te - b);
      DistItem r;
      r.tile = (uint32_t)t;
      r.first = (uint32_t)c0;
      r.range = b | n << 48 | single;
      recs[q] = r;
    }
  }
}

// Segment bytes (SegLayout) of line l of axis ax from its finished summary
// li: per segment, the uniformity test of the sign kernel on the summary
// (seg_uniform on the crossings' segment ranges, or inline_uniform on the
// segment's coordinate range when the line has none), so the sign kernel
// reads one byte per line and brick.  With segment ranges and at most 128
// segments the test runs on all segments at once (seg_words_masks).
// Segments of word w (bytes) whose byte is 1 (uniform, both rays even) or 7
// (uniform, both odd), into 128-bit masks m1 = (m[0], m[1]), m7 = (m[2], m[3]).
__device__ inline void word_same_masks(uint32_t word, int w, uint64_t m[4]) {
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const uint32_t byte = (word >> (8 * b)) & 0xFFu;
    const int s = 4 * w + b;
    const uint64_t bit = s < 128 ? 1ull << (s & 63) : 0ull;
    const bool hi = s >= 64;
    if (byte == 1u) {
      m[0] |= hi ? 0ull : bit;
      m[1] |= hi ? bit : 0ull;
    }
    if (byte == 7u) {
      m[2] |= hi ? 0ull : bit;
      m[3] |= hi ? bit : 0ull;
    }
  }
}

// lmask (nullable): the line's byte-1 and byte-7 segment masks (4 words),
// zero beyond 128 segments.
__device__ inline void write_seg_words(const GridP &g, const LineInline &li, int ax, int64_t l,
                                       const SegLayout &sl, uint32_t *__restrict__ segw,
                                       uint64_t *__restrict__ lmask) {
  LineTail r;
  memcpy(&r, reinterpret_cast<const char *>(&li) + 48, sizeof(r));
  uint32_t *out = segw + sl.off[ax] + l;
  const int64_t stride = sl.nl[ax];
  const int ns = sl.ns[ax];
  uint64_t m[4] = {0ull, 0ull, 0ull, 0ull};
  if ((r.info & kLineSeg) && ns <= 128) {
    seg_words_masks(r, ns, [&](int w, uint32_t word) { out[(int64_t)w * stride] = word; }, m);
    if (lmask)
      for (int q = 0; q < 4; q++) lmask[q] = m[q];
    return;
  }
  const double oa = ax == 0 ? g.ox : (ax == 1 ? g.oy : g.oz);
  const int32_t na = ax == 0 ? g.nx : (ax == 1 ? g.ny : g.nz);
  const int len = kSegLen[ax];
  const bool seg = (r.info & kLineSeg) != 0u;
  uint32_t word = 0u;
  for (int sgi = 0; sgi < ns; sgi++) {
    int p, q;
    bool u;
    if (seg) {
      u = seg_uniform(r, sgi, &p, &q);
    } else {
      const int64_t e = (int64_t)sgi * len + len - 1;
      const double lo = lattice_coord(oa, g.h, (int64_t)sgi * len);
      const double hi = lattice_coord(oa, g.h, e < na - 1 ? e : na - 1);
      u = inline_uniform(li, lo, hi, &p, &q);
    }
    word |= (u ? (uint32_t)(1 | (p << 1) | (q << 2)) : 0u) << (8 * (sgi & 3));
    if ((sgi & 3) == 3 || sgi == ns - 1) {
      out[(int64_t)(sgi >> 2) * stride] = word;
      word_same_masks(word, sgi >> 2, m);
      word = 0u;
    }
  }
  if (lmask)
    for (int q = 0; q < 4; q++) lmask[q] = m[q];
}

// Fast-path summaries of each axis line (one thread per line).
//
// Error model (FP64 without contraction, unit roundoff u = 2^-53): for a hit
// with record (v1, e1, e2, f) on a line along axis a, the reference's
// t(x) = f * dot(e2, cross(o - v1, e1)) at the point with coordinate x is
// compared with c - x, where c = o0_a + t(o0) is formed once at the line's
// first point o0.  In exact arithmetic on the FP64 inputs, t is affine in x
// with slope -f * a_exact = -1 + beta, |beta| <= 2u (M_a |f| + 1), M_a the
// two products forming a.  The evaluation error of t is at most
// 8u (|f| M_t + |t|), M_t = sum_i |e2_i| (S_j |e1_k| + S_k |e1_j|) with S the
// largest |o - v1| components on the line.  Hence
//   |t(x) - (c - x)| <= beta L + 2 E + u (|c| + |t0| + 2L) =: delta / 2
// (L the line length); delta stores twice that, rounded up to FP32.  Two
// consecutive sorted crossings c_k, c_k+1 belong to one cluster when
// c_k+1 - c_k < eps - 2 Dmax and start a new one when the gap exceeds
// eps + 2 Dmax (Dmax the line's largest delta); any gap in between makes
// the whole line take the exact path.
// Geometry of axis line L: its axis, first and last lattice points, length
// and the sign bricks' segments along it (kSegLen points each).
struct LineGeom {
  int ax, seglen;
  int32_t nax;
  int64_t nseg, lidx;
  double o0[3], on[3], len, oa;
  double o0a, ona;  // o0[ax], on[ax]
};
// v[i] by selects (a private array indexed at run time would live in scratch)
__device__ inline double sel3(const double *v, int i) {
  return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}
__device__ inline LineGeom line_geom(const GridP &g, int64_t L, int64_t base1, int64_t base2) {
  LineGeom G;
  int64_t a, b;
  if (L < base1) {
    G.ax = 0;
    G.lidx = L;
    a = L % g.ny;
    b = L / g.ny;
  } else if (L < base2) {
    G.ax = 1;
    G.lidx = L - base1;
    a = G.lidx % g.nx;
    b = G.lidx / g.nx;
  } else {
    G.ax = 2;
    G.lidx = L - base2;
    a = G.lidx % g.nx;
    b = G.lidx / g.nx;
  }
  G.nax = G.ax == 0 ? g.nx : (G.ax == 1 ? g.ny : g.nz);
  G.seglen = kSegLen[G.ax];
  G.nseg = ((int64_t)G.nax + G.seglen - 1) / G.seglen;
  if (G.ax == 0) {
    G.o0[1] = G.on[1] = lattice_coord(g.oy, g.h, a);
    G.o0[2] = G.on[2] = lattice_coord(g.oz, g.h, b);
    G.o0[0] = lattice_coord(g.ox, g.h, 0);
    G.on[0] = lattice_coord(g.ox, g.h, G.nax - 1);
  } else if (G.ax == 1) {
    G.o0[0] = G.on[0] = lattice_coord(g.ox, g.h, a);
    G.o0[2] = G.on[2] = lattice_coord(g.oz, g.h, b);
    G.o0[1] = lattice_coord(g.oy, g.h, 0);
    G.on[1] = lattice_coord(g.oy, g.h, G.nax - 1);
  } else {
    G.o0[0] = G.on[0] = lattice_coord(g.ox, g.h, a);
    G.o0[1] = G.on[1] = lattice_coord(g.oy, g.h, b);
    G.o0[2] = lattice_coord(g.oz, g.h, 0);
    G.on[2] = lattice_coord(g.oz, g.h, G.nax - 1);
  }
  G.o0a = sel3(G.o0, G.ax);
  G.ona = sel3(G.on, G.ax);
  G.len = G.ona - G.o0a;
  G.oa = G.ax == 0 ? g.ox : (G.ax == 1 ? g.oy : g.oz);
  return G;
}

// Per hit (one thread each, hits in their lines' slots): the crossing
// coordinate c, its certified bound delta and, on lines of at most 255
// segments, the segments the crossing leaves uniform (below a: crossing
// above them; from b on: crossing below them, by the comparisons the sign
// kernel makes, found by bisection as they are monotone in the segment),
// written as fast[i] = {c, delta, 0, a | b << 8}; line_prep_kernel puts each
// line's hits in order.  The hit's record is formed from its triangle by
// line_hit_test's arithmetic (the emit pass stores only the triangle's id).
__global__ void hit_prep_kernel(GridP g, const uint32_t *__restrict__ hit_tri,
                                const uint32_t *__restrict__ hit_line,
                                const uint64_t *__restrict__ n_dev, uint64_t cap,
                                const Tri64 *__restrict__ tri, int64_t base1, int64_t base2,
                                LineFast *__restrict__ fast) {
#pragma clang fp contract(off)
  const double u = 1.1102230246251565e-16;
  const uint64_t n = *n_dev < cap ? *n_dev : cap;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const LineGeom G = line_geom(g, (int64_t)hit_line[i], base1, base2);
    const int ax = G.ax, j = (ax + 1) % 3, k = (ax + 2) % 3;
    LineHit r;
    {
      const Tri64 &tr = tri[hit_tri[i]];
      line_hit_record(ax, tr.v, tr.v + 3, tr.v + 6, &r);
    }
    const double t0 = line_hit_t(r, G.o0);
    const double c = G.o0a + t0;
    const double v1a = sel3(r.v1, ax);
    const double sa = fmax(fabs(G.o0a - v1a), fabs(G.ona - v1a));
    double S[3];
#pragma unroll
    for (int d = 0; d < 3; d++) S[d] = d == ax ? sa : fabs(G.o0[d] - r.v1[d]);
    double Mt = 0.0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int q1 = (q + 1) % 3, q2 = (q + 2) % 3;
      Mt += fabs(r.e2[q]) * (S[q1] * fabs(r.e1[q2]) + S[q2] * fabs(r.e1[q1]));
    }
    const double af = fabs(r.f);
    const double Ma = fabs(sel3(r.e1, j) * sel3(r.e2, k)) + fabs(sel3(r.e1, k) * sel3(r.e2, j));
    const double beta = 2.0 * u * (Ma * af + 1.0);
    const double E = 8.0 * u * (af * Mt + fabs(t0) + 2.0 * G.len);
    double delta = 2.0 * (beta * G.len + 2.0 * E + u * (fabs(c) + fabs(t0) + 2.0 * G.len)) +
                   1e-300;
    delta *= 1.0 + 1e-6;
    LineFast lf;
    lf.c = c;
    lf.delta = __double2float_ru(delta);
    lf.cid = 0;
    lf.pad_ = 0;
    if (G.nseg <= 255) {
      const double D = (double)lf.delta;
      auto seg_lo = [&](int64_t q) { return lattice_coord(G.oa, g.h, q * G.seglen); };
      auto seg_hi = [&](int64_t q) {
        const int64_t e = q * G.seglen + G.seglen - 1;
        return lattice_coord(G.oa, g.h, e < G.nax - 1 ? e : G.nax - 1);
      };
      int64_t lo = 0, hi = G.nseg;
      while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (c - seg_hi(mid) > kRayEps + D)
          lo = mid + 1;
        else
          hi = mid;
      }
      const int64_t sa = lo;
      lo = 0;
      hi = G.nseg;
      while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (c - seg_lo(mid) < -kRayEps - D)
          hi = mid;
        else
          lo = mid + 1;
      }
      lf.pad_ = (uint16_t)(sa | (lo << 8));
    }
    fast[i] = lf;
  }
}

// Per line (one thread each): its hits (hit_prep_kernel's summaries) in
// order of c, their clusters, the inline summary and the segment bytes.
__global__ void line_prep_kernel(GridP g, const uint64_t *__restrict__ line_off, uint64_t cap,
                                 int64_t base1, int64_t base2, int64_t nlines,
                                 LineFast *__restrict__ fast,
                                 uint32_t *__restrict__ meta,
                                 LineInline *__restrict__ inl, SegLayout sl,
                                 uint32_t *__restrict__ segw, uint64_t *__restrict__ lmask) {
#pragma clang fp contract(off)
  const double u = 1.1102230246251565e-16;
  // the sentinel read by lanes whose line serves no point (byte 1
  // everywhere), and the one of byte 7 (bricks whose lines all carry 7)
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    segw[sl.total] = 0x01010101u;
    segw[sl.total + 1] = 0x07070707u;
  }
  for (int64_t L = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; L < nlines;
       L += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t hb = line_off[L], he = line_off[L + 1];
    const int64_t n = (int64_t)(he - hb);
    LineInline li;
    for (int q = 0; q < kInlineHits; q++) {
      li.c[q] = 0.0;
      li.delta[q] = 0.f;
      li.cid[q] = 0;
      li.seg[q] = 0;
    }
    // sign-brick segments along the line (kSegLen points each)
    const int lax = L < base1 ? 0 : (L < base2 ? 1 : 2);
    const int32_t lnax = lax == 0 ? g.nx : (lax == 1 ? g.ny : g.nz);
    const int seglen = kSegLen[lax];
    const int64_t nseg = ((int64_t)lnax + seglen - 1) / seglen;
    const uint32_t segflag = nseg <= 255 ? kLineSeg : 0u;
    const int64_t lidx = L - (lax == 0 ? 0 : (lax == 1 ? base1 : base2));
    if (n == 0) {
      meta[L] = 0u;
      li.info = segflag;
      inl[L] = li;
      write_seg_words(g, li, lax, lidx, sl, segw, lmask ? lmask + 4 * L : nullptr);
      continue;
    }
    if (n > kLineFastMax || he > cap) {  // he > cap: overflow, the call is re-run
      meta[L] = kLineExact;
      li.info = kLineExact | 255u;
      inl[L] = li;
      write_seg_words(g, li, lax, lidx, sl, segw, lmask ? lmask + 4 * L : nullptr);
      continue;
    }
    // clusters of the sorted hits: a gap beyond eps by more than the bounds
    // starts a new cluster; a gap within the bounds of eps has no fast path
    auto gap_class = [&](double gap, double dmax) -> int {
      const double m = 2.0 * dmax + 4.0 * u * fabs(gap) + 1e-300;
      if (gap - kRayEps > m) return 1;
      return (gap - kRayEps < -m) ? 0 : -1;
    };
    uint32_t cl = 0;
    bool ok = true;
    if (n <= kInlineHits) {
      // short line: sorted in registers (stable bubble passes), no global
      // round trips beyond the one read
      double cs[kInlineHits];
      float ds[kInlineHits];
      uint16_t ss[kInlineHits];
#pragma unroll
      for (int q = 0; q < kInlineHits; q++) {
        if (q < n) {
          const LineFast lf = fast[hb + q];
          cs[q] = lf.c;
          ds[q] = lf.delta;
          ss[q] = lf.pad_;
        } else {
          cs[q] = __builtin_huge_val();
          ds[q] = 0.f;
          ss[q] = 0;
        }
      }
#pragma unroll
      for (int pass = kInlineHits - 1; pass > 0; pass--)
#pragma unroll
        for (int q = 0; q < pass; q++)
          if (cs[q] > cs[q + 1]) {
            const double tc = cs[q];
            cs[q] = cs[q + 1];
            cs[q + 1] = tc;
            const float td = ds[q];
            ds[q] = ds[q + 1];
            ds[q + 1] = td;
            const uint16_t ts = ss[q];
            ss[q] = ss[q + 1];
            ss[q + 1] = ts;
          }
      double dmax = 0.0;
#pragma unroll
      for (int q = 0; q < kInlineHits; q++) dmax = fmax(dmax, (double)ds[q]);
      uint8_t cid[kInlineHits] = {0, 0, 0, 0};
#pragma unroll
      for (int q = 1; q < kInlineHits; q++) {
        if (q < n && ok) {
          const int gc = gap_class(cs[q] - cs[q - 1], dmax);
          if (gc < 0) ok = false;
          cl += gc > 0 ? 1u : 0u;
          cid[q] = (uint8_t)cl;
        }
      }
#pragma unroll
      for (int q = 0; q < kInlineHits; q++) {
        if (q < n) {
          LineFast lf;
          lf.c = cs[q];
          lf.delta = ds[q];
          lf.cid = cid[q];
          lf.pad_ = 0;
          fast[hb + q] = lf;
          if (ok) {
            li.c[q] = cs[q];
            li.delta[q] = ds[q];
            li.cid[q] = cid[q];
            if (segflag) li.seg[q] = ss[q];
          }
        }
      }
    } else {
      // longer line: insertion sort in place (global memory)
      double dmax = 0.0;
      for (int64_t q = 0; q < n; q++) {
        LineFast lf = fast[hb + q];
        lf.cid = 0;
        lf.pad_ = 0;
        int64_t p = q;
        while (p > 0 && fast[hb + p - 1].c > lf.c) {
          fast[hb + p] = fast[hb + p - 1];
          p--;
        }
        fast[hb + p] = lf;
        dmax = fmax(dmax, (double)lf.delta);
      }
      double prev = fast[hb].c;
      for (int64_t q = 1; q < n && ok; q++) {
        const double c = fast[hb + q].c;
        const int gc = gap_class(c - prev, dmax);
        if (gc < 0) ok = false;
        cl += gc > 0 ? 1u : 0u;
        fast[hb + q].cid = (uint16_t)cl;
        prev = c;
      }
    }
    meta[L] = ok ? (cl + 1u) : kLineExact;
    li.info = (uint32_t)(n < 255 ? n : 255) | ((cl + 1u) & 0xFFu) << 8 | (ok ? 0u : kLineExact) |
              segflag;
    inl[L] = li;
    write_seg_words(g, li, lax, lidx, sl, segw, lmask ? lmask + 4 * L : nullptr);
  }
}

// Per brick row of each axis, the AND over the brick's lines of that axis
// (inside the grid) of their byte-1 and byte-7 segment masks, one wave per
// row: part 0 the x-lines of bricks (., by, bz), a mask over bx; part 1 the
// y-lines of (bx, ., bz) over by; part 2 the z-lines of (bx, by, .) over bz.
__global__ __launch_bounds__(64) void brick_same_reduce_kernel(GridP g,
                                                               const uint64_t *__restrict__ lmask,
                                                               int64_t base1, int64_t base2,
                                                               SameRows rows, int32_t nbz) {
  const int lane = threadIdx.x;
  const int32_t nbx = rows.nbx, nby = rows.nby;
  const int64_t n0 = (int64_t)nby * nbz, n1 = (int64_t)nbx * nbz, n2 = (int64_t)nbx * nby;
  for (int64_t w = blockIdx.x; w < n0 + n1 + n2; w += gridDim.x) {
    int64_t L = 0;
    bool valid;
    uint64_t *dst;
    if (w < n0) {
      const int32_t by = (int32_t)(w % nby), bz = (int32_t)(w / nby);
      const int32_t j = by * 8 + (lane & 7), k = bz * kSignZ + (lane >> 3);
      valid = lane < 8 * kSignZ && j < g.ny && k < g.nz;
      L = (int64_t)k * g.ny + j;
      dst = rows.x + 4 * w;
    } else if (w < n0 + n1) {
      const int64_t r = w - n0;
      const int32_t bx = (int32_t)(r % nbx), bz = (int32_t)(r / nbx);
      const int32_t i = bx * 8 + (lane & 7), k = bz * kSignZ + (lane >> 3);
      valid = lane < 8 * kSignZ && i < g.nx && k < g.nz;
      L = base1 + (int64_t)k * g.nx + i;
      dst = rows.y + 4 * r;
    } else {
      const int64_t r = w - n0 - n1;
      const int32_t bx = (int32_t)(r % nbx), by = (int32_t)(r / nbx);
      const int32_t i = bx * 8 + (lane & 7), j = by * 8 + (lane >> 3);
      valid = i < g.nx && j < g.ny;
      L = base2 + (int64_t)j * g.nx + i;
      dst = rows.z + 4 * r;
    }
    uint64_t v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = valid ? lmask[4 * L + q] : ~0ull;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
      for (int q = 0; q < 4; q++) v[q] &= __shfl_xor(v[q], off);
    if (lane == 0)
      for (int q = 0; q < 4; q++) dst[q] = v[q];
  }
}

// Bits 2 (every line byte 1: outside) and 3 (every line byte 7: inside) of
// each brick's plan flags, from the three row masks.
// rest (nullable): the other bricks' launch positions, in rest[0, *nrest)
// (*nrest zero on entry; one atomic per wave).
__global__ __launch_bounds__(256) void brick_same_flags_kernel(
    BrickPlan *__restrict__ plan, int64_t nbricks, const uint64_t *__restrict__ xm,
    const uint64_t *__restrict__ ym, const uint64_t *__restrict__ zm, int32_t nbx, int32_t nby,
    uint32_t *__restrict__ rest, uint32_t *nrest) {
  // kPer bricks per thread (m0 + r * 256 + thread), and the unmarked ones
  // listed with one counter atomic per workgroup: a wave's atomic on the one
  // counter each (5 K of them on the star surface's 322 K bricks) took most
  // of this pass's 20 us
  constexpr int kPer = 4;
  __shared__ uint32_t wcnt[4], bbase;
  auto bit = [](const uint64_t *M, int32_t s, int which) -> uint32_t {
    const uint64_t w = s < 64 ? M[2 * which] : M[2 * which + 1];
    return (uint32_t)((w >> (s & 63)) & 1ull);
  };
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int64_t m0 = (int64_t)blockIdx.x * (256 * kPer); m0 < nbricks;
       m0 += (int64_t)gridDim.x * (256 * kPer)) {
    bool other[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
      const int64_t m = m0 + r * 256 + threadIdx.x;
      other[r] = false;
      if (m < nbricks) {
        // xy, z, flags, pad in one 16-byte load
        const uint4 t = reinterpret_cast<const uint4 *>(plan + m)[1];
        const uint32_t flags = t.z;
        if (flags & 2u) {  // not a padding position
          const int32_t bx = (int32_t)(t.x & 0xFFFFu), by = (int32_t)(t.x >> 16),
                        bz = (int32_t)t.y;
          const uint64_t *X = xm + 4 * ((int64_t)bz * nby + by);
          const uint64_t *Y = ym + 4 * ((int64_t)bz * nbx + bx);
          const uint64_t *Z = zm + 4 * ((int64_t)by * nbx + bx);
          const uint32_t s1 = bit(X, bx, 0) & bit(Y, by, 0) & bit(Z, bz, 0);
          const uint32_t s7 = bit(X, bx, 1) & bit(Y, by, 1) & bit(Z, bz, 1);
          if (s1 | s7)
            plan[m].flags = flags | (s1 << 2) | (s7 << 3);
          else
            other[r] = true;
        }
      }
    }
    if (rest) {  // uniform
      uint64_t bal[kPer];
      uint32_t tot = 0;
#pragma unroll
      for (int r = 0; r < kPer; r++) {
        bal[r] = __ballot(other[r]);
        tot += (uint32_t)__popcll(bal[r]);
      }
      if (lane == 0) wcnt[wv] = tot;
      __syncthreads();
      if (threadIdx.x == 0) {
        const uint32_t sum = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        bbase = sum ? atomicAdd(nrest, sum) : 0u;
      }
      __syncthreads();
      uint32_t base = bbase;
      for (int k = 0; k < wv; k++) base += wcnt[k];
      const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
      for (int r = 0; r < kPer; r++) {
        if (other[r])
          rest[base + (uint32_t)__popcll(bal[r] & below)] = (uint32_t)(m0 + r * 256 + threadIdx.x);
        base += (uint32_t)__popcll(bal[r]);
      }
      __syncthreads();  // wcnt and bbase are rewritten by the next round
    }
  }
}

uint64_t read_u64(hipStream_t s, const uint64_t *d) {
  uint64_t v = 0;
  SDF_HIP(hipMemcpyAsync(&v, d, sizeof(v), hipMemcpyDeviceToHost, s));
  SDF_HIP(hipStreamSynchronize(s));
  return v;
}

}  // namespace

void launch_status(hipStream_t s, const StatusSrc &src, uint64_t *out) {
  hipLaunchKernelGGL(status_kernel, dim3(1), dim3(64), 0, s, src, out);
  SDF_HIP(hipGetLastError());
}

void launch_fill_u32(hipStream_t s, uint32_t *p, int64_t n, uint32_t v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     p, n, v);
  SDF_HIP(hipGetLastError());
}
void launch_fill_u64(hipStream_t s, uint64_t *p, int64_t n, uint64_t v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     p, n, v);
  SDF_HIP(hipGetLastError());
}
void launch_fill_f64(hipStream_t s, double *p, int64_t n, double v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_f64_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     p, n, v);
  SDF_HIP(hipGetLastError());
}

void launch_gather_tris(hipStream_t s, const double *X, int64_t nnp,
                        const int64_t *IEN, int64_t nel, Tri64 *tri,
                        uint32_t *err) {
  if (nel <= 0) return;
  hipLaunchKernelGGL(gather_tris_kernel, dim3(grid_for(nel, kGatherThreads)),
                     dim3(kGatherThreads), 0, s, X, nnp, IEN, nel, tri, err);
  SDF_HIP(hipGetLastError());
}

void scan_counts(hipStream_t s, Arena &ar, const uint32_t *in, uint64_t *out,
                 int64_t n) {
  scan_rec<uint32_t>(s, ar, in, n, out, n + 1, 0);
}

void tile_bins_begin(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                     const GridP &g, double radius, TileBuild &tb) {
  TileDims td = tile_dims(g);
  tb.R = radius * (1.0 + 1e-6);
  tb.tcnt = ar.get<uint32_t>("bin_tri_cnt", nel + 1);
  tb.toff = ar.get<uint64_t>("bin_tri_off", nel + 2);
  tb.big = ar.get<uint32_t>("bin_big", nel + 1);
  tb.bounds = ar.get<uint64_t>("bin_tile_bounds", 2 * (td.ntiles + 1));
  tb.nbig = ar.get<uint32_t>("bin_nbig", 1);
  launch_fill_u32(s, tb.nbig, 1, 0u);
  if (nel > 0) {
    hipLaunchKernelGGL(tile_count_small_kernel, dim3(grid_for(nel, 256)), dim3(256), 0, s,
                       tri, nel, g, td, tb.R, tb.tcnt, tb.big, tb.nbig, tb.bounds,
                       2 * (td.ntiles + 1));
    hipLaunchKernelGGL(tile_big_kernel, dim3(1024), dim3(256), 0, s, tri, g, td, tb.R, 0,
                       (const uint32_t *)tb.big, (const uint32_t *)tb.nbig, tb.tcnt,
                       (const uint64_t *)nullptr, (uint32_t *)nullptr, (uint32_t *)nullptr,
                       (uint64_t)0);
    SDF_HIP(hipGetLastError());
    scan_counts(s, ar, tb.tcnt, tb.toff, nel);
  }
}

void tile_bins_finish(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                      const GridP &g, TileBuild &tb, Bins &b, int64_t cap, float *d2tile,
                      bool defer_bounds) {
  TileDims td = tile_dims(g);
  uint64_t *tile_off = tb.bounds, *tile_end = tb.bounds + td.ntiles + 1;
  const uint64_t *n_dev = nel > 0 ? tb.toff + nel : nullptr;
  if (cap < 0) cap = nel > 0 ? (int64_t)read_u64(s, tb.toff + nel) : 0;
  uint32_t *keys = ar.get<uint32_t>("bin_tile_keys", cap + 1);
  uint32_t *tile_tris = ar.get<uint32_t>("bin_tile_tris", cap + 1);
  if (cap > 0 && nel > 0) {
    hipLaunchKernelGGL(tile_emit_small_kernel, dim3(grid_for(nel, 256)), dim3(256), 0, s,
                       tri, nel, g, td, tb.R, (const uint64_t *)tb.toff, keys, tile_tris,
                       (uint64_t)cap);
    hipLaunchKernelGGL(tile_big_kernel, dim3(1024), dim3(256), 0, s, tri, g, td, tb.R, 1,
                       (const uint32_t *)tb.big, (const uint32_t *)tb.nbig, tb.tcnt,
                       (const uint64_t *)tb.toff, keys, tile_tris, (uint64_t)cap);
    SDF_HIP(hipGetLastError());
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) <= td.ntiles) bits++;
    // own scratch names: the sorted lists may stay in the scratch pair
    const std::string scope = ar.scope;
    ar.scope = scope + "tile/";
    radix_sort_u32_pairs(s, ar, keys, tile_tris, cap, bits, n_dev, &keys, &tile_tris);
    ar.scope = scope;
    if (defer_bounds && d2tile) {
      const int64_t nwin = (cap + kTilePts - 1) / kTilePts;
      hipLaunchKernelGGL(split_slots_kernel, dim3(grid_for(nwin, 256)), dim3(256), 0, s,
                         (const uint32_t *)keys, cap, n_dev, d2tile, td.ntiles);
    } else {
      hipLaunchKernelGGL(tile_bounds_kernel, dim3(grid_for(cap, 256)), dim3(256), 0, s,
                         (const uint32_t *)keys, cap, n_dev, tile_off, tile_end, d2tile,
                         td.ntiles);
    }
    SDF_HIP(hipGetLastError());
  } else {
    if (nel <= 0) launch_fill_u64(s, tb.bounds, 2 * (td.ntiles + 1), 0ull);  // no counting pass
    if (d2tile)  // the spare slot
      launch_fill_u32(s, (uint32_t *)d2tile + td.ntiles * kTilePts, kTilePts, 0x7F800000u);
  }
  b.tile_off = tile_off;
  b.tile_end = tile_end;
  b.tile_tris = tile_tris;
  b.tile_keys = keys;
  b.d2_by_tile = d2tile != nullptr;
  b.ntile_pairs = cap;
  b.npairs_dev = n_dev;
  b.bounds_deferred = defer_bounds && d2tile && cap > 0 && nel > 0;
}

void tile_bins_bounds(hipStream_t s, const GridP &g, const Bins &b) {
  if (!b.bounds_deferred) return;
  hipLaunchKernelGGL(tile_bounds_kernel, dim3(grid_for(b.ntile_pairs, 256)), dim3(256), 0, s,
                     b.tile_keys, b.ntile_pairs, b.npairs_dev, b.tile_off, b.tile_end,
                     (float *)nullptr, tile_dims(g).ntiles);
  SDF_HIP(hipGetLastError());
}

void build_tile_bins(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                     const GridP &g, double radius, Bins &b) {
  TileBuild tb;
  tile_bins_begin(s, ar, tri, nel, g, radius, tb);
  tile_bins_finish(s, ar, tri, nel, g, tb, b);
}

void build_dist_chunks(hipStream_t s, Arena &ar, const GridP &g, Bins &b, int64_t cap,
                       int chunk_len, float **d2buf) {
  if (chunk_len < 1 || chunk_len > kTilePts) chunk_len = kTilePts;
  TileDims td = tile_dims(g);
  uint32_t *nch = ar.get<uint32_t>("chunk_cnt", td.ntiles + 1);
  uint64_t *coff = ar.get<uint64_t>("chunk_off", td.ntiles + 2);
  hipLaunchKernelGGL(tile_chunks_kernel, dim3(grid_for(td.ntiles, 256)), dim3(256),
                     0, s, (const uint64_t *)b.tile_off, (const uint64_t *)b.tile_end,
                     td.ntiles, chunk_len, nch);
  SDF_HIP(hipGetLastError());
  scan_counts(s, ar, nch, coff, td.ntiles);
  if (cap < 0) cap = (int64_t)read_u64(s, coff + td.ntiles);
  uint32_t *items = ar.get<uint32_t>("chunk_tile", cap + 1);
  DistItem *recs = ar.get<DistItem>("chunk_item", cap + 1);
  float *d2 = nullptr;
  if (d2buf) {
    d2 = ar.get<float>("d2buf", ((size_t)cap + 1) * kTilePts);
    *d2buf = d2;
    if (cap <= 0) launch_fill_u32(s, (uint32_t *)d2, kTilePts, 0x7F800000u);  // spare slot
  }
  if (cap > 0) {
    // the packed header keeps triangle-list indices below 2^48
    if (b.ntile_pairs >= ((int64_t)1 << 48)) throw Error{-1, "too many tile-triangle pairs"};
    hipLaunchKernelGGL(chunk_items_kernel, dim3(grid_for(td.ntiles, 256)), dim3(256),
                       0, s, (const uint64_t *)coff, (const uint64_t *)b.tile_off,
                       (const uint64_t *)b.tile_end, td.ntiles, chunk_len, items, recs,
                       (uint64_t)cap, d2);
    SDF_HIP(hipGetLastError());
  }
  b.chunk_off = coff;
  b.chunk_tile = items;
  b.chunk_item = recs;
  b.nchunks = cap;
  b.nchunks_dev = coff + td.ntiles;
  b.chunk_len = chunk_len;
}

void line_hits_begin(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                     const GridP &g, LineBuild &lb, Bins &b) {
  lb.nlx = (int64_t)g.ny * g.nz;
  lb.nly = (int64_t)g.nx * g.nz;
  const int64_t nlz = (int64_t)g.nx * g.ny;
  lb.nlines = lb.nlx + lb.nly + nlz;
  if (lb.nlines >= (int64_t)0xFFFFFFFFll) throw Error{-1, "too many grid lines"};
  b.line_base[0] = 0;
  b.line_base[1] = lb.nlx;
  b.line_base[2] = lb.nlx + lb.nly;
  b.line_base[3] = lb.nlines;
  // lcnt: the first hits of each triangle (and, from the emit pass on, the
  // cursor of the others); lcnt2: the other hits; one fill clears both
  lb.lcnt = ar.get<uint32_t>("line_cnt", 2 * (lb.nlines + 1));
  lb.lcnt2 = lb.lcnt + lb.nlines + 1;
  lb.loff = ar.get<uint64_t>("line_off", lb.nlines + 2);
  lb.big = ar.get<uint32_t>("line_big", nel + 1);
  lb.nbig = ar.get<uint32_t>("line_nbig", 1);
  launch_fill_u32(s, lb.lcnt, 2 * (lb.nlines + 1), 0u);
  launch_fill_u32(s, lb.nbig, 1, 0u);
  if (nel > 0) {
    lb.hitn = ar.get<uint8_t>("line_hitn", nel + 1);
    lb.hitl = ar.get<uint32_t>("line_hitl", (size_t)kEmitBatch * nel + 1);
    lb.hslot = ar.get<uint32_t>("line_hslot", (size_t)kEmitBatch * nel + 1);
    hipLaunchKernelGGL(line_count_small_kernel, dim3(capped_grid(nel, 256, lb.max_blocks)),
                       dim3(256), 0, s, tri, nel, g, lb.nlx, lb.nlx + lb.nly, lb.lcnt, lb.lcnt2,
                       lb.big, lb.nbig, lb.hitn, lb.hitl, lb.hslot);
    hipLaunchKernelGGL(line_bin_big_kernel, dim3(1024), dim3(256), 0, s, tri, g, lb.nlx,
                       lb.nlx + lb.nly, 0, lb.lcnt2, (const uint64_t *)nullptr,
                       (uint32_t *)nullptr, (uint32_t *)nullptr, (uint64_t)0,
                       (const uint32_t *)lb.big,
                       (const uint32_t *)lb.nbig);
    SDF_HIP(hipGetLastError());
  }
  scan_rec<uint32_t>(s, ar, lb.lcnt, lb.nlines, lb.loff, lb.nlines + 1, 0,
                     (const uint32_t *)lb.lcnt2);
}

void line_hits_finish(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                      const GridP &g, LineBuild &lb, Bins &b, int64_t cap) {
  const uint64_t *n_dev = lb.loff + lb.nlines;
  if (cap < 0) cap = (int64_t)read_u64(s, lb.loff + lb.nlines);
  uint32_t *hits = ar.get<uint32_t>("line_hit_tri", cap + 1);
  uint32_t *hline = ar.get<uint32_t>("line_hit_line", cap + 1);
  if (cap > 0 && nel > 0) {
    // lcnt continues as the cursor of the hits without a slot
    hipLaunchKernelGGL(line_emit_small_kernel, dim3(capped_grid(nel, 256, lb.max_blocks)),
                       dim3(256), 0, s, tri, nel, g, lb.nlx, lb.nlx + lb.nly,
                       (const uint8_t *)lb.hitn, (const uint32_t *)lb.hitl,
                       (const uint32_t *)lb.hslot, lb.lcnt, (const uint64_t *)lb.loff, hits,
                       hline, (uint64_t)cap);
    hipLaunchKernelGGL(line_bin_big_kernel, dim3(1024), dim3(256), 0, s, tri, g, lb.nlx,
                       lb.nlx + lb.nly, 1, lb.lcnt, (const uint64_t *)lb.loff, hits, hline,
                       (uint64_t)cap, (const uint32_t *)lb.big, (const uint32_t *)lb.nbig);
    SDF_HIP(hipGetLastError());
  }
  LineFast *fast = ar.get<LineFast>("line_fast", cap + 1);
  uint32_t *meta = ar.get<uint32_t>("line_meta", lb.nlines + 1);
  LineInline *inl = ar.get<LineInline>("line_inline", lb.nlines + 1);
  // the summaries, then each line's segment bytes (SegLayout) from them
  const SegLayout sl = seg_layout(g);
  uint32_t *segw = ar.get<uint32_t>("line_segw", (size_t)sl.total + 2);
  if (cap > 0 && nel > 0)
    hipLaunchKernelGGL(hit_prep_kernel, dim3(capped_grid(cap, 256, lb.max_blocks)), dim3(256), 0,
                       s, g, (const uint32_t *)hits, (const uint32_t *)hline, n_dev,
                       (uint64_t)cap, tri, lb.nlx, lb.nlx + lb.nly, fast);
  // the lines' byte-1 / byte-7 segment masks (launch_brick_same; lines of
  // at most 128 segments)
  const bool masks = lb.same_masks && sl.ns[0] <= 128 && sl.ns[1] <= 128 && sl.ns[2] <= 128;
  uint64_t *lmask = masks ? ar.get<uint64_t>("line_same_mask", 4 * (size_t)lb.nlines + 4) : nullptr;
  hipLaunchKernelGGL(line_prep_kernel, dim3(capped_grid(lb.nlines, 256, lb.max_blocks)),
                     dim3(256), 0, s, g, (const uint64_t *)lb.loff, (uint64_t)cap, lb.nlx,
                     lb.nlx + lb.nly, lb.nlines, fast, meta, inl, sl, segw, lmask);
  SameRows rows;
  if (masks) {
    const SignDims sd = sign_dims(g);
    rows.nbx = sd.sx;
    rows.nby = sd.sy;
    const int64_t nx4 = 4 * (int64_t)sd.sy * sd.sz, ny4 = 4 * (int64_t)sd.sx * sd.sz,
                  nz4 = 4 * (int64_t)sd.sx * sd.sy;
    rows.x = ar.get<uint64_t>("brick_same_rows", (size_t)(nx4 + ny4 + nz4));
    rows.y = rows.x + nx4;
    rows.z = rows.y + ny4;
    const int64_t nrows = (nx4 + ny4 + nz4) / 4;
    hipLaunchKernelGGL(brick_same_reduce_kernel, dim3((unsigned)(nrows < 65536 ? nrows : 65536)),
                       dim3(64), 0, s, g, (const uint64_t *)lmask, lb.nlx, lb.nlx + lb.nly, rows,
                       sd.sz);
  }
  b.same_rows = rows;

  SDF_HIP(hipGetLastError());
  b.line_off = lb.loff;
  b.seg_words = segw;
  b.line_hit_tri = hits;
  b.line_fast = fast;
  b.line_meta = meta;
  b.line_inline = inl;
  b.nline_hits = cap;
  b.nhits_dev = n_dev;
}

bool launch_brick_same(hipStream_t s, const GridP &g, const Bins &b, BrickPlan *plan,
                       uint32_t *rest, uint32_t *nrest) {
  if (!b.same_rows.x || !plan) return false;
  const int64_t nbricks = brick_plan_size(g);
  if (nbricks <= 0) return false;
  hipLaunchKernelGGL(brick_same_flags_kernel, dim3(grid_for((nbricks + 3) / 4, 256)), dim3(256), 0,
                     s, plan,
                     nbricks, (const uint64_t *)b.same_rows.x, (const uint64_t *)b.same_rows.y,
                     (const uint64_t *)b.same_rows.z, b.same_rows.nbx, b.same_rows.nby, rest,
                     nrest);
  SDF_HIP(hipGetLastError());
  return true;
}

void build_line_hits(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                     const GridP &g, Bins &b) {
  LineBuild lb;
  line_hits_begin(s, ar, tri, nel, g, lb, b);
  line_hits_finish(s, ar, tri, nel, g, lb, b);
}

}  // namespace sdf
