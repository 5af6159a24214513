// Linear BVH over triangles: 30-bit Morton codes of the centroids, an LDS
// radix sort (4 passes of 8 bits, stable), the Karras (2012) hierarchy and
// FP32 node boxes rounded outward, refit from a segment tree over the leaves.
//
// Replaces the reference's ImplicitBVH bounding-sphere tree
// (RaycastingSignDetection.jl:35-56) for ray casting and the triangle-centre
// KD-tree (dfOnTriangularMesh.jl:27-35) for distance queries at arbitrary
// points.  Boxes enclose their triangles exactly (outward rounding), so
// pruning never drops a triangle that the FP64 predicates would accept.
#include "sdf_internal.h"
#include "sdf_lbvh.h"

namespace sdf {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

inline int blocks_for(int64_t n, int t) {
  int64_t b = (n + t - 1) / t;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (int)b;
}

__device__ inline uint32_t float_order(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float order_float(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Triangle boxes (outward-rounded FP32) and the centroid bounds, reduced
// per block in LDS into part[6 * block + k] (no atomics: the blocks' bounds
// would all hit the same six words).
__global__ __launch_bounds__(256) void leaf_boxes_kernel(
    const Tri64 *__restrict__ tri, int64_t n, float *__restrict__ box,
    uint32_t *__restrict__ part) {
  __shared__ uint32_t red[6][256];
  uint32_t cmin[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
  uint32_t cmax[3] = {0u, 0u, 0u};
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const Tri64 r = tri[t];
    for (int d = 0; d < 3; d++) {
      double a = r.v[d], b = r.v[3 + d], c = r.v[6 + d];
      double lo = fmin(a, fmin(b, c)), hi = fmax(a, fmax(b, c));
      box[6 * t + d] = __double2float_rd(lo);
      box[6 * t + 3 + d] = __double2float_ru(hi);
      uint32_t cc = float_order((float)((a + b + c) / 3.0));
      cmin[d] = min(cmin[d], cc);
      cmax[d] = max(cmax[d], cc);
    }
  }
  const int tid = threadIdx.x;
  for (int d = 0; d < 3; d++) {
    red[d][tid] = cmin[d];
    red[3 + d][tid] = cmax[d];
  }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off)
      for (int d = 0; d < 3; d++) {
        red[d][tid] = min(red[d][tid], red[d][tid + off]);
        red[3 + d][tid] = max(red[3 + d][tid], red[3 + d][tid + off]);
      }
    __syncthreads();
  }
  if (tid < 6) part[6 * (int64_t)blockIdx.x + tid] = red[tid][0];
}

// The centroid bounds from the per-block parts (one workgroup).
__global__ __launch_bounds__(256) void bounds_reduce_kernel(const uint32_t *__restrict__ part,
                                                            int64_t nparts,
                                                            uint32_t *__restrict__ cbounds) {
  __shared__ uint32_t red[6][256];
  const int tid = threadIdx.x;
  uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
  for (int64_t b = tid; b < nparts; b += blockDim.x)
    for (int k = 0; k < 6; k++)
      v[k] = k < 3 ? min(v[k], part[6 * b + k]) : max(v[k], part[6 * b + k]);
  for (int k = 0; k < 6; k++) red[k][tid] = v[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off)
      for (int k = 0; k < 6; k++)
        red[k][tid] = k < 3 ? min(red[k][tid], red[k][tid + off])
                            : max(red[k][tid], red[k][tid + off]);
    __syncthreads();
  }
  if (tid < 6) cbounds[tid] = red[tid][0];
}

__device__ inline uint32_t expand10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void morton_kernel(const float *__restrict__ box, int64_t n,
                              const uint32_t *__restrict__ cbounds,
                              uint32_t *__restrict__ keys,
                              uint32_t *__restrict__ vals) {
  float lo[3], ext[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = order_float(cbounds[d]);
    float hi = order_float(cbounds[3 + d]);
    ext[d] = hi > lo[d] ? hi - lo[d] : 1.0f;
  }
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    uint32_t q[3];
    for (int d = 0; d < 3; d++) {
      float c = 0.5f * (box[6 * t + d] + box[6 * t + 3 + d]);
      float x = (c - lo[d]) / ext[d] * 1024.0f;
      x = fminf(fmaxf(x, 0.0f), 1023.0f);
      q[d] = (uint32_t)x;
    }
    keys[t] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
    vals[t] = (uint32_t)t;
  }
}

// Leaf boxes and Morton codes (bits per axis: 8 or 10) in one pass,
// quantised over a box given by the caller (a lattice's box, which covers
// the mesh it was built around; codes of centres outside it are clamped,
// which only coarsens the tree).
__global__ void leaf_morton_kernel(const Tri64 *__restrict__ tri, int64_t n, QBox qb, int bits,
                                   float *__restrict__ box, uint32_t *__restrict__ keys,
                                   uint32_t *__restrict__ vals) {
  const float cells = (float)(1u << bits);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const Tri64 r = tri[t];
    uint32_t q[3];
    for (int d = 0; d < 3; d++) {
      const double a = r.v[d], b = r.v[3 + d], c = r.v[6 + d];
      const float lo = __double2float_rd(fmin(a, fmin(b, c)));
      const float hi = __double2float_ru(fmax(a, fmax(b, c)));
      box[6 * t + d] = lo;
      box[6 * t + 3 + d] = hi;
      float x = (0.5f * (lo + hi) - qb.lo[d]) / qb.ext[d] * cells;
      x = fminf(fmaxf(x, 0.0f), cells - 1.0f);
      q[d] = (uint32_t)x;
    }
    keys[t] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
    vals[t] = (uint32_t)t;
  }
}

// ---- radix sort (LSD, stable) ----
//
// A pass: per-block digit counts (digit-major, hist[d * nblocks + b]), the
// exclusive scan of each digit's row (one workgroup per digit, which also
// writes the row total), then the stable scatter, whose blocks add the
// digit bases (exclusive scan of the 256 totals) themselves.
static_assert(kSortThreads == kRadix, "one thread per digit in the scatter prologue");

// Exclusive prefix sum over the 256 threads of a block; *total gets the
// block's sum.  Contains barriers: every thread of the block calls it.
__device__ inline uint64_t block_scan_256(uint64_t v, uint64_t *total) {
  __shared__ uint64_t s_wsum[kSortThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wsum[w] = x;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kSortThreads / 64; k++) {
    const uint64_t q = s_wsum[k];
    if (k < w) pre += q;
    tot += q;
  }
  __syncthreads();
  *total = tot;
  return pre + x - v;
}

__global__ __launch_bounds__(kSortThreads) void radix_rows_kernel(
    const uint32_t *__restrict__ hist, int64_t nblocks, uint64_t *__restrict__ off,
    uint64_t *__restrict__ total) {
  constexpr int kItems = 8;
  const uint32_t *row = hist + (int64_t)blockIdx.x * nblocks;
  uint64_t *orow = off + (int64_t)blockIdx.x * nblocks;
  uint64_t carry = 0;
  for (int64_t base = 0; base < nblocks; base += (int64_t)kSortThreads * kItems) {
    const int64_t b0 = base + (int64_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    uint64_t sum = 0;
#pragma unroll
    for (int q = 0; q < kItems; q++) {
      v[q] = b0 + q < nblocks ? row[b0 + q] : 0u;
      sum += v[q];
    }
    uint64_t tot;
    uint64_t pre = block_scan_256(sum, &tot) + carry;
#pragma unroll
    for (int q = 0; q < kItems; q++) {
      if (b0 + q < nblocks) orow[b0 + q] = pre;
      pre += v[q];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) total[blockIdx.x] = carry;
}

__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(
    const uint32_t *__restrict__ keys, int64_t n, const uint64_t *__restrict__ n_dev,
    int shift, int64_t nblocks, uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[kRadix];
  if (n_dev && (int64_t)*n_dev < n) n = (int64_t)*n_dev;
  for (int i = threadIdx.x; i < kRadix; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  const int lane = threadIdx.x & 63;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kSortItems; r++) {
    const int64_t i = base + (int64_t)r * kSortThreads + threadIdx.x;
    const bool valid = i < n;
    const uint32_t dig = valid ? (keys[i] >> shift) & (kRadix - 1) : 0u;
    // lanes of the wave with the same digit (keys are often coherent): the
    // first of them adds the group's count, one LDS atomic per distinct digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; b++) {
      const uint64_t bal = __ballot((dig >> b) & 1u);
      peers &= ((dig >> b) & 1u) ? bal : ~bal;
    }
    if (valid && (peers & below) == 0ull) atomicAdd(&h[dig], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += blockDim.x)
    hist[(int64_t)d * nblocks + blockIdx.x] = h[d];
}

// Items are taken round-major (index = base + r*256 + thread), so the
// rank of an item among equal digits is: digits seen in earlier rounds +
// earlier waves of this round + earlier lanes of this wave.
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, int64_t n,
    const uint64_t *__restrict__ n_dev, int shift, int64_t nblocks,
    const uint64_t *__restrict__ off, const uint64_t *__restrict__ total) {
  if (n_dev && (int64_t)*n_dev < n) n = (int64_t)*n_dev;
  constexpr int kWaves = kSortThreads / 64;
  __shared__ uint32_t running[kRadix];
  __shared__ uint32_t wave_cnt[kWaves][kRadix];
  __shared__ uint64_t base_off[kRadix];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  {
    uint64_t all;
    const uint64_t digit_base = block_scan_256(total[tid], &all);
    running[tid] = 0;
    base_off[tid] = digit_base + off[(int64_t)tid * nblocks + blockIdx.x];
  }
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kSortItems; r++) {
    for (int q = tid; q < kWaves * kRadix; q += blockDim.x)
      (&wave_cnt[0][0])[q] = 0;
    __syncthreads();
    const int64_t i = base + (int64_t)r * kSortThreads + tid;
    const bool valid = i < n;
    uint32_t key = valid ? kin[i] : 0u;
    uint32_t val = valid ? vin[i] : 0u;
    uint32_t dig = (key >> shift) & (kRadix - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; b++) {
      uint64_t bal = __ballot((dig >> b) & 1u);
      peers &= ((dig >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t wrank = (uint32_t)__popcll(peers & below);
    if (valid && wrank == 0) wave_cnt[wave][dig] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t rank = running[dig] + wrank;
      for (int w = 0; w < wave; w++) rank += wave_cnt[w][dig];
      const uint64_t pos = base_off[dig] + rank;
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    for (int d = tid; d < kRadix; d += blockDim.x) {
      uint32_t s = 0;
      for (int w = 0; w < kWaves; w++) s += wave_cnt[w][d];
      running[d] += s;
    }
    __syncthreads();
  }
}

// ---- Karras hierarchy ----

// Leaf i's code is codes[i * stride] (stride 1: triangles; kRayLeaf: the
// first triangle of each run of the ray hierarchy).
__device__ inline int delta(const uint32_t *__restrict__ codes, int64_t stride, int64_t n,
                            int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = codes[i * stride], b = codes[j * stride];
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}

__global__ void karras_kernel(const uint32_t *__restrict__ codes, int64_t stride, int64_t n,
                              BvhNode *__restrict__ nodes,
                              int32_t *__restrict__ range) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n - 1;
       i += (int64_t)gridDim.x * blockDim.x) {
    auto dl = [&](int64_t a, int64_t b) { return delta(codes, stride, n, a, b); };
    const int d = (dl(i, i + 1) - dl(i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = dl(i, i - d);
    int64_t lmax = 2;
    while (dl(i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
      if (dl(i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = dl(i, j);
    int64_t s = 0, step = l;
    do {
      step = (step + 1) >> 1;
      if (dl(i, i + (s + step) * d) > dnode) s += step;
    } while (step > 1);
    const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    nodes[i].child[0] = (lo == gamma) ? -(int32_t)(gamma + 1) : (int32_t)gamma;
    nodes[i].child[1] = (hi == gamma + 1) ? -(int32_t)(gamma + 2) : (int32_t)(gamma + 1);
    range[2 * i] = (int32_t)lo;
    range[2 * i + 1] = (int32_t)hi;
  }
}

// Box refit without inter-workgroup synchronisation.  Karras node i covers
// the sorted leaf range [lo_i, hi_i]; its box is the union of the leaf boxes
// of that range, taken from a segment tree over the sorted leaves (level L
// entry k covers leaves [k 2^L, (k+1) 2^L)), O(log n) boxes per node.
// Unions are exact, so the result equals a bottom-up refit.
struct SegLevels {
  int64_t off[64];  // float offset of each level inside the segment array
  int nlev;
};

// Levels lev_in .. lev_in + nlev - 1 (nlev <= 6) of the segment tree in one
// launch, without LDS (so it runs beside kernels that hold the LDS): each
// wavefront takes 64 consecutive entries of level lev_in - 1, and after r
// cross-lane butterfly steps every lane holds the union of its aligned
// block of 2^r entries, i.e. the level lev_in - 1 + r entry, which the
// block's first lane writes.  Entries past a level's end are empty boxes,
// and a union with an empty box returns the other operand exactly, so
// each entry equals the pairwise union of its two children.
// leaf_tri != nullptr: the input is the leaf boxes in sorted order, also
// written as level 0.
constexpr int kSegSpan = 64;
constexpr int kSegLevelsPerLaunch = 6;  // log2(kSegSpan)

__global__ __launch_bounds__(64) void seg_levels_kernel(float *__restrict__ seg, SegLevels L,
                                                        int lev_in, int nlev, int64_t m_in,
                                                        const float *__restrict__ leaf_box,
                                                        const uint32_t *__restrict__ leaf_tri) {
  const int lane = threadIdx.x;
  const int64_t gi = (int64_t)blockIdx.x * kSegSpan + lane;
  float b[6];
  if (gi < m_in) {
    if (leaf_tri) {
      const int64_t t = leaf_tri[gi];
      for (int q = 0; q < 6; q++) b[q] = leaf_box[6 * t + q];
      for (int q = 0; q < 6; q++) seg[6 * gi + q] = b[q];
    } else {
      const float *below = seg + L.off[lev_in - 1];
      for (int q = 0; q < 6; q++) b[q] = below[6 * gi + q];
    }
  } else {
    b[0] = b[1] = b[2] = __builtin_huge_valf();
    b[3] = b[4] = b[5] = -__builtin_huge_valf();
  }
  int64_t size = m_in;
  for (int r = 1; r <= nlev; r++) {
    size = (size + 1) / 2;
    const int w = 1 << (r - 1);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      b[q] = fminf(b[q], __shfl_xor(b[q], w));
      b[3 + q] = fmaxf(b[3 + q], __shfl_xor(b[3 + q], w));
    }
    if ((lane & ((w << 1) - 1)) == 0) {
      const int64_t k = gi >> r;
      if (k < size) {
        float *out = seg + L.off[lev_in - 1 + r];
        for (int q = 0; q < 6; q++) out[6 * k + q] = b[q];
      }
    }
  }
}

__device__ inline void seg_query(const float *__restrict__ seg, const SegLevels &L,
                                 int64_t lo, int64_t hi, float *b) {
  for (int q = 0; q < 3; q++) {
    b[q] = 3.4e38f;
    b[3 + q] = -3.4e38f;
  }
  int64_t l = lo, r = hi + 1;
  for (int lev = 0; l < r && lev < L.nlev; lev++) {
    const float *S = seg + L.off[lev];
    if (l & 1) {
      for (int q = 0; q < 3; q++) {
        b[q] = fminf(b[q], S[6 * l + q]);
        b[3 + q] = fmaxf(b[3 + q], S[6 * l + 3 + q]);
      }
      l++;
    }
    if (r & 1) {
      r--;
      for (int q = 0; q < 3; q++) {
        b[q] = fminf(b[q], S[6 * r + q]);
        b[3 + q] = fmaxf(b[3 + q], S[6 * r + 3 + q]);
      }
    }
    l >>= 1;
    r >>= 1;
  }
}

// Child boxes of every internal node; a leaf child's reference changes from
// its sorted position to its triangle id (traversals skip the leaf_tri load).
__global__ void node_boxes_kernel(const float *__restrict__ seg, SegLevels L,
                                  const int32_t *__restrict__ range, int64_t n,
                                  const uint32_t *__restrict__ leaf_tri,
                                  BvhNode *__restrict__ nodes) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n - 1;
       i += (int64_t)gridDim.x * blockDim.x) {
    for (int c = 0; c < 2; c++) {
      const int32_t ch = nodes[i].child[c];
      float b[6];
      if (ch < 0) {
        const int64_t leaf = -(int64_t)(ch + 1);
        for (int q = 0; q < 6; q++) b[q] = seg[6 * leaf + q];
        nodes[i].child[c] = -(int32_t)leaf_tri[leaf] - 1;
      } else {
        seg_query(seg, L, range[2 * ch], range[2 * ch + 1], b);
      }
      for (int q = 0; q < 6; q++) nodes[i].box[c][q] = b[q];
    }
  }
}

// Child boxes of the ray hierarchy's nodes: a leaf (run c of kRayLeaf
// sorted triangles) is segment-tree level log2(kRayLeaf) entry c; an
// internal child covers the runs [lo, hi], i.e. the sorted leaves
// [kRayLeaf lo, min(kRayLeaf (hi + 1) - 1, n - 1)].
constexpr int kRayLeafLevel = kRayLeaf == 4 ? 2 : (kRayLeaf == 8 ? 3 : (kRayLeaf == 16 ? 4 : -1));
static_assert(kRayLeafLevel > 0, "runs are a segment-tree level");
__global__ void run_node_boxes_kernel(const float *__restrict__ seg, SegLevels L,
                                      const int32_t *__restrict__ range, int64_t nc, int64_t n,
                                      BvhNode *__restrict__ nodes) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nc - 1;
       i += (int64_t)gridDim.x * blockDim.x) {
    for (int c = 0; c < 2; c++) {
      const int32_t ch = nodes[i].child[c];
      float b[6];
      if (ch < 0) {
        const float *lv = seg + L.off[kRayLeafLevel] + 6 * (int64_t)(-(ch + 1));
        for (int q = 0; q < 6; q++) b[q] = lv[q];
      } else {
        const int64_t lo = (int64_t)range[2 * ch] * kRayLeaf;
        const int64_t hi0 = (int64_t)range[2 * ch + 1] * kRayLeaf + kRayLeaf - 1;
        seg_query(seg, L, lo, hi0 < n - 1 ? hi0 : n - 1, b);
      }
      for (int q = 0; q < 6; q++) nodes[i].box[c][q] = b[q];
    }
  }
}

// Wide node i from binary node i: the two children of each internal child
// and each leaf child as it is; unused slots get an empty box.  Four lanes
// per node, one slot each, so a wave writes 16 whole nodes.
__global__ void wide_nodes_kernel(const BvhNode *__restrict__ nodes, int64_t nn,
                                  BvhWide *__restrict__ wide) {
  const int64_t total = 4 * nn;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t >> 2;
    const int slot = (int)(t & 3);
    const int32_t c0 = nodes[i].child[0], c1 = nodes[i].child[1];
    const int n0 = c0 >= 0 ? 2 : 1, n1 = c1 >= 0 ? 2 : 1;
    const float *src = nullptr;
    int32_t ch = kNoChild;
    if (slot < n0) {
      src = c0 >= 0 ? nodes[c0].box[slot] : nodes[i].box[0];
      ch = c0 >= 0 ? nodes[c0].child[slot] : c0;
    } else if (slot < n0 + n1) {
      const int k = slot - n0;
      src = c1 >= 0 ? nodes[c1].box[k] : nodes[i].box[1];
      ch = c1 >= 0 ? nodes[c1].child[k] : c1;
    }
    float *dst = wide[i].box[slot];
    for (int q = 0; q < 3; q++) {
      dst[q] = src ? src[q] : __builtin_huge_valf();
      dst[3 + q] = src ? src[3 + q] : -__builtin_huge_valf();
    }
    wide[i].child[slot] = ch;
    wide[i].pad_[slot] = 0;
  }
}

// Sorted pairs end in (keys, vals), or, when keys_out is given, in the pair
// of buffers the last pass wrote (*keys_out, *vals_out; no copy back).
void radix_sort_pairs(hipStream_t s, Arena &ar, uint32_t *keys, uint32_t *vals,
                      int64_t n, int key_bits, const uint64_t *n_dev,
                      uint32_t **keys_out = nullptr, uint32_t **vals_out = nullptr) {
  if (keys_out) {
    *keys_out = keys;
    *vals_out = vals;
  }
  if (n <= 1) return;
  int64_t nblocks = (n + kSortTile - 1) / kSortTile;
  uint32_t *k2 = ar.get<uint32_t>(ar.tmp("sort_k2"), n);
  uint32_t *v2 = ar.get<uint32_t>(ar.tmp("sort_v2"), n);
  uint32_t *hist = ar.get<uint32_t>(ar.tmp("sort_hist"), kRadix * nblocks + 1);
  uint64_t *off = ar.get<uint64_t>(ar.tmp("sort_off"), kRadix * nblocks + 2);
  uint64_t *total = ar.get<uint64_t>(ar.tmp("sort_total"), kRadix);
  uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
  int passes = 0;
  for (int shift = 0; shift < key_bits; shift += kRadixBits, passes++) {
    hipLaunchKernelGGL(radix_hist_kernel, dim3((unsigned)nblocks),
                       dim3(kSortThreads), 0, s, ka, n, n_dev, shift, nblocks, hist);
    SDF_HIP(hipGetLastError());
    hipLaunchKernelGGL(radix_rows_kernel, dim3(kRadix), dim3(kSortThreads), 0, s,
                       (const uint32_t *)hist, nblocks, off, total);
    hipLaunchKernelGGL(radix_scatter_kernel, dim3((unsigned)nblocks),
                       dim3(kSortThreads), 0, s, ka, va, kb, vb, n, n_dev, shift,
                       nblocks, (const uint64_t *)off, (const uint64_t *)total);
    SDF_HIP(hipGetLastError());
    uint32_t *t = ka;
    ka = kb;
    kb = t;
    t = va;
    va = vb;
    vb = t;
  }
  if (keys_out) {
    *keys_out = ka;
    *vals_out = va;
  } else if (passes & 1) {  // odd pass count: the result sits in the scratch pair
    SDF_HIP(hipMemcpyAsync(keys, ka, (size_t)n * sizeof(uint32_t),
                           hipMemcpyDeviceToDevice, s));
    SDF_HIP(hipMemcpyAsync(vals, va, (size_t)n * sizeof(uint32_t),
                           hipMemcpyDeviceToDevice, s));
  }
}

}  // namespace

void radix_sort_u32_pairs(hipStream_t s, Arena &ar, uint32_t *keys,
                          uint32_t *vals, int64_t n, int key_bits, const uint64_t *n_dev,
                          uint32_t **keys_out, uint32_t **vals_out) {
  radix_sort_pairs(s, ar, keys, vals, n, key_bits, n_dev, keys_out, vals_out);
}

Bvh build_lbvh(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t n, const QBox *qbox,
               bool full) {
  Bvh b;
  b.n = n;
  if (n <= 0) return b;
  b.leaf_box = ar.get<float>("bvh_leaf_box", 6 * n);
  uint32_t *codes = ar.get<uint32_t>("bvh_codes", n);
  uint32_t *leaf_tri = ar.get<uint32_t>("bvh_leaf_tri", n);
  // the ray hierarchy alone groups runs of kRayLeaf triangles, for which
  // 8 bits per axis (3 sort passes) order space finely enough; nearest-point
  // queries use the per-triangle hierarchy, with 10
  const int bits = qbox && !full ? 8 : 10;
  if (qbox) {
    hipLaunchKernelGGL(leaf_morton_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, tri, n,
                       *qbox, bits, b.leaf_box, codes, leaf_tri);
    SDF_HIP(hipGetLastError());
  } else {
    uint32_t *cb = ar.get<uint32_t>("bvh_cbounds", 8);
    const int lb_blocks = blocks_for(n, 256);
    uint32_t *part = ar.get<uint32_t>("bvh_cbounds_part", 6 * (size_t)lb_blocks);
    hipLaunchKernelGGL(leaf_boxes_kernel, dim3(lb_blocks), dim3(256), 0, s, tri, n, b.leaf_box,
                       part);
    hipLaunchKernelGGL(bounds_reduce_kernel, dim3(1), dim3(256), 0, s, (const uint32_t *)part,
                       (int64_t)lb_blocks, cb);
    hipLaunchKernelGGL(morton_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                       b.leaf_box, n, cb, codes, leaf_tri);
    SDF_HIP(hipGetLastError());
  }
  // the sorted pairs stay in whichever buffers the last pass wrote (the
  // arena scope keeps the sort's scratch pair this hierarchy's own)
  radix_sort_pairs(s, ar, codes, leaf_tri, n, 3 * bits, nullptr, &codes, &leaf_tri);
  b.leaf_tri = leaf_tri;
  b.nc = (n + kRayLeaf - 1) / kRayLeaf;
  if (n == 1) return b;
  SegLevels L;
  int64_t total = 0, m = n;
  L.nlev = 0;
  while (true) {
    L.off[L.nlev++] = 6 * total;
    total += m;
    if (m == 1 || L.nlev == 64) break;
    m = (m + 1) / 2;
  }
  float *seg = ar.get<float>("bvh_seg", 6 * total);
  // level 0 (sorted leaf boxes) and the levels above, 6 per launch
  m = n;
  for (int lev = 1; lev < L.nlev;) {
    const int k = L.nlev - lev < kSegLevelsPerLaunch ? L.nlev - lev : kSegLevelsPerLaunch;
    hipLaunchKernelGGL(seg_levels_kernel, dim3((unsigned)((m + kSegSpan - 1) / kSegSpan)),
                       dim3(kSegSpan), 0, s, seg, L, lev, k, m, (const float *)b.leaf_box,
                       lev == 1 ? (const uint32_t *)leaf_tri : (const uint32_t *)nullptr);
    SDF_HIP(hipGetLastError());
    for (int q = 0; q < k; q++) m = (m + 1) / 2;
    lev += k;
  }
  if (b.nc > 1) {
    // the ray hierarchy over runs of kRayLeaf sorted triangles
    b.cnodes = ar.get<BvhNode>("bvh_cnodes", b.nc - 1);
    int32_t *crange = ar.get<int32_t>("bvh_crange", 2 * (b.nc - 1));
    hipLaunchKernelGGL(karras_kernel, dim3(blocks_for(b.nc - 1, 256)), dim3(256), 0, s,
                       (const uint32_t *)codes, (int64_t)kRayLeaf, b.nc, b.cnodes, crange);
    hipLaunchKernelGGL(run_node_boxes_kernel, dim3(blocks_for(b.nc - 1, 256)), dim3(256), 0, s,
                       (const float *)seg, L, (const int32_t *)crange, b.nc, n, b.cnodes);
    b.cwide = ar.get<BvhWide>("bvh_cwide", b.nc - 1);
    hipLaunchKernelGGL(wide_nodes_kernel, dim3(blocks_for(4 * (b.nc - 1), 256)), dim3(256), 0,
                       s, (const BvhNode *)b.cnodes, b.nc - 1, b.cwide);
    SDF_HIP(hipGetLastError());
  }
  if (full) {
    b.nodes = ar.get<BvhNode>("bvh_nodes", n - 1);
    int32_t *range = ar.get<int32_t>("bvh_range", 2 * (n - 1));
    hipLaunchKernelGGL(karras_kernel, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s,
                       (const uint32_t *)codes, (int64_t)1, n, b.nodes, range);
    hipLaunchKernelGGL(node_boxes_kernel, dim3(blocks_for(n - 1, 256)), dim3(256),
                       0, s, seg, L, range, n, (const uint32_t *)leaf_tri, b.nodes);
    SDF_HIP(hipGetLastError());
  }
  return b;
}

}  // namespace sdf
