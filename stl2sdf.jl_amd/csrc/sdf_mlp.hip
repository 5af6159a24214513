// Learned search-radius MLP on the CDNA4 matrix cores (BF16 operands, FP32
// accumulation): operand packing, inference over tile points, and the
// training step (forward + backward, weight-gradient GEMM, Adam).  The
// network and parameter layout are described in sdf_mlp.h.
//
// Matrix-core layout (v_mfma_f32_16x16x32_bf16, one 64-lane wavefront):
//   A (16 x 32): lane l holds row l%16, k = 8*(l/16) + j, j < 8
//   B (32 x 16): lane l holds column l%16, k = 8*(l/16) + j
//   D (16 x 16): lane l holds column l%16, rows 4*(l/16) + i, i < 4
// Layers compute Z^T = W . H^T with neurons as rows and 16 points as
// columns.  The contraction index of every layer is permuted,
//   k = kperm(s, q, j) = 32 s + 16 (j / 4) + 4 q + j % 4   (k-step s, q = l/16),
// so that the D fragments of row tiles 2s and 2s+1 are, lane for lane, the B
// fragment of k-step s of the next layer: activations never leave registers.
// The weights are stored pre-permuted to match.
#include <hip/hip_runtime.h>

#include <cmath>

#include "sdf_mlp.h"

namespace sdf {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

inline int blocks_for(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > 65535 * 16) b = 65535 * 16;
  return (int)b;
}

// round to nearest even; NaN stays NaN (quietened), infinities stay; no branches
__host__ __device__ inline uint16_t bf16_bits(float x) {
  const uint32_t b = __builtin_bit_cast(uint32_t, x);
  const uint32_t rounded = b + 0x7FFFu + ((b >> 16) & 1u);
  const uint32_t special = (b & 0x007FFFFFu) ? (b | 0x00400000u) : b;
  return (uint16_t)((((b & 0x7F800000u) == 0x7F800000u) ? special : rounded) >> 16);
}
// A zero the compiler cannot see through: indexing the weights with it keeps
// their loads inside the per-tile loop (hoisted, they would need 384 VGPRs).
__device__ inline int opaque_zero() {
  int z;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z));
  return z;
}
__device__ inline float bf16_float(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}
__device__ inline float bf16_round(float x) { return bf16_float(bf16_bits(x)); }

__device__ inline bf16x8 as_bf16x8(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }

// 2^(0.4 k) rounded to FP32
__constant__ float kFreq[16] = {
    0x1.000000p+0f, 0x1.51cb46p+0f, 0x1.bdb8cep+0f, 0x1.261118p+1f,
    0x1.840600p+1f, 0x1.000000p+2f, 0x1.51cb46p+2f, 0x1.bdb8cep+2f,
    0x1.261118p+3f, 0x1.840600p+3f, 0x1.000000p+4f, 0x1.51cb46p+4f,
    0x1.bdb8cep+4f, 0x1.261118p+5f, 0x1.840600p+5f, 0x1.000000p+6f};

// B fragments (4 k-steps) of the features of this lane's point.  Slots
// (2m, 2m+1) of a fragment are the consecutive features (f, f+1), which for
// f >= 4 are the sine and cosine of one argument.
__device__ inline void mlp_features(const float u[3], int q, u16x8 B[4]) {
#pragma unroll
  for (int s = 0; s < 4; s++) {
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const int f = 32 * s + 16 * (m >> 1) + 4 * q + 2 * (m & 1);
      const int gg = f < 4 ? 0 : (f < 100 ? f - 4 : 0);
      const int k = gg / 6, a = (gg % 6) >> 1;
      const float ua = a == 0 ? u[0] : (a == 1 ? u[1] : u[2]);
      const float r = (kFreq[k] * ua) * 0.5f;  // revolutions: sin(pi f u) = sin(2 pi r)
      const float sn = __builtin_amdgcn_sinf(r), cs = __builtin_amdgcn_cosf(r);
      const bool four = f >= 4 && f < 100;
      const float v0 = four ? sn : (f == 0 ? u[0] : (f == 2 ? u[2] : 0.f));
      const float v1 = four ? cs : (f == 0 ? u[1] : 0.f);
      B[s][2 * m] = bf16_bits(v0);
      B[s][2 * m + 1] = bf16_bits(v1);
    }
  }
}

__device__ inline f32x4 mfma(u16x8 a, u16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a), as_bf16x8(b), c, 0, 0, 0);
}

// One layer for NT point tiles: acc[t][nt] = W . B[t] + bias (rows nt*16+4q+i).
// A is the packed layer (lane-indexed, 8 row tiles x 4 k-steps).
template <int NT, typename LoadA>
__device__ inline void layer_fwd(LoadA A, const float *bias, int lane,
                                 const u16x8 (&B)[NT][4], f32x4 (&acc)[NT][8]) {
  const int q = lane >> 4;
#pragma unroll
  for (int nt = 0; nt < 8; nt++) {
    f32x4 b0;
#pragma unroll
    for (int i = 0; i < 4; i++) b0[i] = bias[nt * 16 + 4 * q + i];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t][nt] = b0;
  }
#pragma unroll
  for (int s = 0; s < 4; s++) {
#pragma unroll
    for (int nt = 0; nt < 8; nt++) {
      const u16x8 a = A((nt * 4 + s) * 64 + lane);
#pragma unroll
      for (int t = 0; t < NT; t++) acc[t][nt] = mfma(a, B[t][s], acc[t][nt]);
    }
    // keep the next k-step's operand loads from being hoisted (register pressure)
    __builtin_amdgcn_sched_barrier(0);
  }
}

// ReLU, round to BF16, and reinterpret D fragments as next-layer B fragments.
template <int NT>
__device__ inline void relu_pack(const f32x4 (&acc)[NT][8], u16x8 (&B)[NT][4]) {
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        B[t][s][i] = bf16_bits(fmaxf(acc[t][2 * s][i], 0.f));
        B[t][s][4 + i] = bf16_bits(fmaxf(acc[t][2 * s + 1][i], 0.f));
      }
}

// Output layer: y for the point of this lane's column (all four lane groups
// return the same value).  Fixed summation order: per lane over nt, i; then
// the xor-16 and xor-32 partners.
template <int NT>
__device__ inline void out_layer(const f32x4 (&acc)[NT][8], const float *w4b, float b4,
                                 int lane, float (&y)[NT]) {
  const int q = lane >> 4;
#pragma unroll
  for (int t = 0; t < NT; t++) {
    float s = 0.f;
#pragma unroll
    for (int nt = 0; nt < 8; nt++)
#pragma unroll
      for (int i = 0; i < 4; i++)
        s += w4b[nt * 16 + 4 * q + i] * bf16_round(fmaxf(acc[t][nt][i], 0.f));
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    y[t] = s + b4;
  }
}

// ---------------------------------------------------------------- packing
__global__ void mlp_pack_kernel(const float *__restrict__ params, uint16_t *__restrict__ afwd,
                                uint16_t *__restrict__ abwd, float *__restrict__ w4b) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < 3 * 16384) {
    const int L = (int)(e >> 14), n = (int)((e >> 7) & 127), k = (int)(e & 127);
    const int off = L == 0 ? kOffW1 : (L == 1 ? kOffW2 : kOffW3);
    const uint16_t b = bf16_bits(params[off + n * 128 + k]);
    {  // forward: row n, contraction slot of k
      const int s = k >> 5, r = k & 31, j = ((r >> 4) << 2) | (r & 3), q = (r >> 2) & 3;
      const int lane = q * 16 + (n & 15);
      afwd[(int64_t)L * kMlpPacked + (((n >> 4) * 4 + s) * 64 + lane) * 8 + j] = b;
    }
    if (L > 0) {  // backward: row k, contraction slot of n
      const int s = n >> 5, r = n & 31, j = ((r >> 4) << 2) | (r & 3), q = (r >> 2) & 3;
      const int lane = q * 16 + (k & 15);
      abwd[(int64_t)(L - 1) * kMlpPacked + (((k >> 4) * 4 + s) * 64 + lane) * 8 + j] = b;
    }
  } else if (e < 3 * 16384 + 128) {
    const int k = (int)(e - 3 * 16384);
    w4b[k] = bf16_float(bf16_bits(params[kOffW4 + k]));
  }
}

// ---------------------------------------------------------------- inference
// Generic points: u as FP32 triples; one wave = 16 points per pass.
__global__ __launch_bounds__(256) void mlp_forward_kernel(
    const uint16_t *__restrict__ afwd, const float *__restrict__ params,
    const float *__restrict__ w4b, const float *__restrict__ u, int stride, int64_t n,
    float *__restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t base = wave * 16; base < n; base += nwaves * 16) {
    const u16x8 *A = reinterpret_cast<const u16x8 *>(afwd) + opaque_zero();
    const int64_t p = base + (lane & 15);
    float uu[3] = {0.f, 0.f, 0.f};
    if (p < n) {
      uu[0] = u[stride * p];
      uu[1] = u[stride * p + 1];
      uu[2] = u[stride * p + 2];
    }
    u16x8 B[1][4];
    mlp_features(uu, lane >> 4, B[0]);
    f32x4 acc[1][8];
    layer_fwd<1>([&](int i) { return A[i]; }, params + kOffB1, lane, B, acc);
    relu_pack<1>(acc, B);
    layer_fwd<1>([&](int i) { return A[kMlpPacked / 8 + i]; }, params + kOffB2, lane, B, acc);
    relu_pack<1>(acc, B);
    layer_fwd<1>([&](int i) { return A[2 * kMlpPacked / 8 + i]; }, params + kOffB3, lane, B,
                 acc);
    float yy[1];
    out_layer<1>(acc, w4b, params[kOffB4], lane, yy);
    if (lane < 16 && p < n) y[p] = yy[0];
  }
}

// Tile points: each wave takes the 64 points of one tile that has distance
// work (tile lane = 16 t + l%16, t < 4), NT point tiles of 16 per pass, with
// the packed weights of all three layers in LDS (96 KB).
template <int NT, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void mlp_tile_radius_kernel(
    GridP g, TileDims td, MlpNorm nm, const uint16_t *__restrict__ afwd,
    const float *__restrict__ params, const float *__restrict__ w4b,
    const uint64_t *__restrict__ chunk_off, const uint32_t *__restrict__ chunk_tile,
    int64_t nchunks, const uint64_t *__restrict__ nchunks_dev, float scale, float delta,
    float h, float *__restrict__ r2out) {
  __shared__ u16x8 sA[3 * kMlpPacked / 8];
  __shared__ float sb[3 * 128 + 128 + 1];
  for (int i = threadIdx.x; i < 3 * kMlpPacked / 8; i += blockDim.x)
    sA[i] = reinterpret_cast<const u16x8 *>(afwd)[i];
  for (int i = threadIdx.x; i < 128; i += blockDim.x) {
    sb[i] = params[kOffB1 + i];
    sb[128 + i] = params[kOffB2 + i];
    sb[256 + i] = params[kOffB3 + i];
    sb[384 + i] = w4b[i];
  }
  if (threadIdx.x == 0) sb[512] = params[kOffB4];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4;
  const int64_t nw = (int64_t)gridDim.x * WAVES;
  nchunks = chunk_count(nchunks, nchunks_dev);
  for (int64_t item = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6); item < nchunks;
       item += nw) {
    const uint32_t tile = chunk_tile[item];
    const uint64_t first = chunk_off[tile];
    if ((uint64_t)item != first) continue;  // one pass per tile
    int32_t tx, ty, tz;
    tile_coords((uint32_t)tile, td, &tx, &ty, &tz);
#pragma unroll 1
    for (int pass = 0; pass < 4 / NT; pass++) {
      const u16x8 *W = sA + opaque_zero();
      u16x8 B[NT][4];
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const int L = 16 * (pass * NT + t) + (lane & 15);
        const Vec3d p = lattice_point(g, tx * kTileX + (L & 7), ty * kTileY + ((L >> 3) & 3),
                                      tz * kTileZ + (L >> 5));
        const float uu[3] = {(float)((p.x - nm.c[0]) * nm.s), (float)((p.y - nm.c[1]) * nm.s),
                             (float)((p.z - nm.c[2]) * nm.s)};
        mlp_features(uu, q, B[t]);
      }
      f32x4 acc[NT][8];
      layer_fwd<NT>([&](int i) { return W[i]; }, sb, lane, B, acc);
      relu_pack<NT>(acc, B);
      layer_fwd<NT>([&](int i) { return W[kMlpPacked / 8 + i]; }, sb + 128, lane, B, acc);
      relu_pack<NT>(acc, B);
      layer_fwd<NT>([&](int i) { return W[2 * kMlpPacked / 8 + i]; }, sb + 256, lane, B, acc);
      float y[NT];
      out_layer<NT>(acc, sb + 384, sb[512], lane, y);
#pragma unroll
      for (int t = 0; t < NT; t++) {
        if (q == t) {  // lane group t writes point tile pass*NT + t
          // the prediction of a distance, clamped at 0, plus the calibration
          const float r = fminf(fmaxf((fmaxf(y[t], 0.f) + delta) * scale, 0.01f * h), h);
          r2out[first * kTilePts + 16 * (pass * NT + t) + (lane & 15)] = r * r;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- training
// Forward + backward of one batch: each wave takes 16 points.  Writes the
// BF16 layer inputs (phi, h1, h2) and pre-activation gradients (dz1..dz3)
// as [row][point] arrays for the weight-gradient GEMM, and per-block sums of
// the bias / output-layer gradients and of the squared error.
__global__ __launch_bounds__(256) void mlp_train_kernel(
    const uint16_t *__restrict__ afwd, const uint16_t *__restrict__ abwd,
    const float *__restrict__ params, const float *__restrict__ w4b,
    const float4 *__restrict__ samples, const uint32_t *__restrict__ idx, int64_t n,
    int64_t P, uint16_t *__restrict__ acts, float *__restrict__ bpart) {
  __shared__ float red[4][kMlpBPart];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, q = lane >> 4;
  const int64_t col = (int64_t)blockIdx.x * 64 + w * 16 + (lane & 15);
  const bool valid = col < n;
  float4 smp = make_float4(0.f, 0.f, 0.f, 0.f);
  if (valid) smp = samples[idx ? (int64_t)idx[col] : col];
  const float uu[3] = {smp.x, smp.y, smp.z};
  const u16x8 *AF = reinterpret_cast<const u16x8 *>(afwd);
  const u16x8 *AB = reinterpret_cast<const u16x8 *>(abwd);
  uint16_t *PHI = acts, *H1 = acts + 128 * P, *H2 = acts + 2 * 128 * P;
  uint16_t *DZ1 = acts + 3 * 128 * P, *DZ2 = acts + 4 * 128 * P, *DZ3 = acts + 5 * 128 * P;
  auto store_b = [&](uint16_t *dst, const u16x8(&Bf)[1][4]) {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int row = 32 * s + 16 * (j >> 2) + 4 * q + (j & 3);
        dst[(int64_t)row * P + col] = Bf[0][s][j];
      }
  };

  u16x8 B0[1][4], B1[1][4], B2[1][4];
  mlp_features(uu, q, B0[0]);
  f32x4 z1[1][8], z2[1][8], z3[1][8];
  layer_fwd<1>([&](int i) { return AF[i]; }, params + kOffB1, lane, B0, z1);
  relu_pack<1>(z1, B1);
  layer_fwd<1>([&](int i) { return AF[kMlpPacked / 8 + i]; }, params + kOffB2, lane, B1, z2);
  relu_pack<1>(z2, B2);
  layer_fwd<1>([&](int i) { return AF[2 * kMlpPacked / 8 + i]; }, params + kOffB3, lane, B2,
               z3);
  float y[1];
  out_layer<1>(z3, w4b, params[kOffB4], lane, y);
  const float r = valid ? y[0] - smp.w : 0.f;
  const float dy = valid ? 2.0f * r / (float)n : 0.f;
  store_b(PHI, B0);
  store_b(H1, B1);
  store_b(H2, B2);

  // per-lane partial sums: rows nt*16 + 4q + i of this lane's point
  float pdb3[8][4], pdb2[8][4], pdb1[8][4], pdw4[8][4];
  f32x4 g[1][8];
  u16x8 DB[1][4];
#pragma unroll
  for (int nt = 0; nt < 8; nt++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = nt * 16 + 4 * q + i;
      pdw4[nt][i] = dy * bf16_round(fmaxf(z3[0][nt][i], 0.f));
      const float d = z3[0][nt][i] > 0.f ? w4b[row] * dy : 0.f;
      pdb3[nt][i] = d;
      g[0][nt][i] = d;
    }
  auto pack_grad = [&](const f32x4(&gg)[1][8], u16x8(&Bf)[1][4]) {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        Bf[0][s][i] = bf16_bits(gg[0][2 * s][i]);
        Bf[0][s][4 + i] = bf16_bits(gg[0][2 * s + 1][i]);
      }
  };
  pack_grad(g, DB);
  store_b(DZ3, DB);
  // dh2 = W3^T dz3
  {
    f32x4 acc[1][8];
#pragma unroll
    for (int kt = 0; kt < 8; kt++) acc[0][kt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int kt = 0; kt < 8; kt++)
        acc[0][kt] = mfma(AB[kMlpPacked / 8 + (kt * 4 + s) * 64 + lane], DB[0][s], acc[0][kt]);
#pragma unroll
    for (int nt = 0; nt < 8; nt++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float d = z2[0][nt][i] > 0.f ? acc[0][nt][i] : 0.f;
        pdb2[nt][i] = d;
        g[0][nt][i] = d;
      }
  }
  pack_grad(g, DB);
  store_b(DZ2, DB);
  // dh1 = W2^T dz2
  {
    f32x4 acc[1][8];
#pragma unroll
    for (int kt = 0; kt < 8; kt++) acc[0][kt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int kt = 0; kt < 8; kt++)
        acc[0][kt] = mfma(AB[(kt * 4 + s) * 64 + lane], DB[0][s], acc[0][kt]);
#pragma unroll
    for (int nt = 0; nt < 8; nt++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float d = z1[0][nt][i] > 0.f ? acc[0][nt][i] : 0.f;
        pdb1[nt][i] = d;
        g[0][nt][i] = d;
      }
  }
  pack_grad(g, DB);
  store_b(DZ1, DB);

  // reduce over the 16 points of the wave (lanes with equal q), then waves
  auto red16 = [&](float v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
  };
#pragma unroll
  for (int nt = 0; nt < 8; nt++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = nt * 16 + 4 * q + i;
      const float a1 = red16(pdb1[nt][i]), a2 = red16(pdb2[nt][i]);
      const float a3 = red16(pdb3[nt][i]), a4 = red16(pdw4[nt][i]);
      if ((lane & 15) == 0) {
        red[w][row] = a1;
        red[w][128 + row] = a2;
        red[w][256 + row] = a3;
        red[w][384 + row] = a4;
      }
    }
  {
    float e = red16(dy), l2 = red16(r * r);
    if (lane == 0) {
      red[w][512] = e;
      red[w][513] = l2;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kMlpBPart; i += blockDim.x)
    bpart[(int64_t)blockIdx.x * kMlpBPart + i] =
        (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
}

// dW_L = DZ_L . Hin_L^T over one slice of the batch (K = points).  Block =
// 4 waves; wave w computes row tiles 2w, 2w+1 x all 8 column tiles.
__global__ __launch_bounds__(256) void mlp_wgrad_kernel(const uint16_t *__restrict__ acts,
                                                        int64_t P,
                                                        float *__restrict__ wpart) {
  const int slice = blockIdx.x, L = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, q = lane >> 4;
  const uint16_t *Hin = acts + (int64_t)L * 128 * P;        // phi, h1, h2
  const uint16_t *DZ = acts + (int64_t)(3 + L) * 128 * P;   // dz1, dz2, dz3
  f32x4 acc[2][8];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int kt = 0; kt < 8; kt++) acc[a][kt] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int64_t p_begin = (int64_t)slice * kMlpSlice;
  const int64_t p_end = p_begin + kMlpSlice < P ? p_begin + kMlpSlice : P;
  for (int64_t p0 = p_begin; p0 < p_end; p0 += 32) {
    u16x8 a[2], b[8];
#pragma unroll
    for (int t = 0; t < 2; t++)
      a[t] = *reinterpret_cast<const u16x8 *>(DZ + (int64_t)((2 * w + t) * 16 + (lane & 15)) * P +
                                             p0 + 8 * q);
#pragma unroll
    for (int kt = 0; kt < 8; kt++)
      b[kt] = *reinterpret_cast<const u16x8 *>(Hin + (int64_t)(kt * 16 + (lane & 15)) * P + p0 +
                                              8 * q);
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int kt = 0; kt < 8; kt++) acc[t][kt] = mfma(a[t], b[kt], acc[t][kt]);
  }
  float *out = wpart + ((int64_t)slice * 3 + L) * 16384;
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int kt = 0; kt < 8; kt++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int nrow = (2 * w + t) * 16 + 4 * q + i, kcol = kt * 16 + (lane & 15);
        out[nrow * 128 + kcol] = acc[t][kt][i];
      }
}

// Sum the partials into the gradient (fixed order), optionally store it,
// and optionally take one Adam step (FP32, no contraction, bias corrections
// c1 = 1 - 0.9^t and c2 = 1 - 0.999^t from the host).
__global__ void mlp_reduce_adam_kernel(float *__restrict__ params, float *__restrict__ m,
                                       float *__restrict__ v, const float *__restrict__ wpart,
                                       int nslices, const float *__restrict__ bpart,
                                       int nblocks, float *__restrict__ grad_out,
                                       float *__restrict__ loss_out, float inv_n, float lr,
                                       float c1, float c2) {
#pragma clang fp contract(off)
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i == 0 && loss_out) {
    float l = 0.f;
    for (int b = 0; b < nblocks; b++) l += bpart[(int64_t)b * kMlpBPart + 513];
    *loss_out = l * inv_n;
  }
  if (i >= kMlpParams) return;
  float gsum = 0.f;
  int L = -1, e = 0;
  if (i < kOffB1) {
    L = 0;
    e = (int)i;
  } else if (i >= kOffW2 && i < kOffB2) {
    L = 1;
    e = (int)(i - kOffW2);
  } else if (i >= kOffW3 && i < kOffB3) {
    L = 2;
    e = (int)(i - kOffW3);
  }
  if (L >= 0) {
    for (int sl = 0; sl < nslices; sl++) gsum += wpart[((int64_t)sl * 3 + L) * 16384 + e];
  } else {
    int pi;
    if (i < kOffW2)
      pi = (int)(i - kOffB1);
    else if (i < kOffW3)
      pi = 128 + (int)(i - kOffB2);
    else if (i < kOffW4)
      pi = 256 + (int)(i - kOffB3);
    else if (i < kOffB4)
      pi = 384 + (int)(i - kOffW4);
    else
      pi = 512;
    for (int b = 0; b < nblocks; b++) gsum += bpart[(int64_t)b * kMlpBPart + pi];
  }
  if (grad_out) grad_out[i] = gsum;
  if (lr > 0.f) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mi = b1 * m[i] + (1.0f - b1) * gsum;
    const float vi = b2 * v[i] + (1.0f - b2) * (gsum * gsum);
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    params[i] = params[i] - lr * mh / (sqrtf(vh) + eps);
  }
}


// ---------------------------------------------------------------- training set
__device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// counts[0] += points with d < hc (the band), counts[1] += the others
__global__ void fit_count_kernel(const double *__restrict__ dc, int64_t n, double hc,
                                 unsigned long long *__restrict__ counts) {
  __shared__ unsigned int sb[2];
  if (threadIdx.x < 2) sb[threadIdx.x] = 0;
  __syncthreads();
  unsigned int nb = 0, nf = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (dc[i] < hc)
      nb++;
    else
      nf++;
  }
  atomicAdd(&sb[0], nb);
  atomicAdd(&sb[1], nf);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(counts, (unsigned long long)sb[0]);
    atomicAdd(counts + 1, (unsigned long long)sb[1]);
  }
}

// keep[i]: band points, and far points whose hash falls below far_thr
__global__ void fit_flag_kernel(const double *__restrict__ dc, int64_t n, double hc,
                                uint32_t far_thr, uint32_t *__restrict__ keep) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    keep[i] = (dc[i] < hc || mix32((uint32_t)i) < far_thr) ? 1u : 0u;
}

__global__ void fit_write_kernel(GridP g, MlpNorm nm, const double *__restrict__ dc, int64_t n,
                                 double hc, const uint32_t *__restrict__ keep,
                                 const uint64_t *__restrict__ off, float4 *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (!keep[i]) continue;
    const int64_t ix = i % g.nx, iy = (i / g.nx) % g.ny, iz = i / ((int64_t)g.nx * g.ny);
    const Vec3d p = lattice_point(g, ix, iy, iz);
    const double t = dc[i] < hc ? dc[i] / hc : 1.0;
    out[off[i]] = make_float4((float)((p.x - nm.c[0]) * nm.s), (float)((p.y - nm.c[1]) * nm.s),
                              (float)((p.z - nm.c[2]) * nm.s), (float)t);
  }
}
}  // namespace

void mlp_alloc(Arena &ar, MlpDev &m) {
  m.params = ar.get<float>("mlp_params", kMlpParams);
  m.m = ar.get<float>("mlp_m", kMlpParams);
  m.v = ar.get<float>("mlp_v", kMlpParams);
  m.afwd = ar.get<uint16_t>("mlp_afwd", 3 * kMlpPacked);
  m.abwd = ar.get<uint16_t>("mlp_abwd", 2 * kMlpPacked);
  m.w4b = ar.get<float>("mlp_w4b", 128);
}

void mlp_pack(hipStream_t s, MlpDev &m) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3(blocks_for(3 * 16384 + 128, 256)), dim3(256), 0, s,
                     (const float *)m.params, m.afwd, m.abwd, m.w4b);
  SDF_HIP(hipGetLastError());
}

void mlp_forward(hipStream_t s, const MlpDev &m, const float *u, int stride, int64_t n,
                 float *y) {
  if (n <= 0) return;
  int64_t waves = (n + 15) / 16;
  int blocks = blocks_for(waves * 64, 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(256), 0, s,
                     (const uint16_t *)m.afwd, (const float *)m.params, (const float *)m.w4b,
                     u, stride, n, y);
  SDF_HIP(hipGetLastError());
}

void mlp_step(hipStream_t s, Arena &ar, MlpDev &m, const float4 *samples,
              const uint32_t *idx, int64_t n, float *grad, float *loss, float lr) {
  if (n <= 0) throw Error{-1, "empty training batch"};
  const int64_t P = (n + 63) / 64 * 64;
  const int nblocks = (int)(P / 64);
  const int nslices = (int)((P + kMlpSlice - 1) / kMlpSlice);
  uint16_t *acts = ar.get<uint16_t>("mlp_acts", (size_t)6 * 128 * P);
  float *bpart = ar.get<float>("mlp_bpart", (size_t)nblocks * kMlpBPart);
  float *wpart = ar.get<float>("mlp_wpart", (size_t)nslices * 3 * 16384);
  hipLaunchKernelGGL(mlp_train_kernel, dim3(nblocks), dim3(256), 0, s,
                     (const uint16_t *)m.afwd, (const uint16_t *)m.abwd,
                     (const float *)m.params, (const float *)m.w4b, samples, idx, n, P, acts,
                     bpart);
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(nslices, 3), dim3(256), 0, s,
                     (const uint16_t *)acts, P, wpart);
  float c1 = 1.f, c2 = 1.f;
  if (lr > 0.f) {
    const int t = m.step + 1;
    c1 = (float)(1.0 - std::pow(0.9, t));
    c2 = (float)(1.0 - std::pow(0.999, t));
  }
  hipLaunchKernelGGL(mlp_reduce_adam_kernel, dim3(blocks_for(kMlpParams, 256)), dim3(256), 0,
                     s, m.params, m.m, m.v, (const float *)wpart, nslices,
                     (const float *)bpart, nblocks, grad, loss, 1.0f / (float)n, lr, c1, c2);
  SDF_HIP(hipGetLastError());
  if (lr > 0.f) {
    m.step++;
    mlp_pack(s, m);
  }
}

void mlp_tile_radius(hipStream_t s, const MlpDev &m, const MlpNorm &nm, const GridP &g,
                     const Bins &b, float scale, float delta, float *r2) {
  if (b.nchunks <= 0) return;
  const TileDims td = tile_dims(g);
  constexpr int kNT = 2, kWaves = 8;
  // persistent: one block per CU, each loading the 96 KB of weights once
  int blocks = (int)((b.nchunks + kWaves - 1) / kWaves);
  if (blocks > 256) blocks = 256;
  hipLaunchKernelGGL((mlp_tile_radius_kernel<kNT, kWaves>), dim3(blocks), dim3(kWaves * 64), 0,
                     s, g, td, nm, (const uint16_t *)m.afwd, (const float *)m.params,
                     (const float *)m.w4b, (const uint64_t *)b.chunk_off,
                     (const uint32_t *)b.chunk_tile, b.nchunks, b.nchunks_dev, scale, delta,
                     (float)g.h, r2);
  SDF_HIP(hipGetLastError());
}

}  // namespace sdf

namespace sdf {

int64_t mlp_fit_samples(hipStream_t s, Arena &ar, const GridP &gc, const MlpNorm &nm,
                        const double *dc, double hc, float4 **out, int64_t *nband) {
  const int64_t n = (int64_t)gc.nx * gc.ny * gc.nz;
  unsigned long long *cnt = ar.get<unsigned long long>("fit_cnt", 2);
  launch_fill_u64(s, (uint64_t *)cnt, 2, 0ull);
  int blocks = blocks_for(n, 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fit_count_kernel, dim3(blocks), dim3(256), 0, s, dc, n, hc, cnt);
  unsigned long long h[2] = {0, 0};
  SDF_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
  SDF_HIP(hipStreamSynchronize(s));
  *nband = (int64_t)h[0];
  // as many far points as band points (all of them when fewer)
  double frac = h[1] > 0 ? (double)h[0] / (double)h[1] : 0.0;
  if (frac > 1.0) frac = 1.0;
  const uint32_t thr = frac >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(frac * 4294967296.0);
  uint32_t *keep = ar.get<uint32_t>("fit_keep", n + 1);
  uint64_t *off = ar.get<uint64_t>("fit_off", n + 2);
  hipLaunchKernelGGL(fit_flag_kernel, dim3(blocks), dim3(256), 0, s, dc, n, hc, thr, keep);
  SDF_HIP(hipGetLastError());
  scan_counts(s, ar, keep, off, n);
  uint64_t total = 0;
  SDF_HIP(hipMemcpyAsync(&total, off + n, sizeof(total), hipMemcpyDeviceToHost, s));
  SDF_HIP(hipStreamSynchronize(s));
  float4 *smp = ar.get<float4>("fit_samples", total + 1);
  hipLaunchKernelGGL(fit_write_kernel, dim3(blocks), dim3(256), 0, s, gc, nm, dc, n, hc,
                     (const uint32_t *)keep, (const uint64_t *)off, smp);
  SDF_HIP(hipGetLastError());
  *out = smp;
  return (int64_t)total;
}

}  // namespace sdf
