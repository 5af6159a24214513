// RBF smoothing on the lattice (src/SdfSmoothing/RBFs4Smoothing.jl:275-320).
//
// The reference assembles K with a KD-tree over the coarse points
// (compute_sparse_kernel_matrix, :142-176): K_ij = exp(-(r/h)^2) for pairs
// within r <= h sqrt(-ln threshold) whose value exceeds the threshold.  On a
// uniform lattice that is a translation-invariant stencil (81 taps for the
// default 1e-3) cut at the grid boundary, applied here matrix-free.
//
// Weights (compute_rbf_weights, :191-202): the reference runs IterativeSolvers
// cg in Float32 from x0 = 0 until ||r|| <= reltol ||b||.  Here: FP32 CG on the
// correction equation inside an FP64 iterative refinement whose residual
// b - K x is formed in FP64; the stopping test uses that true residual.
//
// Fine evaluation (rbf_interpolation_kdtree, :219-253): every coarse point
// within the truncation radius of a fine point contributes
// w_j exp(-(d/h)^2); fine points fall in s^3 classes by their offset inside
// a coarse cell, each class with its own precomputed tap list.
#include <math.h>

#include <vector>

#include "sdf_internal.h"

namespace sdf {

namespace {

inline int blocks_for(int64_t n, int t) {
  int64_t b = (n + t - 1) / t;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (int)b;
}

constexpr int kRedBlocks = 1024;

struct Tap {
  int32_t di, dj, dk;
  float w;
};

// process_vector (:15-22): Float32 values; entries within isapprox of 1e10
// become sign(x) * max|x| over |x| < 1e9.
__global__ void to_f32_kernel(const double *__restrict__ d, int64_t n,
                              float *__restrict__ f,
                              unsigned int *__restrict__ maxbits,
                              unsigned int *__restrict__ nbig) {
  unsigned int m = 0, big = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float x = (float)d[i];
    f[i] = x;
    float ax = fabsf(x);
    if (ax < 1.0e9f) m = max(m, __float_as_uint(ax));
    if (fabsf(ax - 1.0e10f) <= 3.4526698e-4f * 1.0e10f) big++;
  }
  atomicMax(maxbits, m);
  if (big) atomicAdd(nbig, big);
}

__global__ void fix_big_kernel(float *__restrict__ f, int64_t n,
                               const unsigned int *__restrict__ maxbits) {
  const float mv = __uint_as_float(*maxbits);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float ax = fabsf(f[i]);
    if (fabsf(ax - 1.0e10f) <= 3.4526698e-4f * 1.0e10f) f[i] = copysignf(mv, f[i]);
  }
}

// y = K x over the lattice (zero outside); optional per-block partial of
// dot(x, y) into part[blockIdx].
template <typename T>
__global__ __launch_bounds__(256) void stencil_kernel(
    GridP g, const Tap *__restrict__ taps, int ntaps, const T *__restrict__ x,
    T *__restrict__ y, double *__restrict__ part) {
  __shared__ double red[256];
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  double acc_dot = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = (int32_t)(i % g.nx), b = (int32_t)((i / g.nx) % g.ny),
                  c = (int32_t)(i / nxy);
    T s = T(0);
    for (int t = 0; t < ntaps; t++) {
      const Tap tp = taps[t];
      const int32_t aa = a + tp.di, bb = b + tp.dj, cc = c + tp.dk;
      if (aa < 0 || bb < 0 || cc < 0 || aa >= g.nx || bb >= g.ny || cc >= g.nz)
        continue;
      s += (T)tp.w * x[(int64_t)cc * nxy + (int64_t)bb * g.nx + aa];
    }
    y[i] = s;
    if (part) acc_dot += (double)x[i] * (double)s;
  }
  if (part) {
    red[threadIdx.x] = acc_dot;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
  }
}

// Deterministic final sum of block partials (single block, fixed order).
__global__ void sum_parts_kernel(const double *__restrict__ part, int np,
                                 double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// CG state scalars on the device: [0] rr, [1] pq, [2] rr_new, [3] alpha,
// [4] beta.  x += alpha p; r -= alpha q; partial of r.r
__global__ __launch_bounds__(256) void cg_update_kernel(
    int64_t n, float *__restrict__ x, float *__restrict__ r,
    const float *__restrict__ p, const float *__restrict__ q,
    const double *__restrict__ sc, double *__restrict__ part) {
  __shared__ double red[256];
  const double pq = sc[1];
  const float alpha = pq != 0.0 ? (float)(sc[0] / pq) : 0.0f;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    float ri = r[i] - alpha * q[i];
    r[i] = ri;
    acc += (double)ri * (double)ri;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void cg_direction_kernel(int64_t n, const float *__restrict__ r,
                                    float *__restrict__ p,
                                    double *__restrict__ sc) {
  const double rr = sc[0], rr2 = sc[2];
  const float beta = rr != 0.0 ? (float)(rr2 / rr) : 0.0f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = r[i] + beta * p[i];
}

__global__ void cg_shift_kernel(double *sc) { sc[0] = sc[2]; }

__global__ __launch_bounds__(256) void dot32_kernel(int64_t n,
                                                    const float *__restrict__ a,
                                                    const float *__restrict__ b,
                                                    double *__restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += (double)a[i] * (double)b[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// r64 = b - K x64 (K x64 precomputed in kx); r32 = (float)(r64 * scale);
// partial of r64.r64
__global__ __launch_bounds__(256) void residual_kernel(
    int64_t n, const float *__restrict__ b, const double *__restrict__ kx,
    double *__restrict__ r64, double *__restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double ri = (double)b[i] - kx[i];
    r64[i] = ri;
    acc += ri * ri;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void scale_to_f32_kernel(int64_t n, const double *__restrict__ r64,
                                    double scale, float *__restrict__ r32,
                                    float *__restrict__ p32,
                                    float *__restrict__ d32) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float v = (float)(r64[i] * scale);
    r32[i] = v;
    p32[i] = v;
    d32[i] = 0.0f;
  }
}

__global__ void accumulate_kernel(int64_t n, const float *__restrict__ d32,
                                  double inv_scale, double *__restrict__ x64) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x64[i] += (double)d32[i] * inv_scale;
}

__global__ void f64_to_f32_kernel(int64_t n, const double *__restrict__ a,
                                  float *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// Fine-grid evaluation: fine point (a, b, c) of the (N s + 1)^3 lattice; its
// class (a mod s, b mod s, c mod s) selects the tap list, offsets are
// relative to the coarse point (a / s, b / s, c / s).
__global__ __launch_bounds__(256) void fine_eval_kernel(
    GridP cg, int32_t fx, int32_t fy, int32_t fz, int32_t s,
    const Tap *__restrict__ taps, const int32_t *__restrict__ cls_off,
    const float *__restrict__ w, float *__restrict__ out) {
  const int64_t n = (int64_t)fx * fy * fz;
  const int64_t cxy = (int64_t)cg.nx * cg.ny;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = (int32_t)(i % fx), b = (int32_t)((i / fx) % fy),
                  c = (int32_t)(i / ((int64_t)fx * fy));
    const int cls = ((c % s) * s + (b % s)) * s + (a % s);
    const int32_t i0 = a / s, j0 = b / s, k0 = c / s;
    float acc = 0.0f;
    for (int t = cls_off[cls]; t < cls_off[cls + 1]; t++) {
      const Tap tp = taps[t];
      const int32_t ii = i0 + tp.di, jj = j0 + tp.dj, kk = k0 + tp.dk;
      if (ii < 0 || jj < 0 || kk < 0 || ii >= cg.nx || jj >= cg.ny || kk >= cg.nz)
        continue;
      acc += w[(int64_t)kk * cxy + (int64_t)jj * cg.nx + ii] * tp.w;
    }
    out[i] = acc;
  }
}

// Taps of the coarse kernel matrix: offsets with m = di^2+dj^2+dk^2 inside
// the truncation radius (m <= -ln(threshold)) and exp(-m) > threshold.
std::vector<Tap> coarse_taps(double threshold) {
  std::vector<Tap> t;
  const double R2 = -log(threshold);
  const int r = (int)ceil(sqrt(R2)) + 1;
  for (int dk = -r; dk <= r; dk++)
    for (int dj = -r; dj <= r; dj++)
      for (int di = -r; di <= r; di++) {
        const double m = (double)(di * di + dj * dj + dk * dk);
        if (m > R2) continue;
        const double v = exp(-m);
        if (!(v > threshold)) continue;
        t.push_back(Tap{di, dj, dk, (float)v});
      }
  return t;
}

}  // namespace

int rbf_weights(hipStream_t st, Arena &ar, const GridP &g, const float *b,
                double threshold, double reltol, int max_outer,
                int max_inner, float *w_out, double *final_rel) {
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  std::vector<Tap> taps = coarse_taps(threshold);
  Tap *d_taps = ar.get<Tap>("rbf_taps", taps.size() + 1);
  SDF_HIP(hipMemcpyAsync(d_taps, taps.data(), taps.size() * sizeof(Tap),
                         hipMemcpyHostToDevice, st));
  const int nt = (int)taps.size();
  double *x64 = ar.get<double>("rbf_x64", n + 1);
  double *kx = ar.get<double>("rbf_kx", n + 1);
  double *r64 = ar.get<double>("rbf_r64", n + 1);
  float *r32 = ar.get<float>("rbf_r32", n + 1);
  float *p32 = ar.get<float>("rbf_p32", n + 1);
  float *q32 = ar.get<float>("rbf_q32", n + 1);
  float *d32 = ar.get<float>("rbf_d32", n + 1);
  double *part = ar.get<double>("rbf_part", kRedBlocks + 1);
  double *sc = ar.get<double>("rbf_scalars", 8);
  const int nb = blocks_for(n, 256) < kRedBlocks ? blocks_for(n, 256) : kRedBlocks;
  launch_fill_f64(st, x64, n, 0.0);

  auto read_scalar = [&](int k) {
    double v = 0.0;
    SDF_HIP(hipMemcpyAsync(&v, sc + k, sizeof(double), hipMemcpyDeviceToHost, st));
    SDF_HIP(hipStreamSynchronize(st));
    return v;
  };
  // ||b||
  hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(256), 0, st, n, b,
                     (const double *)x64, r64, part);  // x64 = 0: r = b
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, st, part, nb, sc + 5);
  SDF_HIP(hipGetLastError());
  const double bnorm = sqrt(read_scalar(5));
  int total_inner = 0;
  double rel = bnorm > 0 ? 1.0 : 0.0;
  for (int outer = 0; outer < max_outer && bnorm > 0; outer++) {
    // true residual in FP64
    hipLaunchKernelGGL(stencil_kernel<double>, dim3(blocks_for(n, 256)), dim3(256),
                       0, st, g, d_taps, nt, (const double *)x64, kx,
                       (double *)nullptr);
    hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(256), 0, st, n, b,
                       (const double *)kx, r64, part);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, st, part, nb, sc + 5);
    SDF_HIP(hipGetLastError());
    const double rn = sqrt(read_scalar(5));
    rel = rn / bnorm;
    if (rn <= reltol * bnorm) break;
    // FP32 CG on K d = r / ||r|| (unit right-hand side)
    const double scale = 1.0 / rn;
    hipLaunchKernelGGL(scale_to_f32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
                       st, n, (const double *)r64, scale, r32, p32, d32);
    hipLaunchKernelGGL(dot32_kernel, dim3(nb), dim3(256), 0, st, n,
                       (const float *)r32, (const float *)r32, part);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, st, part, nb, sc);
    // inner tolerance: enough to make progress in one or two outer steps
    const double inner_tol = fmax(0.5 * reltol * bnorm / rn, 1e-4);
    int it = 0;
    for (; it < max_inner; it++) {
      hipLaunchKernelGGL(stencil_kernel<float>, dim3(nb), dim3(256), 0, st, g,
                         d_taps, nt, (const float *)p32, q32, part);
      hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, st, part, nb, sc + 1);
      hipLaunchKernelGGL(cg_update_kernel, dim3(nb), dim3(256), 0, st, n, d32, r32,
                         (const float *)p32, (const float *)q32, (const double *)sc,
                         part);
      hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, st, part, nb, sc + 2);
      hipLaunchKernelGGL(cg_direction_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
                         st, n, (const float *)r32, p32, sc);
      hipLaunchKernelGGL(cg_shift_kernel, dim3(1), dim3(1), 0, st, sc);
      SDF_HIP(hipGetLastError());
      if ((it & 7) == 7) {
        if (sqrt(read_scalar(0)) <= inner_tol) {
          it++;
          break;
        }
      }
    }
    total_inner += it;
    hipLaunchKernelGGL(accumulate_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
                       st, n, (const float *)d32, rn, x64);
    SDF_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                     n, (const double *)x64, w_out);
  SDF_HIP(hipGetLastError());
  if (final_rel) *final_rel = rel;
  return total_inner;
}

void rbf_prepare(hipStream_t st, Arena &ar, const double *d, int64_t n,
                 float *f32) {
  unsigned int *w = ar.get<unsigned int>("rbf_pv", 4);
  launch_fill_u32(st, w, 2, 0u);
  hipLaunchKernelGGL(to_f32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, d,
                     n, f32, w, w + 1);
  SDF_HIP(hipGetLastError());
  unsigned int nbig = 0;
  SDF_HIP(hipMemcpyAsync(&nbig, w + 1, 4, hipMemcpyDeviceToHost, st));
  SDF_HIP(hipStreamSynchronize(st));
  if (nbig) {
    hipLaunchKernelGGL(fix_big_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                       f32, n, (const unsigned int *)w);
    SDF_HIP(hipGetLastError());
  }
}

void rbf_fine(hipStream_t st, Arena &ar, const GridP &g, int smooth,
              double threshold, const float *w, float *fine) {
  // per-class tap lists: fine offset (ca, cb, cc) / s inside the coarse cell
  const double R2 = -log(threshold);
  const int r = (int)ceil(sqrt(R2)) + 2;
  std::vector<Tap> taps;
  std::vector<int32_t> off;
  for (int cc = 0; cc < smooth; cc++)
    for (int cb = 0; cb < smooth; cb++)
      for (int ca = 0; ca < smooth; ca++) {
        off.push_back((int32_t)taps.size());
        for (int dk = -r; dk <= r; dk++)
          for (int dj = -r; dj <= r; dj++)
            for (int di = -r; di <= r; di++) {
              const double ex = (double)ca / smooth - di, ey = (double)cb / smooth - dj,
                           ez = (double)cc / smooth - dk;
              const double m = ex * ex + ey * ey + ez * ez;
              if (m > R2) continue;
              taps.push_back(Tap{di, dj, dk, (float)exp(-m)});
            }
      }
  off.push_back((int32_t)taps.size());
  Tap *d_taps = ar.get<Tap>("rbf_fine_taps", taps.size() + 1);
  int32_t *d_off = ar.get<int32_t>("rbf_fine_off", off.size() + 1);
  SDF_HIP(hipMemcpyAsync(d_taps, taps.data(), taps.size() * sizeof(Tap),
                         hipMemcpyHostToDevice, st));
  SDF_HIP(hipMemcpyAsync(d_off, off.data(), off.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, st));
  const int32_t fx = (g.nx - 1) * smooth + 1, fy = (g.ny - 1) * smooth + 1,
                fz = (g.nz - 1) * smooth + 1;
  const int64_t nf = (int64_t)fx * fy * fz;
  hipLaunchKernelGGL(fine_eval_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, st,
                     g, fx, fy, fz, smooth, d_taps, d_off, w, fine);
  SDF_HIP(hipGetLastError());
  // the host vectors must outlive the async copies
  SDF_HIP(hipStreamSynchronize(st));
}

}  // namespace sdf
