// Connected components of the interior mask (sdf >= threshold) with
// 6-connectivity, and the small-component flip of remove_sdf_artifacts!.
//
// Reference: src/SignedDistances/SdfArtifactRemoval.jl:134-248 (removal) and
// :259-318 (analysis).  The reference unions from all Julia threads into one
// unsynchronised UnionFind; here every union is a lock-free CAS that hooks
// the larger root under the smaller one, so each component ends with its
// smallest linear index as root -- a deterministic, canonical label.
//
// Invariant during hooking: lab[x] <= x and lab[x] is an ancestor of x.
// Roots change only by CAS; non-roots only by path halving to an ancestor.
// Reads of lab during hooking use device-scope atomic loads (coherent across
// XCDs); a stale value is still an ancestor, and a failed CAS returns the
// current value, so every loop makes progress and ends.
#include "sdf_internal.h"

namespace sdf {

namespace {

inline int blocks_for(int64_t n, int t) {
  int64_t b = (n + t - 1) / t;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (int)b;
}

__device__ inline int32_t ld(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline int32_t find_root(int32_t *lab, int32_t x) {
  int32_t p = ld(lab + x);
  while (p != x) {
    int32_t g = ld(lab + p);
    if (g != p) {
      // path halving: x -> grandparent (an ancestor, never a root change)
      __hip_atomic_store(lab + x, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    x = p;
    p = g;
  }
  return x;
}

__device__ inline void unite(int32_t *lab, int32_t a, int32_t b) {
  while (true) {
    a = find_root(lab, a);
    b = find_root(lab, b);
    if (a == b) return;
    if (a < b) {
      int32_t t = a;
      a = b;
      b = t;
    }
    // hook the larger root a under b
    int32_t expected = a;
    if (__hip_atomic_compare_exchange_strong(lab + a, &expected, b,
                                             __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
    a = expected;  // a was hooked meanwhile; retry from its new parent
  }
}

__global__ void ccl_init_kernel(const double *__restrict__ sdf, int64_t n,
                                double thr, int32_t *__restrict__ lab) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    lab[i] = (sdf[i] >= thr) ? (int32_t)i : -1;
}

__global__ void ccl_hook_kernel(GridP g, int32_t *lab) {
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (lab[i] < 0) continue;
    const int64_t x = i % g.nx, y = (i / g.nx) % g.ny, z = i / nxy;
    if (x + 1 < g.nx && lab[i + 1] >= 0) unite(lab, (int32_t)i, (int32_t)(i + 1));
    if (y + 1 < g.ny && lab[i + g.nx] >= 0)
      unite(lab, (int32_t)i, (int32_t)(i + g.nx));
    if (z + 1 < g.nz && lab[i + nxy] >= 0)
      unite(lab, (int32_t)i, (int32_t)(i + nxy));
  }
}

// Full compression (after hooking; plain accesses are coherent again), then
// component sizes with one atomic per distinct root per wavefront.
__global__ void ccl_compress_count_kernel(int64_t n, int32_t *__restrict__ lab,
                                          uint32_t *__restrict__ size) {
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    int32_t r = -1;
    if (i < n && lab[i] >= 0) {
      r = lab[i];
      while (lab[r] != r) r = lab[r];
      lab[i] = r;
    }
    uint64_t pending = __ballot(r >= 0);
    while (pending) {
      const int leader = __ffsll((unsigned long long)pending) - 1;
      const int32_t R = __shfl(r, leader);
      const uint64_t same = __ballot(r == R) & pending;
      if ((int)(threadIdx.x & 63) == leader) atomicAdd(&size[R], (uint32_t)__popcll(same));
      pending &= ~same;
    }
  }
}

// Largest component: max of (size << 32 | ~root) -- largest size, smallest
// root among equal sizes.
__global__ void ccl_largest_kernel(int64_t n, const int32_t *__restrict__ lab,
                                   const uint32_t *__restrict__ size,
                                   unsigned long long *best,
                                   unsigned long long *ncomp) {
  __shared__ unsigned long long s[256];
  __shared__ unsigned long long c[256];
  unsigned long long m = 0, cnt = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (lab[i] == (int32_t)i) {
      unsigned long long v = ((unsigned long long)size[i] << 32) |
                             (unsigned long long)(0xFFFFFFFFu - (uint32_t)i);
      m = v > m ? v : m;
      cnt++;
    }
  }
  s[threadIdx.x] = m;
  c[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      s[threadIdx.x] = s[threadIdx.x] > s[threadIdx.x + off] ? s[threadIdx.x] : s[threadIdx.x + off];
      c[threadIdx.x] += c[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicMax(best, s[0]);
    atomicAdd(ncomp, c[0]);
  }
}

__global__ void ccl_flip_kernel(int64_t n, const int32_t *__restrict__ lab,
                                const uint32_t *__restrict__ size,
                                int32_t largest, uint32_t min_size,
                                double *__restrict__ sdf,
                                unsigned long long *flipped) {
  __shared__ unsigned long long s[256];
  unsigned long long cnt = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = lab[i];
    if (r >= 0 && r != largest && size[r] < min_size) {
      sdf[i] = -fabs(sdf[i]);
      cnt++;
    }
  }
  s[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0 && s[0]) atomicAdd(flipped, s[0]);
}

__global__ void ccl_roots_kernel(int64_t n, const int32_t *__restrict__ lab,
                                 const uint32_t *__restrict__ size,
                                 int64_t *__restrict__ roots,
                                 int64_t *__restrict__ sizes, uint32_t *nroots,
                                 uint32_t cap) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (lab[i] == (int32_t)i) {
      uint32_t slot = atomicAdd(nroots, 1u);
      if (slot < cap) {
        roots[slot] = i;
        sizes[slot] = size[i];
      }
    }
  }
}

__global__ void ccl_labels_out_kernel(int64_t n, const int32_t *__restrict__ lab,
                                      int64_t *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = lab[i];
}

}  // namespace

CclResult run_ccl(hipStream_t s, Arena &ar, const GridP &g, const double *sdf,
                  double thr) {
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  CclResult r;
  r.lab = ar.get<int32_t>("ccl_lab", n + 1);
  r.size = ar.get<uint32_t>("ccl_size", n + 1);
  r.words = ar.get<unsigned long long>("ccl_words", 4);
  run_ccl_into(s, g, sdf, thr, r);
  return r;
}

void run_ccl_into(hipStream_t s, const GridP &g, const double *sdf, double thr,
                  CclResult &r) {
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  r.n = n;
  // every word the kernels accumulate into starts at zero (fill kernels, so
  // a buffer that kept an earlier call's contents is cleared the same way)
  launch_fill_u64(s, (uint64_t *)r.words, 4, 0ull);
  launch_fill_u32(s, r.size, n + 1, 0u);
  if (n == 0) return;
  hipLaunchKernelGGL(ccl_init_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                     sdf, n, thr, r.lab);
  SDF_HIP(hipGetLastError());
  hipLaunchKernelGGL(ccl_hook_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                     g, r.lab);
  SDF_HIP(hipGetLastError());
  hipLaunchKernelGGL(ccl_compress_count_kernel, dim3(blocks_for(n, 256)),
                     dim3(256), 0, s, n, r.lab, r.size);
  SDF_HIP(hipGetLastError());
  hipLaunchKernelGGL(ccl_largest_kernel, dim3(blocks_for(n, 256) < 2048 ? blocks_for(n, 256) : 2048),
                     dim3(256), 0, s, n, r.lab, r.size, r.words, r.words + 1);
  SDF_HIP(hipGetLastError());
}

void ccl_flip(hipStream_t s, const CclResult &r, int32_t largest,
              uint32_t min_size, double *sdf) {
  if (r.n == 0) return;
  hipLaunchKernelGGL(ccl_flip_kernel, dim3(blocks_for(r.n, 256) < 4096 ? blocks_for(r.n, 256) : 4096),
                     dim3(256), 0, s, r.n, r.lab, r.size, largest, min_size, sdf,
                     r.words + 2);
  SDF_HIP(hipGetLastError());
}

void ccl_roots(hipStream_t s, const CclResult &r, int64_t *roots, int64_t *sizes,
               uint32_t *nroots, uint32_t cap) {
  if (r.n == 0) return;
  hipLaunchKernelGGL(ccl_roots_kernel, dim3(blocks_for(r.n, 256)), dim3(256), 0,
                     s, r.n, r.lab, r.size, roots, sizes, nroots, cap);
  SDF_HIP(hipGetLastError());
}

void ccl_labels(hipStream_t s, const CclResult &r, int64_t *out) {
  if (r.n == 0) return;
  hipLaunchKernelGGL(ccl_labels_out_kernel, dim3(blocks_for(r.n, 256)), dim3(256),
                     0, s, r.n, r.lab, out);
  SDF_HIP(hipGetLastError());
}

}  // namespace sdf
