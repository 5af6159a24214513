// Learned search-radius model: a coordinate MLP (16 Fourier frequencies,
// three hidden layers of 128, ReLU) in BF16 on the CDNA4 matrix cores with
// FP32 accumulation, trained per mesh by the kernels in sdf_mlp.hip.
//
// Network (u = normalised point):
//   phi[0..2] = u, phi[3] = 0, phi[4+6k+2a] = sin(pi f_k u_a),
//   phi[5+6k+2a] = cos(pi f_k u_a), f_k = 2^(0.4k) (k < 16, a < 3),
//   phi[100..127] = 0
//   h1 = relu(W1 phi + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3),
//   y = w4 . h3 + b4
// Parameters (FP32 masters, W stored out x in, row-major):
//   W1 @0, b1 @16384, W2 @16512, b2 @32896, W3 @33024, b3 @49408,
//   w4 @49536, b4 @49664 (49665 floats).
// Features, W1..W3, w4 and hidden activations are BF16 (round to nearest
// even); back-propagated pre-activation gradients are rounded to BF16
// before they multiply weights or activations.
#pragma once

#include "sdf_internal.h"

namespace sdf {

constexpr int kMlpH = 128;
constexpr int kMlpParams = 49665;
enum : int {
  kOffW1 = 0,
  kOffB1 = 16384,
  kOffW2 = 16512,
  kOffB2 = 32896,
  kOffW3 = 33024,
  kOffB3 = 49408,
  kOffW4 = 49536,
  kOffB4 = 49664
};
// BF16 elements of one packed 128x128 operand (8 row tiles x 4 k-steps x
// 64 lanes x 8 values).
constexpr int kMlpPacked = 16384;
// Per-block partial sums of the training kernel: db1, db2, db3, dw4 (128
// each), db4, loss.
constexpr int kMlpBPart = 4 * 128 + 2;
constexpr int kMlpSlice = 512;  // points per weight-gradient block

// Normalisation of points: u = (p - c) * s, computed in FP64, rounded to FP32.
struct MlpNorm {
  double c[3];
  double s;
};

struct MlpDev {
  float *params = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *afwd = nullptr;  // 3 packed layers: forward A operands
  uint16_t *abwd = nullptr;  // layers 2, 3 packed transposed: backward A operands
  float *w4b = nullptr;      // w4 rounded to BF16 (as floats)
  int step = 0;              // Adam steps taken
};

void mlp_alloc(Arena &ar, MlpDev &m);
// Pack the BF16 operands from the FP32 masters.
void mlp_pack(hipStream_t s, MlpDev &m);
// y = network(u) for n points; point i is u[stride*i .. +3] (normalised).
void mlp_forward(hipStream_t s, const MlpDev &m, const float *u, int stride, int64_t n,
                 float *y);
// Training set from exact distances dc on the lattice gc: every point with
// dc < hc and a hash-selected subset of the others (about as many), as
// (u.x, u.y, u.z, min(dc / hc, 1)) in lattice order.  Returns the count.
int64_t mlp_fit_samples(hipStream_t s, Arena &ar, const GridP &gc, const MlpNorm &nm,
                        const double *dc, double hc, float4 **out, int64_t *nband);
// Loss and gradient of the batch samples[idx[i]] (idx may be null: identity),
// i < n.  samples: (u.x, u.y, u.z, target).  When lr > 0 an Adam step with
// that learning rate follows (step counter m.step advanced); grad / loss may
// be null.
void mlp_step(hipStream_t s, Arena &ar, MlpDev &m, const float4 *samples,
              const uint32_t *idx, int64_t n, float *grad, float *loss, float lr);
// Squared search radius of every point of each tile that has distance work:
// r = min(max((y + delta) * scale, h / 100), h), written as r^2 to
// r2[chunk_off[t]*64 + lane].
void mlp_tile_radius(hipStream_t s, const MlpDev &m, const MlpNorm &nm, const GridP &g,
                     const Bins &b, float scale, float delta, float *r2);

}  // namespace sdf
