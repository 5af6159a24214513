/* CPU oracle of the learned-radius coordinate MLP (test infrastructure only).
 *
 * The reference (jezekon/stl2sdf.jl) has no learned component; this file is
 * the specification the device kernels are checked against.  Network:
 *   u = normalised point (3 floats)
 *   phi[0..2] = u, phi[3] = 0,
 *   phi[4 + 6k + 2a] = sin(pi f_k u_a), phi[5 + 6k + 2a] = cos(pi f_k u_a),
 *     f_k = 2^(0.4 k) (float), k = 0..15, a = 0..2;  phi[100..127] = 0
 *   h1 = relu(W1 phi + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3)
 *   y  = w4 . h3 + b4
 * with W* stored out x in, row-major.  Parameter vector (floats):
 *   W1 @0, b1 @16384, W2 @16512, b2 @32896, W3 @33024, b3 @49408,
 *   w4 @49536, b4 @49664; 49665 in total.
 * Loss over a batch of n points: (1/n) sum (y - t)^2.
 *
 * mode ORC_MLP_BF16 mirrors the device arithmetic: features, weights (W*, w4)
 * and hidden activations rounded to bfloat16 (round to nearest even), the
 * back-propagated pre-activation gradients rounded to bfloat16 before they
 * multiply weights or activations, biases and accumulation in higher
 * precision (FP64 here, FP32 on the device).
 * mode ORC_MLP_EXACT is the same network in FP64 with no rounding (features
 * from double sin/cos); its gradient is checked by finite differences. */
#ifndef MLP_ORACLE_H
#define MLP_ORACLE_H
#include <stdint.h>

#define ORC_MLP_H 128
#define ORC_MLP_NPARAM 49665
#define ORC_MLP_EXACT 0
#define ORC_MLP_BF16 1

float orc_bf16(float x);
void orc_mlp_freqs(float f[16]);
void orc_mlp_features(const float u[3], int mode, double phi[ORC_MLP_H]);
/* y_out[i] for n points (u: 3 x n, column-major as xyz triples). */
void orc_mlp_forward(const double *params, const float *u, int64_t n, int mode,
                     double *y_out);
/* Loss and its gradient w.r.t. every parameter (grad: ORC_MLP_NPARAM). */
double orc_mlp_loss_grad(const double *params, const float *u, const float *t,
                         int64_t n, int mode, double *grad);
/* One Adam step in FP32 (beta1 0.9, beta2 0.999, eps 1e-8), step t >= 1. */
void orc_mlp_adam(float *params, float *m, float *v, const float *grad,
                  int64_t n, int t, float lr);
#endif
