/* CPU oracle of the learned-radius coordinate MLP.  TEST INFRASTRUCTURE ONLY:
 * tests/, the smoke check and the benchmark's CPU leg use it as the checker;
 * the product library never links it.  Formulas: see mlp_oracle.h. */
#include "mlp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { H = ORC_MLP_H };
enum {
  OW1 = 0, OB1 = 16384, OW2 = 16512, OB2 = 32896, OW3 = 33024, OB3 = 49408,
  OW4 = 49536, OB4 = 49664
};

float orc_bf16(float x) {
  uint32_t b;
  memcpy(&b, &x, 4);
  if ((b & 0x7F800000u) == 0x7F800000u) {  /* inf / nan */
    if (b & 0x007FFFFFu) b |= 0x00400000u;
    b &= 0xFFFF0000u;
  } else {
    uint32_t lsb = (b >> 16) & 1u;
    b += 0x7FFFu + lsb;
    b &= 0xFFFF0000u;
  }
  float r;
  memcpy(&r, &b, 4);
  return r;
}

void orc_mlp_freqs(float f[16]) {
  for (int k = 0; k < 16; k++) f[k] = (float)pow(2.0, 0.4 * k);
}

static double rnd(double x, int mode) { return mode == ORC_MLP_BF16 ? (double)orc_bf16((float)x) : x; }

void orc_mlp_features(const float u[3], int mode, double phi[H]) {
  float f[16];
  orc_mlp_freqs(f);
  for (int i = 0; i < H; i++) phi[i] = 0.0;
  for (int a = 0; a < 3; a++) phi[a] = rnd(u[a], mode);
  const double twopi = 6.283185307179586476925286766559;
  for (int k = 0; k < 16; k++)
    for (int a = 0; a < 3; a++) {
      double s, c;
      if (mode == ORC_MLP_BF16) {
        /* argument in revolutions, rounded like the device: fl(f_k u_a) / 2 */
        float r = (f[k] * u[a]) * 0.5f;
        s = (double)orc_bf16((float)sin(twopi * (double)r));
        c = (double)orc_bf16((float)cos(twopi * (double)r));
      } else {
        double x = 3.14159265358979323846 * (double)f[k] * (double)u[a];
        s = sin(x);
        c = cos(x);
      }
      phi[4 + 6 * k + 2 * a] = s;
      phi[5 + 6 * k + 2 * a] = c;
    }
}

/* weights as used by the arithmetic (bf16-rounded in BF16 mode) */
static double *weights(const double *params, int mode) {
  double *w = (double *)malloc(sizeof(double) * ORC_MLP_NPARAM);
  for (int i = 0; i < ORC_MLP_NPARAM; i++) {
    int is_w = (i < OB1) || (i >= OW2 && i < OB2) || (i >= OW3 && i < OB3) ||
               (i >= OW4 && i < OB4);
    w[i] = is_w ? rnd(params[i], mode) : params[i];
  }
  return w;
}

typedef struct {
  double phi[H], z1[H], h1[H], z2[H], h2[H], z3[H], h3[H], y;
} Acts;

static void layer(const double *W, const double *b, const double *in, double *z,
                  double *h, int mode) {
  for (int n = 0; n < H; n++) {
    double s = 0.0;
    for (int k = 0; k < H; k++) s += W[n * H + k] * in[k];
    z[n] = s + b[n];
    h[n] = rnd(z[n] > 0.0 ? z[n] : 0.0, mode);
  }
}

static void forward1(const double *w, const float *u, int mode, Acts *A) {
  orc_mlp_features(u, mode, A->phi);
  layer(w + OW1, w + OB1, A->phi, A->z1, A->h1, mode);
  layer(w + OW2, w + OB2, A->h1, A->z2, A->h2, mode);
  layer(w + OW3, w + OB3, A->h2, A->z3, A->h3, mode);
  double s = 0.0;
  for (int n = 0; n < H; n++) s += w[OW4 + n] * A->h3[n];
  A->y = s + w[OB4];
}

void orc_mlp_forward(const double *params, const float *u, int64_t n, int mode,
                     double *y_out) {
  double *w = weights(params, mode);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    Acts A;
    forward1(w, u + 3 * i, mode, &A);
    y_out[i] = A.y;
  }
  free(w);
}

double orc_mlp_loss_grad(const double *params, const float *u, const float *t,
                         int64_t n, int mode, double *grad) {
  double *w = weights(params, mode);
  memset(grad, 0, sizeof(double) * ORC_MLP_NPARAM);
  double loss = 0.0;
  for (int64_t i = 0; i < n; i++) {
    Acts A;
    forward1(w, u + 3 * i, mode, &A);
    const double r = A.y - (double)t[i];
    loss += r * r;
    const double dy = 2.0 * r / (double)n;
    grad[OB4] += dy;
    double dz3[H], dz2[H], dz1[H];
    for (int k = 0; k < H; k++) {
      grad[OW4 + k] += dy * A.h3[k];
      const double d = A.z3[k] > 0.0 ? w[OW4 + k] * dy : 0.0;
      grad[OB3 + k] += d;
      dz3[k] = rnd(d, mode);
    }
    for (int k = 0; k < H; k++) {
      double s = 0.0;
      for (int m = 0; m < H; m++) s += w[OW3 + m * H + k] * dz3[m];
      const double d = A.z2[k] > 0.0 ? s : 0.0;
      grad[OB2 + k] += d;
      dz2[k] = rnd(d, mode);
    }
    for (int k = 0; k < H; k++) {
      double s = 0.0;
      for (int m = 0; m < H; m++) s += w[OW2 + m * H + k] * dz2[m];
      const double d = A.z1[k] > 0.0 ? s : 0.0;
      grad[OB1 + k] += d;
      dz1[k] = rnd(d, mode);
    }
    for (int m = 0; m < H; m++)
      for (int k = 0; k < H; k++) {
        grad[OW3 + m * H + k] += dz3[m] * A.h2[k];
        grad[OW2 + m * H + k] += dz2[m] * A.h1[k];
        grad[OW1 + m * H + k] += dz1[m] * A.phi[k];
      }
  }
  free(w);
  return loss / (double)n;
}

void orc_mlp_adam(float *params, float *m, float *v, const float *grad,
                  int64_t n, int t, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  /* bias corrections in FP64, rounded once (the device takes them the same way) */
  const float c1 = (float)(1.0 - pow(0.9, t)), c2 = (float)(1.0 - pow(0.999, t));
  for (int64_t i = 0; i < n; i++) {
    const float g = grad[i];
    m[i] = b1 * m[i] + (1.0f - b1) * g;
    v[i] = b2 * v[i] + (1.0f - b2) * (g * g);
    const float mh = m[i] / c1, vh = v[i] / c2;
    params[i] = params[i] - lr * mh / (sqrtf(vh) + eps);
  }
}
