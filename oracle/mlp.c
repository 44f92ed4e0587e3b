/* mlp.c - CPU oracle of the learned skip predictor's network, in double.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/ may build, load or call this file.  It
 * restates the 132-64-64-1 ReLU network of the product's skip predictor (flat
 * parameters W1[64][132], b1, W2[64][64], b2, W3[64], b3) and the gradient of the
 * batch mean squared error by plain reverse-mode differentiation, so the GPU's
 * forward and backward kernels can be checked against it.  The network has no
 * counterpart in the reference.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define IN 132
#define H 64
#define OW1 0
#define OB1 (OW1 + H * IN)
#define OW2 (OB1 + H)
#define OB2 (OW2 + H * H)
#define OW3 (OB2 + H)
#define OB3 (OW3 + H)
#define NPARAM (OB3 + 1)

int orc_mlp_nparams(void) { return NPARAM; }

/* y[b] for normalised inputs x (n x 132) */
void orc_mlp_forward(const float* p, const float* x, int n, double* y) {
  double h1[H], h2[H];
  for (int b = 0; b < n; b++) {
    const float* xb = x + (size_t)b * IN;
    for (int j = 0; j < H; j++) {
      double s = p[OB1 + j];
      for (int k = 0; k < IN; k++) s += (double)p[OW1 + j * IN + k] * xb[k];
      h1[j] = s > 0.0 ? s : 0.0;
    }
    for (int j = 0; j < H; j++) {
      double s = p[OB2 + j];
      for (int k = 0; k < H; k++) s += (double)p[OW2 + j * H + k] * h1[k];
      h2[j] = s > 0.0 ? s : 0.0;
    }
    double o = p[OB3];
    for (int k = 0; k < H; k++) o += (double)p[OW3 + k] * h2[k];
    y[b] = o;
  }
}

/* loss = mean_b (y_b - t_b)^2 and its gradient with respect to every parameter */
double orc_mlp_grad(const float* p, const float* x, const float* t, int n, double* g) {
  double h1[H], h2[H], d1[H], d2[H];
  double loss = 0.0;
  memset(g, 0, sizeof(double) * NPARAM);
  for (int b = 0; b < n; b++) {
    const float* xb = x + (size_t)b * IN;
    for (int j = 0; j < H; j++) {
      double s = p[OB1 + j];
      for (int k = 0; k < IN; k++) s += (double)p[OW1 + j * IN + k] * xb[k];
      h1[j] = s > 0.0 ? s : 0.0;
    }
    for (int j = 0; j < H; j++) {
      double s = p[OB2 + j];
      for (int k = 0; k < H; k++) s += (double)p[OW2 + j * H + k] * h1[k];
      h2[j] = s > 0.0 ? s : 0.0;
    }
    double o = p[OB3];
    for (int k = 0; k < H; k++) o += (double)p[OW3 + k] * h2[k];
    const double r = o - t[b];
    loss += r * r / n;
    const double dy = 2.0 * r / n;
    g[OB3] += dy;
    for (int k = 0; k < H; k++) {
      g[OW3 + k] += dy * h2[k];
      d2[k] = h2[k] > 0.0 ? dy * p[OW3 + k] : 0.0;
    }
    for (int i = 0; i < H; i++) {
      g[OB2 + i] += d2[i];
      for (int k = 0; k < H; k++) g[OW2 + i * H + k] += d2[i] * h1[k];
    }
    for (int k = 0; k < H; k++) {
      double s = 0.0;
      for (int i = 0; i < H; i++) s += (double)p[OW2 + i * H + k] * d2[i];
      d1[k] = h1[k] > 0.0 ? s : 0.0;
    }
    for (int j = 0; j < H; j++) {
      g[OB1 + j] += d1[j];
      for (int k = 0; k < IN; k++) g[OW1 + j * IN + k] += d1[j] * xb[k];
    }
  }
  return loss;
}
