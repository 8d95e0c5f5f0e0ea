/*
 * oracle_mlp.c -- FP64 reference of the warm-start MLP.
 * TEST INFRASTRUCTURE ONLY.
 */
#include "oracle_mlp.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

long orc_mlp_param_count(int n_layers, const int *dims) {
  long n = 0;
  for (int l = 0; l < n_layers; l++) n += (long)dims[l + 1] * dims[l] + dims[l + 1];
  return n;
}

/* acts[l] holds the batch x dims[l] activations of layer l (acts[0] = X). */
static double **forward_all(int n_layers, const int *dims,
                            const double *params, int batch, const double *X) {
  double **acts = (double **)malloc(sizeof(double *) * (n_layers + 1));
  acts[0] = (double *)malloc(sizeof(double) * batch * dims[0]);
  memcpy(acts[0], X, sizeof(double) * batch * dims[0]);
  const double *p = params;
  for (int l = 0; l < n_layers; l++) {
    int din = dims[l], dout = dims[l + 1];
    const double *W = p, *b = p + (long)dout * din;
    p = b + dout;
    acts[l + 1] = (double *)malloc(sizeof(double) * batch * dout);
    for (int s = 0; s < batch; s++)
      for (int o = 0; o < dout; o++) {
        double z = b[o];
        for (int i = 0; i < din; i++) z += W[(long)o * din + i] * acts[l][(long)s * din + i];
        acts[l + 1][(long)s * dout + o] = (l < n_layers - 1) ? tanh(z) : z;
      }
  }
  return acts;
}

static void free_acts(double **acts, int n_layers) {
  for (int l = 0; l <= n_layers; l++) free(acts[l]);
  free(acts);
}

void orc_mlp_forward(int n_layers, const int *dims, const double *params,
                     int batch, const double *X, double *Y) {
  double **acts = forward_all(n_layers, dims, params, batch, X);
  memcpy(Y, acts[n_layers], sizeof(double) * batch * dims[n_layers]);
  free_acts(acts, n_layers);
}

/* loss = mean over batch and outputs of (y - t)^2; grad = d loss / d params */
double orc_mlp_loss_grad(int n_layers, const int *dims, const double *params,
                         int batch, const double *X, const double *T,
                         double *grad) {
  double **acts = forward_all(n_layers, dims, params, batch, X);
  int dlast = dims[n_layers];
  double loss = 0.0;
  double *delta = (double *)malloc(sizeof(double) * batch * dlast);
  double scale = 2.0 / ((double)batch * dlast);
  for (long k = 0; k < (long)batch * dlast; k++) {
    double e = acts[n_layers][k] - T[k];
    loss += e * e;
    delta[k] = scale * e;
  }
  loss /= (double)batch * dlast;
  /* offsets of each layer's parameters */
  long *off = (long *)malloc(sizeof(long) * (n_layers + 1));
  off[0] = 0;
  for (int l = 0; l < n_layers; l++) off[l + 1] = off[l] + (long)dims[l + 1] * dims[l] + dims[l + 1];
  for (int l = n_layers - 1; l >= 0; l--) {
    int din = dims[l], dout = dims[l + 1];
    const double *W = params + off[l];
    double *gW = grad + off[l], *gb = gW + (long)dout * din;
    for (int o = 0; o < dout; o++) {
      double sb = 0.0;
      for (int s = 0; s < batch; s++) sb += delta[(long)s * dout + o];
      gb[o] = sb;
      for (int i = 0; i < din; i++) {
        double sw = 0.0;
        for (int s = 0; s < batch; s++) sw += delta[(long)s * dout + o] * acts[l][(long)s * din + i];
        gW[(long)o * din + i] = sw;
      }
    }
    if (l > 0) {
      double *nd = (double *)malloc(sizeof(double) * batch * din);
      for (int s = 0; s < batch; s++)
        for (int i = 0; i < din; i++) {
          double g = 0.0;
          for (int o = 0; o < dout; o++) g += W[(long)o * din + i] * delta[(long)s * dout + o];
          double a = acts[l][(long)s * din + i]; /* tanh output */
          nd[(long)s * din + i] = g * (1.0 - a * a);
        }
      free(delta);
      delta = nd;
    }
  }
  free(delta);
  free(off);
  free_acts(acts, n_layers);
  return loss;
}

/* Adam (Kingma & Ba 2015, Algorithm 1) with bias correction at step t>=1. */
void orc_adam_step(long n, double *params, const double *grad, double *m,
                   double *v, int t, double lr, double b1, double b2,
                   double eps) {
  double c1 = 1.0 - pow(b1, t), c2 = 1.0 - pow(b2, t);
  for (long k = 0; k < n; k++) {
    m[k] = b1 * m[k] + (1.0 - b1) * grad[k];
    v[k] = b2 * v[k] + (1.0 - b2) * grad[k] * grad[k];
    double mh = m[k] / c1, vh = v[k] / c2;
    params[k] -= lr * mh / (sqrt(vh) + eps);
  }
}
