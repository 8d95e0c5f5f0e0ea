// host_math.cpp -- host-side numerics of libscft_hip.so: Romberg pairing
// weights, the EXP target profile and uniform meshes.
#include "host_math.h"

#include <cmath>
#include <vector>

namespace scft {

// Value at h = 0 of the polynomial through (h[k], s[k]) (Neville), the
// extrapolation step of romint (DEALII_SCFT/src/romint.c:50, polint.c:5-41).
static double neville_zero(const double* h, const double* s, int n) {
  double p[16];
  for (int k = 0; k < n; k++) p[k] = s[k];
  for (int d = 1; d < n; d++)
    for (int k = 0; k + d < n; k++)
      p[k] = (h[k] * p[k + 1] - h[k + d] * p[k]) / (h[k] - h[k + d]);
  return p[0];
}

// Weights c_0..c_m with romint(f, m, hh) == sum c_j f_j: romint.c:21-57 is a
// fixed linear functional (trapezoid levels with 2^l panels, the last K = 5
// extrapolated to h = 0 along h_l = 4^-l), so its weights are the Neville
// combination of the five trapezoid rules.  m must be 2^k >= 16.
bool romberg_weights(int m, double hh, std::vector<double>& c) {
  const int K = 5;
  if (m < 16 || (m & (m - 1)) != 0) return false;
  const int levels = static_cast<int>(std::log(m * 1.0) / 0.6931471805599453 + 1.5);
  double h[K], e[K], lam[K];
  for (int k = 0; k < K; k++) h[k] = std::pow(0.25, levels - K + k);
  for (int k = 0; k < K; k++) {
    for (int t = 0; t < K; t++) e[t] = (t == k) ? 1.0 : 0.0;
    lam[k] = neville_zero(h, e, K);
  }
  c.assign(m + 1, 0.0);
  for (int k = 0; k < K; k++) {
    const int level = levels - K + k;
    const int stride = m >> level;
    const double panel = stride * hh;
    for (int j = 0; j <= m; j += stride)
      c[j] += lam[k] * ((j == 0 || j == m) ? 0.5 * panel : panel);
  }
  return true;
}

// Pairing weights of the folded sum phi = sum_j c_j q_j q_{N-j}: the weights
// are symmetric (c_j = c_{N-j}), so phi = c_{N/2} q_{N/2}^2 +
// sum_{j > N/2} 2 c_j q_j q_{N-j}; pw[t] belongs to step N/2 + t.
bool pairing_weights(int N, std::vector<double>& pw) {
  std::vector<double> c;
  if (!romberg_weights(N, 1.0 / N, c)) return false;
  const int half = N / 2;
  pw.assign(half + 1, 0.0);
  pw[0] = c[half];
  for (int t = 1; t <= half; t++) pw[t] = 2.0 * c[half + t];
  return true;
}

// EXP wall profile as the reference evaluates it, (e^u-1)^2/(e^u+1)^2 with
// u = 4 tau x/(tau^2-x^2), NaN -> 1 (DEALII_SCFT/src/scft.cc:201-207).
double exp_profile(double x, double tau) {
  const double u = 4.0 * tau * x / (tau * tau - x * x);
  const double e = std::exp(u);
  double v = std::pow(e - 1.0, 2) / std::pow(e + 1.0, 2);
  if (std::isnan(v)) v = 1.0;
  return v;
}

// Position-based target on all nodes (scft.cc:188-215), mirrored for DD.
void target_position(int M, const double* x, double tau, bool mirror, double* phi_t) {
  for (int i = 0; i < M; i++) phi_t[i] = 1.0;
  for (int i = 0; i < M; i++) {
    if (x[i] <= tau) {
      phi_t[i] = exp_profile(x[i], tau);
      if (mirror) phi_t[M - i - 1] = phi_t[i];
    } else {
      break;
    }
  }
}

// x_i = i L / (M-1) (1D_FEM.c:310).
void mesh_uniform(int M, double L, double* x) {
  for (int i = 0; i < M; i++) x[i] = i * L / (M - 1);
}

}  // namespace scft
