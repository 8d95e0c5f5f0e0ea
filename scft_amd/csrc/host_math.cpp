// host_math.cpp -- host-side numerics of libscft_hip.so: Romberg pairing
// weights, the EXP target profile and uniform meshes.
#include "host_math.h"

#include <algorithm>
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
  if (N < 16 || (N & (N - 1)) != 0) return false;
  pw.assign(N / 2 + 1, 0.0);
  return pairing_weights_into(N, pw.data());
}

// The same into a caller's array of N/2 + 1 doubles, with no allocation:
// each c_j summed over the five levels in the order romberg_weights uses, so
// the weights are the same bits.
bool pairing_weights_into(int N, double* pw) {
  const int K = 5;
  if (N < 16 || (N & (N - 1)) != 0) return false;
  const double hh = 1.0 / N;
  const int levels = static_cast<int>(std::log(N * 1.0) / 0.6931471805599453 + 1.5);
  double h[K], e[K], lam[K];
  for (int k = 0; k < K; k++) h[k] = std::pow(0.25, levels - K + k);
  for (int k = 0; k < K; k++) {
    for (int t = 0; t < K; t++) e[t] = (t == k) ? 1.0 : 0.0;
    lam[k] = neville_zero(h, e, K);
  }
  const int half = N / 2;
  for (int t = 0; t <= half; t++) {
    const int j = half + t;
    double c = 0.0;
    for (int k = 0; k < K; k++) {
      const int stride = N >> (levels - K + k);
      if (j % stride) continue;
      const double panel = stride * hh;
      c += lam[k] * ((j == 0 || j == N) ? 0.5 * panel : panel);
    }
    pw[t] = t == 0 ? c : 2.0 * c;
  }
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

namespace scft {

// Second derivatives M of the not-a-knot cubic spline through (x, y), the
// interpolant of DEALII_SCFT/src/spline_chen.c:13-107 (m == NULL).  The
// not-a-knot rows (third derivative continuous across x_1 and x_{n-2}) give
//   M_0     = (1 + h_0/h_1) M_1 - (h_0/h_1) M_2,
//   M_{n-1} = (1 + h_{n-2}/h_{n-3}) M_{n-2} - (h_{n-2}/h_{n-3}) M_{n-3};
// substituted into the first and last interior rows they leave a
// diagonally dominant tridiagonal system in M_1 .. M_{n-2}, solved by
// Thomas.  Same solution as the reference's dense Gauss-Jordan solve.
bool spline_not_a_knot(int n, const double* x, const double* y, std::vector<double>& M) {
  if (n < 4) return false;
  for (int i = 0; i + 1 < n; i++)
    if (!(x[i + 1] > x[i])) return false;
  const int k = n - 2;  // unknowns M_1 .. M_{n-2}
  std::vector<double> lo(k), di(k), up(k), rhs(k);
  for (int i = 1; i <= n - 2; i++) {
    const double hl = x[i] - x[i - 1], hr = x[i + 1] - x[i];
    lo[i - 1] = hl / 6.0;
    di[i - 1] = (hl + hr) / 3.0;
    up[i - 1] = hr / 6.0;
    rhs[i - 1] = (y[i + 1] - y[i]) / hr - (y[i] - y[i - 1]) / hl;
  }
  const double h0 = x[1] - x[0], h1 = x[2] - x[1];
  const double ha = x[n - 2] - x[n - 3], hb = x[n - 1] - x[n - 2];
  // row of M_1: lo * M_0 folded in
  di[0] += lo[0] * (1.0 + h0 / h1);
  if (k > 1) up[0] -= lo[0] * (h0 / h1);
  lo[0] = 0.0;
  // row of M_{n-2}: up * M_{n-1} folded in
  di[k - 1] += up[k - 1] * (1.0 + hb / ha);
  if (k > 1) lo[k - 1] -= up[k - 1] * (hb / ha);
  up[k - 1] = 0.0;
  // Thomas
  for (int i = 1; i < k; i++) {
    const double w = lo[i] / di[i - 1];
    di[i] -= w * up[i - 1];
    rhs[i] -= w * rhs[i - 1];
  }
  std::vector<double> m(k);
  m[k - 1] = rhs[k - 1] / di[k - 1];
  for (int i = k - 2; i >= 0; i--) m[i] = (rhs[i] - up[i] * m[i + 1]) / di[i];
  M.assign(n, 0.0);
  for (int i = 0; i < k; i++) M[i + 1] = m[i];
  M[0] = (1.0 + h0 / h1) * M[1] - (h0 / h1) * M[2];
  M[n - 1] = (1.0 + hb / ha) * M[n - 2] - (hb / ha) * M[n - 3];
  return true;
}

// Second derivatives of the natural cubic spline (M_0 = M_{n-1} = 0, the
// reference's m = 0 case, spline_chen.c:60-65): the interior rows alone, a
// diagonally dominant tridiagonal system solved by Thomas.
bool spline_natural(int n, const double* x, const double* y, std::vector<double>& M) {
  if (n < 2) return false;
  for (int i = 0; i + 1 < n; i++)
    if (!(x[i + 1] > x[i])) return false;
  M.assign(n, 0.0);
  const int k = n - 2;
  if (k < 1) return true;
  std::vector<double> lo(k), di(k), up(k), rhs(k);
  for (int i = 1; i <= n - 2; i++) {
    const double hl = x[i] - x[i - 1], hr = x[i + 1] - x[i];
    lo[i - 1] = hl / 6.0;
    di[i - 1] = (hl + hr) / 3.0;
    up[i - 1] = hr / 6.0;
    rhs[i - 1] = (y[i + 1] - y[i]) / hr - (y[i] - y[i - 1]) / hl;
  }
  for (int i = 1; i < k; i++) {
    const double w = lo[i] / di[i - 1];
    di[i] -= w * up[i - 1];
    rhs[i] -= w * rhs[i - 1];
  }
  M[k] = rhs[k - 1] / di[k - 1];
  for (int i = k - 2; i >= 0; i--) M[i + 1] = (rhs[i] - up[i] * M[i + 2]) / di[i];
  return true;
}

// Field at the walls as the deal.II evaluator sets it before assembling
// (DEALII_SCFT/src/scft.cc:454-480): the natural spline through the interior
// nodes 1..M-2, evaluated (extrapolated) at x_0 and x_{M-1}.
bool wall_values_natural(int M, const double* x, const double* w_interior, double* wall0,
                         double* wallL) {
  if (M < 4) return false;
  std::vector<double> Mi;
  if (!spline_natural(M - 2, x + 1, w_interior, Mi)) return false;
  *wall0 = spline_eval(M - 2, x + 1, w_interior, Mi.data(), x[0]);
  *wallL = spline_eval(M - 2, x + 1, w_interior, Mi.data(), x[M - 1]);
  return true;
}

// The same for a field given on the unknowns of either boundary problem:
// DD, nodes 1..M-2 (both walls extrapolated); DN, nodes 1..M-1 (x = 0
// extrapolated from the natural spline through them, *wallL the field at
// x = L itself, which is an unknown there).
bool wall_values_from_unknowns(int M, const double* x, const double* w_unknowns, bool dd,
                               double* wall0, double* wallL) {
  if (dd) return wall_values_natural(M, x, w_unknowns, wall0, wallL);
  if (M < 3) return false;
  std::vector<double> Mi;
  if (!spline_natural(M - 1, x + 1, w_unknowns, Mi)) return false;
  *wall0 = spline_eval(M - 1, x + 1, w_unknowns, Mi.data(), x[0]);
  *wallL = w_unknowns[M - 2];
  return true;
}

// Value at xp of the spline (x, y, M): bisection for the bracketing
// interval (the end intervals outside the table) and the cubic in the
// form of spline_chen.c:88-100.
double spline_eval(int n, const double* x, const double* y, const double* M, double xp) {
  int klo = 0, khi = n - 1;
  while (khi - klo > 1) {
    const int k = (khi + klo) >> 1;
    if (x[k] > xp)
      khi = k;
    else
      klo = k;
  }
  const double h = x[khi] - x[klo];
  const double a = (x[khi] - xp) / h, b = (xp - x[klo]) / h;
  return a * y[klo] + b * y[khi] +
         ((a * a * a - a) * M[klo] + (b * b * b - b) * M[khi]) * (h * h) / 6.0;
}

}  // namespace scft

namespace scft {

// Cells flagged by the gradient rule of Matlab_files/refine_mesh.m:1-36:
// e_c = |w_{c+1} - w_c| / (x_{c+1} - x_c) over the cells of the node field
// (cells touching a Dirichlet wall count as +inf, as the reference pads the
// interior solution with Inf; the far end is a wall only for DD), refined
// when e_c >= factor * median(e).  Returns the new nodes (flagged cells
// bisected).
std::vector<double> refine_by_gradient(int M, const double* x, const double* w_nodes,
                                       double factor, bool wall_right) {
  std::vector<double> e(M - 1);
  for (int c = 0; c < M - 1; c++) {
    const bool wall = (c == 0 || (wall_right && c == M - 2));
    e[c] = wall ? INFINITY : std::fabs((w_nodes[c + 1] - w_nodes[c]) / (x[c + 1] - x[c]));
  }
  std::vector<double> s(e);
  std::sort(s.begin(), s.end());
  // MATLAB median: mean of the two middle values for an even count
  const size_t k = s.size();
  const double med = (k % 2) ? s[k / 2] : 0.5 * (s[k / 2 - 1] + s[k / 2]);
  const double threshold = factor * med;
  std::vector<double> xn;
  xn.reserve(2 * M);
  for (int c = 0; c < M - 1; c++) {
    xn.push_back(x[c]);
    if (e[c] >= threshold) xn.push_back(0.5 * (x[c] + x[c + 1]));
  }
  xn.push_back(x[M - 1]);
  return xn;
}

// Mean-field free energy of DEALII_SCFT/src/scft.cc:404-450
// (set_mean_field_free_energy_romint): the field, linear between its nodes
// (the finite-element field sampled at y = 0), times the position-based
// target (mirrored at the far wall when `mirror`), on 2^k + 1 uniform points
// of [0, L]; Romberg (romint) over them; F = (I / f0bar / L + ln f0bar) / -1000.
bool free_energy(int M, const double* x, const double* w_nodes, double L, double tau, bool mirror,
                 int k, double f0bar, double* F) {
  const int m = 1 << k;
  std::vector<double> c;
  if (!romberg_weights(m, L / m, c)) return false;
  std::vector<double> xp(m + 1), ft(m + 1);
  for (int i = 0; i <= m; i++) xp[i] = L * i / m;
  target_position(m + 1, xp.data(), tau, mirror, ft.data());
  double I = 0.0;
  int seg = 0;
  for (int i = 0; i <= m; i++) {
    const double xi = xp[i];
    while (seg + 2 < M && x[seg + 1] < xi) seg++;
    const double t = (xi - x[seg]) / (x[seg + 1] - x[seg]);
    const double wv = w_nodes[seg] + t * (w_nodes[seg + 1] - w_nodes[seg]);
    I += c[i] * wv * ft[i];
  }
  *F = (I / f0bar / L + std::log(f0bar)) / (-1000.0);
  return true;
}

}  // namespace scft
