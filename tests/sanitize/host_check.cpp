// host_check.cpp -- drives the host C++ of libscft_hip.so (host_math.cpp and
// the Broyden iteration of broyden.cpp through its callback entry) under
// AddressSanitizer and UndefinedBehaviorSanitizer.  Built and run by
// tests/test_sanitizers.py with g++; exits nonzero on a failed check (a
// sanitizer report aborts the run by itself).  The device entry points that
// broyden.cpp also references are stubbed: nothing here reaches them.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/scft.h"
#include "../../scft_amd/csrc/host_math.h"

extern "C" {
int scft_batch_create(scft_batch**, const scft_batch_desc*) { return SCFT_E_STATE; }
int scft_batch_destroy(scft_batch*) { return SCFT_E_STATE; }
int scft_batch_eval(scft_batch*, const double*, double*, double*, double*, unsigned*) {
  return SCFT_E_STATE;
}
int scft_batch_set_walls(scft_batch*, const double*, const double*) { return SCFT_E_STATE; }
}

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
      failures++;                                                     \
    }                                                                 \
  } while (0)

static double lcg(unsigned long long* s) {
  *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
  return ((*s >> 11) + 0.5) / 9007199254740992.0;
}

// F(w) = A w - b on n unknowns (A diagonally dominant), and a discretised
// nonlinear two-point problem -w'' + exp(w)/10 = 1 with w(0) = w(1) = 0
struct Linear {
  int n;
  std::vector<double> A, b;
};
static int linear_fn(int n, const double* w, double* f, void* ctx) {
  const Linear* L = static_cast<const Linear*>(ctx);
  for (int i = 0; i < n; i++) {
    double s = -L->b[i];
    for (int j = 0; j < n; j++) s += L->A[i * n + j] * w[j];
    f[i] = s;
  }
  return 0;
}
static int bvp_fn(int n, const double* w, double* f, void*) {
  const double h = 1.0 / (n + 1);
  for (int i = 0; i < n; i++) {
    const double wl = i > 0 ? w[i - 1] : 0.0, wr = i + 1 < n ? w[i + 1] : 0.0;
    f[i] = (2.0 * w[i] - wl - wr) / (h * h) + std::exp(w[i]) / 10.0 - 1.0;
  }
  return 0;
}
static int failing_fn(int n, const double*, double* f, void*) {
  for (int i = 0; i < n; i++) f[i] = NAN;
  return 1;
}

int main() {
  // Romberg pairing weights: both forms, every supported N; sum c_j = 1
  for (int N = 16; N <= 16384; N *= 2) {
    std::vector<double> c, pw;
    CHECK(scft::romberg_weights(N, 1.0 / N, c));
    CHECK(scft::pairing_weights(N, pw));
    std::vector<double> pw2(N / 2 + 1);
    CHECK(scft::pairing_weights_into(N, pw2.data()));
    double s = 0.0;
    for (double v : c) s += v;
    CHECK(std::fabs(s - 1.0) < 1e-12);
    for (int t = 0; t <= N / 2; t++) CHECK(pw[t] == pw2[t]);
  }
  std::vector<double> bad(3);
  CHECK(!scft::romberg_weights(24, 1.0, bad));
  CHECK(!scft::pairing_weights_into(24, bad.data()));

  // mesh, target, splines, wall values
  unsigned long long seed = 7;
  for (int M : {4, 5, 17, 300}) {
    std::vector<double> x(M), y(M), pt(M);
    scft::mesh_uniform(M, 3.7, x.data());
    scft::target_position(M, x.data(), 0.53, true, pt.data());
    for (int i = 0; i < M; i++) y[i] = std::sin(2.0 * x[i]) + 0.01 * lcg(&seed);
    std::vector<double> m1, m2;
    CHECK(scft::spline_not_a_knot(M, x.data(), y.data(), m1));
    CHECK(scft::spline_natural(M, x.data(), y.data(), m2));
    for (int i = 0; i < M; i++) {
      CHECK(std::fabs(scft::spline_eval(M, x.data(), y.data(), m1.data(), x[i]) - y[i]) < 1e-9);
      CHECK(std::fabs(scft::spline_eval(M, x.data(), y.data(), m2.data(), x[i]) - y[i]) < 1e-9);
    }
    double w0 = 0.0, wl = 0.0;
    CHECK(scft::wall_values_natural(M, x.data(), y.data() + 1, &w0, &wl));
    CHECK(scft::wall_values_from_unknowns(M, x.data(), y.data() + 1, false, &w0, &wl));
    CHECK(std::isfinite(w0) && wl == y[M - 1]);
    std::vector<double> xr = scft::refine_by_gradient(M, x.data(), y.data(), 10.0, true);
    CHECK(static_cast<int>(xr.size()) >= M);
    double F = 0.0;
    CHECK(scft::free_energy(M, x.data(), y.data(), x[M - 1], 0.53, true, 12, 0.8925, &F));
    CHECK(std::isfinite(F));
  }

  // Broyden through the callback entry
  scft_broyden_stats st;
  {
    Linear L{12, std::vector<double>(144), std::vector<double>(12)};
    for (int i = 0; i < 12; i++) {
      L.b[i] = lcg(&seed) - 0.5;
      for (int j = 0; j < 12; j++) L.A[i * 12 + j] = (i == j ? 4.0 : 0.0) + 0.3 * (lcg(&seed) - 0.5);
    }
    std::vector<double> w(12, 0.0);
    CHECK(scft_broyden_solve_fn(12, linear_fn, &L, 1e-10, w.data(), &st) == SCFT_OK);
    CHECK(st.final_err < 1e-10 && st.iterations == 1);
  }
  {
    std::vector<double> w(60, 0.0);
    CHECK(scft_broyden_solve_fn(60, bvp_fn, nullptr, 1e-9, w.data(), &st) == SCFT_OK);
    CHECK(st.final_err < 1e-9);
    std::vector<double> f(60);
    bvp_fn(60, w.data(), f.data(), nullptr);
    double e = 0.0;
    for (double v : f) e = std::fmax(e, std::fabs(v));
    CHECK(e < 1e-9);
  }
  {
    std::vector<double> w(5, 0.0);
    CHECK(scft_broyden_solve_fn(5, failing_fn, nullptr, 1e-9, w.data(), &st) == SCFT_E_NAN);
  }
  CHECK(scft_broyden_solve_fn(0, bvp_fn, nullptr, 1e-9, nullptr, &st) == SCFT_E_ARG);
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("host_check: all checks passed\n");
  return 0;
}
