// broyden.cpp -- Broyden's method with a line search for one SCFT instance,
// F(w) = phi[w] - phi_t, every finite-difference Jacobian evaluated as ONE
// batched sweep of the device march.
//
// This is the reference's outer solver (broydn.c:44-292 with
// DEALII_SCFT/src/{fdjac,lnsrch,qrdcmp,qrupdt,rotate,rsolv}.c), written from
// the published method:
//   J. E. Dennis and R. B. Schnabel, Numerical Methods for Unconstrained
//   Optimization and Nonlinear Equations (Prentice-Hall 1983; SIAM 1996).
//   * secant update J+ = J + (y - J s) s^T / s^T s of a factored J = Q R
//     (sec. 8.3), done as a rank-one change of the factors with Jacobi
//     rotations (sec. 3.4, Alg. A3.4.1); components of y - J s below the
//     noise of F are dropped first;
//   * a fresh forward-difference Jacobian (sec. 5.4) factored by Householder
//     reflections;
//   * backtracking on f = |F|_2^2 / 2 (sec. 6.3, Alg. A6.3.1): the full
//     step capped at maxstep, then a quadratic and afterwards cubic model of
//     f(lambda), each new lambda kept in [0.1, 0.5] times the last one,
//     accepted at sufficient decrease f+ <= f + alpha lambda g^T p;
//   * the restart strategy of the reference's driver: when the search stalls
//     after secant updates take a new Jacobian, when it stalls on a fresh one
//     stop; stop on |F|_inf < tolf, on a vanishing scaled gradient of f (a
//     spurious minimum) and on a vanishing step.
// The reference's constants are kept so that evaluation counts compare:
// 400 iterations, noise floor 1e-14 for the update, step cap
// 100 max(|w|_2, n), difference step 1e-7 |w_j| (at least 1e-7), sufficient
// decrease 1e-4 and relative step floor 1e-7 in the search, the caller's tolf
// as both the |F| tolerance and the gradient tolerance.
//
// Where the evaluations run: the n perturbed fields of a Jacobian
// (fdjac.c:18-34, one call after another in the reference) are instances
// 1..n of one (n+1)-instance batch whose instance 0 is the unperturbed field,
// and each column is differenced against instance 0 of the same launch: the
// same kernel variant gives the same rounding, while an F(w) from another
// launch shape can differ by ~1e-13, which divided by steps down to 1e-7|w_j|
// would spoil the columns.  Single evaluations use a one-instance batch.
// With the consistent w-mass the wall values are extrapolated from each
// field by a natural spline before it is evaluated (DEALII_SCFT/src/scft.cc:
// 454-480), the perturbed fields included.  The dense n x n algebra is host
// code, O(n^3) per Jacobian, meant for n up to ~1000.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/scft.h"
#include "host_math.h"

namespace {

constexpr int kMaxIterations = 400;
constexpr double kUpdateNoise = 1e-14;    // relative size below which y - J s is noise
constexpr double kStepVanishes = 1e-14;   // relative step that ends the iteration
constexpr double kMaxStepScale = 100.0;   // maxstep = 100 max(|w|_2, n)
constexpr double kDiffStep = 1e-7;        // forward-difference step, relative
constexpr double kSufficientDecrease = 1e-4;
constexpr double kSearchStepFloor = 1e-7; // smallest relative step of a search

double dot(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

double max_abs(const double* a, int n) {
  double m = 0.0;
  for (int i = 0; i < n; i++) m = std::max(m, std::fabs(a[i]));
  return m;
}

// ---------------------------------------------------------------------------
// Residual sources: where F and the difference Jacobian come from.
class Residual {
 public:
  virtual ~Residual() = default;
  // F(w) into F; false when the evaluation failed (NaN, pivot, device error)
  virtual bool eval(const double* w, double* F) = 0;
  // forward-difference Jacobian at w, where F(w) = F (row-major,
  // J[i * n + j] = dF_i/dw_j)
  virtual bool jacobian(const double* w, const double* F, double* J) = 0;
  long evaluations = 0;
  long sweeps = 0;
};

// the perturbed value of a component wj for its difference column; *h
// receives the step that value really carries (wp - wj after rounding)
double perturbed(double wj, double* h) {
  double step = kDiffStep * std::fabs(wj);
  if (step < kDiffStep) step = kDiffStep;
  const double wp = wj + (wj >= 0.0 ? step : -step);
  *h = wp - wj;
  return wp;
}

class DeviceResidual final : public Residual {
 public:
  ~DeviceResidual() override {
    if (single_) scft_batch_destroy(single_);
    if (columns_) scft_batch_destroy(columns_);
  }

  int create(int M, int N, const double* x, const double* phi_t, int bc, int cmat) {
    M_ = M;
    bc_ = bc;
    consistent_ = cmat == SCFT_C_CONSISTENT;
    n_ = bc == SCFT_BC_DD ? M - 2 : M - 1;
    x_.assign(x, x + M);
    scft_batch_desc d{};
    long long zero = 0;
    d.n_inst = 1;
    d.M = &M;
    d.N = &N;
    d.xoff = &zero;
    d.x = x;
    d.phi_t = phi_t;
    d.bc = bc;
    d.cmat = cmat;
    if (int rc = scft_batch_create(&single_, &d)) return rc;
    const int copies = n_ + 1;
    std::vector<int> Ms(copies, M), Ns(copies, N);
    std::vector<long long> xoff(copies);
    std::vector<double> xs, pts;
    xs.reserve(static_cast<size_t>(copies) * M);
    pts.reserve(static_cast<size_t>(copies) * n_);
    for (int c = 0; c < copies; c++) {
      xoff[c] = static_cast<long long>(c) * M;
      xs.insert(xs.end(), x, x + M);
      pts.insert(pts.end(), phi_t, phi_t + n_);
    }
    d.n_inst = copies;
    d.M = Ms.data();
    d.N = Ns.data();
    d.xoff = xoff.data();
    d.x = xs.data();
    d.phi_t = pts.data();
    if (int rc = scft_batch_create(&columns_, &d)) return rc;
    fields_.resize(static_cast<size_t>(copies) * n_);
    values_.resize(static_cast<size_t>(copies) * n_);
    rnorm_.resize(copies);
    status_.resize(copies);
    wall0_.resize(copies);
    wallL_.resize(copies);
    return SCFT_OK;
  }

  bool eval(const double* w, double* F) override {
    evaluations++;
    sweeps++;
    if (consistent_ && !set_walls(single_, w, 1)) return false;
    double rn = 0.0;
    unsigned st = 0;
    if (scft_batch_eval(single_, w, nullptr, F, &rn, &st) != SCFT_OK) return false;
    return !(st & (SCFT_ST_NAN | SCFT_ST_PIVOT)) && std::isfinite(rn);
  }

  // F is not used: the columns are differenced against the base field
  // evaluated in the same launch (see the file comment)
  bool jacobian(const double* w, const double*, double* J) override {
    const int n = n_;
    std::vector<double> h(n);
    for (int c = 0; c <= n; c++) std::copy(w, w + n, fields_.begin() + static_cast<size_t>(c) * n);
    for (int j = 0; j < n; j++) {
      fields_[static_cast<size_t>(j + 1) * n + j] = perturbed(w[j], &h[j]);
    }
    evaluations += n;
    sweeps++;
    if (consistent_ && !set_walls(columns_, fields_.data(), n + 1)) return false;
    if (scft_batch_eval(columns_, fields_.data(), nullptr, values_.data(), rnorm_.data(),
                        status_.data()) != SCFT_OK)
      return false;
    for (int c = 0; c <= n; c++)
      if (status_[c] & (SCFT_ST_NAN | SCFT_ST_PIVOT)) return false;
    const double* base = values_.data();
    for (int j = 0; j < n; j++) {
      const double* col = values_.data() + static_cast<size_t>(j + 1) * n;
      for (int i = 0; i < n; i++) J[static_cast<size_t>(i) * n + j] = (col[i] - base[i]) / h[j];
    }
    return true;
  }

 private:
  // wall values of `count` fields (consecutive, n each) by natural-spline
  // extrapolation of the unknowns, as the deal.II evaluator assembles
  bool set_walls(scft_batch* b, const double* w, int count) {
    for (int c = 0; c < count; c++) {
      double w0 = 0.0, wl = 0.0;
      if (!scft::wall_values_from_unknowns(M_, x_.data(), w + static_cast<size_t>(c) * n_,
                                           bc_ == SCFT_BC_DD, &w0, &wl))
        return false;
      wall0_[c] = w0;
      wallL_[c] = wl;
    }
    return scft_batch_set_walls(b, wall0_.data(), wallL_.data()) == SCFT_OK;
  }

  int M_ = 0, n_ = 0, bc_ = SCFT_BC_DD;
  bool consistent_ = false;
  std::vector<double> x_;
  scft_batch* single_ = nullptr;   // F at one field
  scft_batch* columns_ = nullptr;  // the base field and its n perturbations
  std::vector<double> fields_, values_, rnorm_, wall0_, wallL_;
  std::vector<unsigned> status_;
};

// F from a caller's function, the Jacobian column by column.
class CallbackResidual final : public Residual {
 public:
  CallbackResidual(int n, int (*fn)(int, const double*, double*, void*), void* ctx)
      : n_(n), fn_(fn), ctx_(ctx), base_(n), col_(n), wp_(n) {}

  bool eval(const double* w, double* F) override {
    evaluations++;
    sweeps++;
    if (fn_(n_, w, F, ctx_) != 0) return false;
    for (int i = 0; i < n_; i++)
      if (!std::isfinite(F[i])) return false;
    return true;
  }

  bool jacobian(const double* w, const double* F, double* J) override {
    const int n = n_;
    std::copy(F, F + n, base_.begin());
    std::copy(w, w + n, wp_.begin());
    for (int j = 0; j < n; j++) {
      double h;
      wp_[j] = perturbed(w[j], &h);
      const bool ok = eval(wp_.data(), col_.data());
      wp_[j] = w[j];
      if (!ok) return false;
      for (int i = 0; i < n; i++) J[static_cast<size_t>(i) * n + j] = (col_[i] - base_[i]) / h;
    }
    return true;
  }

 private:
  int n_;
  int (*fn_)(int, const double*, double*, void*);
  void* ctx_;
  std::vector<double> base_, col_, wp_;
};

// ---------------------------------------------------------------------------
// J = Q R: Q orthogonal (row-major), R upper triangular (row-major).
class QRFactors {
 public:
  explicit QRFactors(int n)
      : n_(n), q_(static_cast<size_t>(n) * n), r_(static_cast<size_t>(n) * n), t_(n) {}

  // Householder reflections H_k = I - (2 / v^T v) v v^T turn J into R; Q is
  // their product.  False when a column has nothing left (R singular).
  bool factor(std::vector<double>& J) {
    const int n = n_;
    std::fill(q_.begin(), q_.end(), 0.0);
    for (int i = 0; i < n; i++) Q(i, i) = 1.0;
    std::vector<double> v(n);
    bool singular = false;
    for (int k = 0; k < n; k++) {
      const int len = n - k;
      double big = 0.0;
      for (int i = k; i < n; i++) big = std::max(big, std::fabs(J[idx(i, k)]));
      if (big == 0.0) {
        singular = true;
        continue;
      }
      double ss = 0.0;
      for (int i = k; i < n; i++) {
        const double y = J[idx(i, k)] / big;
        ss += y * y;
      }
      const double x0 = J[idx(k, k)];
      const double alpha = (x0 > 0.0 ? -1.0 : 1.0) * big * std::sqrt(ss);  // R(k,k)
      if (k == n - 1) {
        J[idx(k, k)] = x0;  // nothing below the diagonal: no reflection
        break;
      }
      for (int i = 0; i < len; i++) v[i] = J[idx(k + i, k)];
      v[0] -= alpha;
      const double vtv = 2.0 * alpha * (alpha - x0);  // |x - alpha e1|^2
      if (vtv == 0.0) continue;
      const double beta = 2.0 / vtv;
      for (int j = k + 1; j < n; j++) {
        double s = 0.0;
        for (int i = 0; i < len; i++) s += v[i] * J[idx(k + i, j)];
        s *= beta;
        for (int i = 0; i < len; i++) J[idx(k + i, j)] -= s * v[i];
      }
      J[idx(k, k)] = alpha;
      for (int i = k + 1; i < n; i++) J[idx(i, k)] = 0.0;
      for (int row = 0; row < n; row++) {  // Q <- Q H_k
        double* qr = &Q(row, k);
        double s = 0.0;
        for (int i = 0; i < len; i++) s += qr[i] * v[i];
        s *= beta;
        for (int i = 0; i < len; i++) qr[i] -= s * v[i];
      }
    }
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) R(i, j) = j >= i ? J[idx(i, j)] : 0.0;
    for (int i = 0; i < n && !singular; i++) singular = R(i, i) == 0.0;
    return !singular;
  }

  // out = Q^T v
  void qt_times(const double* v, double* out) const {
    std::fill(out, out + n_, 0.0);
    for (int i = 0; i < n_; i++) {
      const double vi = v[i];
      const double* qi = &q_[idx(i, 0)];
      for (int j = 0; j < n_; j++) out[j] += qi[j] * vi;
    }
  }

  // out = J v = Q (R v)
  void j_times(const double* v, double* out) {
    for (int i = 0; i < n_; i++) {
      double s = 0.0;
      for (int j = i; j < n_; j++) s += R(i, j) * v[j];
      t_[i] = s;
    }
    for (int i = 0; i < n_; i++) out[i] = dot(&q_[idx(i, 0)], t_.data(), n_);
  }

  // out = J^T F = R^T (Q^T F)
  void jt_times(const double* F, double* out) {
    qt_times(F, t_.data());
    for (int j = 0; j < n_; j++) {
      double s = 0.0;
      for (int i = 0; i <= j; i++) s += R(i, j) * t_[i];
      out[j] = s;
    }
  }

  // x = J^{-1} b = R^{-1} Q^T b
  void solve(const double* b, double* x) {
    qt_times(b, t_.data());
    for (int i = n_ - 1; i >= 0; i--) {
      double s = t_[i];
      for (int j = i + 1; j < n_; j++) s -= R(i, j) * x[j];
      x[i] = s / R(i, i);
    }
  }

  // J <- J + a b^T.  With u = Q^T a, rotations from the bottom fold u into
  // its first entry (R turns upper Hessenberg), the first row of R takes
  // u_0 b^T, and rotations from the top restore the triangle.  Every
  // rotation of two rows of R is matched on the same two columns of Q.
  // False if R ends with a zero diagonal entry.
  bool rank_one(const double* a, const double* b) {
    const int n = n_;
    std::vector<double> u(n);
    qt_times(a, u.data());
    int last = n - 1;
    while (last > 0 && u[last] == 0.0) last--;
    for (int i = last - 1; i >= 0; i--) {
      const double rho = std::hypot(u[i], u[i + 1]);
      if (rho == 0.0) continue;
      rotate(i, i, u[i] / rho, u[i + 1] / rho);
      u[i] = rho;
      u[i + 1] = 0.0;
    }
    for (int j = 0; j < n; j++) R(0, j) += u[0] * b[j];
    for (int i = 0; i < last; i++) {
      const double rho = std::hypot(R(i, i), R(i + 1, i));
      if (rho == 0.0) continue;
      rotate(i, i, R(i, i) / rho, R(i + 1, i) / rho);
      R(i + 1, i) = 0.0;
    }
    for (int i = 0; i < n; i++)
      if (R(i, i) == 0.0) return false;
    return true;
  }

 private:
  size_t idx(int i, int j) const { return static_cast<size_t>(i) * n_ + j; }
  double& Q(int i, int j) { return q_[idx(i, j)]; }
  double& R(int i, int j) { return r_[idx(i, j)]; }

  // rows i, i+1 of R (columns from `from`) <- G (rows), columns i, i+1 of
  // Q <- (columns) G^T, with G = [c s; -s c]: Q R is unchanged
  void rotate(int i, int from, double c, double s) {
    double* ra = &R(i, 0);
    double* rb = &R(i + 1, 0);
    for (int j = from; j < n_; j++) {
      const double p = ra[j], q = rb[j];
      ra[j] = c * p + s * q;
      rb[j] = c * q - s * p;
    }
    for (int row = 0; row < n_; row++) {
      double* qr = &Q(row, i);
      const double p = qr[0], q = qr[1];
      qr[0] = c * p + s * q;
      qr[1] = c * q - s * p;
    }
  }

  int n_;
  std::vector<double> q_, r_, t_;
};

// ---------------------------------------------------------------------------
enum class Search { kAccepted, kStalled, kFailed, kUphill };

struct Iterate {
  std::vector<double> w, F;
  double f = 0.0;  // |F|_2^2 / 2
};

// Backtracking along p from `cur` (Alg. A6.3.1); `next` receives the new
// iterate, or a copy of `cur` when the search stalls.
Search line_search(Residual& res, int n, const Iterate& cur, const double* g,
                   std::vector<double>& p, double maxstep, Iterate& next) {
  const double len = std::sqrt(dot(p.data(), p.data(), n));
  if (len > maxstep)
    for (int i = 0; i < n; i++) p[i] *= maxstep / len;
  const double slope = dot(g, p.data(), n);
  if (slope >= 0.0) return Search::kUphill;
  double rel = 0.0;
  for (int i = 0; i < n; i++)
    rel = std::max(rel, std::fabs(p[i]) / std::max(std::fabs(cur.w[i]), 1.0));
  const double lambda_min = kSearchStepFloor / rel;
  double lambda = 1.0, lambda_prev = 0.0, f_prev = 0.0;
  for (;;) {
    for (int i = 0; i < n; i++) next.w[i] = cur.w[i] + lambda * p[i];
    if (!res.eval(next.w.data(), next.F.data())) return Search::kFailed;
    next.f = 0.5 * dot(next.F.data(), next.F.data(), n);
    if (next.f <= cur.f + kSufficientDecrease * lambda * slope) return Search::kAccepted;
    if (lambda < lambda_min) {
      next = cur;
      return Search::kStalled;
    }
    double trial;
    if (lambda == 1.0) {  // minimiser of the quadratic through f(0), f'(0), f(1)
      trial = -slope / (2.0 * (next.f - cur.f - slope));
    } else {              // of the cubic through the last two trials
      const double e1 = (next.f - cur.f - lambda * slope) / (lambda * lambda);
      const double e2 = (f_prev - cur.f - lambda_prev * slope) / (lambda_prev * lambda_prev);
      const double a = (e1 - e2) / (lambda - lambda_prev);
      const double b = (lambda * e2 - lambda_prev * e1) / (lambda - lambda_prev);
      if (a == 0.0) {
        trial = -slope / (2.0 * b);
      } else {
        const double disc = b * b - 3.0 * a * slope;
        trial = disc < 0.0 ? 0.5 * lambda : (-b + std::sqrt(disc)) / (3.0 * a);
      }
      trial = std::min(trial, 0.5 * lambda);
    }
    lambda_prev = lambda;
    f_prev = next.f;
    lambda = std::max(trial, 0.1 * lambda);
  }
}

struct Outcome {
  int rc = SCFT_E_NOCONV;
  int check = 1;
  int iterations = 0;
  double final_err = 0.0;
};

Outcome broyden(Residual& res, int n, double* w, double tolf) {
  Outcome out;
  Iterate cur, next;
  cur.w.assign(w, w + n);
  cur.F.resize(n);
  next.w.resize(n);
  next.F.resize(n);
  auto finish = [&](int rc, int check) {
    out.rc = rc;
    out.check = check;
    out.final_err = max_abs(cur.F.data(), n);
    std::copy(cur.w.begin(), cur.w.end(), w);
    return out;
  };
  if (!res.eval(cur.w.data(), cur.F.data())) return finish(SCFT_E_NAN, 1);
  cur.f = 0.5 * dot(cur.F.data(), cur.F.data(), n);
  if (max_abs(cur.F.data(), n) < tolf) return finish(SCFT_OK, 0);
  const double maxstep =
      kMaxStepScale * std::max(std::sqrt(dot(cur.w.data(), cur.w.data(), n)), double(n));

  QRFactors qr(n);
  std::vector<double> J(static_cast<size_t>(n) * n), g(n), p(n), s(n), y(n);
  bool fresh = true;  // the next iteration takes a new Jacobian
  bool have_step = false;
  for (int it = 1; it <= kMaxIterations; it++) {
    out.iterations = it;
    if (fresh) {
      if (!res.jacobian(cur.w.data(), cur.F.data(), J.data())) return finish(SCFT_E_NAN, 1);
      if (!qr.factor(J)) return finish(SCFT_E_PIVOT, 1);
    } else if (have_step) {
      // secant condition J+ s = y: J += (y - J s) s^T / s^T s
      qr.j_times(s.data(), p.data());
      bool any = false;
      for (int i = 0; i < n; i++) {
        double d = y[i] - p[i];
        if (std::fabs(d) < kUpdateNoise * (std::fabs(cur.F[i]) + std::fabs(cur.F[i] - y[i])))
          d = 0.0;
        y[i] = d;
        any = any || d != 0.0;
      }
      if (any) {
        const double ss = dot(s.data(), s.data(), n);
        for (int i = 0; i < n; i++) s[i] /= ss;
        if (!qr.rank_one(y.data(), s.data())) return finish(SCFT_E_PIVOT, 1);
      }
    }
    qr.jt_times(cur.F.data(), g.data());  // gradient of f
    qr.solve(cur.F.data(), p.data());     // Newton step -J^{-1} F
    for (int i = 0; i < n; i++) p[i] = -p[i];
    const Search sr = line_search(res, n, cur, g.data(), p, maxstep, next);
    if (sr == Search::kFailed) return finish(SCFT_E_NAN, 1);
    if (sr == Search::kUphill) return finish(SCFT_E_NOCONV, 1);
    for (int i = 0; i < n; i++) {
      s[i] = next.w[i] - cur.w[i];
      y[i] = next.F[i] - cur.F[i];
    }
    have_step = true;
    double step = 0.0;
    for (int i = 0; i < n; i++)
      step = std::max(step, std::fabs(s[i]) / std::max(std::fabs(next.w[i]), 1.0));
    std::swap(cur, next);
    if (max_abs(cur.F.data(), n) < tolf) return finish(SCFT_OK, 0);
    if (sr == Search::kStalled) {
      if (fresh) return finish(SCFT_E_NOCONV, 1);  // no progress on a fresh Jacobian
      // a minimum of f that is not a root: the scaled gradient vanishes
      const double den = std::max(cur.f, 0.5 * n);
      double grad = 0.0;
      for (int i = 0; i < n; i++)
        grad = std::max(grad, std::fabs(g[i]) * std::max(std::fabs(cur.w[i]), 1.0) / den);
      if (grad < tolf) return finish(SCFT_OK, 0);
      fresh = true;
      have_step = false;
    } else {
      fresh = false;
      if (step < kStepVanishes) return finish(SCFT_OK, 0);
    }
  }
  return finish(SCFT_E_NOCONV, 1);
}

void fill_stats(scft_broyden_stats* stats, const Outcome& o, const Residual& r) {
  if (!stats) return;
  stats->iterations = o.iterations;
  stats->evaluations = r.evaluations;
  stats->sweeps = r.sweeps;
  stats->final_err = o.final_err;
  stats->check = o.check;
}

}  // namespace

extern "C" int scft_broyden_solve(int M, int N, const double* x, const double* phi_t, int bc,
                                  int cmat, double tolf, double* w, scft_broyden_stats* stats) {
  if (M < 4 || !x || !phi_t || !w || !(tolf > 0.0)) return SCFT_E_ARG;
  if (bc != SCFT_BC_DD && bc != SCFT_BC_DN) return SCFT_E_ARG;
  if (cmat != SCFT_C_ROWSCALED && cmat != SCFT_C_CONSISTENT) return SCFT_E_ARG;
  const int n = bc == SCFT_BC_DD ? M - 2 : M - 1;
  if (n < 2 || n > 4096) return SCFT_E_ARG;
  DeviceResidual res;
  if (int rc = res.create(M, N, x, phi_t, bc, cmat)) return rc;
  const Outcome o = broyden(res, n, w, tolf);
  fill_stats(stats, o, res);
  return o.rc;
}

extern "C" int scft_broyden_solve_fn(int n, int (*F)(int, const double*, double*, void*),
                                     void* ctx, double tolf, double* w,
                                     scft_broyden_stats* stats) {
  if (n < 1 || !F || !w || !(tolf > 0.0)) return SCFT_E_ARG;
  CallbackResidual res(n, F, ctx);
  const Outcome o = broyden(res, n, w, tolf);
  fill_stats(stats, o, res);
  return o.rc;
}
