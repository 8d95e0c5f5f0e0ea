// broyden.cpp -- globally convergent Broyden on one SCFT instance with the
// finite-difference Jacobian evaluated as one batched sweep.
//
// The iteration is Numerical Recipes 9.7 as the reference ships it
// (broydn.c:44-292 with DEALII_SCFT/src/{fdjac,lnsrch,qrdcmp,qrupdt,rotate,
// rsolv}.c): Householder QR of a forward-difference Jacobian, Broyden rank-1
// updates of Q R by Givens rotations, a backtracking line search on
// f = |F|^2 / 2, Jacobian restarts when the search stalls, and the same
// constants (MAXITS 400, EPS 1e-14, TOLX = EPS, STPMX 100; fdjac EPS 1e-7,
// lnsrch ALF 1e-4, TOLX 1e-7).  What changes is where the evaluations run:
// fdjac's n perturbed evaluations (fdjac.c:18-34), one after another in the
// reference, are one evaluation of an (n+1)-instance batch here: instance 0
// is the unperturbed field and instance j the field with component j
// perturbed, and the differences are taken against instance 0 of the same
// launch (the same kernel variant, so the same rounding: against an F(x)
// from another launch configuration the ~1e-13 disagreement, divided by
// steps down to 1e-7 |x_j|, would corrupt the columns).  The other
// evaluations use a one-instance batch.  The dense n x n linear algebra
// stays on the host.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/scft.h"

namespace {

double sgn(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

struct Evaluators {
  int n = 0;
  scft_batch* one = nullptr;   // F at a single field
  scft_batch* jac = nullptr;   // n + 1 copies of the instance (fdjac)
  std::vector<double> r1, rn;
  long evaluations = 0, sweeps = 0;
  bool failed = false;

  ~Evaluators() {
    if (one) scft_batch_destroy(one);
    if (jac) scft_batch_destroy(jac);
  }

  int create(int M, int N, const double* x, const double* phi_t, int bc, int cmat, int n_) {
    n = n_;
    scft_batch_desc d{};
    long long off0 = 0;
    d.n_inst = 1;
    d.M = &M;
    d.N = &N;
    d.xoff = &off0;
    d.x = x;
    d.phi_t = phi_t;
    d.bc = bc;
    d.cmat = cmat;
    d.tune_file = nullptr;
    if (int rc = scft_batch_create(&one, &d)) return rc;
    const int nj = n + 1;
    std::vector<int> Ms(nj, M), Ns(nj, N);
    std::vector<long long> xoff(nj);
    std::vector<double> xs(static_cast<size_t>(nj) * M), pts(static_cast<size_t>(nj) * n);
    for (int j = 0; j < nj; j++) {
      xoff[j] = static_cast<long long>(j) * M;
      std::copy(x, x + M, xs.begin() + xoff[j]);
      std::copy(phi_t, phi_t + n, pts.begin() + static_cast<size_t>(j) * n);
    }
    d.n_inst = nj;
    d.M = Ms.data();
    d.N = Ns.data();
    d.xoff = xoff.data();
    d.x = xs.data();
    d.phi_t = pts.data();
    if (int rc = scft_batch_create(&jac, &d)) return rc;
    r1.resize(n);
    rn.resize(static_cast<size_t>(nj) * n);
    return SCFT_OK;
  }

  // F(x) into f; f = |F|^2 / 2 returned (0 after a failure, as fminbrd)
  double fmin(const double* xv, double* fv) {
    double rnorm = 0.0;
    unsigned st = 0;
    const int rc = scft_batch_eval(one, xv, nullptr, fv, &rnorm, &st);
    evaluations++;
    sweeps++;
    if (rc != SCFT_OK || (st & (SCFT_ST_NAN | SCFT_ST_PIVOT))) failed = true;
    if (failed) return 0.0;
    double s = 0.0;
    for (int i = 0; i < n; i++) s += fv[i] * fv[i];
    return 0.5 * s;
  }

  // forward-difference Jacobian (row-major df[i * n + j]) at x
  void fdjac(const double* xv, double* df) {
    const double EPS = 1.0e-7;
    const int nj = n + 1;
    std::vector<double> W(static_cast<size_t>(nj) * n), h(n);
    std::copy(xv, xv + n, W.begin());  // instance 0: the base point
    for (int j = 0; j < n; j++) {
      double* wj = W.data() + static_cast<size_t>(j + 1) * n;
      std::copy(xv, xv + n, wj);
      const double temp = xv[j];
      double hj = EPS * temp;
      if (std::fabs(hj) < EPS) hj = sgn(EPS, temp);
      wj[j] = temp + hj;
      h[j] = wj[j] - temp;  // the step actually representable
    }
    std::vector<double> rnorm(nj);
    std::vector<unsigned> st(nj);
    const int rc = scft_batch_eval(jac, W.data(), nullptr, rn.data(), rnorm.data(), st.data());
    evaluations += n;
    sweeps++;
    if (rc != SCFT_OK) failed = true;
    for (int j = 0; j < nj && !failed; j++)
      if (st[j] & (SCFT_ST_NAN | SCFT_ST_PIVOT)) failed = true;
    if (failed) return;
    const double* f0 = rn.data();
    for (int j = 0; j < n; j++) {
      const double* fj = rn.data() + static_cast<size_t>(j + 1) * n;
      for (int i = 0; i < n; i++) df[static_cast<size_t>(i) * n + j] = (fj[i] - f0[i]) / h[j];
    }
  }
};

// lnsrch.c:8-61; returns false on the reference's "Roundoff problem" exit
bool line_search(Evaluators& ev, int n, const double* xold, double fold, const double* g,
                 double* p, double* x, double* f, double* fvec, double stpmax, int* check) {
  const double ALF = 1.0e-4, TOLX = 1.0e-7;
  *check = 0;
  double sum = 0.0;
  for (int i = 0; i < n; i++) sum += p[i] * p[i];
  sum = std::sqrt(sum);
  if (sum > stpmax)
    for (int i = 0; i < n; i++) p[i] *= stpmax / sum;
  double slope = 0.0;
  for (int i = 0; i < n; i++) slope += g[i] * p[i];
  if (slope >= 0.0) return false;
  double test = 0.0;
  for (int i = 0; i < n; i++) test = std::max(test, std::fabs(p[i]) / std::max(std::fabs(xold[i]), 1.0));
  const double alamin = TOLX / test;
  double alam = 1.0, alam2 = 0.0, f2 = 0.0, tmplam = 0.0;
  for (;;) {
    for (int i = 0; i < n; i++) x[i] = xold[i] + alam * p[i];
    *f = ev.fmin(x, fvec);
    if (alam < alamin) {
      for (int i = 0; i < n; i++) x[i] = xold[i];
      *check = 1;
      return true;
    }
    if (*f <= fold + ALF * alam * slope) return true;
    if (alam == 1.0) {
      tmplam = -slope / (2.0 * (*f - fold - slope));
    } else {
      const double rhs1 = *f - fold - alam * slope;
      const double rhs2 = f2 - fold - alam2 * slope;
      const double a = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2);
      const double b =
          (-alam2 * rhs1 / (alam * alam) + alam * rhs2 / (alam2 * alam2)) / (alam - alam2);
      if (a == 0.0) {
        tmplam = -slope / (2.0 * b);
      } else {
        const double disc = b * b - 3.0 * a * slope;
        if (disc < 0.0)
          tmplam = 0.5 * alam;
        else if (b <= 0.0)
          tmplam = (-b + std::sqrt(disc)) / (3.0 * a);
        else
          tmplam = -slope / (b + std::sqrt(disc));
      }
      tmplam = std::min(tmplam, 0.5 * alam);
    }
    alam2 = alam;
    f2 = *f;
    alam = std::max(tmplam, 0.1 * alam);
  }
}

// Householder QR with scaled columns (qrdcmp.c:6-34); R's diagonal in d,
// the reflectors below it in a, c[k] their normalisers
bool qr_decompose(std::vector<double>& a, int n, std::vector<double>& c, std::vector<double>& d) {
  bool sing = false;
  auto A = [&](int i, int j) -> double& { return a[static_cast<size_t>(i) * n + j]; };
  for (int k = 0; k < n - 1; k++) {
    double scale = 0.0;
    for (int i = k; i < n; i++) scale = std::max(scale, std::fabs(A(i, k)));
    if (scale == 0.0) {
      sing = true;
      c[k] = d[k] = 0.0;
      continue;
    }
    double sum = 0.0;
    for (int i = k; i < n; i++) {
      A(i, k) /= scale;
      sum += A(i, k) * A(i, k);
    }
    const double sigma = sgn(std::sqrt(sum), A(k, k));
    A(k, k) += sigma;
    c[k] = sigma * A(k, k);
    d[k] = -scale * sigma;
    for (int j = k + 1; j < n; j++) {
      double s2 = 0.0;
      for (int i = k; i < n; i++) s2 += A(i, k) * A(i, j);
      const double tau = s2 / c[k];
      for (int i = k; i < n; i++) A(i, j) -= tau * A(i, k);
    }
  }
  d[n - 1] = A(n - 1, n - 1);
  return sing || d[n - 1] == 0.0;
}

// Givens rotation of rows i, i+1 of r and qt (rotate.c:6-35)
void givens(std::vector<double>& r, std::vector<double>& qt, int n, int i, double a, double b) {
  double c, s;
  if (a == 0.0) {
    c = 0.0;
    s = b >= 0.0 ? 1.0 : -1.0;
  } else if (std::fabs(a) > std::fabs(b)) {
    const double fact = b / a;
    c = sgn(1.0 / std::sqrt(1.0 + fact * fact), a);
    s = fact * c;
  } else {
    const double fact = a / b;
    s = sgn(1.0 / std::sqrt(1.0 + fact * fact), b);
    c = fact * s;
  }
  double* ri = r.data() + static_cast<size_t>(i) * n;
  double* rj = ri + n;
  for (int j = i; j < n; j++) {
    const double y = ri[j], w = rj[j];
    ri[j] = c * y - s * w;
    rj[j] = s * y + c * w;
  }
  double* qi = qt.data() + static_cast<size_t>(i) * n;
  double* qj = qi + n;
  for (int j = 0; j < n; j++) {
    const double y = qi[j], w = qj[j];
    qi[j] = c * y - s * w;
    qj[j] = s * y + c * w;
  }
}

// Q R <- Q (R + u v^T) (qrupdt.c:6-25)
void qr_update(std::vector<double>& r, std::vector<double>& qt, int n, std::vector<double>& u,
               const std::vector<double>& v) {
  int k = n - 1;
  while (k > 0 && u[k] == 0.0) k--;
  for (int i = k - 1; i >= 0; i--) {
    givens(r, qt, n, i, u[i], -u[i + 1]);
    if (u[i] == 0.0)
      u[i] = std::fabs(u[i + 1]);
    else if (std::fabs(u[i]) > std::fabs(u[i + 1]))
      u[i] = std::fabs(u[i]) * std::sqrt(1.0 + (u[i + 1] / u[i]) * (u[i + 1] / u[i]));
    else
      u[i] = std::fabs(u[i + 1]) * std::sqrt(1.0 + (u[i] / u[i + 1]) * (u[i] / u[i + 1]));
  }
  for (int j = 0; j < n; j++) r[j] += u[0] * v[j];
  for (int i = 0; i < k; i++)
    givens(r, qt, n, i, r[static_cast<size_t>(i) * n + i], -r[static_cast<size_t>(i + 1) * n + i]);
}

}  // namespace

extern "C" int scft_broyden_solve(int M, int N, const double* x, const double* phi_t, int bc,
                                  int cmat, double tolf, double* w, scft_broyden_stats* stats) {
  if (M < 4 || !x || !phi_t || !w || !(tolf > 0.0)) return SCFT_E_ARG;
  const int n = (bc == SCFT_BC_DD) ? M - 2 : M - 1;
  if (n < 2 || n > 4096) return SCFT_E_ARG;
  const int MAXITS = 400;
  const double EPS = 1.0e-14, TOLX = EPS, STPMX = 100.0, TOLMIN = tolf;
  Evaluators ev;
  if (int rc = ev.create(M, N, x, phi_t, bc, cmat, n)) return rc;
  const size_t nn = static_cast<size_t>(n) * n;
  std::vector<double> qt(nn), r(nn), c(n), d(n), fvcold(n), g(n), p(n), s(n), t(n), wv(n),
      xold(n), fvec(n);
  int check = 1, its = 0;
  bool restrt = true;
  double err = 0.0;
  auto R = [&](int i, int j) -> double& { return r[static_cast<size_t>(i) * n + j]; };
  auto QT = [&](int i, int j) -> double& { return qt[static_cast<size_t>(i) * n + j]; };
  auto finish = [&](int code) {
    if (stats) {
      stats->iterations = its;
      stats->evaluations = ev.evaluations;
      stats->sweeps = ev.sweeps;
      stats->final_err = err;
      stats->check = check;
    }
    return code;
  };

  double f = ev.fmin(w, fvec.data());
  if (ev.failed) return finish(SCFT_E_NAN);
  for (int i = 0; i < n; i++) err = std::max(err, std::fabs(fvec[i]));
  if (err < tolf) {
    check = 0;
    return finish(SCFT_OK);
  }
  double sum = 0.0;
  for (int i = 0; i < n; i++) sum += w[i] * w[i];
  const double stpmax = STPMX * std::max(std::sqrt(sum), static_cast<double>(n));
  for (its = 1; its <= MAXITS; its++) {
    if (restrt) {
      ev.fdjac(w, r.data());
      if (ev.failed) {
        check = 1;
        return finish(SCFT_E_NAN);
      }
      if (qr_decompose(r, n, c, d)) {
        check = 1;
        return finish(SCFT_E_PIVOT);
      }
      std::fill(qt.begin(), qt.end(), 0.0);
      for (int i = 0; i < n; i++) QT(i, i) = 1.0;
      for (int k = 0; k < n - 1; k++) {  // Q^T from the stored reflectors
        if (c[k] == 0.0) continue;
        for (int j = 0; j < n; j++) {
          double sm = 0.0;
          for (int i = k; i < n; i++) sm += R(i, k) * QT(i, j);
          sm /= c[k];
          for (int i = k; i < n; i++) QT(i, j) -= sm * R(i, k);
        }
      }
      for (int i = 0; i < n; i++) {
        R(i, i) = d[i];
        for (int j = 0; j < i; j++) R(i, j) = 0.0;
      }
    } else if (its > 1) {  // Broyden update of Q R
      for (int i = 0; i < n; i++) s[i] = w[i] - xold[i];
      for (int i = 0; i < n; i++) {
        double sm = 0.0;
        for (int j = i; j < n; j++) sm += R(i, j) * s[j];
        t[i] = sm;
      }
      bool skip = true;
      for (int i = 0; i < n; i++) {
        double sm = 0.0;
        for (int j = 0; j < n; j++) sm += QT(j, i) * t[j];
        wv[i] = fvec[i] - fvcold[i] - sm;
        if (std::fabs(wv[i]) >= EPS * (std::fabs(fvec[i]) + std::fabs(fvcold[i])))
          skip = false;
        else
          wv[i] = 0.0;
      }
      if (!skip) {
        for (int i = 0; i < n; i++) {
          double sm = 0.0;
          for (int j = 0; j < n; j++) sm += QT(i, j) * wv[j];
          t[i] = sm;
        }
        double den = 0.0;
        for (int j = 0; j < n; j++) den += s[j] * s[j];
        for (int i = 0; i < n; i++) s[i] /= den;
        qr_update(r, qt, n, t, s);
        for (int i = 0; i < n; i++) {
          if (R(i, i) == 0.0) {
            check = 1;
            return finish(SCFT_E_PIVOT);
          }
          d[i] = R(i, i);
        }
      }
    }
    // step: p = -R^{-1} Q^T F, gradient g = (Q R)^T F
    for (int i = 0; i < n; i++) {
      double sm = 0.0;
      for (int j = 0; j < n; j++) sm += QT(i, j) * fvec[j];
      p[i] = -sm;
    }
    for (int i = n - 1; i >= 0; i--) {
      double sm = 0.0;
      for (int j = 0; j <= i; j++) sm -= R(j, i) * p[j];
      g[i] = sm;
    }
    std::copy(w, w + n, xold.begin());
    fvcold = fvec;
    const double fold = f;
    p[n - 1] /= d[n - 1];  // rsolv.c:3-12
    for (int i = n - 2; i >= 0; i--) {
      double sm = 0.0;
      for (int j = i + 1; j < n; j++) sm += R(i, j) * p[j];
      p[i] = (p[i] - sm) / d[i];
    }
    if (!line_search(ev, n, xold.data(), fold, g.data(), p.data(), w, &f, fvec.data(), stpmax,
                     &check)) {
      check = 1;
      return finish(SCFT_E_NOCONV);
    }
    if (ev.failed) {
      check = 1;
      return finish(SCFT_E_NAN);
    }
    err = 0.0;
    for (int i = 0; i < n; i++) err = std::max(err, std::fabs(fvec[i]));
    if (err < tolf) {
      check = 0;
      return finish(SCFT_OK);
    }
    if (check) {  // line search stalled
      if (restrt) return finish(SCFT_E_NOCONV);
      double test = 0.0;
      const double den = std::max(f, 0.5 * n);
      for (int i = 0; i < n; i++)
        test = std::max(test, std::fabs(g[i]) * std::max(std::fabs(w[i]), 1.0) / den);
      if (test < TOLMIN) {
        check = 0;
        return finish(SCFT_OK);
      }
      restrt = true;
    } else {
      restrt = false;
      double test = 0.0;
      for (int i = 0; i < n; i++)
        test = std::max(test, std::fabs(w[i] - xold[i]) / std::max(std::fabs(w[i]), 1.0));
      if (test < TOLX) return finish(SCFT_OK);
    }
  }
  check = 1;
  its = MAXITS;
  return finish(SCFT_E_NOCONV);
}
