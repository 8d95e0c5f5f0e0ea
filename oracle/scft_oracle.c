/*
 * scft_oracle.c -- CPU restatement of giantsda/SCFT's inner field evaluation.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): used by tests/, by the smoke check
 * and by bench.py's CPU-baseline leg; never by the product library.
 *
 * Each function cites the reference lines it restates.  "Mode R" is the
 * parity definition: 1D_FEM.c matrices and boundary rows, N implicit-Euler
 * steps with a direct tridiagonal solve, the q_j * q_{N-j} pairing and
 * romint(v, N, 1/N) of DEALII_SCFT/drivescft.cc:184-193, the position-based
 * target of DEALII_SCFT/src/scft.cc:188-215, and r = phi - phi_t
 * (1D_FEM.c:274-277).
 *
 * How each part is pinned (tests/test_oracle.py):
 *   target      .res `phi` columns (Exp_m32/m1024_n2048_IE.res)
 *   romint      the reference romint.c/polint.c compiled under oracle/_ref,
 *               polynomial exactness, f0bar KAT 0.892581217773656
 *   assembly    closed-form stencils, K*1 = 0, row sums, consistent C
 *               reduces to c*A for constant w
 *   march       exact sine-mode solution for constant w (DD), q == 1 for
 *               w == 0 with Neumann at both ends, dense LU on tiny M,
 *               mirror symmetry, half-domain DN == full-domain DD
 *   Mode M      solution_matlab_N=33 solves the MATLAB residual to ~2.3e-7
 *   IRK4 march  (tests/test_oracle_irk4.py) the Gauss (2,2)-Pade factor on
 *               sine modes, fourth-order convergence to expm, a dense solve
 *               of the stage system, and the reference's N=33_for_read.txt:
 *               a field deal.II's IRK solver converged to 1.42e-9 evaluates
 *               here to 1.7e-9 (consistent C, N = 2048), 1.1e-4 with IE
 */
#include "scft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* target density                                                            */
/* ------------------------------------------------------------------------ */

/* EXP surface profile, written exactly as the reference evaluates it
 * ((e^u-1)^2/(e^u+1)^2, u = 4 tau x/(tau^2-x^2)), with the NaN -> 1 guard of
 * DEALII_SCFT/src/scft.cc:201-207 (e^u overflows to inf near x = tau). */
double orc_exp_profile(double x, double tau) {
  double u = 4.0 * tau * x / (tau * tau - x * x);
  double e = exp(u);
  double v = pow(e - 1.0, 2) / pow(e + 1.0, 2);
  if (isnan(v)) v = 1.0;
  return v;
}

/* Position-based target: DEALII_SCFT/src/scft.cc:188-215.  Nodes with
 * x_i <= tau get the profile; with `mirror` (two walls, DD) node M-1-i gets
 * the same value (index mirror, as the reference does). */
void orc_target_position(int M, const double *x, double tau, int mirror,
                         double *phi_t) {
  for (int i = 0; i < M; i++) phi_t[i] = 1.0;
  for (int i = 0; i < M; i++) {
    if (x[i] <= tau) {
      phi_t[i] = orc_exp_profile(x[i], tau);
      if (mirror) phi_t[M - i - 1] = phi_t[i];
    } else {
      break;
    }
  }
}

/* Count-based target: 1D_FEM.c:303-319 (x_left = ceil(M tau / L), mirrored,
 * no NaN guard) and Matlab_files/set_f0_given.m:1-14 (same with NaN -> 1). */
void orc_target_count(int M, const double *x, double tau, double L,
                      int nan_guard, double *phi_t) {
  int x_left = (int)ceil(M * tau / L);
  for (int i = 0; i < M; i++) phi_t[i] = 1.0;
  for (int i = 0; i < x_left && i < M; i++) {
    double u = 4.0 * tau * x[i] / (tau * tau - x[i] * x[i]);
    double v = pow(exp(u) - 1.0, 2) / pow(exp(u) + 1.0, 2);
    if (nan_guard && isnan(v)) v = 1.0;
    phi_t[i] = v;
    phi_t[M - i - 1] = v;
  }
}

/* ------------------------------------------------------------------------ */
/* meshes                                                                    */
/* ------------------------------------------------------------------------ */

/* Uniform nodes x_i = i L / (M-1): 1D_FEM.c:61,310. */
void orc_mesh_uniform(int M, double L, double *x) {
  for (int i = 0; i < M; i++) x[i] = i * L / (M - 1);
}

/* Graded mesh with element size h(x) = h0 (1 + x/ell) (production mix; no
 * reference counterpart -- the MATLAB general-mesh stencils
 * simple_FEM_1D_transient.m:35-63 are what make it usable).  Integrating
 * dx/di = h0 (1 + x/ell) gives x(i) = ell (exp(i h0/ell) - 1); the element
 * count is the number that reaches L, clamped to [m_min, m_max]. */
int orc_graded_elements(double L, double h0, double ell, int m_min,
                        int m_max) {
  double m = ell * log1p(L / ell) / h0;
  long mi = lround(m);
  if (mi < m_min) mi = m_min;
  if (mi > m_max) mi = m_max;
  return (int)mi;
}

/* Nodes of the graded mesh with M-1 elements spanning [0, L] exactly. */
void orc_mesh_graded(int M, double L, double ell, double *x) {
  int m = M - 1;
  double g = log1p(L / ell) / m;
  for (int i = 0; i <= m; i++) x[i] = ell * expm1(g * i);
  x[0] = 0.0;
  x[m] = L;
}

/* ------------------------------------------------------------------------ */
/* assembly                                                                  */
/* ------------------------------------------------------------------------ */

/* Linear-FEM mass A, stiffness B and w-mass C as three diagonals
 * (X3[3*i+0] = sub, X3[3*i+1] = diag, X3[3*i+2] = super) for every row,
 * boundary rows included in their natural (one-sided) form.
 *   A: (h_{i-1}/6, h_{i-1}/3 + h_i/3, h_i/6)  simple_FEM_1D_transient.m:35-47
 *      (uniform: h/6, 2h/3, h/6, 1D_FEM.c:95-96)
 *   B: (-1/h_{i-1}, 1/h_{i-1} + 1/h_i, -1/h_i) simple_FEM_1D_transient.m:51-63
 *   C row-scaled: w_i * A_i,:                1D_FEM.c:104-105, .m:66-70
 *   C consistent: element (a,b) contributes h/12 [3wa+wb, wa+wb; wa+wb, wa+3wb]
 *      (exact for linear w_h, equal to deal.II's QGauss(2), scft.cc:641-651) */
void orc_assemble(int M, const double *x, const double *w, int cmat,
                  double *A3, double *B3, double *C3) {
  memset(A3, 0, sizeof(double) * 3 * M);
  memset(B3, 0, sizeof(double) * 3 * M);
  memset(C3, 0, sizeof(double) * 3 * M);
  for (int e = 0; e < M - 1; e++) {
    int a = e, b = e + 1;
    double h = x[b] - x[a];
    /* element mass / stiffness into rows a and b */
    A3[3 * a + 1] += h / 3.0;
    A3[3 * a + 2] += h / 6.0;
    A3[3 * b + 0] += h / 6.0;
    A3[3 * b + 1] += h / 3.0;
    B3[3 * a + 1] += 1.0 / h;
    B3[3 * a + 2] += -1.0 / h;
    B3[3 * b + 0] += -1.0 / h;
    B3[3 * b + 1] += 1.0 / h;
    if (cmat == ORC_C_CONSISTENT) {
      double wa = w[a], wb = w[b];
      C3[3 * a + 1] += h * (3.0 * wa + wb) / 12.0;
      C3[3 * a + 2] += h * (wa + wb) / 12.0;
      C3[3 * b + 0] += h * (wa + wb) / 12.0;
      C3[3 * b + 1] += h * (wa + 3.0 * wb) / 12.0;
    }
  }
  if (cmat != ORC_C_CONSISTENT) {
    for (int i = 0; i < M; i++)
      for (int k = 0; k < 3; k++) C3[3 * i + k] = w[i] * A3[3 * i + k];
  }
}

/* Implicit-Euler system D = A + dt (B + C) and right-hand-side operator R = A,
 * with Dirichlet rows replaced by identity rows in D and zero rows in R
 * (1D_FEM.c:177-182 and the zeroed RHS entries :213-214). */
void orc_system(int M, const double *x, const double *w, double dt, int bc,
                int cmat, double *D3, double *R3) {
  double *A3 = (double *)malloc(sizeof(double) * 9 * M);
  double *B3 = A3 + 3 * M, *C3 = A3 + 6 * M;
  orc_assemble(M, x, w, cmat, A3, B3, C3);
  for (int i = 0; i < 3 * M; i++) {
    D3[i] = A3[i] + dt * (B3[i] + C3[i]);
    R3[i] = A3[i];
  }
  /* Dirichlet wall at x = 0 */
  D3[0] = 0.0; D3[1] = 1.0; D3[2] = 0.0;
  R3[0] = R3[1] = R3[2] = 0.0;
  /* the first row has no sub-diagonal, the last no super-diagonal */
  if (bc == ORC_BC_DD) {
    int l = 3 * (M - 1);
    D3[l] = 0.0; D3[l + 1] = 1.0; D3[l + 2] = 0.0;
    R3[l] = R3[l + 1] = R3[l + 2] = 0.0;
  }
  D3[3 * (M - 1) + 2] = 0.0;
  R3[3 * (M - 1) + 2] = 0.0;
  free(A3);
}

/* ------------------------------------------------------------------------ */
/* march                                                                     */
/* ------------------------------------------------------------------------ */

/* Direct tridiagonal (Thomas) solve of D out = rhs.  The reference solves the
 * same system with PETSc CG+ICC at default tolerance (1D_FEM.c:191-206,215);
 * a direct solve is the exact answer that iteration approximates. */
int orc_thomas(int M, const double *D3, const double *rhs, double *out) {
  double *u = (double *)malloc(sizeof(double) * 2 * M);
  double *y = u + M;
  u[0] = D3[1];
  if (u[0] == 0.0) { free(u); return ORC_E_PIVOT; }
  y[0] = rhs[0];
  for (int i = 1; i < M; i++) {
    double l = D3[3 * i] / u[i - 1];
    u[i] = D3[3 * i + 1] - l * D3[3 * (i - 1) + 2];
    if (u[i] == 0.0) { free(u); return ORC_E_PIVOT; }
    y[i] = rhs[i] - l * y[i - 1];
  }
  out[M - 1] = y[M - 1] / u[M - 1];
  for (int i = M - 2; i >= 0; i--)
    out[i] = (y[i] - D3[3 * i + 2] * out[i + 1]) / u[i];
  free(u);
  return ORC_OK;
}

/* Propagator march: q^0 = 1 with 0 on Dirichlet nodes (1D_FEM.c:114-128),
 * then N steps of D q^{j+1} = R q^j (1D_FEM.c:208-228 with the DEALII step
 * count N, drivescft.cc:130).  The Thomas factors are computed once and
 * reused for all N steps.  q must hold (N+1)*M values. */
int orc_march(int M, int N, const double *x, const double *w, int bc,
              int cmat, double *q) {
  if (M < 3 || N < 1) return ORC_E_ARG;
  double dt = 1.0 / N;
  double *D3 = (double *)malloc(sizeof(double) * 6 * M);
  double *R3 = D3 + 3 * M;
  double *lmul = (double *)malloc(sizeof(double) * 3 * M);
  double *inv = lmul + M, *y = lmul + 2 * M;
  int rc = ORC_OK;
  orc_system(M, x, w, dt, bc, cmat, D3, R3);
  /* factor */
  double u = D3[1];
  inv[0] = 1.0 / u;
  lmul[0] = 0.0;
  for (int i = 1; i < M; i++) {
    lmul[i] = D3[3 * i] / u;
    u = D3[3 * i + 1] - lmul[i] * D3[3 * (i - 1) + 2];
    if (u == 0.0) { rc = ORC_E_PIVOT; goto done; }
    inv[i] = 1.0 / u;
  }
  for (int i = 0; i < M; i++) q[i] = 1.0;
  q[0] = 0.0;
  if (bc == ORC_BC_DD) q[M - 1] = 0.0;
  for (int j = 0; j < N; j++) {
    const double *qo = q + (size_t)j * M;
    double *qn = q + (size_t)(j + 1) * M;
    for (int i = 0; i < M; i++) {
      double r = R3[3 * i + 1] * qo[i];
      if (i > 0) r += R3[3 * i] * qo[i - 1];
      if (i < M - 1) r += R3[3 * i + 2] * qo[i + 1];
      y[i] = (i > 0) ? r - lmul[i] * y[i - 1] : r;
    }
    qn[M - 1] = y[M - 1] * inv[M - 1];
    for (int i = M - 2; i >= 0; i--)
      qn[i] = (y[i] - D3[3 * i + 2] * qn[i + 1]) * inv[i];
  }
  for (size_t k = 0; k < (size_t)(N + 1) * M; k++)
    if (!isfinite(q[k])) { rc = ORC_E_NAN; break; }
done:
  free(D3);
  free(lmul);
  return rc;
}

/* The same march carried in x87 extended precision (64-bit significand):
 * the higher-precision truth that the FP64 paths (this oracle's own
 * orc_march and the device kernel) are both measured against.  The system
 * is assembled in FP64 exactly as orc_march does; only the factorisation
 * and the steps run in long double.  q receives the result rounded to FP64. */
int orc_march_ld(int M, int N, const double *x, const double *w, int bc,
                 int cmat, double *q) {
  if (M < 3 || N < 1) return ORC_E_ARG;
  double dt = 1.0 / N;
  double *D3 = (double *)malloc(sizeof(double) * 6 * M);
  double *R3 = D3 + 3 * M;
  long double *lmul = (long double *)malloc(sizeof(long double) * 4 * M);
  long double *inv = lmul + M, *y = lmul + 2 * M, *qq = lmul + 3 * M;
  long double *qn = (long double *)malloc(sizeof(long double) * M);
  int rc = ORC_OK;
  orc_system(M, x, w, dt, bc, cmat, D3, R3);
  long double u = D3[1];
  inv[0] = 1.0L / u;
  lmul[0] = 0.0L;
  for (int i = 1; i < M; i++) {
    lmul[i] = (long double)D3[3 * i] / u;
    u = (long double)D3[3 * i + 1] - lmul[i] * (long double)D3[3 * (i - 1) + 2];
    if (u == 0.0L) { rc = ORC_E_PIVOT; goto done; }
    inv[i] = 1.0L / u;
  }
  for (int i = 0; i < M; i++) qq[i] = 1.0L;
  qq[0] = 0.0L;
  if (bc == ORC_BC_DD) qq[M - 1] = 0.0L;
  for (int i = 0; i < M; i++) q[i] = (double)qq[i];
  for (int j = 0; j < N; j++) {
    for (int i = 0; i < M; i++) {
      long double r = (long double)R3[3 * i + 1] * qq[i];
      if (i > 0) r += (long double)R3[3 * i] * qq[i - 1];
      if (i < M - 1) r += (long double)R3[3 * i + 2] * qq[i + 1];
      y[i] = (i > 0) ? r - lmul[i] * y[i - 1] : r;
    }
    qn[M - 1] = y[M - 1] * inv[M - 1];
    for (int i = M - 2; i >= 0; i--)
      qn[i] = (y[i] - (long double)D3[3 * i + 2] * qn[i + 1]) * inv[i];
    for (int i = 0; i < M; i++) {
      qq[i] = qn[i];
      q[(size_t)(j + 1) * M + i] = (double)qn[i];
    }
  }
done:
  free(D3);
  free(lmul);
  free(qn);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* quadrature                                                                */
/* ------------------------------------------------------------------------ */

/* Value at 0 of the polynomial through (h[k], s[k]), k < n (Neville's
 * tableau); this is what polint(h, s, K, 0.0, ...) returns
 * (DEALII_SCFT/src/polint.c:5-41). */
static double neville_at_zero(const double *h, const double *s, int n) {
  double p[16];
  for (int k = 0; k < n; k++) p[k] = s[k];
  for (int d = 1; d < n; d++)
    for (int k = 0; k + d < n; k++)
      p[k] = (h[k] * p[k + 1] - h[k + d] * p[k]) / (h[k] - h[k + d]);
  return p[0];
}

/* Romberg integral of f[0..m] with spacing hh: DEALII_SCFT/src/romint.c:21-57.
 * Level count M = floor(log2 m + 1.5) (requires M >= 5, i.e. m >= 16), the
 * level-1 trapezoid m*hh*(f0+fm)/2, each later level adds the midpoints of
 * the previous stride, the step-size sequence is h_k = 4^-k, and the last
 * K = 5 levels are extrapolated to h = 0.  Returns NAN where the reference
 * exits (m < 16). */
double orc_romint(const double *f, int m, double hh) {
  enum { K = 5 };
  int levels = (int)(log(m * 1.0) / 0.6931471805599453 + 1.5);
  if (levels < K || levels > 40) return NAN;
  double s[48], h[48];
  h[0] = 1.0;
  s[0] = m * hh * (f[0] + f[m]) / 2;
  int np = 1;
  for (int j = 1; j < levels; j++) {
    int ii = m / np;
    double sum = 0.0;
    for (int i = ii / 2; i < m; i += ii) sum += f[i];
    s[j] = (s[j - 1] + ii * hh * sum) / 2;
    np += np;
    h[j] = h[j - 1] / 4;
  }
  return neville_at_zero(&h[levels - K], &s[levels - K], K);
}

/* Weights c[0..m] with orc_romint(f, m, hh) == sum_j c_j f_j for every f
 * (romint is linear in f).  Built as the Neville combination of the five
 * trapezoid levels, each level's weights written out directly. */
int orc_romberg_weights(int m, double hh, double *c) {
  enum { K = 5 };
  int levels = (int)(log(m * 1.0) / 0.6931471805599453 + 1.5);
  if (levels < K || levels > 40) return ORC_E_ARG;
  if ((m & (m - 1)) != 0) return ORC_E_ARG; /* trapezoid levels need 2^k */
  double h[K], e[K], lam[K];
  for (int k = 0; k < K; k++) h[k] = pow(0.25, levels - K + k);
  for (int k = 0; k < K; k++) {
    for (int t = 0; t < K; t++) e[t] = (t == k) ? 1.0 : 0.0;
    lam[k] = neville_at_zero(h, e, K);
  }
  for (int j = 0; j <= m; j++) c[j] = 0.0;
  for (int k = 0; k < K; k++) {
    int level = levels - K + k; /* trapezoid with 2^level panels */
    int stride = m >> level;
    double panel = stride * hh;
    for (int j = 0; j <= m; j += stride) {
      double wt = (j == 0 || j == m) ? 0.5 * panel : panel;
      c[j] += lam[k] * wt;
    }
  }
  return ORC_OK;
}

/* Plain trapezoid of f[0..m]; Mode M's quadrature
 * (simple_FEM_1D_transient.m:119-124). */
double orc_trapezoid(const double *f, int m, double hh) {
  double s = 0.5 * (f[0] + f[m]);
  for (int j = 1; j < m; j++) s += f[j];
  return s * hh;
}

/* ------------------------------------------------------------------------ */
/* density, residual                                                         */
/* ------------------------------------------------------------------------ */

/* phi_i = romint(v, N, 1/N) with v_j = q_j(i) q_{N-j}(i), node by node:
 * DEALII_SCFT/drivescft.cc:184-193. */
int orc_phi_romint(int M, int N, const double *q, double *phi) {
  double *v = (double *)malloc(sizeof(double) * (N + 1));
  for (int i = 0; i < M; i++) {
    for (int j = 0; j <= N; j++)
      v[j] = q[(size_t)j * M + i] * q[(size_t)(N - j) * M + i];
    phi[i] = orc_romint(v, N, 1.0 / N);
  }
  free(v);
  return isnan(phi[0]) ? ORC_E_ARG : ORC_OK;
}

/* The same functional with precomputed weights (the evaluation order the
 * device path uses). */
int orc_phi_weights(int M, int N, const double *q, const double *c,
                    double *phi) {
  for (int i = 0; i < M; i++) phi[i] = 0.0;
  for (int j = 0; j <= N; j++) {
    const double *a = q + (size_t)j * M, *b = q + (size_t)(N - j) * M;
    for (int i = 0; i < M; i++) phi[i] += c[j] * a[i] * b[i];
  }
  return ORC_OK;
}

int orc_n_unknowns(int M, int bc) { return bc == ORC_BC_DD ? M - 2 : M - 1; }
int orc_first_unknown(void) { return 1; }

/* Whole Mode-R evaluation: march, Romberg density, r = phi - phi_t on all
 * nodes (1D_FEM.c:274-277 sign; zero on Dirichlet nodes), and the infinity
 * norm over the unknowns (adm.c:114-122).  q_out (nullable) receives the
 * (N+1)*M slabs. */
static int eval_stepper(const orc_problem *p, int stepper, const double *w,
                        double *q_out, double *phi, double *r, double *rnorm) {
  int M = p->M, N = p->N;
  double *q = q_out ? q_out : (double *)malloc(sizeof(double) * (size_t)(N + 1) * M);
  double *c = (double *)malloc(sizeof(double) * (N + 1));
  if (!q || !c) return ORC_E_ALLOC;
  int rc = orc_romberg_weights(N, 1.0 / N, c);
  if (rc == ORC_OK)
    rc = (stepper == ORC_STEP_IRK4) ? orc_march_irk4(M, N, p->x, w, p->bc, p->cmat, q)
                                    : orc_march(M, N, p->x, w, p->bc, p->cmat, q);
  if (rc == ORC_OK) rc = orc_phi_weights(M, N, q, c, phi);
  if (rc == ORC_OK) {
    int last = (p->bc == ORC_BC_DD) ? M - 2 : M - 1;
    double nrm = 0.0;
    for (int i = 0; i < M; i++) {
      int unk = (i >= 1 && i <= last);
      r[i] = unk ? phi[i] - p->phi_t[i] : 0.0;
      if (unk && fabs(r[i]) > nrm) nrm = fabs(r[i]);
      if (unk && isnan(r[i])) nrm = NAN;
    }
    if (rnorm) *rnorm = nrm;
    if (isnan(nrm)) rc = ORC_E_NAN;
  }
  free(c);
  if (!q_out) free(q);
  return rc;
}

int orc_eval(const orc_problem *p, const double *w, double *q_out,
             double *phi, double *r, double *rnorm) {
  return eval_stepper(p, ORC_STEP_IE, w, q_out, phi, r, rnorm);
}

/* orc_eval with the march chosen: ORC_STEP_IE or ORC_STEP_IRK4 (the same
 * pairing, Romberg weights and residual). */
int orc_eval_stepper(const orc_problem *p, int stepper, const double *w,
                     double *q_out, double *phi, double *r, double *rnorm) {
  if (stepper != ORC_STEP_IE && stepper != ORC_STEP_IRK4) return ORC_E_ARG;
  return eval_stepper(p, stepper, w, q_out, phi, r, rnorm);
}

/* Mode M: Matlab_files/simple_FEM_1D_transient.m:1-128.  Walls w = 0
 * (`in=[0; in; 0]`, :8), general-mesh stencils, identity Dirichlet rows,
 * `steps` implicit-Euler steps (time_step = 2049, dt = 1/2048 in the
 * reference), then the offset trapezoid over v_k = q_k q_{steps-k},
 * k = 0..steps (:119-124), and out = phi_t - phi on the interior (:126-128).
 * w is the full-node field; w[0] and w[M-1] are ignored (taken as 0). */
int orc_eval_matlab(int M, const double *x, const double *w,
                    const double *phi_t, int steps, double dt, double *out) {
  double *wf = (double *)malloc(sizeof(double) * M);
  double *D3 = (double *)malloc(sizeof(double) * 6 * M);
  double *R3 = D3 + 3 * M;
  double *q = (double *)malloc(sizeof(double) * (size_t)(steps + 1) * M);
  double *rhs = (double *)malloc(sizeof(double) * 2 * M);
  double *v = rhs + M;
  int rc = ORC_OK;
  for (int i = 0; i < M; i++) wf[i] = w[i];
  wf[0] = 0.0;
  wf[M - 1] = 0.0;
  orc_system(M, x, wf, dt, ORC_BC_DD, ORC_C_ROWSCALED, D3, R3);
  for (int i = 0; i < M; i++) q[i] = 1.0;
  q[0] = q[M - 1] = 0.0;
  for (int t = 0; t < steps && rc == ORC_OK; t++) {
    const double *qo = q + (size_t)t * M;
    for (int i = 0; i < M; i++) {
      double r = R3[3 * i + 1] * qo[i];
      if (i > 0) r += R3[3 * i] * qo[i - 1];
      if (i < M - 1) r += R3[3 * i + 2] * qo[i + 1];
      rhs[i] = r;
    }
    rc = orc_thomas(M, D3, rhs, q + (size_t)(t + 1) * M);
  }
  double *vv = (double *)malloc(sizeof(double) * (steps + 1));
  for (int i = 1; i < M - 1 && rc == ORC_OK; i++) {
    for (int k = 0; k <= steps; k++)
      vv[k] = q[(size_t)k * M + i] * q[(size_t)(steps - k) * M + i];
    double phi = orc_trapezoid(vv, steps, dt);
    out[i - 1] = phi_t[i] - phi;
  }
  (void)v;
  free(vv);
  free(wf);
  free(D3);
  free(q);
  free(rhs);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* two-stage Gauss-Legendre (IRK4) march                                     */
/* ------------------------------------------------------------------------ */

/* inverse of a 2x2 block g (row-major) into gi; 0 when singular */
static int inv2(const double *g, double *gi) {
  double det = g[0] * g[3] - g[1] * g[2];
  if (det == 0.0 || !isfinite(det)) return 0;
  gi[0] = g[3] / det;
  gi[1] = -g[1] / det;
  gi[2] = -g[2] / det;
  gi[3] = g[0] / det;
  return 1;
}
static void mul2(const double *a, const double *b, double *c) {
  double c0 = a[0] * b[0] + a[1] * b[2], c1 = a[0] * b[1] + a[1] * b[3];
  double c2 = a[2] * b[0] + a[3] * b[2], c3 = a[2] * b[1] + a[3] * b[3];
  c[0] = c0; c[1] = c1; c[2] = c2; c[3] = c3;
}

/* The deal.II path's stepper (DEALII_SCFT/src/scft.cc:671-695,
 * drivescft.cc:130-146; Matlab_files/IRK4.m:96-120): with D' = B + C, each
 * step solves for the two stage slopes
 *   [ A + dt/4 D'               (1/4 - sqrt3/6) dt D' ] [k1]   [-D' q]
 *   [ (1/4 + sqrt3/6) dt D'     A + dt/4 D'           ] [k2] = [-D' q]
 * and sets q <- q + dt/2 (k1 + k2).  Dirichlet nodes keep k = 0 (IRK4.m:
 * 81-95 identity rows with a zero right-hand side), so only the unknowns
 * enter.  The 2n x 2n system is solved the way it is written: unknown i
 * carries the pair (k1_i, k2_i), which makes it block tridiagonal with 2x2
 * blocks, factored once (block Thomas, explicit 2x2 inverses) and reused for
 * every step.  Same initial condition, step count and slab layout as
 * orc_march. */
int orc_march_irk4(int M, int N, const double *x, const double *w, int bc,
                   int cmat, double *q) {
  if (M < 3 || N < 1) return ORC_E_ARG;
  const double dt = 1.0 / N;
  const double c01 = (0.25 - sqrt(3.0) / 6.0) * dt, c10 = (0.25 + sqrt(3.0) / 6.0) * dt;
  const int last = (bc == ORC_BC_DD) ? M - 2 : M - 1; /* unknowns 1..last */
  const int n = last;
  double *A3 = (double *)malloc(sizeof(double) * 9 * M);
  double *B3 = A3 + 3 * M, *C3 = A3 + 6 * M;
  double *blk = (double *)malloc(sizeof(double) * (size_t)n * 3 * 4); /* sub, diag, sup */
  double *Lm = (double *)malloc(sizeof(double) * (size_t)n * 4);
  double *Gi = (double *)malloc(sizeof(double) * (size_t)n * 4);
  double *y = (double *)malloc(sizeof(double) * (size_t)n * 2);
  double *Dp = (double *)malloc(sizeof(double) * 3 * M);
  int rc = ORC_OK;
  orc_assemble(M, x, w, cmat, A3, B3, C3);
  for (int k = 0; k < 3 * M; k++) Dp[k] = B3[k] + C3[k];
  for (int j = 0; j < n; j++) {
    int i = j + 1;
    for (int e = 0; e < 3; e++) {
      double a = A3[3 * i + e], d = Dp[3 * i + e];
      /* columns outside the unknowns are Dirichlet nodes (k = 0) */
      if ((e == 0 && j == 0) || (e == 2 && j == n - 1)) a = d = 0.0;
      double *b = blk + ((size_t)j * 3 + e) * 4;
      b[0] = a + dt / 4.0 * d;
      b[1] = c01 * d;
      b[2] = c10 * d;
      b[3] = a + dt / 4.0 * d;
    }
  }
  /* block factorisation: G_0 = D_0, L_j = S_j G_{j-1}^{-1}, G_j = D_j - L_j U_{j-1} */
  {
    double G[4];
    memcpy(G, blk + 4, sizeof G);
    if (!inv2(G, Gi)) { rc = ORC_E_PIVOT; goto done; }
    for (int j = 1; j < n; j++) {
      const double *S = blk + ((size_t)j * 3 + 0) * 4;
      const double *Dj = blk + ((size_t)j * 3 + 1) * 4;
      const double *Up = blk + ((size_t)(j - 1) * 3 + 2) * 4;
      double T[4];
      mul2(S, Gi + 4 * (j - 1), Lm + 4 * j);
      mul2(Lm + 4 * j, Up, T);
      for (int e = 0; e < 4; e++) G[e] = Dj[e] - T[e];
      if (!inv2(G, Gi + 4 * j)) { rc = ORC_E_PIVOT; goto done; }
    }
  }
  for (int i = 0; i < M; i++) q[i] = 1.0;
  q[0] = 0.0;
  if (bc == ORC_BC_DD) q[M - 1] = 0.0;
  for (int s = 0; s < N; s++) {
    const double *qo = q + (size_t)s * M;
    double *qn = q + (size_t)(s + 1) * M;
    /* forward: y_j = F_j - L_j y_{j-1}, F_j = (f_j, f_j), f = -D' q */
    for (int j = 0; j < n; j++) {
      int i = j + 1;
      double f = Dp[3 * i + 1] * qo[i] + Dp[3 * i] * qo[i - 1];
      if (i + 1 < M) f += Dp[3 * i + 2] * qo[i + 1];
      f = -f;
      double y0 = f, y1 = f;
      if (j > 0) {
        const double *L = Lm + 4 * j;
        y0 -= L[0] * y[2 * (j - 1)] + L[1] * y[2 * (j - 1) + 1];
        y1 -= L[2] * y[2 * (j - 1)] + L[3] * y[2 * (j - 1) + 1];
      }
      y[2 * j] = y0;
      y[2 * j + 1] = y1;
    }
    /* backward: k_j = G_j^{-1} (y_j - U_j k_{j+1}), stored over y */
    for (int j = n - 1; j >= 0; j--) {
      double t0 = y[2 * j], t1 = y[2 * j + 1];
      if (j < n - 1) {
        const double *U = blk + ((size_t)j * 3 + 2) * 4;
        t0 -= U[0] * y[2 * (j + 1)] + U[1] * y[2 * (j + 1) + 1];
        t1 -= U[2] * y[2 * (j + 1)] + U[3] * y[2 * (j + 1) + 1];
      }
      const double *G = Gi + 4 * j;
      y[2 * j] = G[0] * t0 + G[1] * t1;
      y[2 * j + 1] = G[2] * t0 + G[3] * t1;
    }
    for (int i = 0; i < M; i++) qn[i] = qo[i];
    for (int j = 0; j < n; j++) qn[j + 1] = qo[j + 1] + 0.5 * dt * (y[2 * j] + y[2 * j + 1]);
  }
  for (size_t k = 0; k < (size_t)(N + 1) * M; k++)
    if (!isfinite(q[k])) { rc = ORC_E_NAN; break; }
done:
  free(A3);
  free(blk);
  free(Lm);
  free(Gi);
  free(y);
  free(Dp);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* exact solution for constant w                                             */
/* ------------------------------------------------------------------------ */

/* Constant w = c, uniform mesh, DD.  Sine vectors v_k(i) = sin(k pi i/(M-1))
 * satisfy K v_k = lam_k A v_k with lam_k = (6/h^2)(1-cos t)/(2+cos t),
 * t = k pi/(M-1), so implicit Euler multiplies mode k by
 * g_k = 1/(1 + dt (lam_k + c)) per step.  Expanding q^0 (ones on the
 * interior) in the discrete sine basis gives q^j in closed form. */
void orc_sine_mode_exact(int M, double L, double c, int N, int j,
                         double *q_j) {
  int n = M - 1;
  double h = L / n, dt = 1.0 / N;
  for (int i = 0; i < M; i++) q_j[i] = 0.0;
  for (int k = 1; k < n; k++) {
    double t = k * M_PI / n;
    double a = 0.0;
    for (int i = 1; i < n; i++) a += sin(t * i);
    a *= 2.0 / n;
    double lam = 6.0 / (h * h) * (1.0 - cos(t)) / (2.0 + cos(t));
    double g = pow(1.0 / (1.0 + dt * (lam + c)), j);
    for (int i = 1; i < n; i++) q_j[i] += a * g * sin(t * i);
  }
}
