/*
 * oracle_solvers.c -- CPU restatement of the reference's outer nonlinear
 * solvers: gaussj (Gauss-Jordan, full pivoting), adm (fixed-point Anderson
 * mixing, adm.c) and broydn (globally convergent Broyden, broydn.c with its
 * helpers), plus the safeguarded residual-form Anderson mixer that the
 * batched device solver implements.
 *
 * TEST INFRASTRUCTURE ONLY (the checker); arrays here are 0-based, the
 * reference's are 1-based Numerical-Recipes offsets.
 */
#include "oracle_solvers.h"
#include "scft_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* gaussj: gaussj.c:7-60 (zero pivot += 1e-18) and                            */
/* DEALII_SCFT/src/gaussj.c:7-78 (zero pivot -> return 1).                    */
/* a is n x n row-major, b is n x m row-major; on exit a holds the inverse    */
/* and b the solution.  Pivot search visits rows then columns and keeps the  */
/* last entry with |a| >= big, as the reference loop does.                   */
/* ------------------------------------------------------------------------ */
int orc_gaussj(double *a, int n, double *b, int m, int zero_pivot_mode) {
  int ipiv[64], indxr[64], indxc[64];
  if (n > 64) return 2;
  for (int j = 0; j < n; j++) ipiv[j] = 0;
  for (int i = 0; i < n; i++) {
    double big = 0.0;
    int irow = 0, icol = 0;
    for (int j = 0; j < n; j++)
      if (ipiv[j] != 1)
        for (int k = 0; k < n; k++)
          if (ipiv[k] == 0 && fabs(a[j * n + k]) >= big) {
            big = fabs(a[j * n + k]);
            irow = j;
            icol = k;
          }
    ++ipiv[icol];
    if (irow != icol) {
      for (int l = 0; l < n; l++) {
        double t = a[irow * n + l]; a[irow * n + l] = a[icol * n + l]; a[icol * n + l] = t;
      }
      for (int l = 0; l < m; l++) {
        double t = b[irow * m + l]; b[irow * m + l] = b[icol * m + l]; b[icol * m + l] = t;
      }
    }
    indxr[i] = irow;
    indxc[i] = icol;
    if (a[icol * n + icol] == 0.0) {
      if (zero_pivot_mode == ORC_GJ_RETURN) return 1;
      a[icol * n + icol] += 1e-18;
    }
    double pivinv = 1.0 / a[icol * n + icol];
    a[icol * n + icol] = 1.0;
    for (int l = 0; l < n; l++) a[icol * n + l] *= pivinv;
    for (int l = 0; l < m; l++) b[icol * m + l] *= pivinv;
    for (int ll = 0; ll < n; ll++)
      if (ll != icol) {
        double dum = a[ll * n + icol];
        a[ll * n + icol] = 0.0;
        for (int l = 0; l < n; l++) a[ll * n + l] -= a[icol * n + l] * dum;
        for (int l = 0; l < m; l++) b[ll * m + l] -= b[icol * m + l] * dum;
      }
  }
  for (int l = n - 1; l >= 0; l--)
    if (indxr[l] != indxc[l])
      for (int k = 0; k < n; k++) {
        double t = a[k * n + indxr[l]];
        a[k * n + indxr[l]] = a[k * n + indxc[l]];
        a[k * n + indxc[l]] = t;
      }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* adm: adm.c:24-313, Anderson branch (the SPM branch is jumped over,         */
/* adm.c:47).  Ring of NRMAX+1 = 11 (x, d = f(x)-x) slots; lambda = 0.05 on   */
/* the first step, 1 - 0.95^its afterwards; ||d||_inf < tolf stops and        */
/* returns f(x).  The least-squares unknown i (1-based) belongs to the i-th   */
/* most recent earlier slot, which is the order adm.c:232-276 fills U, b in.  */
/* ------------------------------------------------------------------------ */
#define ADM_NRMAX 10
int orc_adm(double *x, int n, orc_fixed_point_fn f, void *ctx, double tolf,
            int maxits, double *xnew, int *its_out, double *err_out) {
  const int S = ADM_NRMAX + 1;
  double *xh = (double *)malloc(sizeof(double) * (size_t)n * S);
  double *dh = (double *)malloc(sizeof(double) * (size_t)n * S);
  double U[ADM_NRMAX * ADM_NRMAX], bb[ADM_NRMAX];
  int check = 1, its = 1, nc;
  double err = 0.0, lambda = 0.05;

  f(n, x, xnew, ctx);
  for (int i = 0; i < n; i++) {
    xh[i] = x[i];
    dh[i] = xnew[i] - x[i];
  }
  nc = 1;
  for (int i = 0; i < n; i++) err = fmax(err, fabs(dh[i]));
  if (err < tolf) { check = 0; goto out; }
  for (int i = 0; i < n; i++) x[i] = xh[i] + lambda * dh[i];

  for (its = 2; its <= maxits; ++its) {
    int nr = its - 1 < ADM_NRMAX ? its - 1 : ADM_NRMAX;
    lambda = 1.0 - pow(0.95, its);
    f(n, x, xnew, ctx);
    if (nc == S) nc = 0;
    int s = nc;
    double *xs = xh + (size_t)n * s, *ds = dh + (size_t)n * s;
    for (int i = 0; i < n; i++) {
      ds[i] = xnew[i] - x[i];
      xs[i] = x[i];
    }
    ++nc;
    err = 0.0;
    for (int i = 0; i < n; i++) err = fmax(err, fabs(ds[i]));
    if (err < tolf) { check = 0; goto out; }
    for (int a = 0; a < nr; a++) {
      const double *da = dh + (size_t)n * ((s - (a + 1) + S) % S);
      for (int b = 0; b <= a; b++) {
        const double *db = dh + (size_t)n * ((s - (b + 1) + S) % S);
        double t = 0.0;
        for (int i = 0; i < n; i++) t += (ds[i] - da[i]) * (ds[i] - db[i]);
        U[a * nr + b] = U[b * nr + a] = t;
      }
      double t = 0.0;
      for (int i = 0; i < n; i++) t += (ds[i] - da[i]) * ds[i];
      bb[a] = t;
    }
    orc_gaussj(U, nr, bb, 1, ORC_GJ_PERTURB);
    for (int i = 0; i < n; i++) {
      double tx = 0.0, td = 0.0;
      for (int a = 0; a < nr; a++) {
        int sa = (s - (a + 1) + S) % S;
        tx += bb[a] * (xh[(size_t)n * sa + i] - xs[i]);
        td += bb[a] * (dh[(size_t)n * sa + i] - ds[i]);
      }
      x[i] = xs[i] + tx + lambda * (ds[i] + td);
    }
  }
  its = maxits;
out:
  if (its_out) *its_out = its;
  if (err_out) *err_out = err;
  free(xh);
  free(dh);
  return check;
}
#undef ADM_NRMAX

/* ------------------------------------------------------------------------ */
/* Safeguarded residual-form Anderson mixer (the algorithm of the device      */
/* mixer; adm.c is the fixed-point form of the same update).                  */
/*   r_k = F(w_k) = phi[w_k] - phi_t,   stop when ||r_k||_inf <= tol           */
/*   theta = argmin || r_k - sum_i theta_i (r_k - r_{k-i}) ||_2  (normal eqns, */
/*           Gauss-Jordan with full pivoting as adm.c:278)                     */
/*   w_{k+1} = wbar + beta * rbar, wbar/rbar the same combination of w / r     */
/*   r is mixed as is, relative to phi_t, or as log(phi/phi_t) (transform)     */
/*   the step w_{k+1} - w_k is scaled down to ||.||_inf <= cap                 */
/*   history is cleared when ||r_k|| > restart * best so far, or U singular    */
/* ------------------------------------------------------------------------ */
int orc_precond_stencil(const double *x, int M, int bc, double *st) {
  int n = orc_n_unknowns(M, bc);
  if (n < 1) return ORC_E_ARG;
  if (n == 1) { st[0] = st[1] = st[2] = 0.0; return ORC_OK; }
  for (int u = 0; u < n; u++) {
    int i = u + 1; /* node of unknown u */
    double hl = (u > 0) ? x[i] - x[i - 1] : 0.0;
    double hr = (u + 1 < n) ? x[i + 1] - x[i] : 0.0;
    double m = 0.5 * (hl + hr);
    st[3 * u + 0] = (u > 0) ? (1.0 / hl) / m : 0.0;
    st[3 * u + 2] = (u + 1 < n) ? (1.0 / hr) / m : 0.0;
    st[3 * u + 1] = st[3 * u + 0] + st[3 * u + 2];
  }
  return ORC_OK;
}

int orc_gs_stencil(const double *x, int M, int bc, double *st) {
  int n = orc_n_unknowns(M, bc);
  if (n < 1) return ORC_E_ARG;
  for (int u = 0; u < n; u++) {
    int i = u + 1; /* node of unknown u; node i - 1 always exists */
    double hl = x[i] - x[i - 1];
    double hr = (i + 1 < M) ? x[i + 1] - x[i] : 0.0;
    double m = 0.5 * (hl + hr);
    st[3 * u + 0] = (1.0 / hl) / m;
    st[3 * u + 1] = 0.0;
    st[3 * u + 2] = (i + 1 < M) ? (1.0 / hr) / m : 0.0;
  }
  return ORC_OK;
}

int orc_anderson(double *w, int n, orc_residual_fn F, void *ctx,
                 const orc_anderson_opts *o, const double *phi_t_unk,
                 orc_anderson_stats *st) {
  if (o->precond > 0.0 && !o->pstencil) return 3;
  if (o->gs_precond > 0.0 && (!o->gstencil || !phi_t_unk)) return 3;
  int mh = o->history;
  int S = mh + 1;
  double *wh = (double *)calloc((size_t)n * S, sizeof(double));
  double *rh = (double *)calloc((size_t)n * S, sizeof(double));
  double *r = (double *)malloc(sizeof(double) * n);
  double *wn = (double *)malloc(sizeof(double) * n);
  double *wa = (double *)malloc(sizeof(double) * n); /* last accepted iterate */
  double U[64 * 64], bb[64];
  int stored = 0, head = 0, k, rc = 1, restarts = 0, halvings = 0;
  double best = INFINITY, err = INFINITY, eacc = INFINITY;
  int gs_on = o->gs_precond > 0.0;
  int rat_on = o->gs_shift_below > 0.0 ? 0 : 1;
  for (k = 0; k < o->max_iter; k++) {
    int ferr = F(n, w, r, ctx);
    err = 0.0;
    for (int i = 0; i < n; i++) {
      if (isnan(r[i])) { err = NAN; break; }
      err = fmax(err, fabs(r[i]));
    }
    if (st && st->err_trace && k < st->trace_len) st->err_trace[k] = err;
    /* backtracking: halve the step that led here, drop the history */
    if (o->backtrack > 0.0 && k > 0 &&
        (ferr != 0 || isnan(err) || err > o->backtrack * eacc)) {
      if (halvings >= o->max_halvings) { rc = 2; break; }
      for (int i = 0; i < n; i++) w[i] = wa[i] + 0.5 * (w[i] - wa[i]);
      halvings++;
      restarts++;
      stored = 0;
      continue;
    }
    if (ferr != 0 || isnan(err)) { rc = 2; break; }
    halvings = 0;
    eacc = err;
    memcpy(wa, w, sizeof(double) * n);
    if (err <= o->tol) { rc = 0; break; }
    if (err > o->restart_factor * best && stored > 0) {
      stored = 0;
      restarts++;
    }
    if (gs_on && ((o->gs_switch > 0 && k + 1 > o->gs_switch) ||
                    (o->gs_switch_growth > 0.0 && err > o->gs_switch_growth * best))) {
      gs_on = 0; /* leave the ground-state form: plain form, history cleared */
      stored = 0;
      restarts++;
    }
    if (err < best) best = err;
    if (gs_on && o->gs_shift > 0.0 && !rat_on && err < o->gs_shift_below) {
      rat_on = 1; /* near the solution: the rational form, history cleared */
      stored = 0;
      restarts++;
    }
    /* mixing residual: r itself, r/max(phi_t, floor), or log(phi/phi_t) */
    if (o->transform != ORC_MIX_PLAIN && phi_t_unk) {
      for (int i = 0; i < n; i++) {
        double pt = fmax(phi_t_unk[i], o->floor);
        if (o->transform == ORC_MIX_RELATIVE) r[i] = r[i] / pt;
        else r[i] = log(fmax(pt + r[i], o->floor * 1e-3) / pt);
      }
    }
    /* preconditioned mixing residual g = f + a M_L^{-1} K f, or the
     * ground-state form for starts near the solution:
     *   g_u = f_u + a_w [ l_u (d_{u-1}/d_u)(f_u - f_{u-1})
     *                   + r_u (d_{u+1}/d_u)(f_u - f_{u+1}) + c max(w_u,0) f_u ]
     * with d = sqrt(max(phi_t, floor)) and d = 0 beyond the unknowns (the
     * wall nodes; a natural end has r_u = 0) */
    if (gs_on && o->gs_shift > 0.0 && rat_on) {
      /* the rational form (1 - a) f + a G f + a g (g I + G0)^{-1} f, G0 = G
       * without the field term: tridiagonal with |lower| + |upper| = its
       * diagonal, so g I + G0 is strictly diagonally dominant and Thomas
       * needs no pivoting; its factors depend on the mesh and phi_t only */
      const double *gs = o->gstencil;
      const double a = o->gs_precond, g = o->gs_shift;
      double *lo = (double *)malloc(sizeof(double) * 4 * (size_t)n);
      double *up = lo + n, *cp = up + n, *y = cp + n;
      for (int i = 0; i < n; i++) {
        const double di = sqrt(fmax(phi_t_unk[i], o->floor));
        const double dl = i > 0 ? sqrt(fmax(phi_t_unk[i - 1], o->floor)) : 0.0;
        const double dr = (i + 1 < n) ? sqrt(fmax(phi_t_unk[i + 1], o->floor)) : 0.0;
        lo[i] = gs[3 * i] * (dl / di);
        up[i] = gs[3 * i + 2] * (dr / di);
        const double fl = i > 0 ? r[i - 1] : 0.0, fr = (i + 1 < n) ? r[i + 1] : 0.0;
        double t = lo[i] * (r[i] - fl) + up[i] * (r[i] - fr);
        t += o->gs_field_weight * fmax(w[i], 0.0) * r[i];
        wn[i] = (1.0 - a) * r[i] + a * t;
      }
      for (int i = 0; i < n; i++) { /* forward sweep of (g I + G0) y = f */
        const double diag = g + lo[i] + up[i];
        const double den = i > 0 ? diag + lo[i] * cp[i - 1] : diag;
        const double rinv = 1.0 / den;
        cp[i] = -up[i] * rinv;
        y[i] = (i > 0 ? r[i] + lo[i] * y[i - 1] : r[i]) * rinv;
      }
      for (int i = n - 2; i >= 0; i--) y[i] -= cp[i] * y[i + 1];
      for (int i = 0; i < n; i++) wn[i] += a * g * y[i];
      free(lo);
      memcpy(r, wn, sizeof(double) * n);
    } else if (gs_on) {
      const double *gs = o->gstencil;
      for (int i = 0; i < n; i++) {
        const double di = sqrt(fmax(phi_t_unk[i], o->floor));
        const double dl = i > 0 ? sqrt(fmax(phi_t_unk[i - 1], o->floor)) : 0.0;
        const double dr = (i + 1 < n) ? sqrt(fmax(phi_t_unk[i + 1], o->floor)) : 0.0;
        const double fl = i > 0 ? r[i - 1] : 0.0, fr = (i + 1 < n) ? r[i + 1] : 0.0;
        double t = gs[3 * i] * (dl / di) * (r[i] - fl) + gs[3 * i + 2] * (dr / di) * (r[i] - fr);
        t += o->gs_field_weight * fmax(w[i], 0.0) * r[i];
        wn[i] = r[i] + o->gs_precond * t;
      }
      memcpy(r, wn, sizeof(double) * n);
    } else if (o->precond > 0.0) {
      const double *ps = o->pstencil;
      double fl = 0.0, fc = r[0];
      for (int i = 0; i < n; i++) {
        double fr = (i + 1 < n) ? r[i + 1] : 0.0;
        wn[i] = fc + o->precond * (ps[3 * i + 1] * fc - ps[3 * i] * fl - ps[3 * i + 2] * fr);
        fl = fc;
        fc = fr;
      }
      memcpy(r, wn, sizeof(double) * n);
    }
    /* push (w_k, r_k) */
    memcpy(wh + (size_t)n * head, w, sizeof(double) * n);
    memcpy(rh + (size_t)n * head, r, sizeof(double) * n);
    int cur = head;
    head = (head + 1) % S;
    int m = stored < mh ? stored : mh;
    if (stored < S) stored++;
    double *theta = bb;
    int singular = 0;
    if (m > 0) {
      for (int a = 0; a < m; a++) {
        const double *ra = rh + (size_t)n * ((cur - (a + 1) + S) % S);
        for (int b = 0; b <= a; b++) {
          const double *rb = rh + (size_t)n * ((cur - (b + 1) + S) % S);
          double t = 0.0;
          for (int i = 0; i < n; i++) t += (r[i] - ra[i]) * (r[i] - rb[i]);
          U[a * m + b] = U[b * m + a] = t;
        }
        double t = 0.0;
        for (int i = 0; i < n; i++) t += (r[i] - ra[i]) * r[i];
        bb[a] = t;
      }
      double tr = 0.0;
      for (int a = 0; a < m; a++) tr += U[a * m + a];
      for (int a = 0; a < m; a++) U[a * m + a] += o->regularization * tr;
      if (tr == 0.0 || orc_gaussj(U, m, bb, 1, ORC_GJ_RETURN) != 0) singular = 1;
      for (int a = 0; a < m && !singular; a++)
        if (!isfinite(theta[a])) singular = 1;
      if (singular) {
        restarts++;
        stored = 1; /* keep only the current pair */
        m = 0;
      }
    }
    double step = 0.0;
    for (int i = 0; i < n; i++) {
      double wb = w[i], rb = r[i];
      for (int a = 0; a < m; a++) {
        int sa = (cur - (a + 1) + S) % S;
        wb -= theta[a] * (w[i] - wh[(size_t)n * sa + i]);
        rb -= theta[a] * (r[i] - rh[(size_t)n * sa + i]);
      }
      wn[i] = wb + o->beta * rb;
      step = fmax(step, fabs(wn[i] - w[i]));
    }
    double scale = (o->step_cap > 0.0 && step > o->step_cap) ? o->step_cap / step : 1.0;
    for (int i = 0; i < n; i++) w[i] += scale * (wn[i] - w[i]);
    if (o->field_floor > 0.0)
      for (int i = 0; i < n; i++) w[i] = fmax(w[i], o->field_lower);
  }
  if (st) {
    st->iterations = k + (rc == 0 ? 1 : 0);
    st->final_err = err;
    st->restarts = restarts;
  }
  free(wh);
  free(rh);
  free(r);
  free(wn);
  free(wa);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Broyden (Numerical Recipes 9.7 as shipped in broydn.c:44-292,             */
/* DEALII_SCFT/src/{fdjac,lnsrch,qrdcmp,qrupdt,rotate,rsolv}.c).             */
/* ------------------------------------------------------------------------ */
#define SGN(a, b) ((b) >= 0.0 ? fabs(a) : -fabs(a))

typedef struct {
  int n;
  orc_vecfunc_fn F;
  void *ctx;
  double *fvec;
  int funcerr;
  long nevals;
} brd_state;

/* fminbrd: broydn.c:30-42 */
static double brd_fmin(brd_state *b, const double *x) {
  if (b->F(b->n, x, b->fvec, b->ctx) != 0) b->funcerr = 1;
  b->nevals++;
  if (b->funcerr) return 0.0;
  double sum = 0.0;
  for (int i = 0; i < b->n; i++) sum += b->fvec[i] * b->fvec[i];
  return 0.5 * sum;
}

/* fdjac: DEALII_SCFT/src/fdjac.c:9-41 (forward differences, EPS 1e-7) */
static void brd_fdjac(brd_state *b, double *x, const double *fvec, double *df) {
  int n = b->n;
  const double EPS = 1.0e-7;
  double *f = (double *)malloc(sizeof(double) * n);
  for (int j = 0; j < n; j++) {
    double temp = x[j];
    double h = EPS * temp;
    if (fabs(h) < EPS) h = SGN(EPS, temp);
    x[j] = temp + h;
    h = x[j] - temp;
    if (b->F(n, x, f, b->ctx) != 0) b->funcerr = 1;
    b->nevals++;
    if (b->funcerr) { x[j] = temp; break; }
    x[j] = temp;
    for (int i = 0; i < n; i++) df[i * n + j] = (f[i] - fvec[i]) / h;
  }
  free(f);
}

/* lnsrch: DEALII_SCFT/src/lnsrch.c:8-61 (ALF 1e-4, TOLX 1e-7); returns -1 on
 * the reference's "Roundoff problem" exit. */
static int brd_lnsrch(brd_state *bs, const double *xold, double fold,
                      const double *g, double *p, double *x, double *f,
                      double stpmax, int *check) {
  int n = bs->n;
  const double ALF = 1.0e-4, TOLX = 1.0e-7;
  double a, alam, alam2 = 0.0, alamin, b, disc, f2 = 0.0, rhs1, rhs2, slope,
                  sum, temp, test, tmplam;
  *check = 0;
  sum = 0.0;
  for (int i = 0; i < n; i++) sum += p[i] * p[i];
  sum = sqrt(sum);
  if (sum > stpmax)
    for (int i = 0; i < n; i++) p[i] *= stpmax / sum;
  slope = 0.0;
  for (int i = 0; i < n; i++) slope += g[i] * p[i];
  if (slope >= 0.0) return -1;
  test = 0.0;
  for (int i = 0; i < n; i++) {
    temp = fabs(p[i]) / fmax(fabs(xold[i]), 1.0);
    if (temp > test) test = temp;
  }
  alamin = TOLX / test;
  alam = 1.0;
  for (;;) {
    for (int i = 0; i < n; i++) x[i] = xold[i] + alam * p[i];
    *f = brd_fmin(bs, x);
    if (alam < alamin) {
      for (int i = 0; i < n; i++) x[i] = xold[i];
      *check = 1;
      return 0;
    } else if (*f <= fold + ALF * alam * slope) {
      return 0;
    } else {
      if (alam == 1.0) {
        tmplam = -slope / (2.0 * (*f - fold - slope));
      } else {
        rhs1 = *f - fold - alam * slope;
        rhs2 = f2 - fold - alam2 * slope;
        a = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2);
        b = (-alam2 * rhs1 / (alam * alam) + alam * rhs2 / (alam2 * alam2)) /
            (alam - alam2);
        if (a == 0.0) {
          tmplam = -slope / (2.0 * b);
        } else {
          disc = b * b - 3.0 * a * slope;
          if (disc < 0.0) tmplam = 0.5 * alam;
          else if (b <= 0.0) tmplam = (-b + sqrt(disc)) / (3.0 * a);
          else tmplam = -slope / (b + sqrt(disc));
        }
        if (tmplam > 0.5 * alam) tmplam = 0.5 * alam;
      }
    }
    alam2 = alam;
    f2 = *f;
    alam = fmax(tmplam, 0.1 * alam);
  }
}

/* qrdcmp: DEALII_SCFT/src/qrdcmp.c:6-34 (Householder, scaled columns) */
static void brd_qrdcmp(double *a, int n, double *c, double *d, int *sing) {
  *sing = 0;
  for (int k = 0; k < n - 1; k++) {
    double scale = 0.0;
    for (int i = k; i < n; i++) scale = fmax(scale, fabs(a[i * n + k]));
    if (scale == 0.0) {
      *sing = 1;
      c[k] = d[k] = 0.0;
    } else {
      double sum = 0.0;
      for (int i = k; i < n; i++) a[i * n + k] /= scale;
      for (int i = k; i < n; i++) sum += a[i * n + k] * a[i * n + k];
      double sigma = SGN(sqrt(sum), a[k * n + k]);
      a[k * n + k] += sigma;
      c[k] = sigma * a[k * n + k];
      d[k] = -scale * sigma;
      for (int j = k + 1; j < n; j++) {
        double s2 = 0.0;
        for (int i = k; i < n; i++) s2 += a[i * n + k] * a[i * n + j];
        double tau = s2 / c[k];
        for (int i = k; i < n; i++) a[i * n + j] -= tau * a[i * n + k];
      }
    }
  }
  d[n - 1] = a[(n - 1) * n + n - 1];
  if (d[n - 1] == 0.0) *sing = 1;
}

/* rotate: DEALII_SCFT/src/rotate.c:6-35 (Givens on rows i, i+1 of r, qt) */
static void brd_rotate(double *r, double *qt, int n, int i, double a, double b) {
  double c, fact, s, w, y;
  if (a == 0.0) {
    c = 0.0;
    s = (b >= 0.0 ? 1.0 : -1.0);
  } else if (fabs(a) > fabs(b)) {
    fact = b / a;
    c = SGN(1.0 / sqrt(1.0 + (fact * fact)), a);
    s = fact * c;
  } else {
    fact = a / b;
    s = SGN(1.0 / sqrt(1.0 + (fact * fact)), b);
    c = fact * s;
  }
  for (int j = i; j < n; j++) {
    y = r[i * n + j];
    w = r[(i + 1) * n + j];
    r[i * n + j] = c * y - s * w;
    r[(i + 1) * n + j] = s * y + c * w;
  }
  for (int j = 0; j < n; j++) {
    y = qt[i * n + j];
    w = qt[(i + 1) * n + j];
    qt[i * n + j] = c * y - s * w;
    qt[(i + 1) * n + j] = s * y + c * w;
  }
}

/* qrupdt: DEALII_SCFT/src/qrupdt.c:6-25 (rank-1 update of Q R) */
static void brd_qrupdt(double *r, double *qt, int n, double *u, const double *v) {
  int k;
  for (k = n - 1; k >= 0; k--)
    if (u[k] != 0.0) break;
  if (k < 0) k = 0;
  for (int i = k - 1; i >= 0; i--) {
    brd_rotate(r, qt, n, i, u[i], -u[i + 1]);
    if (u[i] == 0.0) u[i] = fabs(u[i + 1]);
    else if (fabs(u[i]) > fabs(u[i + 1]))
      u[i] = fabs(u[i]) * sqrt(1.0 + (u[i + 1] / u[i]) * (u[i + 1] / u[i]));
    else
      u[i] = fabs(u[i + 1]) * sqrt(1.0 + (u[i] / u[i + 1]) * (u[i] / u[i + 1]));
  }
  for (int j = 0; j < n; j++) r[j] += u[0] * v[j];
  for (int i = 0; i < k; i++)
    brd_rotate(r, qt, n, i, r[i * n + i], -r[(i + 1) * n + i]);
}

/* rsolv: DEALII_SCFT/src/rsolv.c:3-12 */
static void brd_rsolv(const double *a, int n, const double *d, double *b) {
  b[n - 1] /= d[n - 1];
  for (int i = n - 2; i >= 0; i--) {
    double sum = 0.0;
    for (int j = i + 1; j < n; j++) sum += a[i * n + j] * b[j];
    b[i] = (b[i] - sum) / d[i];
  }
}

/* broydn: broydn.c:44-292.  *jc carries the reference's Jacobian-reuse flag
 * (0: none, 1: previous, 2: current); tolf is its `err` input.  Returns the
 * reference's `check` (0 converged, 1 failed) or -1 where the reference
 * would exit (singular Jacobian, lnsrch roundoff). */
int orc_broydn(double *x, int n, orc_vecfunc_fn F, void *ctx, double tolf,
               int *jc, double *qt, double *r, double *d,
               orc_broydn_stats *st) {
  const int MAXITS = 400;
  const double EPS = 1e-14, TOLX = EPS, STPMX = 100.0;
  double TOLF = tolf, TOLMIN = TOLF;
  double *c = (double *)malloc(sizeof(double) * n * 10);
  double *fvcold = c + n, *g = c + 2 * n, *p = c + 3 * n, *s = c + 4 * n,
         *t = c + 5 * n, *w = c + 6 * n, *xold = c + 7 * n, *fvec = c + 8 * n;
  brd_state bs = {n, F, ctx, fvec, 0, 0};
  int check = 1, restrt = 0, sing, its = 0, rc;
  double f, fold, err = 0.0, sum, den, test, temp, stpmax;

  f = brd_fmin(&bs, x);
  if (bs.funcerr) goto error;
  err = 0.0;
  for (int i = 0; i < n; i++) err = fmax(err, fabs(fvec[i]));
  if (st && st->err_trace && st->trace_len > 0) st->err_trace[0] = err;
  if (err < TOLF) { check = 0; its = 0; goto done; }
  sum = 0.0;
  for (int i = 0; i < n; i++) sum += x[i] * x[i];
  stpmax = STPMX * fmax(sqrt(sum), (double)n);
  if (!*jc) restrt = 1;
  for (its = 1; its <= MAXITS; its++) {
    if (restrt) {
      brd_fdjac(&bs, x, fvec, r);
      if (bs.funcerr) goto error;
      brd_qrdcmp(r, n, c, d, &sing);
      if (sing) { check = -1; goto done; }
      for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) qt[i * n + j] = 0.0;
        qt[i * n + i] = 1.0;
      }
      for (int k = 0; k < n - 1; k++) {
        if (c[k] != 0.0) {
          for (int j = 0; j < n; j++) {
            sum = 0.0;
            for (int i = k; i < n; i++) sum += r[i * n + k] * qt[i * n + j];
            sum /= c[k];
            for (int i = k; i < n; i++) qt[i * n + j] -= sum * r[i * n + k];
          }
        }
      }
      for (int i = 0; i < n; i++) {
        r[i * n + i] = d[i];
        for (int j = 0; j < i; j++) r[i * n + j] = 0.0;
      }
      *jc = 2;
    } else if (its > 1) {
      for (int i = 0; i < n; i++) s[i] = x[i] - xold[i];
      for (int i = 0; i < n; i++) {
        for (sum = 0.0, rc = i; rc < n; rc++) sum += r[i * n + rc] * s[rc];
        t[i] = sum;
      }
      int skip = 1;
      for (int i = 0; i < n; i++) {
        for (sum = 0.0, rc = 0; rc < n; rc++) sum += qt[rc * n + i] * t[rc];
        w[i] = fvec[i] - fvcold[i] - sum;
        if (fabs(w[i]) >= EPS * (fabs(fvec[i]) + fabs(fvcold[i]))) skip = 0;
        else w[i] = 0.0;
      }
      if (!skip) {
        for (int i = 0; i < n; i++) {
          for (sum = 0.0, rc = 0; rc < n; rc++) sum += qt[i * n + rc] * w[rc];
          t[i] = sum;
        }
        for (den = 0.0, rc = 0; rc < n; rc++) den += s[rc] * s[rc];
        for (int i = 0; i < n; i++) s[i] /= den;
        brd_qrupdt(r, qt, n, t, s);
        for (int i = 0; i < n; i++) {
          if (r[i * n + i] == 0.0) { check = -1; goto done; }
          d[i] = r[i * n + i];
        }
      }
    }
    for (int i = 0; i < n; i++) {
      for (sum = 0.0, rc = 0; rc < n; rc++) sum += qt[i * n + rc] * fvec[rc];
      p[i] = -sum;
    }
    for (int i = n - 1; i >= 0; i--) {
      for (sum = 0.0, rc = 0; rc <= i; rc++) sum -= r[rc * n + i] * p[rc];
      g[i] = sum;
    }
    for (int i = 0; i < n; i++) {
      xold[i] = x[i];
      fvcold[i] = fvec[i];
    }
    fold = f;
    brd_rsolv(r, n, d, p);
    if (brd_lnsrch(&bs, xold, fold, g, p, x, &f, stpmax, &check) < 0) {
      check = -1;
      goto done;
    }
    if (bs.funcerr) goto error;
    err = 0.0;
    for (int i = 0; i < n; i++) err = fmax(err, fabs(fvec[i]));
    if (st && st->err_trace && its < st->trace_len) st->err_trace[its] = err;
    if (err < TOLF) { check = 0; *jc = 1; goto done; }
    if (check) {
      if (restrt) goto done;
      test = 0.0;
      den = fmax(f, 0.5 * n);
      for (int i = 0; i < n; i++) {
        temp = fabs(g[i]) * fmax(fabs(x[i]), 1.0) / den;
        if (temp > test) test = temp;
      }
      if (test < TOLMIN) { check = 0; *jc = 1; goto done; }
      restrt = 1;
    } else {
      restrt = 0;
      test = 0.0;
      for (int i = 0; i < n; i++) {
        temp = fabs(x[i] - xold[i]) / fmax(fabs(x[i]), 1.0);
        if (temp > test) test = temp;
      }
      if (test < TOLX) { *jc = 1; goto done; }
    }
  }
error:
  check = 1;
done:
  if (st) {
    st->iterations = its;
    st->evaluations = bs.nevals;
    st->final_err = err;
  }
  free(c);
  return check;
}
#undef SGN
