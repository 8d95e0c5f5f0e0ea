/*
 * oracle_batch.c -- CPU drivers around the oracle evaluator: residual and
 * fixed-point callbacks with the reference's conventions, a single-instance
 * Anderson solve, and an OpenMP loop over independent instances (the CPU
 * baseline that bench.py times next to the device path).
 * TEST INFRASTRUCTURE ONLY.
 */
#include "oracle_batch.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_ctx_init(orc_eval_ctx *c, const orc_problem *p, double beta) {
  memset(c, 0, sizeof(*c));
  c->p = *p;
  c->beta = beta;
  c->wfull = (double *)calloc(p->M, sizeof(double));
  c->phi = (double *)malloc(sizeof(double) * p->M);
  c->rfull = (double *)malloc(sizeof(double) * p->M);
  c->q = (double *)malloc(sizeof(double) * (size_t)(p->N + 1) * p->M);
  c->nevals = 0;
  return (c->wfull && c->phi && c->rfull && c->q) ? ORC_OK : ORC_E_ALLOC;
}

void orc_ctx_free(orc_eval_ctx *c) {
  free(c->wfull);
  free(c->phi);
  free(c->rfull);
  free(c->q);
  memset(c, 0, sizeof(*c));
}

/* r = phi[w] - phi_t on the unknowns 1..n (1D_FEM.c:274-277 sign); the
 * walls keep whatever c->wfull holds there (0 in Mode R). */
int orc_residual_cb(int n, const double *w, double *r, void *ctx) {
  orc_eval_ctx *c = (orc_eval_ctx *)ctx;
  memcpy(c->wfull + 1, w, sizeof(double) * n);
  double nrm;
  int rc = orc_eval(&c->p, c->wfull, c->q, c->phi, c->rfull, &nrm);
  memcpy(r, c->rfull + 1, sizeof(double) * n);
  c->nevals++;
  return rc;
}

/* adm.c's fixed-point form: xnew = x + beta (phi[x] - phi_t). */
void orc_fixed_point_cb(int n, const double *x, double *xnew, void *ctx) {
  orc_eval_ctx *c = (orc_eval_ctx *)ctx;
  orc_residual_cb(n, x, xnew, ctx);
  for (int i = 0; i < n; i++) xnew[i] = x[i] + c->beta * xnew[i];
}

/* adm_chen / deal.II sign: out = phi_t - phi (drivescft.cc:210-215). */
int orc_residual_dealii_cb(int n, const double *w, double *r, void *ctx) {
  int rc = orc_residual_cb(n, w, r, ctx);
  for (int i = 0; i < n; i++) r[i] = -r[i];
  return rc;
}

int orc_solve_anderson(const orc_problem *p, double *w_full,
                       const orc_anderson_opts *o, orc_anderson_stats *st) {
  orc_eval_ctx c;
  if (orc_ctx_init(&c, p, o->beta) != ORC_OK) return ORC_E_ALLOC;
  int n = orc_n_unknowns(p->M, p->bc);
  memcpy(c.wfull, w_full, sizeof(double) * p->M);
  orc_anderson_opts oo = *o;
  if (oo.field_floor > 0.0) oo.field_lower = -oo.field_floor * p->N;
  double *ps = NULL, *gs = NULL;
  if (oo.precond > 0.0 && !oo.pstencil) { /* this mesh's stencils */
    ps = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    if (!ps) { orc_ctx_free(&c); return ORC_E_ALLOC; }
    orc_precond_stencil(p->x, p->M, p->bc, ps);
    oo.pstencil = ps;
  }
  if (oo.gs_precond > 0.0 && !oo.gstencil) {
    gs = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    if (!gs) { free(ps); orc_ctx_free(&c); return ORC_E_ALLOC; }
    orc_gs_stencil(p->x, p->M, p->bc, gs);
    oo.gstencil = gs;
  }
  int rc = orc_anderson(w_full + 1, n, orc_residual_cb, &c, &oo, p->phi_t + 1, st);
  free(ps);
  free(gs);
  orc_ctx_free(&c);
  return rc;
}

/* Independent instances over OpenMP threads.  Instance k has M[k] nodes,
 * N[k] steps, nodes x + xoff[k], target phi_t + xoff[k], initial/returned
 * field w + xoff[k] (full-node layout).  iters[k] / status[k] receive the
 * evaluation count and the solver's return code. */
int orc_batch_solve(int n_inst, const int *M, const int *N, const long *xoff,
                    const double *x, const double *phi_t, double *w, int bc,
                    int cmat, const orc_anderson_opts *o, int threads,
                    int *iters, int *status, double *final_err) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < n_inst; k++) {
    orc_problem p = {M[k], N[k], bc, cmat, x + xoff[k], phi_t + xoff[k]};
    orc_anderson_stats st = {0, 0, 0.0, NULL, 0};
    status[k] = orc_solve_anderson(&p, w + xoff[k], o, &st);
    iters[k] = st.iterations;
    if (final_err) final_err[k] = st.final_err;
  }
  return ORC_OK;
}

/* One evaluation of each instance (no solve), OpenMP over instances:
 * r + xoff[k] receives phi - phi_t on every node, rnorm[k] the norm. */
int orc_batch_eval(int n_inst, const int *M, const int *N, const long *xoff,
                   const double *x, const double *phi_t, const double *w,
                   int bc, int cmat, int threads, double *phi, double *r,
                   double *rnorm) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int k = 0; k < n_inst; k++) {
    orc_problem p = {M[k], N[k], bc, cmat, x + xoff[k], phi_t + xoff[k]};
    double *q = (double *)malloc(sizeof(double) * (size_t)(N[k] + 1) * M[k]);
    int rc = q ? orc_eval(&p, w + xoff[k], q, phi + xoff[k], r + xoff[k],
                          &rnorm[k])
               : ORC_E_ALLOC;
    free(q);
    bad |= (rc != ORC_OK);
  }
  return bad ? ORC_E_NAN : ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Callbacks with the reference's exact signatures (no context argument), so */
/* the reference's compiled adm/broydn can drive the oracle evaluator:       */
/*   adm.c:25-27    void funcvmix(int n, double *x, double *xnew)  (flag 0)  */
/*   broydn.c:45-46 void vecfunc(int n, double v[], double f[])   (1-based)  */
/* Every call's input is appended to a trace for trajectory comparisons;     */
/* after `limit` calls the fixed-point form returns x itself, which ends     */
/* adm (its tolerance is hard-coded, adm.c:29).                              */
/* ------------------------------------------------------------------------ */
static orc_eval_ctx *g_ctx = NULL;
static double *g_trace = NULL;
static long g_trace_cap = 0, g_calls = 0, g_limit = 0;

void orc_global_setup(orc_eval_ctx *c, double *trace, long trace_cap, long limit) {
  g_ctx = c;
  g_trace = trace;
  g_trace_cap = trace_cap;
  g_calls = 0;
  g_limit = limit;
}

long orc_global_calls(void) { return g_calls; }

static void record(int n, const double *x) {
  if (g_trace && (g_calls + 1) * n <= g_trace_cap)
    memcpy(g_trace + g_calls * n, x, sizeof(double) * n);
  g_calls++;
}

void orc_global_fixed_point(int n, double *x, double *xnew) {
  record(n, x);
  if (g_limit > 0 && g_calls > g_limit) {
    memcpy(xnew, x, sizeof(double) * n);
    return;
  }
  orc_fixed_point_cb(n, x, xnew, g_ctx);
}

void orc_global_vecfunc_nr(int n, double *v, double *f) {
  record(n, v + 1);
  orc_residual_cb(n, v + 1, f + 1, g_ctx);
}

/* The traced callbacks in the oracle's own calling conventions. */
void orc_global_fixed_point_ctx(int n, const double *x, double *xnew, void *ctx) {
  (void)ctx;
  orc_global_fixed_point(n, (double *)x, xnew);
}

int orc_global_residual_ctx(int n, const double *x, double *f, void *ctx) {
  (void)ctx;
  record(n, x);
  return orc_residual_cb(n, x, f, g_ctx);
}
