/*
 * oracle_batch.h -- CPU drivers around the oracle evaluator.
 * TEST INFRASTRUCTURE ONLY.
 */
#ifndef SCFT_ORACLE_BATCH_H
#define SCFT_ORACLE_BATCH_H

#include "oracle_solvers.h"
#include "scft_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_eval_ctx {
  orc_problem p;
  double beta;
  double *wfull, *phi, *rfull, *q;
  long nevals;
} orc_eval_ctx;

int orc_ctx_init(orc_eval_ctx *c, const orc_problem *p, double beta);
void orc_ctx_free(orc_eval_ctx *c);
int orc_residual_cb(int n, const double *w, double *r, void *ctx);
void orc_fixed_point_cb(int n, const double *x, double *xnew, void *ctx);
int orc_residual_dealii_cb(int n, const double *w, double *r, void *ctx);

int orc_solve_anderson(const orc_problem *p, double *w_full,
                       const orc_anderson_opts *o, orc_anderson_stats *st);

int orc_batch_solve(int n_inst, const int *M, const int *N, const long *xoff,
                    const double *x, const double *phi_t, double *w, int bc,
                    int cmat, const orc_anderson_opts *o, int threads,
                    int *iters, int *status, double *final_err);

int orc_batch_eval(int n_inst, const int *M, const int *N, const long *xoff,
                   const double *x, const double *phi_t, const double *w,
                   int bc, int cmat, int threads, double *phi, double *r,
                   double *rnorm);

void orc_global_setup(orc_eval_ctx *c, double *trace, long trace_cap, long limit);
long orc_global_calls(void);
void orc_global_fixed_point(int n, double *x, double *xnew);
void orc_global_vecfunc_nr(int n, double *v, double *f);
void orc_global_fixed_point_ctx(int n, const double *x, double *xnew, void *ctx);
int orc_global_residual_ctx(int n, const double *x, double *f, void *ctx);

#ifdef __cplusplus
}
#endif
#endif
