/*
 * oracle_solvers.h -- CPU restatement of the reference's outer solvers
 * (gaussj, adm, broydn) and of the safeguarded Anderson mixer.
 * TEST INFRASTRUCTURE ONLY.
 */
#ifndef SCFT_ORACLE_SOLVERS_H
#define SCFT_ORACLE_SOLVERS_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_GJ_PERTURB = 0, ORC_GJ_RETURN = 1 };

int orc_gaussj(double *a, int n, double *b, int m, int zero_pivot_mode);

/* fixed-point map x -> xnew (adm.c's funcvmix, 0-based) */
typedef void (*orc_fixed_point_fn)(int n, const double *x, double *xnew,
                                   void *ctx);
int orc_adm(double *x, int n, orc_fixed_point_fn f, void *ctx, double tolf,
            int maxits, double *xnew, int *its_out, double *err_out);

/* residual map w -> r (nonzero return = evaluation failure) */
typedef int (*orc_residual_fn)(int n, const double *w, double *r, void *ctx);

typedef struct orc_anderson_opts {
  int history;          /* previous iterates kept (10) */
  double beta;          /* mixing (0.1) */
  double tol;           /* stop at ||r||_inf <= tol */
  int max_iter;         /* evaluations allowed */
  double step_cap;      /* ||w_{k+1}-w_k||_inf cap (<= 0: none) */
  double restart_factor;/* clear history when ||r|| > factor * best */
  double regularization;/* U += reg * trace(U) * I before the solve */
  int transform;        /* ORC_MIX_PLAIN / _RELATIVE / _LOG */
  double floor;         /* phi_t floor used by the transforms */
  double precond;       /* a >= 0: mix g = (I + a M_L^{-1} K) f instead of f */
  const double *pstencil; /* 3 n doubles (l, d, r per unknown) when a > 0 */
  double backtrack;     /* > 0: an evaluation that is NaN or whose error
                           exceeds backtrack * (last accepted error) halves
                           the last step and clears the history */
  int max_halvings;     /* consecutive halvings allowed before giving up */
  double field_floor;   /* > 0: after each update w >= -field_floor / dt,
                           so 1 + dt w >= 1 - field_floor and the implicit-
                           Euler matrix stays strictly diagonally dominant */
  double field_lower;   /* the bound itself, -field_floor * N: set from the
                           instance's N by orc_solve_anderson (a caller of
                           orc_anderson sets it directly) */
  double gs_precond;  /* a_w > 0: the ground-state form
                           g_u = f_u + a_w [ l_u (d_{u-1}/d_u)(f_u - f_{u-1})
                                           + r_u (d_{u+1}/d_u)(f_u - f_{u+1})
                                           + c max(w_u, 0) f_u ],
                           d = sqrt(max(phi_t, floor)), d = 0 on a wall node,
                           l, r from gstencil, instead of the plain form (needs
                           gstencil and phi_t_unk) */
  double gs_field_weight; /* c */
  int gs_switch;      /* leave the ground-state form (history cleared) after this
                           many evaluations without converging (0: never) */
  double gs_switch_growth; /* ... or when err > growth * best (0: never) */
  const double *gstencil; /* 3 n doubles (l, unused, r per unknown) of the
                             ground-state form when gs_precond > 0 */
  double gs_shift;      /* g > 0: the rational ground-state form
                           g_f = (1 - a) f + a G f + a g (g I + G0)^{-1} f,
                           G the operator above, G0 = G without its field
                           term (0: the linear form f + a G f) */
  double gs_shift_below; /* > 0: the linear form until ||r||_inf first falls
                           below this, then the rational one (history cleared
                           at the change); 0: the rational form throughout */
} orc_anderson_opts;

/* Stencil of M_L^{-1} K on the grid of unknowns alone (linear elements
 * between consecutive unknowns, lumped mass, natural rows at both ends: the
 * mixing residual is not zero at the walls):
 * (M_L^{-1} K f)_u = d_u f_u - l_u f_{u-1} - r_u f_{u+1}, d_u = l_u + r_u.
 * st receives 3 n doubles. */
int orc_precond_stencil(const double *x, int M, int bc, double *st);

/* Coefficients of the ground-state form's Laplacian, -(1/d^2) (d^2 f')',
 * on the nodes of the mesh: unknown u (node i = u + 1) couples to nodes
 * i - 1 and i + 1 with l_u = (1/h_l)/m_u and r_u = (1/h_r)/m_u, m_u = (h_l +
 * h_r)/2, where the wall nodes count as neighbours (their d is zero, so
 * their terms vanish) and a natural end (DN at x = L) has h_r = 0, r_u = 0.
 * st receives 3 n doubles (l, 0, r). */
int orc_gs_stencil(const double *x, int M, int bc, double *st);

enum { ORC_MIX_PLAIN = 0, ORC_MIX_RELATIVE = 1, ORC_MIX_LOG = 2 };

typedef struct orc_anderson_stats {
  int iterations;   /* residual evaluations made */
  int restarts;
  double final_err;
  double *err_trace; /* optional, trace_len entries */
  int trace_len;
} orc_anderson_stats;

int orc_anderson(double *w, int n, orc_residual_fn F, void *ctx,
                 const orc_anderson_opts *o, const double *phi_t_unk,
                 orc_anderson_stats *st);

/* vector function for Broyden (nonzero return = the reference's funcerr) */
typedef int (*orc_vecfunc_fn)(int n, const double *x, double *f, void *ctx);

typedef struct orc_broydn_stats {
  int iterations;
  long evaluations;
  double final_err;
  double *err_trace;
  int trace_len;
} orc_broydn_stats;

int orc_broydn(double *x, int n, orc_vecfunc_fn F, void *ctx, double tolf,
               int *jc, double *qt, double *r, double *d,
               orc_broydn_stats *st);

#ifdef __cplusplus
}
#endif
#endif
