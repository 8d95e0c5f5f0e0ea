/*
 * scft_oracle.h -- CPU restatement of giantsda/SCFT's inner field evaluation.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg use it, and only as the checker.
 *
 * Conventions (all FP64):
 *   M      number of mesh nodes (m = M-1 linear elements)
 *   N      number of implicit-Euler contour steps, dt = 1/N (Mode R)
 *   x[M]   node coordinates, x[0] = 0 is the hard wall
 *   w[M]   field at every node (the reference's yita/eta)
 *   q      propagator slabs, slab-major: q[j*M + i] = q(x_i, s_j)
 *   bc     ORC_BC_DD: Dirichlet q=0 at both ends (1D_FEM.c:179-182)
 *          ORC_BC_DN: Dirichlet at x=0, natural (zero-flux) row at x=L
 *                     (Matlab_files/simple_FEM_1D_transient.m:45-47,61-63)
 *   cmat   ORC_C_ROWSCALED: C_ij = w_i A_ij      (1D_FEM.c:104-105)
 *          ORC_C_CONSISTENT: C_ij = int w_h phi_i phi_j, exact 2-point Gauss
 *                            (DEALII_SCFT/src/scft.cc:610-666)
 * Unknowns are the non-Dirichlet nodes: 1..M-2 (DD) or 1..M-1 (DN).
 */
#ifndef SCFT_ORACLE_H
#define SCFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_DD = 0, ORC_BC_DN = 1 };
enum { ORC_C_ROWSCALED = 0, ORC_C_CONSISTENT = 1 };
enum { ORC_STEP_IE = 0, ORC_STEP_IRK4 = 1 };

#define ORC_OK 0
#define ORC_E_ARG 1
#define ORC_E_PIVOT 2
#define ORC_E_NAN 3
#define ORC_E_ALLOC 4

/* ---- target density (a1) ---- */
double orc_exp_profile(double x, double tau);
void orc_target_position(int M, const double *x, double tau, int mirror,
                         double *phi_t);
void orc_target_count(int M, const double *x, double tau, double L,
                      int nan_guard, double *phi_t);

/* ---- meshes (a2) ---- */
void orc_mesh_uniform(int M, double L, double *x);
int orc_graded_elements(double L, double h0, double ell, int m_min, int m_max);
void orc_mesh_graded(int M, double L, double ell, double *x);

/* ---- assembly (a3) ---- */
void orc_assemble(int M, const double *x, const double *w, int cmat,
                  double *A3, double *B3, double *C3);
void orc_system(int M, const double *x, const double *w, double dt, int bc,
                int cmat, double *D3, double *R3);

/* ---- march (a4, a5) ---- */
int orc_thomas(int M, const double *D3, const double *rhs, double *out);
int orc_march(int M, int N, const double *x, const double *w, int bc,
              int cmat, double *q);
int orc_march_ld(int M, int N, const double *x, const double *w, int bc,
                 int cmat, double *q);
/* Two-stage Gauss-Legendre (IRK4) march of the deal.II path: the 2x2-block
 * stage system solved as written (block Thomas), q <- q + dt/2 (k1 + k2). */
int orc_march_irk4(int M, int N, const double *x, const double *w, int bc,
                   int cmat, double *q);

/* ---- quadrature (a6) ---- */
double orc_romint(const double *f, int m, double hh);
int orc_romberg_weights(int m, double hh, double *c);
double orc_trapezoid(const double *f, int m, double hh);

/* ---- field evaluation (a3-a8) ---- */
typedef struct orc_problem {
  int M, N, bc, cmat;
  const double *x;
  const double *phi_t;
} orc_problem;

int orc_phi_romint(int M, int N, const double *q, double *phi);
int orc_phi_weights(int M, int N, const double *q, const double *c,
                    double *phi);
int orc_eval(const orc_problem *p, const double *w, double *q_out,
             double *phi, double *r, double *rnorm);
int orc_eval_stepper(const orc_problem *p, int stepper, const double *w,
                     double *q_out, double *phi, double *r, double *rnorm);
int orc_n_unknowns(int M, int bc);
int orc_first_unknown(void);

/* Mode M (MATLAB-compatible): `steps` IE steps of size dt, offset trapezoid,
 * out = phi_t - phi at the interior nodes 1..M-2 (length M-2). */
int orc_eval_matlab(int M, const double *x, const double *w,
                    const double *phi_t, int steps, double dt, double *out);

/* ---- exact sine-mode propagator for constant w, uniform mesh, DD ---- */
void orc_sine_mode_exact(int M, double L, double c, int N, int j,
                         double *q_j);

#ifdef __cplusplus
}
#endif
#endif
