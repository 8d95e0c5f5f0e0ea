/*
 * scft.h -- C ABI of libscft_hip.so, the MI355X (gfx950) implementation of
 * giantsda/SCFT's inner field evaluation: linear-FEM assembly, the
 * implicit-Euler propagator march, the Romberg density phi(x) and the
 * residual phi - phi_t, plus the outer Anderson/Broyden loops that drive the
 * residual to zero.  All arithmetic on the solver path is FP64.
 *
 * Plain C types only (pointers, sizes, doubles); every function returns an
 * int status (SCFT_OK = 0) unless it replaces a reference symbol whose
 * signature says otherwise.  Nothing here calls exit().
 *
 * Index conventions
 *   node i = 0..M-1 of a mesh with M nodes (m = M-1 elements); x_0 = 0 is the
 *   hard wall.  The unknowns are the nodes that are not Dirichlet:
 *     SCFT_BC_DD (walls at both ends, the reference problem): nodes 1..M-2
 *     SCFT_BC_DN (wall at 0, zero-flux end at L):              nodes 1..M-1
 *   Arrays "on unknowns" hold n values, entry u <-> node u+1.
 */
#ifndef SCFT_H
#define SCFT_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SCFT_OK 0
#define SCFT_E_ARG 1      /* invalid argument (sizes, N not 2^k >= 16, ...) */
#define SCFT_E_HIP 2      /* HIP runtime error (no device, launch failure)  */
#define SCFT_E_NAN 3      /* non-finite value produced by an evaluation     */
#define SCFT_E_PIVOT 4    /* non-positive tridiagonal pivot (D not an M-matrix) */
#define SCFT_E_NOMEM 5    /* device or host allocation failed               */
#define SCFT_E_STATE 6    /* call out of order (no init / no context)       */
#define SCFT_E_NOCONV 7   /* outer solver hit its iteration limit           */

#define SCFT_BC_DD 0
#define SCFT_BC_DN 1
#define SCFT_C_ROWSCALED 0  /* C_ij = w_i A_ij (1D_FEM.c:104-105)            */
#define SCFT_C_CONSISTENT 1 /* C_ij = int w_h phi_i phi_j (scft.cc:641-651)   */
/* contour stepper of the propagator march */
#define SCFT_STEP_IE 0   /* implicit Euler (1D_FEM.c:208-228, IE march of the MATLAB files) */
#define SCFT_STEP_IRK4 1 /* two-stage Gauss-Legendre (DEALII_SCFT/src/scft.cc:671-695,
                            drivescft.cc:130-146, Matlab_files/IRK4.m:96-120): fourth
                            order in s, A- but not L-stable (the step factor of a mode
                            tends to 1 as dt*lambda grows), so meant for
                            dt * 12/h_min^2 up to ~25 as in the reference (33 nodes,
                            N = 2048); at most 2048 unknowns per instance */

/* per-instance status word bits (scft_batch_* outputs) */
#define SCFT_ST_CONVERGED 1u
#define SCFT_ST_NAN 2u
#define SCFT_ST_PIVOT 4u
#define SCFT_ST_MAXITER 8u
#define SCFT_ST_FALLBACK 16u /* warm start rejected, restarted cold */
#define SCFT_ST_FAILED 32u   /* evaluation failed (NaN / pivot) beyond recovery */
#define SCFT_ST_PLAIN_FORM 64u /* left the ground-state form (gs_switch) */
#define SCFT_ST_RATIONAL 128u  /* reached the rational form (gs_shift_below) */

/* ---- library lifetime ---------------------------------------------------
 * Replaces PetscInitialize/PetscFinalize around the evaluator
 * (1D_FEM.c:295,363).  scft_init is optional: the drop-in evaluator
 * initialises lazily with the defaults below. */
typedef struct scft_config {
  int device;        /* HIP device ordinal (default 0)                      */
  int N;             /* contour steps, 2^k >= 16 (default 2048, drivescft.cc:130) */
  double L;          /* domain length (default 3.72374357332160)            */
  double tau;        /* target wall width (default 0.530252230020752)       */
  int bc;            /* SCFT_BC_DD (default, the reference problem) / _DN   */
  int cmat;          /* SCFT_C_ROWSCALED (default) / SCFT_C_CONSISTENT      */
  double beta;       /* mixing used by scft_fixed_point (default 0.1)       */
  int verbose;       /* print each evaluation like 1D_FEM.c:280-283 (default 0) */
  int stepper;       /* SCFT_STEP_IE (default) / SCFT_STEP_IRK4 (the deal.II path) */
} scft_config;

void scft_default_config(scft_config *cfg);
int scft_init(const scft_config *cfg);
int scft_finalize(void);
const char *scft_strerror(int code);
int scft_last_status(void); /* status of the last drop-in evaluation */

/* ---- reference drop-in evaluator ----------------------------------------
 * Replaces `int simple_FEM_1D_transient(int n, double *w, double *out)`
 * (1D_FEM.c:47-287): n interior unknowns, NR 1-based w[1..n] in and
 * out[1..n] = phi - phi_t out (1D_FEM.c:274-277), walls at w = 0
 * (1D_FEM.c:65-66).  Mode R numerics: N steps paired q_j q_{N-j} and
 * Romberg-integrated as DEALII_SCFT/drivescft.cc:184-193.  The target is the
 * caller's global `f0_given` (0-based, n+2 values, 1D_FEM.c:14) when that
 * pointer is non-NULL, else the position-based profile of
 * DEALII_SCFT/src/scft.cc:188-215 (or whatever scft_set_target installed).
 * Increments `local_iteration` (1D_FEM.c:279).  On failure sets the
 * caller-visible `funcerr` (broydn.c:25,37,82) and returns nonzero.
 * Synchronous: returns after the device work and the copy back. */
int simple_FEM_1D_transient(int n, double *w, double *out);
extern double *f0_given;
extern int local_iteration;
extern int funcerr;

/* Install the target used by the drop-in entries when f0_given is NULL:
 * phi_t at all n+2 nodes (0-based).  NULL restores the built-in profile. */
int scft_set_target(int n_nodes, const double *phi_t);

/* adm.c callback form (adm.c:25-27, funcvmix): xnew = x + beta (phi[x] - phi_t)
 * on the n unknowns, x[0..n-1] (flag 0) or x[1..n] (flag 1, NR offset). */
void scft_fixed_point(int n, double *x, double *xnew);
void scft_fixed_point_nr(int n, double *x, double *xnew);

/* adm_chen / deal.II callback form (NR_chen.h:31-34, drivescft.cc:218-243):
 * 0-based, out = phi_t - phi (drivescft.cc:210-215 sign). */
void scft_adm_chen_f(int n, double *in, double *out);

/* ---- batched context ----------------------------------------------------
 * Many independent instances of uneven size evaluated in one persistent
 * launch per bucket of sizes.  Instance k: M[k] nodes at x + xoff[k]
 * (xoff in doubles into the concatenated node array), N[k] contour steps
 * (2^k >= 16), target phi_t on its unknowns (concatenated, instance k at
 * uoff[k] as reported by scft_batch_layout).  All inputs are copied. */
typedef struct scft_batch scft_batch;

typedef struct scft_batch_desc {
  int n_inst;
  const int *M;        /* nodes per instance                                */
  const int *N;        /* contour steps per instance                        */
  const long long *xoff;/* offset of instance k's nodes in x                */
  const double *x;     /* concatenated node coordinates                     */
  const double *phi_t; /* concatenated targets on unknowns (uoff layout)    */
  int bc;              /* SCFT_BC_DD / SCFT_BC_DN (all instances)           */
  int cmat;            /* SCFT_C_ROWSCALED / SCFT_C_CONSISTENT              */
  const char *tune_file; /* autotuner table to read/write, or NULL          */
  int stepper;         /* SCFT_STEP_IE (0, also a zero-initialised desc) /
                          SCFT_STEP_IRK4; a tune file holds one stepper's table */
} scft_batch_desc;

int scft_batch_create(scft_batch **out, const scft_batch_desc *desc);
int scft_batch_destroy(scft_batch *b);
/* offsets of each instance's unknowns (length n_inst) and the total */
int scft_batch_layout(const scft_batch *b, long long *uoff, long long *total);

/* One field evaluation of every instance (host buffers on unknowns):
 * phi (nullable), r = phi - phi_t, rnorm[k] = ||r_k||_inf, status[k]. */
int scft_batch_eval(scft_batch *b, const double *w, double *phi, double *r,
                    double *rnorm, unsigned *status);

/* Field values at the two walls, used by the consistent w-mass
 * (SCFT_C_CONSISTENT) of the wall elements: the deal.II path extrapolates
 * them from the interior field by a natural spline (DEALII_SCFT/src/scft.cc:
 * 454-480); the default is 0, the convention of 1D_FEM.c:65-66.  One value
 * per instance in each array (wallL is not read under SCFT_BC_DN). */
int scft_batch_set_walls(scft_batch *b, const double *wall0, const double *wallL);

/* Diagnostic for parity tests: the q-slabs 1..N/2-1 kept in device scratch
 * by the last evaluation of a single-instance batch, in node order
 * (out[(s-1)*n + u] = q(x_{u+1}, s/N)); out holds (N/2-1)*n doubles. */
int scft_debug_slabs(scft_batch *b, int inst, double *out);

/* Device-resident safeguarded Anderson mixing of every instance until
 * ||r||_inf <= tol or max_iter evaluations:
 *   w_{k+1} = wbar + beta rbar, theta from the history Gram system,
 *   r mixed as is / relative / log(phi/phi_t), the step capped in inf-norm,
 *   history cleared on growth past restart_factor * best or singular U. */
typedef struct scft_anderson_opts {
  int history;            /* previous iterates in the Gram system, 1..16 (10) */
  double beta;            /* 0.1 */
  double tol;             /* 1e-6 */
  int max_iter;           /* evaluation limit per instance */
  double step_cap;        /* inf-norm cap of w_{k+1} - w_k (<= 0: none) */
  double restart_factor;  /* growth restart (huge = never) */
  double regularization;  /* U += reg * trace(U) * I */
  int transform;          /* 0 plain, 1 relative, 2 log */
  double floor;           /* phi_t floor used by the transforms */
  double fallback_factor; /* warm starts whose first ||r|| exceed
                             fallback_factor * median restart cold (<=0: off) */
  int check_every;        /* host polls the active count every k sweeps */
  double progress_seconds;/* > 0: print a progress line to stderr this often */
  double precond;         /* a >= 0: mix g = (I + a M_L^{-1} K) f, the lumped-
                             mass Laplacian of the unknowns' grid with natural
                             ends (0: off, the reference's plain mixing) */
  double backtrack;       /* > 0: a NaN / failed evaluation, or one whose
                             ||r|| exceeds backtrack * the last accepted one,
                             halves the last step and clears the history */
  int max_halvings;       /* consecutive halvings before the instance fails */
  double field_floor;     /* in [0, 1): > 0 bounds every update below,
                             w >= -field_floor / dt (dt = 1/N of the
                             instance), so 1 + dt w >= 1 - field_floor and the
                             implicit-Euler matrix stays strictly diagonally
                             dominant (0: off, the reference's update) */
  double gs_precond;      /* a > 0: mix the ground-state form instead of
                             the plain one,
                               g_u = f_u + a [ l_u (d_{u-1}/d_u)(f_u - f_{u-1})
                                             + r_u (d_{u+1}/d_u)(f_u - f_{u+1})
                                             + c max(w_u, 0) f_u ],
                             d = sqrt(max(phi_t, floor)) (0 on a wall node),
                             l_u, r_u the linear-element Laplacian's
                             couplings to the neighbouring nodes: the
                             operator -(1/d^2)(d^2 f')' + c w+, which maps
                             constants to themselves.  With a = 0.45 and
                             c = 0.25: from predicted starts ~2.4x fewer
                             evaluations than the plain form, from w0 ~2x
                             fewer (0: off) */
  double gs_field_weight; /* c of gs_precond (0.25) */
  int gs_switch;          /* an instance leaves the ground-state form for
                             the plain one (history cleared,
                             SCFT_ST_PLAIN_FORM) after this many evaluations
                             without converging (0: never) ... */
  double gs_switch_growth;/* ... or when its ||r||_inf exceeds this factor
                             times its best so far (0: never) */
  double gs_shift;        /* g > 0: the rational ground-state form
                               g = (1 - a) f + a G f + a g (g I + G0)^{-1} f,
                             G the operator above, G0 = G without its field
                             term (so the factors of g I + G0 depend on the
                             mesh and phi_t only and are computed once per
                             batch, ~ms): its symbol follows the
                             inverse Debye function of the chain's response
                             to ~2% (a = 1/2, g = 3) where the linear form
                             f + a G f follows it to ~13%.  ~20% fewer
                             evaluations from predicted starts, none fewer
                             from w0 (0: the linear form) */
  double gs_shift_below;  /* > 0: the linear form until an instance's
                             ||r||_inf first falls below this, then the
                             rational one (history cleared, status
                             SCFT_ST_RATIONAL); from w0 with 0.5: ~25% fewer
                             evaluations than the linear form throughout.
                             0: the rational form from the first update */
  int inner_sweeps;       /* k >= 1 (0 counts as 1): within each sweep after
                             the first, the instances of the size buckets
                             with at most a third of the work of the
                             largest go through k evaluations and mixer
                             updates, on their own streams while the
                             largest marches once.  Every instance takes
                             the same evaluations and updates in the same
                             order as with 1 (the results are the same
                             bits); the solve needs fewer sweeps when small
                             instances need more evaluations than large
                             ones.  At most 16 (1) */
} scft_anderson_opts;

void scft_default_anderson(scft_anderson_opts *o);

typedef struct scft_solve_stats {
  int sweeps;             /* batched sweeps launched */
  int converged;          /* instances with ||r|| <= tol */
  double seconds;         /* wall time of the solve (device resident) */
  double eval_seconds;    /* part spent in the march kernels (events) */
  double node_steps;      /* sum over evaluations of n * N */
  double bytes;           /* sum over evaluations of the algorithmic bytes */
} scft_solve_stats;

/* w: in = initial fields, out = solutions (host, on unknowns); w_cold
 * (nullable) = cold-start fields used by the fallback; iters[k] / status[k]
 * (nullable) per instance.  Equivalent to upload + solve_resident + download;
 * stats->seconds covers the device-resident part only.  Returns SCFT_OK when
 * every instance converged, SCFT_E_NOCONV otherwise. */
int scft_batch_solve(scft_batch *b, const scft_anderson_opts *o, double *w,
                     const double *w_cold, int *iters, unsigned *status,
                     scft_solve_stats *stats);
/* The three phases separately: host fields into HBM; the solve with inputs
 * and outputs resident in HBM (graph-launched sweeps, the host polls a
 * 4-byte active count every check_every sweeps); results back to the host. */
int scft_batch_upload(scft_batch *b, const double *w, const double *w_cold);
int scft_batch_solve_resident(scft_batch *b, const scft_anderson_opts *o,
                              scft_solve_stats *stats);
int scft_batch_download(scft_batch *b, double *w, int *iters, unsigned *status,
                        scft_solve_stats *stats);

/* Timing helpers used by bench.py: time `reps` evaluations of the current
 * device field with HIP events on the launch stream. */
int scft_batch_time_eval(scft_batch *b, const double *w, int reps,
                         double *ms_per_eval, double *march_ms);
/* Per size bucket (in the order of scft_batch_tuning): average duration of one persistent march
 * launch over `reps` (HIP events on the launch stream), its algorithmic HBM
 * bytes (q-slabs written and read back plus O(n) I/O) and instance count. */
int scft_batch_time_buckets(scft_batch *b, const double *w, int reps,
                            int max_buckets, int *n_buckets, double *ms,
                            double *bytes, int *n_items);

/* Autotuner: time every (nodes-per-lane, waves) variant that fits each size
 * bucket and keep the fastest; written to desc.tune_file when given.  Ends
 * with scft_batch_tune_placement. */
int scft_batch_autotune(scft_batch *b, int reps);
/* No counterpart in the reference.  The march's speed depends on which
 * physical pages back the slab scratch (on MI355X the same launch ran 20-25%
 * faster on some allocations than on others, whatever the virtual address).
 * Tries up to `tries` further allocations of the scratch, each made while the
 * earlier ones are held so that it gets other pages, times a whole sweep of
 * the current device field on each (`reps` sweeps) and keeps the fastest;
 * the others are freed.  Stops early when free device memory would fall
 * below a quarter of the scratch plus 1 GiB.  Cannot be stored in a tune
 * file: call it after creating (or re-tuning) a batch. */
int scft_batch_tune_placement(scft_batch *b, int tries, int reps);
int scft_batch_tuning(const scft_batch *b, int *n_buckets, int *bucket_nmax,
                      int *K, int *W, int max_buckets);

/* ---- evaluation without allocation ----------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): scft_batch_eval of ONE instance with inputs and outputs
 * in the caller's DEVICE memory -- d_x (M node coordinates), d_w and d_phi_t
 * (n unknowns) in; d_phi (n, nullable), d_r = phi - phi_t (n), d_rnorm
 * (1 double) and d_status (1 word, SCFT_ST_* bits) out.  Walls at w = 0, the
 * implicit-Euler stepper, the march variant of a one-instance batch (so the
 * same bits as scft_batch_eval).  Calls no allocating routine: the working
 * memory (descriptor, pairing weights, work counter, q-slabs) is a fixed
 * device array of 64 MiB + 256 KiB per GPU that the library holds until the
 * process ends.  The slabs need (N/2) * cap * 8 bytes, cap = the power of
 * two >= max(n, 64): up to n = 4096 with N = 4096, n = 2048 with N = 8192,
 * n = 1024 with N = 16384; larger instances return SCFT_E_ARG (message in
 * scft_noalloc_error()).  Every call on a GPU uses that GPU's one array, so
 * the library serialises calls within a process and calls that overlap on
 * one GPU from elsewhere would race.  `stream`: a hipStream_t (NULL: the
 * null stream); returns after the work on it has finished. */
int scft_eval_noalloc(int M, int N, int bc, int cmat, const double *d_x,
                      const double *d_w, const double *d_phi_t, double *d_phi,
                      double *d_r, double *d_rnorm, unsigned *d_status,
                      void *stream);
const char *scft_noalloc_error(void);

/* ---- Broyden with a batched Jacobian ------------------------------------
 * Replaces broydn(x, n, &check, simple_FEM_1D_transient) (1D_FEM.c:356,
 * broydn.c:44-292 with DEALII_SCFT/src/{fdjac,lnsrch,qrdcmp,qrupdt,rotate,
 * rsolv}.c; written from Dennis & Schnabel's secant method with the
 * reference's constants) for one instance: F(w) = phi[w] - phi_t, the
 * n-column forward-difference Jacobian of fdjac.c:9-41 evaluated as ONE
 * batched sweep of n perturbed fields, the other evaluations on the device
 * one at a time, the n x n QR / rank-1 updates on the host (O(n^3) per
 * Jacobian: meant for n up to ~1000).  Instance as in scft_batch_desc (M
 * nodes at x, N steps, phi_t on the n unknowns, bc, cmat; with
 * SCFT_C_CONSISTENT every field evaluated gets its wall values by natural-
 * spline extrapolation, DEALII_SCFT/src/scft.cc:454-480); w on the unknowns
 * (0-based): in = start, out = result; tolf = the reference's `err`.
 * Returns SCFT_OK when broydn's check would be 0, SCFT_E_NOCONV where it
 * would be 1 (or where the reference would exit), SCFT_E_NAN / SCFT_E_PIVOT
 * for a failed evaluation / singular Jacobian. */
typedef struct scft_broyden_stats {
  int iterations;     /* Broyden iterations */
  long evaluations;   /* residual evaluations, Jacobian columns included */
  long sweeps;        /* device launches: one per single evaluation, one per Jacobian */
  double final_err;   /* ||F||_inf at return */
  int check;          /* broydn's check flag */
} scft_broyden_stats;
int scft_broyden_solve(int M, int N, const double *x, const double *phi_t,
                       int bc, int cmat, double tolf, double *w,
                       scft_broyden_stats *stats);
/* The same iteration for a residual the caller computes: replaces
 * broydn(x, n, &check, vecfunc) (broydn.c:44-46) with the callback
 * F(n, w, f, ctx) on 0-based arrays, returning nonzero where the reference's
 * vecfunc would set funcerr (broydn.c:37,82).  Each Jacobian is n further
 * calls of F, one per column (fdjac.c:18-34).  Host only: no device work. */
int scft_broyden_solve_fn(int n, int (*F)(int n, const double *w, double *f, void *ctx),
                          void *ctx, double tolf, double *w,
                          scft_broyden_stats *stats);

/* ---- reference file formats ---------------------------------------------
 * .res files (Exp_m32_n2048_IE.res: the reader of 1D_FEM.c:321-342 skips
 * 10 lines and takes column 3 as the interior field): header values and
 * every table row (x/l, phi, eta, phie, phij), the wall rows included, so
 * rows 1..M-2 are 1D_FEM.c's yita_middle.  *rows receives the row count
 * (arrays nullable, at most max_rows filled). */
typedef struct scft_res_header {
  int m, n;        /* elements, contour steps ("m = 32   n = 2048") */
  int iters;       /* "N=1000" */
  double l, cp, mphi, t, Z, f, fmax;
} scft_res_header;
int scft_read_res(const char *path, int max_rows, double *x_over_l, double *phi,
                  double *eta, double *phie, double *phij, int *rows,
                  scft_res_header *hdr);
/* solution_yita_1D_N=*.txt (DEALII_SCFT/src/scft.cc:318-334): "N= M, ERROR=
 * e", "mean_field_free_energy, F", then "i,x_i,w_i" per node. */
int scft_write_solution(const char *path, int M, const double *x,
                        const double *w, double err, double free_energy);
/* Reader of that format (scft_util.cc:13-41); files without the free-energy
 * line (older MATLAB output) are accepted and *free_energy is NaN. */
int scft_read_solution(const char *path, int cap, int *M, double *x, double *w,
                       double *err, double *free_energy);
/* Mean-field free energy (DEALII_SCFT/src/scft.cc:404-450): the field linear
 * between its M nodes, times the position-based target (mirrored when
 * `mirror`), Romberg-integrated on 2^log2_points + 1 uniform points of
 * [0, L] (the reference uses 2^18 + 1); F = (I / f0bar / L + ln f0bar) /
 * -1000 with f0bar = 0.892581217773656 (scft.cc:414). */
int scft_free_energy(int M, const double *x, const double *w_nodes, double L,
                     double tau, int mirror, int log2_points, double *F);

/* ---- mesh continuation ---------------------------------------------------
 * One refinement by the rule of Matlab_files/refine_mesh.m:1-36: cell
 * errors |dw/dx| (wall cells +inf), cells with error >= factor * median
 * (the reference: 10) bisected; the field moved to the new unknowns by the
 * not-a-knot spline through the old ones (refine_mesh.m:33, spline_chen).
 * *M_new receives the new node count; x_new / w_new (capacity cap) the new
 * nodes and field (wall values kept); with x_new or w_new NULL only the count
 * is returned.  bc: the far end is a wall for DD. */
int scft_refine_mesh(int M, const double *x, const double *w_nodes,
                     double factor, int bc, int cap, int *M_new, double *x_new,
                     double *w_new);

/* ---- warm-start MLP (no reference counterpart) -------------------------
 * Dense FP32 network trained and evaluated on the device by hand-written
 * kernels (matrix-core GEMMs with fused bias / tanh / tanh' epilogues, MSE,
 * Adam): tanh hidden layers, linear output.  dims[0] inputs ... dims[n_layers]
 * outputs; parameters packed per layer as W (dims[l+1] x dims[l], row-major)
 * then b (dims[l+1]); weights start Glorot-uniform from splitmix64(seed),
 * biases at 0. */
typedef struct scft_mlp scft_mlp;
int scft_mlp_create(scft_mlp **out, int n_layers, const int *dims,
                    unsigned long long seed);
int scft_mlp_destroy(scft_mlp *m);
int scft_mlp_param_count(const scft_mlp *m, long long *count);
int scft_mlp_get_params(const scft_mlp *m, float *params);
int scft_mlp_set_params(scft_mlp *m, const float *params); /* resets Adam */
/* x' = (x - x_mean) / x_std on inputs and y = y' y_std + y_mean on outputs,
 * applied by scft_mlp_forward and scft_mlp_train; all NULL clears them. */
int scft_mlp_set_normalisation(scft_mlp *m, const float *x_mean,
                               const float *x_std, const float *y_mean,
                               const float *y_std);
int scft_mlp_forward(scft_mlp *m, int batch, const float *X, float *Y);
/* loss = mean((Y - T)^2) over batch x outputs and its gradient (grad
 * nullable), on raw rows (no normalisation): for gradient checks. */
int scft_mlp_loss_grad(scft_mlp *m, int batch, const float *X, const float *T,
                       double *loss, float *grad);
int scft_mlp_adam_step(scft_mlp *m, float lr); /* b1 0.9, b2 0.999, eps 1e-8 */
/* Mini-batch Adam, reshuffled each epoch from splitmix64(seed); mean loss of
 * each epoch (normalised units) into epoch_loss (nullable). */
int scft_mlp_train(scft_mlp *m, int n_samples, const float *X, const float *Y,
                   int epochs, int batch, float lr, unsigned long long seed,
                   double *epoch_loss);

/* Warm start written into a batch on the device (no reference counterpart):
 * the network's forward pass on inputs kept in HBM, then one kernel that maps
 * each instance's outputs onto its unknowns and writes them into the batch's
 * resident field (what scft_batch_upload's w would have written), so no field
 * crosses PCIe between the prediction and scft_batch_solve_resident.  Output
 * row of instance k (after the network's output normalisation): n_grid
 * values at the points grid_x[k * n_grid ...] (increasing), interpolated by
 * the not-a-knot spline of scft_resample_spline, then n_node samples
 * asinh(w / node_scale) at the 1-based node numbers node_idx (increasing);
 * the first wall_count[k] unknowns (nodes 1..wall_count[k], at most
 * node_idx[n_node-1]) take the samples interpolated linearly in the node
 * number.  inputs: n_inst rows of n_in raw features (normalised here with the
 * network's input normalisation as it stands: create the plan after
 * training).  n_grid <= 256, n_node <= 32, dims[last] = n_grid + n_node.
 * The plan keeps pointers to the batch and the network: destroy it first. */
typedef struct scft_warm_plan scft_warm_plan;
int scft_warm_plan_create(scft_warm_plan **out, scft_batch *b, scft_mlp *net, int n_in,
                          const float *inputs, int n_grid, const double *grid_x, int n_node,
                          const int *node_idx, double node_scale, const int *wall_count);
/* The same for a subset of the batch's instances (one network per range of
 * mesh sizes, one plan each): row k of inputs, grid_x and wall_count belongs
 * to instance sel[k] (0-based, increasing, 1 <= n_sel <= n_inst); the other
 * instances' fields are left as they are. */
int scft_warm_plan_create_sel(scft_warm_plan **out, scft_batch *b, scft_mlp *net, int n_sel,
                              const int *sel, int n_in, const float *inputs, int n_grid,
                              const double *grid_x, int n_node, const int *node_idx,
                              double node_scale, const int *wall_count);
int scft_warm_plan_predict(scft_warm_plan *p); /* synchronous on the batch's stream */
int scft_warm_plan_destroy(scft_warm_plan *p);

/* Cubic spline through (x_src, y_src) (x increasing, n_src >= 4) with the
 * not-a-knot end conditions of DEALII_SCFT/src/spline_chen.c:13-107 (the
 * reference's m == NULL case), evaluated at x_dst; points outside the table
 * use the end intervals' cubics, as there.  O(n) banded solve instead of the
 * reference's dense gaussj. */
int scft_resample_spline(int n_src, const double *x_src, const double *y_src,
                         int n_dst, const double *x_dst, double *y_dst);
/* The same for n_tables independent tables in one call: table k has
 * n_src[k] points at x_src + src_off[k] (values y_src + src_off[k]) and is
 * evaluated at the n_dst[k] points x_dst + dst_off[k] into y_dst + dst_off[k]. */
int scft_resample_spline_batch(int n_tables, const int *n_src,
                               const long long *src_off, const double *x_src,
                               const double *y_src, const int *n_dst,
                               const long long *dst_off, const double *x_dst,
                               double *y_dst);

#ifdef __cplusplus
}
#endif
#endif
