/*
 * ref_globals.c -- the globals that the reference's broydn.c requires its
 * calling program to define (broydn.c:22-28; defined by the reference at
 * 1D_FEM.c:16-21), plus an allocator for them.  Compiled only into
 * oracle/_ref/libnrref_dealii.so next to the reference's own sources.
 * TEST INFRASTRUCTURE ONLY.
 */
double **dmatrix(long nrl, long nrh, long ncl, long nch);
double *dvector(long nl, long nh);

double **qt, **r, *d, err;
int funcerr, jc, PRINT;

/* Allocate qt[1..n][1..n], r[1..n][1..n], d[1..n] the way 1D_FEM.c:346-348
 * does, and set the tolerance and Jacobian flag. */
void nrref_broydn_setup(int n, double tolf) {
  qt = dmatrix(1, n, 1, n);
  r = dmatrix(1, n, 1, n);
  d = dvector(1, n);
  jc = 0;
  funcerr = 0;
  err = tolf;
}

double nrref_get_err(void) { return err; }
int nrref_get_jc(void) { return jc; }
