/*
 * chen_shim.cpp -- C-linkage entry points for the reference's C++-built
 * routines adm_chen (DEALII_SCFT/src/ADM_chen_C.c:18-147) and spline_chen
 * (DEALII_SCFT/src/spline_chen.c:12-106), so tests can call them through
 * ctypes.  Compiled only into oracle/_ref/libnrref_chen.so.
 * TEST INFRASTRUCTURE ONLY.
 */
#include "NR_chen.h"

extern "C" int nrref_adm_chen(void (*f)(int, double *, double *), double *x,
                              double tol, int max_it, int n, double lmd,
                              int nn, int final_flag) {
  return adm_chen(f, x, tol, max_it, n, lmd, nn, final_flag != 0);
}

extern "C" void nrref_spline_chen(double *x, double *y, double *xp,
                                  double *yp, int nx, int nxp) {
  spline_chen(x, y, xp, yp, nx, nxp, nullptr);
}
