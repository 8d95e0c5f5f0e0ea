// irk4.hip -- the two-stage Gauss-Legendre (IRK4) propagator march on gfx950,
// FP64: the contour stepper of the deal.II path (DEALII_SCFT/src/scft.cc:
// 671-695, DEALII_SCFT/drivescft.cc:130-146, Matlab_files/IRK4.m:96-120).
//
// The reference solves, every step, the 2n x 2n stage system
//     (I (x) A + dt Ah (x) D') [k1; k2] = [1; 1] (x) (-D' q),   D' = B + C,
//     Ah = [[1/4, 1/4 - sqrt3/6], [1/4 + sqrt3/6, 1/4]],
// and sets q <- q + dt/2 (k1 + k2).  Ah has the complex pair of eigenvalues
// lambda = 1/4 +- i sqrt3/12.  In Ah's eigenbasis the stage system splits into
// two complex-conjugate tridiagonal systems, and the update becomes
//     q <- q + dt Re(omega z),   (A + dt lambda D') z = -D' q,
// omega = 1 - i sqrt3 (the product of the first entries of 1^T V and
// V^{-1} 1 for Ah's eigenvector matrix V; on a scalar problem q' = -mu q this
// reproduces the Gauss method's factor (1 - y/2 + y^2/12) / (1 + y/2 + y^2/12),
// y = mu dt).  So a step is one complex tridiagonal solve, with the
// implicit-Euler march's structure (march.hip) carried over complex numbers:
//
//   setup   rows of A and D' from the mesh and the field, S = A + dt lambda D'
//           (complex), Thomas pivots of S from a wave/workgroup prefix product
//           of complex 2x2 Moebius matrices in double-double (the pivot
//           recurrence is nearly neutral when dt/h^2 >> 1, exactly as for
//           implicit Euler), then per row alpha = a/u, gamma = c/u and
//           rho = -dt omega / u;
//   step    r'_i = rho_i (D' q)_i, z_i = r'_i - alpha_i z_{i-1},
//           x_i = z_i - gamma_i x_{i+1}, q_i += Re x_i; each recurrence as a
//           lane-local pass, a 16-lane DPP row scan, two cross-row levels
//           (broadcast DPP forward, LDS-crossbar permutes backward) and a
//           fold over waves through LDS, with every multiplier fixed per
//           instance;
//   density and residual exactly as march.hip: slabs 1 .. N/2-1 streamed to
//           HBM and paired with steps N/2+1 .. N through a register ring,
//           Romberg pairing weights in LDS, r = phi - phi_t, ||r||_inf.
//
// Per step and row this is 14 flops against implicit Euler's ~6, and twice
// the lane moves; the HBM traffic per evaluation is the same.
#include "scft_internal.h"
#include "device_common.h"

namespace scft {

// ------------------------------------------------------------ complex FP64
struct cd {
  double re, im;
};
__device__ __forceinline__ cd c_add(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd c_mul(cd a, cd b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * b + c
__device__ __forceinline__ cd c_fma(cd a, cd b, cd c) {
  return {fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im))};
}
// c - a * b
__device__ __forceinline__ cd c_fnma(cd a, cd b, cd c) {
  return {fma(-a.re, b.re, fma(a.im, b.im, c.re)), fma(-a.re, b.im, fma(-a.im, b.re, c.im))};
}
__device__ __forceinline__ cd c_div(cd a, cd b) {
  const double s = 1.0 / (b.re * b.re + b.im * b.im);
  return {(a.re * b.re + a.im * b.im) * s, (a.im * b.re - a.re * b.im) * s};
}
__device__ __forceinline__ cd c_sel(bool p, cd a, cd b) { return {p ? a.re : b.re, p ? a.im : b.im}; }
template <int S>
__device__ __forceinline__ cd c_row_shr(cd v) { return {row_shr<S>(v.re), row_shr<S>(v.im)}; }
template <int S>
__device__ __forceinline__ cd c_row_shl(cd v) { return {row_shl<S>(v.re), row_shl<S>(v.im)}; }
__device__ __forceinline__ cd c_readlane(cd v, int l) { return {readlane_d(v.re, l), readlane_d(v.im, l)}; }
__device__ __forceinline__ cd c_shfl_up(cd v, int s) { return {__shfl_up(v.re, s), __shfl_up(v.im, s)}; }
__device__ __forceinline__ cd c_row_bcast15(cd v) { return {row_bcast15(v.re), row_bcast15(v.im)}; }
__device__ __forceinline__ cd c_row_bcast31(cd v) { return {row_bcast31(v.re), row_bcast31(v.im)}; }
__device__ __forceinline__ cd c_wave_shr1(cd v) { return {wave_shr1(v.re), wave_shr1(v.im)}; }
__device__ __forceinline__ cd c_wave_shl1(cd v) { return {wave_shl1(v.re), wave_shl1(v.im)}; }
__device__ __forceinline__ cd c_bpermute(int byte_addr, cd v) {
  return {bpermute_d(byte_addr, v.re), bpermute_d(byte_addr, v.im)};
}

// ------------------------------------------------ complex double-double
struct cdd {
  dd re, im;
};
__device__ __forceinline__ cdd cdd_from(cd a) { return {dd_from(a.re), dd_from(a.im)}; }
__device__ __forceinline__ dd dd_neg(dd a) { return {-a.h, -a.l}; }
__device__ __forceinline__ cdd cdd_add(cdd a, cdd b) { return {dd_add(a.re, b.re), dd_add(a.im, b.im)}; }
__device__ __forceinline__ cdd cdd_mul(cdd a, cdd b) {
  return {dd_add(dd_mul(a.re, b.re), dd_neg(dd_mul(a.im, b.im))),
          dd_add(dd_mul(a.re, b.im), dd_mul(a.im, b.re))};
}
// -(a * c) of two complex doubles, each real product exact (two_prod)
__device__ __forceinline__ cdd cdd_neg_prod(cd a, cd c) {
  return {dd_add(two_prod(-a.re, c.re), two_prod(a.im, c.im)),
          dd_add(two_prod(-a.re, c.im), two_prod(-a.im, c.re))};
}
__device__ __forceinline__ cdd cdd_shfl_up(cdd v, int s) {
  return {shfl_up_dd(v.re, s), shfl_up_dd(v.im, s)};
}
// 2x2 complex double-double product R = X * Y rescaled by a power of two
// (the pivot is a ratio of entries)
__device__ __forceinline__ void cmat_mul_norm(const cdd X[4], const cdd Y[4], cdd R[4]) {
  const cdd r0 = cdd_add(cdd_mul(X[0], Y[0]), cdd_mul(X[1], Y[2]));
  const cdd r1 = cdd_add(cdd_mul(X[0], Y[1]), cdd_mul(X[1], Y[3]));
  const cdd r2 = cdd_add(cdd_mul(X[2], Y[0]), cdd_mul(X[3], Y[2]));
  const cdd r3 = cdd_add(cdd_mul(X[2], Y[1]), cdd_mul(X[3], Y[3]));
  double mx = fmax(fmax(fabs(r0.re.h), fabs(r0.im.h)), fmax(fabs(r1.re.h), fabs(r1.im.h)));
  mx = fmax(mx, fmax(fmax(fabs(r2.re.h), fabs(r2.im.h)), fmax(fabs(r3.re.h), fabs(r3.im.h))));
  const double sc = (mx > 0.0 && isfinite(mx)) ? ldexp(1.0, -ilogb(mx)) : 1.0;
  R[0] = {dd_scale(r0.re, sc), dd_scale(r0.im, sc)};
  R[1] = {dd_scale(r1.re, sc), dd_scale(r1.im, sc)};
  R[2] = {dd_scale(r2.re, sc), dd_scale(r2.im, sc)};
  R[3] = {dd_scale(r3.re, sc), dd_scale(r3.im, sc)};
}
// a / b rounded to complex double: q1 from the leading parts, then one
// correction with the double-double remainder a - b q1
__device__ __forceinline__ cd cdd_div_to_cd(cdd a, cdd b) {
  const cd bh = {b.re.h, b.im.h};
  const cd q1 = c_div({a.re.h, a.im.h}, bh);
  const cdd bq = cdd_mul(b, cdd_from(q1));
  const cdd r = {dd_add(a.re, dd_neg(bq.re)), dd_add(a.im, dd_neg(bq.im))};
  return c_add(q1, c_div({r.re.h, r.im.h}, bh));
}

template <int W>
struct Irk4Shared {
  cd Bw[W];          // forward-scan wave totals (per step)
  cd Cw[W];          // backward-scan wave totals (per step)
  cd Aw[W];          // forward wave transfer products (per instance)
  cd Gw[W];          // backward wave transfer products (per instance)
  cd cx[W];          // last super-diagonal of S in each wave (setup)
  cd gx[W];          // last gamma of each wave (setup)
  cd fz[W][W];       // forward fold multipliers of each wave (per instance)
  cd fx[W][W];       // backward fold multipliers of each wave (per instance)
  double Mx[W][4][4];  // wave products of the pivot matrices (setup)
  double red[W];     // norm reduction
  unsigned bad[W];
};

// Rows of the mass matrix A and of D' = B + C for unknown u (node u+1), from
// its two adjacent elements; the same expressions and boundary handling as
// march.hip's assemble_row (no FMA contraction).  Padding rows (u >= n) are
// zero in both, and the couplings to Dirichlet nodes or past the last
// unknown are dropped.
struct SplitRow {
  double al, ad, ar;  // A
  double dl, dd, dr;  // D'
};
__device__ __forceinline__ SplitRow assemble_split(const double* __restrict__ x,
                                                   const double* __restrict__ w, int n, int M,
                                                   int u, int cmat, double wall0, double wallL) {
#pragma clang fp contract(off)
  const bool live = u < n;
  const int i = live ? u + 1 : 1;
  const double xl = x[i - 1], xc = x[i];
  const bool has_r = (i + 1 < M);
  const double xr = has_r ? x[i + 1] : xc;
  const double hl = xc - xl, hr = xr - xc;
  const double wi = live ? w[live ? u : 0] : 0.0;
  const double Al = hl / 6.0, Ar = hr / 6.0, Ad = hl / 3.0 + hr / 3.0;
  const double Bl = -1.0 / hl;
  const double Br = has_r ? -1.0 / hr : 0.0;
  const double Bd = 1.0 / hl + (has_r ? 1.0 / hr : 0.0);
  double Cl, Cd, Cr;
  if (cmat == 0) {
    Cl = wi * Al;
    Cd = wi * Ad;
    Cr = wi * Ar;
  } else {
    const double wl = (u > 0 && live) ? w[u - 1] : wall0;
    const double wr = (u + 1 < n) ? w[u + 1] : wallL;
    Cl = hl * (wl + wi) / 12.0;
    Cd = hl * (wl + 3.0 * wi) / 12.0 + hr * (3.0 * wi + wr) / 12.0;
    Cr = hr * (wi + wr) / 12.0;
  }
  const bool has_l = live && u > 0, has_rr = live && u < n - 1;
  SplitRow s;
  s.al = has_l ? Al : 0.0;
  s.ad = live ? Ad : 0.0;
  s.ar = has_rr ? Ar : 0.0;
  s.dl = has_l ? Bl + Cl : 0.0;
  s.dd = live ? Bd + Cd : 0.0;
  s.dr = has_rr ? Br + Cr : 0.0;
  return s;
}

template <int K, int W>
__device__ void irk4_one(const MarchArgs& a, const int inst, const unsigned st0,
                         double* __restrict__ slab, Irk4Shared<W>& sh,
                         double* __restrict__ pw_lds, const int tid) {
  constexpr int P = 64 * W;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const InstDesc d = a.desc[inst];
  const int n = __builtin_amdgcn_readfirstlane(d.n);
  const int N = __builtin_amdgcn_readfirstlane(d.N);
  const int M = __builtin_amdgcn_readfirstlane(d.M);
  const double dt = 1.0 / N;
  const double* __restrict__ x = a.x + d.xo;
  const double* __restrict__ w = a.w + d.uo;
  const double* __restrict__ pw = a.pw + d.po;
  const int cmat = __builtin_amdgcn_readfirstlane(d.cmat);
  const int u0 = tid * K;
  // lambda = 1/4 + i sqrt3/12, -dt omega = -dt (1 - i sqrt3)
  const double lam_re = 0.25, lam_im = 0.14433756729740644;  // sqrt(3)/12
  (void)lam_re;
  const cd ndo = {-dt, dt * 1.7320508075688772};

  wg_barrier();
  for (int i = tid; i <= (N >> 1); i += P) pw_lds[i] = pw[i];
  wg_barrier();

  // Rows of S = A + dt lambda D' (padding rows: identity), re-assembled in
  // each setup phase rather than held across the double-double scan.
  auto srow = [&](const int t, cd& sa, cd& sb, cd& sc, double& dl_, double& dm_,
                  double& dr_) __attribute__((always_inline)) {
    const SplitRow s = assemble_split(x, w, n, M, u0 + t, cmat, d.wall0, d.wallL);
    const bool live = u0 + t < n;
    sa = {s.al + dt * lam_re * s.dl, dt * lam_im * s.dl};
    sb = {live ? s.ad + dt * lam_re * s.dd : 1.0, dt * lam_im * s.dd};
    sc = {s.ar + dt * lam_re * s.dr, dt * lam_im * s.dr};
    dl_ = s.dl;
    dm_ = s.dd;
    dr_ = s.dr;
  };

  // ------------------------------------------------------------ factor scan
  cd cprev;
  {
    cd sa, sb, sc;
    double t0, t1, t2;
    srow(K - 1, sa, sb, sc, t0, t1, t2);
    cprev = c_shfl_up(sc, 1);
    if constexpr (W > 1) {
      sh.cx[wave] = c_readlane(sc, 63);
      wg_barrier();
      const cd cx_prev = sh.cx[wave > 0 ? wave - 1 : 0];
      cprev = c_sel(lane == 0, c_sel(wave > 0, cx_prev, cd{0.0, 0.0}), cprev);
    } else {
      cprev = c_sel(lane == 0, cd{0.0, 0.0}, cprev);
    }
  }
  const cdd one = {dd_from(1.0), dd_from(0.0)}, zero = {dd_from(0.0), dd_from(0.0)};
  cdd T[4] = {one, zero, zero, one};
  {
    cd cp = cprev;
#pragma unroll
    for (int t = 0; t < K; t++) {
      cd sa, sb, sc;
      double t0, t1, t2;
      srow(t, sa, sb, sc, t0, t1, t2);
      const cdd Mt[4] = {cdd_from(sb), cdd_neg_prod(sa, cp), one, zero};
      cmat_mul_norm(Mt, T, T);
      cp = sc;
    }
  }
#pragma unroll
  for (int dd_ = 0; dd_ < 6; dd_++) {
    const int s = 1 << dd_;
    cdd Tp[4];
#pragma unroll
    for (int e = 0; e < 4; e++) Tp[e] = cdd_shfl_up(T[e], s);
    if (lane >= s) cmat_mul_norm(T, Tp, T);
  }
  cdd E[4];
#pragma unroll
  for (int e = 0; e < 4; e++) E[e] = cdd_shfl_up(T[e], 1);
  cdd Pre[4] = {one, zero, zero, one};
  if constexpr (W > 1) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      sh.Mx[wave][e][0] = readlane_d(T[e].re.h, 63);
      sh.Mx[wave][e][1] = readlane_d(T[e].re.l, 63);
      sh.Mx[wave][e][2] = readlane_d(T[e].im.h, 63);
      sh.Mx[wave][e][3] = readlane_d(T[e].im.l, 63);
    }
    wg_barrier();
    for (int v = 0; v < wave; v++) {
      cdd Mv[4];
#pragma unroll
      for (int e = 0; e < 4; e++)
        Mv[e] = {{sh.Mx[v][e][0], sh.Mx[v][e][1]}, {sh.Mx[v][e][2], sh.Mx[v][e][3]}};
      cmat_mul_norm(Mv, Pre, Pre);
    }
  }
  if (lane > 0) {
    cmat_mul_norm(E, Pre, E);
  } else {
#pragma unroll
    for (int e = 0; e < 4; e++) E[e] = Pre[e];
  }
  cd up = (tid > 0) ? cdd_div_to_cd(E[0], E[2]) : cd{1.0, 0.0};

  unsigned bad = 0;
  double dl[K], dm[K], dr[K];
  cd al[K], ga[K], rho[K];
  {
    cd cp = cprev;
#pragma unroll
    for (int t = 0; t < K; t++) {
      cd sa, sb, sc;
      srow(t, sa, sb, sc, dl[t], dm[t], dr[t]);
      const bool has_a = (sa.re != 0.0) || (sa.im != 0.0);
      const cd l = has_a ? c_div(sa, up) : cd{0.0, 0.0};
      const cd uu = c_fnma(l, cp, sb);
      const double m2 = uu.re * uu.re + uu.im * uu.im;
      if (!(m2 > 0.0) || !isfinite(m2)) bad |= ST_PIVOT;
      const cd inv = c_div({1.0, 0.0}, uu);
      al[t] = c_mul(sa, inv);
      ga[t] = c_mul(sc, inv);
      rho[t] = c_mul(ndo, inv);
      cp = sc;
      up = uu;
    }
  }

  // -------------------------------------------- scan constants of the march
  const int rl = lane & 15;
  const int row = lane >> 4;
  const cd c0 = {0.0, 0.0}, c1v = {1.0, 0.0};
  cd Af = c1v, Gb = c1v;
#pragma unroll
  for (int t = 0; t < K; t++) {
    Af = c_mul(Af, {-al[t].re, -al[t].im});
    Gb = c_mul(Gb, {-ga[t].re, -ga[t].im});
  }
  cd mf[4], mb[4];
  {
    cd Ap, Gp;
    Ap = c_row_shr<1>(Af); Gp = c_row_shl<1>(Gb);
    mf[0] = c_sel(rl >= 1, Af, c0); mb[0] = c_sel(rl + 1 < 16, Gb, c0);
    Af = c_sel(rl >= 1, c_mul(Af, Ap), Af); Gb = c_sel(rl + 1 < 16, c_mul(Gb, Gp), Gb);
    Ap = c_row_shr<2>(Af); Gp = c_row_shl<2>(Gb);
    mf[1] = c_sel(rl >= 2, Af, c0); mb[1] = c_sel(rl + 2 < 16, Gb, c0);
    Af = c_sel(rl >= 2, c_mul(Af, Ap), Af); Gb = c_sel(rl + 2 < 16, c_mul(Gb, Gp), Gb);
    Ap = c_row_shr<4>(Af); Gp = c_row_shl<4>(Gb);
    mf[2] = c_sel(rl >= 4, Af, c0); mb[2] = c_sel(rl + 4 < 16, Gb, c0);
    Af = c_sel(rl >= 4, c_mul(Af, Ap), Af); Gb = c_sel(rl + 4 < 16, c_mul(Gb, Gp), Gb);
    Ap = c_row_shr<8>(Af); Gp = c_row_shl<8>(Gb);
    mf[3] = c_sel(rl >= 8, Af, c0); mb[3] = c_sel(rl + 8 < 16, Gb, c0);
    Af = c_sel(rl >= 8, c_mul(Af, Ap), Af); Gb = c_sel(rl + 8 < 16, c_mul(Gb, Gp), Gb);
  }
  const cd Arow0 = c_readlane(Af, 15), Arow1 = c_readlane(Af, 31);
  const cd Arow2 = c_readlane(Af, 47), Arow3 = c_readlane(Af, 63);
  const cd Grow0 = c_readlane(Gb, 0), Grow1 = c_readlane(Gb, 16);
  const cd Grow2 = c_readlane(Gb, 32), Grow3 = c_readlane(Gb, 48);
  // DPP reads with every lane active (a DPP move under a partial EXEC mask
  // reads 0 from the disabled lanes)
  const cd Af_sh = c_row_shr<1>(Af), Gb_sh = c_row_shl<1>(Gb);
  const cd Ashift = c_sel(rl > 0, Af_sh, c0);
  const cd Gshift = c_sel(rl < 15, Gb_sh, c0);
  // selects on values throughout: a ?: between struct lvalues becomes a
  // select of addresses, which puts its operands in scratch memory
  const cd A01 = c_mul(Arow0, Arow1), G23 = c_mul(Grow2, Grow3);
  const cd Abefore = c_sel(row == 0, c1v, c_sel(row == 1, Arow0,
                            c_sel(row == 2, A01, c_mul(A01, Arow2))));
  const cd Gafter = c_sel(row == 3, c1v, c_sel(row == 2, Grow3,
                           c_sel(row == 1, G23, c_mul(Grow1, G23))));
  const cd Apre_m1 = c_mul(c_sel(rl > 0, Ashift, c1v), Abefore);
  const cd Gsuf_p1 = c_mul(c_sel(rl < 15, Gshift, c1v), Gafter);
  // multipliers of the cross-row levels (march.hip): forward through the
  // broadcast moves, backward through ds_bpermute
  const cd mf5 = Af;
  const cd mf6 = c_sel(row == 3, c_mul(Arow2, Af), Af);
  const cd mbA = c_sel(row == 0 || row == 2, Gb, c0);
  const cd mbB = c_sel(row == 0, c_mul(Gb, Grow1), c_sel(row == 1, Gb, c0));
  const int srcA = ((row + 1) & 3) * 64;
  const int srcB = 32 * 4;
  cd gl = c_shfl_up(ga[K - 1], 1);
  if constexpr (W > 1) {
    sh.Aw[wave] = c_mul(A01, c_mul(Arow2, Arow3));
    sh.gx[wave] = c_readlane(ga[K - 1], 63);
    sh.Gw[wave] = c_mul(c_mul(Grow0, Grow1), G23);
    wg_barrier();
    const cd gx_prev = sh.gx[wave > 0 ? wave - 1 : 0];
    gl = c_sel(lane == 0, c_sel(wave > 0, gx_prev, c0), gl);
    // fold multipliers (as in march.hip; complex, so kept in LDS rather than
    // scalar registers): Z = sum_{v < wave} fz[v] Bw[v], X = sum_{v > wave} fx[v] Cw[v]
    cd p = c1v;
    for (int v = W - 1; v >= 0; v--) {
      sh.fz[wave][v] = c_sel(v < wave, p, c0);
      p = c_sel(v < wave, c_mul(p, sh.Aw[v]), p);
    }
    p = c1v;
    for (int v = 0; v < W; v++) {
      sh.fx[wave][v] = c_sel(v > wave, p, c0);
      p = c_sel(v > wave, c_mul(p, sh.Gw[v]), p);
    }
  } else {
    gl = c_sel(lane == 0, c0, gl);
  }

  // ------------------------------------------------------------------ march
  double q[K], acc[K];
#pragma unroll
  for (int t = 0; t < K; t++) {
    q[t] = (u0 + t < n) ? 1.0 : 0.0;
    acc[t] = 0.0;
  }
  // q of the unknowns left of q[0] and right of q[K-1], advanced with the
  // same x values their owners add (recomputed here from the scan carries)
  double ql = (u0 - 1 >= 0 && u0 - 1 < n) ? 1.0 : 0.0;
  double qr = (u0 + K < n) ? 1.0 : 0.0;
  const int half = N >> 1;

  auto advance = [&]() __attribute__((always_inline)) {
    cd rp[K];
#pragma unroll
    for (int t = 0; t < K; t++) {
      const double qm = (t == 0) ? ql : q[t - 1];
      const double qp = (t == K - 1) ? qr : q[t + 1];
      const double dq = fma(dl[t], qm, fma(dm[t], q[t], dr[t] * qp));
      rp[t] = {rho[t].re * dq, rho[t].im * dq};
    }
    // forward z_i = r'_i - alpha_i z_{i-1}
    cd B = c0;
#pragma unroll
    for (int t = 0; t < K; t++) B = c_fnma(al[t], B, rp[t]);
    B = c_fma(mf[0], c_row_shr<1>(B), B);
    B = c_fma(mf[1], c_row_shr<2>(B), B);
    B = c_fma(mf[2], c_row_shr<4>(B), B);
    B = c_fma(mf[3], c_row_shr<8>(B), B);
    B = c_fma(mf5, c_row_bcast15(B), B);
    B = c_fma(mf6, c_row_bcast31(B), B);
    cd zin = c_wave_shr1(B);
    if constexpr (W > 1) {
      sh.Bw[wave] = c_readlane(B, 63);
      wg_barrier();
      cd Z = c0;
#pragma unroll
      for (int v = 0; v + 1 < W; v++) Z = c_fma(sh.fz[wave][v], sh.Bw[v], Z);
      zin = c_fma(Apre_m1, Z, zin);
    }
    {
      cd z = zin;
#pragma unroll
      for (int t = 0; t < K; t++) {
        z = c_fnma(al[t], z, rp[t]);
        rp[t] = z;
      }
    }
    // backward x_i = z_i - gamma_i x_{i+1}
    cd Cc = c0;
#pragma unroll
    for (int t = K - 1; t >= 0; t--) Cc = c_fnma(ga[t], Cc, rp[t]);
    Cc = c_fma(mb[0], c_row_shl<1>(Cc), Cc);
    Cc = c_fma(mb[1], c_row_shl<2>(Cc), Cc);
    Cc = c_fma(mb[2], c_row_shl<4>(Cc), Cc);
    Cc = c_fma(mb[3], c_row_shl<8>(Cc), Cc);
    Cc = c_fma(mbA, c_bpermute(srcA, Cc), Cc);
    Cc = c_fma(mbB, c_bpermute(srcB, Cc), Cc);
    cd xin = c_wave_shl1(Cc);
    if constexpr (W > 1) {
      sh.Cw[wave] = c_readlane(Cc, 0);
      wg_barrier();
      cd X = c0;
#pragma unroll
      for (int v = 1; v < W; v++) X = c_fma(sh.fx[wave][v], sh.Cw[v], X);
      xin = c_fma(Gsuf_p1, X, xin);
    }
    cd xx = xin;
#pragma unroll
    for (int t = K - 1; t >= 0; t--) {
      xx = c_fnma(ga[t], xx, rp[t]);
      q[t] += xx.re;
    }
    ql += c_fnma(gl, xx, zin).re;
    qr += xin.re;
  };

  // slabs, pairing and residual: the layout and order of march.hip
  const bool lane_live = u0 < n;
  const __amdgpu_buffer_rsrc_t srsrc =
      slab_rsrc(slab, static_cast<unsigned>(a.slab_stride * sizeof(double)));
  const unsigned off_dst = static_cast<unsigned>(tid) * slab_lane_bytes<K>();
  const unsigned off_src = lane_live ? off_dst : 0u;
  auto store_slab = [&](const int j) __attribute__((always_inline)) {
    if (lane_live) slab_put<K>(q, srsrc, off_dst, static_cast<unsigned>(j * (P * K) * 8), P);
  };
  store_slab(0);
  for (int j = 1; j < half; j++) {
    advance();
    store_slab(j);
  }
  advance();
  {
    const double cw0 = pw_lds[0];
#pragma unroll
    for (int t = 0; t < K; t++) acc[t] = cw0 * q[t] * q[t];
  }
  constexpr int PD = (K <= 2) ? 8 : (K == 4) ? 4 : 2;
  double ring[PD][K];
  auto fetch = [&](const int s, double (&dstv)[K]) __attribute__((always_inline)) {
    slab_get<K>(dstv, srsrc, off_src, static_cast<unsigned>(max(s, 0) * (P * K) * 8), P);
  };
  // The loop below waits on each ring set by the count of loads issued after
  // it.  That count is exact only if nothing else is in flight (the first
  // half's slab stores are drained here) and the sets are issued in ring
  // order (the scheduler may not reorder them across the barrier).
  vmem_drain();
#pragma unroll
  for (int e = 0; e < PD; e++) {
    fetch(half - 1 - e, ring[e]);
    __builtin_amdgcn_sched_barrier(0);
  }
  for (int i0 = 1; i0 <= half; i0 += PD) {
#pragma unroll
    for (int e = 0; e < PD; e++) {
      const int i = i0 + e;
      advance();
      const double cj = pw_lds[i];
#pragma unroll
      for (int t = 0; t < K; t++) acc[t] = fma(cj * q[t], ring[e][t], acc[t]);
      fetch(half - i - PD, ring[e]);
    }
  }

  const double* __restrict__ pt = a.phit + d.uo;
  double* __restrict__ r = a.r + d.uo;
  double* __restrict__ phi = a.phi ? a.phi + d.uo : nullptr;
  double mx = 0.0;
#pragma unroll
  for (int t = 0; t < K; t++) {
    const int u = u0 + t;
    if (u < n) {
      const double res = acc[t] - pt[u];
      r[u] = res;
      if (phi) phi[u] = acc[t];
      if (!isfinite(res)) bad |= ST_NAN;
      mx = fmax(mx, fabs(res));
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, off));
    bad |= __shfl_xor(bad, off);
  }
  mx = readlane_d(mx, 0);
  bad = __builtin_amdgcn_readfirstlane(bad);
  if constexpr (W > 1) {
    sh.red[wave] = mx;
    sh.bad[wave] = bad;
    wg_barrier();
    for (int v = 0; v < W; v++) {
      mx = fmax(mx, sh.red[v]);
      bad |= sh.bad[v];
    }
    wg_barrier();
  }
  if (wave == 0) {
    a.rnorm[inst] = (bad & ST_NAN) ? __builtin_nan("") : mx;
    a.status[inst] = st0 | bad;
  }
}

// The work-queue loop of march.hip's march_kernel.
template <int K, int W>
__global__ void __launch_bounds__(64 * W) irk4_kernel(MarchArgs a) {
  __shared__ Irk4Shared<W> sh;
  __shared__ int s_next;
  extern __shared__ double pw_lds[];
  const int tid = threadIdx.x;
  double* slab = a.slabs + (long long)blockIdx.x * a.slab_stride;
  for (;;) {
    if (tid == 0) s_next = atomicAdd(a.queue, 1);
    wg_barrier();
    const int it = __builtin_amdgcn_readfirstlane(s_next);
    wg_barrier();
    if (it >= a.n_items) break;
    const int inst = __builtin_amdgcn_readfirstlane(a.items[it]);
    const unsigned st = __builtin_amdgcn_readfirstlane(a.status[inst]);
    if (st & ST_DONE_MASK) continue;
    irk4_one<K, W>(a, inst, st, slab, sh, pw_lds, tid);
  }
}

// ------------------------------------------------------------------ dispatch
// (K, W) variants.  About 11 doubles per row stay in registers for the
// whole march, so K = 8 needs the AGPR half of a 1-wave-per-SIMD workgroup
// (W <= 4; 20 bytes per lane still spill), and every variant up to 2048
// unknowns (64 * 8 * 4) keeps its rows in registers.  The 4096-unknown
// bucket has only (8, 8): two waves per SIMD leave 256 registers per wave,
// and ~800 bytes per lane of row state spill to scratch memory, re-read
// every step.  It is correct and slower per node-step than the register
// variants; it exists so that the stepper takes every mesh the implicit
// Euler march takes.
#define SCFT_IRK4_VARIANTS(X) \
  X(1, 1) X(2, 1) X(4, 1) X(8, 1)   \
  X(1, 2) X(2, 2) X(4, 2) X(8, 2)   \
  X(1, 4) X(2, 4) X(4, 4) X(8, 4)   \
  X(1, 8) X(2, 8) X(8, 8)

static size_t irk4_lds_bytes(int half_max) {
  return static_cast<size_t>(half_max + 1) * sizeof(double);
}

static const void* irk4_fn(int K, int W) {
#define X(k, w) \
  if (K == k && W == w) return reinterpret_cast<const void*>(&irk4_kernel<k, w>);
  SCFT_IRK4_VARIANTS(X)
#undef X
  return nullptr;
}

bool irk4_variant_exists(int K, int W) { return irk4_fn(K, W) != nullptr; }

hipError_t launch_irk4(int K, int W, const MarchArgs& a, int grid, hipStream_t s) {
#define X(k, w)                                                                 \
  if (K == k && W == w) {                                                       \
    hipLaunchKernelGGL((irk4_kernel<k, w>), dim3(grid), dim3(64 * w),             \
                       irk4_lds_bytes(a.half_max), s, a);                       \
    return hipGetLastError();                                                   \
  }
  SCFT_IRK4_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

int irk4_max_grid(int K, int W, int half_max) {
  const void* f = irk4_fn(K, W);
  if (!f) return 0;
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, 64 * W, irk4_lds_bytes(half_max)) !=
      hipSuccess)
    return 0;
  if (per < 1) return 0;
  return per * cus;
}

}  // namespace scft
