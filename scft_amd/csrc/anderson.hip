// anderson.hip -- device-resident safeguarded Anderson mixing, one workgroup
// per instance, one launch per sweep for the whole batch.
//
// Restates the update of adm.c:232-306 in residual form: with r_k the mixing
// residual of w_k and r_{k-1..k-m} the stored ones,
//   theta = argmin || r_k - sum_a theta_a (r_k - r_{k-a}) ||_2
// from the Gram system U theta = b (adm.c:232-276) solved by Gauss-Jordan
// with full pivoting (gaussj.c:7-60, singular -> history restart as
// ADM_chen_C.c:103-112), then
//   w_{k+1} = w_k - sum theta_a (w_k - w_{k-a}) + beta (r_k - sum theta_a (r_k - r_{k-a}))
// (adm.c:289-306 with mixing beta).  Safeguards: the step is scaled to
// ||w_{k+1} - w_k||_inf <= step_cap, the history is cleared when ||r_k||
// exceeds restart_factor times the best so far, and a warm start whose first
// residual is fallback_factor times the batch median is replaced by the cold
// start.  Convergence is judged on the plain residual phi - phi_t.
//
// Extensions (off by default, not in the reference):
//  * preconditioned mixing residual g = f + a M_L^{-1} K f, with M_L^{-1} K
//    the lumped-mass linear-element Laplacian of the unknowns' grid with
//    natural ends.  phi responds to a field perturbation of wavenumber k like
//    the Debye function, ~2/k^2 for k >> 1 (contour length 1, q_s = q_xx),
//    so I + K/2 undoes the response at short wavelengths;
//  * backtracking: a failed evaluation (NaN / zero pivot) or one whose ||r||
//    exceeds backtrack times the last accepted one halves the step toward the
//    last accepted iterate (ring slot head - 1) and clears the history;
//  * the ground-state form of the preconditioner (below, at the push of the
//    history): the Laplacian conjugated by d = sqrt(phi_t), the propagator's
//    profile next to the wall, with the potential that makes constants its
//    null space, -(1/d^2)(d^2 f')'.  In the log transform the Jacobian of f
//    near the solution is ~ -d^-1 L^-1 d with L = -d_xx + w, so the form
//    inverts it where the plain Laplacian does not: the wall layer, where
//    phi_t falls to zero.  The preconditioned Jacobian of production-mix
//    instances has every eigenvalue within 0.14 of 1 with it (the plain
//    form: ~half of them beyond 0.2);
//  * a lower bound w >= -field_floor / dt on the updated field.  Near a wall
//    on a fine mesh the converged field spans decades (w at the first node
//    1e6-1e8 when dt/h^2 ~ 1e4-1e5), and the preconditioned step from a smooth
//    start overshoots to large negative values at the next nodes, where
//    1 + dt w < 0 makes the implicit-Euler matrix indefinite and the density
//    explode; with 1 + dt w >= 1 - field_floor every row stays strictly
//    diagonally dominant.
#include "scft_internal.h"

namespace scft {

constexpr int AM_THREADS = 256;
constexpr int AM_MAXH = 16;  // previous iterates supported (history <= 16)

// lane l's value of a double, l wave-uniform
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane(static_cast<unsigned>(b), l);
  const unsigned hi = __builtin_amdgcn_readlane(static_cast<unsigned>(b >> 32), l);
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

__device__ __forceinline__ double mix_residual(double r, double pt, int transform,
                                               double floor) {
  if (transform == 0) return r;
  const double p = fmax(pt, floor);
  if (transform == 1) return r / p;
  return log(fmax(p + r, floor * 1e-3) / p);
}

__global__ void __launch_bounds__(AM_THREADS) anderson_kernel(AndersonArgs a) {
  // the largest instances are dispatched first, so the small ones fill the
  // CUs behind them instead of each round waiting on a large one
  const int inst = a.order ? a.order[blockIdx.x] : static_cast<int>(blockIdx.x);
  const int tid = threadIdx.x;
  __shared__ double U[AM_MAXH * AM_MAXH];
  __shared__ double bvec[AM_MAXH];
  __shared__ double red[AM_THREADS / 64][16 * 16 + 16];
  __shared__ double red1[AM_THREADS / 64];
  __shared__ int s_m, s_mode;
  __shared__ double s_scale, s_tr;
  __shared__ double s_pv[AM_THREADS / 64];  // per-wave pivot candidates
  __shared__ int s_pi[AM_THREADS / 64];

  const unsigned st0 = __builtin_amdgcn_readfirstlane(a.status[inst]);
  if (st0 & (ST_CONVERGED | ST_MAXITER | ST_FAILED)) return;
  const InstDesc d = a.desc[inst];
  const int n = __builtin_amdgcn_readfirstlane(d.n);
  const int S = a.history + 1;
  double* __restrict__ w = a.w + d.uo;
  const double* __restrict__ r = a.r + d.uo;
  const double* __restrict__ pt = a.phit + d.uo;
  double* __restrict__ hw = a.hw + d.ho;
  double* __restrict__ hr = a.hr + d.ho;
  MixState ms = a.st[inst];
  const double err = a.rnorm[inst];
  const int k = ms.k + 1;  // evaluations including this one

  // ------------------------------------------------------------ stop tests
  if (tid == 0) {
    s_mode = 0;
    const bool bad_eval = !(err == err) || (st0 & (ST_NAN | ST_PIVOT));
    const bool have_acc = a.backtrack > 0.0 && ms.eacc < INFINITY;
    const bool can_halve = have_acc && ms.halvings < a.max_halvings;
    if (bad_eval) {
      s_mode = can_halve ? 5 : 1;  // NaN / pivot flagged by the march
    } else if (have_acc && err > a.backtrack * ms.eacc) {
      s_mode = can_halve ? 5 : 1;
    } else if (err <= a.tol) {
      s_mode = 2;
    } else if (ms.k == 0 && a.fallback_thresh > 0.0 && err > a.fallback_thresh &&
               a.wcold != nullptr) {
      s_mode = 3;  // reject the warm start
    } else if (k >= a.max_iter) {
      s_mode = 4;
    }
  }
  __syncthreads();
  const int mode = __builtin_amdgcn_readfirstlane(s_mode);
  if (mode == 1 || mode == 2 || mode == 4) {
    if (tid == 0) {
      if (mode == 1) a.status[inst] |= ST_FAILED;
      if (mode == 2) a.status[inst] |= ST_CONVERGED;
      if (mode == 4) a.status[inst] |= ST_MAXITER;
      a.iters[inst] = k;
      ms.k = k;
      a.st[inst] = ms;
    }
    return;
  }
  if (mode == 3) {  // restart from the cold-start field, history empty
    const double* wc = a.wcold + d.uo;
    for (int u = tid; u < n; u += AM_THREADS) w[u] = wc[u];
    if (tid == 0) {
      a.status[inst] |= ST_FALLBACK;
      ms.k = k;  // the rejected evaluation counts
      ms.stored = 0;
      ms.head = 0;
      ms.best = INFINITY;
      ms.eacc = INFINITY;
      ms.halvings = 0;
      a.st[inst] = ms;
      atomicAdd(a.active, 1);
    }
    return;
  }
  if (mode == 5) {  // halve the last step toward the last accepted iterate
    const int sa = __builtin_amdgcn_readfirstlane((ms.head - 1 + S) % S);
    for (int u = tid; u < n; u += AM_THREADS) {
      const double wa = hw[(long long)sa * n + u];
      w[u] = wa + 0.5 * (w[u] - wa);
    }
    if (tid == 0) {
      a.status[inst] = st0 & ~(ST_NAN | ST_PIVOT);
      ms.k = k;
      ms.halvings++;
      ms.restarts++;
      ms.stored = 0;
      a.st[inst] = ms;
      atomicAdd(a.active, 1);
    }
    return;
  }
  ms.halvings = 0;
  ms.eacc = err;

  // ----------------------------------------------- history bookkeeping
  int stored = ms.stored;
  if (err > a.restart_factor * ms.best && stored > 0) {
    stored = 0;
    ms.restarts++;
  }
  // the ground-state form of the preconditioner (below) until the instance
  // leaves it for the plain form (history cleared, status PLAIN_FORM): after
  // gs_switch evaluations, or when its error grows past gs_switch_growth
  // times the best so far
  bool gs = a.gs_precond > 0.0 && !(st0 & ST_PLAIN_FORM);
  unsigned st_new = st0;
  if (gs && ((a.gs_switch > 0 && k > a.gs_switch) ||
             (a.gs_switch_growth > 0.0 && err > a.gs_switch_growth * ms.best))) {
    gs = false;
    stored = 0;
    ms.restarts++;
    st_new |= ST_PLAIN_FORM;
  }
  // the rational form (below) from the start, or, with gs_shift_below, once
  // the error first falls below that threshold (history cleared then)
  bool rat_ok = a.gs_shift_below <= 0.0 || (st0 & ST_RATIONAL);
  if (gs && a.gs_shift > 0.0 && !rat_ok && err < a.gs_shift_below) {
    rat_ok = true;
    stored = 0;
    ms.restarts++;
    st_new |= ST_RATIONAL;
  }
  if (tid == 0 && st_new != st0) a.status[inst] = st_new;
  if (err < ms.best) ms.best = err;
  const int cur = __builtin_amdgcn_readfirstlane(ms.head);
  int m = __builtin_amdgcn_readfirstlane(stored < a.history ? stored : a.history);

  // push (w_k, g_k) into slot `cur`
  const double* __restrict__ xn = a.x + d.xo;
  // the preconditioned mixing residual g: the plain form
  //   g = f + a M_L^-1 K f   (lumped-mass Laplacian of the unknowns' grid,
  //                           natural rows at both ends)
  // or the ground-state form
  //   g_u = f_u + a_g [ l_u (d_{u-1}/d_u)(f_u - f_{u-1})
  //                   + r_u (d_{u+1}/d_u)(f_u - f_{u+1}) + c max(w_u, 0) f_u ]
  // with d = sqrt(max(phi_t, floor)), zero beyond the unknowns (the wall
  // nodes), l_u = (1/h_l)/m_u, r_u = (1/h_r)/m_u, m_u = (h_l + h_r)/2 over
  // the mesh's nodes (a natural end has no right neighbour): the operator
  // -(1/d^2)(d^2 f')' + c w+, which maps constants to themselves.  With
  // gs_shift = g > 0 the rational form
  //   g = (1 - a_g) f + a_g G f + a_g g (g I + G0)^-1 f,  G0 = G with c = 0,
  // whose symbol (1 - a) + a x + a g / (g + x) follows 1/g_D(x), the
  // inverse Debye function of the chain's linear response, to ~2% for
  // a = 1/2, g = 3 (the linear form: to ~13%).  The Thomas factors of
  // g I + G0 depend on the mesh and phi_t only (gs_pivots_kernel, once per
  // batch): per sweep the solve is its two first-order recurrences, run as
  // scans below.
  const double coef = gs ? a.gs_precond : a.precond;
  const bool pc = coef > 0.0 && (gs || n > 1);
  const bool rational = pc && gs && a.gs_shift > 0.0 && rat_ok;
  const double* __restrict__ t_ri = rational ? a.tri + d.uo : nullptr;  // 1 / pivot
  const double* __restrict__ t_al = rational ? t_ri + a.tri_stride : nullptr;  // lower / pivot
  const double* __restrict__ t_mu = rational ? t_al + a.tri_stride : nullptr;  // upper / pivot
  double* __restrict__ t_rh = rational ? a.tri + 3 * a.tri_stride + d.uo : nullptr;
  // (whole-wave trips: the neighbours' values come from the lanes beside,
  // and only a wave's edge lanes evaluate theirs again)
  for (int u0 = 0; u0 < n; u0 += AM_THREADS) {
    const int u = u0 + tid, lane = tid & 63;
    const bool act = u < n;
    double f = act ? mix_residual(r[u], pt[u], a.transform, a.floor) : 0.0;
    if (pc && gs) {
      const double dv = act ? sqrt(fmax(pt[u], a.floor)) : 1.0;
      double fl = __shfl_up(f, 1), fr = __shfl_down(f, 1);
      double dl = __shfl_up(dv, 1), dr = __shfl_down(dv, 1);
      if (act) {
        const int i = u + 1;  // node of unknown u; node i - 1 always exists
        const bool has_l = u > 0, has_r = u + 1 < n;
        if (lane == 0 && has_l) {
          fl = mix_residual(r[u - 1], pt[u - 1], a.transform, a.floor);
          dl = sqrt(fmax(pt[u - 1], a.floor));
        }
        if (lane == 63 && has_r) {
          fr = mix_residual(r[u + 1], pt[u + 1], a.transform, a.floor);
          dr = sqrt(fmax(pt[u + 1], a.floor));
        }
        if (!has_l) fl = dl = 0.0;
        if (!has_r) fr = dr = 0.0;
        const bool node_r = i + 1 < d.M;
        const double hl = xn[i] - xn[i - 1];
        const double hrr = node_r ? xn[i + 1] - xn[i] : 0.0;
        const double mm = 0.5 * (hl + hrr);
        const double cl = (1.0 / hl) / mm;
        const double cr = node_r ? (1.0 / hrr) / mm : 0.0;
        const double lo = cl * (dl / dv), up = cr * (dr / dv);
        const double fw = a.gs_field_weight * fmax(w[u], 0.0);
        double t = lo * (f - fl) + up * (f - fr);
        t += fw * f;
        if (rational) {
          t_rh[u] = f;
          f = (1.0 - coef) * f + coef * t;
        } else {
          f = f + coef * t;
        }
      }
    } else if (pc) {
      double fl = __shfl_up(f, 1), fr = __shfl_down(f, 1);
      if (act) {
        const int i = u + 1;  // node of unknown u
        const bool has_l = u > 0, has_r = u + 1 < n;
        if (lane == 0 && has_l) fl = mix_residual(r[u - 1], pt[u - 1], a.transform, a.floor);
        if (lane == 63 && has_r) fr = mix_residual(r[u + 1], pt[u + 1], a.transform, a.floor);
        fl = has_l ? fl : 0.0;
        fr = has_r ? fr : 0.0;
        const double hl = has_l ? xn[i] - xn[i - 1] : 0.0;
        const double hrr = has_r ? xn[i + 1] - xn[i] : 0.0;
        const double mm = 0.5 * (hl + hrr);
        const double cl = has_l ? (1.0 / hl) / mm : 0.0;
        const double cr = has_r ? (1.0 / hrr) / mm : 0.0;
        f = f + coef * ((cl + cr) * f - cl * fl - cr * fr);
      }
    }
    if (act) {
      hw[(long long)cur * n + u] = w[u];
      hr[(long long)cur * n + u] = f;
    }
  }
  __syncthreads();

  if (rational) {
    // (g I + G0) y = f with the stored factors: z_i = ri_i f_i + al_i z_{i-1},
    // then y_i = z_i + mu_i y_{i+1}.  Both are affine recurrences with
    // coefficients in [0, 1) (diagonal dominance), scanned in parallel: each
    // thread composes the maps of a contiguous chunk of rows, the chunk maps
    // are scanned across the wave (shuffles) and the waves (LDS), and each
    // thread replays its chunk from its incoming value.
    __shared__ double s_mapA[AM_THREADS / 64], s_mapB[AM_THREADS / 64];
    const int lane = tid & 63, wave = tid >> 6;
    const int C = (n + AM_THREADS - 1) / AM_THREADS;
    const int r0 = min(n, tid * C), r1 = min(n, r0 + C);
    // forward: chunk map z_out = A z_in + B
    double A = 1.0, B = 0.0;
    for (int i = r0; i < r1; i++) {
      const double al = t_al[i];
      B = t_ri[i] * t_rh[i] + al * B;
      A = al * A;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {  // inclusive, lane order
      const double Ae = __shfl_up(A, off), Be = __shfl_up(B, off);
      if (lane >= off) {
        B = A * Be + B;
        A = A * Ae;
      }
    }
    if (lane == 63) {
      s_mapA[wave] = A;
      s_mapB[wave] = B;
    }
    __syncthreads();
    double zin = 0.0;  // z entering this wave (z_{-1} = 0)
    for (int v = 0; v < wave; v++) zin = s_mapA[v] * zin + s_mapB[v];
    {
      const double Ap = __shfl_up(A, 1), Bp = __shfl_up(B, 1);
      double z = lane > 0 ? Ap * zin + Bp : zin;
      for (int i = r0; i < r1; i++) {
        z = t_ri[i] * t_rh[i] + t_al[i] * z;
        t_rh[i] = z;
      }
    }
    __syncthreads();
    // backward: chunk map y_out(first row) = A y_in(after the chunk) + B
    A = 1.0;
    B = 0.0;
    for (int i = r1 - 1; i >= r0; i--) {
      const double mu = t_mu[i];
      B = t_rh[i] + mu * B;
      A = mu * A;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {  // inclusive, reverse lane order
      const double Al = __shfl_down(A, off), Bl = __shfl_down(B, off);
      if (lane + off < 64) {
        B = A * Bl + B;
        A = A * Al;
      }
    }
    if (lane == 0) {
      s_mapA[wave] = A;
      s_mapB[wave] = B;
    }
    __syncthreads();
    double yin = 0.0;  // y entering this wave from above (y_n = 0)
    for (int v = AM_THREADS / 64 - 1; v > wave; v--) yin = s_mapA[v] * yin + s_mapB[v];
    {
      const double An = __shfl_down(A, 1), Bn = __shfl_down(B, 1);
      double y = lane < 63 ? An * yin + Bn : yin;
      for (int i = r1 - 1; i >= r0; i--) {
        y = t_rh[i] + t_mu[i] * y;
        t_rh[i] = y;
      }
    }
    __syncthreads();
    for (int u = tid; u < n; u += AM_THREADS)
      hr[(long long)cur * n + u] += coef * a.gs_shift * t_rh[u];
    __syncthreads();
  }

  // ----------------------------------------------------- Gram system
  // U_ij = sum_u dv_i dv_j and b_i = sum_u dv_i r_k with dv_i = r_k - r_{k-1-i}
  // on the FP64 matrix cores: v_mfma_f64_16x16x4f64 on a 16 x 4 block of dv
  // (unknowns u0 .. u0+3), lane l holding dv_{l%16}(u0 + l/16), which is both
  // its A element (row l%16, k = l/16) and its B element (k = l/16, column
  // l%16); a second product against r_k in column 0 gives b.  One pass over
  // the history, 8 accumulator doubles per lane, m <= 16.
  if (m > 0) {
    typedef double d4 __attribute__((ext_vector_type(4)));
    const int lane = tid & 63, wave = tid >> 6;
    const int vi = lane & 15, ku = lane >> 4;
    const int slot = ((cur - (vi + 1)) % S + S) % S;
    const double* __restrict__ hk = hr + (long long)cur * n;
    const double* __restrict__ hi = hr + (long long)slot * n;
    const bool use = vi < m;
    // each wave takes a contiguous range of unknowns, four at a time
    constexpr int NW = AM_THREADS / 64;
    const int per = ((n + 4 * NW - 1) / (4 * NW)) * 4;
    const int ub = wave * per, ue = min(n, ub + per);
    d4 accG = {0.0, 0.0, 0.0, 0.0}, accB = {0.0, 0.0, 0.0, 0.0};
    // GU blocks of 4 unknowns per trip: their loads are all in flight before
    // the first product (blocks past the range multiply zeros)
    constexpr int GU = 8;
    for (int u0 = ub; u0 < ue; u0 += 4 * GU) {
      double rkv[GU], hvv[GU];
#pragma unroll
      for (int g = 0; g < GU; g++) {
        const int u = u0 + 4 * g + ku;
        const int uc = (u < ue) ? u : ub;  // in-range index for the masked lanes
        rkv[g] = hk[uc];
        hvv[g] = hi[uc];
      }
#pragma unroll
      for (int g = 0; g < GU; g++) {
        const bool ok = u0 + 4 * g + ku < ue;
        const double rk = ok ? rkv[g] : 0.0;
        const double dv = (ok && use) ? rk - hvv[g] : 0.0;
        accG = __builtin_amdgcn_mfma_f64_16x16x4f64(dv, dv, accG, 0, 0, 0);
        accB = __builtin_amdgcn_mfma_f64_16x16x4f64(dv, vi == 0 ? rk : 0.0, accB, 0, 0, 0);
      }
    }
    // result layout: lane l, element e holds row l/16 + 4 e, column l%16
#pragma unroll
    for (int e = 0; e < 4; e++) {
      red[wave][(ku + 4 * e) * 16 + vi] = accG[e];
      if (vi == 0) red[wave][256 + ku + 4 * e] = accB[e];
    }
    __syncthreads();
    for (int e = tid; e < m * m + m; e += AM_THREADS) {
      const int idx = (e < m * m) ? (e / m) * 16 + (e % m) : 256 + (e - m * m);
      double t = 0.0;
      for (int v = 0; v < NW; v++) t += red[v][idx];
      if (e < m * m)
        U[e] = t;
      else
        bvec[e - m * m] = t;
    }
    __syncthreads();
    if (tid == 0) {
      double tr = 0.0;
      for (int i = 0; i < m; i++) tr += U[i * m + i];
      for (int i = 0; i < m; i++) U[i * m + i] += a.reg * tr;
      s_tr = tr;
    }
    __syncthreads();
    // Gauss-Jordan with full pivoting (gaussj.c:7-60) on all threads: thread
    // t owns element (t / 16, t % 16) of U and threads t < m the entry b_t.
    // Each step takes the pivot the serial scan takes (the last largest |U_jk|
    // in row-major order over rows and columns not yet pivoted), swaps,
    // scales and eliminates with the same operation per element, so the
    // solution is the serial one bit for bit.  A zero pivot = singular.
    bool sing = !(s_tr > 0.0);
    {
      const int gj = tid >> 4, gk = tid & 15;
      const bool own = gj < m && gk < m;
      unsigned pivoted = 0;  // bit c: column c holds a pivot
      for (int i = 0; i < m && !sing; i++) {
        double v = -2.0;
        int idx = -1;
        if (own && !((pivoted >> gj) & 1u) && !((pivoted >> gk) & 1u)) {
          const double av = fabs(U[gj * m + gk]);
          v = (av == av) ? av : -1.0;  // a NaN never wins, as in the scan
          idx = gj * m + gk;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          const double ov = __shfl_xor(v, off);
          const int oi = __shfl_xor(idx, off);
          if (ov > v || (ov == v && oi > idx)) {
            v = ov;
            idx = oi;
          }
        }
        if ((tid & 63) == 0) {
          s_pv[tid >> 6] = v;
          s_pi[tid >> 6] = idx;
        }
        __syncthreads();
        v = s_pv[0];
        idx = s_pi[0];
#pragma unroll
        for (int w2 = 1; w2 < AM_THREADS / 64; w2++)
          if (s_pv[w2] > v || (s_pv[w2] == v && s_pi[w2] > idx)) {
            v = s_pv[w2];
            idx = s_pi[w2];
          }
        const int irow = idx / m, icol = idx - (idx / m) * m;
        pivoted |= 1u << icol;
        if (irow != icol) {
          if (tid < m) {
            const double t = U[irow * m + tid];
            U[irow * m + tid] = U[icol * m + tid];
            U[icol * m + tid] = t;
          }
          if (tid == 0) {
            const double t = bvec[irow];
            bvec[irow] = bvec[icol];
            bvec[icol] = t;
          }
        }
        __syncthreads();
        const double piv = U[icol * m + icol];
        if (piv == 0.0) {
          sing = true;
          break;
        }
        const double pivinv = 1.0 / piv;
        double cur_v = 0.0, dum = 0.0, prow = 0.0, bj = 0.0, dumb = 0.0;
        if (own) {
          cur_v = U[gj * m + gk];
          dum = U[gj * m + icol];
          prow = U[icol * m + gk];
        }
        if (tid < m) {
          bj = bvec[tid];
          dumb = U[tid * m + icol];
        }
        const double bpiv = bvec[icol];
        __syncthreads();
        if (own) {
          const double pr = (gk == icol ? 1.0 : prow) * pivinv;  // scaled pivot row
          U[gj * m + gk] = (gj == icol) ? pr : (gk == icol ? 0.0 : cur_v) - pr * dum;
        }
        if (tid < m) {
          const double bp = bpiv * pivinv;
          bvec[tid] = (tid == icol) ? bp : bj - bp * dumb;
        }
        __syncthreads();
      }
    }
    for (int i = 0; i < m && !sing; i++)
      if (!isfinite(bvec[i])) sing = true;
    if (sing) ms.restarts++;
    if (tid == 0) s_m = sing ? 0 : m;
    __syncthreads();
    const int sm = __builtin_amdgcn_readfirstlane(s_m);
    if (sm == 0) stored = 0;  // keep only the current pair
    m = sm;
  }

  // ----------------------------------------------------------- update
  // lower bound of the field (field_floor > 0): 1 + dt w >= 1 - field_floor
  // keeps every row of the implicit-Euler matrix strictly diagonally dominant
  const double lo = a.field_floor > 0.0 ? -a.field_floor * d.N : -INFINITY;
  double stepmax = 0.0;
  for (int u = tid; u < n; u += AM_THREADS) {
    const double wk = w[u];
    const double rk = hr[(long long)cur * n + u];
    double wb = wk, rb = rk;
    for (int i = 0; i < m; i++) {
      int sl = cur - (i + 1);  // ring slot of iterate k-1-i (m < S)
      sl += (sl < 0) ? S : 0;
      wb -= bvec[i] * (wk - hw[(long long)sl * n + u]);
      rb -= bvec[i] * (rk - hr[(long long)sl * n + u]);
    }
    const double wn = wb + a.beta * rb;
    stepmax = fmax(stepmax, fabs(wn - wk));
    w[u] = a.step_cap > 0.0 ? wn : fmax(wn, lo);  // capped: rescaled and bounded below
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) stepmax = fmax(stepmax, __shfl_xor(stepmax, off));
  if ((tid & 63) == 0) red1[tid >> 6] = stepmax;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int v = 0; v < AM_THREADS / 64; v++) s = fmax(s, red1[v]);
    s_scale = (a.step_cap > 0.0 && s > a.step_cap) ? a.step_cap / s : 1.0;
  }
  __syncthreads();
  const double scale = s_scale;
  if (scale != 1.0 || (a.step_cap > 0.0 && a.field_floor > 0.0)) {
    for (int u = tid; u < n; u += AM_THREADS) {
      const double wk = hw[(long long)cur * n + u];
      w[u] = fmax(wk + scale * (w[u] - wk), lo);
    }
  }
  if (tid == 0) {
    ms.k = k;
    ms.head = (cur + 1) % S;
    ms.stored = (stored < S) ? stored + 1 : S;
    a.st[inst] = ms;
    atomicAdd(a.active, 1);
  }
}

// zero the per-sweep counters (bucket queues and the active count) with a
// kernel rather than a memset node, so the sweep can be captured in a graph
__global__ void reset_counters_kernel(int* counters, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) counters[i] = 0;
}

// The Thomas factors of g I + G0 for the rational ground-state form, one
// wave per instance: rows (lower l_u d_{u-1}/d_u, upper r_u d_{u+1}/d_u,
// diagonal g + lower + upper) as in anderson_kernel; the pivot recurrence
// den_u = diag_u - lower_u upper_{u-1} / den_{u-1} walked in blocks of 64
// rows, each lane holding one row and every lane following the
// (wave-uniform) recurrence through readlane.  Stores 1/den, lower/den and
// upper/den at tri[0..2] x tri_stride.
__global__ void __launch_bounds__(64) gs_pivots_kernel(AndersonArgs a) {
  const int inst = blockIdx.x, lane = threadIdx.x;
  const InstDesc d = a.desc[inst];
  const int n = d.n;
  const double* __restrict__ pt = a.phit + d.uo;
  const double* __restrict__ xn = a.x + d.xo;
  double* __restrict__ t_ri = a.tri + d.uo;
  double* __restrict__ t_al = t_ri + a.tri_stride;
  double* __restrict__ t_mu = t_al + a.tri_stride;
  double cpv = 0.0;  // -upper_{u-1} / den_{u-1}
  for (int b0 = 0; b0 < n; b0 += 64) {
    const int u = b0 + lane;
    double lo = 0.0, up = 0.0;
    if (u < n) {
      const int i = u + 1;
      const double dv = sqrt(fmax(pt[u], a.floor));
      const double dl = u > 0 ? sqrt(fmax(pt[u - 1], a.floor)) : 0.0;
      const double dr = u + 1 < n ? sqrt(fmax(pt[u + 1], a.floor)) : 0.0;
      const bool node_r = i + 1 < d.M;
      const double hl = xn[i] - xn[i - 1];
      const double hrr = node_r ? xn[i + 1] - xn[i] : 0.0;
      const double mm = 0.5 * (hl + hrr);
      const double cl = (1.0 / hl) / mm;
      const double cr = node_r ? (1.0 / hrr) / mm : 0.0;
      lo = cl * (dl / dv);
      up = cr * (dr / dv);
    }
    const double dg = a.gs_shift + lo + up;
    const int cnt = min(64, n - b0);
    double myri = 0.0;
    for (int j = 0; j < cnt; j++) {
      const double loj = readlane_f64(lo, j), dgj = readlane_f64(dg, j);
      const double upj = readlane_f64(up, j);
      const double den = b0 + j == 0 ? dgj : dgj + loj * cpv;
      const double ri = 1.0 / den;
      cpv = -upj * ri;
      if (lane == j) myri = ri;
    }
    if (u < n) {
      t_ri[u] = myri;
      t_al[u] = lo * myri;
      t_mu[u] = up * myri;
    }
  }
}

hipError_t launch_gs_pivots(const AndersonArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(gs_pivots_kernel, dim3(a.n_inst), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_anderson(const AndersonArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(anderson_kernel, dim3(a.n_inst), dim3(AM_THREADS), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_reset(int* counters, int n, hipStream_t s) {
  hipLaunchKernelGGL(reset_counters_kernel, dim3((n + 255) / 256), dim3(256), 0, s, counters,
                     n);
  return hipGetLastError();
}

}  // namespace scft
