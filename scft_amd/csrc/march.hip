// march.hip -- the hot path of the SCFT field evaluation on gfx950, FP64.
//
// One workgroup of W wavefronts (P = 64 W lanes) evaluates one instance at a
// time; lane p owns K consecutive unknowns, so n <= P K.  Per instance:
//
//   1. assembly  (1D_FEM.c:84-186, simple_FEM_1D_transient.m:28-79,
//                 scft.cc:641-651): every lane builds its own rows of
//                 D = A + dt (B + C) and of the right-hand-side operator A from
//                 the node coordinates and the field (no atomics: a row reads
//                 its two adjacent elements itself).
//   2. factor:   the Thomas pivots u_i = b_i - a_i c_{i-1}/u_{i-1} are a
//                linear-fractional recurrence; each lane multiplies its K
//                2x2 matrices [[b_i, -a_i c_{i-1}], [1, 0]], a wave-wide
//                prefix product (plus a fold over waves) gives every lane the
//                pivot entering its chunk, then the lane finishes its pivots
//                sequentially.  The solve then reads
//                    z_i = r'_i - alpha_i z_{i-1},   x_i = z_i - gamma_i x_{i+1}
//                with alpha = a/u, gamma = c/u and r' = (A q)/u, all kept in
//                VGPRs for the whole march.
//   3. march:    N implicit-Euler steps D q^{j+1} = A q^j (1D_FEM.c:208-228,
//                step count of DEALII_SCFT/drivescft.cc:130).  Each first-order
//                recurrence is split into a lane-local pass, a scan of the
//                lanes' affine maps over the wave (four DPP row levels, then
//                two cross-row levels: broadcast DPP moves forward, LDS-
//                crossbar permutes backward; every multiplier is fixed per
//                instance, so a level costs one FMA per step), a fold over
//                waves through LDS in one round trip, and a lane-local re-run.
//   4. density:  phi_i = sum_j c_j q_j(i) q_{N-j}(i) (drivescft.cc:184-193 with
//                romint.c's weights, symmetric in j).  Slabs q_1..q_{N/2-1}
//                stream to HBM (coalesced: element t of lane p at t*P + p,
//                raw-buffer stores and loads with non-temporal hints); from
//                step N/2 on each new slab is paired with the stored slab
//                N-j, loaded a few steps ahead through a register ring.
//   5. residual: r = phi - phi_t (1D_FEM.c:274-277), ||r||_inf per instance.
//
// No MFMA: nothing here is a dense contraction.
#include "scft_internal.h"
#include "device_common.h"

namespace scft {

template <int W>
struct MarchShared {
  double Bw[W];     // forward-scan wave totals (per step)
  double Cw[W];     // backward-scan wave totals (per step)
  double Aw[W];     // forward wave transfer products (per instance)
  double Gw[W];     // backward wave transfer products (per instance)
  double cx[W];     // last super-diagonal of each wave (setup)
  double gx[W];     // last gamma of each wave (setup)
  double Mx[W][4][2];  // wave products of the pivot matrices, double-double (setup)
  double red[W];    // norm reduction
  unsigned bad[W];
};


// One row of D = A + dt (B + C) and of the right-hand-side operator A for
// unknown u (node u+1), from its two adjacent elements.  Evaluated without
// FMA contraction and in the same operation order as the CPU restatement:
// the rows of D are nearly identical, so a contracted rounding would shift
// every row the same way and move the diagonal-dominance margin
// b - |a| - |c| ~ h by eps * dt/h^2 relative per step, which adds up over the
// march (~1e-10 after a few hundred steps when dt/h^2 ~ 1e3).
struct Row {
  double a, b, c;  // D = A + dt (B + C): sub-, main and super-diagonal
  double ra, rc;   // A q = ra (q_{i-1} + 2 q_i) + rc (2 q_i + q_{i+1})
};

// Returned by value with every field assigned once (selects, no stores
// through references): conditional stores to two different outputs get
// merged into one store through a pointer phi, which keeps the row arrays
// out of registers.
__device__ __forceinline__ Row assemble_row(const double* __restrict__ x,
                                            const double* __restrict__ w, int n, int M, int u,
                                            double dt, int cmat, double wall0, double wallL) {
#pragma clang fp contract(off)
  const bool live = u < n;
  const int i = live ? u + 1 : 1;  // padding rows read node 1 and are discarded
  const double xl = x[i - 1], xc = x[i];
  const bool has_r = (i + 1 < M);
  const double xr = has_r ? x[i + 1] : xc;
  const double hl = xc - xl, hr = xr - xc;
  const double wi = live ? w[live ? u : 0] : 0.0;  // index clamped: no load past the field
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
  const double av = Al + dt * (Bl + Cl);
  const double bv = Ad + dt * (Bd + Cd);
  const double cv = Ar + dt * (Br + Cr);
  // left neighbour of unknown 0 is the Dirichlet wall; right neighbour of
  // unknown n-1 is a wall (DD) or none (DN); padding rows are decoupled
  // identity rows with a zero right-hand side (q stays 0 there)
  Row r;
  r.a = (live && u > 0) ? av : 0.0;
  r.c = (live && u < n - 1) ? cv : 0.0;
  r.b = live ? bv : 1.0;
  r.ra = live ? Al : 0.0;
  r.rc = live ? Ar : 0.0;
  return r;
}

template <int K, int W>
__device__ void march_one(const MarchArgs& a, const int inst, const unsigned st0,
                          double* __restrict__ slab, MarchShared<W>& sh,
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

  // pairing weights of this N into LDS (the previous instance's reads of the
  // table ended at its residual barrier / in program order for one wave)
  wg_barrier();
  for (int i = tid; i <= (N >> 1); i += P) pw_lds[i] = pw[i];
  wg_barrier();

  // ---------------------------------------------------------------- assembly
  double av[K], bv[K], cv[K], ra[K], rc[K];
#pragma unroll
  for (int t = 0; t < K; t++) {
    const Row rw = assemble_row(x, w, n, M, u0 + t, dt, cmat, d.wall0, d.wallL);
    av[t] = rw.a;
    bv[t] = rw.b;
    cv[t] = rw.c;
    ra[t] = rw.ra;
    rc[t] = rw.rc;
  }

  // ------------------------------------------------------------ factor scan
  double cprev = __shfl_up(cv[K - 1], 1);
  if constexpr (W > 1) {
    sh.cx[wave] = readlane_d(cv[K - 1], 63);
    wg_barrier();
    const double cx_prev = (wave > 0) ? sh.cx[wave - 1] : 0.0;
    cprev = (lane == 0) ? cx_prev : cprev;
  } else {
    cprev = (lane == 0) ? 0.0 : cprev;
  }
  dd T[4] = {dd_from(1.0), dd_from(0.0), dd_from(0.0), dd_from(1.0)};
  {
    double cp = cprev;
#pragma unroll
    for (int t = 0; t < K; t++) {
      const dd Mt[4] = {dd_from(bv[t]), two_prod(-av[t], cp), dd_from(1.0), dd_from(0.0)};
      mat_mul_norm(Mt, T, T);
      cp = cv[t];
    }
  }
#pragma unroll
  for (int dd_ = 0; dd_ < 6; dd_++) {
    const int s = 1 << dd_;
    dd Tp[4];
#pragma unroll
    for (int e = 0; e < 4; e++) Tp[e] = shfl_up_dd(T[e], s);
    if (lane >= s) mat_mul_norm(T, Tp, T);
  }
  dd E[4];
#pragma unroll
  for (int e = 0; e < 4; e++) E[e] = shfl_up_dd(T[e], 1);
  dd Pre[4] = {dd_from(1.0), dd_from(0.0), dd_from(0.0), dd_from(1.0)};
  if constexpr (W > 1) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      sh.Mx[wave][e][0] = readlane_d(T[e].h, 63);
      sh.Mx[wave][e][1] = readlane_d(T[e].l, 63);
    }
    wg_barrier();
    for (int v = 0; v < wave; v++) {
      const dd Mv[4] = {{sh.Mx[v][0][0], sh.Mx[v][0][1]}, {sh.Mx[v][1][0], sh.Mx[v][1][1]},
                        {sh.Mx[v][2][0], sh.Mx[v][2][1]}, {sh.Mx[v][3][0], sh.Mx[v][3][1]}};
      mat_mul_norm(Mv, Pre, Pre);
    }
  }
  if (lane > 0) {
    mat_mul_norm(E, Pre, E);
  } else {
#pragma unroll
    for (int e = 0; e < 4; e++) E[e] = Pre[e];
  }
  // pivot of the row just before this lane's chunk (unused for tid 0: a = 0)
  double up = (tid > 0) ? dd_div_to_double(E[0], E[2]) : 1.0;

  unsigned bad = 0;
  double al[K], ga[K];
  {
#pragma clang fp contract(off)
    double cp = cprev;
#pragma unroll
    for (int t = 0; t < K; t++) {
      // Thomas order: l = a / u_prev, u = b - l c_prev
      const double l = (av[t] != 0.0) ? av[t] / up : 0.0;
      const double uu = bv[t] - l * cp;
      if (!(uu > 0.0) || !isfinite(uu)) bad |= ST_PIVOT;
      const double inv = 1.0 / uu;
      al[t] = av[t] * inv;
      ga[t] = cv[t] * inv;
      ra[t] *= inv;
      rc[t] *= inv;
      cp = cv[t];
      up = uu;
    }
  }

  // -------------------------------------------- scan constants of the march
  // Each lane's chunk is an affine map: forward z_out = Af z_in + B, backward
  // x_first = Gb x_next + C, with Af = prod(-alpha), Gb = prod(-gamma) fixed
  // for the instance.  The per-step scans combine these maps over the wave:
  // a Hillis-Steele scan inside each 16-lane row (DPP row_shr / row_shl by
  // 1, 2, 4, 8), then two row-to-row levels (forward: DPP row_bcast:15 and
  // row_bcast:31; backward: ds_bpermute from the first lane of the next
  // rows), then (several waves) a fold over wave totals in LDS.  All
  // multipliers are precomputed here; a step costs one FMA per level.
  const int rl = lane & 15;   // lane within its row
  const int row = lane >> 4;  // row within the wave
  double Af = 1.0, Gb = 1.0;
#pragma unroll
  for (int t = 0; t < K; t++) {
    Af *= -al[t];
    Gb *= -ga[t];
  }
  // Within the chunk, z_t = zloc_t + Pf[t] zin with zloc the local pass from
  // zin = 0 and Pf[t] = prod_{s <= t} (-alpha_s); so the lane's backward offset
  // (the local backward pass over z) is C0 + Sz zin, C0 the local backward
  // pass over zloc and Sz = sum_t prod_{s < t} (-gamma_s) Pf[t].  C0 needs no
  // zin: it runs while the forward scan is in flight, and the final forward
  // pass leaves the critical path.
  double Pf[K], Sz = 0.0;
  {
    double p = 1.0, g = 1.0;
#pragma unroll
    for (int t = 0; t < K; t++) {
      p *= -al[t];
      Pf[t] = p;
      Sz = fma(g, p, Sz);
      g *= -ga[t];
    }
  }
  double mf[4], mb[4];
  {
    double Ap, Gp;
    Ap = row_shr<1>(Af); Gp = row_shl<1>(Gb);
    mf[0] = (rl >= 1) ? Af : 0.0; mb[0] = (rl + 1 < 16) ? Gb : 0.0;
    Af = (rl >= 1) ? Af * Ap : Af; Gb = (rl + 1 < 16) ? Gb * Gp : Gb;
    Ap = row_shr<2>(Af); Gp = row_shl<2>(Gb);
    mf[1] = (rl >= 2) ? Af : 0.0; mb[1] = (rl + 2 < 16) ? Gb : 0.0;
    Af = (rl >= 2) ? Af * Ap : Af; Gb = (rl + 2 < 16) ? Gb * Gp : Gb;
    Ap = row_shr<4>(Af); Gp = row_shl<4>(Gb);
    mf[2] = (rl >= 4) ? Af : 0.0; mb[2] = (rl + 4 < 16) ? Gb : 0.0;
    Af = (rl >= 4) ? Af * Ap : Af; Gb = (rl + 4 < 16) ? Gb * Gp : Gb;
    Ap = row_shr<8>(Af); Gp = row_shl<8>(Gb);
    mf[3] = (rl >= 8) ? Af : 0.0; mb[3] = (rl + 8 < 16) ? Gb : 0.0;
    Af = (rl >= 8) ? Af * Ap : Af; Gb = (rl + 8 < 16) ? Gb * Gp : Gb;
  }
  // Af = product over the row from its first lane to this lane, Gb = product
  // from this lane to the row's last lane; row totals as uniform scalars
  const double Arow0 = readlane_d(Af, 15), Arow1 = readlane_d(Af, 31);
  const double Arow2 = readlane_d(Af, 47), Arow3 = readlane_d(Af, 63);
  const double Grow0 = readlane_d(Gb, 0), Grow1 = readlane_d(Gb, 16);
  const double Grow2 = readlane_d(Gb, 32), Grow3 = readlane_d(Gb, 48);
  // DPP reads are made with every lane active: a DPP move in a conditional
  // arm runs under a partial EXEC mask and reads 0 from disabled lanes.
  const double Af_sh = row_shr<1>(Af), Gb_sh = row_shl<1>(Gb);
  const double Ashift = (rl > 0) ? Af_sh : 0.0;   // row product to lane-1
  const double Gshift = (rl < 15) ? Gb_sh : 0.0;  // row product from lane+1
  // products over whole rows before / after this lane's row
  const double Abefore = (row == 0) ? 1.0 : (row == 1) ? Arow0
                         : (row == 2) ? Arow0 * Arow1 : Arow0 * Arow1 * Arow2;
  const double Gafter = (row == 3) ? 1.0 : (row == 2) ? Grow3
                        : (row == 1) ? Grow2 * Grow3 : Grow1 * Grow2 * Grow3;
  // wave prefix over lanes 0..lane-1 and suffix over lane+1..63 (wave folds)
  const double Apre_m1 = ((rl > 0) ? Ashift : 1.0) * Abefore;
  // multipliers of the forward scan's two broadcast levels: rows 1 and 3 take
  // the prefix ending at lane 15 / 47 times their own row product up to the
  // lane; then rows 2 and 3 take the prefix ending at lane 31 times the
  // product from lane 32 up to the lane
  const double mf5 = Af;
  const double mf6 = (row == 3) ? Arow2 * Af : Af;
  // the backward scan's two cross-row levels, mirror images of those (DPP has
  // no downward broadcast: the values come through ds_bpermute): rows 0 and 2
  // take the suffix starting at lane 16 / 48 times their own row product
  // from the lane on; then rows 0 and 1 take the suffix starting at lane 32
  // times the product from the lane to lane 31.  0 for rows that do not take.
  const double mbA = (row == 0 || row == 2) ? Gb : 0.0;
  const double mbB = (row == 0) ? Gb * Grow1 : (row == 1) ? Gb : 0.0;
  const int srcA = ((row + 1) & 3) * 64;  // byte address of lane 16 (row + 1)
  const int srcB = 32 * 4;
  const double Gsuf_p1 = ((rl < 15) ? Gshift : 1.0) * Gafter;
  double gl = __shfl_up(ga[K - 1], 1);  // gamma of the row left of me
  // Fold multipliers.  The forward carry into this wave from the waves before
  // it is Z = sum_{v < wave} fz[v] Bw[v] with fz[v] = prod_{v < u < wave} Aw[u];
  // the backward carry from the waves after it is X = sum_{v > wave} fx[v] Cw[v]
  // with fx[v] = prod_{wave < u < v} Gw[u].  Fixed per instance (0 where the
  // wave does not read) and held as wave-uniform values, so a step reads all
  // wave totals at once: one LDS round trip instead of one per earlier wave.
  [[maybe_unused]] double fz[W], fx[W];
  if constexpr (W > 1) {
    sh.Aw[wave] = Arow0 * Arow1 * Arow2 * Arow3;
    sh.gx[wave] = readlane_d(ga[K - 1], 63);
    sh.Gw[wave] = Grow0 * Grow1 * Grow2 * Grow3;
    wg_barrier();
    const double gx_prev = (wave > 0) ? sh.gx[wave - 1] : 0.0;
    gl = (lane == 0) ? gx_prev : gl;
    double p = 1.0;
#pragma unroll
    for (int v = W - 1; v >= 0; v--) {
      fz[v] = readlane_d((v < wave) ? p : 0.0, 0);
      p = (v < wave) ? p * sh.Aw[v] : p;
    }
    p = 1.0;
#pragma unroll
    for (int v = 0; v < W; v++) {
      fx[v] = readlane_d((v > wave) ? p : 0.0, 0);
      p = (v > wave) ? p * sh.Gw[v] : p;
    }
  } else {
    gl = (lane == 0) ? 0.0 : gl;
  }

  // ------------------------------------------------------------------ march
  double q[K], acc[K];
#pragma unroll
  for (int t = 0; t < K; t++) {
    q[t] = (u0 + t < n) ? 1.0 : 0.0;
    acc[t] = 0.0;
  }
  double ql = (u0 - 1 >= 0 && u0 - 1 < n) ? 1.0 : 0.0;
  double qr = (u0 + K < n) ? 1.0 : 0.0;
  const int half = N >> 1;

  // One implicit-Euler step q <- D^{-1} (A q) for the whole instance; ql / qr
  // become the new values of the unknowns left of q[0] and right of q[K-1].
  auto advance = [&]() __attribute__((always_inline)) {
    double rp[K];
#pragma unroll
    for (int t = 0; t < K; t++) {
      const double qm = (t == 0) ? ql : q[t - 1];
      const double qp = (t == K - 1) ? qr : q[t + 1];
      rp[t] = ra[t] * (qm + 2.0 * q[t]) + rc[t] * (2.0 * q[t] + qp);
    }
    // forward recurrence z_i = r'_i - alpha_i z_{i-1}: the local pass from 0
    // (zloc kept in rp), and the local backward pass over it
    double B = 0.0;
#pragma unroll
    for (int t = 0; t < K; t++) {
      B = rp[t] - al[t] * B;
      rp[t] = B;
    }
    double C0 = 0.0;
#pragma unroll
    for (int t = K - 1; t >= 0; t--) C0 = rp[t] - ga[t] * C0;
    B = fma(mf[0], row_shr<1>(B), B);
    B = fma(mf[1], row_shr<2>(B), B);
    B = fma(mf[2], row_shr<4>(B), B);
    B = fma(mf[3], row_shr<8>(B), B);
    // rows to the whole wave: broadcast levels instead of readlane carries
    B = fma(mf5, row_bcast15(B), B);
    B = fma(mf6, row_bcast31(B), B);
    // B is now the wave-inclusive prefix; this lane's input is lane - 1's
    double zin = wave_shr1(B);
    if constexpr (W > 1) {
      sh.Bw[wave] = readlane_d(B, 63);
      wg_barrier();
      double Z0 = 0.0, Z1 = 0.0;
#pragma unroll
      for (int v = 0; v + 1 < W; v += 2) {
        Z0 = fma(fz[v], sh.Bw[v], Z0);
        if (v + 2 < W) Z1 = fma(fz[v + 1], sh.Bw[v + 1], Z1);
      }
      zin = fma(Apre_m1, Z0 + Z1, zin);
    }
    // backward recurrence x_i = z_i - gamma_i x_{i+1}: this lane's offset
    double Cc = fma(Sz, zin, C0);
    Cc = fma(mb[0], row_shl<1>(Cc), Cc);
    Cc = fma(mb[1], row_shl<2>(Cc), Cc);
    Cc = fma(mb[2], row_shl<4>(Cc), Cc);
    Cc = fma(mb[3], row_shl<8>(Cc), Cc);
    // carries into rows 0..2: the suffix from the first lane of the row after
    Cc = fma(mbA, bpermute_d(srcA, Cc), Cc);
    Cc = fma(mbB, bpermute_d(srcB, Cc), Cc);
    // Cc is now the wave-inclusive suffix; this lane's input is lane + 1's
    double xin = wave_shl1(Cc);
    if constexpr (W > 1) {
      sh.Cw[wave] = readlane_d(Cc, 0);
      wg_barrier();
      double X0 = 0.0, X1 = 0.0;
#pragma unroll
      for (int v = 1; v < W; v += 2) {
        X0 = fma(fx[v], sh.Cw[v], X0);
        if (v + 1 < W) X1 = fma(fx[v + 1], sh.Cw[v + 1], X1);
      }
      xin = fma(Gsuf_p1, X0 + X1, xin);
    }
    {
      double xx = xin;
#pragma unroll
      for (int t = K - 1; t >= 0; t--) {
        xx = fma(Pf[t], zin, rp[t]) - ga[t] * xx;  // z_t - gamma_t x_{t+1}
        q[t] = xx;
      }
    }
    ql = zin - gl * q[0];
    qr = xin;
  };

  // Slab s (q after step s, s = 0 .. N/2 - 1; slab 0 is the initial
  // condition) starts at s * P * K; within it row t of lane tid sits at
  // (t / 2) * 2P + 2 tid + t % 2 (K >= 2: 16 bytes per lane and access) or at
  // tid (K = 1).  Lanes whose rows are all
  // padding (u0 >= n) store nothing (one exec mask for the whole loop) and
  // load from lane 0's address, so a wave's dead lanes share one cache line;
  // what they load only ever multiplies their q = 0.  The region is accessed
  // as a raw buffer (slab_rsrc): the step and element select a scalar
  // offset, the lane a fixed 32-bit offset.
  // First half: nothing here waits on memory.
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
  // step N/2 pairs with itself
  advance();
  {
    const double c0 = pw_lds[0];
#pragma unroll
    for (int t = 0; t < K; t++) acc[t] = c0 * q[t] * q[t];
  }

  // Second half: step N/2 + i pairs with slab N/2 - i (i = 1 .. N/2).  Slabs
  // come through a ring of PD register sets, each loaded PD steps before its
  // use (indices below 0 reload slab 0 and are never used).  PD divides N/2
  // (N is a power of two >= 16).
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

  // --------------------------------------------------------------- residual
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
    wg_barrier();  // sh.red / sh.bad are reused by the next instance
  }
  // every lane of wave 0 stores the same (uniform) values: no lane branch
  if (wave == 0) {
    a.rnorm[inst] = (bad & ST_NAN) ? __builtin_nan("") : mx;
    a.status[inst] = st0 | bad;
  }
}

// Persistent workgroups take items (sorted by cost, largest first) from a
// per-bucket counter: greedy list scheduling, which also rebalances as
// instances converge.  Thread 0 takes the next index and broadcasts it
// through LDS; the exit test is on that workgroup-uniform value, so every
// wave leaves the loop together once the counter passes n_items.
template <int K, int W>
__global__ void __launch_bounds__(64 * W) march_kernel(MarchArgs a) {
  __shared__ MarchShared<W> sh;
  __shared__ int s_next;
  extern __shared__ double pw_lds[];  // half_max + 1 pairing weights
  const int tid = threadIdx.x;
  double* slab = a.slabs + (long long)blockIdx.x * a.slab_stride;
  for (;;) {
    if (tid == 0) s_next = atomicAdd(a.queue, 1);
    wg_barrier();
    const int it = __builtin_amdgcn_readfirstlane(s_next);
    wg_barrier();  // everyone has read s_next before thread 0 overwrites it
    if (it >= a.n_items) break;
    const int inst = __builtin_amdgcn_readfirstlane(a.items[it]);
    const unsigned st = __builtin_amdgcn_readfirstlane(a.status[inst]);
    if (st & ST_DONE_MASK) continue;
    march_one<K, W>(a, inst, st, slab, sh, pw_lds, tid);
  }
}

// ------------------------------------------------------------------ dispatch
// (K, W) variants: K * 8 VGPRs of coefficients per lane must leave room for
// the register file share of a 64W-lane workgroup (no spills).
#define SCFT_MARCH_VARIANTS(X) \
  X(1, 1) X(2, 1) X(4, 1) X(8, 1) X(16, 1) \
  X(1, 2) X(2, 2) X(4, 2) X(8, 2) X(16, 2) \
  X(1, 4) X(2, 4) X(4, 4) X(8, 4) X(16, 4) \
  X(1, 8) X(2, 8) X(4, 8) X(8, 8)          \
  X(1, 16) X(2, 16)

static size_t march_lds_bytes(int half_max) {
  return static_cast<size_t>(half_max + 1) * sizeof(double);
}

static const void* march_fn(int K, int W) {
#define X(k, w) \
  if (K == k && W == w) return reinterpret_cast<const void*>(&march_kernel<k, w>);
  SCFT_MARCH_VARIANTS(X)
#undef X
  return nullptr;
}

bool march_variant_exists(int K, int W) { return march_fn(K, W) != nullptr; }

hipError_t launch_march(int K, int W, const MarchArgs& a, int grid, hipStream_t s) {
#define X(k, w)                                                                 \
  if (K == k && W == w) {                                                       \
    hipLaunchKernelGGL((march_kernel<k, w>), dim3(grid), dim3(64 * w),           \
                       march_lds_bytes(a.half_max), s, a);                      \
    return hipGetLastError();                                                   \
  }
  SCFT_MARCH_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

int march_max_grid(int K, int W, int half_max) {
  const void* f = march_fn(K, W);
  if (!f) return 0;
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, 64 * W, march_lds_bytes(half_max)) !=
      hipSuccess)
    return 0;
  if (per < 1) per = 1;
  return per * cus;
}

}  // namespace scft
