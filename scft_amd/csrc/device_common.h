// device_common.h -- lane moves and double-double arithmetic shared by the
// propagator marches (march.hip: implicit Euler, irk4.hip: two-stage Gauss).
#pragma once
#include <hip/hip_runtime.h>

namespace scft {

__device__ __forceinline__ void wg_barrier() { __syncthreads(); }

// Wait until every vector-memory operation of this wave has completed
// (s_waitcnt vmcnt(0); gfx9 encoding: vmcnt in bits 3:0 and 15:14, expcnt in
// 6:4 and lgkmcnt in 11:8 left at their maxima).  Issued once before a
// software-pipelined load loop, it leaves only the loop's own loads in
// flight.  The compiler's per-step waits then count exactly the loads newer
// than the one being consumed, instead of draining the queue every trip.
__device__ __forceinline__ void vmem_drain() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// DPP lane moves on a double (two 32-bit halves).  CTRL is a gfx9 DPP
// control: 0x110 + s = row_shr:s (lane p reads p - s of its 16-lane row),
// 0x100 + s = row_shl:s (reads p + s).  Lanes whose source leaves the row
// read 0 (bound_ctrl); every use multiplies them by a zero weight.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}
template <int S>
__device__ __forceinline__ double row_shr(double v) { return dpp_d<0x110 + S>(v); }
template <int S>
__device__ __forceinline__ double row_shl(double v) { return dpp_d<0x100 + S>(v); }

// GFX9 broadcast and wave-shift DPP moves on a double.  Rows outside
// ROWMASK are not written and read 0 (the `old` operand); bound_ctrl makes a
// lane without a source read 0.
template <int CTRL, int ROWMASK, bool BOUND>
__device__ __forceinline__ double dpp_rows_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, ROWMASK, 0xF, BOUND);
  const int hi =
      __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, ROWMASK, 0xF, BOUND);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}
// rows 1 and 3 read lane 15 of the row before (rows 0, 2: 0)
__device__ __forceinline__ double row_bcast15(double v) { return dpp_rows_d<0x142, 0xA, false>(v); }
// rows 2 and 3 read lane 31 (rows 0, 1: 0)
__device__ __forceinline__ double row_bcast31(double v) { return dpp_rows_d<0x143, 0xC, false>(v); }
// lane p reads lane p - 1 across the whole wave (lane 0: 0)
__device__ __forceinline__ double wave_shr1(double v) { return dpp_rows_d<0x138, 0xF, true>(v); }

// lane p reads lane p + 1 across the whole wave (lane 63: 0)
__device__ __forceinline__ double wave_shl1(double v) { return dpp_rows_d<0x130, 0xF, true>(v); }

// value of lane (byte_addr / 4) mod 64 through the LDS crossbar (ds_bpermute:
// no VALU issue, no LDS storage)
__device__ __forceinline__ double bpermute_d(int byte_addr, double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(byte_addr, static_cast<int>(b));
  const int hi = __builtin_amdgcn_ds_bpermute(byte_addr, static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// value of `v` in lane `l`, as a wave-uniform scalar
__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// Slab scratch as a raw buffer: a scalar resource (base, size) plus a
// wave-uniform byte offset and a per-lane byte offset, one instruction per
// access with no per-lane address arithmetic.  Accesses past `bytes` read 0
// and are dropped by the hardware.  aux = 2 sets the non-temporal hint: each
// slab element is written once and read once, steps apart.  The third
// descriptor word is the raw-buffer format word of the GFX9 family.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slab_rsrc(double* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(base, 0, static_cast<int>(bytes), 0x00020000);
}
__device__ __forceinline__ void slab_store(double v, __amdgpu_buffer_rsrc_t r, unsigned lane_off,
                                           unsigned uni_off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, lane_off, uni_off, 2);
}
__device__ __forceinline__ double slab_load(__amdgpu_buffer_rsrc_t r, unsigned lane_off,
                                            unsigned uni_off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, lane_off, uni_off, 2));
}

// One slab holds the K rows of each of the P lanes of a workgroup after one
// step.  For K >= 2 the rows go in pairs: rows 2i and 2i+1 of lane p sit at
// i * 2P + 2p and the double after it, so a wave moves 16 contiguous bytes
// per lane (1 KB per instruction).  For K = 1 it is one row of P.
//   element (row t, lane p):  K >= 2: (t / 2) * 2P + 2p + t % 2;  K = 1: p.
// lane_off = p * slab_lane_bytes<K>(); slab_off = byte offset of the slab.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <int K>
__device__ constexpr unsigned slab_lane_bytes() {
  return K >= 2 ? 16u : 8u;
}
template <int K>
__device__ __forceinline__ void slab_put(const double (&q)[K], __amdgpu_buffer_rsrc_t r,
                                         unsigned lane_off, unsigned slab_off, int P) {
  if constexpr (K >= 2) {
#pragma unroll
    for (int t = 0; t < K; t += 2) {
      const f64x2 v = {q[t], q[t + 1]};
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, lane_off,
                                             slab_off + static_cast<unsigned>(t * P * 8), 2);
    }
  } else {
    slab_store(q[0], r, lane_off, slab_off);
  }
}
template <int K>
__device__ __forceinline__ void slab_get(double (&d)[K], __amdgpu_buffer_rsrc_t r,
                                         unsigned lane_off, unsigned slab_off, int P) {
  if constexpr (K >= 2) {
#pragma unroll
    for (int t = 0; t < K; t += 2) {
      const f64x2 v = __builtin_bit_cast(
          f64x2, __builtin_amdgcn_raw_buffer_load_b128(
                     r, lane_off, slab_off + static_cast<unsigned>(t * P * 8), 2));
      d[t] = v.x;
      d[t + 1] = v.y;
    }
  } else {
    d[0] = slab_load(r, lane_off, slab_off);
  }
}

// ---- double-double arithmetic (hi + lo, |lo| <= ulp(hi)/2) for the pivot
// scan.  The pivots are a projective product of 2x2 matrices whose two
// eigenvalues nearly coincide when diffusion dominates (dt/h^2 >> 1); in FP64
// the lane-entry pivots come out ~1e-14 off, and D's conditioning turns that
// into ~1e-11 per step in q.  Carrying the scan in double-double (once per
// evaluation) brings the entry pivots back to FP64 accuracy, so the computed
// LU is as backward-stable as a sequential Thomas factorisation.  Every
// routine below turns FMA contraction off: fusing a product into a two-sum
// (s = fma(a, b, c), then s - a*b as another fma) breaks the error-free
// transformations, and the device compiler does exactly that by default.
struct dd {
  double h, l;
};
__device__ __forceinline__ dd dd_from(double a) { return {a, 0.0}; }
__device__ __forceinline__ dd two_sum(double a, double b) {
#pragma clang fp contract(off)
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
__device__ __forceinline__ dd quick_two_sum(double a, double b) {
#pragma clang fp contract(off)
  const double s = a + b;
  return {s, b - (s - a)};
}
__device__ __forceinline__ dd two_prod(double a, double b) {
#pragma clang fp contract(off)
  const double p = a * b;
  return {p, fma(a, b, -p)};
}
__device__ __forceinline__ dd dd_add(dd a, dd b) {
#pragma clang fp contract(off)
  dd s = two_sum(a.h, b.h);
  dd t = two_sum(a.l, b.l);
  s.l += t.h;
  s = quick_two_sum(s.h, s.l);
  s.l += t.l;
  return quick_two_sum(s.h, s.l);
}
__device__ __forceinline__ dd dd_mul(dd a, dd b) {
#pragma clang fp contract(off)
  dd p = two_prod(a.h, b.h);
  p.l += a.h * b.l + a.l * b.h;
  return quick_two_sum(p.h, p.l);
}
__device__ __forceinline__ dd dd_scale(dd a, double s) { return {a.h * s, a.l * s}; }  // s = 2^k
__device__ __forceinline__ double dd_div_to_double(dd a, dd b) {
#pragma clang fp contract(off)
  const double q1 = a.h / b.h;
  const dd r = dd_add(a, dd_mul(b, dd_from(-q1)));
  return q1 + r.h / b.h;
}

// 2x2 double-double product R = X * Y, rescaled by a power of two so the
// largest entry is in [1, 2): the pivot is a ratio of entries, so scaling is
// harmless and long products neither overflow nor underflow.
__device__ __forceinline__ void mat_mul_norm(const dd X[4], const dd Y[4], dd R[4]) {
#pragma clang fp contract(off)
  const dd r0 = dd_add(dd_mul(X[0], Y[0]), dd_mul(X[1], Y[2]));
  const dd r1 = dd_add(dd_mul(X[0], Y[1]), dd_mul(X[1], Y[3]));
  const dd r2 = dd_add(dd_mul(X[2], Y[0]), dd_mul(X[3], Y[2]));
  const dd r3 = dd_add(dd_mul(X[2], Y[1]), dd_mul(X[3], Y[3]));
  const double mx = fmax(fmax(fabs(r0.h), fabs(r1.h)), fmax(fabs(r2.h), fabs(r3.h)));
  const double sc = (mx > 0.0 && isfinite(mx)) ? ldexp(1.0, -ilogb(mx)) : 1.0;
  R[0] = dd_scale(r0, sc);
  R[1] = dd_scale(r1, sc);
  R[2] = dd_scale(r2, sc);
  R[3] = dd_scale(r3, sc);
}

__device__ __forceinline__ dd shfl_up_dd(dd v, int s) {
  return {__shfl_up(v.h, s), __shfl_up(v.l, s)};
}

}  // namespace scft
