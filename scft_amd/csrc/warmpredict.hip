// warmpredict.hip -- the learned warm start's last step on the device: the
// network's outputs mapped onto every instance's unknowns and written into
// the batch's resident field, so no field crosses PCIe between the
// prediction and the solve.  No counterpart in the reference.
//
// Per instance (one 256-thread workgroup): the outputs y' ys + ym (FP32, as
// the network produces them) become FP64 values; the grid part is
// interpolated by the not-a-knot cubic spline of DEALII_SCFT/src/spline_chen.c
// (the same tridiagonal form and evaluation as host_math.cpp's
// spline_not_a_knot / spline_eval, one thread solving the <= 256-point system
// in LDS), the leading wall unknowns by the samples asinh(w / s) interpolated
// linearly in the node number.
#include "scft_internal.h"

namespace scft {

namespace {

constexpr int WM_THREADS = 256;

__device__ double spline_value(int n, const double* gx, const double* gy, const double* M,
                               double xp) {
  int klo = 0, khi = n - 1;
  while (khi - klo > 1) {
    const int k = (khi + klo) >> 1;
    if (gx[k] > xp)
      khi = k;
    else
      klo = k;
  }
  const double h = gx[khi] - gx[klo];
  const double a = (gx[khi] - xp) / h, b = (xp - gx[klo]) / h;
  return a * gy[klo] + b * gy[khi] +
         ((a * a * a - a) * M[klo] + (b * b * b - b) * M[khi]) * (h * h) / 6.0;
}

__global__ void __launch_bounds__(WM_THREADS) warm_map_kernel(WarmMapArgs a) {
  const int row = blockIdx.x;  // row of the plan; the batch instance it writes:
  const int inst = a.sel ? a.sel[row] : row;
  const int tid = threadIdx.x;
  __shared__ double gx[WARM_MAX_GRID], gy[WARM_MAX_GRID], M[WARM_MAX_GRID];
  __shared__ double lo[WARM_MAX_GRID], di[WARM_MAX_GRID], up[WARM_MAX_GRID],
      rhs[WARM_MAX_GRID];
  __shared__ double sv[WARM_MAX_NODES];
  const int ng = a.n_grid;
  const float* yr = a.y + static_cast<long long>(row) * a.n_out;
  for (int i = tid; i < ng; i += WM_THREADS) {
    gx[i] = a.grid[static_cast<long long>(row) * ng + i];
    const float v = a.ys ? yr[i] * a.ys[i] + a.ym[i] : yr[i];
    gy[i] = static_cast<double>(v);
  }
  for (int j = tid; j < a.n_node; j += WM_THREADS) {
    const int o = ng + j;
    const float v = a.ys ? yr[o] * a.ys[o] + a.ym[o] : yr[o];
    sv[j] = static_cast<double>(v);
  }
  __syncthreads();
  if (tid == 0) {  // second derivatives, not-a-knot ends folded into the end rows
    const int n = ng, k = n - 2;
    for (int i = 1; i <= n - 2; i++) {
      const double hl = gx[i] - gx[i - 1], hr = gx[i + 1] - gx[i];
      lo[i - 1] = hl / 6.0;
      di[i - 1] = (hl + hr) / 3.0;
      up[i - 1] = hr / 6.0;
      rhs[i - 1] = (gy[i + 1] - gy[i]) / hr - (gy[i] - gy[i - 1]) / hl;
    }
    const double h0 = gx[1] - gx[0], h1 = gx[2] - gx[1];
    const double ha = gx[n - 2] - gx[n - 3], hb = gx[n - 1] - gx[n - 2];
    di[0] += lo[0] * (1.0 + h0 / h1);
    if (k > 1) up[0] -= lo[0] * (h0 / h1);
    lo[0] = 0.0;
    di[k - 1] += up[k - 1] * (1.0 + hb / ha);
    if (k > 1) lo[k - 1] -= up[k - 1] * (hb / ha);
    up[k - 1] = 0.0;
    for (int i = 1; i < k; i++) {
      const double wgt = lo[i] / di[i - 1];
      di[i] -= wgt * up[i - 1];
      rhs[i] -= wgt * rhs[i - 1];
    }
    M[k] = rhs[k - 1] / di[k - 1];
    for (int i = k - 2; i >= 0; i--) M[i + 1] = (rhs[i] - up[i] * M[i + 2]) / di[i];
    M[0] = (1.0 + h0 / h1) * M[1] - (h0 / h1) * M[2];
    M[n - 1] = (1.0 + hb / ha) * M[n - 2] - (hb / ha) * M[n - 3];
  }
  __syncthreads();
  const InstDesc d = a.desc[inst];
  const double* xn = a.x + d.xo;
  double* w = a.w + d.uo;
  const int nwall = a.wall_count[row];
  for (int u = tid; u < d.n; u += WM_THREADS) {
    const int node = u + 1;
    double v;
    if (u < nwall) {
      // bracket of the node number among the samples (clamped to the ends)
      int j = 0;
      while (j + 2 < a.n_node && a.node_idx[j + 1] <= node) j++;
      const double t = static_cast<double>(node - a.node_idx[j]) /
                       static_cast<double>(a.node_idx[j + 1] - a.node_idx[j]);
      const double tc = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
      v = a.node_scale * sinh((1.0 - tc) * sv[j] + tc * sv[j + 1]);
    } else {
      v = spline_value(ng, gx, gy, M, xn[node]);
    }
    w[u] = v;
  }
}

}  // namespace

hipError_t launch_warm_map(const WarmMapArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(warm_map_kernel, dim3(a.n_inst), dim3(WM_THREADS), 0, s, a);
  return hipGetLastError();
}

}  // namespace scft
