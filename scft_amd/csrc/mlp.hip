// mlp.hip -- the warm-start MLP: dense FP32 layers (tanh on hidden layers,
// linear output), mean-squared-error loss and Adam, trained and evaluated on
// the device by the kernels below.  No reference counterpart: the reference
// has no learned warm start.  Parameters are packed as the FP64 restatement in
// oracle/oracle_mlp.c packs them: per layer W (dout x din, row-major), then
// b (dout).
//
// Every product is one kernel, `gemm_kernel`: C[m][n] = sum_k A(m,k) B(k,n)
// with arbitrary element strides for A and B, FP32 operands and FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x4f32), and a fused
// epilogue (bias + tanh for hidden layers, bias for the output layer, the
// tanh derivative for back-propagated deltas, or nothing).  Workgroups of 4
// waves own a 32 x 32 tile of C, each wave one 16 x 16 block; K advances in
// chunks of 64 staged in LDS, each operand loaded along whichever of its
// strides is 1.  The training products are 256 x 256 x 256 at batch 256:
// small tiles give 64 workgroups for them, and deep chunks four round trips
// to memory per tile instead of sixteen.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/scft.h"

namespace scft {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int TILE = 32;
constexpr int KC = 64;  // k chunk staged per iteration
enum Epilogue { EP_NONE = 0, EP_BIAS = 1, EP_BIAS_TANH = 2, EP_DTANH = 3 };

struct GemmArgs {
  const float* A;
  long long a_m, a_k;  // A(m, k) = A[m * a_m + k * a_k]
  const float* B;
  long long b_k, b_n;  // B(k, n) = B[k * b_k + n * b_n]
  float* C;            // C[m * ldc + n]
  long long ldc;
  int M, N, K;
  int epilogue;
  const float* bias;   // EP_BIAS*: bias[n]
  const float* act;    // EP_DTANH: C *= 1 - act[m * ldc + n]^2
};

__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
  __shared__ float As[TILE][KC + 1];
  __shared__ float Bs[KC][TILE + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * TILE, n0 = blockIdx.x * TILE;
  const int wm = (wave >> 1) * 16, wn = (wave & 1) * 16;  // the wave's block
  // consecutive threads walk the operand's contiguous index
  const bool a_kfast = g.a_k == 1, b_nfast = g.b_n == 1;
  floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < g.K; k0 += KC) {
    // stage A[m0:m0+32, k0:k0+64] and B[k0:k0+64, n0:n0+32], zero outside
#pragma unroll
    for (int e = tid; e < TILE * KC; e += 256) {
      const int mm = a_kfast ? e / KC : e % TILE;
      const int kk = a_kfast ? e % KC : e / TILE;
      const int m = m0 + mm, k = k0 + kk;
      As[mm][kk] = (m < g.M && k < g.K) ? g.A[m * g.a_m + k * g.a_k] : 0.f;
    }
#pragma unroll
    for (int e = tid; e < TILE * KC; e += 256) {
      const int kk = b_nfast ? e / TILE : e % KC;
      const int nn = b_nfast ? e % TILE : e / KC;
      const int k = k0 + kk, n = n0 + nn;
      Bs[kk][nn] = (n < g.N && k < g.K) ? g.B[k * g.b_k + n * g.b_n] : 0.f;
    }
    __syncthreads();
    // 16 x 16 x 4 MFMA operands: lane l supplies A(row l % 16, k l / 16) and
    // B(k l / 16, col l % 16); the result's lane l holds rows
    // 4 (l / 16) + 0..3 of column l % 16.
    const int r = lane & 15, q = lane >> 4;
#pragma unroll
    for (int ks = 0; ks < KC; ks += 4)
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(As[wm + r][ks + q], Bs[ks + q][wn + r], acc,
                                                 0, 0, 0);
    __syncthreads();
  }

  const int col = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int m = m0 + wm + rq + e;
    const int n = n0 + wn + col;
    if (m >= g.M || n >= g.N) continue;
    float v = acc[e];
    if (g.epilogue == EP_BIAS || g.epilogue == EP_BIAS_TANH) v += g.bias[n];
    if (g.epilogue == EP_BIAS_TANH) v = tanhf(v);
    if (g.epilogue == EP_DTANH) {
      const float a = g.act[m * g.ldc + n];
      v *= 1.f - a * a;
    }
    g.C[m * g.ldc + n] = v;
  }
}

// gb[o] = sum_b delta[b][o]: a workgroup takes 64 columns, its 4 waves
// every 4th row each (lanes on consecutive columns, loads coalesced), then
// the 4 partial sums meet in LDS
__global__ void __launch_bounds__(256) colsum_kernel(const float* __restrict__ delta, int batch,
                                                     int dout, float* __restrict__ gb) {
  __shared__ float part[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int o = blockIdx.x * 64 + lane;
  float s = 0.f;
  if (o < dout)
    for (int b = wave; b < batch; b += 4) s += delta[(long long)b * dout + o];
  part[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && o < dout) gb[o] = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// delta = 2 (y - t) / (batch * dout); loss partial sums into *loss (FP64)
__global__ void mse_kernel(const float* __restrict__ y, const float* __restrict__ t, long long n,
                           float scale, float* __restrict__ delta, double* __restrict__ loss) {
  __shared__ double red[4];
  double s = 0.0;
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const float e = y[k] - t[k];
    delta[k] = scale * e;
    s += (double)e * e;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int v = 0; v < (int)(blockDim.x >> 6); v++) tot += red[v];
    atomicAdd(loss, tot);
  }
}

// Adam (Kingma & Ba 2015, Algorithm 1) with bias corrections c1, c2
__global__ void adam_kernel(long long n, float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
  }
}

// rows idx[0..batch) of src (row length len) into dst
__global__ void gather_kernel(const float* __restrict__ src, const int* __restrict__ idx,
                              int batch, int len, float* __restrict__ dst) {
  const long long n = (long long)batch * len;
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const long long b = k / len, c = k % len;
    dst[k] = src[(long long)idx[b] * len + c];
  }
}

}  // namespace
}  // namespace scft

// ------------------------------------------------------------------ host
using namespace scft;

struct scft_mlp {
  std::vector<int> dims;           // dims[0] inputs ... dims[L] outputs
  std::vector<long long> off;      // parameter offset of each layer
  long long n_params = 0;
  int cap_batch = 0;               // activation buffers sized for this batch
  float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
  std::vector<float*> d_act;       // L + 1 activation buffers [cap_batch x dims[l]]
  float *d_delta0 = nullptr, *d_delta1 = nullptr, *d_t = nullptr;
  double* d_loss = nullptr;
  int adam_t = 0;
  hipStream_t stream = nullptr;
  // input / output normalisation: x' = (x - xm) / xs, y = y' * ys + ym
  std::vector<float> xm, xs, ym, ys;
};

namespace {
#define MLPCHK(expr)                                                         \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      fprintf(stderr, "scft mlp: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
      return SCFT_E_HIP;                                                     \
    }                                                                        \
  } while (0)

int mlp_dalloc(float** p, long long n) {
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(float) * (n > 0 ? n : 1)) == hipSuccess
             ? SCFT_OK
             : SCFT_E_NOMEM;
}

int ensure_batch(scft_mlp* m, int batch) {
  if (batch <= m->cap_batch) return SCFT_OK;
  for (auto p : m->d_act) (void)hipFree(p);
  (void)hipFree(m->d_delta0);
  (void)hipFree(m->d_delta1);
  (void)hipFree(m->d_t);
  m->d_act.assign(m->dims.size(), nullptr);
  int widest = 0;
  for (size_t l = 0; l < m->dims.size(); l++) {
    if (mlp_dalloc(&m->d_act[l], (long long)batch * m->dims[l])) return SCFT_E_NOMEM;
    widest = std::max(widest, m->dims[l]);
  }
  if (mlp_dalloc(&m->d_delta0, (long long)batch * widest) ||
      mlp_dalloc(&m->d_delta1, (long long)batch * widest) ||
      mlp_dalloc(&m->d_t, (long long)batch * m->dims.back()))
    return SCFT_E_NOMEM;
  m->cap_batch = batch;
  return SCFT_OK;
}

hipError_t gemm(hipStream_t s, const float* A, long long a_m, long long a_k, const float* B,
                long long b_k, long long b_n, float* C, long long ldc, int M, int N, int K,
                int ep, const float* bias, const float* act) {
  GemmArgs g{A, a_m, a_k, B, b_k, b_n, C, ldc, M, N, K, ep, bias, act};
  dim3 grid((N + TILE - 1) / TILE, (M + TILE - 1) / TILE);
  hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, s, g);
  return hipGetLastError();
}

// forward of `batch` rows already in d_act[0]; activations kept for backward
int forward(scft_mlp* m, int batch) {
  const int L = static_cast<int>(m->dims.size()) - 1;
  for (int l = 0; l < L; l++) {
    const int din = m->dims[l], dout = m->dims[l + 1];
    const float* W = m->d_p + m->off[l];
    const float* b = W + (long long)dout * din;
    MLPCHK(gemm(m->stream, m->d_act[l], din, 1, W, 1, din, m->d_act[l + 1], dout, batch, dout,
                din, l < L - 1 ? EP_BIAS_TANH : EP_BIAS, b, nullptr));
  }
  return SCFT_OK;
}

// loss and gradient for the rows in d_act[0] against targets in d_t; the
// squared error is added to *d_loss, which is cleared first unless the
// caller accumulates over several batches
int backward(scft_mlp* m, int batch, bool clear_loss = true) {
  const int L = static_cast<int>(m->dims.size()) - 1;
  const int dl = m->dims[L];
  const long long n = (long long)batch * dl;
  if (clear_loss) MLPCHK(hipMemsetAsync(m->d_loss, 0, sizeof(double), m->stream));
  hipLaunchKernelGGL(mse_kernel, dim3(64), dim3(256), 0, m->stream, m->d_act[L], m->d_t, n,
                     2.f / static_cast<float>(n), m->d_delta0, m->d_loss);
  MLPCHK(hipGetLastError());
  float* delta = m->d_delta0;
  float* other = m->d_delta1;
  for (int l = L - 1; l >= 0; l--) {
    const int din = m->dims[l], dout = m->dims[l + 1];
    const float* W = m->d_p + m->off[l];
    float* gW = m->d_g + m->off[l];
    float* gb = gW + (long long)dout * din;
    // gW[o][i] = sum_b delta[b][o] act_l[b][i]
    MLPCHK(gemm(m->stream, delta, 1, dout, m->d_act[l], din, 1, gW, din, dout, din, batch,
                EP_NONE, nullptr, nullptr));
    hipLaunchKernelGGL(colsum_kernel, dim3((dout + 63) / 64), dim3(256), 0, m->stream, delta,
                       batch, dout, gb);
    MLPCHK(hipGetLastError());
    if (l > 0) {  // delta_prev[b][i] = sum_o delta[b][o] W[o][i] * (1 - act_l[b][i]^2)
      MLPCHK(gemm(m->stream, delta, dout, 1, W, din, 1, other, din, batch, din, dout, EP_DTANH,
                  nullptr, m->d_act[l]));
      std::swap(delta, other);
    }
  }
  return SCFT_OK;
}

void normalise_rows(const scft_mlp* m, const float* X, int rows, std::vector<float>& out) {
  const int din = m->dims[0];
  out.resize((size_t)rows * din);
  for (int r = 0; r < rows; r++)
    for (int i = 0; i < din; i++) {
      const float v = X[(size_t)r * din + i];
      out[(size_t)r * din + i] = m->xm.empty() ? v : (v - m->xm[i]) / m->xs[i];
    }
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
}  // namespace

extern "C" int scft_mlp_create(scft_mlp** out, int n_layers, const int* dims,
                               unsigned long long seed) {
  if (!out || n_layers < 1 || !dims) return SCFT_E_ARG;
  for (int l = 0; l <= n_layers; l++)
    if (dims[l] < 1 || dims[l] > 4096) return SCFT_E_ARG;
  scft_mlp* m = new scft_mlp;
  m->dims.assign(dims, dims + n_layers + 1);
  m->off.resize(n_layers + 1);
  long long o = 0;
  for (int l = 0; l < n_layers; l++) {
    m->off[l] = o;
    o += (long long)dims[l + 1] * dims[l] + dims[l + 1];
  }
  m->off[n_layers] = o;
  m->n_params = o;
  // Glorot-uniform weights, zero biases, from splitmix64(seed)
  std::vector<float> p(o, 0.f);
  uint64_t s = seed;
  for (int l = 0; l < n_layers; l++) {
    const double lim = std::sqrt(6.0 / (dims[l] + dims[l + 1]));
    for (long long k = 0; k < (long long)dims[l + 1] * dims[l]; k++) {
      const double u = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
      p[m->off[l] + k] = static_cast<float>((2.0 * u - 1.0) * lim);
    }
  }
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      mlp_dalloc(&m->d_p, o) || mlp_dalloc(&m->d_g, o) || mlp_dalloc(&m->d_m, o) ||
      mlp_dalloc(&m->d_v, o) ||
      hipMalloc(reinterpret_cast<void**>(&m->d_loss), sizeof(double)) != hipSuccess) {
    delete m;
    return SCFT_E_NOMEM;
  }
  if (hipMemcpy(m->d_p, p.data(), sizeof(float) * o, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->d_m, 0, sizeof(float) * o) != hipSuccess ||
      hipMemset(m->d_v, 0, sizeof(float) * o) != hipSuccess) {
    delete m;
    return SCFT_E_HIP;
  }
  *out = m;
  return SCFT_OK;
}

extern "C" int scft_mlp_destroy(scft_mlp* m) {
  if (!m) return SCFT_OK;
  (void)hipStreamSynchronize(m->stream);
  for (auto p : m->d_act) (void)hipFree(p);
  (void)hipFree(m->d_delta0);
  (void)hipFree(m->d_delta1);
  (void)hipFree(m->d_t);
  (void)hipFree(m->d_p);
  (void)hipFree(m->d_g);
  (void)hipFree(m->d_m);
  (void)hipFree(m->d_v);
  (void)hipFree(m->d_loss);
  (void)hipStreamDestroy(m->stream);
  delete m;
  return SCFT_OK;
}

extern "C" int scft_mlp_param_count(const scft_mlp* m, long long* count) {
  if (!m || !count) return SCFT_E_ARG;
  *count = m->n_params;
  return SCFT_OK;
}

extern "C" int scft_mlp_get_params(const scft_mlp* m, float* out) {
  if (!m || !out) return SCFT_E_ARG;
  MLPCHK(hipStreamSynchronize(m->stream));
  MLPCHK(hipMemcpy(out, m->d_p, sizeof(float) * m->n_params, hipMemcpyDeviceToHost));
  return SCFT_OK;
}

extern "C" int scft_mlp_set_params(scft_mlp* m, const float* in) {
  if (!m || !in) return SCFT_E_ARG;
  MLPCHK(hipStreamSynchronize(m->stream));
  MLPCHK(hipMemcpy(m->d_p, in, sizeof(float) * m->n_params, hipMemcpyHostToDevice));
  MLPCHK(hipMemset(m->d_m, 0, sizeof(float) * m->n_params));
  MLPCHK(hipMemset(m->d_v, 0, sizeof(float) * m->n_params));
  m->adam_t = 0;
  return SCFT_OK;
}

extern "C" int scft_mlp_set_normalisation(scft_mlp* m, const float* x_mean, const float* x_std,
                                          const float* y_mean, const float* y_std) {
  if (!m) return SCFT_E_ARG;
  const int din = m->dims.front(), dout = m->dims.back();
  if (!x_mean || !x_std || !y_mean || !y_std) {
    m->xm.clear();
    m->xs.clear();
    m->ym.clear();
    m->ys.clear();
    return SCFT_OK;
  }
  for (int i = 0; i < din; i++)
    if (!(x_std[i] > 0.f)) return SCFT_E_ARG;
  for (int i = 0; i < dout; i++)
    if (!(y_std[i] > 0.f)) return SCFT_E_ARG;
  m->xm.assign(x_mean, x_mean + din);
  m->xs.assign(x_std, x_std + din);
  m->ym.assign(y_mean, y_mean + dout);
  m->ys.assign(y_std, y_std + dout);
  return SCFT_OK;
}

extern "C" int scft_mlp_forward(scft_mlp* m, int batch, const float* X, float* Y) {
  if (!m || batch < 1 || !X || !Y) return SCFT_E_ARG;
  if (int rc = ensure_batch(m, batch)) return rc;
  std::vector<float> xn;
  normalise_rows(m, X, batch, xn);
  MLPCHK(hipMemcpyAsync(m->d_act[0], xn.data(), sizeof(float) * xn.size(),
                        hipMemcpyHostToDevice, m->stream));
  if (int rc = forward(m, batch)) return rc;
  const int dout = m->dims.back();
  MLPCHK(hipMemcpyAsync(Y, m->d_act.back(), sizeof(float) * (size_t)batch * dout,
                        hipMemcpyDeviceToHost, m->stream));
  MLPCHK(hipStreamSynchronize(m->stream));
  if (!m->ym.empty())
    for (int r = 0; r < batch; r++)
      for (int o = 0; o < dout; o++) Y[(size_t)r * dout + o] = Y[(size_t)r * dout + o] *
                                                                    m->ys[o] + m->ym[o];
  return SCFT_OK;
}

// Forward pass for the device-side warm-start prediction (scft_internal.h).
int scft_mlp_forward_resident(scft_mlp* m, int batch, const float* d_xn, const float** d_out,
                              const float** ym, const float** ys, int* n_out, int* n_in) {
  if (!m || batch < 1 || !d_xn || !d_out) return SCFT_E_ARG;
  if (int rc = ensure_batch(m, batch)) return rc;
  const int din = m->dims.front();
  MLPCHK(hipMemcpyAsync(m->d_act[0], d_xn, sizeof(float) * (size_t)batch * din,
                        hipMemcpyDeviceToDevice, m->stream));
  if (int rc = forward(m, batch)) return rc;
  MLPCHK(hipStreamSynchronize(m->stream));
  *d_out = m->d_act.back();
  if (ym) *ym = m->ym.empty() ? nullptr : m->ym.data();
  if (ys) *ys = m->ys.empty() ? nullptr : m->ys.data();
  if (n_out) *n_out = m->dims.back();
  if (n_in) *n_in = din;
  return SCFT_OK;
}

// x' = (x - xm) / xs of `rows` host rows, as scft_mlp_forward applies it.
int scft_mlp_normalise_inputs(const scft_mlp* m, const float* X, int rows, float* out) {
  if (!m || !X || !out || rows < 0) return SCFT_E_ARG;
  std::vector<float> xn;
  normalise_rows(m, X, rows, xn);
  std::copy(xn.begin(), xn.end(), out);
  return SCFT_OK;
}

// Loss and gradient on raw (already normalised) rows, for gradient checks:
// no normalisation is applied here.
extern "C" int scft_mlp_loss_grad(scft_mlp* m, int batch, const float* X, const float* T,
                                  double* loss, float* grad) {
  if (!m || batch < 1 || !X || !T) return SCFT_E_ARG;
  if (int rc = ensure_batch(m, batch)) return rc;
  const int din = m->dims.front(), dout = m->dims.back();
  MLPCHK(hipMemcpyAsync(m->d_act[0], X, sizeof(float) * (size_t)batch * din,
                        hipMemcpyHostToDevice, m->stream));
  MLPCHK(hipMemcpyAsync(m->d_t, T, sizeof(float) * (size_t)batch * dout, hipMemcpyHostToDevice,
                        m->stream));
  if (int rc = forward(m, batch)) return rc;
  if (int rc = backward(m, batch)) return rc;
  double l = 0.0;
  MLPCHK(hipMemcpyAsync(&l, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (grad)
    MLPCHK(hipMemcpyAsync(grad, m->d_g, sizeof(float) * m->n_params, hipMemcpyDeviceToHost,
                          m->stream));
  MLPCHK(hipStreamSynchronize(m->stream));
  if (loss) *loss = l / ((double)batch * dout);
  return SCFT_OK;
}

// One Adam step on the gradient currently held (after scft_mlp_loss_grad).
extern "C" int scft_mlp_adam_step(scft_mlp* m, float lr) {
  if (!m) return SCFT_E_ARG;
  const float b1 = 0.9f, b2 = 0.999f;
  m->adam_t++;
  const float c1 = 1.f - std::pow(b1, (float)m->adam_t);
  const float c2 = 1.f - std::pow(b2, (float)m->adam_t);
  hipLaunchKernelGGL(adam_kernel, dim3(256), dim3(256), 0, m->stream, m->n_params, m->d_p,
                     m->d_g, m->d_m, m->d_v, lr, b1, b2, 1e-8f, c1, c2);
  MLPCHK(hipGetLastError());
  MLPCHK(hipStreamSynchronize(m->stream));
  return SCFT_OK;
}

// Mini-batch Adam over n_samples rows (normalised with the object's
// statistics when set), reshuffled every epoch from splitmix64(seed); the
// last partial batch of an epoch is used as is.  epoch_loss[e] (nullable)
// receives the mean loss of epoch e in normalised units.
extern "C" int scft_mlp_train(scft_mlp* m, int n_samples, const float* X, const float* Y,
                              int epochs, int batch, float lr, unsigned long long seed,
                              double* epoch_loss) {
  if (!m || n_samples < 1 || !X || !Y || epochs < 1 || batch < 1 || !(lr > 0.f))
    return SCFT_E_ARG;
  const int din = m->dims.front(), dout = m->dims.back();
  if (int rc = ensure_batch(m, batch)) return rc;
  std::vector<float> xn, yn((size_t)n_samples * dout);
  normalise_rows(m, X, n_samples, xn);
  for (int r = 0; r < n_samples; r++)
    for (int o = 0; o < dout; o++) {
      const float v = Y[(size_t)r * dout + o];
      yn[(size_t)r * dout + o] = m->ym.empty() ? v : (v - m->ym[o]) / m->ys[o];
    }
  float *dX = nullptr, *dY = nullptr;
  int* dIdx = nullptr;
  if (mlp_dalloc(&dX, (long long)n_samples * din) || mlp_dalloc(&dY, (long long)n_samples * dout) ||
      hipMalloc(reinterpret_cast<void**>(&dIdx), sizeof(int) * n_samples) != hipSuccess) {
    (void)hipFree(dX);
    (void)hipFree(dY);
    return SCFT_E_NOMEM;
  }
  int rc = SCFT_OK;
  std::vector<int> idx(n_samples);
  for (int i = 0; i < n_samples; i++) idx[i] = i;
  uint64_t s = seed;
  const float b1 = 0.9f, b2 = 0.999f;
  if (hipMemcpy(dX, xn.data(), sizeof(float) * xn.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dY, yn.data(), sizeof(float) * yn.size(), hipMemcpyHostToDevice) != hipSuccess)
    rc = SCFT_E_HIP;
  for (int e = 0; e < epochs && rc == SCFT_OK; e++) {
    for (int i = n_samples - 1; i > 0; i--) {  // Fisher-Yates
      const int j = static_cast<int>(splitmix(s) % static_cast<uint64_t>(i + 1));
      std::swap(idx[i], idx[j]);
    }
    if (hipMemcpyAsync(dIdx, idx.data(), sizeof(int) * n_samples, hipMemcpyHostToDevice,
                       m->stream) != hipSuccess) {
      rc = SCFT_E_HIP;
      break;
    }
    // the epoch's squared error accumulates on the device and is read once
    // at its end: the steps queue without a host round trip between them
    if (hipMemsetAsync(m->d_loss, 0, sizeof(double), m->stream) != hipSuccess) {
      rc = SCFT_E_HIP;
      break;
    }
    for (int b0 = 0; b0 < n_samples && rc == SCFT_OK; b0 += batch) {
      const int nb = std::min(batch, n_samples - b0);
      hipLaunchKernelGGL(gather_kernel, dim3(128), dim3(256), 0, m->stream, dX, dIdx + b0, nb,
                         din, m->d_act[0]);
      hipLaunchKernelGGL(gather_kernel, dim3(128), dim3(256), 0, m->stream, dY, dIdx + b0, nb,
                         dout, m->d_t);
      if (hipGetLastError() != hipSuccess || forward(m, nb) || backward(m, nb, false)) {
        rc = SCFT_E_HIP;
        break;
      }
      m->adam_t++;
      const float c1 = 1.f - std::pow(b1, (float)m->adam_t);
      const float c2 = 1.f - std::pow(b2, (float)m->adam_t);
      hipLaunchKernelGGL(adam_kernel, dim3(256), dim3(256), 0, m->stream, m->n_params, m->d_p,
                         m->d_g, m->d_m, m->d_v, lr, b1, b2, 1e-8f, c1, c2);
      if (hipGetLastError() != hipSuccess) rc = SCFT_E_HIP;
    }
    // also orders the next epoch's host shuffle after this epoch's index copy
    double esum = 0.0;
    if (rc == SCFT_OK &&
        (hipMemcpyAsync(&esum, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, m->stream) !=
             hipSuccess ||
         hipStreamSynchronize(m->stream) != hipSuccess))
      rc = SCFT_E_HIP;
    if (epoch_loss && rc == SCFT_OK)
      epoch_loss[e] = esum / (static_cast<double>(n_samples) * dout);
  }
  (void)hipStreamSynchronize(m->stream);
  (void)hipFree(dX);
  (void)hipFree(dY);
  (void)hipFree(dIdx);
  return rc;
}
