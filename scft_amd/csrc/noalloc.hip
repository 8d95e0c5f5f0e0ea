// noalloc.hip -- one field evaluation that allocates nothing, for callers
// that cannot call hipMalloc (code under stream capture, tight serving loops,
// harnesses that forbid device allocation).  No counterpart in the reference.
//
// It is scft_batch_eval for one instance (the same march kernel and variant
// as a one-instance batch, so the same bits), with the caller's device
// buffers for inputs and outputs and the library's working memory in the
// file-scope device array `scratch` below:
//   [0, 256)            the instance descriptor
//   [256, 260)          the bucket's item list (instance 0)
//   [512, 516)          the work-queue counter
//   [1024, ...)         the N/2 + 1 pairing weights (N <= 16384: < 64 KiB)
//   [kHeadBytes, ...)   the q-slab region of the one persistent workgroup,
//                       (N/2) * cap doubles (cap = 64 W K >= n)
// The array's device address differs per GPU; it is looked up once per
// device (hipGetSymbolAddress) and only that address is passed to kernels
// and copies.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/scft.h"
#include "host_math.h"
#include "scft_internal.h"

using namespace scft;

namespace {

constexpr size_t kHeadBytes = size_t(256) << 10;   // descriptor, items, counter, weights
constexpr size_t kSlabBytes = size_t(64) << 20;    // q-slabs
constexpr size_t kScratchBytes = kHeadBytes + kSlabBytes;
constexpr size_t kDescOff = 0, kItemsOff = 256, kQueueOff = 512, kWeightsOff = 1024;
constexpr int kMaxDevices = 64;

static_assert(sizeof(InstDesc) <= kItemsOff, "descriptor slot");
static_assert(kWeightsOff + (16384 / 2 + 1) * sizeof(double) <= kHeadBytes, "weights slot");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;                         // one call at a time in this process
void* g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch` per GPU
alignas(256) unsigned char g_staging[kWeightsOff + (16384 / 2 + 1) * sizeof(double)];
thread_local char g_msg[256];

int fail(int code, const char* what, hipError_t e = hipSuccess) {
  if (e != hipSuccess)
    snprintf(g_msg, sizeof g_msg, "scft_eval_noalloc: %s: %s", what, hipGetErrorString(e));
  else
    snprintf(g_msg, sizeof g_msg, "scft_eval_noalloc: %s", what);
  return code;
}

// the device address of `scratch` on the current device, looked up on the
// first call there (caller holds g_mu)
int scratch_address(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(SCFT_E_HIP, "hipGetDevice", e);
  if (dev < 0 || dev >= kMaxDevices) return fail(SCFT_E_HIP, "device ordinal out of range");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(SCFT_E_HIP, "hipGetSymbolAddress(scratch)", e);
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return SCFT_OK;
}

}  // namespace

extern "C" const char* scft_noalloc_error(void) { return g_msg; }

extern "C" int scft_eval_noalloc(int M, int N, int bc, int cmat, const double* d_x,
                                 const double* d_w, const double* d_phi_t, double* d_phi,
                                 double* d_r, double* d_rnorm, unsigned* d_status,
                                 void* stream) {
  g_msg[0] = 0;
  if (bc != SCFT_BC_DD && bc != SCFT_BC_DN) return fail(SCFT_E_ARG, "bc");
  if (cmat != SCFT_C_ROWSCALED && cmat != SCFT_C_CONSISTENT) return fail(SCFT_E_ARG, "cmat");
  if (!d_x || !d_w || !d_phi_t || !d_r || !d_rnorm || !d_status)
    return fail(SCFT_E_ARG, "null device buffer");
  const int n = bc == SCFT_BC_DD ? M - 2 : M - 1;
  if (M < 3 || n < 1 || n > 4096) return fail(SCFT_E_ARG, "M out of range (1..4096 unknowns)");
  if (N < 16 || N > 16384 || (N & (N - 1)) != 0)
    return fail(SCFT_E_ARG, "N must be a power of two in [16, 16384]");
  const int cap = bucket_capacity(n);
  const int K = bucket_default_K(cap, false), W = cap / 64 / K;
  const long long slab_doubles = static_cast<long long>(N / 2) * cap;
  if (static_cast<size_t>(slab_doubles) * sizeof(double) > kSlabBytes) {
    snprintf(g_msg, sizeof g_msg,
             "scft_eval_noalloc: n=%d (bucket %d) with N=%d needs %lld bytes of slabs, more "
             "than the %zu of the fixed scratch", n, cap, N,
             slab_doubles * static_cast<long long>(sizeof(double)), kSlabBytes);
    return SCFT_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  unsigned char* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);

  // descriptor, item list, zeroed counter and pairing weights, staged on the
  // host and copied in one transfer
  std::memset(g_staging, 0, kWeightsOff);
  InstDesc d;
  std::memset(&d, 0, sizeof d);
  d.n = n;
  d.N = N;
  d.M = M;
  d.bc = bc;
  d.cmat = cmat;
  std::memcpy(g_staging + kDescOff, &d, sizeof d);
  const int item = 0;
  std::memcpy(g_staging + kItemsOff, &item, sizeof item);
  if (!pairing_weights_into(N, reinterpret_cast<double*>(g_staging + kWeightsOff)))
    return fail(SCFT_E_ARG, "pairing weights");
  const size_t bytes = kWeightsOff + (N / 2 + 1) * sizeof(double);
  hipError_t e = hipMemcpyAsync(base, g_staging, bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return fail(SCFT_E_HIP, "hipMemcpyAsync", e);
  // the caller's status word starts at zero (a fill kernel, not a memset)
  e = launch_reset(reinterpret_cast<int*>(d_status), 1, s);
  if (e != hipSuccess) return fail(SCFT_E_HIP, "status reset", e);

  MarchArgs a;
  std::memset(&a, 0, sizeof a);
  a.desc = reinterpret_cast<const InstDesc*>(base + kDescOff);
  a.x = d_x;
  a.w = d_w;
  a.phit = d_phi_t;
  a.pw = reinterpret_cast<const double*>(base + kWeightsOff);
  a.phi = d_phi;
  a.r = d_r;
  a.rnorm = d_rnorm;
  a.status = d_status;
  a.slabs = reinterpret_cast<double*>(base + kHeadBytes);
  a.slab_stride = slab_doubles;
  a.items = reinterpret_cast<const int*>(base + kItemsOff);
  a.queue = reinterpret_cast<int*>(base + kQueueOff);
  a.n_items = 1;
  a.half_max = N / 2;
  e = launch_march(K, W, a, 1, s);
  if (e != hipSuccess) return fail(SCFT_E_HIP, "march launch", e);
  // the host staging buffer is reused by the next call: wait for this one
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(SCFT_E_HIP, "hipStreamSynchronize", e);
  return SCFT_OK;
}
