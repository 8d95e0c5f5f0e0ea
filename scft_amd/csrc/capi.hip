// capi.hip -- host side of libscft_hip.so: batched contexts, the size-bucket
// scheduler and autotuner, the graph-captured solve loop, and the reference
// drop-in entry points.  Everything exported here is declared in scft.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scft.h"
#include "host_math.h"
#include "scft_internal.h"

using namespace scft;

// ------------------------------------------------------------ error plumbing
static thread_local char g_errmsg[512];

#define HIPCHK(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      snprintf(g_errmsg, sizeof g_errmsg, "%s: %s (%s:%d)", #expr,            \
               hipGetErrorString(e_), __FILE__, __LINE__);                    \
      return e_ == hipErrorOutOfMemory ? SCFT_E_NOMEM : SCFT_E_HIP;           \
    }                                                                         \
  } while (0)

#define SCFT_TRY(expr)        \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != SCFT_OK) return rc_; \
  } while (0)

template <typename T>
static int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  return SCFT_OK;
}

template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// ------------------------------------------------------------------ buckets
namespace {

struct Bucket {
  int cap = 0;  // 64 * W * K
  int K = 1, W = 1;
  int nmax = 0, Nmax = 0;
  std::vector<int> items;  // instance ids, largest cost first
  int* d_items = nullptr;
  int grid = 0;
  int wpc = 0;  // workgroups per CU of the persistent grid (0: as many as fit)
  long long slab_stride = 0;  // doubles per workgroup
  long long slab_base = 0;    // this bucket's region in the scratch (doubles)
  int stream = 0;             // which of the batch's streams launches it
  double work = 0.0;          // sum of n N over its instances
};

// rows per lane of a new bucket (implicit Euler keeps ~8 doubles per row in
// registers, IRK4 ~11: its variants stop at K = 8) and the bucket of n
// unknowns: shared with the allocation-free entry point (scft_internal.h)
int default_K(int cap, int stepper) { return bucket_default_K(cap, stepper == SCFT_STEP_IRK4); }
int cap_for(int n) { return bucket_capacity(n); }

}  // namespace

struct scft_batch {
  int device = 0;
  int n_inst = 0;
  int bc = SCFT_BC_DD, cmat = SCFT_C_ROWSCALED;
  int stepper = SCFT_STEP_IE;
  int cus = 1;  // compute units of the device
  std::vector<InstDesc> desc;
  std::vector<long long> uoff;
  long long total_u = 0, total_x = 0, total_pw = 0, total_h = 0;
  std::string tune_file;
  std::vector<Bucket> buckets;
  hipStream_t stream = nullptr;  // main stream: copies, resets, the mixer
  // buckets run concurrently on the main stream plus kSide side streams:
  // one stream per bucket for the seven buckets of meshes up to 4096 nodes
  // (a shared stream serialises its buckets: a sweep of 4096 uniform meshes,
  // m 32-4096, took 13.2-13.7 ms on four streams against 12.5 ms on one
  // stream each)
  static constexpr int kSide = 7;
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr, join[kSide] = {};
  // device state
  InstDesc* d_desc = nullptr;
  double *d_x = nullptr, *d_w = nullptr, *d_phit = nullptr, *d_pw = nullptr;
  double *d_phi = nullptr, *d_r = nullptr, *d_rnorm = nullptr;
  unsigned* d_status = nullptr;
  double* d_slabs = nullptr;
  long long slab_doubles = 0;
  int* d_counters = nullptr;  // bucket queues, the active count, then a slot the
                              // inner mixer rounds count into (never read)
  // mixer
  int hist_S = 0;
  double *d_hw = nullptr, *d_hr = nullptr, *d_wcold = nullptr;
  double* d_tri = nullptr;  // rational preconditioner factors + work, 4 x total_u
  bool tri_valid = false;    // factors computed for tri_shift / tri_floor
  double tri_shift = 0.0, tri_floor = 0.0;
  MixState* d_ms = nullptr;
  int* d_iters = nullptr;
  int* d_mix_order = nullptr;  // instances by decreasing n (mixer dispatch order)
  // algorithmic work per full sweep (all instances)
  double node_steps_all = 0.0, bytes_all = 0.0;
  bool has_wcold = false;
  std::vector<MixState> ms0;
  int* h_active = nullptr;  // pinned
  int last_sweeps = 0;
  // the sweep graph of the last resident solve, reused while its baked kernel
  // arguments (bucket launches, mixer arguments) are unchanged
  hipGraph_t sweep_graph = nullptr;
  hipGraphExec_t sweep_exec = nullptr;
  AndersonArgs sweep_args;
  int sweep_inner = 1;
  int top_bucket = -1;  // the non-empty bucket of most work (n N summed)
};

// Forget the cached sweep graph: the bucket launches or mixer buffers it
// bakes in are about to change.
static void drop_sweep_graph(scft_batch* b) {
  if (b->sweep_exec) (void)hipGraphExecDestroy(b->sweep_exec);
  if (b->sweep_graph) (void)hipGraphDestroy(b->sweep_graph);
  b->sweep_exec = nullptr;
  b->sweep_graph = nullptr;
}

// The march of the batch's stepper: implicit Euler (march.hip) or IRK4
// (irk4.hip); both take the same arguments, slabs and buckets.
static bool variant_exists(const scft_batch* b, int K, int W) {
  return b->stepper == SCFT_STEP_IRK4 ? irk4_variant_exists(K, W) : march_variant_exists(K, W);
}
static hipError_t launch_eval(const scft_batch* b, int K, int W, const MarchArgs& a, int grid,
                              hipStream_t s) {
  return b->stepper == SCFT_STEP_IRK4 ? launch_irk4(K, W, a, grid, s)
                                      : launch_march(K, W, a, grid, s);
}
static int eval_max_grid(const scft_batch* b, int K, int W, int half_max) {
  return b->stepper == SCFT_STEP_IRK4 ? irk4_max_grid(K, W, half_max)
                                      : march_max_grid(K, W, half_max);
}

// Algorithmic HBM bytes of one evaluation: q-slabs 1..N/2-1 written once and
// read back once, plus the O(n) inputs/outputs (x, w, phi_t in; phi, r out).
static double eval_bytes(int n, int N) {
  return 8.0 * n * (2.0 * (N / 2 - 1)) + 40.0 * n;
}

// Every non-empty bucket owns a scratch region (the buckets of one sweep run
// concurrently): grid_b workgroups x N_max/2 slabs x 64 W K doubles.
// Buckets are spread over the main and side streams, largest work first.
static int size_scratch(scft_batch* b) {
  drop_sweep_graph(b);
  long long need = 0;
  std::vector<std::pair<double, int>> order;
  for (size_t i = 0; i < b->buckets.size(); i++) {
    Bucket& bk = b->buckets[i];
    if (bk.items.empty()) continue;
    if (!variant_exists(b, bk.K, bk.W)) {
      snprintf(g_errmsg, sizeof g_errmsg, "no march variant K=%d W=%d", bk.K, bk.W);
      return SCFT_E_ARG;
    }
    const int P = 64 * bk.W;
    const int mg = eval_max_grid(b, bk.K, bk.W, bk.Nmax / 2);
    if (mg < 1) {
      snprintf(g_errmsg, sizeof g_errmsg, "march variant K=%d W=%d does not fit a CU", bk.K,
               bk.W);
      return SCFT_E_ARG;
    }
    // fewer co-resident workgroups per CU can finish a bucket sooner: its
    // longest instances set the makespan, and they step faster alone
    const int cap_grid = bk.wpc > 0 ? std::min(mg, bk.wpc * b->cus) : mg;
    bk.grid = std::max(1, std::min<int>(static_cast<int>(bk.items.size()), cap_grid));
    bk.slab_stride = static_cast<long long>(std::max(1, bk.Nmax / 2)) * P * bk.K;
    bk.slab_base = need;
    need += bk.slab_stride * bk.grid;
    double work = 0.0;
    for (int k : bk.items) work += static_cast<double>(b->desc[k].n) * b->desc[k].N;
    bk.work = work;
    order.push_back({work, static_cast<int>(i)});
  }
  std::sort(order.begin(), order.end(), std::greater<std::pair<double, int>>());
  b->top_bucket = order.empty() ? -1 : order[0].second;
  // greedy: next bucket onto the least loaded stream
  double load[1 + scft_batch::kSide] = {};
  for (auto& o : order) {
    int s = 0;
    for (int t = 1; t <= scft_batch::kSide; t++)
      if (load[t] < load[s]) s = t;
    b->buckets[o.second].stream = s;
    load[s] += o.first;
  }
  if (need > b->slab_doubles) {
    dfree(b->d_slabs);
    SCFT_TRY(dalloc(&b->d_slabs, static_cast<size_t>(need)));
    b->slab_doubles = need;
  }
  return SCFT_OK;
}

static void read_tune_file(scft_batch* b) {
  if (b->tune_file.empty()) return;
  std::ifstream f(b->tune_file);
  std::string line;
  while (std::getline(f, line)) {  // "cap K W [workgroups per CU]"
    std::istringstream ls(line);
    int cap, K, W, wpc = 0;
    if (!(ls >> cap >> K >> W)) continue;
    if (!(ls >> wpc) || wpc < 0) wpc = 0;
    for (auto& bk : b->buckets)
      if (bk.cap == cap && 64 * K * W >= cap && variant_exists(b, K, W)) {
        bk.K = K;
        bk.W = W;
        bk.wpc = wpc;
      }
  }
}

static void write_tune_file(const scft_batch* b) {
  if (b->tune_file.empty()) return;
  std::ofstream f(b->tune_file);
  for (auto& bk : b->buckets)
    if (!bk.items.empty()) f << bk.cap << " " << bk.K << " " << bk.W << " " << bk.wpc << "\n";
}

static MarchArgs march_args(const scft_batch* b, const Bucket& bk, int bucket_idx) {
  MarchArgs a;
  a.desc = b->d_desc;
  a.x = b->d_x;
  a.w = b->d_w;
  a.phit = b->d_phit;
  a.pw = b->d_pw;
  a.phi = b->d_phi;
  a.r = b->d_r;
  a.rnorm = b->d_rnorm;
  a.status = b->d_status;
  a.slabs = b->d_slabs + bk.slab_base;
  a.slab_stride = bk.slab_stride;
  a.half_max = bk.Nmax / 2;
  a.items = bk.d_items;
  a.queue = b->d_counters + bucket_idx;
  a.n_items = static_cast<int>(bk.items.size());
  return a;
}

// One evaluation sweep of every instance not marked done: counters reset on
// the main stream, then one persistent launch per non-empty bucket, the
// buckets spread over the main and side streams (fork/join through events,
// which stream capture turns into parallel graph branches).  The buckets are
// launched smallest first, so that their short, latency-bound grids become
// resident beside the large buckets' persistent grids.  That order matters
// when buckets share a stream (a sweep of 4096 uniform meshes on four
// streams: 13.2 ms against 14.0 ms largest first); with a stream per bucket
// both orders measure 12.5 ms.
static hipStream_t stream_of(scft_batch* b, int s) { return s == 0 ? b->stream : b->side[s - 1]; }

static int enqueue_sweep(scft_batch* b) {
  const int nb = static_cast<int>(b->buckets.size());
  HIPCHK(launch_reset(b->d_counters, nb + 1, b->stream));
  bool used[1 + scft_batch::kSide] = {true};
  for (auto& bk : b->buckets)
    if (!bk.items.empty()) used[bk.stream] = true;
  HIPCHK(hipEventRecord(b->fork, b->stream));
  for (int s = 1; s <= scft_batch::kSide; s++)
    if (used[s]) HIPCHK(hipStreamWaitEvent(b->side[s - 1], b->fork, 0));
  for (int i = 0; i < nb; i++) {
    const Bucket& bk = b->buckets[i];
    if (bk.items.empty()) continue;
    HIPCHK(launch_eval(b, bk.K, bk.W, march_args(b, bk, i), bk.grid, stream_of(b, bk.stream)));
  }
  for (int s = 1; s <= scft_batch::kSide; s++)
    if (used[s]) {
      HIPCHK(hipEventRecord(b->join[s - 1], b->side[s - 1]));
      HIPCHK(hipStreamWaitEvent(b->stream, b->join[s - 1], 0));
    }
  return SCFT_OK;
}

// One sweep of a solve: the march of every bucket as in enqueue_sweep, the
// mixer of every instance after the join.  With inner > 1 every bucket with
// at most a third of the work of the largest (whose march is the sweep's
// critical path) goes on, on its own stream, through inner - 1 further
// rounds of [mixer over its own instances, march]; a bucket with more
// would make the full sweeps longer than its rounds save (on 4096 uniform
// meshes the 2048 bucket, half the largest's work: its four rounds ran to
// 14.7 ms of a sweep whose largest march took 7.3).  Each instance still
// alternates one evaluation and one update, and instances already done are
// skipped by both kernels, so the results are the same bits as with
// inner = 1 and only the sweep in which an evaluation happens changes.  The inner mixer rounds count active instances into a
// slot of their own: the host's stop test reads the final mixer's count.
static int enqueue_solve_sweep(scft_batch* b, const AndersonArgs& a, int inner) {
  const int nb = static_cast<int>(b->buckets.size());
  HIPCHK(launch_reset(b->d_counters, nb + 1, b->stream));
  bool used[1 + scft_batch::kSide] = {true};
  for (auto& bk : b->buckets)
    if (!bk.items.empty()) used[bk.stream] = true;
  HIPCHK(hipEventRecord(b->fork, b->stream));
  for (int s = 1; s <= scft_batch::kSide; s++)
    if (used[s]) HIPCHK(hipStreamWaitEvent(b->side[s - 1], b->fork, 0));
  for (int i = 0; i < nb; i++) {  // every bucket's first march, smallest first
    const Bucket& bk = b->buckets[i];
    if (bk.items.empty()) continue;
    HIPCHK(launch_eval(b, bk.K, bk.W, march_args(b, bk, i), bk.grid, stream_of(b, bk.stream)));
  }
  // then the inner rounds, queued behind every first march (the largest
  // bucket's included) on streams that share a hardware queue
  for (int i = 0; i < nb; i++) {
    const Bucket& bk = b->buckets[i];
    if (bk.items.empty() || inner < 2 || i == b->top_bucket ||
        3.0 * bk.work > b->buckets[b->top_bucket].work)
      continue;
    const hipStream_t s = stream_of(b, bk.stream);
    const MarchArgs ma = march_args(b, bk, i);
    AndersonArgs ai = a;
    ai.order = bk.d_items;
    ai.n_inst = static_cast<int>(bk.items.size());
    ai.active = b->d_counters + nb + 1;
    for (int r = 1; r < inner; r++) {
      HIPCHK(launch_anderson(ai, s));
      HIPCHK(launch_reset(b->d_counters + i, 1, s));  // the bucket's work queue
      HIPCHK(launch_eval(b, bk.K, bk.W, ma, bk.grid, s));
    }
  }
  for (int s = 1; s <= scft_batch::kSide; s++)
    if (used[s]) {
      HIPCHK(hipEventRecord(b->join[s - 1], b->side[s - 1]));
      HIPCHK(hipStreamWaitEvent(b->stream, b->join[s - 1], 0));
    }
  HIPCHK(launch_anderson(a, b->stream));
  return SCFT_OK;
}

// ------------------------------------------------------------- public: batch
extern "C" int scft_batch_create(scft_batch** out, const scft_batch_desc* dsc) {
  if (!out || !dsc || dsc->n_inst <= 0 || !dsc->M || !dsc->N || !dsc->xoff || !dsc->x ||
      !dsc->phi_t)
    return SCFT_E_ARG;
  if (dsc->bc != SCFT_BC_DD && dsc->bc != SCFT_BC_DN) return SCFT_E_ARG;
  if (dsc->cmat != SCFT_C_ROWSCALED && dsc->cmat != SCFT_C_CONSISTENT) return SCFT_E_ARG;
  if (dsc->stepper != SCFT_STEP_IE && dsc->stepper != SCFT_STEP_IRK4) return SCFT_E_ARG;
  *out = nullptr;
  auto* b = new scft_batch();
  b->n_inst = dsc->n_inst;
  b->bc = dsc->bc;
  b->cmat = dsc->cmat;
  b->stepper = dsc->stepper;
  if (dsc->tune_file) b->tune_file = dsc->tune_file;
  if (hipGetDevice(&b->device) != hipSuccess ||
      hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device) !=
          hipSuccess) {
    delete b;
    return SCFT_E_HIP;
  }
  b->cus = std::max(1, b->cus);
  // descriptors, offsets, pairing weights per distinct N
  std::vector<int> Ns;
  std::vector<std::vector<double>> pws;
  std::vector<long long> pwo;
  long long pw_total = 0, x_total = 0;
  b->desc.resize(b->n_inst);
  b->uoff.resize(b->n_inst);
  for (int k = 0; k < b->n_inst; k++) {
    const int M = dsc->M[k], N = dsc->N[k];
    // N: a power of two (Romberg) in [16, 16384] (the N/2 + 1 pairing weights
    // of a bucket are held in LDS)
    // up to 4096 unknowns (both steppers)
    const int n_max = 4096;
    if (M < 3 || (b->bc == SCFT_BC_DD ? M - 2 : M - 1) > n_max || N < 16 || N > 16384 ||
        (N & (N - 1)) != 0) {
      snprintf(g_errmsg, sizeof g_errmsg, "instance %d: M=%d N=%d unsupported", k, M, N);
      delete b;
      return SCFT_E_ARG;
    }
    InstDesc& d = b->desc[k];
    d.M = M;
    d.N = N;
    d.n = (b->bc == SCFT_BC_DD) ? M - 2 : M - 1;
    d.bc = b->bc;
    d.cmat = b->cmat;
    d.pad = 0;
    d.xo = dsc->xoff[k];
    d.uo = b->total_u;
    d.wall0 = 0.0;
    d.wallL = 0.0;
    b->uoff[k] = b->total_u;
    b->total_u += d.n;
    x_total = std::max<long long>(x_total, dsc->xoff[k] + M);
    auto it = std::find(Ns.begin(), Ns.end(), N);
    if (it == Ns.end()) {
      std::vector<double> pw;
      pairing_weights(N, pw);
      Ns.push_back(N);
      pwo.push_back(pw_total);
      pw_total += static_cast<long long>(pw.size());
      pws.push_back(pw);
      d.po = pwo.back();
    } else {
      d.po = pwo[it - Ns.begin()];
    }
    d.ho = 0;
    b->node_steps_all += static_cast<double>(d.n) * N;
    b->bytes_all += eval_bytes(d.n, N);
  }
  b->total_x = x_total;
  b->total_pw = pw_total;
  // buckets by power-of-two capacity, items sorted by cost n*N descending
  std::vector<int> order(b->n_inst);
  for (int k = 0; k < b->n_inst; k++) order[k] = k;
  std::sort(order.begin(), order.end(), [&](int i, int j) {
    const double ci = static_cast<double>(b->desc[i].n) * b->desc[i].N;
    const double cj = static_cast<double>(b->desc[j].n) * b->desc[j].N;
    return ci > cj || (ci == cj && i < j);
  });
  for (int cap = 64; cap <= 4096; cap *= 2) {
    Bucket bk;
    bk.cap = cap;
    bk.K = default_K(cap, b->stepper);
    bk.W = cap / 64 / bk.K;
    b->buckets.push_back(bk);
  }
  for (int k : order) {
    const int c = cap_for(b->desc[k].n);
    for (auto& bk : b->buckets)
      if (bk.cap == c) {
        bk.items.push_back(k);
        bk.nmax = std::max(bk.nmax, b->desc[k].n);
        bk.Nmax = std::max(bk.Nmax, b->desc[k].N);
      }
  }
  read_tune_file(b);

  int rc = SCFT_OK;
  auto fail = [&](int code) {
    scft_batch_destroy(b);
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(SCFT_E_HIP);
  for (int s = 0; s < scft_batch::kSide; s++)
    if (hipStreamCreateWithFlags(&b->side[s], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->join[s], hipEventDisableTiming) != hipSuccess)
      return fail(SCFT_E_HIP);
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
    return fail(SCFT_E_HIP);
  if ((rc = dalloc(&b->d_desc, b->n_inst)) || (rc = dalloc(&b->d_x, b->total_x)) ||
      (rc = dalloc(&b->d_w, b->total_u)) || (rc = dalloc(&b->d_phit, b->total_u)) ||
      (rc = dalloc(&b->d_pw, b->total_pw)) || (rc = dalloc(&b->d_phi, b->total_u)) ||
      (rc = dalloc(&b->d_r, b->total_u)) || (rc = dalloc(&b->d_rnorm, b->n_inst)) ||
      (rc = dalloc(&b->d_status, b->n_inst)) ||
      (rc = dalloc(&b->d_counters, b->buckets.size() + 2)))
    return fail(rc);
  for (auto& bk : b->buckets) {
    if (bk.items.empty()) continue;
    if ((rc = dalloc(&bk.d_items, bk.items.size()))) return fail(rc);
    if (hipMemcpy(bk.d_items, bk.items.data(), bk.items.size() * sizeof(int),
                  hipMemcpyHostToDevice) != hipSuccess)
      return fail(SCFT_E_HIP);
  }
  std::vector<double> pw_all(std::max<long long>(1, b->total_pw));
  for (size_t i = 0; i < Ns.size(); i++)
    std::copy(pws[i].begin(), pws[i].end(), pw_all.begin() + pwo[i]);
  if (hipMemcpy(b->d_desc, b->desc.data(), sizeof(InstDesc) * b->n_inst,
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_x, dsc->x, sizeof(double) * b->total_x, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_phit, dsc->phi_t, sizeof(double) * b->total_u, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_pw, pw_all.data(), sizeof(double) * b->total_pw,
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(b->d_status, 0, sizeof(unsigned) * b->n_inst) != hipSuccess)
    return fail(SCFT_E_HIP);
  if ((rc = size_scratch(b))) return fail(rc);
  *out = b;
  return SCFT_OK;
}

extern "C" int scft_batch_destroy(scft_batch* b) {
  if (!b) return SCFT_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_sweep_graph(b);
  for (auto& bk : b->buckets) dfree(bk.d_items);
  dfree(b->d_desc);
  dfree(b->d_x);
  dfree(b->d_w);
  dfree(b->d_phit);
  dfree(b->d_pw);
  dfree(b->d_phi);
  dfree(b->d_r);
  dfree(b->d_rnorm);
  dfree(b->d_status);
  dfree(b->d_slabs);
  dfree(b->d_counters);
  dfree(b->d_hw);
  dfree(b->d_hr);
  dfree(b->d_wcold);
  dfree(b->d_tri);
  dfree(b->d_ms);
  dfree(b->d_iters);
  dfree(b->d_mix_order);
  if (b->h_active) (void)hipHostFree(b->h_active);
  for (int s = 0; s < scft_batch::kSide; s++) {
    if (b->side[s]) (void)hipStreamSynchronize(b->side[s]);
    if (b->side[s]) (void)hipStreamDestroy(b->side[s]);
    if (b->join[s]) (void)hipEventDestroy(b->join[s]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return SCFT_OK;
}

extern "C" int scft_batch_layout(const scft_batch* b, long long* uoff, long long* total) {
  if (!b) return SCFT_E_STATE;
  if (uoff) std::copy(b->uoff.begin(), b->uoff.end(), uoff);
  if (total) *total = b->total_u;
  return SCFT_OK;
}

extern "C" int scft_batch_set_walls(scft_batch* b, const double* wall0, const double* wallL) {
  if (!b || !wall0 || !wallL) return SCFT_E_ARG;
  for (int k = 0; k < b->n_inst; k++) {
    if (!std::isfinite(wall0[k]) || !std::isfinite(wallL[k])) return SCFT_E_ARG;
    b->desc[k].wall0 = wall0[k];
    b->desc[k].wallL = wallL[k];
  }
  HIPCHK(hipMemcpyAsync(b->d_desc, b->desc.data(), sizeof(InstDesc) * b->n_inst,
                        hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return SCFT_OK;
}

// Replace the target of every instance (host array on unknowns).
static int batch_set_target(scft_batch* b, const double* phit) {
  HIPCHK(hipMemcpyAsync(b->d_phit, phit, sizeof(double) * b->total_u, hipMemcpyHostToDevice,
                        b->stream));
  b->tri_valid = false;  // the rational preconditioner's factors depend on phi_t
  return SCFT_OK;
}

static int batch_eval_impl(scft_batch* b, const double* w, double* phi, double* r,
                           double* rnorm, unsigned* status) {
  HIPCHK(hipMemcpyAsync(b->d_w, w, sizeof(double) * b->total_u, hipMemcpyHostToDevice,
                        b->stream));
  HIPCHK(hipMemsetAsync(b->d_status, 0, sizeof(unsigned) * b->n_inst, b->stream));
  SCFT_TRY(enqueue_sweep(b));
  if (phi)
    HIPCHK(hipMemcpyAsync(phi, b->d_phi, sizeof(double) * b->total_u, hipMemcpyDeviceToHost,
                          b->stream));
  if (r)
    HIPCHK(hipMemcpyAsync(r, b->d_r, sizeof(double) * b->total_u, hipMemcpyDeviceToHost,
                          b->stream));
  if (rnorm)
    HIPCHK(hipMemcpyAsync(rnorm, b->d_rnorm, sizeof(double) * b->n_inst, hipMemcpyDeviceToHost,
                          b->stream));
  if (status)
    HIPCHK(hipMemcpyAsync(status, b->d_status, sizeof(unsigned) * b->n_inst,
                          hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return SCFT_OK;
}

extern "C" int scft_batch_eval(scft_batch* b, const double* w, double* phi, double* r,
                               double* rnorm, unsigned* status) {
  if (!b || !w) return SCFT_E_ARG;
  return batch_eval_impl(b, w, phi, r, rnorm, status);
}

static int time_sweeps(scft_batch* b, int reps, float* ms);

extern "C" int scft_batch_time_eval(scft_batch* b, const double* w, int reps,
                                    double* ms_per_eval, double* march_ms) {
  if (!b || !w || reps < 1) return SCFT_E_ARG;
  HIPCHK(hipMemcpyAsync(b->d_w, w, sizeof(double) * b->total_u, hipMemcpyHostToDevice,
                        b->stream));
  HIPCHK(hipMemsetAsync(b->d_status, 0, sizeof(unsigned) * b->n_inst, b->stream));
  float ms = 0.f;
  SCFT_TRY(time_sweeps(b, reps, &ms));
  if (ms_per_eval) *ms_per_eval = ms;
  if (march_ms) *march_ms = ms;
  return SCFT_OK;
}

extern "C" int scft_batch_time_buckets(scft_batch* b, const double* w, int reps,
                                       int max_buckets, int* n_buckets, double* ms,
                                       double* bytes, int* n_items) {
  if (!b || !w || reps < 1) return SCFT_E_ARG;
  HIPCHK(hipMemcpyAsync(b->d_w, w, sizeof(double) * b->total_u, hipMemcpyHostToDevice,
                        b->stream));
  HIPCHK(hipMemsetAsync(b->d_status, 0, sizeof(unsigned) * b->n_inst, b->stream));
  const int nb = static_cast<int>(b->buckets.size());
  int c = 0;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (int i = 0; i < nb; i++) {  // the order of scft_batch_tuning
    const Bucket& bk = b->buckets[i];
    if (bk.items.empty()) continue;
    const MarchArgs a = march_args(b, bk, i);
    HIPCHK(launch_reset(b->d_counters, nb + 1, b->stream));
    HIPCHK(launch_eval(b, bk.K, bk.W, a, bk.grid, b->stream));  // warm-up
    HIPCHK(hipEventRecord(e0, b->stream));
    for (int r = 0; r < reps; r++) {
      // the counter reset is a ~2 us kernel inside the timed region
      HIPCHK(launch_reset(b->d_counters, nb + 1, b->stream));
      HIPCHK(launch_eval(b, bk.K, bk.W, a, bk.grid, b->stream));
    }
    HIPCHK(hipEventRecord(e1, b->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    if (c < max_buckets) {
      if (ms) ms[c] = t / reps;
      double by = 0.0;
      for (int k : bk.items) by += eval_bytes(b->desc[k].n, b->desc[k].N);
      if (bytes) bytes[c] = by;
      if (n_items) n_items[c] = static_cast<int>(bk.items.size());
    }
    c++;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (n_buckets) *n_buckets = c;
  return SCFT_OK;
}

// Wall time of `reps` whole sweeps of the current device field, launched as
// the solve launches them (one sweep captured into a graph and replayed;
// HIP events on the main stream, after one untimed replay).
static int time_sweeps(scft_batch* b, int reps, float* ms) {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_sweep(b);
  hipError_t ce = hipStreamEndCapture(b->stream, &graph);
  if (rc == SCFT_OK && ce != hipSuccess) rc = SCFT_E_HIP;
  if (rc == SCFT_OK && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
    rc = SCFT_E_HIP;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == SCFT_OK &&
      (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = SCFT_E_HIP;
  if (rc == SCFT_OK && (hipGraphLaunch(exec, b->stream) != hipSuccess ||
                        hipEventRecord(e0, b->stream) != hipSuccess))
    rc = SCFT_E_HIP;
  for (int i = 0; i < reps && rc == SCFT_OK; i++)
    if (hipGraphLaunch(exec, b->stream) != hipSuccess) rc = SCFT_E_HIP;
  if (rc == SCFT_OK && (hipEventRecord(e1, b->stream) != hipSuccess ||
                        hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(ms, e0, e1) != hipSuccess))
    rc = SCFT_E_HIP;
  (void)hipStreamSynchronize(b->stream);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc == SCFT_OK) *ms /= static_cast<float>(reps);
  return rc;
}

extern "C" int scft_batch_tune_placement(scft_batch* b, int tries, int reps) {
  if (!b || !b->d_slabs || tries < 0) return SCFT_E_ARG;
  if (reps < 1) reps = 1;
  HIPCHK(hipMemsetAsync(b->d_status, 0, sizeof(unsigned) * b->n_inst, b->stream));
  const size_t bytes = static_cast<size_t>(b->slab_doubles) * sizeof(double);
  std::vector<double*> cand{b->d_slabs};
  float best_ms = 0.f;
  int best = 0, rc = time_sweeps(b, reps, &best_ms);
  for (int c = 0; c < tries && rc == SCFT_OK; c++) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess ||
        free_b < bytes + bytes / 4 + (size_t(1) << 30))
      break;
    double* p = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&p), bytes) != hipSuccess) {
      (void)hipGetLastError();
      break;
    }
    cand.push_back(p);
    b->d_slabs = p;
    drop_sweep_graph(b);
    float ms = 0.f;
    rc = time_sweeps(b, reps, &ms);
    if (rc == SCFT_OK && ms < best_ms) {
      best_ms = ms;
      best = static_cast<int>(cand.size()) - 1;
    }
  }
  (void)hipStreamSynchronize(b->stream);
  for (size_t i = 0; i < cand.size(); i++)
    if (static_cast<int>(i) != best) (void)hipFree(cand[i]);
  b->d_slabs = cand[best];
  drop_sweep_graph(b);
  return rc;
}

extern "C" int scft_batch_autotune(scft_batch* b, int reps) {
  if (!b) return SCFT_E_STATE;
  if (reps < 1) reps = 1;
  const int Ks[] = {1, 2, 4, 8, 16};
  const int Wsz[] = {1, 2, 4, 8, 16};
  // current field: leave d_w as is (callers upload first); clear statuses
  HIPCHK(hipMemsetAsync(b->d_status, 0, sizeof(unsigned) * b->n_inst, b->stream));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int nb = static_cast<int>(b->buckets.size());
  for (int bi = 0; bi < nb; bi++) {
    Bucket& bk = b->buckets[bi];
    if (bk.items.empty()) continue;
    float best = 1e30f;
    int bestK = bk.K, bestW = bk.W, bestP = bk.wpc;
    for (int K : Ks)
      for (int W : Wsz) {
        if (64 * K * W < bk.nmax || 64 * K * W > 2 * bk.cap) continue;
        if (!variant_exists(b, K, W)) continue;
        const int mg = eval_max_grid(b, K, W, bk.Nmax / 2);
        if (mg < 1) continue;
        const int occ = std::max(1, mg / b->cus);
        for (int wpc = 1; wpc <= occ; wpc++) {
          bk.K = K;
          bk.W = W;
          bk.wpc = (wpc == occ) ? 0 : wpc;
          SCFT_TRY(size_scratch(b));
          MarchArgs a = march_args(b, bk, bi);
          HIPCHK(launch_reset(b->d_counters, nb + 1, b->stream));
          HIPCHK(launch_eval(b, K, W, a, bk.grid, b->stream));  // warm-up
          HIPCHK(hipEventRecord(e0, b->stream));
          for (int r = 0; r < reps; r++) {
            HIPCHK(launch_reset(b->d_counters, nb + 1, b->stream));
            HIPCHK(launch_eval(b, K, W, a, bk.grid, b->stream));
          }
          HIPCHK(hipEventRecord(e1, b->stream));
          HIPCHK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, e0, e1));
          if (ms < best) {
            best = ms;
            bestK = K;
            bestW = W;
            bestP = bk.wpc;
          }
        }
      }
    bk.K = bestK;
    bk.W = bestW;
    bk.wpc = bestP;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SCFT_TRY(size_scratch(b));
  write_tune_file(b);
  return scft_batch_tune_placement(b, 5, std::max(2, reps));
}

extern "C" int scft_batch_tuning(const scft_batch* b, int* n_buckets, int* bucket_nmax, int* K,
                                 int* W, int max_buckets) {
  if (!b) return SCFT_E_STATE;
  int c = 0;
  for (auto& bk : b->buckets) {
    if (bk.items.empty()) continue;
    if (c < max_buckets) {
      if (bucket_nmax) bucket_nmax[c] = bk.nmax;
      if (K) K[c] = bk.K;
      if (W) W[c] = bk.W;
    }
    c++;
  }
  if (n_buckets) *n_buckets = c;
  return SCFT_OK;
}

// Diagnostic: the q-slabs 1..N/2-1 left in the scratch by the last
// evaluation of a single-instance batch, returned in node order
// (out[(s-1)*n + u] = q_s at unknown u).
extern "C" int scft_debug_slabs(scft_batch* b, int inst, double* out) {
  if (!b || !out || b->n_inst != 1 || inst != 0) return SCFT_E_ARG;
  const InstDesc& d = b->desc[0];
  for (auto& bk : b->buckets) {
    if (bk.items.empty()) continue;
    const int P = 64 * bk.W, K = bk.K, half = d.N / 2;
    std::vector<double> raw(static_cast<size_t>(half) * P * K);
    HIPCHK(hipMemcpy(raw.data(), b->d_slabs + bk.slab_base, raw.size() * sizeof(double),
                     hipMemcpyDeviceToHost));
    for (int s = 1; s < half; s++)
      for (int u = 0; u < d.n; u++) {
        const int p = u / K, t = u % K;  // lane p, row t (slab_put's layout)
        const size_t e = K >= 2 ? static_cast<size_t>(t / 2) * 2 * P + 2 * p + t % 2
                                : static_cast<size_t>(p);
        out[static_cast<size_t>(s - 1) * d.n + u] = raw[static_cast<size_t>(s) * P * K + e];
      }
    return SCFT_OK;
  }
  return SCFT_E_STATE;
}

// ------------------------------------------------------------ Anderson solve
extern "C" void scft_default_anderson(scft_anderson_opts* o) {
  o->history = 10;
  o->beta = 0.1;
  o->tol = 1e-6;
  o->max_iter = 5000;
  o->step_cap = 0.5;
  o->restart_factor = 1e30;
  o->regularization = 1e-14;
  o->transform = 0;
  o->floor = 1e-8;
  o->fallback_factor = 0.0;
  o->check_every = 8;
  o->progress_seconds = 0.0;
  o->precond = 0.0;
  o->backtrack = 0.0;
  o->max_halvings = 30;
  o->field_floor = 0.0;
  o->gs_precond = 0.0;
  o->gs_field_weight = 0.25;
  o->gs_switch = 0;
  o->gs_switch_growth = 0.0;
  o->gs_shift = 0.0;
  o->gs_shift_below = 0.0;
  o->inner_sweeps = 1;
}

static int ensure_mixer(scft_batch* b, int S) {
  if (b->hist_S == S && b->d_hw) return SCFT_OK;
  drop_sweep_graph(b);
  dfree(b->d_hw);
  dfree(b->d_hr);
  dfree(b->d_ms);
  dfree(b->d_iters);
  dfree(b->d_mix_order);
  long long off = 0;
  for (auto& d : b->desc) {
    d.ho = off;
    off += static_cast<long long>(S) * d.n;
  }
  b->total_h = off;
  SCFT_TRY(dalloc(&b->d_hw, off));
  SCFT_TRY(dalloc(&b->d_hr, off));
  SCFT_TRY(dalloc(&b->d_ms, b->n_inst));
  SCFT_TRY(dalloc(&b->d_iters, b->n_inst));
  std::vector<int> order(b->n_inst);
  for (int k = 0; k < b->n_inst; k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](int p, int q) { return b->desc[p].n > b->desc[q].n; });
  SCFT_TRY(dalloc(&b->d_mix_order, b->n_inst));
  HIPCHK(hipMemcpy(b->d_mix_order, order.data(), sizeof(int) * b->n_inst,
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_desc, b->desc.data(), sizeof(InstDesc) * b->n_inst,
                   hipMemcpyHostToDevice));
  b->hist_S = S;
  return SCFT_OK;
}

extern "C" int scft_batch_upload(scft_batch* b, const double* w, const double* w_cold) {
  if (!b || !w) return SCFT_E_ARG;
  HIPCHK(hipMemcpyAsync(b->d_w, w, sizeof(double) * b->total_u, hipMemcpyHostToDevice,
                        b->stream));
  if (w_cold) {
    if (!b->d_wcold) SCFT_TRY(dalloc(&b->d_wcold, b->total_u));
    HIPCHK(hipMemcpyAsync(b->d_wcold, w_cold, sizeof(double) * b->total_u,
                          hipMemcpyHostToDevice, b->stream));
    b->has_wcold = true;
  } else {
    b->has_wcold = false;
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return SCFT_OK;
}

extern "C" int scft_batch_solve_resident(scft_batch* b, const scft_anderson_opts* o,
                                         scft_solve_stats* stats) {
  if (!b || !o) return SCFT_E_ARG;
  if (o->history < 1 || o->history > MIX_MAX_HISTORY || o->max_iter < 1) return SCFT_E_ARG;
  if (!(o->precond >= 0.0) || !(o->backtrack >= 0.0) || o->max_halvings < 0) return SCFT_E_ARG;
  if (!(o->field_floor >= 0.0 && o->field_floor < 1.0)) return SCFT_E_ARG;
  if (!(o->gs_precond >= 0.0) || !(o->gs_field_weight >= 0.0) || o->gs_switch < 0 ||
      !(o->gs_switch_growth >= 0.0) || !(o->gs_shift >= 0.0) || !(o->gs_shift_below >= 0.0))
    return SCFT_E_ARG;
  if (o->inner_sweeps < 0 || o->inner_sweeps > 16) return SCFT_E_ARG;
  const int inner = std::max(1, o->inner_sweeps);
  SCFT_TRY(ensure_mixer(b, o->history + 1));
  if (o->gs_shift > 0.0 && o->gs_precond > 0.0 && !b->d_tri) {
    drop_sweep_graph(b);
    SCFT_TRY(dalloc(&b->d_tri, 4 * b->total_u));
  }
  const auto t0 = std::chrono::steady_clock::now();
  const int nb = static_cast<int>(b->buckets.size());
  AndersonArgs a;
  std::memset(&a, 0, sizeof a);  // padding included: the struct is compared bytewise
  a.order = b->d_mix_order;
  a.desc = b->d_desc;
  a.w = b->d_w;
  a.r = b->d_r;
  a.rnorm = b->d_rnorm;
  a.phit = b->d_phit;
  a.x = b->d_x;
  a.wcold = b->has_wcold ? b->d_wcold : nullptr;
  a.hw = b->d_hw;
  a.hr = b->d_hr;
  a.st = b->d_ms;
  a.status = b->d_status;
  a.iters = b->d_iters;
  a.active = b->d_counters + nb;
  a.n_inst = b->n_inst;
  a.history = o->history;
  a.transform = o->transform;
  a.max_iter = o->max_iter;
  a.beta = o->beta;
  a.tol = o->tol;
  a.step_cap = o->step_cap;
  a.restart_factor = o->restart_factor;
  a.reg = o->regularization;
  a.floor = o->floor;
  a.fallback_thresh = 0.0;
  a.precond = o->precond;
  a.backtrack = o->backtrack;
  a.max_halvings = o->max_halvings;
  a.field_floor = o->field_floor;
  a.gs_precond = o->gs_precond;
  a.gs_field_weight = o->gs_field_weight;
  a.gs_switch = o->gs_switch;
  a.gs_switch_growth = o->gs_switch_growth;
  a.gs_shift = o->gs_precond > 0.0 ? o->gs_shift : 0.0;
  a.gs_shift_below = o->gs_shift_below;
  a.tri = b->d_tri;
  a.tri_stride = b->total_u;
  if (a.gs_shift > 0.0 &&
      (!b->tri_valid || b->tri_shift != a.gs_shift || b->tri_floor != a.floor)) {
    // the factors of g I + G0: mesh and phi_t only, so once per batch
    HIPCHK(launch_gs_pivots(a, b->stream));
    b->tri_valid = true;
    b->tri_shift = a.gs_shift;
    b->tri_floor = a.floor;
  }
  // fresh mixer state (a small kernel-free host copy of zeros is not needed:
  // the per-instance state array is rewritten from a pinned staging copy)
  if (static_cast<int>(b->ms0.size()) != b->n_inst)
    b->ms0.assign(b->n_inst, MixState{0, 0, 0, 0, INFINITY, INFINITY, 0, 0});
  HIPCHK(hipMemcpyAsync(b->d_ms, b->ms0.data(), sizeof(MixState) * b->n_inst,
                        hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemsetAsync(b->d_status, 0, sizeof(unsigned) * b->n_inst, b->stream));
  HIPCHK(hipMemsetAsync(b->d_iters, 0, sizeof(int) * b->n_inst, b->stream));

  // first sweep eagerly: the warm-start fallback needs the median first residual
  SCFT_TRY(enqueue_sweep(b));
  if (b->has_wcold && o->fallback_factor > 0.0) {
    std::vector<double> rn(b->n_inst);
    HIPCHK(hipMemcpyAsync(rn.data(), b->d_rnorm, sizeof(double) * b->n_inst,
                          hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    std::vector<double> fin;
    for (double v : rn)
      if (std::isfinite(v)) fin.push_back(v);
    if (!fin.empty()) {
      std::nth_element(fin.begin(), fin.begin() + fin.size() / 2, fin.end());
      a.fallback_thresh = o->fallback_factor * fin[fin.size() / 2];
    }
  }
  HIPCHK(launch_anderson(a, b->stream));
  a.fallback_thresh = 0.0;  // only the first evaluation may be rejected

  // remaining sweeps: [reset, march per bucket (inner rounds beside the
  // largest bucket), mixer] captured once
  if (b->sweep_exec &&
      (std::memcmp(&a, &b->sweep_args, sizeof a) != 0 || b->sweep_inner != inner))
    drop_sweep_graph(b);
  if (!b->sweep_exec) {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_solve_sweep(b, a, inner);
    hipError_t ce = hipStreamEndCapture(b->stream, &graph);
    if (rc != SCFT_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIPCHK(ce);
    b->sweep_graph = graph;
    HIPCHK(hipGraphInstantiate(&b->sweep_exec, graph, nullptr, nullptr, 0));
    b->sweep_args = a;
    b->sweep_inner = inner;
  }
  hipGraphExec_t exec = b->sweep_exec;
  if (!b->h_active) HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&b->h_active), sizeof(int)));
  int sweeps = 1;
  const int check = std::max(1, o->check_every);
  double next_report = o->progress_seconds;
  for (;;) {
    HIPCHK(hipMemcpyAsync(b->h_active, b->d_counters + nb, sizeof(int), hipMemcpyDeviceToHost,
                          b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (o->progress_seconds > 0.0) {
      const double el =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (el >= next_report) {
        fprintf(stderr, "scft: %.1f s, %d sweeps, %d of %d instances active\n", el, sweeps,
                *b->h_active, b->n_inst);
        next_report = el + o->progress_seconds;
      }
    }
    if (*b->h_active == 0 || sweeps >= o->max_iter) break;
    for (int i = 0; i < check && sweeps < o->max_iter; i++, sweeps++)
      HIPCHK(hipGraphLaunch(exec, b->stream));
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  const auto t1 = std::chrono::steady_clock::now();
  b->last_sweeps = sweeps;
  if (stats) {
    stats->sweeps = sweeps;
    stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    stats->eval_seconds = 0.0;
  }
  return SCFT_OK;
}

extern "C" int scft_batch_download(scft_batch* b, double* w, int* iters, unsigned* status,
                                   scft_solve_stats* stats) {
  if (!b) return SCFT_E_ARG;
  if (w)
    HIPCHK(hipMemcpyAsync(w, b->d_w, sizeof(double) * b->total_u, hipMemcpyDeviceToHost,
                          b->stream));
  std::vector<int> it(b->n_inst, 0);
  std::vector<unsigned> stv(b->n_inst, 0);
  if (b->d_iters)
    HIPCHK(hipMemcpyAsync(it.data(), b->d_iters, sizeof(int) * b->n_inst,
                          hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipMemcpyAsync(stv.data(), b->d_status, sizeof(unsigned) * b->n_inst,
                        hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  int conv = 0;
  double ns = 0.0, by = 0.0;
  for (int k = 0; k < b->n_inst; k++) {
    if (stv[k] & ST_CONVERGED) conv++;
    ns += static_cast<double>(it[k]) * b->desc[k].n * b->desc[k].N;
    by += it[k] * eval_bytes(b->desc[k].n, b->desc[k].N);
  }
  if (iters) std::copy(it.begin(), it.end(), iters);
  if (status) std::copy(stv.begin(), stv.end(), status);
  if (stats) {
    stats->sweeps = b->last_sweeps;
    stats->converged = conv;
    stats->node_steps = ns;
    stats->bytes = by;
  }
  return conv == b->n_inst ? SCFT_OK : SCFT_E_NOCONV;
}

extern "C" int scft_batch_solve(scft_batch* b, const scft_anderson_opts* o, double* w,
                                const double* w_cold, int* iters, unsigned* status,
                                scft_solve_stats* stats) {
  if (!b || !o || !w) return SCFT_E_ARG;
  SCFT_TRY(scft_batch_upload(b, w, w_cold));
  scft_solve_stats st{};
  SCFT_TRY(scft_batch_solve_resident(b, o, &st));
  const double secs = st.seconds;
  const int rc = scft_batch_download(b, w, iters, status, &st);
  if (rc != SCFT_OK && rc != SCFT_E_NOCONV) return rc;
  st.seconds = secs;
  if (stats) *stats = st;
  return rc;
}

// ------------------------------------------- warm start into the batch
// The network's outputs mapped onto the batch's unknowns on the device
// (warmpredict.hip): the prediction writes the resident field directly.  A
// plan covers every instance of the batch or a chosen subset (one network
// per range of mesh sizes: one plan each).
struct scft_warm_plan {
  scft_batch* b = nullptr;
  scft_mlp* net = nullptr;
  int n_inst = 0, n_in = 0, n_out = 0, n_grid = 0, n_node = 0;  // n_inst: rows
  int* d_sel = nullptr;      // batch instance of each row (null: all, in order)
  double node_scale = 1.0;
  float* d_xn = nullptr;     // normalised inputs, n_inst x n_in
  double* d_grid = nullptr;  // n_inst x n_grid
  int* d_node_idx = nullptr;
  int* d_wall = nullptr;
  float* d_norm = nullptr;   // output shift, then scale (2 n_out)
};

extern "C" int scft_warm_plan_destroy(scft_warm_plan* p) {
  if (!p) return SCFT_E_ARG;
  dfree(p->d_xn);
  dfree(p->d_grid);
  dfree(p->d_node_idx);
  dfree(p->d_wall);
  dfree(p->d_norm);
  dfree(p->d_sel);
  delete p;
  return SCFT_OK;
}

extern "C" int scft_warm_plan_create_sel(scft_warm_plan** out, scft_batch* b, scft_mlp* net,
                                         int n_sel, const int* sel, int n_in,
                                         const float* inputs, int n_grid, const double* grid_x,
                                         int n_node, const int* node_idx, double node_scale,
                                         const int* wall_count) {
  if (!out || !b || !net || !inputs || !grid_x || !wall_count || n_in < 1) return SCFT_E_ARG;
  if (n_grid < 4 || n_grid > WARM_MAX_GRID || n_node < 0 || n_node > WARM_MAX_NODES ||
      (n_node > 0 && (!node_idx || n_node < 2)))
    return SCFT_E_ARG;
  if (sel ? (n_sel < 1 || n_sel > b->n_inst) : n_sel != b->n_inst) return SCFT_E_ARG;
  *out = nullptr;
  for (int j = 0; j < n_node; j++)
    if (node_idx[j] < 1 || (j > 0 && node_idx[j] <= node_idx[j - 1])) return SCFT_E_ARG;
  for (int k = 0; k < n_sel; k++) {
    // the instances of a subset: increasing, so each is written once
    if (sel && (sel[k] < 0 || sel[k] >= b->n_inst || (k > 0 && sel[k] <= sel[k - 1])))
      return SCFT_E_ARG;
    const int inst = sel ? sel[k] : k;
    if (wall_count[k] < 0 || wall_count[k] > b->desc[inst].n ||
        (wall_count[k] > 0 && (n_node < 2 || wall_count[k] > node_idx[n_node - 1])))
      return SCFT_E_ARG;
    const double* g = grid_x + static_cast<size_t>(k) * n_grid;
    for (int i = 1; i < n_grid; i++)
      if (!(g[i] > g[i - 1])) return SCFT_E_ARG;
  }
  auto* p = new scft_warm_plan();
  p->b = b;
  p->net = net;
  p->n_inst = n_sel;
  p->n_in = n_in;
  p->n_grid = n_grid;
  p->n_node = n_node;
  p->n_out = n_grid + n_node;
  p->node_scale = node_scale;
  std::vector<float> xn(static_cast<size_t>(p->n_inst) * n_in);
  int rc = scft_mlp_normalise_inputs(net, inputs, p->n_inst, xn.data());
  auto fail = [&](int code) {
    scft_warm_plan_destroy(p);
    return code;
  };
  if (rc) return fail(rc);
  if ((rc = dalloc(&p->d_xn, xn.size())) || (rc = dalloc(&p->d_grid, (size_t)p->n_inst * n_grid)) ||
      (rc = dalloc(&p->d_node_idx, std::max(1, n_node))) || (rc = dalloc(&p->d_wall, p->n_inst)) ||
      (rc = dalloc(&p->d_norm, 2 * static_cast<size_t>(p->n_out))) ||
      (sel && (rc = dalloc(&p->d_sel, p->n_inst))))
    return fail(rc);
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream) == hipSuccess;
  };
  if (!up(p->d_xn, xn.data(), sizeof(float) * xn.size()) ||
      !up(p->d_grid, grid_x, sizeof(double) * p->n_inst * n_grid) ||
      (n_node > 0 && !up(p->d_node_idx, node_idx, sizeof(int) * n_node)) ||
      !up(p->d_wall, wall_count, sizeof(int) * p->n_inst) ||
      (sel && !up(p->d_sel, sel, sizeof(int) * p->n_inst)) ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return fail(SCFT_E_HIP);
  *out = p;
  return SCFT_OK;
}

extern "C" int scft_warm_plan_create(scft_warm_plan** out, scft_batch* b, scft_mlp* net,
                                     int n_in, const float* inputs, int n_grid,
                                     const double* grid_x, int n_node, const int* node_idx,
                                     double node_scale, const int* wall_count) {
  if (!b) return SCFT_E_ARG;
  return scft_warm_plan_create_sel(out, b, net, b->n_inst, nullptr, n_in, inputs, n_grid, grid_x,
                                   n_node, node_idx, node_scale, wall_count);
}

extern "C" int scft_warm_plan_predict(scft_warm_plan* p) {
  if (!p) return SCFT_E_ARG;
  scft_batch* b = p->b;
  const float *d_y = nullptr, *ym = nullptr, *ys = nullptr;
  int n_out = 0, n_in = 0;
  SCFT_TRY(scft_mlp_forward_resident(p->net, p->n_inst, p->d_xn, &d_y, &ym, &ys, &n_out, &n_in));
  if (n_out != p->n_out || n_in != p->n_in) return SCFT_E_ARG;
  const bool norm = ym && ys;
  if (norm) {
    HIPCHK(hipMemcpyAsync(p->d_norm, ym, sizeof(float) * n_out, hipMemcpyHostToDevice,
                          b->stream));
    HIPCHK(hipMemcpyAsync(p->d_norm + n_out, ys, sizeof(float) * n_out, hipMemcpyHostToDevice,
                          b->stream));
  }
  WarmMapArgs a;
  a.desc = b->d_desc;
  a.x = b->d_x;
  a.w = b->d_w;
  a.y = d_y;
  a.ym = norm ? p->d_norm : nullptr;
  a.ys = norm ? p->d_norm + n_out : nullptr;
  a.grid = p->d_grid;
  a.node_idx = p->d_node_idx;
  a.wall_count = p->d_wall;
  a.sel = p->d_sel;
  a.node_scale = p->node_scale;
  a.n_inst = p->n_inst;
  a.n_out = n_out;
  a.n_grid = p->n_grid;
  a.n_node = p->n_node;
  HIPCHK(launch_warm_map(a, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return SCFT_OK;
}

// ------------------------------------------------------- drop-in evaluator
extern "C" {
double* f0_given = nullptr;
int local_iteration = 0;
int funcerr = 0;
}

namespace {
std::mutex g_mu;
bool g_init = false;
scft_config g_cfg;
scft_batch* g_one = nullptr;  // single-instance context of the drop-in entries
int g_one_n = -1;
std::vector<double> g_target;  // from scft_set_target (all nodes)
std::vector<double> g_x, g_ptall, g_ptu, g_r;
int g_last_status = SCFT_OK;
}  // namespace

extern "C" void scft_default_config(scft_config* c) {
  c->device = 0;
  c->N = 2048;
  c->L = 3.72374357332160;
  c->tau = 0.530252230020752;
  c->bc = SCFT_BC_DD;
  c->cmat = SCFT_C_ROWSCALED;
  c->beta = 0.1;
  c->verbose = 0;
  c->stepper = SCFT_STEP_IE;
}

static int init_locked(const scft_config* cfg) {
  if (cfg) {
    if (cfg->N < 16 || (cfg->N & (cfg->N - 1)) != 0 || !(cfg->L > 0.0) ||
        (cfg->bc != SCFT_BC_DD && cfg->bc != SCFT_BC_DN) ||
        (cfg->stepper != SCFT_STEP_IE && cfg->stepper != SCFT_STEP_IRK4))
      return SCFT_E_ARG;
    g_cfg = *cfg;
  } else {
    scft_default_config(&g_cfg);
  }
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (g_cfg.device < 0 || g_cfg.device >= ndev) return SCFT_E_HIP;
  HIPCHK(hipSetDevice(g_cfg.device));
  if (g_one) scft_batch_destroy(g_one);
  g_one = nullptr;
  g_one_n = -1;
  g_init = true;
  return SCFT_OK;
}

extern "C" int scft_init(const scft_config* cfg) {
  std::lock_guard<std::mutex> lk(g_mu);
  return init_locked(cfg);
}

extern "C" int scft_finalize(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_one) scft_batch_destroy(g_one);
  g_one = nullptr;
  g_one_n = -1;
  g_init = false;
  return SCFT_OK;
}

extern "C" const char* scft_strerror(int code) {
  switch (code) {
    case SCFT_OK: return "ok";
    case SCFT_E_ARG: return "invalid argument";
    case SCFT_E_HIP: return g_errmsg[0] ? g_errmsg : "HIP runtime error";
    case SCFT_E_NAN: return "non-finite value in the evaluation";
    case SCFT_E_PIVOT: return "non-positive tridiagonal pivot";
    case SCFT_E_NOMEM: return "out of memory";
    case SCFT_E_STATE: return "library not initialised / no context";
    case SCFT_E_NOCONV: return "outer solver did not converge";
    default: return "unknown error";
  }
}

extern "C" int scft_last_status(void) { return g_last_status; }

extern "C" int scft_set_target(int n_nodes, const double* phi_t) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!phi_t) {
    g_target.clear();
    return SCFT_OK;
  }
  if (n_nodes < 3) return SCFT_E_ARG;
  g_target.assign(phi_t, phi_t + n_nodes);
  return SCFT_OK;
}

// Evaluate r = phi[w] - phi_t for the drop-in entries (w, r on n unknowns).
static int dropin_eval(int n, const double* w, double* r) {
  if (!g_init) SCFT_TRY(init_locked(nullptr));
  const bool dd = (g_cfg.bc == SCFT_BC_DD);
  const int M = dd ? n + 2 : n + 1;
  if (n < 1) return SCFT_E_ARG;
  if (g_one_n != n) {
    if (g_one) scft_batch_destroy(g_one);
    g_one = nullptr;
    g_x.assign(M, 0.0);
    mesh_uniform(M, g_cfg.L, g_x.data());
    g_ptu.assign(n, 1.0);
    const long long xoff = 0;
    scft_batch_desc d{};
    d.n_inst = 1;
    d.M = &M;
    d.N = &g_cfg.N;
    d.xoff = &xoff;
    d.x = g_x.data();
    d.phi_t = g_ptu.data();
    d.bc = g_cfg.bc;
    d.cmat = g_cfg.cmat;
    d.stepper = g_cfg.stepper;
    d.tune_file = nullptr;
    SCFT_TRY(scft_batch_create(&g_one, &d));
    g_one_n = n;
  }
  // target: the caller's global (0-based, all nodes), else the installed one,
  // else the built-in profile
  g_ptall.assign(M, 1.0);
  if (f0_given) {
    std::copy(f0_given, f0_given + M, g_ptall.begin());
  } else if (static_cast<int>(g_target.size()) == M) {
    g_ptall = g_target;
  } else {
    target_position(M, g_x.data(), g_cfg.tau, dd, g_ptall.data());
  }
  for (int u = 0; u < n; u++) g_ptu[u] = g_ptall[u + 1];
  SCFT_TRY(batch_set_target(g_one, g_ptu.data()));
  // the consistent w-mass reads the field at the walls: extrapolated from the
  // unknowns by a natural spline as the deal.II evaluator does (DD: both
  // walls; DN: the wall at x = 0; a single unknown keeps zero walls)
  if (g_cfg.cmat == SCFT_C_CONSISTENT && n >= 2) {
    double w0 = 0.0, wl = 0.0;
    if (!wall_values_from_unknowns(M, g_x.data(), w, dd, &w0, &wl)) return SCFT_E_ARG;
    SCFT_TRY(scft_batch_set_walls(g_one, &w0, &wl));
  }
  double rn = 0.0;
  unsigned st = 0;
  SCFT_TRY(batch_eval_impl(g_one, w, nullptr, r, &rn, &st));
  if (st & ST_PIVOT) return SCFT_E_PIVOT;
  if ((st & ST_NAN) || !std::isfinite(rn)) return SCFT_E_NAN;
  return SCFT_OK;
}

extern "C" int simple_FEM_1D_transient(int n, double* w, double* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = dropin_eval(n, w + 1, out + 1);
  local_iteration++;
  g_last_status = rc;
  if (rc != SCFT_OK) funcerr = 1;
  if (g_cfg.verbose)
    for (int i = 1; i <= n; i++)
      printf("in[%d]=%2.15f ; out[%d]=%2.15f ; local_iteration:%d \n", i, w[i], i, out[i],
             local_iteration);
  return rc;
}

static void fixed_point_impl(int n, const double* x, double* xnew) {
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = dropin_eval(n, x, xnew);
  g_last_status = rc;
  local_iteration++;
  if (rc != SCFT_OK) {
    funcerr = 1;
    for (int i = 0; i < n; i++) xnew[i] = NAN;
    return;
  }
  for (int i = 0; i < n; i++) xnew[i] = x[i] + g_cfg.beta * xnew[i];
}

extern "C" void scft_fixed_point(int n, double* x, double* xnew) { fixed_point_impl(n, x, xnew); }

extern "C" void scft_fixed_point_nr(int n, double* x, double* xnew) {
  fixed_point_impl(n, x + 1, xnew + 1);
}

extern "C" void scft_adm_chen_f(int n, double* in, double* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = dropin_eval(n, in, out);
  g_last_status = rc;
  local_iteration++;
  if (rc != SCFT_OK) {
    funcerr = 1;
    for (int i = 0; i < n; i++) out[i] = NAN;
    return;
  }
  for (int i = 0; i < n; i++) out[i] = -out[i];
}

// ------------------------------------------------------------ resampling
extern "C" int scft_resample_spline(int n_src, const double* x_src, const double* y_src,
                                    int n_dst, const double* x_dst, double* y_dst) {
  if (n_src < 4 || n_dst < 0 || !x_src || !y_src || (n_dst > 0 && (!x_dst || !y_dst)))
    return SCFT_E_ARG;
  std::vector<double> M;
  if (!scft::spline_not_a_knot(n_src, x_src, y_src, M)) return SCFT_E_ARG;
  for (int i = 0; i < n_dst; i++) y_dst[i] = scft::spline_eval(n_src, x_src, y_src, M.data(), x_dst[i]);
  return SCFT_OK;
}

extern "C" int scft_resample_spline_batch(int n_tables, const int* n_src,
                                          const long long* src_off, const double* x_src,
                                          const double* y_src, const int* n_dst,
                                          const long long* dst_off, const double* x_dst,
                                          double* y_dst) {
  if (n_tables < 0 || (n_tables > 0 && (!n_src || !src_off || !x_src || !y_src || !n_dst ||
                                        !dst_off || !x_dst || !y_dst)))
    return SCFT_E_ARG;
  // tables are independent: contiguous ranges of roughly equal work on up
  // to 16 host threads
  double total = 0.0;
  for (int k = 0; k < n_tables; k++) total += n_src[k] + n_dst[k];
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int nt = (total < 2e5) ? 1 : static_cast<int>(std::min<unsigned>(hw, n_tables));
  std::vector<int> bounds(1, 0);
  double acc = 0.0;
  for (int k = 0; k < n_tables && static_cast<int>(bounds.size()) < nt; k++) {
    acc += n_src[k] + n_dst[k];
    if (acc >= total * bounds.size() / nt) bounds.push_back(k + 1);
  }
  bounds.push_back(n_tables);
  std::vector<int> rcs(bounds.size(), SCFT_OK);
  auto work = [&](size_t part) {
    for (int k = bounds[part]; k < bounds[part + 1]; k++) {
      const int rc = scft_resample_spline(n_src[k], x_src + src_off[k], y_src + src_off[k],
                                          n_dst[k], x_dst + dst_off[k], y_dst + dst_off[k]);
      if (rc != SCFT_OK) {
        rcs[part] = rc;
        return;
      }
    }
  };
  std::vector<std::thread> pool;
  for (size_t p = 1; p + 1 < bounds.size(); p++) pool.emplace_back(work, p);
  work(0);
  for (auto& t : pool) t.join();
  for (int rc : rcs)
    if (rc != SCFT_OK) return rc;
  return SCFT_OK;
}

// ------------------------------------------------- reference file formats
extern "C" int scft_read_res(const char* path, int max_rows, double* x_over_l, double* phi,
                             double* eta, double* phie, double* phij, int* rows,
                             scft_res_header* hdr) {
  if (!path || max_rows < 0 || !rows) return SCFT_E_ARG;
  FILE* f = fopen(path, "r");
  if (!f) return SCFT_E_ARG;
  char buf[512];
  scft_res_header h{};
  int r = 0, line = 0;
  bool table = false;
  while (fgets(buf, sizeof buf, f)) {
    line++;
    if (!table) {
      int a, b, nn;
      double v1, v2, v3, v4, v5;
      if (sscanf(buf, " m = %d n = %d", &a, &b) == 2) {
        h.m = a;
        h.n = b;
      } else if (sscanf(buf, " N=%d l=%lf cp=%lf mphi=%lf t=%lf", &nn, &v1, &v2, &v3, &v4) == 5) {
        h.iters = nn;
        h.l = v1;
        h.cp = v2;
        h.mphi = v3;
        h.t = v4;
      } else if (sscanf(buf, " Z=%lf f=%lf |f|max=%lf", &v1, &v2, &v3) == 3) {
        h.Z = v1;
        h.f = v2;
        h.fmax = v3;
      } else if (buf[0] == '-') {
        table = true;  // the dashes under the column names
      }
      (void)v5;
      continue;
    }
    double c[5];
    if (sscanf(buf, "%lf %lf %lf %lf %lf", &c[0], &c[1], &c[2], &c[3], &c[4]) != 5) continue;
    if (r < max_rows) {
      if (x_over_l) x_over_l[r] = c[0];
      if (phi) phi[r] = c[1];
      if (eta) eta[r] = c[2];
      if (phie) phie[r] = c[3];
      if (phij) phij[r] = c[4];
    }
    r++;
  }
  fclose(f);
  *rows = r;
  if (hdr) *hdr = h;
  return (table && r > 0) ? SCFT_OK : SCFT_E_ARG;
}

extern "C" int scft_write_solution(const char* path, int M, const double* x, const double* w,
                                   double err, double free_energy) {
  if (!path || M < 1 || !x || !w) return SCFT_E_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return SCFT_E_ARG;
  fprintf(f, "N= %d, ", M);
  fprintf(f, "ERROR= %e \n", err);
  fprintf(f, "mean_field_free_energy, %2.15f \n", free_energy);
  for (int i = 0; i < M; i++) fprintf(f, "%d,%2.15f,%2.15f\n", i, x[i], w[i]);
  const bool ok = fflush(f) == 0;
  fclose(f);
  return ok ? SCFT_OK : SCFT_E_ARG;
}

extern "C" int scft_read_solution(const char* path, int cap, int* M, double* x, double* w,
                                  double* err, double* free_energy) {
  if (!path || cap < 0 || !M) return SCFT_E_ARG;
  FILE* f = fopen(path, "r");
  if (!f) return SCFT_E_ARG;
  char buf[512];
  int n = 0;
  double e = NAN, fe = NAN;
  if (!fgets(buf, sizeof buf, f) || sscanf(buf, "N= %d", &n) != 1) {
    fclose(f);
    return SCFT_E_ARG;
  }
  const char* pe = strstr(buf, "ERROR=");
  if (pe) sscanf(pe, "ERROR= %lf", &e);
  int count = 0;
  while (fgets(buf, sizeof buf, f)) {
    if (strncmp(buf, "mean_field_free_energy", 22) == 0) {
      sscanf(buf, "mean_field_free_energy, %lf", &fe);
      continue;
    }
    int i;
    double xv, wv;
    if (sscanf(buf, "%d ,%lf, %lf", &i, &xv, &wv) != 3) continue;
    if (i < 0 || i >= n) continue;
    if (i < cap) {
      if (x) x[i] = xv;
      if (w) w[i] = wv;
    }
    count++;
  }
  fclose(f);
  *M = n;
  if (err) *err = e;
  if (free_energy) *free_energy = fe;
  return (count == n) ? SCFT_OK : SCFT_E_ARG;
}

extern "C" int scft_free_energy(int M, const double* x, const double* w_nodes, double L,
                                double tau, int mirror, int log2_points, double* F) {
  if (M < 2 || !x || !w_nodes || !F || !(L > 0.0) || log2_points < 4 || log2_points > 24)
    return SCFT_E_ARG;
  return scft::free_energy(M, x, w_nodes, L, tau, mirror != 0, log2_points, 0.892581217773656, F)
             ? SCFT_OK
             : SCFT_E_ARG;
}

// ------------------------------------------------------- mesh continuation
extern "C" int scft_refine_mesh(int M, const double* x, const double* w_nodes, double factor,
                                int bc, int cap, int* M_new, double* x_new, double* w_new) {
  if (M < 5 || !x || !w_nodes || !M_new || !(factor > 0.0)) return SCFT_E_ARG;
  std::vector<double> wpad(w_nodes, w_nodes + M);
  const std::vector<double> xn = scft::refine_by_gradient(M, x, wpad.data(), factor, bc == SCFT_BC_DD);
  const int Mn = static_cast<int>(xn.size());
  *M_new = Mn;
  if (!x_new || !w_new) return SCFT_OK;  // size query
  if (Mn > cap) return SCFT_E_ARG;
  std::copy(xn.begin(), xn.end(), x_new);
  // the field on the new unknowns: not-a-knot spline through the old
  // unknowns (refine_mesh.m:33, spline_chen); the walls keep their values
  const int u_end = (bc == SCFT_BC_DD) ? M - 1 : M;
  const int un_end = (bc == SCFT_BC_DD) ? Mn - 1 : Mn;
  std::vector<double> Ms;
  if (!scft::spline_not_a_knot(u_end - 1, x + 1, w_nodes + 1, Ms)) return SCFT_E_ARG;
  for (int i = 1; i < un_end; i++)
    w_new[i] = scft::spline_eval(u_end - 1, x + 1, w_nodes + 1, Ms.data(), xn[i]);
  w_new[0] = w_nodes[0];
  if (bc == SCFT_BC_DD) w_new[Mn - 1] = w_nodes[M - 1];
  return SCFT_OK;
}
