// scft_internal.h -- device data layout shared by the kernels and the host
// side of libscft_hip.so (not part of the public C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace scft {

// One SCFT instance as the kernels see it.  Unknown u <-> mesh node u+1.
struct InstDesc {
  int n;          // unknowns (M-2 for DD, M-1 for DN)
  int N;          // implicit-Euler contour steps, 2^k >= 16
  int M;          // mesh nodes
  int bc;         // SCFT_BC_DD / SCFT_BC_DN
  int cmat;       // SCFT_C_ROWSCALED / SCFT_C_CONSISTENT
  int pad;
  long long xo;   // offset of the M node coordinates in `x`
  long long uo;   // offset of the n unknowns in w / phi_t / phi / r
  long long po;   // offset of the N/2+1 pairing weights in `pw`
  long long ho;   // offset of the history ring (S*n doubles) of the mixer
  double wall0;   // field at node 0 (only the consistent C reads it)
  double wallL;   // field at node M-1 for DD (consistent C only)
};

// Arguments of one persistent march launch (one size bucket).
struct MarchArgs {
  const InstDesc* desc;
  const double* x;        // node coordinates
  const double* w;        // field on unknowns
  const double* phit;     // target on unknowns
  const double* pw;       // Romberg pairing weights
  double* phi;            // out: density on unknowns (nullable)
  double* r;              // out: residual phi - phi_t on unknowns
  double* rnorm;          // out: ||r||_inf per instance
  unsigned* status;       // per-instance status word
  double* slabs;          // q-slab scratch, one region per workgroup
  long long slab_stride;  // doubles per workgroup region (N_max/2 slabs)
  const int* items;       // instance ids of this bucket, largest first
  int* queue;             // next item to take (zeroed before every launch)
  int n_items;
  int half_max;           // largest N/2 of the bucket (LDS weight table size - 1)
};

constexpr unsigned ST_CONVERGED = 1u;
constexpr unsigned ST_NAN = 2u;
constexpr unsigned ST_PIVOT = 4u;
constexpr unsigned ST_MAXITER = 8u;
constexpr unsigned ST_FALLBACK = 16u;
constexpr unsigned ST_FAILED = 32u;
constexpr unsigned ST_PLAIN_FORM = 64u;  // left the ground-state form
constexpr unsigned ST_RATIONAL = 128u;   // reached the rational form's threshold
// instances the march skips: converged, out of iterations, or given up by the
// mixer (FAILED without a NaN: growth with no halvings left)
constexpr unsigned ST_DONE_MASK = ST_CONVERGED | ST_NAN | ST_PIVOT | ST_MAXITER | ST_FAILED;

struct MixState {
  int k;         // evaluations so far
  int head;      // next ring slot
  int stored;    // valid ring entries
  int restarts;
  double best;   // best ||r||_inf so far
  double eacc;   // ||r||_inf of the last accepted iterate (inf: none yet)
  int halvings;  // consecutive backtracking halvings
  int pad;
};

struct AndersonArgs {
  const int* order;  // instance of workgroup g: largest n first (nullable: g)
  const InstDesc* desc;
  double* w;            // field on unknowns (updated in place)
  const double* r;      // plain residual phi - phi_t from the march
  const double* rnorm;  // ||r||_inf from the march
  const double* phit;   // target on unknowns
  const double* x;      // node coordinates (preconditioner stencil)
  const double* wcold;  // cold-start fields (nullable)
  double* hw;           // history of fields   (S * n per instance at ho)
  double* hr;           // history of residuals
  MixState* st;
  unsigned* status;
  int* iters;           // evaluations used (written at convergence / stop)
  int* active;          // count of instances still running after this sweep
  int n_inst;
  int history;          // m (ring of m+1 slots)
  int transform;        // 0 plain, 1 relative, 2 log
  int max_iter;
  double beta, tol, step_cap, restart_factor, reg, floor;
  double fallback_thresh;  // absolute threshold (<= 0: off), first sweep only
  double precond;          // Laplacian preconditioner coefficient (0: off)
  double backtrack;        // growth factor that triggers a halving (0: off)
  int max_halvings;
  double field_floor;      // w >= -field_floor * N after each update (0: off)
  double gs_precond;       // ground-state form coefficient (0: off)
  double gs_field_weight;  // its field weight c
  int gs_switch;           // evaluations allowed in that form (0: no limit)
  double gs_switch_growth; // leave it when ||r|| > growth * best (0: never)
  double gs_shift;         // g > 0: the rational ground-state form (0: linear)
  double gs_shift_below;   // > 0: rational only once ||r|| < this (linear before)
  double* tri;             // 4 x tri_stride doubles at desc.uo: the factors of
  long long tri_stride;    // g I + G0 (1/pivot, lower/pivot, upper/pivot), work
};

// Size bucket of an instance with n unknowns: the power of two >= n (at
// least 64) = 64 * W * K lanes-times-rows of the workgroup that marches it.
inline int bucket_capacity(int n) {
  int c = 64;
  while (c < n) c *= 2;
  return c;
}
// Rows per lane of a bucket before any tuning: implicit Euler keeps ~8
// doubles per row in registers, IRK4 ~11 (its variants stop at K = 8).
inline int bucket_default_K(int cap, bool irk4) {
  const int k = cap / 64, kmax = irk4 ? 8 : 16;
  return k < kmax ? k : kmax;
}

// Launch one bucket.  Returns hipSuccess or the launch error.
hipError_t launch_march(int K, int W, const MarchArgs& a, int grid,
                        hipStream_t s);
// Largest grid of co-resident workgroups for a variant (persistent launch).
int march_max_grid(int K, int W, int half_max);
bool march_variant_exists(int K, int W);
// The same for the two-stage Gauss (IRK4) march (irk4.hip).
hipError_t launch_irk4(int K, int W, const MarchArgs& a, int grid, hipStream_t s);
int irk4_max_grid(int K, int W, int half_max);  // 0: the workgroup does not fit
bool irk4_variant_exists(int K, int W);
// The Thomas factors of the rational preconditioner's g I + G0 (once per
// batch and shift).
hipError_t launch_gs_pivots(const AndersonArgs& a, hipStream_t s);
hipError_t launch_anderson(const AndersonArgs& a, hipStream_t s);
hipError_t launch_reset(int* counters, int n, hipStream_t s);
constexpr int MIX_MAX_HISTORY = 16;

// Mapping of the warm-start network's outputs onto a batch's unknowns
// (warmpredict.hip): instance k's row of raw outputs y (before the output
// normalisation y' ys + ym) holds n_grid values at its grid points, then
// n_node wall samples asinh(w / node_scale) at the 1-based node numbers
// node_idx; its first wall_count[k] unknowns come from the samples
// (linear in the node number), the others from the not-a-knot cubic spline
// through the grid values.  One workgroup per instance.
constexpr int WARM_MAX_GRID = 256;
constexpr int WARM_MAX_NODES = 32;
struct WarmMapArgs {
  const InstDesc* desc;
  const double* x;         // the batch's node coordinates
  double* w;               // out: the batch's field on unknowns
  const float* y;          // network outputs, one row per instance of the plan
  const float* ym;         // output shift and scale (nullable: none)
  const float* ys;
  const double* grid;      // n_grid grid points per row, increasing
  const int* node_idx;     // n_node sample node numbers, increasing
  const int* wall_count;   // per row
  const int* sel;          // batch instance of each row (nullable: row k is instance k)
  double node_scale;
  int n_inst, n_out, n_grid, n_node;  // n_inst: rows
};
hipError_t launch_warm_map(const WarmMapArgs& a, hipStream_t s);

}  // namespace scft

// The network's forward pass on rows already normalised and resident in
// device memory (mlp.hip): *d_out receives its raw output activations
// (batch x dims.back(), before the output normalisation), *ym / *ys its
// output shift and scale as host arrays (nullptr: none).  Waits for the
// network's stream.
struct scft_mlp;
int scft_mlp_forward_resident(scft_mlp* m, int batch, const float* d_xn, const float** d_out,
                              const float** ym, const float** ys, int* n_out, int* n_in);
int scft_mlp_normalise_inputs(const scft_mlp* m, const float* X, int rows, float* out);

namespace scft {

}  // namespace scft
