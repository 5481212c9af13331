// Shared internals of the device library: plan / workspace layout, error
// plumbing and launch helpers. All device code targets gfx950 (wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "grico_mf.h"

namespace gmf {

struct Error : std::runtime_error {
  gmf_status status;
  Error(gmf_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define GMF_HIP(call)                                                                        \
  do {                                                                                       \
    hipError_t _e = (call);                                                                  \
    if (_e != hipSuccess)                                                                    \
      throw ::gmf::Error(GMF_EDEVICE, std::string(#call) + ": " + hipGetErrorString(_e));    \
  } while (0)

inline void require(bool ok, gmf_status s, const std::string& msg) {
  if (!ok) throw Error(s, msg);
}

constexpr int kMaxDim = 3;
constexpr int kMaxExtent = 16;  // per-axis modes / points handled by the generic kernel

// 2x2x2-brick decomposition of a fast-path bucket (cell_fast.hip). All
// bricks run in one launch. A dof held by one brick only is stored by that
// brick; every occurrence of a dof held by several bricks (the bricks' shared
// surfaces) writes its own halo slot, and k_brick_halo sums each such dof's
// slots, which are contiguous and in brick order.
struct BrickPath {
  bool enabled = false;
  int64_t bricks = 0;
  // bricks x (2N-1)^3 lattice entries (dof << 1) | exclusive; -1: no dof;
  // -2: no dof (a constrained dof of the reduced layout) at a shared position,
  // which still takes its halo slot.
  int32_t* map_full = nullptr;
  int32_t* map_red = nullptr;  // same with reduced dofs
  // Shared positions of brick b, in ascending lattice position, write the
  // halo slots halo_slot[halo_base[b] .. halo_base[b + 1]) (bricks + 1 offsets).
  int32_t* halo_base = nullptr;
  int32_t* halo_slot = nullptr;
  int64_t halo_slots = 0;      // halo slots of the bucket
  // Shared dof i sums slots halo_ptr[i] .. halo_ptr[i+1]-1 into halo_full[i] /
  // halo_red[i] (-1: no reduced dof).
  int64_t shared_count = 0;
  int32_t* halo_full = nullptr;
  int32_t* halo_red = nullptr;
  int32_t* halo_ptr = nullptr;
  int64_t slot_base = 0;       // offset of this bucket's slots in the workspace array
  // bricks x 8 batch cells (-1: empty position) when the bricks come from an
  // explicit grouping; null: brick b holds cells 8 b .. 8 b + 7.
  int32_t* cell_of = nullptr;
};

// Sum-factorized path of a point-cloud batch of conforming faces
// (face_tensor.hip): every item's points are the same tensor grid of
// tangential Gauss coordinates on both sides, at a fixed normal coordinate
// +-1 per side. The point action is stored compressed to the nonzero
// (test channel, trial channel) pattern shared by all points.
struct FacePath {
  bool enabled = false;
  int axis = 0;                 // normal axis
  int T[2] = {-1, -1};          // tangential axes (-1: absent)
  int Qt[2] = {1, 1};           // tangential points per slot
  int toff[2][6] = {};          // per side: offsets of psi(+-1), psi'(+-1), tv0, td0, tv1, td1 in tables
  int table_doubles = 0;
  double* tables = nullptr;
  int nnz = 0;
  unsigned char pr[64] = {}, pc[64] = {};  // pattern: test channel, trial channel
  double* pdata = nullptr;      // points x nnz
  size_t lds = 0;
};

// Device-resident cell batch.
constexpr int kCloudUnitPoints = 512;  // default points per work unit of the point-cloud kernels

// Kernel family and launch shape a bucket runs with, as picked by the
// plan-time tuner (tune.hip); mirrors gmf_bucket_tuning.
enum TuneKernel {
  kTuneNone = 0,
  kTuneBrick = 1,
  kTuneCell = 2,
  kTuneGeneric = 3,
  kTuneCloud = 4,
  kTuneFace = 5,
  kTuneCloudMatrix = 6
};
struct TuneInfo {
  int kernel = kTuneNone;
  int param0 = 0, param1 = 0;  // brick: -, -; cell / generic: lanes per workgroup / cell; cloud: unit points, LDS KB
  int candidates = 0;          // shapes timed (0: not tuned)
  float best_us = 0.0f;        // time per launch of the chosen shape
  float default_us = 0.0f;     // time per launch of the default shape
};

struct Bucket {
  int dim = 0;
  int N[kMaxDim] = {1, 1, 1};
  int Q[kMaxDim] = {1, 1, 1};
  int form = GMF_FORM_DIFFUSION;
  int64_t cells = 0;
  int Ntot = 1, Qtot = 1;
  int point_stride = 1;  // coefficients per point (1 or (D+1)^2)
  double* values = nullptr;   // per-axis tables, concatenated
  double* derivs = nullptr;
  double* jinv = nullptr;     // cells x dim
  double* pdata = nullptr;    // cells x Qtot x point_stride
  int32_t* lidx = nullptr;    // cells x Ntot (full layout)
  int32_t* ridx = nullptr;    // cells x Ntot (reduced layout; only for selection maps)
  double* lcoef = nullptr;    // cells x Ntot or null
  int64_t multi_count = 0;
  int64_t* moff = nullptr;
  int32_t* midx = nullptr;    // full layout
  int32_t* mridx = nullptr;   // reduced layout (-1 = fixed)
  double* mcoef = nullptr;
  int table_offset[kMaxDim] = {0, 0, 0};  // offset of axis a in values/derivs
  std::vector<double> h_values, h_derivs;
  // Fast-path data (3D, N == Q, diffusion): collocation derivative matrix
  // Dc[q*N+s] = l_s'(xi_q), passed to the kernel by value.
  std::vector<double> h_colloc;
  bool fast = false;
  BrickPath bricks;
  // Lanes per cell of the generic kernel and lanes per workgroup of the
  // per-cell fast kernel, picked by the plan-time tuner.
  int generic_threads = 64;
  int cell_threads = 256;
  TuneInfo tune;
  // FP32 copies of the per-cell arrays for FP32 multigrid levels (prepare_fp32).
  bool fp32 = false;
  float* jinv32 = nullptr;
  float* pdata32 = nullptr;
  float* lcoef32 = nullptr;
  // Point-cloud batch: item c owns points poff[c]..poff[c+1]-1 at biunit
  // reference coordinates rpts (points x dim); pdata rows per point.
  bool cloud = false;
  // Interface items (sides = 2): the local box is the minus box (N) followed
  // by the plus box (Np); rpts and jinv hold both sides per point / item and
  // pdata rows are (2(D+1))^2.
  int sides = 1;
  int Np[kMaxDim] = {1, 1, 1};
  FacePath face;
  int64_t total_points = 0;
  int max_points = 0;
  int64_t* poff = nullptr;
  double* rpts = nullptr;
  // Work units of the apply kernels: items with more points than
  // kCloudUnitPoints are split into point ranges (unit u covers item
  // unit_item[u], points unit_start[u] .. min(unit_start[u] + kCloudUnitPoints,
  // poff[item + 1]) - 1), so heavy cut cells spread over several workgroups.
  int64_t units = 0;
  int32_t* unit_item = nullptr;
  int64_t* unit_start = nullptr;
  int unit_points = kCloudUnitPoints;  // points per work unit (tuned)
  int lds_kb = 64;                     // LDS budget of one workgroup (tuned)
  // Local matrices (items x Ntot x Ntot, column-major per item), built from
  // the point data at plan creation when the tuner finds their apply faster;
  // null: the apply runs over the points.
  double* emat = nullptr;
  std::vector<int64_t> h_poff;         // host copy of poff (work-unit rebuilds)
  // Exact reduced diagonal on cells whose modes couple through multi-term
  // expansions or repeat a dof: `coupled` flags those cells (null if none) and
  // each (cell, reduced dof) pair lists the dof's local coefficient vector.
  uint8_t* coupled = nullptr;
  int64_t pair_count = 0;
  int64_t* pair_cell = nullptr;
  int32_t* pair_dof = nullptr;
  int64_t* pair_off = nullptr;
  int32_t* pair_loc = nullptr;
  double* pair_coef = nullptr;
};

// Generic-kernel buckets of a plan applied in one launch (cell_generic.hip):
// per-bucket kernel arguments on the device (full and reduced index maps),
// cumulative cell offsets, and the bucket ids covered.
struct GenericBatch {
  int threads = 64;          // lanes per cell of this launch
  int count = 0;
  void* args_full = nullptr;
  void* args_red = nullptr;  // null when a bucket has no reduced index map
  int64_t* start = nullptr;  // count + 1 cumulative cell offsets (device)
  int64_t cells = 0;
  size_t lds = 0;
  std::vector<int> buckets;
};

struct Csr {
  int64_t rows = 0;
  int64_t* offsets = nullptr;
  int64_t* indices = nullptr;
  double* coefficients = nullptr;
};

}  // namespace gmf

struct gmf_plan {
  int device = 0;
  int dim = 0;
  int64_t n_full = 0;
  int64_t n_red = 0;
  std::vector<gmf::Bucket> buckets;
  int64_t fixed_count = 0;
  int64_t* fixed = nullptr;
  int64_t affine_count = 0;
  int64_t* affine_pivots = nullptr;
  gmf::Csr affine;
  int64_t mean_count = 0;
  int64_t* mean_pivots = nullptr;
  gmf::Csr mean;
  gmf::Csr map;      // full rows -> reduced
  gmf::Csr map_t;    // reduced rows -> full (transpose, for P^T)
  double* shift = nullptr;
  bool selection_map = false;  // P is a 0/1 selection without shift terms in apply
  bool multi_terms = false;
  int64_t device_bytes = 0;
  int64_t brick_slots = 0;  // halo slots of all brick buckets (workspace scratch)
  int fp32_state = 0;       // 0: not prepared, 1: every bucket has FP32 kernels, -1: FP64 with conversion
  std::vector<gmf::GenericBatch> generic_batches;  // one launch per lanes-per-cell tier
  gmf::GenericBatch face_batch;                    // every face-tensor batch in one launch
  std::vector<void*> allocations;
};

struct gmf_workspace {
  gmf_plan* plan = nullptr;
  hipStream_t stream = nullptr;
  double* full_in = nullptr;
  double* full_out = nullptr;
  double* partials = nullptr;   // reduction partial sums
  double* scalars = nullptr;    // device scalars for solvers
  double* host_scalars = nullptr;  // pinned mirror
  double* pipelined_scalars = nullptr;  // pinned mirror of the CG loop's overlapped reads
  hipEvent_t scalar_event = nullptr;    // marks the arrival of pipelined_scalars
  int64_t partial_capacity = 0;
  double* brick_slots = nullptr;   // plan->brick_slots halo slots of the brick kernels
  double* red64_in = nullptr;      // FP64 staging of FP32 reduced applies on plans without FP32 kernels
  double* red64_out = nullptr;
};

namespace gmf {

// Kernel launchers implemented in the .hip files.
void launch_cell_generic(const gmf_plan& plan, const Bucket& b, const double* x, double* y, bool reduced,
                         hipStream_t stream);
// Right-hand side of one tensor bucket from lowered point loads (cells x Q x
// channels, channels 1 or 1 + D, weights included), added into y (full layout).
void launch_cell_rhs(const Bucket& b, const double* loads, int channels, double* y, hipStream_t stream);
// Per-cell verification sums (w (u_h - e)^2, w e^2) into partial (2 per cell).
void launch_cell_l2(const Bucket& b, const double* u, const double* w, const double* e, double* partial,
                    hipStream_t stream);
// Value (+ physical gradient) of full-layout u at every point of one tensor
// bucket: cells x Q x channels (channels 1 or 1 + D).
void launch_cell_interp(const Bucket& b, const double* u, int channels, double* out, hipStream_t stream);
// Sum-factorized verification sums on fast buckets (3D, N = Q, isotropic
// tables): false when the bucket has no fast path (the caller then runs the
// generic kernel).
bool launch_cell_l2_fast(const Bucket& b, const double* u, const double* w, const double* e, double* partial,
                         hipStream_t stream);
// out[0..1] = column sums of partial (cells x 2) in a fixed order; partial is
// overwritten (its chunk sums).
void launch_sum_pairs(double* partial, int64_t cells, double* out, hipStream_t stream);
// Plan-time autotuner (tune.hip): at plan creation every bucket's kernel
// launch shape is timed over its candidates on scratch vectors (HIP events,
// one warm-up and `reps` timed launches each) and the fastest is kept.
struct TuneTimer {
  double* x = nullptr;  // scratch input (n_full, filled with 1)
  double* y = nullptr;  // scratch output (n_full)
  double* slots = nullptr;  // scratch brick halo slots (plan.brick_slots)
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int reps = 3;
  template <class F>
  float time_us(F&& launch) {
    launch();
    GMF_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) launch();
    GMF_HIP(hipEventRecord(e1, nullptr));
    GMF_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    GMF_HIP(hipEventElapsedTime(&ms, e0, e1));
    return 1000.0f * ms / float(reps);
  }
};
// Tune every bucket of the plan (GMF_TUNE=0: keep the default shapes).
void tune_plan(gmf_plan& plan);
// Per-family tuners: generic kernel lanes per cell (64, 128, 256;
// GMF_GENERIC_THREADS forces one), fast kernels (brick vs per-cell and the
// per-cell workgroup size), point-cloud work-unit size and LDS budget.
void tune_generic(gmf_plan& plan, TuneTimer& t);
void tune_fast(gmf_plan& plan, TuneTimer& t);
void tune_cloud(gmf_plan& plan, TuneTimer& t);
// Work units of a point-cloud bucket at `unit_points` points per unit into
// device arrays owned by the caller (item, first point); returns the count.
int64_t cloud_units(const Bucket& b, int unit_points, std::vector<int32_t>& item, std::vector<int64_t>& start);
// Batch the plan's generic-kernel buckets into one launch per lanes-per-cell
// tier (tiers holding at least two buckets).
void prepare_generic_batch(gmf_plan& plan);
// Launch every batch; `done[i]` (one flag per bucket, cleared by the caller)
// is set for the buckets they covered. Clear entries must be launched per
// bucket (no batch, or no reduced maps).
void launch_generic_batches(const gmf_plan& plan, const double* x, double* y, bool reduced, hipStream_t stream,
                            char* done);
// Per-bucket flags for the batch launchers: on the stack for plans of at most
// kStackBuckets buckets (no host allocation per cell pass), else on the heap.
struct BucketFlags {
  static constexpr size_t kStackBuckets = 64;
  char stack[kStackBuckets] = {};
  std::vector<char> heap;
  char* data;
  explicit BucketFlags(size_t n) : data(stack) {
    if (n > kStackBuckets) {
      heap.assign(n, 0);
      data = heap.data();
    }
  }
  char operator[](size_t i) const { return data[i]; }
};
void launch_cell_diagonal_generic(const gmf_plan& plan, const Bucket& b, double* diag, bool reduced,
                                  hipStream_t stream);
// Exact reduced-diagonal contributions of the bucket's coupled cells (and the
// local diagonal of the others, which the caller launches with skip_coupled).
void launch_cell_diagonal_pairs(const gmf_plan& plan, const Bucket& b, double* diag, hipStream_t stream);
void launch_cell_diagonal_generic_skip(const gmf_plan& plan, const Bucket& b, double* diag, hipStream_t stream);
// Point-cloud batches (cell_cloud.hip).
void launch_cell_cloud(const gmf_plan& plan, const Bucket& b, const double* x, double* y, bool reduced,
                       hipStream_t stream);
void launch_cell_cloud_diagonal(const gmf_plan& plan, const Bucket& b, double* diag, bool reduced, hipStream_t stream);
// Detect conforming-face batches and prepare their sum-factorized path
// (face_tensor.hip; GMF_FACE_TENSOR=0 keeps every batch on the point-cloud kernel).
void prepare_face_path(gmf_plan& plan, Bucket& b, const gmf_bucket_desc& bd);
bool launch_face_tensor(const gmf_plan& plan, const Bucket& b, const double* x, double* y, bool reduced,
                        hipStream_t stream);
// All face-tensor batches of the plan in one launch (after every bucket is
// prepared); launch_face_batch marks the buckets it covered in `done`.
void prepare_face_batch(gmf_plan& plan);
// Face loads (items x points x sides x (1 + D)) of one face-tensor batch added into y.
void launch_face_rhs(const Bucket& b, const double* loads, double* y, hipStream_t stream);
// Point loads of a point-cloud batch off the face path projected and added
// into y (full layout): per item, point and side (W v, W g_1, .., W g_D).
void launch_cloud_rhs(const Bucket& b, const double* loads, double* y, hipStream_t stream);
// Verification sums of a one-sided point-cloud batch off the face path: one
// pair (sum w (u_h - e)^2, sum w e^2) per work unit into partial.
void launch_cloud_l2(const Bucket& b, const double* u_full, const double* w, const double* e, double* partial,
                     hipStream_t stream);
void launch_face_batch(const gmf_plan& plan, const double* x, double* y, bool reduced, hipStream_t stream,
                       char* done);
// slots: the workspace's brick halo array (plan.brick_slots entries).
bool launch_cell_fast(const gmf_plan& plan, const Bucket& b, const double* x, double* y, bool reduced, double* slots,
                      hipStream_t stream);
// lidx: host copy of the bucket's full-layout index map; red_of_full: full dof
// -> reduced dof (empty unless the reduced map is a selection).
void prepare_fast_path(gmf_plan& plan, Bucket& b, const int32_t* lidx, const std::vector<int64_t>& red_of_full,
                       const int32_t* brick_cells = nullptr, int64_t brick_count = 0);
bool plan_writes_every_dof(const gmf_plan& plan);
// FP32 variant of launch_cell_fast for buckets prepared by prepare_fp32.
bool launch_cell_fast32(const gmf_plan& plan, const Bucket& b, const float* x, float* y, bool reduced, float* slots,
                        hipStream_t stream);
// Make FP32 copies of the fast buckets' per-cell arrays; true when every
// bucket of the plan then has FP32 kernels (and the reduced map is a selection).
bool prepare_fp32(gmf_plan& plan);

// One Chebyshev smoother step applied where the operator's result of dof i is
// stored (v = (A u)_i, u_i the apply input at i), with the arithmetic of the
// separate vector kernels (no contraction of a product into a following sum in
// either), so fused and unfused steps round alike:
//   start (u is x):  r = b - v, d = c2 (D^-1 r), x = u + d
//   step  (u is d):  r = r - v, d = c1 u + c2 (D^-1 r), x = x + d
// Each reduced dof is stored exactly once per apply (by the brick holding it
// alone, or by the halo pass after every brick finished), so updating u in
// place there races with no other read of u.
template <class R>
struct ChebFused {
  R* x;
  R* d;
  R* r;
  const R* b;
  const R* dinv;
  R c1, c2;
  int start;
  __device__ void operator()(int64_t i, R v, R u) const {
#pragma clang fp contract(off)
    if (start) {
      const R ri = b[i] - v;
      const R di = c2 * (dinv[i] * ri);
      r[i] = ri;
      d[i] = di;
      x[i] = u + di;
    } else {
      const R ri = r[i] - v;
      const R di = fma(c1, u, c2 * (dinv[i] * ri));
      r[i] = ri;
      d[i] = di;
      x[i] += di;
    }
  }
};
// The Chebyshev step fused into the reduced apply of u when the plan is one
// brick bucket on a selection map (u = c.x for start, c.d for step); returns
// false, launching nothing, for any other plan. slots: the workspace's halo.
bool chebyshev_fused(const gmf_plan& plan, const double* u, const ChebFused<double>& c, double* slots,
                     hipStream_t stream);
bool chebyshev_fused(const gmf_plan& plan, const float* u, const ChebFused<float>& c, float* slots,
                     hipStream_t stream);

// Copy a host array to a new device allocation owned by the plan.
template <class T>
T* upload_array(gmf_plan& p, const T* host, size_t n) {
  if (n == 0 || host == nullptr) return nullptr;
  void* d = nullptr;
  GMF_HIP(hipMalloc(&d, n * sizeof(T)));
  p.allocations.push_back(d);
  p.device_bytes += int64_t(n * sizeof(T));
  GMF_HIP(hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice));
  return static_cast<T*>(d);
}

// Vector kernels (blas.hip).
void fill(double* x, double v, int64_t n, hipStream_t s);
void fill(float* x, float v, int64_t n, hipStream_t s);
void copy(double* dst, const double* src, int64_t n, hipStream_t s);
void copy(float* dst, const float* src, int64_t n, hipStream_t s);
// Precision conversions (round to nearest).
void narrow(float* dst, const double* src, int64_t n, hipStream_t s);
void widen(double* dst, const float* src, int64_t n, hipStream_t s);
// partial sums of x.y into ws->partials then final reduce into scalars[slot].
void dot(gmf_workspace* ws, const double* x, const double* y, int64_t n, int slot);
void read_scalars(gmf_workspace* ws, int count);
// Queue a copy of the first `count` device scalars into ws->pipelined_scalars
// (a buffer no other reader uses) and mark it; wait_scalars blocks until it arrived.
void read_scalars_async(gmf_workspace* ws, int count);
void wait_scalars(gmf_workspace* ws);

}  // namespace gmf
