// Several right-hand sides at once: the reduced operator on a block of
// columns and CG run in lockstep on every column with its own scalars.
//
// The reference solves one right-hand side per solve call (src/solve.jl:376-431
// _cg_solve!); a block of k right-hand sides that share the operator is the
// batched form of those k calls. Columns are stored one after another (column
// j of an n x k block starts at j n). Point-cloud batches (cut cells, surface
// pieces), whose per-point basis evaluation dominates their cost, take every
// column in one launch (cell_apply_cloud_block); the other kernels run once per
// column. Each column's CG iterates follow src/solve.jl:376-431: alpha = rz /
// pq, x += alpha p, r -= alpha q, z = M^-1 r, beta = rz_new / rz, p = z + beta p,
// convergence when |r| <= max(atol, rtol |b|). A converged column is frozen on
// the device (its x, r and residual norm stop changing), so its result is the
// iterate at which it converged, as in the one-column solve.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "blas.hpp"
#include "internal.hpp"

namespace gmf {

void launch_cell_cloud_block(const gmf_plan& plan, const Bucket& b, const double* X, double* Y, int64_t ld, int k,
                             bool reduced, hipStream_t stream);

namespace {

constexpr int kThreads = 256;
constexpr int kMaxParts = 512;  // partial sums per column

// Per-column scalar arrays (k entries each) in one device buffer.
enum BlockSlot { kRz = 0, kPq, kRr, kRzNew, kTol2, kLive, kBb, kBlockSlots };

inline int parts_for(int64_t n) {
  return int(std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, kMaxParts)));
}

__device__ inline double block_total(double v) {
  __shared__ double red[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; ++w) s += red[w];
  return s;
}

__device__ inline bool live(const double* s, int k, int j) { return s == nullptr || s[kLive * k + j] != 0.0; }

// partials[j * g + b] = partial sum of X_j . Y_j (skipped columns write 0).
__global__ void __launch_bounds__(kThreads) kb_dot(const double* __restrict__ X, const double* __restrict__ Y,
                                                   int64_t n, const double* __restrict__ s, int k,
                                                   double* __restrict__ partials) {
  const int j = blockIdx.y, g = gridDim.x;
  double acc = 0.0;
  if (live(s, k, j)) {
    const double* x = X + int64_t(j) * n;
    const double* y = Y + int64_t(j) * n;
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(g) * kThreads)
      acc = fma(x[i], y[i], acc);
  }
  const double t = block_total(acc);
  if (threadIdx.x == 0) partials[int64_t(j) * g + blockIdx.x] = t;
}

// out[slot][j] = sum of column j's partials in a fixed order; columns that
// are no longer live keep their value.
__global__ void __launch_bounds__(kThreads) kb_reduce(const double* __restrict__ partials, int g,
                                                      double* __restrict__ s, int k, int slot, int gated) {
  const int j = blockIdx.x;
  if (gated && !live(s, k, j)) return;
  double acc = 0.0;
  for (int i = threadIdx.x; i < g; i += kThreads) acc += partials[int64_t(j) * g + i];
  const double t = block_total(acc);
  if (threadIdx.x == 0) s[slot * k + j] = t;
}

// x_j += alpha_j p_j; r_j -= alpha_j q_j; partial sums of r_j . r_j.
__global__ void __launch_bounds__(kThreads) kb_cg_update(double* __restrict__ X, double* __restrict__ R,
                                                         const double* __restrict__ P, const double* __restrict__ Q,
                                                         int64_t n, const double* __restrict__ s, int k,
                                                         double* __restrict__ partials) {
  const int j = blockIdx.y, g = gridDim.x;
  double acc = 0.0;
  if (live(s, k, j)) {
    const double alpha = s[kRz * k + j] / s[kPq * k + j];
    double* x = X + int64_t(j) * n;
    double* r = R + int64_t(j) * n;
    const double* p = P + int64_t(j) * n;
    const double* q = Q + int64_t(j) * n;
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(g) * kThreads) {
      x[i] = fma(alpha, p[i], x[i]);
      const double ri = fma(-alpha, q[i], r[i]);
      r[i] = ri;
      acc = fma(ri, ri, acc);
    }
  }
  const double t = block_total(acc);
  if (threadIdx.x == 0) partials[int64_t(j) * g + blockIdx.x] = t;
}

// z_j = D^-1 r_j (FP64 or FP32 inverse diagonal, or the identity) and partial
// sums of r_j . z_j.
__global__ void __launch_bounds__(kThreads) kb_jacobi(double* __restrict__ Z, const double* __restrict__ R,
                                                      const double* __restrict__ inv64, const float* __restrict__ inv32,
                                                      int64_t n, const double* __restrict__ s, int k,
                                                      double* __restrict__ partials) {
  const int j = blockIdx.y, g = gridDim.x;
  double acc = 0.0;
  if (live(s, k, j)) {
    double* z = Z + int64_t(j) * n;
    const double* r = R + int64_t(j) * n;
    for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(g) * kThreads) {
      const double ri = r[i];
      const double zi = inv64 ? inv64[i] * ri : inv32 ? double(inv32[i] * float(ri)) : ri;
      z[i] = zi;
      acc = fma(ri, zi, acc);
    }
  }
  const double t = block_total(acc);
  if (threadIdx.x == 0) partials[int64_t(j) * g + blockIdx.x] = t;
}

// p_j = z_j + beta_j p_j, beta_j = rz_new_j / rz_j.
__global__ void __launch_bounds__(kThreads) kb_direction(double* __restrict__ P, const double* __restrict__ Z,
                                                         int64_t n, const double* __restrict__ s, int k) {
  const int j = blockIdx.y, g = gridDim.x;
  if (!live(s, k, j)) return;
  const double beta = s[kRzNew * k + j] / s[kRz * k + j];
  double* p = P + int64_t(j) * n;
  const double* z = Z + int64_t(j) * n;
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(g) * kThreads)
    p[i] = fma(beta, p[i], z[i]);
}

// mode 0: tol2_j = max(atol, rtol |b_j|)^2 and live_j = rr_j > tol2_j;
// mode 1: live_j &= rr_j > tol2_j; mode 2: rz_j = rz_new_j on live columns.
__global__ void kb_scalars(double* __restrict__ s, int k, int mode, double rtol, double atol) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  if (mode == 0) {
    const double tol = fmax(atol, rtol * sqrt(s[kBb * k + j]));
    s[kTol2 * k + j] = tol * tol;
    s[kLive * k + j] = s[kRr * k + j] > tol * tol ? 1.0 : 0.0;
  } else if (mode == 1) {
    if (s[kLive * k + j] != 0.0 && !(s[kRr * k + j] > s[kTol2 * k + j])) s[kLive * k + j] = 0.0;
  } else if (s[kLive * k + j] != 0.0) {
    s[kRz * k + j] = s[kRzNew * k + j];
  }
}

// z = D^-1 r with the inverse diagonal rounded to FP32 and the product formed
// in FP32 (a single-precision preconditioner inside the FP64 Krylov method).
__global__ void __launch_bounds__(kThreads) k_jacobi32(double* __restrict__ z, const double* __restrict__ r,
                                                       const float* __restrict__ inv, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    z[i] = double(inv[i] * float(r[i]));
}

__global__ void __launch_bounds__(kThreads) k_narrow_inv(float* __restrict__ d, const double* __restrict__ s,
                                                         int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    d[i] = float(s[i]);
}

struct PinnedScalars {
  double* p = nullptr;
  explicit PinnedScalars(size_t n) { GMF_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), n * sizeof(double))); }
  ~PinnedScalars() {
    if (p) (void)hipHostFree(p);
  }
};

}  // namespace

struct JacobiPrecond32 : gmf_preconditioner {
  float* invdiag = nullptr;
  int64_t n = 0;
  ~JacobiPrecond32() override {
    if (invdiag) (void)hipFree(invdiag);
  }
  void apply(const double* r, double* z) override {
    hipLaunchKernelGGL(k_jacobi32, dim3(grid1(n)), dim3(kThreads), 0, ws->stream, z, r, invdiag, n);
    GMF_HIP(hipGetLastError());
  }
};

// Y_j = P^T A P X_j for j < k (columns of length n_red, one after another).
void reduced_apply_block(gmf_workspace* ws, const double* X, double* Y, int k) {
  gmf_plan& p = *ws->plan;
  const int64_t n = p.n_red;
  require(X != Y, GMF_EALIAS, "result and coefficient blocks must not alias");
  if (!p.selection_map) {
    // General reduced maps go through the full layout one column at a time.
    for (int j = 0; j < k; ++j) reduced_apply(ws, X + int64_t(j) * n, Y + int64_t(j) * n);
    return;
  }
  if (!plan_writes_every_dof(p)) fill(Y, 0.0, n * k, ws->stream);
  std::vector<char> block(p.buckets.size(), 0);
  for (int j = 0; j < k; ++j) {
    const double* x = X + int64_t(j) * n;
    double* y = Y + int64_t(j) * n;
    BucketFlags batched(p.buckets.size());
    launch_generic_batches(p, x, y, true, ws->stream, batched.data);
    launch_face_batch(p, x, y, true, ws->stream, batched.data);
    for (size_t i = 0; i < p.buckets.size(); ++i) {
      const Bucket& b = p.buckets[i];
      if (b.cloud) {
        if (batched[i] || launch_face_tensor(p, b, x, y, true, ws->stream)) continue;
        block[i] = 1;  // every column in one launch below
        continue;
      }
      if (b.fast && launch_cell_fast(p, b, x, y, true, ws->brick_slots, ws->stream)) continue;
      if (batched[i]) continue;
      launch_cell_generic(p, b, x, y, true, ws->stream);
    }
  }
  for (size_t i = 0; i < p.buckets.size(); ++i)
    if (block[i]) launch_cell_cloud_block(p, p.buckets[i], X, Y, n, k, true, ws->stream);
}

namespace {

struct BlockVectors {
  DeviceBuffer<double> r, z, p, q, s, parts;
  BlockVectors(int64_t n, int k)
      : r(size_t(n) * k), z(size_t(n) * k), p(size_t(n) * k), q(size_t(n) * k), s(size_t(kBlockSlots) * k),
        parts(size_t(kMaxParts) * k) {}
};

void block_dot(gmf_workspace* ws, const double* X, const double* Y, int64_t n, int k, double* s, int slot,
               bool gated, double* parts) {
  const int g = parts_for(n);
  hipLaunchKernelGGL(kb_dot, dim3(unsigned(g), unsigned(k)), dim3(kThreads), 0, ws->stream, X, Y, n,
                     gated ? s : nullptr, k, parts);
  hipLaunchKernelGGL(kb_reduce, dim3(unsigned(k)), dim3(kThreads), 0, ws->stream, parts, g, s, k, slot, gated ? 1 : 0);
  GMF_HIP(hipGetLastError());
}

// z = M^-1 r on every live column and s[slot] = r.z (gated by liveness when
// `gated`; `known_live` is the host's most recent view, used to skip
// preconditioner applications of columns already known to have converged).
void block_precond(gmf_workspace* ws, gmf_preconditioner* M, BlockVectors& v, int64_t n, int k, int slot, bool gated,
                   const std::vector<char>& known_live) {
  const int g = parts_for(n);
  const double* inv64 = nullptr;
  const float* inv32 = nullptr;
  bool fused = M == nullptr;
  if (auto* J = dynamic_cast<JacobiPrecond*>(M)) {
    inv64 = J->invdiag;
    fused = true;
  } else if (auto* J32 = dynamic_cast<JacobiPrecond32*>(M)) {
    inv32 = J32->invdiag;
    fused = true;
  }
  if (fused) {
    hipLaunchKernelGGL(kb_jacobi, dim3(unsigned(g), unsigned(k)), dim3(kThreads), 0, ws->stream, v.z.p, v.r.p, inv64,
                       inv32, n, gated ? v.s.p : nullptr, k, v.parts.p);
    hipLaunchKernelGGL(kb_reduce, dim3(unsigned(k)), dim3(kThreads), 0, ws->stream, v.parts.p, g, v.s.p, k, slot,
                       gated ? 1 : 0);
    GMF_HIP(hipGetLastError());
    return;
  }
  if (!M->apply_block(v.r.p, v.z.p, n, k)) {
    for (int j = 0; j < k; ++j)
      if (known_live[size_t(j)]) M->apply(v.r.p + int64_t(j) * n, v.z.p + int64_t(j) * n);
  }
  block_dot(ws, v.r.p, v.z.p, n, k, v.s.p, slot, gated, v.parts.p);
}

void scalars_step(gmf_workspace* ws, double* s, int k, int mode, double rtol, double atol) {
  hipLaunchKernelGGL(kb_scalars, dim3(unsigned((k + 63) / 64)), dim3(64), 0, ws->stream, s, k, mode, rtol, atol);
  GMF_HIP(hipGetLastError());
}

void cg_block_run(gmf_workspace* ws, gmf_preconditioner* M, const double* B, double* X, int k, bool use_initial,
                  double rtol, double atol, int64_t maxiter, gmf_solve_info* info) {
  const int64_t n = ws->plan->n_red;
  for (int j = 0; j < k; ++j) info[j] = gmf_solve_info{};
  if (n == 0) {
    for (int j = 0; j < k; ++j) info[j].converged = 1;
    return;
  }
  BlockVectors v(n, k);
  const int64_t nk = n * k;
  const int g = parts_for(n);
  double* s = v.s.p;
  copy(v.r.p, B, nk, ws->stream);
  if (use_initial) {
    reduced_apply_block(ws, X, v.q.p, k);
    axpby(ws, v.r.p, -1.0, v.q.p, 1.0, nk);
  } else {
    fill(X, 0.0, nk, ws->stream);
  }
  std::vector<char> known(size_t(k), 1);
  block_precond(ws, M, v, n, k, kRz, false, known);
  copy(v.p.p, v.z.p, nk, ws->stream);
  block_dot(ws, B, B, n, k, s, kBb, false, v.parts.p);
  block_dot(ws, v.r.p, v.r.p, n, k, s, kRr, false, v.parts.p);
  scalars_step(ws, s, k, 0, rtol, atol);
  PinnedScalars hs(size_t(kBlockSlots) * k);
  GMF_HIP(hipMemcpyAsync(hs.p, s, sizeof(double) * size_t(kBlockSlots) * k, hipMemcpyDeviceToHost, ws->stream));
  sync(ws);
  int remaining = 0;
  for (int j = 0; j < k; ++j) {
    gmf_solve_info& f = info[j];
    f.rhs_norm = std::sqrt(hs.p[kBb * k + j]);
    f.initial_residual_norm = f.final_residual_norm = std::sqrt(hs.p[kRr * k + j]);
    f.operator_applications = use_initial ? 1 : 0;
    f.preconditioner_applications = 1;
    known[size_t(j)] = hs.p[kLive * k + j] != 0.0;
    if (known[size_t(j)])
      ++remaining;
    else
      f.converged = 1;
  }
  for (int64_t it = 1; it <= maxiter && remaining > 0; ++it) {
    reduced_apply_block(ws, v.p.p, v.q.p, k);
    block_dot(ws, v.p.p, v.q.p, n, k, s, kPq, true, v.parts.p);
    hipLaunchKernelGGL(kb_cg_update, dim3(unsigned(g), unsigned(k)), dim3(kThreads), 0, ws->stream, X, v.r.p, v.p.p,
                       v.q.p, n, s, k, v.parts.p);
    hipLaunchKernelGGL(kb_reduce, dim3(unsigned(k)), dim3(kThreads), 0, ws->stream, v.parts.p, g, s, k, int(kRr), 1);
    GMF_HIP(hipGetLastError());
    // Host copy of pq and rr before the liveness update, then freeze converged columns.
    GMF_HIP(hipMemcpyAsync(hs.p, s, sizeof(double) * size_t(kBlockSlots) * k, hipMemcpyDeviceToHost, ws->stream));
    scalars_step(ws, s, k, 1, rtol, atol);
    block_precond(ws, M, v, n, k, kRzNew, true, known);
    hipLaunchKernelGGL(kb_direction, dim3(unsigned(g), unsigned(k)), dim3(kThreads), 0, ws->stream, v.p.p, v.z.p, n,
                       s, k);
    scalars_step(ws, s, k, 2, rtol, atol);
    sync(ws);
    for (int j = 0; j < k; ++j) {
      if (!known[size_t(j)]) continue;
      gmf_solve_info& f = info[j];
      f.iterations = it;
      f.operator_applications++;
      require(hs.p[kPq * k + j] > 0.0, GMF_EBREAKDOWN,
              "CG encountered a non-positive operator direction; provide a different linear_solve for this problem");
      const double rr = hs.p[kRr * k + j];
      f.final_residual_norm = std::sqrt(rr);
      if (!(rr > hs.p[kTol2 * k + j])) {
        f.converged = 1;
        known[size_t(j)] = 0;
        --remaining;
      } else {
        f.preconditioner_applications++;
      }
    }
    if (progress_interval() && it % progress_interval() == 0) {
      double worst = 0.0;
      for (int j = 0; j < k; ++j)
        if (known[size_t(j)]) worst = std::max(worst, std::sqrt(hs.p[kRr * k + j] / hs.p[kTol2 * k + j]));
      std::fprintf(stderr, "gmf block cg: iteration %lld, %d of %d columns live, worst residual / tolerance %.3e\n",
                   (long long)it, remaining, k, worst);
    }
  }
  if (remaining > 0)
    throw Error(GMF_ENOCONV, "CG did not converge in " + std::to_string(maxiter) + " iterations on " +
                                 std::to_string(remaining) + " of " + std::to_string(k) + " columns");
}

}  // namespace
}  // namespace gmf

using namespace gmf;

gmf_status gmf_reduced_apply_block(gmf_workspace* ws, const double* X, double* Y, int32_t k) {
  return guarded_call([&] {
    require(ws && X && Y, GMF_EINVAL, "null argument");
    require(k >= 1 && k <= 4096, GMF_EINVAL, "column count must be in 1..4096");
    reduced_apply_block(ws, X, Y, k);
    sync(ws);
  });
}

gmf_status gmf_cg_solve_block(gmf_workspace* ws, gmf_preconditioner* M, const double* B, double* X, int32_t k,
                              int32_t use_initial, double rtol, double atol, int64_t maxiter, gmf_solve_info* info) {
  return guarded_call([&] {
    require(ws && B && X && info, GMF_EINVAL, "null argument");
    require(k >= 1 && k <= 4096, GMF_EINVAL, "column count must be in 1..4096");
    require(maxiter >= 1, GMF_EINVAL, "maxiter must be positive");
    require(std::isfinite(rtol) && rtol >= 0 && std::isfinite(atol) && atol >= 0, GMF_EINVAL,
            "tolerances must be finite and non-negative");
    require(M == nullptr || M->ws == ws, GMF_EINVAL, "preconditioner belongs to a different workspace");
    require(X != B, GMF_EALIAS, "solution and right-hand side blocks must not alias");
    cg_block_run(ws, M, B, X, k, use_initial != 0, rtol, atol, maxiter, info);
  });
}

gmf_status gmf_precond_create_jacobi_f32(gmf_workspace* ws, const double* invdiag, gmf_preconditioner** out) {
  return guarded_call([&] {
    require(ws && invdiag && out, GMF_EINVAL, "null argument");
    auto j = std::make_unique<JacobiPrecond32>();
    j->ws = ws;
    j->n = ws->plan->n_red;
    j->invdiag = dalloc<float>(size_t(j->n));
    hipLaunchKernelGGL(k_narrow_inv, dim3(grid1(j->n)), dim3(kThreads), 0, ws->stream, j->invdiag, invdiag, j->n);
    GMF_HIP(hipGetLastError());
    sync(ws);
    *out = j.release();
  });
}
