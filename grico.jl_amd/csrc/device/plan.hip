// C ABI of the device library (include/grico_mf.h): plan upload, workspaces,
// operator application, reduced maps, diagonal, Jacobi, CG and FGMRES.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "blas.hpp"
#include "common.hpp"
#include "internal.hpp"

namespace gmf {
namespace {

thread_local std::string g_error;

gmf_status fail(gmf_status s, const std::string& m) {
  g_error = m;
  return s;
}

template <class F>
gmf_status guarded(F&& f) {
  try {
    f();
    return GMF_OK;
  } catch (const Error& e) {
    return fail(e.status, e.what());
  } catch (const std::bad_alloc&) {
    return fail(GMF_EDEVICE, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(GMF_EINVAL, e.what());
  }
}

template <class T>
T* upload(gmf_plan& p, const T* host, size_t n) {
  return upload_array(p, host, n);
}


void check_csr(int64_t rows, const int64_t* off, const int64_t* idx, int64_t bound, const char* what) {
  require(rows == 0 || off != nullptr, GMF_EINVAL, std::string(what) + ": offsets missing");
  if (rows == 0) return;
  require(off[0] == 0, GMF_EINVAL, std::string(what) + ": offsets must start at 0");
  for (int64_t r = 0; r < rows; ++r)
    require(off[r + 1] >= off[r], GMF_EINVAL, std::string(what) + ": offsets must be non-decreasing");
  for (int64_t t = 0; t < off[rows]; ++t)
    require(idx[t] >= 0 && idx[t] < bound, GMF_EINVAL, std::string(what) + ": index out of range");
}

Csr upload_csr(gmf_plan& p, int64_t rows, const int64_t* off, const int64_t* idx, const double* coef) {
  Csr c;
  c.rows = rows;
  if (rows == 0) return c;
  c.offsets = upload(p, off, size_t(rows + 1));
  c.indices = upload(p, idx, size_t(off[rows]));
  c.coefficients = upload(p, coef, size_t(off[rows]));
  return c;
}

// Coupled cells of a bucket for the exact reduced diagonal: a cell whose local
// modes reach one reduced dof through several modes (multi-term hanging /
// p-mismatch expansions, repeated dofs) contributes v_j^T A_c v_j per reduced
// dof j, with v_j the local coefficient vector of j, instead of c^2 A_c,ii.
void build_diagonal_pairs(gmf_plan& p, Bucket& b, const std::vector<int32_t>& ridx, const gmf_bucket_desc& bd,
                          const std::vector<int32_t>& mridx) {
  const int64_t N = b.Ntot;
  std::vector<uint8_t> coupled(size_t(b.cells), 0);
  std::vector<int64_t> pcell, poff{0};
  std::vector<int32_t> pdof, ploc;
  std::vector<double> pcoef;
  bool any = false;
  std::vector<std::pair<int32_t, std::pair<int32_t, double>>> ent;  // (dof, (local, coef))
  for (int64_t c = 0; c < b.cells; ++c) {
    ent.clear();
    bool multi = false;
    for (int64_t i = 0; i < N; ++i) {
      const int32_t g = bd.local_index[c * N + i];
      if (g >= 0) {
        const int32_t r = ridx[size_t(c * N + i)];
        if (r >= 0) ent.push_back({r, {int32_t(i), bd.local_coefficient ? bd.local_coefficient[c * N + i] : 1.0}});
      } else if (g <= -2) {
        multi = true;
        const int64_t m = -int64_t(g) - 2;
        for (int64_t t = bd.multi_offsets[m]; t < bd.multi_offsets[m + 1]; ++t) {
          const int32_t r = mridx[size_t(t)];
          if (r >= 0) ent.push_back({r, {int32_t(i), bd.multi_coefficients[t]}});
        }
      }
    }
    std::stable_sort(ent.begin(), ent.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    bool repeated = false;
    for (size_t k = 1; k < ent.size() && !repeated; ++k) repeated = ent[k].first == ent[k - 1].first;
    if (!multi && !repeated) continue;
    coupled[size_t(c)] = 1;
    any = true;
    for (size_t k = 0; k < ent.size();) {
      size_t e = k;
      pcell.push_back(c);
      pdof.push_back(ent[k].first);
      while (e < ent.size() && ent[e].first == ent[k].first) {
        ploc.push_back(ent[e].second.first);
        pcoef.push_back(ent[e].second.second);
        ++e;
      }
      poff.push_back(int64_t(ploc.size()));
      k = e;
    }
  }
  if (!any) return;
  b.coupled = upload_array(p, coupled.data(), coupled.size());
  b.pair_count = int64_t(pcell.size());
  b.pair_cell = upload_array(p, pcell.data(), pcell.size());
  b.pair_dof = upload_array(p, pdof.data(), pdof.size());
  b.pair_off = upload_array(p, poff.data(), poff.size());
  b.pair_loc = upload_array(p, ploc.data(), ploc.size());
  b.pair_coef = upload_array(p, pcoef.data(), pcoef.size());
}

// ---- small kernels ---------------------------------------------------------

__global__ void k_fixed_rows(double* __restrict__ y, const double* __restrict__ x, const int64_t* __restrict__ rows,
                             int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t d = rows[i];
    y[d] = x[d];
  }
}

// affine: y[p] = x[p] - sum c x[j]; mean: y[p] = sum c x[j].
__global__ void k_linear_rows(double* __restrict__ y, const double* __restrict__ x, const int64_t* __restrict__ pivots,
                              const int64_t* __restrict__ off, const int64_t* __restrict__ idx,
                              const double* __restrict__ coef, int64_t n, int affine) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t p = pivots[i];
    double v = affine ? x[p] : 0.0;
    for (int64_t t = off[i]; t < off[i + 1]; ++t) v = affine ? fma(-coef[t], x[idx[t]], v) : fma(coef[t], x[idx[t]], v);
    y[p] = v;
  }
}

__global__ void k_csr_gather(double* __restrict__ out, const double* __restrict__ in, const int64_t* __restrict__ off,
                             const int64_t* __restrict__ idx, const double* __restrict__ coef,
                             const double* __restrict__ shift, int64_t rows) {
  for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r < rows; r += int64_t(gridDim.x) * blockDim.x) {
    double v = shift ? shift[r] : 0.0;
    for (int64_t t = off[r]; t < off[r + 1]; ++t) v = fma(coef[t], in[idx[t]], v);
    out[r] = v;
  }
}

__global__ void k_fill_int(int* __restrict__ p, int v, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = v;
}

__global__ void k_invert_diagonal(const double* __restrict__ d, double* __restrict__ inv, int64_t n,
                                  int* __restrict__ bad) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const double v = d[i];
    if (!(v != 0.0) || !isfinite(v)) {
      atomicOr(bad, 1);
      inv[i] = 0.0;
    } else {
      inv[i] = 1.0 / v;
    }
  }
}

__global__ void k_scale_dev(double* __restrict__ y, const double* __restrict__ x, const double* __restrict__ s, int slot,
                            double sign, int64_t n) {
  const double a = sign * s[slot];
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    y[i] = fma(a, x[i], y[i]);
}


// ---- operator application ---------------------------------------------------

void cell_pass(gmf_workspace* ws, const double* x, double* y, bool reduced) {
  const gmf_plan& p = *ws->plan;
  // Generic-kernel buckets of uneven degree share one launch per tier.
  BucketFlags batched(p.buckets.size());
  launch_generic_batches(p, x, y, reduced, ws->stream, batched.data);
  launch_face_batch(p, x, y, reduced, ws->stream, batched.data);
  for (size_t i = 0; i < p.buckets.size(); ++i) {
    const Bucket& b = p.buckets[i];
    if (b.cloud) {
      if (batched[i]) continue;
      if (!launch_face_tensor(p, b, x, y, reduced, ws->stream)) launch_cell_cloud(p, b, x, y, reduced, ws->stream);
      continue;
    }
    if (b.fast && launch_cell_fast(p, b, x, y, reduced, ws->brick_slots, ws->stream)) continue;
    if (batched[i]) continue;
    launch_cell_generic(p, b, x, y, reduced, ws->stream);
  }
}

void constraint_rows(gmf_workspace* ws, const double* x, double* y) {
  const gmf_plan& p = *ws->plan;
  if (p.fixed_count)
    hipLaunchKernelGGL(k_fixed_rows, dim3(grid1(p.fixed_count)), dim3(256), 0, ws->stream, y, x, p.fixed,
                       p.fixed_count);
  if (p.affine_count)
    hipLaunchKernelGGL(k_linear_rows, dim3(grid1(p.affine_count)), dim3(256), 0, ws->stream, y, x, p.affine_pivots,
                       p.affine.offsets, p.affine.indices, p.affine.coefficients, p.affine_count, 1);
  if (p.mean_count)
    hipLaunchKernelGGL(k_linear_rows, dim3(grid1(p.mean_count)), dim3(256), 0, ws->stream, y, x, p.mean_pivots,
                       p.mean.offsets, p.mean.indices, p.mean.coefficients, p.mean_count, 0);
  GMF_HIP(hipGetLastError());
}

void apply_full(gmf_workspace* ws, const double* x, double* y) {
  require(x != y, GMF_EALIAS, "result and coefficient vector must not alias");
  if (!plan_writes_every_dof(*ws->plan)) fill(y, 0.0, ws->plan->n_full, ws->stream);
  cell_pass(ws, x, y, false);
  constraint_rows(ws, x, y);
}

void expand(gmf_workspace* ws, const double* xr, double* xf, bool shift) {
  const gmf_plan& p = *ws->plan;
  hipLaunchKernelGGL(k_csr_gather, dim3(grid1(p.n_full)), dim3(256), 0, ws->stream, xf, xr, p.map.offsets,
                     p.map.indices, p.map.coefficients, shift ? p.shift : nullptr, p.n_full);
  GMF_HIP(hipGetLastError());
}

void project(gmf_workspace* ws, const double* yf, double* yr) {
  const gmf_plan& p = *ws->plan;
  if (p.n_red == 0) return;
  hipLaunchKernelGGL(k_csr_gather, dim3(grid1(p.n_red)), dim3(256), 0, ws->stream, yr, yf, p.map_t.offsets,
                     p.map_t.indices, p.map_t.coefficients, nullptr, p.n_red);
  GMF_HIP(hipGetLastError());
}

}  // namespace

gmf_status record_error(gmf_status s, const std::string& message) { return fail(s, message); }

void gmf_preconditioner_apply_dot_default(gmf_preconditioner* M, const double* r, double* z, int64_t n, int slot) {
  M->apply(r, z);
  dot(M->ws, r, z, n, slot);
}

void reduced_apply(gmf_workspace* ws, const double* xr, double* yr) {
  gmf_plan& p = *ws->plan;
  require(xr != yr, GMF_EALIAS, "result and reduced coefficient vector must not alias");
  if (p.selection_map) {
    // P is a 0/1 selection: the cell kernels gather from and scatter to the
    // reduced vector directly through reduced index maps.
    if (!plan_writes_every_dof(p)) fill(yr, 0.0, p.n_red, ws->stream);
    cell_pass(ws, xr, yr, true);
    return;
  }
  expand(ws, xr, ws->full_in, false);
  fill(ws->full_out, 0.0, p.n_full, ws->stream);
  cell_pass(ws, ws->full_in, ws->full_out, false);
  project(ws, ws->full_out, yr);
}

void prepare_fp32_workspace(gmf_workspace* ws) {
  gmf_plan& p = *ws->plan;
  if (prepare_fp32(p)) return;
  if (!ws->red64_in) ws->red64_in = dalloc<double>(size_t(std::max<int64_t>(p.n_red, 1)));
  if (!ws->red64_out) ws->red64_out = dalloc<double>(size_t(std::max<int64_t>(p.n_red, 1)));
}

void reduced_apply_f32(gmf_workspace* ws, const float* xr, float* yr) {
  gmf_plan& p = *ws->plan;
  require(xr != yr, GMF_EALIAS, "result and reduced coefficient vector must not alias");
  require(p.fp32_state != 0, GMF_EINVAL, "workspace is not prepared for FP32 applies");
  if (p.fp32_state > 0) {
    if (!plan_writes_every_dof(p)) fill(yr, 0.0f, p.n_red, ws->stream);
    for (const Bucket& b : p.buckets)
      require(launch_cell_fast32(p, b, xr, yr, true, reinterpret_cast<float*>(ws->brick_slots), ws->stream),
              GMF_EINVAL, "bucket without FP32 kernel");
    return;
  }
  // Buckets without FP32 kernels: FP64 apply of the widened input, rounded.
  require(ws->red64_in && ws->red64_out, GMF_EINVAL, "workspace is not prepared for FP32 applies");
  widen(ws->red64_in, xr, p.n_red, ws->stream);
  reduced_apply(ws, ws->red64_in, ws->red64_out);
  narrow(yr, ws->red64_out, p.n_red, ws->stream);
}

void reduced_diagonal(gmf_workspace* ws, double* diag) {
  const gmf_plan& p = *ws->plan;
  require(p.selection_map, GMF_EUNSUPPORTED,
          "exact reduced diagonal needs a selection reduced map (affine / mean rows are not implemented)");
  fill(diag, 0.0, p.n_red, ws->stream);
  for (const Bucket& b : p.buckets) {
    if (b.cloud) {
      require(!b.coupled, GMF_EUNSUPPORTED,
              "exact reduced diagonal of point-cloud items with multi-term expansions is not implemented");
      launch_cell_cloud_diagonal(p, b, diag, true, ws->stream);
      continue;
    }
    if (b.coupled) {
      launch_cell_diagonal_generic_skip(p, b, diag, ws->stream);
      launch_cell_diagonal_pairs(p, b, diag, ws->stream);
    } else {
      launch_cell_diagonal_generic(p, b, diag, true, ws->stream);
    }
  }
}

void invert_diagonal(gmf_workspace* ws, const double* d, double* inv, int64_t n) {
  DeviceBuffer<int> bad(static_cast<size_t>(1));
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ws->stream, bad.p, 0, int64_t(1));
  hipLaunchKernelGGL(k_invert_diagonal, dim3(grid1(n)), dim3(256), 0, ws->stream, d, inv, n, bad.p);
  GMF_HIP(hipGetLastError());
  int h = 0;
  GMF_HIP(hipMemcpyAsync(&h, bad.p, sizeof(int), hipMemcpyDeviceToHost, ws->stream));
  GMF_HIP(hipStreamSynchronize(ws->stream));
  require(h == 0, GMF_EDIAGONAL, "Jacobi preconditioner requires a nonzero finite diagonal");
}

void sync(gmf_workspace* ws) { GMF_HIP(hipStreamSynchronize(ws->stream)); }

int progress_interval() {
  static const int every = [] {
    const char* e = std::getenv("GMF_PROGRESS");
    const int v = e ? std::atoi(e) : 0;
    return v > 0 ? v : 0;
  }();
  return every;
}

// ---- preconditioners ---------------------------------------------------------

void JacobiPrecond::apply(const double* r, double* z) { jacobi_dot(ws, z, r, invdiag, n, kTmp); }
void JacobiPrecond::apply_dot(const double* r, double* z, int64_t nn, int slot) {
  jacobi_dot(ws, z, r, invdiag, nn, slot);
}

void precond_apply_dot(gmf_workspace* ws, gmf_preconditioner* M, const double* r, double* z, int64_t n, int slot) {
  if (M == nullptr) {
    jacobi_dot(ws, z, r, nullptr, n, slot);
  } else {
    M->apply_dot(r, z, n, slot);
  }
}

void precond_apply(gmf_workspace* ws, gmf_preconditioner* M, const double* r, double* z, int64_t n) {
  if (M == nullptr)
    copy(z, r, n, ws->stream);
  else
    M->apply(r, z);
}

void cg_run(gmf_workspace* ws, gmf_preconditioner* M, const double* b, double* x, bool use_initial, double rtol,
            double atol, int64_t maxiter, gmf_solve_info& inf, CgVectors& v) {
  const int64_t n = ws->plan->n_red;
  inf = gmf_solve_info{};
  if (n == 0) {
    inf.converged = 1;
    return;
  }
  copy(v.r.p, b, n, ws->stream);
  if (use_initial) {
    reduced_apply(ws, x, v.q.p);
    inf.operator_applications++;
    axpby(ws, v.r.p, -1.0, v.q.p, 1.0, n);
  } else {
    fill(x, 0.0, n, ws->stream);
  }
  precond_apply_dot(ws, M, v.r.p, v.z.p, n, kRz);
  inf.preconditioner_applications++;
  copy(v.p.p, v.z.p, n, ws->stream);
  dot(ws, b, b, n, kBb);
  dot(ws, v.r.p, v.r.p, n, kRr);
  read_scalars(ws, kTmp);
  const double* hs = ws->host_scalars;
  inf.rhs_norm = std::sqrt(hs[kBb]);
  const double tol = std::max(atol, rtol * inf.rhs_norm);
  inf.initial_residual_norm = std::sqrt(hs[kRr]);
  inf.final_residual_norm = inf.initial_residual_norm;
  if (hs[kRr] <= tol * tol) {
    inf.converged = 1;
    return;
  }
  const double* ps = ws->pipelined_scalars;
  for (int64_t it = 1; it <= maxiter; ++it) {
    reduced_apply(ws, v.p.p, v.q.p);
    inf.operator_applications++;
    dot(ws, v.p.p, v.q.p, n, kPq);
    cg_update(ws, x, v.r.p, v.p.p, v.q.p, n, kRz, kPq, kRr);
    read_scalars_async(ws, kTmp);
    // The next preconditioner application and search direction write only z,
    // p and their scalars, so they are queued before the host looks at this
    // iteration's residual: the device keeps working during the check. x and
    // r, hence the result and the iteration count, are those of the
    // unpipelined loop.
    precond_apply_dot(ws, M, v.r.p, v.z.p, n, kRzNew);
    direction(ws, v.p.p, v.z.p, n, kRzNew, kRz);
    wait_scalars(ws);
    inf.iterations = it;
    require(ps[kPq] > 0.0, GMF_EBREAKDOWN,
            "CG encountered a non-positive operator direction; provide a different linear_solve for this problem");
    inf.final_residual_norm = std::sqrt(ps[kRr]);
    if (progress_interval() && it % progress_interval() == 0)
      std::fprintf(stderr, "gmf cg: iteration %lld, residual %.3e of tolerance %.3e\n", (long long)it,
                   inf.final_residual_norm, tol);
    if (ps[kRr] <= tol * tol) {
      inf.converged = 1;
      sync(ws);
      return;
    }
    inf.preconditioner_applications++;
  }
  sync(ws);
  throw Error(GMF_ENOCONV, "CG did not converge in " + std::to_string(maxiter) + " iterations");
}

// ---- workspace vectors for solvers ------------------------------------------

struct SolverVectors {
  double* r = nullptr;
  double* z = nullptr;
  double* p = nullptr;
  double* q = nullptr;
  double* V = nullptr;
  double* Z = nullptr;
  double* ydev = nullptr;
  ~SolverVectors() {
    for (double* v : {r, z, p, q, V, Z, ydev})
      if (v) (void)hipFree(v);
  }
};

}  // namespace gmf

void gmf_preconditioner::apply_dot(const double* r, double* z, int64_t n, int slot) {
  gmf::gmf_preconditioner_apply_dot_default(this, r, z, n, slot);
}

using namespace gmf;

namespace gmf {

void arnoldi_mgs(gmf_workspace* ws, double* q, const double* V, int64_t n, int col) {
  for (int bcol = 0; bcol <= col; ++bcol) {
    const double* vb = V + int64_t(bcol) * n;
    dot(ws, q, vb, n, kHessian + bcol);
    hipLaunchKernelGGL(k_scale_dev, dim3(grid1(n)), dim3(256), 0, ws->stream, q, vb, ws->scalars, kHessian + bcol,
                       -1.0, n);
  }
  dot(ws, q, q, n, kHessian + col + 1);
  read_scalars(ws, kHessian + col + 2);
}

void fgmres_run(gmf_workspace* ws, gmf_preconditioner* M, const double* b, double* x, bool use_initial, int m,
                double rtol, double atol, int64_t maxiter, gmf_solve_info& inf, FgmresVectors& v) {
  const int64_t n = ws->plan->n_red;
  inf = gmf_solve_info{};
  if (n == 0) {
    inf.converged = 1;
    return;
  }
  require(m >= 1 && m <= v.m && n == v.n, GMF_EINVAL, "FGMRES vectors do not match the restart length");
  require(m + 1 < kScalarCount - kHessian, GMF_EINVAL, "restart too large");
  copy(v.r.p, b, n, ws->stream);
  if (use_initial) {
    reduced_apply(ws, x, v.q.p);
    inf.operator_applications++;
    axpby(ws, v.r.p, -1.0, v.q.p, 1.0, n);
  } else {
    fill(x, 0.0, n, ws->stream);
  }
  dot(ws, b, b, n, kBb);
  dot(ws, v.r.p, v.r.p, n, kRr);
  read_scalars(ws, kTmp);
  const double* hs = ws->host_scalars;
  inf.rhs_norm = std::sqrt(hs[kBb]);
  const double tol = std::max(atol, rtol * inf.rhs_norm);
  inf.initial_residual_norm = std::sqrt(hs[kRr]);
  inf.final_residual_norm = inf.initial_residual_norm;
  if (inf.initial_residual_norm <= tol) {
    inf.converged = 1;
    return;
  }
  const size_t sm = static_cast<size_t>(m);
  std::vector<double> H((sm + 1) * sm), g(sm + 1), cs(sm), sn(sm), y(sm);
  auto Hm = [&](int r, int c) -> double& { return H[size_t(r * m + c)]; };
  auto residual_norm = [&]() {
    dot(ws, v.r.p, v.r.p, n, kRr);
    read_scalars(ws, kTmp);
    return std::sqrt(hs[kRr]);
  };
  auto finish = [&](int count) {
    for (int row = count - 1; row >= 0; --row) {
      double val = g[size_t(row)];
      for (int c = row + 1; c < count; ++c) val -= Hm(row, c) * y[size_t(c)];
      double dg = Hm(row, row);
      require(std::isfinite(dg), GMF_EINVAL, "FGMRES encountered a non-finite Krylov least-squares diagonal");
      require(dg != 0.0, GMF_EINVAL, "FGMRES encountered a singular Krylov least-squares problem");
      y[size_t(row)] = val / dg;
    }
    GMF_HIP(hipMemcpyAsync(v.ydev.p, y.data(), sizeof(double) * size_t(count), hipMemcpyHostToDevice, ws->stream));
    multi_axpy(ws, x, v.Z.p, n, v.ydev.p, count, n, 1.0);
    reduced_apply(ws, x, v.q.p);
    inf.operator_applications++;
    copy(v.r.p, b, n, ws->stream);
    axpby(ws, v.r.p, -1.0, v.q.p, 1.0, n);
    sync(ws);
  };
  int64_t total = 0;
  while (total < maxiter) {
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    std::fill(cs.begin(), cs.end(), 0.0);
    std::fill(sn.begin(), sn.end(), 0.0);
    double rn = residual_norm();
    inf.final_residual_norm = rn;
    if (rn <= tol) {
      inf.converged = 1;
      return;
    }
    fill(v.V.p, 0.0, n, ws->stream);
    axpby(ws, v.V.p, 1.0 / rn, v.r.p, 0.0, n);
    g[0] = rn;
    int inner_limit = int(std::min<int64_t>(m, maxiter - total));
    int inner = 0;
    bool updated = false;
    for (int col = 0; col < inner_limit; ++col) {
      total++;
      inner = col + 1;
      inf.iterations = total;
      double* zc = v.Z.p + int64_t(col) * n;
      double* vc = v.V.p + int64_t(col) * n;
      precond_apply(ws, M, vc, zc, n);
      inf.preconditioner_applications++;
      reduced_apply(ws, zc, v.q.p);
      inf.operator_applications++;
      arnoldi_mgs(ws, v.q.p, v.V.p, n, col);
      for (int bcol = 0; bcol <= col; ++bcol) Hm(bcol, col) = hs[kHessian + bcol];
      const double next = std::sqrt(hs[kHessian + col + 1]);
      Hm(col + 1, col) = next;
      if (next > 0.0 && col + 1 < m + 1) {
        double* vn = v.V.p + int64_t(col + 1) * n;
        fill(vn, 0.0, n, ws->stream);
        axpby(ws, vn, 1.0 / next, v.q.p, 0.0, n);
      }
      for (int row = 0; row < col; ++row) {
        double f = Hm(row, col), s2 = Hm(row + 1, col);
        Hm(row, col) = cs[size_t(row)] * f + sn[size_t(row)] * s2;
        Hm(row + 1, col) = -sn[size_t(row)] * f + cs[size_t(row)] * s2;
      }
      double scale = std::hypot(Hm(col, col), Hm(col + 1, col));
      double c = 1.0, s = 0.0;
      if (scale != 0.0) {
        c = Hm(col, col) / scale;
        s = Hm(col + 1, col) / scale;
      }
      cs[size_t(col)] = c;
      sn[size_t(col)] = s;
      {
        double f = Hm(col, col), s2 = Hm(col + 1, col);
        Hm(col, col) = c * f + s * s2;
        Hm(col + 1, col) = -s * f + c * s2;
        double gf = g[size_t(col)], gs = g[size_t(col) + 1];
        g[size_t(col)] = c * gf + s * gs;
        g[size_t(col) + 1] = -s * gf + c * gs;
      }
      if (std::fabs(g[size_t(col) + 1]) <= tol || next == 0.0) {
        finish(col + 1);
        double fr = residual_norm();
        inf.final_residual_norm = fr;
        if (fr <= tol) {
          inf.converged = 1;
          return;
        }
        updated = true;
        break;
      }
    }
    require(inner > 0, GMF_EINVAL, "FGMRES failed to build a Krylov update");
    if (!updated) finish(inner);
    double fr = residual_norm();
    inf.final_residual_norm = fr;
    if (fr <= tol) {
      inf.converged = 1;
      return;
    }
  }
  throw Error(GMF_ENOCONV, "FGMRES did not converge in " + std::to_string(maxiter) + " iterations");
}

}  // namespace gmf

extern "C" {

const char* gmf_last_error(void) { return g_error.c_str(); }

gmf_status gmf_device_info(char* name, int32_t* cus, int64_t* mem) {
  return guarded([&] {
    int dev = 0;
    GMF_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    GMF_HIP(hipGetDeviceProperties(&prop, dev));
    if (name) {
      std::strncpy(name, prop.gcnArchName, 63);
      name[63] = 0;
    }
    if (cus) *cus = prop.multiProcessorCount;
    if (mem) *mem = int64_t(prop.totalGlobalMem);
  });
}

gmf_status gmf_plan_create(const gmf_plan_desc* desc, gmf_plan** out) {
  return guarded([&] {
    require(desc != nullptr && out != nullptr, GMF_EINVAL, "null descriptor");
    require(desc->dim >= 1 && desc->dim <= 3, GMF_EINVAL, "dimension must be 1, 2 or 3");
    require(desc->full_dof_count >= 0 && desc->full_dof_count < (int64_t(1) << 31), GMF_EINVAL,
            "full dof count must fit int32 indices");
    auto plan = std::make_unique<gmf_plan>();
    gmf_plan& p = *plan;
    GMF_HIP(hipGetDevice(&p.device));
    p.dim = desc->dim;
    p.n_full = desc->full_dof_count;
    p.n_red = desc->reduced_dof_count;
    require(p.n_red >= 0 && p.n_red <= p.n_full, GMF_EINVAL, "reduced dof count out of range");
    try {
      // Reduced map (validated), selection detection, transpose.
      check_csr(p.n_full, desc->map_offsets, desc->map_indices, p.n_red, "reduced map");
      std::vector<int64_t> red_of_full(size_t(p.n_full), -1);
      bool selection = true;
      std::vector<int64_t> seen(size_t(p.n_red), 0);
      for (int64_t d = 0; d < p.n_full; ++d) {
        int64_t a = desc->map_offsets[d], e = desc->map_offsets[d + 1];
        if (e - a > 1) selection = false;
        if (e - a == 1) {
          if (desc->map_coefficients[a] != 1.0) selection = false;
          red_of_full[size_t(d)] = desc->map_indices[a];
          seen[size_t(desc->map_indices[a])]++;
        }
      }
      for (int64_t r = 0; r < p.n_red && selection; ++r) selection = seen[size_t(r)] == 1;
      p.selection_map = selection;
      p.map = upload_csr(p, p.n_full, desc->map_offsets, desc->map_indices, desc->map_coefficients);
      p.shift = upload(p, desc->shift, size_t(p.n_full));
      {
        std::vector<int64_t> toff(size_t(p.n_red) + 1, 0), tidx(size_t(p.n_full ? desc->map_offsets[p.n_full] : 0));
        std::vector<double> tcoef(tidx.size());
        for (int64_t d = 0; d < p.n_full; ++d)
          for (int64_t t = desc->map_offsets[d]; t < desc->map_offsets[d + 1]; ++t) toff[size_t(desc->map_indices[t]) + 1]++;
        for (int64_t r = 0; r < p.n_red; ++r) toff[size_t(r) + 1] += toff[size_t(r)];
        std::vector<int64_t> fillp(toff.begin(), toff.end() - 1);
        for (int64_t d = 0; d < p.n_full; ++d)
          for (int64_t t = desc->map_offsets[d]; t < desc->map_offsets[d + 1]; ++t) {
            int64_t r = desc->map_indices[t];
            tidx[size_t(fillp[size_t(r)])] = d;
            tcoef[size_t(fillp[size_t(r)]++)] = desc->map_coefficients[t];
          }
        p.map_t = upload_csr(p, p.n_red, toff.data(), tidx.data(), tcoef.data());
      }
      // Constraint rows.
      p.fixed_count = desc->fixed_count;
      for (int64_t i = 0; i < p.fixed_count; ++i)
        require(desc->fixed_dofs[i] >= 0 && desc->fixed_dofs[i] < p.n_full, GMF_EINVAL, "fixed dof out of range");
      p.fixed = upload(p, desc->fixed_dofs, size_t(p.fixed_count));
      p.affine_count = desc->affine_count;
      if (p.affine_count) {
        check_csr(p.affine_count, desc->affine_offsets, desc->affine_indices, p.n_full, "affine rows");
        for (int64_t i = 0; i < p.affine_count; ++i)
          require(desc->affine_pivots[i] >= 0 && desc->affine_pivots[i] < p.n_full, GMF_EINVAL, "affine pivot out of range");
        p.affine_pivots = upload(p, desc->affine_pivots, size_t(p.affine_count));
        p.affine = upload_csr(p, p.affine_count, desc->affine_offsets, desc->affine_indices, desc->affine_coefficients);
      }
      p.mean_count = desc->mean_count;
      if (p.mean_count) {
        check_csr(p.mean_count, desc->mean_offsets, desc->mean_indices, p.n_full, "mean rows");
        for (int64_t i = 0; i < p.mean_count; ++i)
          require(desc->mean_pivots[i] >= 0 && desc->mean_pivots[i] < p.n_full, GMF_EINVAL, "mean pivot out of range");
        p.mean_pivots = upload(p, desc->mean_pivots, size_t(p.mean_count));
        p.mean = upload_csr(p, p.mean_count, desc->mean_offsets, desc->mean_indices, desc->mean_coefficients);
      }
      // Cell batches.
      require(desc->bucket_count >= 0 && (desc->bucket_count == 0 || desc->buckets), GMF_EINVAL, "bucket array missing");
      for (int bi = 0; bi < desc->bucket_count; ++bi) {
        const gmf_bucket_desc& bd = desc->buckets[bi];
        Bucket b;
        require(bd.dim == p.dim, GMF_EINVAL, "bucket dimension mismatch");
        b.dim = bd.dim;
        b.form = bd.form;
        require(b.form == GMF_FORM_DIFFUSION || b.form == GMF_FORM_GENERAL, GMF_EINVAL, "unknown action form");
        b.cells = bd.cell_count;
        require(b.cells >= 0 && b.cells < (int64_t(1) << 31), GMF_EINVAL, "cell count out of range");
        int off = 0;
        b.cloud = bd.point_offsets != nullptr;
        require(bd.sides >= 0 && bd.sides <= 2, GMF_EINVAL, "sides must be 0, 1 or 2");
        b.sides = bd.sides == 2 ? 2 : 1;
        require(b.sides == 1 || b.cloud, GMF_EINVAL, "interface items must be point-cloud batches");
        for (int a = 0; a < b.dim; ++a) {
          b.N[a] = bd.modes[a];
          b.Q[a] = b.cloud ? 1 : bd.points[a];
          require(b.N[a] >= 1 && b.N[a] <= kMaxExtent && b.Q[a] >= 1 && b.Q[a] <= kMaxExtent, GMF_EUNSUPPORTED,
                  "per-axis mode and point counts must lie in 1..16");
          b.Ntot *= b.N[a];
          b.Qtot *= b.Q[a];
          b.table_offset[a] = off;
          off += b.N[a] * b.Q[a];
        }
        if (b.sides == 2) {
          int nplus = 1;
          for (int a = 0; a < b.dim; ++a) {
            b.Np[a] = bd.plus_modes[a];
            require(b.Np[a] >= 1 && b.Np[a] <= kMaxExtent, GMF_EUNSUPPORTED,
                    "per-axis mode counts must lie in 1..16");
            nplus *= b.Np[a];
          }
          b.Ntot += nplus;
        }
        b.point_stride = b.form == GMF_FORM_DIFFUSION ? 1 : (b.sides * (b.dim + 1)) * (b.sides * (b.dim + 1));
        if (b.cloud) {
          require(b.form == GMF_FORM_GENERAL, GMF_EINVAL, "point-cloud batches carry the general point action");
          require(bd.ref_points != nullptr, GMF_EINVAL, "point-cloud batch without reference points");
          const int64_t* po = bd.point_offsets;
          require(po[0] == 0, GMF_EINVAL, "point offsets must start at 0");
          for (int64_t c = 0; c < b.cells; ++c) {
            require(po[c + 1] >= po[c], GMF_EINVAL, "point offsets must be non-decreasing");
            b.max_points = std::max<int>(b.max_points, int(po[c + 1] - po[c]));
          }
          b.total_points = po[b.cells];
          for (int64_t k = 0; k < b.total_points * b.sides * b.dim; ++k)
            require(std::isfinite(bd.ref_points[k]) && std::fabs(bd.ref_points[k]) <= 1.0 + 1e-12, GMF_EINVAL,
                    "reference points must lie in the biunit cell");
          b.poff = upload(p, po, size_t(b.cells) + 1);
          b.rpts = upload(p, bd.ref_points, size_t(b.total_points) * size_t(b.sides) * size_t(b.dim));
          require(b.cells <= INT32_MAX, GMF_EINVAL, "too many point-cloud items in one batch");
          b.h_poff.assign(po, po + b.cells + 1);
          std::vector<int32_t> uitem;
          std::vector<int64_t> ustart;
          b.units = cloud_units(b, b.unit_points, uitem, ustart);
          b.unit_item = upload(p, uitem.data(), uitem.size());
          b.unit_start = upload(p, ustart.data(), ustart.size());
        } else {
          b.h_values.assign(bd.values, bd.values + off);
          b.h_derivs.assign(bd.derivatives, bd.derivatives + off);
          b.values = upload(p, bd.values, size_t(off));
          b.derivs = upload(p, bd.derivatives, size_t(off));
        }
        if (b.cells == 0) {
          p.buckets.push_back(std::move(b));
          continue;
        }
        const size_t nmode = size_t(b.cells) * size_t(b.Ntot);
        const size_t njinv = size_t(b.cells) * size_t(b.sides) * size_t(b.dim);
        for (size_t i = 0; i < njinv; ++i)
          require(std::isfinite(bd.inverse_jacobian[i]), GMF_EINVAL, "non-finite inverse Jacobian");
        b.jinv = upload(p, bd.inverse_jacobian, njinv);
        b.pdata = upload(p, bd.point_data,
                         (b.cloud ? size_t(b.total_points) : size_t(b.cells) * size_t(b.Qtot)) * size_t(b.point_stride));
        b.multi_count = bd.multi_count;
        for (size_t i = 0; i < nmode; ++i) {
          int32_t g = bd.local_index[i];
          require(g >= -1 ? g < p.n_full : (-int64_t(g) - 2) < b.multi_count, GMF_EINVAL, "local index out of range");
        }
        if (b.multi_count) {
          std::vector<int64_t> tmp(bd.multi_indices, bd.multi_indices + bd.multi_offsets[b.multi_count]);
          check_csr(b.multi_count, bd.multi_offsets, tmp.data(), p.n_full, "multi-term map");
          p.multi_terms = true;
        }
        b.lidx = upload(p, bd.local_index, nmode);
        b.lcoef = upload(p, bd.local_coefficient, nmode);
        if (b.multi_count) {
          b.moff = upload(p, bd.multi_offsets, size_t(b.multi_count) + 1);
          b.midx = upload(p, bd.multi_indices, size_t(bd.multi_offsets[b.multi_count]));
          b.mcoef = upload(p, bd.multi_coefficients, size_t(bd.multi_offsets[b.multi_count]));
        }
        if (p.selection_map) {
          std::vector<int32_t> r(nmode);
          for (size_t i = 0; i < nmode; ++i) {
            int32_t g = bd.local_index[i];
            r[i] = g >= 0 ? int32_t(red_of_full[size_t(g)]) : g;
          }
          b.ridx = upload(p, r.data(), nmode);
          std::vector<int32_t> mr;
          if (b.multi_count) {
            int64_t nt = bd.multi_offsets[b.multi_count];
            mr.resize(static_cast<size_t>(nt));
            for (int64_t t = 0; t < nt; ++t) mr[size_t(t)] = int32_t(red_of_full[size_t(bd.multi_indices[t])]);
            b.mridx = upload(p, mr.data(), size_t(nt));
          }
          build_diagonal_pairs(p, b, r, bd, mr);
        }
        if (!b.cloud)
          prepare_fast_path(p, b, bd.local_index, p.selection_map ? red_of_full : std::vector<int64_t>(),
                            bd.brick_cells, bd.brick_count);
        else prepare_face_path(p, b, bd);
        p.buckets.push_back(std::move(b));
      }
      GMF_HIP(hipDeviceSynchronize());
      tune_plan(p);
      prepare_generic_batch(p);
      prepare_face_batch(p);
      GMF_HIP(hipDeviceSynchronize());
    } catch (...) {
      for (void* a : p.allocations) (void)hipFree(a);
      throw;
    }
    *out = plan.release();
  });
}

gmf_status gmf_plan_destroy(gmf_plan* plan) {
  return guarded([&] {
    if (!plan) return;
    for (void* a : plan->allocations) (void)hipFree(a);
    delete plan;
  });
}

gmf_status gmf_plan_sizes(const gmf_plan* plan, int64_t* nf, int64_t* nr, int64_t* bytes) {
  return guarded([&] {
    require(plan != nullptr, GMF_EINVAL, "null plan");
    if (nf) *nf = plan->n_full;
    if (nr) *nr = plan->n_red;
    if (bytes) *bytes = plan->device_bytes;
  });
}

gmf_status gmf_plan_describe(const gmf_plan* plan, char* buf, int64_t capacity) {
  return guarded([&] {
    require(plan && buf && capacity > 0, GMF_EINVAL, "invalid argument");
    std::string s = "full_dofs=" + std::to_string(plan->n_full) + " reduced_dofs=" + std::to_string(plan->n_red) +
                    " selection_map=" + std::to_string(int(plan->selection_map)) +
                    " device_bytes=" + std::to_string(plan->device_bytes) + "\n";
    for (size_t i = 0; i < plan->buckets.size(); ++i) {
      const Bucket& b = plan->buckets[i];
      s += "bucket " + std::to_string(i) + ": dim=" + std::to_string(b.dim) + " N=" + std::to_string(b.N[0]) +
           " Q=" + std::to_string(b.Q[0]) + " cells=" + std::to_string(b.cells) + " form=" + std::to_string(b.form) +
           " multi=" + std::to_string(b.multi_count) + " path=" +
           (b.cloud ? (b.sides == 2 ? "interface points=" : "cloud points=") + std::to_string(b.total_points) +
                          (b.face.enabled ? " face-tensor nnz=" + std::to_string(b.face.nnz) : std::string())
                    : (b.bricks.enabled ? "brick" : (b.fast ? "fast-cell" : "generic")));
      if (b.bricks.enabled) {
        s += " bricks=" + std::to_string(b.bricks.bricks) + " halo_slots=" + std::to_string(b.bricks.halo_slots) +
             " shared_dofs=" + std::to_string(b.bricks.shared_count);
      }
      if (!b.cloud && !b.fast) s += " lanes=" + std::to_string(b.generic_threads);
      s += "\n";
    }
    for (const GenericBatch& g : plan->generic_batches)
      s += "generic batch: lanes=" + std::to_string(g.threads) + " buckets=" + std::to_string(g.count) +
           " cells=" + std::to_string(g.cells) + "\n";
    size_t n = std::min(s.size(), size_t(capacity - 1));
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  });
}

gmf_status gmf_workspace_create(gmf_plan* plan, gmf_workspace** out) {
  return guarded([&] {
    require(plan != nullptr && out != nullptr, GMF_EINVAL, "null plan");
    auto ws = std::make_unique<gmf_workspace>();
    ws->plan = plan;
    ws->full_in = dalloc<double>(size_t(plan->n_full));
    ws->full_out = dalloc<double>(size_t(plan->n_full));
    ws->partial_capacity = 64 * 256 + 4096;
    ws->partials = dalloc<double>(size_t(ws->partial_capacity));
    ws->scalars = dalloc<double>(size_t(kScalarCount));
    if (plan->brick_slots > 0) ws->brick_slots = dalloc<double>(size_t(plan->brick_slots));
    GMF_HIP(hipHostMalloc(reinterpret_cast<void**>(&ws->host_scalars), sizeof(double) * kScalarCount));
    GMF_HIP(hipHostMalloc(reinterpret_cast<void**>(&ws->pipelined_scalars), sizeof(double) * kScalarCount));
    GMF_HIP(hipEventCreateWithFlags(&ws->scalar_event, hipEventDisableTiming));
    *out = ws.release();
  });
}

gmf_status gmf_workspace_destroy(gmf_workspace* ws) {
  return guarded([&] {
    if (!ws) return;
    for (double* v : {ws->full_in, ws->full_out, ws->partials, ws->scalars, ws->brick_slots, ws->red64_in, ws->red64_out})
      if (v) (void)hipFree(v);
    if (ws->host_scalars) (void)hipHostFree(ws->host_scalars);
    if (ws->pipelined_scalars) (void)hipHostFree(ws->pipelined_scalars);
    if (ws->scalar_event) (void)hipEventDestroy(ws->scalar_event);
    delete ws;
  });
}

gmf_status gmf_workspace_set_stream(gmf_workspace* ws, void* stream) {
  return guarded([&] {
    require(ws != nullptr, GMF_EINVAL, "null workspace");
    ws->stream = static_cast<hipStream_t>(stream);
  });
}

gmf_status gmf_apply_async(gmf_workspace* ws, const double* x, double* y) {
  return guarded([&] {
    require(ws && x && y, GMF_EINVAL, "null argument");
    apply_full(ws, x, y);
  });
}

gmf_status gmf_apply(gmf_workspace* ws, const double* x, double* y) {
  return guarded([&] {
    require(ws && x && y, GMF_EINVAL, "null argument");
    apply_full(ws, x, y);
    sync(ws);
  });
}

gmf_status gmf_apply_physical(gmf_workspace* ws, const double* x, double* y) {
  return guarded([&] {
    require(ws && x && y, GMF_EINVAL, "null argument");
    require(x != y, GMF_EALIAS, "result and coefficient vector must not alias");
    if (!plan_writes_every_dof(*ws->plan)) fill(y, 0.0, ws->plan->n_full, ws->stream);
    cell_pass(ws, x, y, false);
    sync(ws);
  });
}

gmf_status gmf_rhs_physical(gmf_workspace* ws, const double* loads, int32_t channels, double* y) {
  return guarded([&] {
    require(ws && loads && y, GMF_EINVAL, "null argument");
    const gmf_plan& p = *ws->plan;
    require_device_pointer(loads, "loads");
    require_device_pointer(y, "y");
    fill(y, 0.0, p.n_full, ws->stream);
    int64_t off = 0;
    for (const Bucket& b : p.buckets) {
      if (b.cloud) continue;  // point-cloud loads: gmf_rhs_faces
      launch_cell_rhs(b, loads + off, channels, y, ws->stream);
      off += b.cells * int64_t(b.Qtot) * channels;
    }
    sync(ws);
  });
}

gmf_status gmf_rhs_faces(gmf_workspace* ws, const double* loads, double* y) {
  return guarded([&] {
    require(ws && loads && y, GMF_EINVAL, "null argument");
    const gmf_plan& p = *ws->plan;
    require_device_pointer(loads, "loads");
    require_device_pointer(y, "y");
    int64_t off = 0;
    for (const Bucket& b : p.buckets) {
      if (!b.cloud) continue;
      if (b.face.enabled)
        launch_face_rhs(b, loads + off, y, ws->stream);
      else
        launch_cloud_rhs(b, loads + off, y, ws->stream);
      off += b.total_points * int64_t(b.sides) * (b.dim + 1);
    }
    sync(ws);
  });
}

gmf_status gmf_interpolate_channels(gmf_workspace* ws, const double* u_full, int32_t channels, double* out) {
  return guarded([&] {
    require(ws && u_full && out, GMF_EINVAL, "null argument");
    const gmf_plan& p = *ws->plan;
    for (const Bucket& b : p.buckets)
      require(!b.cloud, GMF_EUNSUPPORTED, "point-channel interpolation covers tensor-product cells only");
    require_device_pointer(u_full, "u_full");
    require_device_pointer(out, "out");
    int64_t off = 0;
    for (const Bucket& b : p.buckets) {
      launch_cell_interp(b, u_full, channels, out + off, ws->stream);
      off += b.cells * int64_t(b.Qtot) * channels;
    }
    sync(ws);
  });
}

gmf_status gmf_plan_set_point_data(gmf_plan* plan, int32_t bucket, const double* point_data) {
  return guarded([&] {
    require(plan && point_data, GMF_EINVAL, "null argument");
    require(bucket >= 0 && size_t(bucket) < plan->buckets.size(), GMF_EINVAL, "bucket index out of range");
    Bucket& b = plan->buckets[size_t(bucket)];
    require(!b.cloud, GMF_EUNSUPPORTED, "point data updates cover tensor-product cells only");
    if (b.cells == 0) return;
    const size_t n = size_t(b.cells) * size_t(b.Qtot) * size_t(b.point_stride);
    for (size_t i = 0; i < n; ++i)
      require(std::isfinite(point_data[i]), GMF_EINVAL, "non-finite point data");
    GMF_HIP(hipMemcpy(b.pdata, point_data, n * sizeof(double), hipMemcpyHostToDevice));
    if (b.pdata32) {
      narrow(b.pdata32, b.pdata, int64_t(n), nullptr);
      GMF_HIP(hipDeviceSynchronize());
    }
  });
}

gmf_status gmf_reduced_rhs(gmf_workspace* ws, const double* rhs_full, double* b_red) {
  return guarded([&] {
    require(ws && rhs_full && b_red, GMF_EINVAL, "null argument");
    const gmf_plan& p = *ws->plan;
    require(rhs_full != ws->full_in && rhs_full != ws->full_out, GMF_EALIAS, "right-hand side aliases workspace");
    if (p.shift) {
      // b - A_phys shift on the full layout, then P^T.
      if (!plan_writes_every_dof(p)) fill(ws->full_out, 0.0, p.n_full, ws->stream);
      cell_pass(ws, p.shift, ws->full_out, false);
      copy(ws->full_in, rhs_full, p.n_full, ws->stream);
      axpby(ws, ws->full_in, -1.0, ws->full_out, 1.0, p.n_full);
      project(ws, ws->full_in, b_red);
    } else {
      project(ws, rhs_full, b_red);
    }
    sync(ws);
  });
}

gmf_status gmf_l2_error(gmf_workspace* ws, const double* u_full, const double* weights, const double* exact,
                        double* sums) {
  return guarded([&] {
    require(ws && u_full && weights && exact && sums, GMF_EINVAL, "null argument");
    const gmf_plan& p = *ws->plan;
    // One partial pair per tensor cell and per work unit of a one-sided
    // point-cloud batch off the face path; two-sided batches are not summed.
    int64_t pairs = 0;
    for (const Bucket& b : p.buckets) {
      if (!b.cloud)
        pairs += b.cells;
      else if (b.sides == 1 && !b.face.enabled)
        pairs += b.units;
    }
    require_device_pointer(u_full, "u_full");
    require_device_pointer(weights, "weights");
    require_device_pointer(exact, "exact");
    DeviceBuffer<double> partial(size_t(std::max<int64_t>(2 * pairs, 2))), out(2);
    fill(partial.p, 0.0, 2 * std::max<int64_t>(pairs, 1), ws->stream);
    int64_t off = 0, coff = 0;
    for (const Bucket& b : p.buckets) {
      if (!b.cloud) {
        launch_cell_l2(b, u_full, weights + off, exact + off, partial.p + 2 * coff, ws->stream);
        off += b.cells * int64_t(b.Qtot);
        coff += b.cells;
      } else if (b.sides == 1) {
        if (!b.face.enabled) {
          launch_cloud_l2(b, u_full, weights + off, exact + off, partial.p + 2 * coff, ws->stream);
          coff += b.units;
        }
        off += b.total_points;
      }
    }
    launch_sum_pairs(partial.p, pairs, out.p, ws->stream);
    GMF_HIP(hipMemcpyAsync(sums, out.p, 2 * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
    sync(ws);
  });
}

gmf_status gmf_apply_timed(gmf_workspace* ws, const double* x, double* y, int32_t reps, double* cell_ms,
                           double* total_ms) {
  return guarded([&] {
    require(ws && x && y && reps >= 1, GMF_EINVAL, "invalid argument");
    require(x != y, GMF_EALIAS, "result and coefficient vector must not alias");
    hipEvent_t e0, e1, e2, e3;
    GMF_HIP(hipEventCreate(&e0));
    GMF_HIP(hipEventCreate(&e1));
    GMF_HIP(hipEventCreate(&e2));
    GMF_HIP(hipEventCreate(&e3));
    double cell = 0.0, total = 0.0;
    for (int r = 0; r < reps; ++r) {
      GMF_HIP(hipEventRecord(e0, ws->stream));
      if (!plan_writes_every_dof(*ws->plan)) fill(y, 0.0, ws->plan->n_full, ws->stream);
      GMF_HIP(hipEventRecord(e1, ws->stream));
      cell_pass(ws, x, y, false);
      GMF_HIP(hipEventRecord(e2, ws->stream));
      constraint_rows(ws, x, y);
      GMF_HIP(hipEventRecord(e3, ws->stream));
      GMF_HIP(hipEventSynchronize(e3));
      float a = 0.f, b = 0.f;
      GMF_HIP(hipEventElapsedTime(&a, e1, e2));
      GMF_HIP(hipEventElapsedTime(&b, e0, e3));
      cell += a;
      total += b;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    (void)hipEventDestroy(e3);
    if (cell_ms) *cell_ms = cell / reps;
    if (total_ms) *total_ms = total / reps;
  });
}

gmf_status gmf_reduced_apply(gmf_workspace* ws, const double* xr, double* yr) {
  return guarded([&] {
    require(ws && xr && yr, GMF_EINVAL, "null argument");
    reduced_apply(ws, xr, yr);
    sync(ws);
  });
}

gmf_status gmf_reduced_apply_f32(gmf_workspace* ws, const float* xr, float* yr) {
  return guarded([&] {
    require(ws && xr && yr, GMF_EINVAL, "null argument");
    require_device_pointer(xr, "x_red");
    require_device_pointer(yr, "y_red");
    prepare_fp32_workspace(ws);
    reduced_apply_f32(ws, xr, yr);
    sync(ws);
  });
}

gmf_status gmf_expand_reduced(gmf_workspace* ws, const double* xr, double* xf, int32_t include_shift) {
  return guarded([&] {
    require(ws && xr && xf, GMF_EINVAL, "null argument");
    expand(ws, xr, xf, include_shift != 0);
    sync(ws);
  });
}

gmf_status gmf_project_reduced(gmf_workspace* ws, const double* yf, double* yr) {
  return guarded([&] {
    require(ws && yf && yr, GMF_EINVAL, "null argument");
    project(ws, yf, yr);
    sync(ws);
  });
}

gmf_status gmf_diagonal_physical(gmf_workspace* ws, double* diag) {
  return guarded([&] {
    require(ws && diag, GMF_EINVAL, "null argument");
    const gmf_plan& p = *ws->plan;
    fill(diag, 0.0, p.n_full, ws->stream);
    for (const Bucket& b : p.buckets) {
      if (b.cloud)
        launch_cell_cloud_diagonal(p, b, diag, false, ws->stream);
      else
        launch_cell_diagonal_generic(p, b, diag, false, ws->stream);
    }
    sync(ws);
  });
}

gmf_status gmf_reduced_diagonal(gmf_workspace* ws, double* diag) {
  return guarded([&] {
    require(ws && diag, GMF_EINVAL, "null argument");
    reduced_diagonal(ws, diag);
    sync(ws);
  });
}

gmf_status gmf_vector_copy(gmf_workspace* ws, double* dst, const double* src, int64_t n) {
  return guarded([&] {
    require(ws && dst && src && n >= 0, GMF_EINVAL, "invalid argument");
    copy(dst, src, n, ws->stream);
    sync(ws);
  });
}

gmf_status gmf_vector_axpby(gmf_workspace* ws, double* y, double a, const double* x, double b, int64_t n) {
  return guarded([&] {
    require(ws && y && x && n >= 0, GMF_EINVAL, "invalid argument");
    axpby(ws, y, a, x, b, n);
    sync(ws);
  });
}

gmf_status gmf_vector_dot(gmf_workspace* ws, const double* x, const double* y, int64_t n, double* out) {
  return guarded([&] {
    require(ws && x && y && out && n >= 0, GMF_EINVAL, "invalid argument");
    dot(ws, x, y, n, kTmp);
    read_scalars(ws, kTmp + 1);
    *out = ws->host_scalars[kTmp];
  });
}

gmf_status gmf_invert_diagonal(gmf_workspace* ws, const double* d, double* inv, int64_t n) {
  return guarded([&] {
    require(ws && d && inv && n >= 0, GMF_EINVAL, "null argument");
    invert_diagonal(ws, d, inv, n);
  });
}

gmf_status gmf_precond_create_jacobi(gmf_workspace* ws, const double* invdiag, gmf_preconditioner** out) {
  return guarded([&] {
    require(ws && invdiag && out, GMF_EINVAL, "null argument");
    auto j = std::make_unique<JacobiPrecond>();
    j->ws = ws;
    j->n = ws->plan->n_red;
    j->invdiag = dalloc<double>(size_t(j->n));
    copy(j->invdiag, invdiag, j->n, ws->stream);
    sync(ws);
    *out = j.release();
  });
}

gmf_status gmf_precond_destroy(gmf_preconditioner* M) {
  return guarded([&] { delete M; });
}

gmf_status gmf_precond_apply(gmf_preconditioner* M, const double* r, double* z) {
  return guarded([&] {
    require(M && r && z, GMF_EINVAL, "null argument");
    M->apply(r, z);
    sync(M->ws);
  });
}

gmf_status gmf_precond_apply_block(gmf_preconditioner* M, const double* R, double* Z, int32_t k) {
  return guarded([&] {
    require(M && R && Z && k >= 1, GMF_EINVAL, "invalid argument");
    require(R != Z, GMF_EALIAS, "result and input blocks must not alias");
    const int64_t n = M->ws->plan->n_red;
    if (!M->apply_block(R, Z, n, k))
      for (int j = 0; j < k; ++j) M->apply(R + int64_t(j) * n, Z + int64_t(j) * n);
    sync(M->ws);
  });
}

gmf_status gmf_cg_solve(gmf_workspace* ws, gmf_preconditioner* M, const double* b, double* x, int32_t use_initial,
                        double rtol, double atol, int64_t maxiter, gmf_solve_info* info) {
  gmf_solve_info local{};
  gmf_solve_info& inf = info ? *info : local;
  inf = gmf_solve_info{};
  return guarded([&] {
    require(ws && b && x, GMF_EINVAL, "null argument");
    require(maxiter >= 1, GMF_EINVAL, "maxiter must be positive");
    require(std::isfinite(rtol) && rtol >= 0 && std::isfinite(atol) && atol >= 0, GMF_EINVAL,
            "tolerances must be finite and non-negative");
    require(M == nullptr || M->ws == ws, GMF_EINVAL, "preconditioner belongs to a different workspace");
    CgVectors v(ws->plan->n_red);
    cg_run(ws, M, b, x, use_initial != 0, rtol, atol, maxiter, inf, v);
  });
}

gmf_status gmf_fgmres_solve(gmf_workspace* ws, gmf_preconditioner* M, const double* b, double* x, int32_t use_initial,
                            int32_t restart, double rtol, double atol, int64_t maxiter, gmf_solve_info* info) {
  gmf_solve_info local{};
  gmf_solve_info& inf = info ? *info : local;
  inf = gmf_solve_info{};
  return guarded([&] {
    require(ws && b && x, GMF_EINVAL, "null argument");
    require(restart >= 1, GMF_EINVAL, "restart must be positive");
    require(maxiter >= 1, GMF_EINVAL, "maxiter must be positive");
    require(M == nullptr || M->ws == ws, GMF_EINVAL, "preconditioner belongs to a different workspace");
    const int64_t n = ws->plan->n_red;
    if (n == 0) {
      inf.converged = 1;
      return;
    }
    const int m = int(std::min<int64_t>(restart, n));
    require(m + 1 < kScalarCount - kHessian, GMF_EINVAL, "restart too large");
    FgmresVectors v(n, m);
    fgmres_run(ws, M, b, x, use_initial != 0, m, rtol, atol, maxiter, inf, v);
  });
}

}  // extern "C"
