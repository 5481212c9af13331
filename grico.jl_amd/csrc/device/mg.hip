// hp geometric multigrid V-cycle on the device (include/grico_mf.h
// gmf_precond_create_mg).
//
// Reference: src/multigrid.jl:858-1087 (_apply_preconditioner!, _vcycle!,
// _smooth_jacobi!, _prolongate_reduced_add!, _restrict_reduced!,
// _coarse_solve!) and 382-408 (_compile_coarse_solver). Every level owns a
// workspace of its rediscretized plan; transfers are CSR tables in both
// directions so prolongation and restriction are both gathers (deterministic,
// no atomics). The first smoothing sweep of each level starts from x = 0 and
// skips the product A * 0. The level vectors, smoothers, transfers and level
// operators run in the level type R (double, or float for the FP32 V-cycle);
// the coarse solve, the eigenvalue estimates and the diagonals stay FP64.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "blas.hpp"
#include <rocsolver/rocsolver.h>

#include "blocks.hpp"
#include "common.hpp"
#include "internal.hpp"

namespace gmf {
namespace {

// ---- kernels -----------------------------------------------------------------

#define GMF_GRID_STRIDE(i, n) \
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < (n); i += int64_t(gridDim.x) * blockDim.x)

template <class R>
__global__ void k_jacobi_first(R* __restrict__ x, const R* __restrict__ b, const R* __restrict__ dinv, R omega,
                               int64_t n, int64_t ld) {
  const int64_t co = int64_t(blockIdx.y) * ld;  // column blockIdx.y of a block of columns
  x += co;
  b += co;
  GMF_GRID_STRIDE(i, n) x[i] = omega * (dinv[i] * b[i]);
}

// x += omega D^-1 (b - q), q = A x.
template <class R>
__global__ void k_jacobi_update(R* __restrict__ x, const R* __restrict__ b, const R* __restrict__ q,
                                const R* __restrict__ dinv, R omega, int64_t n, int64_t ld) {
  const int64_t co = int64_t(blockIdx.y) * ld;
  x += co;
  b += co;
  q += co;
  GMF_GRID_STRIDE(i, n) x[i] = fma(omega, dinv[i] * (b[i] - q[i]), x[i]);
}

// bc[c] = sum_f P[f,c] (b[f] - q[f]) over the coarse row's CSR (q may be null).
template <class R>
__global__ void k_restrict(R* __restrict__ bc, const R* __restrict__ b, const R* __restrict__ q,
                           const int64_t* __restrict__ off, const int32_t* __restrict__ idx,
                           const R* __restrict__ val, int64_t nc, int64_t ldc, int64_t ldf) {
  bc += int64_t(blockIdx.y) * ldc;
  b += int64_t(blockIdx.y) * ldf;
  if (q) q += int64_t(blockIdx.y) * ldf;
  GMF_GRID_STRIDE(c, nc) {
    R s = R(0);
    for (int64_t t = off[c]; t < off[c + 1]; ++t) {
      const int32_t f = idx[t];
      const R r = q ? b[f] - q[f] : b[f];
      s = fma(val[t], r, s);
    }
    bc[c] = s;
  }
}

// x[f] += sum_c P[f,c] xc[c].
template <class R>
__global__ void k_prolong_add(R* __restrict__ x, const R* __restrict__ xc, const int64_t* __restrict__ off,
                              const int32_t* __restrict__ idx, const R* __restrict__ val, int64_t nf, int64_t ldf,
                              int64_t ldc) {
  x += int64_t(blockIdx.y) * ldf;
  xc += int64_t(blockIdx.y) * ldc;
  GMF_GRID_STRIDE(f, nf) {
    R s = R(0);
    for (int64_t t = off[f]; t < off[f + 1]; ++t) s = fma(val[t], xc[idx[t]], s);
    x[f] += s;
  }
}

// The Chebyshev kernels round every product before a following sum (no
// contraction), like the fused brick-store variant (ChebFused).
// Chebyshev start from x = 0: r = b, d = s D^-1 r, x = d.
template <class R>
__global__ void k_cheb_start0(R* __restrict__ x, R* __restrict__ d, R* __restrict__ r, const R* __restrict__ b,
                              const R* __restrict__ dinv, R s, int64_t n, int64_t ld) {
#pragma clang fp contract(off)
  const int64_t co = int64_t(blockIdx.y) * ld;
  x += co;
  d += co;
  r += co;
  b += co;
  GMF_GRID_STRIDE(i, n) {
    const R ri = b[i];
    const R di = s * (dinv[i] * ri);
    r[i] = ri;
    d[i] = di;
    x[i] = di;
  }
}

// Chebyshev start from x with q = A x: r = b - q, d = s D^-1 r, x += d.
template <class R>
__global__ void k_cheb_start(R* __restrict__ x, R* __restrict__ d, R* __restrict__ r, const R* __restrict__ b,
                             const R* __restrict__ q, const R* __restrict__ dinv, R s, int64_t n, int64_t ld) {
#pragma clang fp contract(off)
  const int64_t co = int64_t(blockIdx.y) * ld;
  x += co;
  d += co;
  r += co;
  b += co;
  q += co;
  GMF_GRID_STRIDE(i, n) {
    const R ri = b[i] - q[i];
    const R di = s * (dinv[i] * ri);
    r[i] = ri;
    d[i] = di;
    x[i] += di;
  }
}

// Chebyshev step with q = A d: r -= q, d = c1 d + c2 D^-1 r, x += d.
template <class R>
__global__ void k_cheb_step(R* __restrict__ x, R* __restrict__ d, R* __restrict__ r, const R* __restrict__ q,
                            const R* __restrict__ dinv, R c1, R c2, int64_t n, int64_t ld) {
#pragma clang fp contract(off)
  const int64_t co = int64_t(blockIdx.y) * ld;
  x += co;
  d += co;
  r += co;
  q += co;
  GMF_GRID_STRIDE(i, n) {
    const R ri = r[i] - q[i];
    const R di = fma(c1, d[i], c2 * (dinv[i] * ri));
    r[i] = ri;
    d[i] = di;
    x[i] += di;
  }
}

// x = Ainv b for a dense row-major inverse (one wavefront per row). Each lane
// keeps four partial sums over columns lane + 64 (4 t + u), so that four
// loads of the row are in flight per step; fixed order, deterministic.
__global__ void k_dense_matvec(double* __restrict__ x, const double* __restrict__ A, const double* __restrict__ b,
                               int64_t n) {
  const int64_t row = blockIdx.x * int64_t(blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double* a = A + row * n;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int64_t j = lane;
  for (; j + 192 < n; j += 256) {
    s0 = fma(a[j], b[j], s0);
    s1 = fma(a[j + 64], b[j + 64], s1);
    s2 = fma(a[j + 128], b[j + 128], s2);
    s3 = fma(a[j + 192], b[j + 192], s3);
  }
  for (; j < n; j += 64) s0 = fma(a[j], b[j], s0);
  double s = (s0 + s1) + (s2 + s3);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) x[row] = s;
}

__global__ void k_unit(double* __restrict__ e, int64_t j, int64_t n) {
  GMF_GRID_STRIDE(i, n) e[i] = (i == j) ? 1.0 : 0.0;
}

// out = a * b elementwise; out may alias b.
__global__ void k_mul(double* out, const double* __restrict__ a, const double* b, int64_t n) {
  GMF_GRID_STRIDE(i, n) out[i] = a[i] * b[i];
}

__global__ void k_scale(double* __restrict__ x, double s, int64_t n) {
  GMF_GRID_STRIDE(i, n) x[i] *= s;
}

__global__ void k_sqrt(double* __restrict__ out, const double* __restrict__ a, int64_t n) {
  GMF_GRID_STRIDE(i, n) out[i] = sqrt(a[i]);
}

// Deterministic start vector of the Lanczos estimate.
__global__ void k_start_vector(double* __restrict__ v, int64_t n) {
  GMF_GRID_STRIDE(i, n) {
    const double t = double(i + 1);
    v[i] = 1.0 + 0.5 * sin(0.019 * t) + 0.1 * cos(0.043 * t);
  }
}


// ---- host helpers ------------------------------------------------------------

// Largest eigenvalue of the symmetric tridiagonal (alpha, beta) by Sturm
// bisection.
double tridiagonal_max_eigenvalue(const std::vector<double>& alpha, const std::vector<double>& beta) {
  const size_t m = alpha.size();
  double lo = std::numeric_limits<double>::max(), hi = -lo;
  for (size_t i = 0; i < m; ++i) {
    const double rad = (i > 0 ? std::fabs(beta[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(beta[i]) : 0.0);
    lo = std::min(lo, alpha[i] - rad);
    hi = std::max(hi, alpha[i] + rad);
  }
  auto count_below = [&](double x) {
    int c = 0;
    double q = alpha[0] - x;
    if (q < 0) ++c;
    for (size_t i = 1; i < m; ++i) {
      const double qq = (q == 0.0) ? std::numeric_limits<double>::epsilon() * (std::fabs(beta[i - 1]) + 1e-300) : q;
      q = alpha[i] - x - beta[i - 1] * beta[i - 1] / qq;
      if (q < 0) ++c;
    }
    return c;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-14 * std::max(1.0, std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) >= int(m))
      hi = mid;
    else
      lo = mid;
  }
  return hi;
}

template <class R>
struct Level {
  gmf_workspace* ws = nullptr;
  int64_t n = 0;
  DeviceBuffer<R> invdiag, x, b, q, r, d;  // invdiag: zero on block dofs when the level has blocks
  // The same vectors for a block of columns (column-major, n x kb), allocated
  // on the first block V-cycle with kb columns or more.
  int kb = 0;
  DeviceBuffer<R> xb, bb, qb, rb, db;
  BlockSet<R> blocks;                      // cell blocks of the block smoother (count 0: point smoothing)
  DeviceBuffer<double> gV, gZ, gy;  // GMRES smoother Krylov basis, preconditioned basis, coefficients
  double lambda_max = 0.0;
};

template <class R>
struct TransferCsr {
  int64_t nc = 0, nf = 0;
  DeviceBuffer<int64_t> poff, roff;
  DeviceBuffer<int32_t> pidx, ridx;
  DeviceBuffer<R> pval, rval;
};

template <class T>
DeviceBuffer<T> upload_vec(const std::vector<T>& h) {
  DeviceBuffer<T> d(h.size());
  if (!h.empty()) GMF_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

inline void level_apply(gmf_workspace* ws, const double* x, double* y) { reduced_apply(ws, x, y); }
inline void level_apply(gmf_workspace* ws, const float* x, float* y) { reduced_apply_f32(ws, x, y); }

// dst (type R) = src (FP64), and back.
inline void to_level(double* dst, const double* src, int64_t n, hipStream_t s) { copy(dst, src, n, s); }
inline void to_level(float* dst, const double* src, int64_t n, hipStream_t s) { narrow(dst, src, n, s); }
inline void from_level(double* dst, const double* src, int64_t n, hipStream_t s) { copy(dst, src, n, s); }
inline void from_level(double* dst, const float* src, int64_t n, hipStream_t s) { widen(dst, src, n, s); }

// Common part of the FP64 and FP32 V-cycles (diagnostics).
struct MGBase : gmf_preconditioner {
  std::vector<int64_t> level_n;
  std::vector<double> level_lambda;
};

template <class R>
struct MGPrecond : MGBase {
  std::vector<Level<R>> levels;
  std::vector<TransferCsr<R>> transfers;
  int smoother = GMF_SMOOTHER_JACOBI;
  int pre = 2, post = 2;
  double omega = 0.8;
  int cheb_degree = 4;
  double cheb_range = 20.0;
  int restart = 8;
  bool spd = true;
  bool fuse = true;  // Chebyshev steps fused into brick-level applies
  // coarse solver (FP64)
  bool dense = false;
  DeviceBuffer<double> coarse_inverse;
  DeviceBuffer<double> coarse_invdiag;
  DeviceBuffer<double> coarse_b, coarse_x;  // FP64 coarse vectors when R is float
  std::unique_ptr<JacobiPrecond> coarse_jacobi;
  std::unique_ptr<CgVectors> coarse_vectors;
  std::unique_ptr<FgmresVectors> coarse_fgmres;
  double coarse_rtol = 1e-12, coarse_atol = 0.0;
  int64_t coarse_maxiter = 10000;
  // Top-level vectors in R (FP32 conversion, or the buffers a captured graph
  // reads and writes) and the captured V-cycle.
  bool use_graph = false;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  DeviceBuffer<R> top_b, top_x;
  rocblas_handle blas = nullptr;  // dense coarse products of block V-cycles

  static constexpr bool kDouble = std::is_same<R, double>::value;

  ~MGPrecond() override {
    if (blas) (void)rocblas_destroy_handle(blas);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }

  void set_streams(hipStream_t s) {
    for (Level<R>& L : levels) L.ws->stream = s;
  }

  void capture() {
    const int top = int(levels.size()) - 1;
    const hipStream_t user = ws->stream;
    if (!capture_stream) GMF_HIP(hipStreamCreateWithFlags(&capture_stream, hipStreamNonBlocking));
    set_streams(capture_stream);
    GMF_HIP(hipStreamBeginCapture(capture_stream, hipStreamCaptureModeThreadLocal));
    try {
      vcycle(top, top_x.p, top_b.p);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(capture_stream, &g);
      if (g) (void)hipGraphDestroy(g);
      set_streams(user);
      throw;
    }
    hipError_t e = hipStreamEndCapture(capture_stream, &graph);
    set_streams(user);
    GMF_HIP(e);
    GMF_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
  }

  void apply(const double* r, double* z) override {
    const int top = int(levels.size()) - 1;
    const int64_t n = levels[size_t(top)].n;
    if constexpr (kDouble) {
      if (!use_graph) {
        set_streams(ws->stream);
        vcycle(top, z, r);
        return;
      }
    }
    if (use_graph && !graph_exec) capture();
    to_level(top_b.p, r, n, ws->stream);
    if (use_graph) {
      GMF_HIP(hipGraphLaunch(graph_exec, ws->stream));
    } else {
      set_streams(ws->stream);
      vcycle(top, top_x.p, top_b.p);
    }
    from_level(z, top_x.p, n, ws->stream);
  }

  // One fixed GMRES cycle (src/multigrid.jl:963-1035): count = min(steps,
  // restart, n) Arnoldi steps on A D^-1 from the residual, Givens rotations,
  // x += omega sum_j y_j D^-1 v_j. Host reads per column (not captured).
  void smooth_gmres(int l, double* x, const double* b, int steps, bool x_zero) {
    Level<R>& L = levels[size_t(l)];
    const int64_t n = L.n;
    gmf_workspace* lw = L.ws;
    hipStream_t s = ws->stream;
    const int count = int(std::min<int64_t>(std::min(steps, restart), n));
    // x is updated in place below; on the first sweep of a level its memory
    // holds no meaningful value yet and must be cleared.
    if (x_zero) fill(x, 0.0, n, s);
    if (count <= 0) return;
    if (x_zero) {
      copy(L.r.p, b, n, s);
    } else {
      level_apply(lw, x, L.q.p);
      copy(L.r.p, b, n, s);
      axpby(lw, L.r.p, -1.0, L.q.p, 1.0, n);
    }
    dot(lw, L.r.p, L.r.p, n, kTmp);
    read_scalars(lw, kTmp + 1);
    const double* hs = lw->host_scalars;
    const double beta = std::sqrt(hs[kTmp]);
    if (beta == 0.0) return;
    double* V = L.gV.p;
    double* Z = L.gZ.p;
    fill(V, 0.0, n, s);
    axpby(lw, V, 1.0 / beta, L.r.p, 0.0, n);
    const size_t sc = size_t(count);
    std::vector<double> H((sc + 1) * sc, 0.0), g(sc + 1, 0.0), cs(sc, 0.0), sn(sc, 0.0), y(sc, 0.0);
    auto Hm = [&](int rr, int cc) -> double& { return H[size_t(rr * count + cc)]; };
    g[0] = beta;
    int inner = 0;
    for (int col = 0; col < count; ++col) {
      inner = col + 1;
      double* zc = Z + int64_t(col) * n;
      const double* vc = V + int64_t(col) * n;
      hipLaunchKernelGGL(k_mul, dim3(grid1(n)), dim3(256), 0, s, zc, reinterpret_cast<const double*>(L.invdiag.p),
                         vc, n);
      GMF_HIP(hipGetLastError());
      level_apply(lw, zc, L.q.p);
      arnoldi_mgs(lw, L.q.p, V, n, col);
      for (int bb = 0; bb <= col; ++bb) Hm(bb, col) = hs[kHessian + bb];
      const double next = std::sqrt(std::max(0.0, hs[kHessian + col + 1]));
      Hm(col + 1, col) = next;
      if (next > 0.0) {
        double* vn = V + int64_t(col + 1) * n;
        fill(vn, 0.0, n, s);
        axpby(lw, vn, 1.0 / next, L.q.p, 0.0, n);
      }
      for (int row = 0; row < col; ++row) {
        const double f = Hm(row, col), t = Hm(row + 1, col);
        Hm(row, col) = cs[size_t(row)] * f + sn[size_t(row)] * t;
        Hm(row + 1, col) = -sn[size_t(row)] * f + cs[size_t(row)] * t;
      }
      const double scale = std::hypot(Hm(col, col), Hm(col + 1, col));
      const double c = scale == 0.0 ? 1.0 : Hm(col, col) / scale;
      const double sv = scale == 0.0 ? 0.0 : Hm(col + 1, col) / scale;
      cs[size_t(col)] = c;
      sn[size_t(col)] = sv;
      {
        const double f = Hm(col, col), t = Hm(col + 1, col);
        Hm(col, col) = c * f + sv * t;
        Hm(col + 1, col) = -sv * f + c * t;
        const double gf = g[size_t(col)], gt = g[size_t(col) + 1];
        g[size_t(col)] = c * gf + sv * gt;
        g[size_t(col) + 1] = -sv * gf + c * gt;
      }
      if (next == 0.0) break;
    }
    for (int row = inner - 1; row >= 0; --row) {
      double v = g[size_t(row)];
      for (int c = row + 1; c < inner; ++c) v -= Hm(row, c) * y[size_t(c)];
      const double dg = Hm(row, row);
      require(std::isfinite(dg), GMF_EINVAL, "GMRES smoother encountered a non-finite Krylov least-squares diagonal");
      require(dg != 0.0, GMF_EINVAL, "GMRES smoother encountered a singular Krylov least-squares problem");
      y[size_t(row)] = v / dg;
    }
    for (int c = 0; c < inner; ++c) y[size_t(c)] *= omega;
    GMF_HIP(hipMemcpyAsync(L.gy.p, y.data(), sizeof(double) * size_t(inner), hipMemcpyHostToDevice, s));
    multi_axpy(lw, x, Z, n, L.gy.p, inner, n, 1.0);
    sync(lw);
  }

  // Level work vectors for k columns (k = 1: the level's own vectors).
  struct Work {
    R *x, *b, *q, *r, *d;
  };
  Work work(Level<R>& L, int k) {
    if (k == 1) return Work{L.x.p, L.b.p, L.q.p, L.r.p, L.d.p};
    return Work{L.xb.p, L.bb.p, L.qb.p, L.rb.p, L.db.p};
  }

  void apply_level(Level<R>& L, const R* x, R* q, int k) {
    if (k == 1) {
      level_apply(L.ws, x, q);
      return;
    }
    if constexpr (kDouble) {
      reduced_apply_block(L.ws, x, q, k);
    } else {
      require(false, GMF_EINVAL, "block V-cycles run in FP64");
    }
  }

  // steps smoothing sweeps on k columns (x, b: n x k, column-major).
  void smooth(int l, R* x, const R* b, int steps, bool x_zero, int k = 1) {
    Level<R>& L = levels[size_t(l)];
    const int64_t n = L.n;
    hipStream_t s = ws->stream;
    if (n == 0) return;
    if (smoother == GMF_SMOOTHER_GMRES) {
      if constexpr (kDouble) smooth_gmres(l, x, b, steps, x_zero);
      return;
    }
    const Work w = work(L, k);
    const dim3 grid(grid1(n), unsigned(k));
    if (smoother == GMF_SMOOTHER_JACOBI) {
      for (int it = 0; it < steps; ++it) {
        if (x_zero && it == 0) {
          hipLaunchKernelGGL(k_jacobi_first<R>, grid, dim3(256), 0, s, x, b, L.invdiag.p, R(omega), n, n);
          block_add<R>(L.blocks, x, nullptr, b, nullptr, R(omega), s, k, n);
        } else {
          apply_level(L, x, w.q, k);
          hipLaunchKernelGGL(k_jacobi_update<R>, grid, dim3(256), 0, s, x, b, w.q, L.invdiag.p, R(omega), n, n);
          // The block part uses the same residual b - A x_old (q is unchanged).
          block_add<R>(L.blocks, x, nullptr, b, w.q, R(omega), s, k, n);
        }
        GMF_HIP(hipGetLastError());
      }
      return;
    }
    const double lmax = L.lambda_max, lmin = L.lambda_max / cheb_range;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    // With blocks, each start / step kernel forms the point part of S r and
    // block_add then adds the blocks' part to d and x (r is final by then).
    // The fused brick steps serve one column.
    for (int it = 0; it < steps; ++it) {
      if (x_zero && it == 0) {
        hipLaunchKernelGGL(k_cheb_start0<R>, grid, dim3(256), 0, s, x, w.d, w.r, b, L.invdiag.p, R(1.0 / theta), n,
                           n);
      } else if (k > 1 || !cheb_fused(L, x, ChebFused<R>{x, L.d.p, L.r.p, b, L.invdiag.p, R(0), R(1.0 / theta), 1})) {
        apply_level(L, x, w.q, k);
        hipLaunchKernelGGL(k_cheb_start<R>, grid, dim3(256), 0, s, x, w.d, w.r, b, w.q, L.invdiag.p, R(1.0 / theta),
                           n, n);
      }
      block_add<R>(L.blocks, x, w.d, w.r, nullptr, R(1.0 / theta), s, k, n);
      double rho = 1.0 / sigma;
      for (int i = 1; i < cheb_degree; ++i) {
        const double rho_new = 1.0 / (2.0 * sigma - rho);
        const R c1 = R(rho_new * rho), c2 = R(2.0 * rho_new / delta);
        if (k > 1 || !cheb_fused(L, L.d.p, ChebFused<R>{x, L.d.p, L.r.p, nullptr, L.invdiag.p, c1, c2, 0})) {
          apply_level(L, w.d, w.q, k);
          hipLaunchKernelGGL(k_cheb_step<R>, grid, dim3(256), 0, s, x, w.d, w.r, w.q, L.invdiag.p, c1, c2, n, n);
        }
        block_add<R>(L.blocks, x, w.d, w.r, nullptr, c2, s, k, n);
        rho = rho_new;
      }
      GMF_HIP(hipGetLastError());
    }
  }

  // Chebyshev start or step fused into the level's apply of u (one brick
  // bucket on a selection map); false when the level needs the separate
  // apply and vector kernel. GMF_MG_FUSED=0 turns the fusion off.
  bool cheb_fused(Level<R>& L, const R* u, const ChebFused<R>& c) {
    if (!fuse) return false;
    gmf_workspace* w = L.ws;
    if constexpr (kDouble)
      return chebyshev_fused(*w->plan, u, c, w->brick_slots, w->stream);
    else
      return chebyshev_fused(*w->plan, u, c, reinterpret_cast<float*>(w->brick_slots), w->stream);
  }

  void coarse_solve64(double* x, const double* b) {
    Level<R>& L = levels[0];
    if (dense) {
      const unsigned blocks = unsigned((L.n + 3) / 4);
      hipLaunchKernelGGL(k_dense_matvec, dim3(blocks), dim3(256), 0, ws->stream, x, coarse_inverse.p, b, L.n);
      GMF_HIP(hipGetLastError());
      return;
    }
    gmf_solve_info inf{};
    coarse_jacobi->ws = L.ws;
    if (spd)
      cg_run(L.ws, coarse_jacobi.get(), b, x, false, coarse_rtol, coarse_atol, coarse_maxiter, inf, *coarse_vectors);
    else
      fgmres_run(L.ws, coarse_jacobi.get(), b, x, false, coarse_fgmres->m, coarse_rtol, coarse_atol, coarse_maxiter,
                 inf, *coarse_fgmres);
  }

  void coarse_solve(R* x, const R* b) {
    Level<R>& L = levels[0];
    if (L.n == 0) return;
    if constexpr (kDouble) {
      coarse_solve64(x, b);
    } else {
      widen(coarse_b.p, b, L.n, ws->stream);
      coarse_solve64(coarse_x.p, coarse_b.p);
      narrow(x, coarse_x.p, L.n, ws->stream);
    }
  }

  // Coarse solve of k columns: one product with the dense inverse (rocBLAS
  // GEMM), or column by column.
  void coarse_solve_block(double* x, const double* b, int k) {
    Level<R>& L = levels[0];
    if (L.n == 0) return;
    if (dense) {
      if (!blas) {
        require(rocblas_create_handle(&blas) == rocblas_status_success, GMF_EDEVICE, "rocblas_create_handle failed");
      }
      require(rocblas_set_stream(blas, ws->stream) == rocblas_status_success, GMF_EDEVICE, "rocblas_set_stream failed");
      // coarse_inverse is row-major: as a column-major matrix it is the
      // transpose, so x = A^-1 b is op(A^T) with op = transpose.
      const double one = 1.0, zero = 0.0;
      const rocblas_int m = rocblas_int(L.n);
      require(rocblas_dgemm(blas, rocblas_operation_transpose, rocblas_operation_none, m, rocblas_int(k), m, &one,
                            coarse_inverse.p, m, b, m, &zero, x, m) == rocblas_status_success,
              GMF_EDEVICE, "rocblas_dgemm failed");
      return;
    }
    for (int j = 0; j < k; ++j) coarse_solve64(x + int64_t(j) * L.n, b + int64_t(j) * L.n);
  }

  // x is treated as zero on entry; k columns (column-major, the level's n rows).
  void vcycle(int l, R* x, const R* b, int k = 1) {
    if (l == 0) {
      if (k == 1) {
        coarse_solve(x, b);
      } else if constexpr (kDouble) {
        coarse_solve_block(x, b, k);
      }
      return;
    }
    Level<R>& L = levels[size_t(l)];
    Level<R>& C = levels[size_t(l - 1)];
    TransferCsr<R>& T = transfers[size_t(l - 1)];
    hipStream_t s = ws->stream;
    const Work w = work(L, k), wc = work(C, k);
    if (pre > 0) {
      smooth(l, x, b, pre, true, k);
      apply_level(L, x, w.q, k);
      hipLaunchKernelGGL(k_restrict<R>, dim3(grid1(T.nc), unsigned(k)), dim3(256), 0, s, wc.b, b, w.q, T.roff.p,
                         T.ridx.p, T.rval.p, T.nc, C.n, L.n);
    } else {
      fill(x, R(0), L.n * k, s);
      hipLaunchKernelGGL(k_restrict<R>, dim3(grid1(T.nc), unsigned(k)), dim3(256), 0, s, wc.b, b, nullptr, T.roff.p,
                         T.ridx.p, T.rval.p, T.nc, C.n, L.n);
    }
    GMF_HIP(hipGetLastError());
    vcycle(l - 1, wc.x, wc.b, k);
    hipLaunchKernelGGL(k_prolong_add<R>, dim3(grid1(T.nf), unsigned(k)), dim3(256), 0, s, x, wc.x, T.poff.p,
                       T.pidx.p, T.pval.p, T.nf, L.n, C.n);
    GMF_HIP(hipGetLastError());
    smooth(l, x, b, post, false, k);
  }

  // Z = M^-1 R for k columns at once (FP64 V-cycles with Jacobi or Chebyshev
  // smoothing): every level applies, smooths and transfers all columns per
  // launch, and a dense coarse level solves them with one matrix product.
  bool apply_block(const double* r, double* z, int64_t n, int k) override {
    if constexpr (!kDouble) {
      return false;
    } else {
      if (k < 2 || smoother == GMF_SMOOTHER_GMRES) return false;
      const int top = int(levels.size()) - 1;
      require(levels[size_t(top)].n == n, GMF_EINVAL, "block size does not match the multigrid top level");
      for (Level<R>& L : levels) {
        if (L.kb >= k) continue;
        const size_t m = size_t(L.n) * size_t(k);
        L.xb = DeviceBuffer<R>(m);
        L.bb = DeviceBuffer<R>(m);
        L.qb = DeviceBuffer<R>(m);
        L.rb = DeviceBuffer<R>(m);
        L.db = DeviceBuffer<R>(m);
        L.kb = k;
      }
      set_streams(ws->stream);
      vcycle(top, z, r, k);
      return true;
    }
  }
};

template <class R>
TransferCsr<R> build_transfer(const gmf_transfer_desc& t, int64_t nc, int64_t nf) {
  require(t.nnz >= 0 && (t.nnz == 0 || (t.coarse_indices && t.fine_indices && t.coefficients)), GMF_EINVAL,
          "transfer arrays missing");
  require(nc < (int64_t(1) << 31) && nf < (int64_t(1) << 31), GMF_EUNSUPPORTED, "level too large for 32-bit indices");
  TransferCsr<R> c;
  c.nc = nc;
  c.nf = nf;
  const size_t nnz = size_t(t.nnz);
  std::vector<int64_t> poff(size_t(nf) + 1, 0), roff(size_t(nc) + 1, 0);
  for (size_t k = 0; k < nnz; ++k) {
    const int64_t ci = t.coarse_indices[k], fi = t.fine_indices[k];
    require(ci >= 0 && ci < nc && fi >= 0 && fi < nf, GMF_EINVAL, "transfer index out of range");
    require(std::isfinite(t.coefficients[k]), GMF_EINVAL, "transfer coefficient must be finite");
    if (k > 0)
      require(t.fine_indices[k - 1] < fi || (t.fine_indices[k - 1] == fi && t.coarse_indices[k - 1] < ci),
              GMF_EINVAL, "transfer entries must be sorted by (fine, coarse) without duplicates");
    poff[size_t(fi) + 1]++;
    roff[size_t(ci) + 1]++;
  }
  std::partial_sum(poff.begin(), poff.end(), poff.begin());
  std::partial_sum(roff.begin(), roff.end(), roff.begin());
  std::vector<int32_t> pidx(nnz), ridx(nnz);
  std::vector<R> pval(nnz), rval(nnz);
  std::vector<int64_t> fill_r(roff.begin(), roff.end() - 1);
  for (size_t k = 0; k < nnz; ++k) {
    pidx[k] = int32_t(t.coarse_indices[k]);
    pval[k] = R(t.coefficients[k]);
    const size_t slot = size_t(fill_r[size_t(t.coarse_indices[k])]++);
    ridx[slot] = int32_t(t.fine_indices[k]);
    rval[slot] = R(t.coefficients[k]);
  }
  c.poff = upload_vec(poff);
  c.roff = upload_vec(roff);
  c.pidx = upload_vec(pidx);
  c.ridx = upload_vec(ridx);
  c.pval = upload_vec(pval);
  c.rval = upload_vec(rval);
  return c;
}

// Lanczos estimate of lambda_max(D^-1 A) through the symmetric D^-1/2 A D^-1/2.
// FP64 throughout (invdiag: the level's FP64 inverse diagonal).
double estimate_lambda_max(gmf_workspace* ws, int64_t n, const double* invdiag, int iterations) {
  hipStream_t s = ws->stream;
  const int m = int(std::min<int64_t>(iterations, n));
  if (m <= 0) return 0.0;
  DeviceBuffer<double> v(static_cast<size_t>(n)), vp(static_cast<size_t>(n)), u(static_cast<size_t>(n)),
      w(static_cast<size_t>(n)), sdiag(static_cast<size_t>(n));
  hipLaunchKernelGGL(k_sqrt, dim3(grid1(n)), dim3(256), 0, s, sdiag.p, invdiag, n);
  hipLaunchKernelGGL(k_start_vector, dim3(grid1(n)), dim3(256), 0, s, v.p, n);
  fill(vp.p, 0.0, n, s);
  dot(ws, v.p, v.p, n, kTmp);
  read_scalars(ws, kTmp + 1);
  double nrm = std::sqrt(ws->host_scalars[kTmp]);
  hipLaunchKernelGGL(k_scale, dim3(grid1(n)), dim3(256), 0, s, v.p, 1.0 / nrm, n);
  std::vector<double> alpha, beta;
  double beta_prev = 0.0;
  for (int j = 0; j < m; ++j) {
    hipLaunchKernelGGL(k_mul, dim3(grid1(n)), dim3(256), 0, s, u.p, sdiag.p, v.p, n);
    reduced_apply(ws, u.p, w.p);
    hipLaunchKernelGGL(k_mul, dim3(grid1(n)), dim3(256), 0, s, w.p, sdiag.p, w.p, n);
    GMF_HIP(hipGetLastError());
    dot(ws, w.p, v.p, n, kTmp);
    read_scalars(ws, kTmp + 1);
    const double a = ws->host_scalars[kTmp];
    alpha.push_back(a);
    axpby(ws, w.p, -a, v.p, 1.0, n);
    if (j > 0) axpby(ws, w.p, -beta_prev, vp.p, 1.0, n);
    dot(ws, w.p, w.p, n, kTmp);
    read_scalars(ws, kTmp + 1);
    const double bnext = std::sqrt(std::max(0.0, ws->host_scalars[kTmp]));
    if (j + 1 == m || !(bnext > 1e-14 * std::fabs(a))) break;
    beta.push_back(bnext);
    std::swap(vp.p, v.p);                       // vp = old v
    hipLaunchKernelGGL(k_scale, dim3(grid1(n)), dim3(256), 0, s, w.p, 1.0 / bnext, n);  // w /= beta
    std::swap(v.p, w.p);                        // v = w
    beta_prev = bnext;
  }
  beta.resize(alpha.size() > 0 ? alpha.size() - 1 : 0);
  sync(ws);
  return tridiagonal_max_eigenvalue(alpha, beta);
}

// lambda_max(S A) for the block smoother S = D_u^-1 + sum_b R_b^T B_b^-1 R_b
// from the Lanczos tridiagonal of `iterations` preconditioned CG steps
// (T_jj = 1/alpha_j + beta_{j-1}/alpha_{j-1}, T_j,j+1 = sqrt(beta_j)/alpha_j).
// FP64 throughout (invdiag zero on block dofs).
double estimate_lambda_max_blocks(gmf_workspace* ws, int64_t n, const double* invdiag, const BlockSet<double>& bs,
                                  int iterations) {
  hipStream_t s = ws->stream;
  const int m = int(std::min<int64_t>(iterations, n));
  if (m <= 0) return 0.0;
  DeviceBuffer<double> r(static_cast<size_t>(n)), z(static_cast<size_t>(n)), p(static_cast<size_t>(n)),
      q(static_cast<size_t>(n));
  auto precondition = [&] {
    hipLaunchKernelGGL(k_mul, dim3(grid1(n)), dim3(256), 0, s, z.p, invdiag, r.p, n);
    block_add<double>(bs, z.p, nullptr, r.p, nullptr, 1.0, s);
  };
  hipLaunchKernelGGL(k_start_vector, dim3(grid1(n)), dim3(256), 0, s, r.p, n);
  precondition();
  copy(p.p, z.p, n, s);
  dot(ws, r.p, z.p, n, kTmp);
  read_scalars(ws, kTmp + 1);
  double rz = ws->host_scalars[kTmp];
  require(rz > 0.0 && std::isfinite(rz), GMF_ENOTSPD, "block smoother is not positive definite");
  std::vector<double> alpha, beta, diag, off;
  for (int j = 0; j < m; ++j) {
    reduced_apply(ws, p.p, q.p);
    dot(ws, p.p, q.p, n, kTmp);
    read_scalars(ws, kTmp + 1);
    const double pq = ws->host_scalars[kTmp];
    require(pq > 0.0 && std::isfinite(pq), GMF_ENOTSPD, "Chebyshev smoother needs a positive spectrum of S A");
    const double a = rz / pq;
    diag.push_back(1.0 / a + (j > 0 ? beta.back() / alpha.back() : 0.0));
    alpha.push_back(a);
    axpby(ws, r.p, -a, q.p, 1.0, n);
    precondition();
    dot(ws, r.p, z.p, n, kTmp);
    read_scalars(ws, kTmp + 1);
    const double rz_new = ws->host_scalars[kTmp];
    if (j + 1 == m || !(rz_new > 1e-28 * rz)) break;
    const double bt = rz_new / rz;
    off.push_back(std::sqrt(bt) / a);
    beta.push_back(bt);
    axpby(ws, p.p, 1.0, z.p, bt, n);  // p = z + beta p
    rz = rz_new;
  }
  off.resize(diag.size() > 0 ? diag.size() - 1 : 0);
  sync(ws);
  return tridiagonal_max_eigenvalue(diag, off);
}

// Inverse of a general dense matrix by LU with partial pivoting (the
// reference's non-SPD coarse solve, src/multigrid.jl:382-408, 899-907).
std::vector<double> lu_inverse(std::vector<double> A, int64_t n) {
  std::vector<int64_t> piv(static_cast<size_t>(n));
  for (int64_t k = 0; k < n; ++k) {
    int64_t p = k;
    for (int64_t i = k + 1; i < n; ++i)
      if (std::fabs(A[size_t(i * n + k)]) > std::fabs(A[size_t(p * n + k)])) p = i;
    require(A[size_t(p * n + k)] != 0.0 && std::isfinite(A[size_t(p * n + k)]), GMF_EINVAL,
            "coarse multigrid matrix is singular");
    piv[size_t(k)] = p;
    if (p != k)
      for (int64_t j = 0; j < n; ++j) std::swap(A[size_t(k * n + j)], A[size_t(p * n + j)]);
    for (int64_t i = k + 1; i < n; ++i) {
      const double f = A[size_t(i * n + k)] / A[size_t(k * n + k)];
      A[size_t(i * n + k)] = f;
      for (int64_t j = k + 1; j < n; ++j) A[size_t(i * n + j)] -= f * A[size_t(k * n + j)];
    }
  }
  std::vector<double> inv(size_t(n * n)), col(static_cast<size_t>(n));
  for (int64_t j = 0; j < n; ++j) {
    std::fill(col.begin(), col.end(), 0.0);
    col[size_t(j)] = 1.0;
    for (int64_t k = 0; k < n; ++k) std::swap(col[size_t(k)], col[size_t(piv[size_t(k)])]);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t k = 0; k < i; ++k) col[size_t(i)] -= A[size_t(i * n + k)] * col[size_t(k)];
    for (int64_t i = n - 1; i >= 0; --i) {
      for (int64_t k = i + 1; k < n; ++k) col[size_t(i)] -= A[size_t(i * n + k)] * col[size_t(k)];
      col[size_t(i)] /= A[size_t(i * n + i)];
    }
    for (int64_t i = 0; i < n; ++i) inv[size_t(i * n + j)] = col[size_t(i)];
  }
  return inv;
}

// Lower triangle -> upper triangle of a column-major n x n matrix (symmetric result).
__global__ void k_mirror_lower(double* __restrict__ A, int64_t n) {
  const int64_t total = n * n;
  GMF_GRID_STRIDE(k, total) {
    const int64_t i = k % n, j = k / n;  // entry (i, j) at A[i + j n]
    if (i < j) A[i + j * n] = A[j + i * n];
  }
}

// out[i n + j] = A[i + j n] (column-major -> row-major).
__global__ void k_transpose(double* __restrict__ out, const double* __restrict__ A, int64_t n) {
  const int64_t total = n * n;
  GMF_GRID_STRIDE(k, total) out[(k % n) * n + k / n] = A[k];
}

// Dense inverse on the device for coarse levels above the reference's direct
// limit: Cholesky factorization and inverse (rocSOLVER potrf / potri) for SPD
// operators, LU with partial pivoting (getrf / getri) otherwise. `col` holds
// A column-major (column j = A e_j) and is overwritten; the result is
// row-major as k_dense_matvec reads it.
DeviceBuffer<double> device_dense_inverse(gmf_workspace* ws, int64_t n, bool spd, DeviceBuffer<double>& col) {
  hipStream_t s = ws->stream;
  rocblas_handle h = nullptr;
  require(rocblas_create_handle(&h) == rocblas_status_success, GMF_EDEVICE, "rocblas_create_handle failed");
  struct Guard {
    rocblas_handle h;
    ~Guard() { (void)rocblas_destroy_handle(h); }
  } guard{h};
  require(rocblas_set_stream(h, s) == rocblas_status_success, GMF_EDEVICE, "rocblas_set_stream failed");
  DeviceBuffer<rocblas_int> info(1);
  rocblas_int hinfo = 0;
  const rocblas_int m = rocblas_int(n);
  if (spd) {
    require(rocsolver_dpotrf(h, rocblas_fill_lower, m, col.p, m, info.p) == rocblas_status_success, GMF_EDEVICE,
            "rocsolver_dpotrf failed");
    GMF_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost, s));
    sync(ws);
    require(hinfo == 0, GMF_ENOTSPD, "declared SPD coarse multigrid matrix is not positive definite");
    require(rocsolver_dpotri(h, rocblas_fill_lower, m, col.p, m, info.p) == rocblas_status_success, GMF_EDEVICE,
            "rocsolver_dpotri failed");
    hipLaunchKernelGGL(k_mirror_lower, dim3(grid1(n * n)), dim3(256), 0, s, col.p, n);
    GMF_HIP(hipGetLastError());
    sync(ws);
    return std::move(col);
  }
  DeviceBuffer<rocblas_int> piv(static_cast<size_t>(n));
  require(rocsolver_dgetrf(h, m, m, col.p, m, piv.p, info.p) == rocblas_status_success, GMF_EDEVICE,
          "rocsolver_dgetrf failed");
  GMF_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost, s));
  sync(ws);
  require(hinfo == 0, GMF_EINVAL, "coarse multigrid matrix is singular");
  require(rocsolver_dgetri(h, m, col.p, m, piv.p, info.p) == rocblas_status_success, GMF_EDEVICE,
          "rocsolver_dgetri failed");
  DeviceBuffer<double> inv(static_cast<size_t>(n * n));
  hipLaunchKernelGGL(k_transpose, dim3(grid1(n * n)), dim3(256), 0, s, inv.p, col.p, n);
  GMF_HIP(hipGetLastError());
  sync(ws);
  return inv;
}

// The reference's dense direct-solve limit (src/multigrid.jl:19
// _DEFAULT_COARSE_DIRECT_DOF_LIMIT); larger coarse levels are inverted on the device.
constexpr int64_t kHostInverseLimit = 512;

// Largest dense coarse level: rocSOLVER's potrf / potri / getrf / getri take
// 32-bit sizes and leading dimensions, and n x n elements must stay below 2^31
// for their element offsets (46340^2 < 2^31 <= 46341^2).
constexpr int64_t kDenseCoarseMax = 46340;

// Dense coarse inverse from the probed SPD coarse matrix (src/multigrid.jl:382-424).
DeviceBuffer<double> dense_coarse_inverse(gmf_workspace* ws, int64_t n, bool spd) {
  require(n <= kDenseCoarseMax, GMF_EUNSUPPORTED,
          "dense coarse multigrid level has " + std::to_string(n) + " dofs; the device dense inverse supports at most " +
              std::to_string(kDenseCoarseMax) + " (lower coarse_direct_dof_limit or keep more levels)");
  hipStream_t s = ws->stream;
  DeviceBuffer<double> e(static_cast<size_t>(n)), col(static_cast<size_t>(n * n));
  for (int64_t j = 0; j < n; ++j) {
    hipLaunchKernelGGL(k_unit, dim3(grid1(n)), dim3(256), 0, s, e.p, j, n);
    reduced_apply(ws, e.p, col.p + j * n);  // column j of A (stored row j of A^T)
  }
  GMF_HIP(hipGetLastError());
  std::vector<double> At(size_t(n * n));
  GMF_HIP(hipMemcpyAsync(At.data(), col.p, At.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  sync(ws);
  // A[i][j] = At[j*n + i]
  double scale = 0.0;
  for (double v : At) {
    require(std::isfinite(v), spd ? GMF_ENOTSPD : GMF_EINVAL, "coarse multigrid matrix is not finite");
    scale = std::max(scale, std::fabs(v));
  }
  if (spd) {
    const double allowed = std::sqrt(std::numeric_limits<double>::epsilon()) * scale;
    for (int64_t c = 0; c < n; ++c)
      for (int64_t r = c + 1; r < n; ++r)
        require(std::fabs(At[size_t(r * n + c)] - At[size_t(c * n + r)]) <= allowed, GMF_ENOTSPD,
                "declared SPD coarse multigrid matrix is not numerically symmetric");
  }
  if (n > kHostInverseLimit) return device_dense_inverse(ws, n, spd, col);
  std::vector<double> A(size_t(n * n));
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) A[size_t(i * n + j)] = At[size_t(j * n + i)];
  if (!spd) return upload_vec(lu_inverse(A, n));
  // Cholesky A = L L^T on the lower triangle.
  std::vector<double> Lf(A);
  for (int64_t j = 0; j < n; ++j) {
    double d = Lf[size_t(j * n + j)];
    for (int64_t k = 0; k < j; ++k) d -= Lf[size_t(j * n + k)] * Lf[size_t(j * n + k)];
    require(d > 0.0 && std::isfinite(d), GMF_ENOTSPD, "declared SPD coarse multigrid matrix is not positive definite");
    const double ljj = std::sqrt(d);
    Lf[size_t(j * n + j)] = ljj;
    for (int64_t i = j + 1; i < n; ++i) {
      double v = Lf[size_t(i * n + j)];
      for (int64_t k = 0; k < j; ++k) v -= Lf[size_t(i * n + k)] * Lf[size_t(j * n + k)];
      Lf[size_t(i * n + j)] = v / ljj;
    }
  }
  // Inverse column by column: L y = e_j, L^T x = y.
  std::vector<double> inv(size_t(n * n));
  std::vector<double> y(static_cast<size_t>(n));
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t i = 0; i < n; ++i) {
      double v = (i == j) ? 1.0 : 0.0;
      for (int64_t k = 0; k < i; ++k) v -= Lf[size_t(i * n + k)] * y[size_t(k)];
      y[size_t(i)] = v / Lf[size_t(i * n + i)];
    }
    for (int64_t i = n - 1; i >= 0; --i) {
      double v = y[size_t(i)];
      for (int64_t k = i + 1; k < n; ++k) v -= Lf[size_t(k * n + i)] * inv[size_t(k * n + j)];
      inv[size_t(i * n + j)] = v / Lf[size_t(i * n + i)];
    }
  }
  // Symmetrize (the inverse of a symmetric matrix).
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = i + 1; j < n; ++j) {
      const double v = 0.5 * (inv[size_t(i * n + j)] + inv[size_t(j * n + i)]);
      inv[size_t(i * n + j)] = v;
      inv[size_t(j * n + i)] = v;
    }
  return upload_vec(inv);
}


template <class R>
std::unique_ptr<MGBase> build_mg(const gmf_mg_desc* desc) {
  auto mg = std::make_unique<MGPrecond<R>>();
  const int L = desc->level_count;
  mg->ws = desc->levels[L - 1];
  require(mg->ws != nullptr, GMF_EINVAL, "null level workspace");
  mg->smoother = desc->smoother;
  mg->pre = desc->pre_steps;
  mg->post = desc->post_steps;
  mg->omega = desc->damping;
  mg->cheb_degree = desc->chebyshev_degree;
  mg->cheb_range = desc->chebyshev_range;
  mg->coarse_rtol = desc->coarse_rtol;
  mg->coarse_atol = desc->coarse_atol;
  mg->coarse_maxiter = desc->coarse_maxiter;
  mg->restart = desc->smoother_restart;
  mg->spd = desc->spd != 0;
  if (const char* e = std::getenv("GMF_MG_FUSED")) mg->fuse = std::atoi(e) != 0;
  mg->levels.resize(size_t(L));
  const bool cheb = desc->smoother == GMF_SMOOTHER_CHEBYSHEV;
  const bool gmres = desc->smoother == GMF_SMOOTHER_GMRES;
  require(!gmres || MGPrecond<R>::kDouble, GMF_EUNSUPPORTED, "the GMRES smoother runs on FP64 levels only");
  for (int l = 0; l < L; ++l) {
    Level<R>& lv = mg->levels[size_t(l)];
    lv.ws = desc->levels[l];
    require(lv.ws != nullptr, GMF_EINVAL, "null level workspace");
    lv.ws->stream = mg->ws->stream;
    lv.n = lv.ws->plan->n_red;
    const size_t n = size_t(lv.n);
    if (l > 0 && !MGPrecond<R>::kDouble) prepare_fp32_workspace(lv.ws);
    DeviceBuffer<double> inv64(n);
    {
      DeviceBuffer<double> d(n);
      const double* given = desc->level_diagonals ? desc->level_diagonals[l] : nullptr;
      if (given != nullptr) {
        if (n > 0) GMF_HIP(hipMemcpy(d.p, given, n * sizeof(double), hipMemcpyHostToDevice));
      } else {
        reduced_diagonal(lv.ws, d.p);
      }
      invert_diagonal(lv.ws, d.p, inv64.p, lv.n);
    }
    const gmf_block_desc* bd = (desc->level_blocks && l > 0) ? &desc->level_blocks[l] : nullptr;
    BlockSet<double> blocks64;
    if (bd != nullptr && bd->count > 0) {
      require(!gmres, GMF_EUNSUPPORTED, "block smoothing runs with the Jacobi and Chebyshev smoothers");
      require(bd->inverse != nullptr, GMF_EINVAL, "block inverses missing");
      blocks64 = upload_blocks<double>(*bd, lv.n);
      const int64_t nd = bd->offsets[bd->count];
      hipLaunchKernelGGL(k_zero_at<double>, dim3(grid1(nd)), dim3(256), 0, lv.ws->stream, inv64.p, blocks64.dofs.p,
                         nd);
      GMF_HIP(hipGetLastError());
      if constexpr (MGPrecond<R>::kDouble)
        lv.blocks = std::move(blocks64);
      else
        lv.blocks = upload_blocks<R>(*bd, lv.n);
    }
    if (l > 0 && cheb && lv.n > 0) {
      const BlockSet<double>* b64 = &blocks64;
      if constexpr (MGPrecond<R>::kDouble) b64 = &lv.blocks;
      if (bd != nullptr && bd->count > 0)
        lv.lambda_max = 1.2 * estimate_lambda_max_blocks(lv.ws, lv.n, inv64.p, *b64, desc->eigen_iterations);
      else
        lv.lambda_max = 1.2 * estimate_lambda_max(lv.ws, lv.n, inv64.p, desc->eigen_iterations);
      require(std::isfinite(lv.lambda_max) && lv.lambda_max > 0.0, GMF_ENOTSPD,
              "Chebyshev smoother needs a positive spectrum of D^-1 A");
    }
    if (l > 0) {
      lv.invdiag = DeviceBuffer<R>(n);
      if constexpr (MGPrecond<R>::kDouble)
        copy(lv.invdiag.p, inv64.p, lv.n, lv.ws->stream);
      else
        narrow(lv.invdiag.p, inv64.p, lv.n, lv.ws->stream);
      lv.q = DeviceBuffer<R>(n);
      if (cheb || gmres) lv.r = DeviceBuffer<R>(n);
      if (cheb) lv.d = DeviceBuffer<R>(n);
      if (gmres) {
        const int64_t cnt = std::min<int64_t>(std::min(std::max(desc->pre_steps, desc->post_steps), mg->restart), lv.n);
        lv.gV = DeviceBuffer<double>(n * size_t(cnt + 1));
        lv.gZ = DeviceBuffer<double>(n * size_t(std::max<int64_t>(cnt, 1)));
        lv.gy = DeviceBuffer<double>(size_t(std::max<int64_t>(cnt, 1)));
      }
    } else {
      mg->coarse_invdiag = std::move(inv64);
    }
    if (l < L - 1) {
      lv.x = DeviceBuffer<R>(n);
      lv.b = DeviceBuffer<R>(n);
    }
    mg->level_n.push_back(lv.n);
    mg->level_lambda.push_back(lv.lambda_max);
  }
  for (int l = 0; l + 1 < L; ++l)
    mg->transfers.push_back(
        build_transfer<R>(desc->transfers[l], mg->levels[size_t(l)].n, mg->levels[size_t(l + 1)].n));
  Level<R>& C = mg->levels[0];
  if (C.n <= desc->coarse_direct_limit) {
    mg->dense = true;
    if (C.n > 0) mg->coarse_inverse = dense_coarse_inverse(C.ws, C.n, mg->spd);
  } else {
    mg->coarse_jacobi = std::make_unique<JacobiPrecond>();
    mg->coarse_jacobi->ws = C.ws;
    mg->coarse_jacobi->n = C.n;
    mg->coarse_jacobi->invdiag = mg->coarse_invdiag.p;
    mg->coarse_jacobi->owns = false;
    if (mg->spd)
      mg->coarse_vectors = std::make_unique<CgVectors>(C.n);
    else
      mg->coarse_fgmres = std::make_unique<FgmresVectors>(C.n, int(std::max<int64_t>(1, std::min<int64_t>(30, C.n))));
  }
  if (!MGPrecond<R>::kDouble) {
    mg->coarse_b = DeviceBuffer<double>(size_t(C.n));
    mg->coarse_x = DeviceBuffer<double>(size_t(C.n));
  }
  // The V-cycle has no host synchronization when the coarse solve is dense,
  // so it can be captured once and replayed.
  mg->use_graph = desc->graph != 0 && mg->dense && !gmres;
  if (mg->use_graph || !MGPrecond<R>::kDouble) {
    mg->top_b = DeviceBuffer<R>(size_t(mg->levels.back().n));
    mg->top_x = DeviceBuffer<R>(size_t(mg->levels.back().n));
  }
  sync(mg->ws);
  return mg;
}

}  // namespace
}  // namespace gmf

using namespace gmf;

extern "C" {

gmf_status gmf_precond_create_mg(const gmf_mg_desc* desc, gmf_preconditioner** out) {
  return guarded_call([&] {
    require(desc && out, GMF_EINVAL, "null argument");
    require(desc->level_count >= 2 && desc->levels && desc->transfers, GMF_EINVAL,
            "multigrid needs at least two levels and their transfers");
    require(desc->smoother == GMF_SMOOTHER_JACOBI || desc->smoother == GMF_SMOOTHER_CHEBYSHEV ||
                desc->smoother == GMF_SMOOTHER_GMRES,
            GMF_EINVAL, "unknown smoother");
    require(desc->smoother != GMF_SMOOTHER_GMRES ||
                (desc->smoother_restart >= 1 && desc->smoother_restart >= std::max(desc->pre_steps, desc->post_steps)),
            GMF_EINVAL,
            "GMRES multigrid smoother requires smoother_restart to be at least max(pre_smoothing_steps, "
            "post_smoothing_steps)");
    require(desc->pre_steps >= 0 && desc->post_steps >= 0, GMF_EINVAL, "smoothing steps must be non-negative");
    require(std::isfinite(desc->damping) && desc->damping > 0, GMF_EINVAL, "smoother_damping must be positive");
    require(desc->smoother != GMF_SMOOTHER_CHEBYSHEV ||
                (desc->chebyshev_degree >= 1 && desc->chebyshev_range > 1.0 && desc->eigen_iterations >= 1),
            GMF_EINVAL, "invalid Chebyshev parameters");
    require(std::isfinite(desc->coarse_rtol) && desc->coarse_rtol >= 0 && std::isfinite(desc->coarse_atol) &&
                desc->coarse_atol >= 0 && desc->coarse_maxiter >= 1,
            GMF_EINVAL, "invalid coarse solver tolerances");
    require(desc->precision == 0 || desc->precision == 32 || desc->precision == 64, GMF_EINVAL,
            "multigrid precision must be 32 or 64");
    std::unique_ptr<MGBase> mg = desc->precision == 32 ? build_mg<float>(desc) : build_mg<double>(desc);
    *out = mg.release();
  });
}

gmf_status gmf_mg_level_info(const gmf_preconditioner* M, int32_t level, int64_t* reduced_dofs, double* lambda_max) {
  return guarded_call([&] {
    const auto* mg = dynamic_cast<const MGBase*>(M);
    require(mg != nullptr, GMF_EINVAL, "not a multigrid preconditioner");
    require(level >= 0 && size_t(level) < mg->level_n.size(), GMF_EINVAL, "level out of range");
    if (reduced_dofs) *reduced_dofs = mg->level_n[size_t(level)];
    if (lambda_max) *lambda_max = mg->level_lambda[size_t(level)];
  });
}

}  // extern "C"
