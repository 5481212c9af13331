// Generic tensor-product cell kernels: any dimension <= 3, any per-axis mode
// and point counts <= 16, both lowered action forms. One workgroup per cell
// (64, 128 or 256 lanes, picked per bucket by tune_generic), every
// intermediate and the one-dimensional tables staged in LDS.
//
// The algorithm is the reference's interpolate -> per-point action -> project
// pipeline (src/accumulators.jl:1955-2013 with the sum-factorized helpers of
// src/integration.jl:1140-1399): D passes per channel forward, D transposed
// passes per channel back, gather / scatter through the term maps
// (src/assembly.jl:1751-1803). The diagonal kernel evaluates the reference's
// cell_diagonal! (src/accumulators.jl:1056-1083) by sum factorization with
// products of 1D tables instead of one basis probe per mode.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "termmap.hpp"

namespace gmf {
namespace {

struct GenericArgs {
  int D;
  int N[kMaxDim], Q[kMaxDim];
  int Ntot, Qtot, M;
  int form, pstride;
  int toff[kMaxDim];
  const double* values;
  const double* derivs;
  const double* jinv;
  const double* pdata;
  const int32_t* idx;
  const double* lcoef;
  const int64_t* moff;
  const int32_t* midx;
  const double* mcoef;
  int lds_tables;  // 1: the apply kernels copy the tables to LDS (when they fit)
};

// One sum-factorization pass over the block. `shape` holds the input extents
// (axis 0 fastest); the pass replaces the extent of `axis` by `out_ext`.
// Forward (modes -> points): out[.., q, ..] = sum_k T[k][q] in[.., k, ..].
// Backward (points -> modes): out[.., i, ..] = sum_q T[i][q] in[.., q, ..].
// T2, when given, multiplies T elementwise (products of tables).
__device__ void contract(const double* __restrict__ in, double* __restrict__ out, int D, const int* shape,
                         int axis, int out_ext, const double* T, const double* T2, int tq, bool backward) {
  const int in_ext = shape[axis];
  int stride = 1, outer = 1;
  for (int a = 0; a < D; ++a) {
    if (a < axis) stride *= shape[a];
    if (a > axis) outer *= shape[a];
  }
  const int total = stride * out_ext * outer;
  for (int o = threadIdx.x; o < total; o += blockDim.x) {
    const int lo = o % stride;
    const int t = o / stride;
    const int j = t % out_ext;
    const int hi = t / out_ext;
    const double* src = in + lo + hi * stride * in_ext;
    double acc = 0.0;
    for (int k = 0; k < in_ext; ++k) {
      const int ti = backward ? j * tq + k : k * tq + j;
      double w = T[ti];
      if (T2) w *= T2[ti];
      acc = fma(w, src[k * stride], acc);
    }
    out[o] = acc;
  }
}

// Evaluate the coefficient box `u` at the points with derivative along
// `daxis` (-1: value). Result in `dst` (Qtot entries). bufA/bufB: scratch.
__device__ void interpolate(const GenericArgs& a, const double* u, double* dst, double* bufA, double* bufB,
                            int daxis) {
  int shape[kMaxDim] = {1, 1, 1};
  for (int d = 0; d < a.D; ++d) shape[d] = a.N[d];
  const double* cur = u;
  for (int d = 0; d < a.D; ++d) {
    double* nxt = (d == a.D - 1) ? dst : ((d & 1) ? bufB : bufA);
    const double* T = (d == daxis ? a.derivs : a.values) + a.toff[d];
    contract(cur, nxt, a.D, shape, d, a.Q[d], T, nullptr, a.Q[d], false);
    shape[d] = a.Q[d];
    __syncthreads();
    cur = nxt;
  }
}

// Accumulate the projection of point data `f` (Qtot) into r (Ntot):
// r[i] += sum_q f[q] prod_d T1_d[i_d][q_d] (T2_d...). T1 uses the derivative
// table on axis `daxis1`, T2 (if use_t2) on `daxis2`.
__device__ void project_add(const GenericArgs& a, const double* f, double* r, double* bufA, double* bufB,
                            int daxis1, bool use_t2, int daxis2) {
  int shape[kMaxDim] = {1, 1, 1};
  for (int d = 0; d < a.D; ++d) shape[d] = a.Q[d];
  const double* cur = f;
  for (int d = 0; d < a.D; ++d) {
    double* nxt = (d & 1) ? bufB : bufA;
    const double* T1 = (d == daxis1 ? a.derivs : a.values) + a.toff[d];
    const double* T2 = use_t2 ? (d == daxis2 ? a.derivs : a.values) + a.toff[d] : nullptr;
    contract(cur, nxt, a.D, shape, d, a.N[d], T1, T2, a.Q[d], true);
    shape[d] = a.N[d];
    __syncthreads();
    cur = nxt;
  }
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) r[i] += cur[i];
  __syncthreads();
}

__device__ inline TermMap term_map(const GenericArgs& a) {
  return TermMap{a.idx, a.lcoef, a.moff, a.midx, a.mcoef, a.Ntot};
}

__device__ inline double gather_one(const GenericArgs& a, int64_t cell, int i, const double* __restrict__ x) {
  return gather_term(term_map(a), cell, i, x);
}

__device__ inline void scatter_one(const GenericArgs& a, int64_t cell, int i, double v, double* __restrict__ y,
                                   bool squared) {
  scatter_term(term_map(a), cell, i, v, y, squared);
}

// r = A_c u for one cell (u, r in LDS, Ntot entries); `fields` holds (D+1) x
// Qtot point values, bufA / bufB are contraction scratch.
__device__ void apply_body(const GenericArgs& a, int64_t cell, const double* u, double* r, double* bufA,
                           double* bufB, double* fields) {
  const bool general = a.form == GMF_FORM_GENERAL;
  if (general) interpolate(a, u, fields, bufA, bufB, -1);
  for (int d = 0; d < a.D; ++d) interpolate(a, u, fields + (d + 1) * a.Qtot, bufA, bufB, d);

  const double* jinv = a.jinv + cell * a.D;
  if (!general) {
    for (int q = threadIdx.x; q < a.Qtot; q += blockDim.x) {
      const double t = a.pdata[cell * a.Qtot + q];
      for (int d = 0; d < a.D; ++d) {
        double* f = fields + (d + 1) * a.Qtot;
        f[q] = t * jinv[d] * (jinv[d] * f[q]);
      }
    }
  } else {
    const int C = a.D + 1;
    for (int q = threadIdx.x; q < a.Qtot; q += blockDim.x) {
      double ch[kMaxDim + 1];
      ch[0] = fields[q];
      for (int d = 0; d < a.D; ++d) ch[d + 1] = jinv[d] * fields[(d + 1) * a.Qtot + q];
      const double* Mq = a.pdata + (cell * a.Qtot + q) * a.pstride;
      for (int c = 0; c < C; ++c) {
        double acc = 0.0;
        for (int k = 0; k < C; ++k) acc = fma(Mq[c * C + k], ch[k], acc);
        fields[c * a.Qtot + q] = (c == 0) ? acc : acc * jinv[c - 1];
      }
    }
  }
  __syncthreads();
  if (general) project_add(a, fields, r, bufA, bufB, -1, false, -1);
  for (int d = 0; d < a.D; ++d) project_add(a, fields + (d + 1) * a.Qtot, r, bufA, bufB, d, false, -1);
}

// Length of the concatenated per-axis tables (values or derivatives).
__host__ __device__ inline int table_length(const GenericArgs& a) {
  return a.toff[a.D - 1] + a.N[a.D - 1] * a.Q[a.D - 1];
}

// LDS doubles of the apply kernels: u, bufA, bufB, (D+1) fields, r and,
// when staged, the tables.
__host__ __device__ inline int apply_lds_doubles(const GenericArgs& a) {
  return 3 * a.M + (a.D + 1) * a.Qtot + a.Ntot + (a.lds_tables ? 2 * table_length(a) : 0);
}

// One cell: gather, the body with the tables in LDS, scatter.
__device__ inline void apply_cell(const GenericArgs& a, int64_t cell, double* smem, const double* __restrict__ x,
                                  double* __restrict__ y) {
  double* u = smem;
  double* bufA = u + a.M;
  double* bufB = bufA + a.M;
  double* fields = bufB + a.M;              // (D+1) x Qtot
  double* r = fields + (a.D + 1) * a.Qtot;  // Ntot
  double* tabs = r + a.Ntot;                // values, derivatives (if staged)
  const int tl = table_length(a);
  if (a.lds_tables)
    for (int i = threadIdx.x; i < tl; i += blockDim.x) {
      tabs[i] = a.values[i];
      tabs[tl + i] = a.derivs[i];
    }
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
    u[i] = gather_one(a, cell, i, x);
    r[i] = 0.0;
  }
  __syncthreads();
  GenericArgs b = a;
  if (a.lds_tables) {
    b.values = tabs;
    b.derivs = tabs + tl;
  }
  apply_body(b, cell, u, r, bufA, bufB, fields);
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) scatter_one(a, cell, i, r[i], y, false);
}

__global__ void __launch_bounds__(256) cell_apply_generic(GenericArgs a, const double* __restrict__ x,
                                                          double* __restrict__ y) {
  extern __shared__ double smem[];
  apply_cell(a, blockIdx.x, smem, x, y);
}

// Cells of several buckets in one launch: block b of the grid belongs to the
// bucket whose cumulative cell range holds it (buckets ordered by decreasing
// work so the longest cells start first).
__global__ void __launch_bounds__(256) cell_apply_generic_batched(const GenericArgs* __restrict__ args,
                                                                  const int64_t* __restrict__ start, int count,
                                                                  const double* __restrict__ x,
                                                                  double* __restrict__ y) {
  extern __shared__ double smem[];
  const int64_t blk = blockIdx.x;
  int k = 0;
  while (k + 1 < count && blk >= start[k + 1]) ++k;
  apply_cell(args[k], blk - start[k], smem, x, y);
}

// Right-hand side from lowered point loads: one cell per workgroup, loads
// (Qtot x channels per cell: W v, then W g_d) projected with the value table
// (channel 0) and the derivative table on axis d scaled by J^-1_d (channel
// d + 1), scattered through the term maps (src/assembly.jl:488-507).
__global__ void __launch_bounds__(256) cell_rhs_generic(GenericArgs a, const double* __restrict__ loads, int channels,
                                                        double* __restrict__ y) {
  extern __shared__ double smem[];
  const int64_t cell = blockIdx.x;
  double* bufA = smem;
  double* bufB = bufA + a.M;
  double* f = bufB + a.M;  // Qtot
  double* r = f + a.Qtot;  // Ntot
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) r[i] = 0.0;
  const double* L = loads + cell * int64_t(a.Qtot) * channels;
  const double* jinv = a.jinv + cell * a.D;
  for (int c = 0; c < channels; ++c) {
    const double scale = c == 0 ? 1.0 : jinv[c - 1];
    for (int q = threadIdx.x; q < a.Qtot; q += blockDim.x) f[q] = L[int64_t(q) * channels + c] * scale;
    __syncthreads();
    project_add(a, f, r, bufA, bufB, c - 1, false, -1);
  }
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) scatter_one(a, cell, i, r[i], y, false);
}

// Verification sums of one cell: partial[2 c] = sum_q w_q (u_h(x_q) - e_q)^2,
// partial[2 c + 1] = sum_q w_q e_q^2, u_h interpolated from the full-layout
// coefficients (src/verification.jl:201-230). Fixed-order tree reduction.
__global__ void __launch_bounds__(256) cell_l2_generic(GenericArgs a, const double* __restrict__ u,
                                                       const double* __restrict__ w, const double* __restrict__ e,
                                                       double* __restrict__ partial) {
  extern __shared__ double smem[];
  const int64_t cell = blockIdx.x;
  double* uc = smem;
  double* bufA = uc + a.M;
  double* bufB = bufA + a.M;
  double* f = bufB + a.M;             // Qtot
  double* red = f + a.Qtot;           // 2 x blockDim.x
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) uc[i] = gather_one(a, cell, i, u);
  __syncthreads();
  interpolate(a, uc, f, bufA, bufB, -1);
  double s0 = 0.0, s1 = 0.0;
  const int64_t base = cell * int64_t(a.Qtot);
  for (int q = threadIdx.x; q < a.Qtot; q += blockDim.x) {
    const double ev = e[base + q], wq = w[base + q];
    const double d = f[q] - ev;
    s0 = fma(wq * d, d, s0);
    s1 = fma(wq * ev, ev, s1);
  }
  red[threadIdx.x] = s0;
  red[blockDim.x + threadIdx.x] = s1;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (int(threadIdx.x) < off) {
      red[threadIdx.x] += red[threadIdx.x + off];
      red[blockDim.x + threadIdx.x] += red[blockDim.x + threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * cell] = red[0];
    partial[2 * cell + 1] = red[blockDim.x];
  }
}

// Value and physical gradient of u at every point of a cell (the state
// channels of src/accumulators.jl:752-787 _tensor_state_channels):
// out[(cell Q + q) channels + c], c = 0 the value, 1 + a the derivative along a
// (J^-1 scaled); channels = 1 or 1 + D.
__global__ void __launch_bounds__(256) cell_interp_generic(GenericArgs a, const double* __restrict__ u, int channels,
                                                           double* __restrict__ out) {
  extern __shared__ double smem[];
  const int64_t cell = blockIdx.x;
  double* uc = smem;
  double* bufA = uc + a.M;
  double* bufB = bufA + a.M;
  double* f = bufB + a.M;  // Qtot
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) uc[i] = gather_one(a, cell, i, u);
  __syncthreads();
  double* O = out + cell * int64_t(a.Qtot) * channels;
  const double* jinv = a.jinv + cell * a.D;
  for (int c = 0; c < channels; ++c) {
    interpolate(a, uc, f, bufA, bufB, c - 1);
    const double scale = c == 0 ? 1.0 : jinv[c - 1];
    for (int q = threadIdx.x; q < a.Qtot; q += blockDim.x) O[int64_t(q) * channels + c] = f[q] * scale;
    __syncthreads();
  }
}

// Column sums of partial (cells x 2) in a fixed order, in two stages: stage 1
// sums workgroup b's contiguous chunk of `chunk` cells and writes its two sums
// over the chunk's first pair (read before, by the same workgroup); stage 2
// (one workgroup, step = chunk) sums those chunk results in chunk order.
__global__ void __launch_bounds__(256) k_sum_pairs(double* __restrict__ partial, int64_t cells, int64_t chunk,
                                                   int64_t step, double* __restrict__ out) {
  __shared__ double red[2 * 256];
  const int64_t lo = blockIdx.x * chunk, hi = lo + chunk < cells ? lo + chunk : cells;
  double s0 = 0.0, s1 = 0.0;
  for (int64_t c = lo + threadIdx.x * step; c < hi; c += int64_t(blockDim.x) * step) {
    s0 += partial[2 * c];
    s1 += partial[2 * c + 1];
  }
  red[threadIdx.x] = s0;
  red[256 + threadIdx.x] = s1;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (int(threadIdx.x) < off) {
      red[threadIdx.x] += red[threadIdx.x + off];
      red[256 + threadIdx.x] += red[256 + threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (out) {
      out[0] = red[0];
      out[1] = red[256];
    } else if (lo < hi) {
      partial[2 * lo] = red[0];
      partial[2 * lo + 1] = red[256];
    }
  }
}

// One (cell, reduced dof) pair per workgroup: diag[dof] += v^T A_c v with v the
// dof's local coefficient vector.
__global__ void __launch_bounds__(64) cell_diagonal_pairs(GenericArgs a, const int64_t* __restrict__ pcell,
                                                          const int32_t* __restrict__ pdof,
                                                          const int64_t* __restrict__ poff,
                                                          const int32_t* __restrict__ ploc,
                                                          const double* __restrict__ pcoef,
                                                          double* __restrict__ diag) {
  extern __shared__ double smem[];
  const int64_t pair = blockIdx.x;
  const int64_t cell = pcell[pair];
  double* u = smem;
  double* bufA = u + a.M;
  double* bufB = bufA + a.M;
  double* fields = bufB + a.M;
  double* r = fields + (a.D + 1) * a.Qtot;
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
    u[i] = 0.0;
    r[i] = 0.0;
  }
  __syncthreads();
  const int64_t e0 = poff[pair], e1 = poff[pair + 1];
  if (threadIdx.x == 0)
    for (int64_t e = e0; e < e1; ++e) u[ploc[e]] += pcoef[e];
  __syncthreads();
  apply_body(a, cell, u, r, bufA, bufB, fields);
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int64_t e = e0; e < e1; ++e) s = fma(pcoef[e], r[ploc[e]], s);
    unsafeAtomicAdd(&diag[pdof[pair]], s);
  }
}

__global__ void __launch_bounds__(64) cell_diagonal_generic(GenericArgs a, double* __restrict__ diag,
                                                            const uint8_t* __restrict__ skip) {
  extern __shared__ double smem[];
  const int64_t cell = blockIdx.x;
  if (skip && skip[cell]) return;
  double* bufA = smem;
  double* bufB = bufA + a.M;
  double* w = bufB + a.M;   // Qtot
  double* r = w + a.Qtot;   // Ntot
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) r[i] = 0.0;
  const double* jinv = a.jinv + cell * a.D;
  __syncthreads();
  if (a.form != GMF_FORM_GENERAL) {
    for (int d = 0; d < a.D; ++d) {
      for (int q = threadIdx.x; q < a.Qtot; q += blockDim.x)
        w[q] = a.pdata[cell * a.Qtot + q] * jinv[d] * jinv[d];
      __syncthreads();
      project_add(a, w, r, bufA, bufB, d, true, d);
    }
  } else {
    const int C = a.D + 1;
    for (int c = 0; c < C; ++c)
      for (int k = 0; k < C; ++k) {
        for (int q = threadIdx.x; q < a.Qtot; q += blockDim.x) {
          double s = a.pdata[(cell * a.Qtot + q) * a.pstride + c * C + k];
          if (c > 0) s *= jinv[c - 1];
          if (k > 0) s *= jinv[k - 1];
          w[q] = s;
        }
        __syncthreads();
        project_add(a, w, r, bufA, bufB, c - 1, true, k - 1);
      }
  }
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) scatter_one(a, cell, i, r[i], diag, true);
}

GenericArgs make_args(const Bucket& b, bool reduced) {
  GenericArgs a{};
  a.D = b.dim;
  a.M = 1;
  for (int d = 0; d < kMaxDim; ++d) {
    a.N[d] = b.N[d];
    a.Q[d] = b.Q[d];
    a.toff[d] = b.table_offset[d];
  }
  for (int d = 0; d < b.dim; ++d) a.M *= (b.N[d] > b.Q[d] ? b.N[d] : b.Q[d]);
  a.Ntot = b.Ntot;
  a.Qtot = b.Qtot;
  a.form = b.form;
  a.pstride = b.point_stride;
  a.values = b.values;
  a.derivs = b.derivs;
  a.jinv = b.jinv;
  a.pdata = b.pdata;
  a.idx = reduced ? b.ridx : b.lidx;
  a.lcoef = b.lcoef;
  a.moff = b.moff;
  a.midx = reduced ? b.mridx : b.midx;
  a.mcoef = b.mcoef;
  a.lds_tables = 1;
  if (sizeof(double) * size_t(apply_lds_doubles(a)) > 64 * 1024) a.lds_tables = 0;
  return a;
}

size_t generic_apply_lds(const GenericArgs& a) { return sizeof(double) * size_t(apply_lds_doubles(a)); }

void launch_generic_threads(const Bucket& b, const double* x, double* y, bool reduced, int threads,
                            hipStream_t stream) {
  GenericArgs a = make_args(b, reduced);
  require(a.idx != nullptr, GMF_EUNSUPPORTED, "reduced index maps are not available for this plan");
  const size_t lds = generic_apply_lds(a);
  require(lds <= 64 * 1024, GMF_EUNSUPPORTED, "cell batch exceeds the generic kernel's LDS budget");
  hipLaunchKernelGGL(cell_apply_generic, dim3(unsigned(b.cells)), dim3(unsigned(threads)), lds, stream, a, x, y);
  GMF_HIP(hipGetLastError());
}

}  // namespace

void launch_cell_generic(const gmf_plan&, const Bucket& b, const double* x, double* y, bool reduced,
                         hipStream_t stream) {
  if (b.cells == 0) return;
  launch_generic_threads(b, x, y, reduced, b.generic_threads, stream);
}

void tune_generic(gmf_plan& p, TuneTimer& t) {
  int forced = 0;
  if (const char* env = std::getenv("GMF_GENERIC_THREADS")) forced = std::atoi(env);
  for (Bucket& b : p.buckets) {
    if (b.cloud || b.fast || b.cells == 0) continue;
    b.tune.kernel = kTuneGeneric;
    if (forced == 64 || forced == 128 || forced == 256) {
      b.generic_threads = forced;
      b.tune.param0 = forced;
      continue;
    }
    float best = -1.0f;
    for (int threads : {64, 128, 256}) {
      const float us = t.time_us([&] { launch_generic_threads(b, t.x, t.y, false, threads, nullptr); });
      if (threads == 64) b.tune.default_us = us;
      ++b.tune.candidates;
      if (best < 0.0f || us < best) {
        best = us;
        b.generic_threads = threads;
      }
    }
    b.tune.param0 = b.generic_threads;
    b.tune.best_us = best;
  }
}

void prepare_generic_batch(gmf_plan& p) {
  p.generic_batches.clear();
  for (int threads : {64, 128, 256}) {
    std::vector<int> ids;
    for (size_t i = 0; i < p.buckets.size(); ++i) {
      const Bucket& b = p.buckets[i];
      if (!b.cloud && !b.fast && b.cells > 0 && b.generic_threads == threads) ids.push_back(int(i));
    }
    if (ids.size() < 2) continue;
    // heaviest cells first
    std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) {
      const Bucket &A = p.buckets[size_t(a)], &B = p.buckets[size_t(b)];
      return int64_t(A.Ntot) * A.Qtot > int64_t(B.Ntot) * B.Qtot;
    });
    std::vector<GenericArgs> full, red;
    std::vector<int64_t> start(1, 0);
    bool have_red = true, ok = true;
    size_t lds = 0;
    for (int i : ids) {
      const Bucket& b = p.buckets[size_t(i)];
      GenericArgs a = make_args(b, false);
      GenericArgs ar = make_args(b, true);
      if (ar.idx == nullptr) have_red = false;
      full.push_back(a);
      red.push_back(ar);
      start.push_back(start.back() + b.cells);
      const size_t l = generic_apply_lds(a);
      if (a.idx == nullptr || l > 64 * 1024) ok = false;
      lds = std::max(lds, l);
    }
    if (!ok || start.back() >= (int64_t(1) << 31)) continue;
    GenericBatch g;
    g.threads = threads;
    g.count = int(ids.size());
    g.buckets = ids;
    g.cells = start.back();
    g.lds = lds;
    g.args_full = upload_array(p, full.data(), full.size());
    if (have_red) g.args_red = upload_array(p, red.data(), red.size());
    g.start = upload_array(p, start.data(), start.size());
    p.generic_batches.push_back(std::move(g));
  }
}

void launch_generic_batches(const gmf_plan& p, const double* x, double* y, bool reduced, hipStream_t stream,
                            char* done) {
  for (const GenericBatch& g : p.generic_batches) {
    const void* args = reduced ? g.args_red : g.args_full;
    if (args == nullptr) continue;
    hipLaunchKernelGGL(cell_apply_generic_batched, dim3(unsigned(g.cells)), dim3(unsigned(g.threads)), g.lds, stream,
                       static_cast<const GenericArgs*>(args), g.start, g.count, x, y);
    GMF_HIP(hipGetLastError());
    for (int i : g.buckets) done[size_t(i)] = 1;
  }
}

void launch_cell_rhs(const Bucket& b, const double* loads, int channels, double* y, hipStream_t stream) {
  if (b.cells == 0) return;
  GenericArgs a = make_args(b, false);
  require(channels == 1 || channels == a.D + 1, GMF_EINVAL, "load channels must be 1 or 1 + dimension");
  const size_t lds = sizeof(double) * size_t(2 * a.M + a.Qtot + a.Ntot);
  require(lds <= 64 * 1024, GMF_EUNSUPPORTED, "cell batch exceeds the generic kernel's LDS budget");
  hipLaunchKernelGGL(cell_rhs_generic, dim3(unsigned(b.cells)), dim3(64), lds, stream, a, loads, channels, y);
  GMF_HIP(hipGetLastError());
}

void launch_cell_l2(const Bucket& b, const double* u, const double* w, const double* e, double* partial,
                    hipStream_t stream) {
  if (b.cells == 0) return;
  if (launch_cell_l2_fast(b, u, w, e, partial, stream)) return;
  GenericArgs a = make_args(b, false);
  const int threads = 64;
  const size_t lds = sizeof(double) * size_t(3 * a.M + a.Qtot + 2 * threads);
  require(lds <= 64 * 1024, GMF_EUNSUPPORTED, "cell batch exceeds the generic kernel's LDS budget");
  hipLaunchKernelGGL(cell_l2_generic, dim3(unsigned(b.cells)), dim3(threads), lds, stream, a, u, w, e, partial);
  GMF_HIP(hipGetLastError());
}

void launch_cell_interp(const Bucket& b, const double* u, int channels, double* out, hipStream_t stream) {
  if (b.cells == 0) return;
  GenericArgs a = make_args(b, false);
  require(channels == 1 || channels == a.D + 1, GMF_EINVAL, "state channels must be 1 or 1 + dimension");
  const size_t lds = sizeof(double) * size_t(3 * a.M + a.Qtot);
  require(lds <= 64 * 1024, GMF_EUNSUPPORTED, "cell batch exceeds the generic kernel's LDS budget");
  hipLaunchKernelGGL(cell_interp_generic, dim3(unsigned(b.cells)), dim3(64), lds, stream, a, u, channels, out);
  GMF_HIP(hipGetLastError());
}

void launch_sum_pairs(double* partial, int64_t cells, double* out, hipStream_t stream) {
  // chunks of at least 4096 cells, at most 1024 of them
  const int64_t chunk = std::max<int64_t>(4096, (cells + 1023) / 1024);
  const int64_t blocks = std::max<int64_t>(1, (cells + chunk - 1) / chunk);
  if (blocks > 1) {
    hipLaunchKernelGGL(k_sum_pairs, dim3(unsigned(blocks)), dim3(256), 0, stream, partial, cells, chunk, int64_t(1),
                       static_cast<double*>(nullptr));
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(256), 0, stream, partial, cells, cells, chunk, out);
  } else {
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(256), 0, stream, partial, cells, cells, int64_t(1), out);
  }
  GMF_HIP(hipGetLastError());
}

void launch_cell_diagonal_generic(const gmf_plan&, const Bucket& b, double* diag, bool reduced,
                                  hipStream_t stream) {
  if (b.cells == 0) return;
  GenericArgs a = make_args(b, reduced);
  require(a.idx != nullptr, GMF_EUNSUPPORTED, "reduced index maps are not available for this plan");
  size_t lds = sizeof(double) * size_t(2 * a.M + a.Qtot + a.Ntot);
  require(lds <= 64 * 1024, GMF_EUNSUPPORTED, "cell batch exceeds the generic kernel's LDS budget");
  hipLaunchKernelGGL(cell_diagonal_generic, dim3(unsigned(b.cells)), dim3(64), lds, stream, a, diag,
                     static_cast<const uint8_t*>(nullptr));
  GMF_HIP(hipGetLastError());
}

void launch_cell_diagonal_generic_skip(const gmf_plan&, const Bucket& b, double* diag, hipStream_t stream) {
  if (b.cells == 0) return;
  GenericArgs a = make_args(b, true);
  require(a.idx != nullptr, GMF_EUNSUPPORTED, "reduced index maps are not available for this plan");
  size_t lds = sizeof(double) * size_t(2 * a.M + a.Qtot + a.Ntot);
  require(lds <= 64 * 1024, GMF_EUNSUPPORTED, "cell batch exceeds the generic kernel's LDS budget");
  hipLaunchKernelGGL(cell_diagonal_generic, dim3(unsigned(b.cells)), dim3(64), lds, stream, a, diag,
                     static_cast<const uint8_t*>(b.coupled));
  GMF_HIP(hipGetLastError());
}

void launch_cell_diagonal_pairs(const gmf_plan&, const Bucket& b, double* diag, hipStream_t stream) {
  if (b.pair_count == 0) return;
  GenericArgs a = make_args(b, true);
  size_t lds = sizeof(double) * size_t(3 * a.M + (a.D + 1) * a.Qtot + a.Ntot);
  require(lds <= 64 * 1024, GMF_EUNSUPPORTED, "cell batch exceeds the generic kernel's LDS budget");
  require(b.pair_count < (int64_t(1) << 31), GMF_EUNSUPPORTED, "too many coupled diagonal pairs");
  hipLaunchKernelGGL(cell_diagonal_pairs, dim3(unsigned(b.pair_count)), dim3(64), lds, stream, a, b.pair_cell,
                     b.pair_dof, b.pair_off, b.pair_loc, b.pair_coef, diag);
  GMF_HIP(hipGetLastError());
}

}  // namespace gmf
