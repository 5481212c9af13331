// Point-cloud cell kernels: cells or embedded-surface pieces that carry their
// own quadrature points (cut-cell rules, surface rules, boundary faces)
// instead of a tensor rule, and two-sided interface items. The reference
// evaluates these through full basis tables (src/accumulators.jl:1022-1054
// _accumulate_cell_accumulator_basis!, 1133-1161
// _accumulate_one_sided_apply_basis!, 1576-1616 interface_apply!); here the
// integrated Legendre factors psi_k(xi_a), psi_k'(xi_a) are recomputed at each
// point on the device (src/polynomials.jl:403-434), so only the reference
// coordinates and the lowered per-point action travel from HBM.
//
// An item has S sides (S = 1: one leaf; S = 2: the minus and plus leaves of
// an interface, src/integration.jl:187-207 InterfaceValues). Its local box is
// the concatenation of the sides' mode boxes, every point carries reference
// coordinates in each side's leaf, and the point action is the
// S(D+1) x S(D+1) matrix acting on the channels (value, physical gradient) of
// all sides.
//
// One 64-lane workgroup per item. Points are processed in chunks of at most 64
// (fewer when the LDS tables of large boxes would exceed 64 KB): lanes over
// points evaluate the 1D factors into LDS, interpolate value and gradient on
// each side, apply the point action; then lanes over the local modes project
// the chunk's channels back onto their side.
#include <cmath>
#include <cstdlib>
#include <string>

#include "common.hpp"
#include "internal.hpp"
#include "termmap.hpp"

namespace gmf {
namespace {

constexpr int kCloudLanes = 64;
constexpr int kMaxSides = 2;

struct CloudArgs {
  int D, S;                    // dimension; sides per item
  int N[kMaxSides][kMaxDim];   // modes per axis of each side
  int Ns[kMaxSides];           // box size of each side
  int Ntot, Nmax;              // sum of the box sizes; largest N over sides and axes
  int chunk;                   // points per chunk = LDS stride of the per-point channel arrays
  int ts;                      // row stride of the 1D factor tables: chunk + 1, so rows of one
                               // point column fall in different LDS banks
  const double* jinv;          // items x S*D
  const double* pdata;         // points x (S(D+1))^2
  const int64_t* poff;         // items + 1
  const double* rpts;          // points x S*D
  TermMap map;                 // items x Ntot
  const int32_t* unit_item;    // work unit -> item
  const int64_t* unit_start;   // work unit -> first point
  int unit_points;             // points per work unit
  int lds_kb;                  // LDS budget of one workgroup in KB
};

// Item and point range [pb, pe) of the apply kernels' work unit `u`.
__device__ inline int64_t unit_range(const CloudArgs& a, int64_t u, int64_t& pb, int64_t& pe) {
  const int64_t item = a.unit_item[u];
  pb = a.unit_start[u];
  const int64_t end = a.poff[item + 1];
  pe = end - pb < a.unit_points ? end : pb + a.unit_points;
  return item;
}

// psi_0..psi_{n-1} and derivatives at x into v[k * stride], d[k * stride].
__device__ inline void integrated_legendre(double x, int n, int stride, double* v, double* d) {
  if (n == 1) {
    v[0] = 1.0;
    d[0] = 0.0;
    return;
  }
  v[0] = 0.5 * (1.0 - x);
  d[0] = -0.5;
  v[stride] = 0.5 * (1.0 + x);
  d[stride] = 0.5;
  double pm2 = 1.0, pm1 = x;  // P_{k-2}, P_{k-1}
  for (int k = 2; k < n; ++k) {
    const double pk = ((2.0 * k - 1.0) * x * pm1 - (k - 1.0) * pm2) / double(k);
    v[k * stride] = (pk - pm2) / sqrt(4.0 * k - 2.0);
    d[k * stride] = sqrt((2.0 * k - 1.0) / 2.0) * pm1;
    pm2 = pm1;
    pm1 = pk;
  }
}

// Side of local mode i and its mode multi-index (axis 0 fastest).
__device__ inline int unflatten(const CloudArgs& a, int i, int (&m)[kMaxDim]) {
  const int s = (a.S > 1 && i >= a.Ns[0]) ? 1 : 0;
  const int il = s ? i - a.Ns[0] : i;
  m[0] = il % a.N[s][0];
  const int t = il / a.N[s][0];
  m[1] = a.D > 1 ? t % a.N[s][1] : 0;
  m[2] = a.D > 2 ? t / a.N[s][1] : 0;
  return s;
}

// Offset of side s's tables in tv / td ([side][axis][k][lane]).
__device__ inline int table_base(const CloudArgs& a, int s) { return s * a.D * a.Nmax * a.ts; }

// Basis value phi and reference derivatives dphi[ax] of multi-index m on
// side s at chunk lane q.
__device__ inline void basis_at(const CloudArgs& a, const double* tv, const double* td, int s, const int (&m)[kMaxDim],
                                int q, double& phi, double (&dphi)[kMaxDim]) {
  double f[kMaxDim], fd[kMaxDim];
  const int base = table_base(a, s);
  for (int ax = 0; ax < a.D; ++ax) {
    f[ax] = tv[base + (ax * a.Nmax + m[ax]) * a.ts + q];
    fd[ax] = td[base + (ax * a.Nmax + m[ax]) * a.ts + q];
  }
  phi = 1.0;
  for (int ax = 0; ax < a.D; ++ax) phi *= f[ax];
  for (int ax = 0; ax < a.D; ++ax) {
    double v = fd[ax];
    for (int bx = 0; bx < a.D; ++bx)
      if (bx != ax) v *= f[bx];
    dphi[ax] = v;
  }
}

// Lane q < cnt of a chunk: 1D factors of point p at every side into tv / td.
__device__ inline void point_tables(const CloudArgs& a, int64_t p, int q, double* tv, double* td) {
  for (int s = 0; s < a.S; ++s)
    for (int ax = 0; ax < a.D; ++ax) {
      const int off = table_base(a, s) + ax * a.Nmax * a.ts + q;
      integrated_legendre(a.rpts[p * a.S * a.D + s * a.D + ax], a.N[s][ax], a.ts, tv + off, td + off);
    }
}

// Phase 1 for one chunk (lanes q < cnt): value and physical gradient of u on
// every side, point action into ch[r][lane] (channel r = s(D+1) + c; c = 0
// value, 1 + a gradient along a), gradient channels already scaled by J^-1.
__device__ void chunk_forward(const CloudArgs& a, int64_t item, int64_t p0, int cnt, const double* u, double* tv,
                              double* td, double* ch) {
  const int q = threadIdx.x;
  if (q >= cnt) return;
  const int C = a.D + 1;
  const int SC = a.S * C;
  const double* jinv = a.jinv + item * a.S * a.D;
  point_tables(a, p0 + q, q, tv, td);
  double chn[kMaxSides * (kMaxDim + 1)];
  for (int s = 0; s < a.S; ++s) {
    double val = 0.0, g[kMaxDim] = {0.0, 0.0, 0.0};
    const int first = s ? a.Ns[0] : 0;
    for (int i = 0; i < a.Ns[s]; ++i) {
      int m[kMaxDim];
      unflatten(a, first + i, m);
      double phi, dphi[kMaxDim];
      basis_at(a, tv, td, s, m, q, phi, dphi);
      const double ui = u[first + i];
      val = fma(ui, phi, val);
      for (int ax = 0; ax < a.D; ++ax) g[ax] = fma(ui, dphi[ax], g[ax]);
    }
    chn[s * C] = val;
    for (int ax = 0; ax < a.D; ++ax) chn[s * C + 1 + ax] = jinv[s * a.D + ax] * g[ax];
  }
  const double* M = a.pdata + (p0 + q) * SC * SC;
  for (int r = 0; r < SC; ++r) {
    double acc = 0.0;
    for (int k = 0; k < SC; ++k) acc = fma(M[r * SC + k], chn[k], acc);
    const int s = r / C, c = r % C;
    ch[r * a.chunk + q] = c == 0 ? acc : acc * jinv[s * a.D + c - 1];
  }
}

__global__ void __launch_bounds__(kCloudLanes) cell_apply_cloud(CloudArgs a, const double* __restrict__ x,
                                                                double* __restrict__ y) {
  extern __shared__ double smem[];
  int64_t pb, pe;
  const int64_t item = unit_range(a, blockIdx.x, pb, pe);
  const int C = a.D + 1;
  double* u = smem;
  double* r = u + a.Ntot;
  double* tv = r + a.Ntot;
  double* td = tv + a.S * a.D * a.Nmax * a.ts;
  double* ch = td + a.S * a.D * a.Nmax * a.ts;
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
    u[i] = gather_term(a.map, item, i, x);
    r[i] = 0.0;
  }
  __syncthreads();
  for (int64_t p0 = pb; p0 < pe; p0 += a.chunk) {
    const int cnt = int(pe - p0 < a.chunk ? pe - p0 : a.chunk);
    chunk_forward(a, item, p0, cnt, u, tv, td, ch);
    __syncthreads();
    for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
      int m[kMaxDim];
      const int s = unflatten(a, i, m);
      const double* chs = ch + s * C * a.chunk;
      double acc = 0.0;
      for (int q = 0; q < cnt; ++q) {
        double phi, dphi[kMaxDim];
        basis_at(a, tv, td, s, m, q, phi, dphi);
        acc = fma(chs[q], phi, acc);
        for (int ax = 0; ax < a.D; ++ax) acc = fma(chs[(ax + 1) * a.chunk + q], dphi[ax], acc);
      }
      r[i] += acc;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) scatter_term(a.map, item, i, r[i], y, false);
}

// Point loads projected onto the items' modes (src/accumulators.jl:1022-1054
// with the load channels in place of the point action's output, 1230-1274
// _accumulate_one_sided_rhs!, 1666-1698 interface_rhs!): for every point and
// side the channels (W v, W g_1, .., W g_D) of the load callbacks, gradient
// channels mapped by J^-1 onto the reference derivatives, then the
// projection and the atomic term scatter of cell_apply_cloud. Work units as
// in the apply.
__global__ void __launch_bounds__(kCloudLanes) cell_rhs_cloud(CloudArgs a, const double* __restrict__ loads,
                                                              double* __restrict__ y) {
  extern __shared__ double smem[];
  int64_t pb, pe;
  const int64_t item = unit_range(a, blockIdx.x, pb, pe);
  const int C = a.D + 1;
  double* r = smem;
  double* tv = r + a.Ntot;
  double* td = tv + a.S * a.D * a.Nmax * a.ts;
  double* ch = td + a.S * a.D * a.Nmax * a.ts;
  const double* jinv = a.jinv + item * a.S * a.D;
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) r[i] = 0.0;
  for (int64_t p0 = pb; p0 < pe; p0 += a.chunk) {
    const int cnt = int(pe - p0 < a.chunk ? pe - p0 : a.chunk);
    const int q = threadIdx.x;
    if (q < cnt) {
      point_tables(a, p0 + q, q, tv, td);
      const double* l = loads + (p0 + q) * a.S * C;
      for (int s = 0; s < a.S; ++s) {
        ch[(s * C) * a.chunk + q] = l[s * C];
        for (int ax = 0; ax < a.D; ++ax)
          ch[(s * C + 1 + ax) * a.chunk + q] = l[s * C + 1 + ax] * jinv[s * a.D + ax];
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
      int m[kMaxDim];
      const int s = unflatten(a, i, m);
      const double* chs = ch + s * C * a.chunk;
      double acc = 0.0;
      for (int qq = 0; qq < cnt; ++qq) {
        double phi, dphi[kMaxDim];
        basis_at(a, tv, td, s, m, qq, phi, dphi);
        acc = fma(chs[qq], phi, acc);
        for (int ax = 0; ax < a.D; ++ax) acc = fma(chs[(ax + 1) * a.chunk + qq], dphi[ax], acc);
      }
      r[i] += acc;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) scatter_term(a.map, item, i, r[i], y, false);
}

// Verification sums of one work unit of a one-sided batch
// (src/verification.jl:201-230 _l2_error_components on cells with their own
// points): partial[2u] = sum_q w_q (u_h(x_q) - e_q)^2, partial[2u + 1] =
// sum_q w_q e_q^2, u_h from the item's gathered coefficients. Each lane reads
// only its own point's column of the 1D factor tables.
__global__ void __launch_bounds__(kCloudLanes) cell_l2_cloud(CloudArgs a, const double* __restrict__ u_full,
                                                             const double* __restrict__ w,
                                                             const double* __restrict__ e,
                                                             double* __restrict__ partial) {
  extern __shared__ double smem[];
  int64_t pb, pe;
  const int64_t item = unit_range(a, blockIdx.x, pb, pe);
  double* u = smem;
  double* tv = u + a.Ntot;
  double* td = tv + a.D * a.Nmax * a.ts;
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) u[i] = gather_term(a.map, item, i, u_full);
  __syncthreads();
  double se = 0.0, sr = 0.0;
  const int q = threadIdx.x;
  for (int64_t p0 = pb; p0 < pe; p0 += a.chunk) {
    const int cnt = int(pe - p0 < a.chunk ? pe - p0 : a.chunk);
    if (q < cnt) {
      point_tables(a, p0 + q, q, tv, td);
      double val = 0.0;
      for (int i = 0; i < a.Ntot; ++i) {
        int m[kMaxDim];
        unflatten(a, i, m);
        double phi, dphi[kMaxDim];
        basis_at(a, tv, td, 0, m, q, phi, dphi);
        val = fma(u[i], phi, val);
      }
      const double wq = w[p0 + q], eq = e[p0 + q], d = val - eq;
      se = fma(wq * d, d, se);
      sr = fma(wq * eq, eq, sr);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    se += __shfl_down(se, off);
    sr += __shfl_down(sr, off);
  }
  if (q == 0) {
    partial[2 * blockIdx.x] = se;
    partial[2 * blockIdx.x + 1] = sr;
  }
}

// Local matrices of point-cloud items: column j of item c's matrix is the
// item's local action on the unit vector of local mode j, computed with the
// arithmetic of cell_apply_cloud over all of the item's points (one
// workgroup per item and column). Stored column-major per item:
// E[c][j][i] at emat[(c * Ntot + j) * Ntot + i].
__global__ void __launch_bounds__(kCloudLanes) cell_local_matrix_cloud(CloudArgs a, double* __restrict__ emat) {
  extern __shared__ double smem[];
  const int64_t item = blockIdx.x;
  const int j = int(blockIdx.y);
  const int C = a.D + 1;
  double* u = smem;
  double* r = u + a.Ntot;
  double* tv = r + a.Ntot;
  double* td = tv + a.S * a.D * a.Nmax * a.ts;
  double* ch = td + a.S * a.D * a.Nmax * a.ts;
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
    u[i] = i == j ? 1.0 : 0.0;
    r[i] = 0.0;
  }
  __syncthreads();
  const int64_t pb = a.poff[item], pe = a.poff[item + 1];
  for (int64_t p0 = pb; p0 < pe; p0 += a.chunk) {
    const int cnt = int(pe - p0 < a.chunk ? pe - p0 : a.chunk);
    chunk_forward(a, item, p0, cnt, u, tv, td, ch);
    __syncthreads();
    for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
      int m[kMaxDim];
      const int s = unflatten(a, i, m);
      const double* chs = ch + s * C * a.chunk;
      double acc = 0.0;
      for (int q = 0; q < cnt; ++q) {
        double phi, dphi[kMaxDim];
        basis_at(a, tv, td, s, m, q, phi, dphi);
        acc = fma(chs[q], phi, acc);
        for (int ax = 0; ax < a.D; ++ax) acc = fma(chs[(ax + 1) * a.chunk + q], dphi[ax], acc);
      }
      r[i] += acc;
    }
    __syncthreads();
  }
  double* col = emat + (item * a.Ntot + j) * a.Ntot;
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) col[i] = r[i];
}

constexpr int kMatrixItemsPerGroup = 4;  // one wavefront per item
// Largest local-matrix store of one point-cloud batch, and the largest local
// box (the apply stages 4 items x 8 columns of local coefficients in LDS).
constexpr int64_t kMaxLocalMatrixBytes = int64_t(2) << 30;
constexpr int kMaxLocalMatrixModes = 128;

// Y[:, j0 + c] += P_c E_c P_c^T X[:, j0 + c] for the KC columns of one column
// group: one wavefront per item gathers its local coefficients of every
// column into LDS; lane i forms rows i, i + 64, .. of E x (columns of E are
// contiguous, so the lanes read consecutive addresses), each summed over the
// local modes in order, and scatters them through the term map.
template <int KC>
__global__ void __launch_bounds__(64 * kMatrixItemsPerGroup)
    cell_apply_local_matrix(TermMap map, const double* __restrict__ emat, int64_t items, const double* __restrict__ X,
                            double* __restrict__ Y, int64_t ld, int k) {
  extern __shared__ double smem[];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t item = int64_t(blockIdx.x) * kMatrixItemsPerGroup + w;
  const int j0 = int(blockIdx.y) * KC;
  const int kc = k - j0 < KC ? k - j0 : KC;
  const int n = map.Ntot;
  if (item >= items) return;  // no workgroup barrier below: each wavefront owns its LDS slice
  double* xs = smem + int64_t(w) * KC * n;  // KC x n
  for (int i = lane; i < n; i += 64)
    for (int c = 0; c < kc; ++c) xs[c * n + i] = gather_term(map, item, i, X + int64_t(j0 + c) * ld);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const double* E = emat + item * int64_t(n) * n;
  for (int i = lane; i < n; i += 64) {
    double t[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) t[c] = 0.0;
    for (int jj = 0; jj < n; ++jj) {
      const double e = E[int64_t(jj) * n + i];
#pragma unroll
      for (int c = 0; c < KC; ++c) t[c] = fma(e, xs[c * n + jj], t[c]);
    }
#pragma unroll
    for (int c = 0; c < KC; ++c)
      if (c < kc) scatter_term(map, item, i, t[c], Y + int64_t(j0 + c) * ld, false);
  }
}

// Basis value and reference derivatives with the dimension known at compile
// time (the products of basis_at in the same order, kept in registers).
template <int D>
__device__ inline void basis_static(const CloudArgs& a, const double* tv, const double* td, int s,
                                    const int (&m)[kMaxDim], int q, double& phi, double (&dphi)[kMaxDim]) {
  double f[D], fd[D];
  const int base = table_base(a, s);
#pragma unroll
  for (int ax = 0; ax < D; ++ax) {
    f[ax] = tv[base + (ax * a.Nmax + m[ax]) * a.ts + q];
    fd[ax] = td[base + (ax * a.Nmax + m[ax]) * a.ts + q];
  }
  phi = 1.0;
#pragma unroll
  for (int ax = 0; ax < D; ++ax) phi *= f[ax];
#pragma unroll
  for (int ax = 0; ax < D; ++ax) {
    double v = fd[ax];
#pragma unroll
    for (int bx = 0; bx < D; ++bx)
      if (bx != ax) v *= f[bx];
    dphi[ax] = v;
  }
}

// Several columns per workgroup (block applies over right-hand sides stored
// column after column, leading dimension ld): the 1D factors of a chunk's
// points and the point action serve KC columns, while every column runs the
// arithmetic of cell_apply_cloud in the same order, so each item's local
// result per column equals the one-column kernel's.
template <int KC, int D, int NF>
__global__ void __launch_bounds__(kCloudLanes) cell_apply_cloud_block(CloudArgs a, const double* __restrict__ X,
                                                                      double* __restrict__ Y, int64_t ld, int k) {
  extern __shared__ double smem[];
  constexpr int C = D + 1;
  constexpr int kMaxSC = kMaxSides * C;
  int64_t pb, pe;
  const int64_t item = unit_range(a, blockIdx.x, pb, pe);
  const int j0 = int(blockIdx.y) * KC;
  const int kc = k - j0 < KC ? k - j0 : KC;
  const int SC = a.S * C;
  double* u = smem;                                  // KC x Ntot
  double* r = u + KC * a.Ntot;                       // KC x Ntot
  double* tv = r + KC * a.Ntot;                      // S x D x Nmax x chunk
  double* td = tv + a.S * D * a.Nmax * a.ts;         // S x D x Nmax x ts
  double* ch = td + a.S * D * a.Nmax * a.ts;         // KC x SC x chunk
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x)
    for (int j = 0; j < kc; ++j) {
      u[j * a.Ntot + i] = gather_term(a.map, item, i, X + int64_t(j0 + j) * ld);
      r[j * a.Ntot + i] = 0.0;
    }
  __syncthreads();
  const double* jinv = a.jinv + item * a.S * D;
  for (int64_t p0 = pb; p0 < pe; p0 += a.chunk) {
    const int cnt = int(pe - p0 < a.chunk ? pe - p0 : a.chunk);
    const int q = threadIdx.x;
    if (q < cnt) {
      // Each lane writes and reads only its own point's column of tv / td.
      point_tables(a, p0 + q, q, tv, td);
      if constexpr (NF > 0) {
        // One side, 3D, NF modes per axis: the point's 1D factors stay in
        // registers; the products and the summation order are those of
        // basis_static over the modes in unflatten order.
        static_assert(D == 3, "register forward is 3D");
        double f[3][NF], fd[3][NF];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
#pragma unroll
          for (int kk = 0; kk < NF; ++kk) {
            f[ax][kk] = tv[(ax * a.Nmax + kk) * a.ts + q];
            fd[ax][kk] = td[(ax * a.Nmax + kk) * a.ts + q];
          }
        double val[KC], g[KC][3];
#pragma unroll
        for (int j = 0; j < KC; ++j) {
          val[j] = 0.0;
          g[j][0] = g[j][1] = g[j][2] = 0.0;
        }
#pragma unroll
        for (int i2 = 0; i2 < NF; ++i2)
#pragma unroll
          for (int i1 = 0; i1 < NF; ++i1)
#pragma unroll
            for (int i0 = 0; i0 < NF; ++i0) {
              const double phi = f[0][i0] * f[1][i1] * f[2][i2];
              const double d0 = fd[0][i0] * f[1][i1] * f[2][i2];
              const double d1 = fd[1][i1] * f[0][i0] * f[2][i2];
              const double d2 = fd[2][i2] * f[0][i0] * f[1][i1];
              const int i = i0 + NF * (i1 + NF * i2);
#pragma unroll
              for (int j = 0; j < KC; ++j) {
                if (j < kc) {
                  const double ui = u[j * a.Ntot + i];
                  val[j] = fma(ui, phi, val[j]);
                  g[j][0] = fma(ui, d0, g[j][0]);
                  g[j][1] = fma(ui, d1, g[j][1]);
                  g[j][2] = fma(ui, d2, g[j][2]);
                }
              }
            }
#pragma unroll
        for (int j = 0; j < KC; ++j) {
          if (j < kc) {
            double* cj = ch + j * SC * a.chunk + q;
            cj[0] = val[j];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) cj[(1 + ax) * a.chunk] = jinv[ax] * g[j][ax];
          }
        }
      } else
      for (int s = 0; s < a.S; ++s) {
        double val[KC], g[KC][D];
#pragma unroll
        for (int j = 0; j < KC; ++j) {
          val[j] = 0.0;
#pragma unroll
          for (int ax = 0; ax < D; ++ax) g[j][ax] = 0.0;
        }
        const int first = s ? a.Ns[0] : 0;
        for (int i = 0; i < a.Ns[s]; ++i) {
          int m[kMaxDim];
          unflatten(a, first + i, m);
          double phi, dphi[kMaxDim];
          basis_static<D>(a, tv, td, s, m, q, phi, dphi);
#pragma unroll
          for (int j = 0; j < KC; ++j) {
            if (j < kc) {
              const double ui = u[j * a.Ntot + first + i];
              val[j] = fma(ui, phi, val[j]);
#pragma unroll
              for (int ax = 0; ax < D; ++ax) g[j][ax] = fma(ui, dphi[ax], g[j][ax]);
            }
          }
        }
#pragma unroll
        for (int j = 0; j < KC; ++j) {
          if (j < kc) {
            double* cj = ch + (j * SC + s * C) * a.chunk + q;
            cj[0] = val[j];
#pragma unroll
            for (int ax = 0; ax < D; ++ax) cj[(1 + ax) * a.chunk] = jinv[s * D + ax] * g[j][ax];
          }
        }
      }
      const double* M = a.pdata + (p0 + q) * SC * SC;
      for (int j = 0; j < kc; ++j) {
        double chn[kMaxSC];
        double* cj = ch + j * SC * a.chunk + q;
#pragma unroll
        for (int kk = 0; kk < kMaxSC; ++kk) chn[kk] = kk < SC ? cj[kk * a.chunk] : 0.0;
        for (int rr = 0; rr < SC; ++rr) {
          double acc = 0.0;
#pragma unroll
          for (int kk = 0; kk < kMaxSC; ++kk)
            if (kk < SC) acc = fma(M[rr * SC + kk], chn[kk], acc);
          const int s = rr / C, c = rr % C;
          cj[rr * a.chunk] = c == 0 ? acc : acc * jinv[s * D + c - 1];
        }
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
      int m[kMaxDim];
      const int s = unflatten(a, i, m);
      double acc[KC];
#pragma unroll
      for (int j = 0; j < KC; ++j) acc[j] = 0.0;
      for (int qq = 0; qq < cnt; ++qq) {
        double phi, dphi[kMaxDim];
        basis_static<D>(a, tv, td, s, m, qq, phi, dphi);
#pragma unroll
        for (int j = 0; j < KC; ++j) {
          if (j < kc) {
            const double* chs = ch + (j * SC + s * C) * a.chunk;
            acc[j] = fma(chs[qq], phi, acc[j]);
#pragma unroll
            for (int ax = 0; ax < D; ++ax) acc[j] = fma(chs[(ax + 1) * a.chunk + qq], dphi[ax], acc[j]);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < KC; ++j)
        if (j < kc) r[j * a.Ntot + i] += acc[j];
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x)
    for (int j = 0; j < kc; ++j) scatter_term(a.map, item, i, r[j * a.Ntot + i], Y + int64_t(j0 + j) * ld, false);
}

// Local diagonal d_i = sum_q sum_{c,k} M_q[(s,c)][(s,k)] chi_c chi_k with
// chi_0 = phi_i, chi_{1+a} = J^-1_a d_a phi_i on the side s of mode i (the
// other side's channels of a trace basis function are zero,
// src/accumulators.jl:723-736), scattered with squared coefficients.
__global__ void __launch_bounds__(kCloudLanes) cell_diagonal_cloud(CloudArgs a, double* __restrict__ diag) {
  extern __shared__ double smem[];
  const int64_t item = blockIdx.x;
  const int C = a.D + 1;
  const int SC = a.S * C;
  double* r = smem;
  double* tv = r + a.Ntot;
  double* td = tv + a.S * a.D * a.Nmax * a.ts;
  double* Ms = td + a.S * a.D * a.Nmax * a.ts;  // S x C x C x chunk: diagonal blocks
  const double* jinv = a.jinv + item * a.S * a.D;
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) r[i] = 0.0;
  __syncthreads();
  const int64_t pb = a.poff[item], pe = a.poff[item + 1];
  for (int64_t p0 = pb; p0 < pe; p0 += a.chunk) {
    const int cnt = int(pe - p0 < a.chunk ? pe - p0 : a.chunk);
    const int q = threadIdx.x;
    if (q < cnt) {
      point_tables(a, p0 + q, q, tv, td);
      const double* M = a.pdata + (p0 + q) * SC * SC;
      for (int s = 0; s < a.S; ++s)
        for (int c = 0; c < C; ++c)
          for (int k = 0; k < C; ++k)
            Ms[((s * C + c) * C + k) * a.chunk + q] = M[(s * C + c) * SC + s * C + k];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
      int m[kMaxDim];
      const int s = unflatten(a, i, m);
      double acc = 0.0;
      for (int qq = 0; qq < cnt; ++qq) {
        double phi, dphi[kMaxDim], chi[kMaxDim + 1];
        basis_at(a, tv, td, s, m, qq, phi, dphi);
        chi[0] = phi;
        for (int ax = 0; ax < a.D; ++ax) chi[ax + 1] = jinv[s * a.D + ax] * dphi[ax];
        for (int c = 0; c < C; ++c)
          for (int k = 0; k < C; ++k) acc = fma(Ms[((s * C + c) * C + k) * a.chunk + qq] * chi[c], chi[k], acc);
      }
      r[i] += acc;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) scatter_term(a.map, item, i, r[i], diag, true);
}

size_t apply_lds(const CloudArgs& a) {
  return sizeof(double) * (size_t(2 * a.Ntot) + size_t(2 * a.S * a.D * a.Nmax * a.ts) +
                           size_t(a.S * (a.D + 1) * a.chunk));
}

size_t rhs_lds(const CloudArgs& a) {
  return sizeof(double) * (size_t(a.Ntot) + size_t(2 * a.S * a.D * a.Nmax * a.ts) +
                           size_t(a.S * (a.D + 1) * a.chunk));
}

size_t l2_lds(const CloudArgs& a) {
  return sizeof(double) * (size_t(a.Ntot) + size_t(2 * a.D * a.Nmax * a.ts));
}

size_t diagonal_lds(const CloudArgs& a) {
  return sizeof(double) * (size_t(a.Ntot) + size_t(2 * a.S * a.D * a.Nmax * a.ts) +
                           size_t(a.S * (a.D + 1) * (a.D + 1) * a.chunk));
}

template <int KC>
size_t block_lds(const CloudArgs& a) {
  return sizeof(double) * (size_t(KC) * size_t(2 * a.Ntot + a.S * (a.D + 1) * a.chunk) +
                           size_t(2 * a.S * a.D * a.Nmax * a.ts));
}

// LDS budget per workgroup in KB: GMF_CLOUD_LDS_KB (8..64) when set, else the
// bucket's tuned budget.
size_t cloud_lds_budget(const CloudArgs& a) {
  static const int forced = [] {
    const char* e = std::getenv("GMF_CLOUD_LDS_KB");
    const int v = e ? std::atoi(e) : 0;
    return v >= 8 && v <= 64 ? v : 0;
  }();
  return size_t(forced ? forced : a.lds_kb) * 1024;
}

// Largest chunk (64, 32, .., 8 points) whose LDS footprint fits the budget
// (64 KB at most).
template <class F>
void pick_chunk(CloudArgs& a, F lds_of) {
  a.chunk = kCloudLanes;
  a.ts = a.chunk + 1;
  while (a.chunk > 8 && lds_of(a) > cloud_lds_budget(a)) {
    a.chunk /= 2;
    a.ts = a.chunk + 1;
  }
  while (a.chunk > 8 && lds_of(a) > 64 * 1024) {
    a.chunk /= 2;
    a.ts = a.chunk + 1;
  }
  require(lds_of(a) <= 64 * 1024, GMF_EUNSUPPORTED, "point-cloud batch exceeds the LDS budget");
}

CloudArgs make_cloud_args(const Bucket& b, bool reduced) {
  CloudArgs a{};
  a.D = b.dim;
  a.S = b.sides;
  a.Nmax = 1;
  a.Ntot = b.Ntot;
  for (int s = 0; s < kMaxSides; ++s) {
    a.Ns[s] = 1;
    for (int d = 0; d < kMaxDim; ++d) {
      a.N[s][d] = s == 0 ? b.N[d] : b.Np[d];
      if (d < b.dim) a.Ns[s] *= a.N[s][d];
      if (s < a.S && d < b.dim) a.Nmax = a.N[s][d] > a.Nmax ? a.N[s][d] : a.Nmax;
    }
  }
  require(a.S == 1 ? a.Ns[0] == a.Ntot : a.Ns[0] + a.Ns[1] == a.Ntot, GMF_EINVAL, "interface box sizes do not add up");
  a.jinv = b.jinv;
  a.pdata = b.pdata;
  a.poff = b.poff;
  a.rpts = b.rpts;
  a.map = TermMap{reduced ? b.ridx : b.lidx, b.lcoef, b.moff, reduced ? b.mridx : b.midx, b.mcoef, b.Ntot};
  a.unit_item = b.unit_item;
  a.unit_start = b.unit_start;
  a.unit_points = b.unit_points;
  a.lds_kb = b.lds_kb;
  return a;
}

// Columns per workgroup of the block kernel: 2 by default (measured on the
// finite-cell sphere: 2 columns 4.46 ms, 4 columns 4.73 ms, 8 columns 8.21 ms
// per column), 4 or 8 with GMF_CLOUD_COLUMNS.
int cloud_block_columns() {
  static const int cols = [] {
    const char* e = std::getenv("GMF_CLOUD_COLUMNS");
    const int v = e ? std::atoi(e) : 2;
    return v == 4 || v == 8 ? v : 2;
  }();
  return cols;
}

// One-column applies run the dimension-specialized kernel with one column
// (registers instead of scratch for the per-point arrays);
// GMF_CLOUD_KERNEL=item selects the dimension-generic cell_apply_cloud.
bool cloud_generic_kernel() {
  static const bool generic = [] {
    const char* e = std::getenv("GMF_CLOUD_KERNEL");
    return e && std::string(e) == "item";
  }();
  return generic;
}

template <int KC>
void launch_block(const Bucket& b, const double* X, double* Y, int64_t ld, int k, bool reduced, hipStream_t stream) {
  CloudArgs a = make_cloud_args(b, reduced);
  require(a.map.idx != nullptr, GMF_EUNSUPPORTED, "reduced index maps are not available for this plan");
  pick_chunk(a, block_lds<KC>);
  const dim3 grid(unsigned(b.units), unsigned((k + KC - 1) / KC));
  const size_t lds = block_lds<KC>(a);
  // One-sided 3D items with the same mode count on every axis take the
  // register forward (N = 2..5, degrees 1..4).
  const bool iso = a.S == 1 && a.D == 3 && a.N[0][0] == a.N[0][1] && a.N[0][0] == a.N[0][2];
  const int nf = iso && a.N[0][0] >= 2 && a.N[0][0] <= 5 ? a.N[0][0] : 0;
  switch (a.D * 8 + nf) {
    case 8: hipLaunchKernelGGL((cell_apply_cloud_block<KC, 1, 0>), grid, dim3(kCloudLanes), lds, stream, a, X, Y, ld, k); break;
    case 16: hipLaunchKernelGGL((cell_apply_cloud_block<KC, 2, 0>), grid, dim3(kCloudLanes), lds, stream, a, X, Y, ld, k); break;
    case 26: hipLaunchKernelGGL((cell_apply_cloud_block<KC, 3, 2>), grid, dim3(kCloudLanes), lds, stream, a, X, Y, ld, k); break;
    case 27: hipLaunchKernelGGL((cell_apply_cloud_block<KC, 3, 3>), grid, dim3(kCloudLanes), lds, stream, a, X, Y, ld, k); break;
    case 28: hipLaunchKernelGGL((cell_apply_cloud_block<KC, 3, 4>), grid, dim3(kCloudLanes), lds, stream, a, X, Y, ld, k); break;
    case 29: hipLaunchKernelGGL((cell_apply_cloud_block<KC, 3, 5>), grid, dim3(kCloudLanes), lds, stream, a, X, Y, ld, k); break;
    default: hipLaunchKernelGGL((cell_apply_cloud_block<KC, 3, 0>), grid, dim3(kCloudLanes), lds, stream, a, X, Y, ld, k);
  }
  GMF_HIP(hipGetLastError());
}

}  // namespace

// Local-matrix apply of KC-column groups.
template <int KC>
void launch_local_matrix(const Bucket& b, const double* X, double* Y, int64_t ld, int k, bool reduced,
                         hipStream_t stream) {
  const TermMap map{reduced ? b.ridx : b.lidx, b.lcoef, b.moff, reduced ? b.mridx : b.midx, b.mcoef, b.Ntot};
  require(map.idx != nullptr, GMF_EUNSUPPORTED, "reduced index maps are not available for this plan");
  require(b.emat != nullptr && b.Ntot <= kMaxLocalMatrixModes, GMF_EINVAL, "point-cloud batch has no local matrices");
  const dim3 grid(unsigned((b.cells + kMatrixItemsPerGroup - 1) / kMatrixItemsPerGroup), unsigned((k + KC - 1) / KC));
  const size_t lds = sizeof(double) * size_t(kMatrixItemsPerGroup) * KC * size_t(b.Ntot);
  hipLaunchKernelGGL((cell_apply_local_matrix<KC>), grid, dim3(64 * kMatrixItemsPerGroup), lds, stream, map, b.emat,
                     b.cells, X, Y, ld, k);
  GMF_HIP(hipGetLastError());
}

void apply_local_matrix(const Bucket& b, const double* X, double* Y, int64_t ld, int k, bool reduced,
                        hipStream_t stream) {
  if (k == 1)
    launch_local_matrix<1>(b, X, Y, ld, k, reduced, stream);
  else
    launch_local_matrix<8>(b, X, Y, ld, k, reduced, stream);
}

void launch_cell_cloud_block(const gmf_plan&, const Bucket& b, const double* X, double* Y, int64_t ld, int k,
                             bool reduced, hipStream_t stream) {
  if (b.cells == 0 || k <= 0) return;
  require(k <= 65535 * 2, GMF_EINVAL, "too many columns for one block launch");
  if (b.emat) {
    apply_local_matrix(b, X, Y, ld, k, reduced, stream);
    return;
  }
  switch (cloud_block_columns()) {
    case 4: launch_block<4>(b, X, Y, ld, k, reduced, stream); break;
    case 8: launch_block<8>(b, X, Y, ld, k, reduced, stream); break;
    default: launch_block<2>(b, X, Y, ld, k, reduced, stream);
  }
}

void launch_cell_cloud(const gmf_plan&, const Bucket& b, const double* x, double* y, bool reduced,
                       hipStream_t stream) {
  if (b.cells == 0) return;
  if (b.emat) {
    apply_local_matrix(b, x, y, 0, 1, reduced, stream);
    return;
  }
  CloudArgs a = make_cloud_args(b, reduced);
  require(a.map.idx != nullptr, GMF_EUNSUPPORTED, "reduced index maps are not available for this plan");
  if (!cloud_generic_kernel()) {
    launch_block<1>(b, x, y, 0, 1, reduced, stream);
    return;
  }
  pick_chunk(a, apply_lds);
  hipLaunchKernelGGL(cell_apply_cloud, dim3(unsigned(b.units)), dim3(kCloudLanes), apply_lds(a), stream, a, x, y);
  GMF_HIP(hipGetLastError());
}

int64_t cloud_units(const Bucket& b, int unit_points, std::vector<int32_t>& item, std::vector<int64_t>& start) {
  require(unit_points > 0 && int64_t(b.h_poff.size()) == b.cells + 1, GMF_EINVAL, "invalid point-cloud work units");
  item.clear();
  start.clear();
  const int64_t* po = b.h_poff.data();
  for (int64_t c = 0; c < b.cells; ++c)
    for (int64_t s0 = po[c]; s0 == po[c] || s0 < po[c + 1]; s0 += unit_points) {
      item.push_back(int32_t(c));
      start.push_back(s0);
    }
  return int64_t(item.size());
}

// Work-unit size (256 .. 2048 points) and LDS budget (64 or 32 KB, which
// halves the chunk of large boxes) of every point-cloud batch that runs the
// point-cloud kernel, timed on its one-column apply. Candidates that give the
// same work decomposition and chunk as one already timed are skipped;
// GMF_CLOUD_UNIT_POINTS restricts the work-unit size to one value.

// Local matrices of every item of a point-cloud batch (a new device array of
// cells x Ntot x Ntot doubles, column-major per item).
double* build_local_matrices(Bucket& b) {
  CloudArgs a = make_cloud_args(b, false);
  pick_chunk(a, apply_lds);
  void* e = nullptr;
  GMF_HIP(hipMalloc(&e, size_t(b.cells) * size_t(b.Ntot) * size_t(b.Ntot) * sizeof(double)));
  for (int64_t c0 = 0; c0 < b.cells; c0 += 65535) {
    // grid.x is limited: items in slices, the kernel indexes items from the slice start
    const int64_t nc = b.cells - c0 < 65535 ? b.cells - c0 : 65535;
    CloudArgs as = a;
    as.jinv = a.jinv + c0 * a.S * a.D;
    as.poff = a.poff + c0;
    hipLaunchKernelGGL(cell_local_matrix_cloud, dim3(unsigned(nc), unsigned(b.Ntot)), dim3(kCloudLanes),
                       apply_lds(a), nullptr, as, static_cast<double*>(e) + c0 * int64_t(b.Ntot) * b.Ntot);
    GMF_HIP(hipGetLastError());
  }
  GMF_HIP(hipDeviceSynchronize());
  return static_cast<double*>(e);
}

void tune_cloud(gmf_plan& p, TuneTimer& t) {
  int forced = 0;  // GMF_CLOUD_UNIT_POINTS: time only this work-unit size
  if (const char* e = std::getenv("GMF_CLOUD_UNIT_POINTS")) forced = std::atoi(e);
  int matrix_mode = -1;  // GMF_CLOUD_MATRIX: 0 never, 1 always (when they fit), unset: timed
  if (const char* e = std::getenv("GMF_CLOUD_MATRIX")) matrix_mode = std::atoi(e) ? 1 : 0;
  for (Bucket& b : p.buckets) {
    if (!b.cloud || b.cells == 0 || b.face.enabled) continue;
    TuneInfo& ti = b.tune;
    ti.kernel = kTuneCloud;
    const int64_t def_units = b.units;
    int32_t* def_item = b.unit_item;
    int64_t* def_start = b.unit_start;
    const int def_points = b.unit_points, def_kb = b.lds_kb;
    int best_points = def_points, best_kb = def_kb;
    float best = -1.0f;
    std::vector<int32_t> best_item;
    std::vector<int64_t> best_start;
    int64_t last_units = -1;
    for (int up : {256, 512, 1024, 2048}) {
      if (forced > 0 && up != forced) continue;
      std::vector<int32_t> item;
      std::vector<int64_t> start;
      const int64_t units = cloud_units(b, up, item, start);
      if (units == last_units) continue;  // no item is split differently
      last_units = units;
      DeviceBuffer<int32_t> di(item.size());
      DeviceBuffer<int64_t> ds(start.size());
      GMF_HIP(hipMemcpy(di.p, item.data(), item.size() * sizeof(int32_t), hipMemcpyHostToDevice));
      GMF_HIP(hipMemcpy(ds.p, start.data(), start.size() * sizeof(int64_t), hipMemcpyHostToDevice));
      b.units = units;
      b.unit_item = di.p;
      b.unit_start = ds.p;
      b.unit_points = up;
      int last_chunk = -1;
      for (int kb : {64, 32}) {
        b.lds_kb = kb;
        CloudArgs a = make_cloud_args(b, false);
        pick_chunk(a, block_lds<1>);
        if (a.chunk == last_chunk) continue;  // the budget changes nothing
        last_chunk = a.chunk;
        const float us = t.time_us([&] { launch_cell_cloud(p, b, t.x, t.y, false, nullptr); });
        ++ti.candidates;
        if (up == def_points && kb == def_kb) ti.default_us = us;
        if (best < 0.0f || us < best) {
          best = us;
          best_points = up;
          best_kb = kb;
          best_item = item;
          best_start = start;
        }
      }
    }
    b.units = def_units;
    b.unit_item = def_item;
    b.unit_start = def_start;
    b.unit_points = def_points;
    b.lds_kb = best_kb;
    if (best_points != def_points) {
      b.units = int64_t(best_item.size());
      b.unit_item = upload_array(p, best_item.data(), best_item.size());
      b.unit_start = upload_array(p, best_start.data(), best_start.size());
      b.unit_points = best_points;
    }
    ti.param0 = b.unit_points;
    ti.param1 = b.lds_kb;
    ti.best_us = best;
    // Local matrices: kept when their apply is faster than the best point
    // shape (GMF_CLOUD_MATRIX=0 never builds them, =1 keeps them whenever
    // they fit); at most kMaxLocalMatrixBytes per batch.
    const int64_t bytes = b.cells * int64_t(b.Ntot) * b.Ntot * int64_t(sizeof(double));
    if (matrix_mode == 0 || b.Ntot > kMaxLocalMatrixModes || bytes > kMaxLocalMatrixBytes) continue;
    double* e = build_local_matrices(b);
    b.emat = e;  // the timed apply reads the matrices through the bucket
    const float us = t.time_us([&] { apply_local_matrix(b, t.x, t.y, 0, 1, false, nullptr); });
    ++ti.candidates;
    if (matrix_mode == 1 || us < best) {
      p.allocations.push_back(e);
      p.device_bytes += bytes;
      ti.kernel = kTuneCloudMatrix;
      ti.param0 = b.Ntot;
      ti.param1 = 0;
      ti.best_us = us;
    } else {
      b.emat = nullptr;
      GMF_HIP(hipFree(e));
    }
  }
}

void launch_cloud_rhs(const Bucket& b, const double* loads, double* y, hipStream_t stream) {
  if (b.cells == 0) return;
  CloudArgs a = make_cloud_args(b, false);
  pick_chunk(a, rhs_lds);
  hipLaunchKernelGGL(cell_rhs_cloud, dim3(unsigned(b.units)), dim3(kCloudLanes), rhs_lds(a), stream, a, loads, y);
  GMF_HIP(hipGetLastError());
}

void launch_cloud_l2(const Bucket& b, const double* u_full, const double* w, const double* e, double* partial,
                     hipStream_t stream) {
  if (b.cells == 0) return;
  CloudArgs a = make_cloud_args(b, false);
  require(a.S == 1, GMF_EINVAL, "verification sums cover one-sided items");
  pick_chunk(a, l2_lds);
  hipLaunchKernelGGL(cell_l2_cloud, dim3(unsigned(b.units)), dim3(kCloudLanes), l2_lds(a), stream, a, u_full, w, e,
                     partial);
  GMF_HIP(hipGetLastError());
}

void launch_cell_cloud_diagonal(const gmf_plan&, const Bucket& b, double* diag, bool reduced, hipStream_t stream) {
  if (b.cells == 0) return;
  CloudArgs a = make_cloud_args(b, reduced);
  require(a.map.idx != nullptr, GMF_EUNSUPPORTED, "reduced index maps are not available for this plan");
  pick_chunk(a, diagonal_lds);
  hipLaunchKernelGGL(cell_diagonal_cloud, dim3(unsigned(b.cells)), dim3(kCloudLanes), diagonal_lds(a), stream, a,
                     diag);
  GMF_HIP(hipGetLastError());
}

}  // namespace gmf
