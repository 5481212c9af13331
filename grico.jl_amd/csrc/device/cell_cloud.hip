// Point-cloud cell kernels: cells or embedded-surface pieces that carry their
// own quadrature points (cut-cell rules, surface rules) instead of a tensor
// rule. The reference evaluates these through full basis tables
// (src/accumulators.jl:1022-1054 _accumulate_cell_accumulator_basis!,
// 1133-1161 _accumulate_one_sided_apply_basis!); here the integrated Legendre
// factors psi_k(xi_a), psi_k'(xi_a) are recomputed at each point on the device
// (src/polynomials.jl:403-434), so only the reference coordinates and the
// lowered per-point action travel from HBM.
//
// One 64-lane workgroup per item. Points are processed in chunks of 64: lanes
// over points evaluate the 1D factors into LDS, interpolate value and
// gradient, apply the (D+1) x (D+1) point action; then lanes over modes project
// the chunk's channels back.
#include <cmath>

#include "common.hpp"
#include "termmap.hpp"

namespace gmf {
namespace {

constexpr int kCloudLanes = 64;

struct CloudArgs {
  int D;
  int N[kMaxDim];
  int Ntot, Nmax;
  const double* jinv;   // items x D
  const double* pdata;  // points x (D+1)^2
  const int64_t* poff;  // items + 1
  const double* rpts;   // points x D
  TermMap map;
};

// psi_0..psi_{n-1} and derivatives at x into v[k * 64], d[k * 64] (stride 64).
__device__ inline void integrated_legendre(double x, int n, double* v, double* d) {
  if (n == 1) {
    v[0] = 1.0;
    d[0] = 0.0;
    return;
  }
  v[0] = 0.5 * (1.0 - x);
  d[0] = -0.5;
  v[kCloudLanes] = 0.5 * (1.0 + x);
  d[kCloudLanes] = 0.5;
  double pm2 = 1.0, pm1 = x;  // P_{k-2}, P_{k-1}
  for (int k = 2; k < n; ++k) {
    const double pk = ((2.0 * k - 1.0) * x * pm1 - (k - 1.0) * pm2) / double(k);
    v[k * kCloudLanes] = (pk - pm2) / sqrt(4.0 * k - 2.0);
    d[k * kCloudLanes] = sqrt((2.0 * k - 1.0) / 2.0) * pm1;
    pm2 = pm1;
    pm1 = pk;
  }
}

// Mode multi-index of flat box index i (axis 0 fastest).
__device__ inline void unflatten(const CloudArgs& a, int i, int (&m)[kMaxDim]) {
  m[0] = i % a.N[0];
  const int t = i / a.N[0];
  m[1] = a.D > 1 ? t % a.N[1] : 0;
  m[2] = a.D > 2 ? t / a.N[1] : 0;
}

// Phase 1 for one chunk: 1D tables of each point into tv / td
// ([axis][k][lane]), value and physical gradient of u, point action into
// ch[c][lane] (channel c = 0 value, 1 + a gradient along a). Returns nothing;
// lanes >= cnt write zeros.
__device__ void chunk_forward(const CloudArgs& a, int64_t item, int64_t p0, int cnt, const double* u, double* tv,
                              double* td, double* ch) {
  const int q = threadIdx.x;
  const int C = a.D + 1;
  const double* jinv = a.jinv + item * a.D;
  if (q < cnt) {
    for (int ax = 0; ax < a.D; ++ax)
      integrated_legendre(a.rpts[(p0 + q) * a.D + ax], a.N[ax], tv + ax * a.Nmax * kCloudLanes + q,
                          td + ax * a.Nmax * kCloudLanes + q);
    double val = 0.0, g[kMaxDim] = {0.0, 0.0, 0.0};
    for (int i = 0; i < a.Ntot; ++i) {
      int m[kMaxDim];
      unflatten(a, i, m);
      double f[kMaxDim], fd[kMaxDim];
      for (int ax = 0; ax < a.D; ++ax) {
        f[ax] = tv[(ax * a.Nmax + m[ax]) * kCloudLanes + q];
        fd[ax] = td[(ax * a.Nmax + m[ax]) * kCloudLanes + q];
      }
      double phi = 1.0;
      for (int ax = 0; ax < a.D; ++ax) phi *= f[ax];
      const double ui = u[i];
      val = fma(ui, phi, val);
      for (int ax = 0; ax < a.D; ++ax) {
        double dphi = fd[ax];
        for (int bx = 0; bx < a.D; ++bx)
          if (bx != ax) dphi *= f[bx];
        g[ax] = fma(ui, dphi, g[ax]);
      }
    }
    double chn[kMaxDim + 1];
    chn[0] = val;
    for (int ax = 0; ax < a.D; ++ax) chn[ax + 1] = jinv[ax] * g[ax];
    const double* M = a.pdata + (p0 + q) * C * C;
    for (int c = 0; c < C; ++c) {
      double acc = 0.0;
      for (int k = 0; k < C; ++k) acc = fma(M[c * C + k], chn[k], acc);
      ch[c * kCloudLanes + q] = c == 0 ? acc : acc * jinv[c - 1];
    }
  } else {
    for (int c = 0; c < C; ++c) ch[c * kCloudLanes + q] = 0.0;
  }
}

__global__ void __launch_bounds__(kCloudLanes) cell_apply_cloud(CloudArgs a, const double* __restrict__ x,
                                                                double* __restrict__ y) {
  extern __shared__ double smem[];
  const int64_t item = blockIdx.x;
  double* u = smem;
  double* r = u + a.Ntot;
  double* tv = r + a.Ntot;
  double* td = tv + a.D * a.Nmax * kCloudLanes;
  double* ch = td + a.D * a.Nmax * kCloudLanes;
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
    u[i] = gather_term(a.map, item, i, x);
    r[i] = 0.0;
  }
  __syncthreads();
  const int64_t pb = a.poff[item], pe = a.poff[item + 1];
  for (int64_t p0 = pb; p0 < pe; p0 += kCloudLanes) {
    const int cnt = int(pe - p0 < kCloudLanes ? pe - p0 : kCloudLanes);
    chunk_forward(a, item, p0, cnt, u, tv, td, ch);
    __syncthreads();
    for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
      int m[kMaxDim];
      unflatten(a, i, m);
      double acc = 0.0;
      for (int q = 0; q < cnt; ++q) {
        double f[kMaxDim], fd[kMaxDim];
        for (int ax = 0; ax < a.D; ++ax) {
          f[ax] = tv[(ax * a.Nmax + m[ax]) * kCloudLanes + q];
          fd[ax] = td[(ax * a.Nmax + m[ax]) * kCloudLanes + q];
        }
        double phi = 1.0;
        for (int ax = 0; ax < a.D; ++ax) phi *= f[ax];
        acc = fma(ch[q], phi, acc);
        for (int ax = 0; ax < a.D; ++ax) {
          double dphi = fd[ax];
          for (int bx = 0; bx < a.D; ++bx)
            if (bx != ax) dphi *= f[bx];
          acc = fma(ch[(ax + 1) * kCloudLanes + q], dphi, acc);
        }
      }
      r[i] += acc;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) scatter_term(a.map, item, i, r[i], y, false);
}

// Local diagonal d_i = sum_q sum_{c,k} M_q[c][k] chi_c chi_k with chi_0 = phi_i,
// chi_{1+a} = J^-1_a d_a phi_i, scattered with squared coefficients.
__global__ void __launch_bounds__(kCloudLanes) cell_diagonal_cloud(CloudArgs a, double* __restrict__ diag) {
  extern __shared__ double smem[];
  const int64_t item = blockIdx.x;
  const int C = a.D + 1;
  double* r = smem;
  double* tv = r + a.Ntot;
  double* td = tv + a.D * a.Nmax * kCloudLanes;
  double* Ms = td + a.D * a.Nmax * kCloudLanes;  // C*C x 64
  const double* jinv = a.jinv + item * a.D;
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) r[i] = 0.0;
  __syncthreads();
  const int64_t pb = a.poff[item], pe = a.poff[item + 1];
  for (int64_t p0 = pb; p0 < pe; p0 += kCloudLanes) {
    const int cnt = int(pe - p0 < kCloudLanes ? pe - p0 : kCloudLanes);
    const int q = threadIdx.x;
    if (q < cnt) {
      for (int ax = 0; ax < a.D; ++ax)
        integrated_legendre(a.rpts[(p0 + q) * a.D + ax], a.N[ax], tv + ax * a.Nmax * kCloudLanes + q,
                            td + ax * a.Nmax * kCloudLanes + q);
      for (int e = 0; e < C * C; ++e) Ms[e * kCloudLanes + q] = a.pdata[(p0 + q) * C * C + e];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) {
      int m[kMaxDim];
      unflatten(a, i, m);
      double acc = 0.0;
      for (int qq = 0; qq < cnt; ++qq) {
        double f[kMaxDim], fd[kMaxDim], chi[kMaxDim + 1];
        for (int ax = 0; ax < a.D; ++ax) {
          f[ax] = tv[(ax * a.Nmax + m[ax]) * kCloudLanes + qq];
          fd[ax] = td[(ax * a.Nmax + m[ax]) * kCloudLanes + qq];
        }
        double phi = 1.0;
        for (int ax = 0; ax < a.D; ++ax) phi *= f[ax];
        chi[0] = phi;
        for (int ax = 0; ax < a.D; ++ax) {
          double dphi = fd[ax];
          for (int bx = 0; bx < a.D; ++bx)
            if (bx != ax) dphi *= f[bx];
          chi[ax + 1] = jinv[ax] * dphi;
        }
        for (int c = 0; c < C; ++c)
          for (int k = 0; k < C; ++k) acc = fma(Ms[(c * C + k) * kCloudLanes + qq] * chi[c], chi[k], acc);
      }
      r[i] += acc;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < a.Ntot; i += blockDim.x) scatter_term(a.map, item, i, r[i], diag, true);
}

CloudArgs make_cloud_args(const Bucket& b, bool reduced) {
  CloudArgs a{};
  a.D = b.dim;
  a.Nmax = 1;
  for (int d = 0; d < kMaxDim; ++d) a.N[d] = b.N[d];
  for (int d = 0; d < b.dim; ++d) a.Nmax = b.N[d] > a.Nmax ? b.N[d] : a.Nmax;
  a.Ntot = b.Ntot;
  a.jinv = b.jinv;
  a.pdata = b.pdata;
  a.poff = b.poff;
  a.rpts = b.rpts;
  a.map = TermMap{reduced ? b.ridx : b.lidx, b.lcoef, b.moff, reduced ? b.mridx : b.midx, b.mcoef, b.Ntot};
  return a;
}

}  // namespace

void launch_cell_cloud(const gmf_plan&, const Bucket& b, const double* x, double* y, bool reduced,
                       hipStream_t stream) {
  if (b.cells == 0) return;
  CloudArgs a = make_cloud_args(b, reduced);
  require(a.map.idx != nullptr, GMF_EUNSUPPORTED, "reduced index maps are not available for this plan");
  const size_t lds = sizeof(double) * size_t(2 * a.Ntot + 2 * a.D * a.Nmax * kCloudLanes + (a.D + 1) * kCloudLanes);
  require(lds <= 64 * 1024, GMF_EUNSUPPORTED, "point-cloud batch exceeds the LDS budget");
  hipLaunchKernelGGL(cell_apply_cloud, dim3(unsigned(b.cells)), dim3(kCloudLanes), lds, stream, a, x, y);
  GMF_HIP(hipGetLastError());
}

void launch_cell_cloud_diagonal(const gmf_plan&, const Bucket& b, double* diag, bool reduced, hipStream_t stream) {
  if (b.cells == 0) return;
  CloudArgs a = make_cloud_args(b, reduced);
  require(a.map.idx != nullptr, GMF_EUNSUPPORTED, "reduced index maps are not available for this plan");
  const int C = a.D + 1;
  const size_t lds = sizeof(double) * size_t(a.Ntot + 2 * a.D * a.Nmax * kCloudLanes + C * C * kCloudLanes);
  require(lds <= 64 * 1024, GMF_EUNSUPPORTED, "point-cloud batch exceeds the LDS budget");
  hipLaunchKernelGGL(cell_diagonal_cloud, dim3(unsigned(b.cells)), dim3(kCloudLanes), lds, stream, a, diag);
  GMF_HIP(hipGetLastError());
}

}  // namespace gmf
