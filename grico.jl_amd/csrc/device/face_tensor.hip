// Sum-factorized kernel for boundary faces and interfaces whose points form
// one tangential tensor grid per side.
//
// A boundary face or an interface carries the Gauss points of its tangential
// rule on the shared face patch at the normal coordinate +-1 of each side
// (src/integration.jl:2230-2260 _compile_boundary_face, 2279-2316
// _compile_interface, 2800-2841 _interface_face_points); on a hanging
// interface the larger side sees them mapped into its sub-face, so batches of
// equal sub-face position share one grid per side. Since psi_k(+-1) = 0 for k >= 2 (integrated Legendre,
// src/polynomials.jl:403-434), the trace of a side is a contraction of its
// mode box along the normal axis with psi(+-1) (value) and psi'(+-1) (normal
// derivative), followed by tensor interpolation over the tangential axes; the
// test side runs the same steps transposed. The reference evaluates these
// items through full basis tables (src/accumulators.jl:1133-1161, 1576-1616);
// here only the face-local coefficient planes and the tangential tables live in
// LDS.
//
// One 64-lane workgroup per item: gather -> normal contraction (lanes over
// side x field x face mode) -> tangential interpolation (lanes over side x
// point) -> compressed point action (lanes over points) -> transposed
// tangential pass -> transposed normal pass and scatter (lanes over modes).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "termmap.hpp"

namespace gmf {
namespace {

constexpr int kFaceLanes = 64;

struct FaceArgs {
  int D, S, C;
  int N[2][kMaxDim];  // modes per axis of each side (absent axes 1)
  int Ns[2];          // box size of each side
  int Ntot;
  int axis;           // normal axis
  int T[2];           // tangential axes (-1: absent)
  int Qt[2];          // points per tangential slot
  int Q;              // points per item
  int Nt[2][2];       // per side modes along the tangential slots (1 if absent)
  int NtMax;          // largest face-plane size over the sides
  int toff[2][6];
  int table_doubles;
  int nnz;
  unsigned char pr[64], pc[64];
  const double* tables;
  const double* pdata;  // items x Q x nnz
  const double* jinv;   // items x S*D
  TermMap map;
};

// Side, normal index k and tangential indices of local mode i.
__device__ inline int face_mode(const FaceArgs& a, int i, int& k, int& t1, int& t2) {
  const int s = (a.S > 1 && i >= a.Ns[0]) ? 1 : 0;
  int il = s ? i - a.Ns[0] : i;
  int m[kMaxDim] = {0, 0, 0};
  for (int d = 0; d < a.D; ++d) {
    m[d] = il % a.N[s][d];
    il /= a.N[s][d];
  }
  k = m[a.axis];
  t1 = a.T[0] >= 0 ? m[a.T[0]] : 0;
  t2 = a.T[1] >= 0 ? m[a.T[1]] : 0;
  return s;
}

// Box index on side s of (normal index k, tangential indices t1, t2).
__device__ inline int face_box(const FaceArgs& a, int s, int k, int t1, int t2) {
  int m[kMaxDim] = {0, 0, 0};
  m[a.axis] = k;
  if (a.T[0] >= 0) m[a.T[0]] = t1;
  if (a.T[1] >= 0) m[a.T[1]] = t2;
  return m[0] + a.N[s][0] * (m[1] + a.N[s][1] * m[2]);
}

__device__ void face_project(const FaceArgs& a, int64_t item, const double* tab, const double* ch, double* R,
                             double* __restrict__ y);

// One item: the whole pipeline on one 64-lane workgroup.
__device__ void face_item(const FaceArgs& a, int64_t item, double* sm, const double* __restrict__ x,
                          double* __restrict__ y) {
  const int tid = threadIdx.x;
  const int C = a.C, SC = a.S * a.C, Q = a.Q, NtM = a.NtMax;
  double* tab = sm;
  double* u = tab + a.table_doubles;  // gathered coefficients
  double* U = u + a.Ntot;             // [S][2][NtMax]: psi(+-1) / psi'(+-1) contractions, later their transposes
  double* ch = U + a.S * 2 * NtM;     // [SC][Q] point channels
  double* R = U;
  for (int i = tid; i < a.table_doubles; i += kFaceLanes) tab[i] = a.tables[i];
  for (int i = tid; i < a.Ntot; i += kFaceLanes) u[i] = gather_term(a.map, item, i, x);
  __syncthreads();

  // Normal contraction: U[s][f][t] = sum_k u_s[k, t] v_f[k].
  for (int t = tid; t < a.S * 2 * NtM; t += kFaceLanes) {
    const int s = t / (2 * NtM), f = (t / NtM) % 2, mt = t % NtM;
    if (mt >= a.Nt[s][0] * a.Nt[s][1]) continue;
    const int t1 = mt % a.Nt[s][0], t2 = mt / a.Nt[s][0];
    const double* v = tab + a.toff[s][f];
    const double* us = u + (s ? a.Ns[0] : 0);
    double acc = 0.0;
    for (int k = 0; k < a.N[s][a.axis]; ++k) acc = fma(us[face_box(a, s, k, t1, t2)], v[k], acc);
    U[(s * 2 + f) * NtM + mt] = acc;
  }
  __syncthreads();

  // Tangential interpolation: value, tangential and normal derivatives per point.
  const double* jinv = a.jinv + item * a.S * a.D;
  for (int t = tid; t < a.S * Q; t += kFaceLanes) {
    const int s = t / Q, q = t % Q, q1 = q % a.Qt[0], q2 = q / a.Qt[0];
    const double* tv0 = tab + a.toff[s][2];
    const double* td0 = tab + a.toff[s][3];
    const double* tv1 = tab + a.toff[s][4];
    const double* td1 = tab + a.toff[s][5];
    const double* U0 = U + (s * 2) * NtM;
    const double* U1 = U0 + NtM;
    double v = 0.0, g0 = 0.0, g1 = 0.0, gn = 0.0;
    for (int t2 = 0; t2 < a.Nt[s][1]; ++t2) {
      const double b2 = tv1[t2 * a.Qt[1] + q2], d2 = td1[t2 * a.Qt[1] + q2];
      double sv = 0.0, sd = 0.0, sn = 0.0;
      for (int t1 = 0; t1 < a.Nt[s][0]; ++t1) {
        const double b1 = tv0[t1 * a.Qt[0] + q1], d1 = td0[t1 * a.Qt[0] + q1];
        const double c0 = U0[t1 + a.Nt[s][0] * t2], c1 = U1[t1 + a.Nt[s][0] * t2];
        sv = fma(b1, c0, sv);
        sd = fma(d1, c0, sd);
        sn = fma(b1, c1, sn);
      }
      v = fma(b2, sv, v);
      g0 = fma(b2, sd, g0);
      g1 = fma(d2, sv, g1);
      gn = fma(b2, sn, gn);
    }
    const double* js = jinv + s * a.D;
    ch[(s * C) * Q + q] = v;
    ch[(s * C + 1 + a.axis) * Q + q] = js[a.axis] * gn;
    if (a.T[0] >= 0) ch[(s * C + 1 + a.T[0]) * Q + q] = js[a.T[0]] * g0;
    if (a.T[1] >= 0) ch[(s * C + 1 + a.T[1]) * Q + q] = js[a.T[1]] * g1;
  }
  __syncthreads();

  // Point action on the compressed pattern (rows selected without dynamic
  // register indexing); gradient test channels scaled by J^-1.
  constexpr int kMaxChannels = 2 * (kMaxDim + 1);
  for (int q = tid; q < Q; q += kFaceLanes) {
    double out[kMaxChannels];
#pragma unroll
    for (int c = 0; c < kMaxChannels; ++c) out[c] = 0.0;
    const double* M = a.pdata + (item * Q + q) * a.nnz;
    for (int e = 0; e < a.nnz; ++e) {
      const double v = M[e] * ch[a.pc[e] * Q + q];
      const int row = a.pr[e];
#pragma unroll
      for (int c = 0; c < kMaxChannels; ++c) out[c] += row == c ? v : 0.0;
    }
#pragma unroll
    for (int c = 0; c < kMaxChannels; ++c) {
      if (c >= SC) break;
      const int s = c / C, cc = c % C;
      ch[c * Q + q] = cc == 0 ? out[c] : out[c] * jinv[s * a.D + cc - 1];
    }
  }
  __syncthreads();

  face_project(a, item, tab, ch, R, y);
}

// Transposed passes of one item: the test channels ch[SC][Q] (gradient
// channels already scaled by J^-1) projected onto both sides' modes and
// scattered into y.
__device__ void face_project(const FaceArgs& a, int64_t item, const double* tab, const double* ch, double* R,
                             double* __restrict__ y) {
  const int tid = threadIdx.x;
  const int C = a.C, Q = a.Q, NtM = a.NtMax;
  // Transposed tangential pass: R[s][0] collects value and tangential
  // derivative channels, R[s][1] the normal derivative channel.
  for (int t = tid; t < a.S * 2 * NtM; t += kFaceLanes) {
    const int s = t / (2 * NtM), f = (t / NtM) % 2, mt = t % NtM;
    if (mt >= a.Nt[s][0] * a.Nt[s][1]) continue;
    const int t1 = mt % a.Nt[s][0], t2 = mt / a.Nt[s][0];
    const double* tv0 = tab + a.toff[s][2];
    const double* td0 = tab + a.toff[s][3];
    const double* tv1 = tab + a.toff[s][4];
    const double* td1 = tab + a.toff[s][5];
    const double* cv = ch + (s * C) * Q;
    const double* cn = ch + (s * C + 1 + a.axis) * Q;
    const double* c0 = a.T[0] >= 0 ? ch + (s * C + 1 + a.T[0]) * Q : nullptr;
    const double* c1 = a.T[1] >= 0 ? ch + (s * C + 1 + a.T[1]) * Q : nullptr;
    double acc = 0.0;
    for (int q2 = 0; q2 < a.Qt[1]; ++q2) {
      const double b2 = tv1[t2 * a.Qt[1] + q2], d2 = td1[t2 * a.Qt[1] + q2];
      for (int q1 = 0; q1 < a.Qt[0]; ++q1) {
        const int q = q1 + a.Qt[0] * q2;
        const double b1 = tv0[t1 * a.Qt[0] + q1];
        if (f == 1) {
          acc = fma(b1 * b2, cn[q], acc);
        } else {
          double w = b1 * b2 * cv[q];
          if (c0) w = fma(td0[t1 * a.Qt[0] + q1] * b2, c0[q], w);
          if (c1) w = fma(b1 * d2, c1[q], w);
          acc += w;
        }
      }
    }
    R[(s * 2 + f) * NtM + mt] = acc;
  }
  __syncthreads();

  // Transposed normal pass and scatter.
  for (int i = tid; i < a.Ntot; i += kFaceLanes) {
    int k, t1, t2;
    const int s = face_mode(a, i, k, t1, t2);
    const int mt = t1 + a.Nt[s][0] * t2;
    const double res = fma(tab[a.toff[s][0] + k], R[(s * 2) * NtM + mt], tab[a.toff[s][1] + k] * R[(s * 2 + 1) * NtM + mt]);
    scatter_term(a.map, item, i, res, y, false);
  }
}

__global__ void __launch_bounds__(kFaceLanes) face_apply_tensor(FaceArgs a, const double* __restrict__ x,
                                                                double* __restrict__ y) {
  extern __shared__ double sm[];
  face_item(a, blockIdx.x, sm, x, y);
}

// Every face-tensor batch of a plan in one launch: workgroup w belongs to
// batch k with start[k] <= w < start[k + 1].
__global__ void __launch_bounds__(kFaceLanes) face_apply_tensor_batched(const FaceArgs* __restrict__ args,
                                                                        const int64_t* __restrict__ start, int count,
                                                                        const double* __restrict__ x,
                                                                        double* __restrict__ y) {
  extern __shared__ double sm[];
  const int64_t blk = blockIdx.x;
  int k = 0;
  while (k + 1 < count && blk >= start[k + 1]) ++k;
  face_item(args[k], blk - start[k], sm, x, y);
}

// Face loads of one batch (src/accumulators.jl:1230-1274 one-sided loads,
// 1666-1698 interface loads): per item, point and side the channels (W v,
// W g_1, .., W g_D), projected like the test channels of the operator.
__global__ void __launch_bounds__(kFaceLanes) face_rhs_tensor(FaceArgs a, const double* __restrict__ loads,
                                                              double* __restrict__ y) {
  extern __shared__ double sm[];
  const int64_t item = blockIdx.x;
  const int tid = threadIdx.x;
  const int C = a.C, SC = a.S * a.C, Q = a.Q;
  double* tab = sm;
  double* U = tab + a.table_doubles + a.Ntot;  // same layout as face_item
  double* ch = U + a.S * 2 * a.NtMax;
  for (int i = tid; i < a.table_doubles; i += kFaceLanes) tab[i] = a.tables[i];
  const double* jinv = a.jinv + item * a.S * a.D;
  const double* L = loads + item * int64_t(Q) * SC;
  for (int t = tid; t < SC * Q; t += kFaceLanes) {
    const int q = t / SC, r = t % SC, s = r / C, c = r % C;
    const double v = L[t];
    ch[r * Q + q] = c == 0 ? v : v * jinv[s * a.D + c - 1];
  }
  __syncthreads();
  face_project(a, item, tab, ch, U, y);
}

// psi_0..psi_{n-1} and derivatives at x (src/polynomials.jl:403-434; p = 0 -> 1).
void psi_host(double x, int n, double* v, double* d) {
  if (n == 1) {
    v[0] = 1.0;
    d[0] = 0.0;
    return;
  }
  v[0] = 0.5 * (1.0 - x);
  d[0] = -0.5;
  v[1] = 0.5 * (1.0 + x);
  d[1] = 0.5;
  double pm2 = 1.0, pm1 = x;
  for (int k = 2; k < n; ++k) {
    const double pk = ((2.0 * k - 1.0) * x * pm1 - (k - 1.0) * pm2) / double(k);
    v[k] = (pk - pm2) / std::sqrt(4.0 * k - 2.0);
    d[k] = std::sqrt((2.0 * k - 1.0) / 2.0) * pm1;
    pm2 = pm1;
    pm1 = pk;
  }
}

size_t face_lds(const FaceArgs& a) {
  return sizeof(double) * (size_t(a.table_doubles) + size_t(a.Ntot) + size_t(2 * a.S * a.NtMax) +
                           size_t(a.S * a.C * a.Q));
}

// Batches below this many items share one launch; larger ones fill the chip alone.
constexpr int64_t kFaceBatchMaxItems = 16384;

FaceArgs make_face_args(const Bucket& b, bool reduced) {
  const FacePath& f = b.face;
  FaceArgs a{};
  a.D = b.dim;
  a.S = b.sides;
  a.C = b.dim + 1;
  a.Ntot = b.Ntot;
  a.axis = f.axis;
  a.NtMax = 1;
  for (int s = 0; s < 2; ++s) {
    a.Ns[s] = 1;
    for (int d = 0; d < kMaxDim; ++d) {
      a.N[s][d] = d < b.dim ? (s == 0 ? b.N[d] : b.Np[d]) : 1;
      a.Ns[s] *= a.N[s][d];
    }
    for (int j = 0; j < 2; ++j) a.Nt[s][j] = f.T[j] >= 0 ? a.N[s][f.T[j]] : 1;
    if (s < a.S) a.NtMax = std::max(a.NtMax, a.Nt[s][0] * a.Nt[s][1]);
    for (int k = 0; k < 6; ++k) a.toff[s][k] = f.toff[s][k];
  }
  for (int j = 0; j < 2; ++j) {
    a.T[j] = f.T[j];
    a.Qt[j] = f.Qt[j];
  }
  a.Q = f.Qt[0] * f.Qt[1];
  a.table_doubles = f.table_doubles;
  a.nnz = f.nnz;
  std::memcpy(a.pr, f.pr, sizeof(a.pr));
  std::memcpy(a.pc, f.pc, sizeof(a.pc));
  a.tables = f.tables;
  a.pdata = f.pdata;
  a.jinv = b.jinv;
  a.map = TermMap{reduced ? b.ridx : b.lidx, b.lcoef, b.moff, reduced ? b.mridx : b.midx, b.mcoef, b.Ntot};
  return a;
}

}  // namespace

void prepare_face_path(gmf_plan& p, Bucket& b, const gmf_bucket_desc& bd) {
  const char* env = std::getenv("GMF_FACE_TENSOR");
  if (env && std::strcmp(env, "0") == 0) return;
  if (!b.cloud || b.cells == 0) return;
  const int D = b.dim, S = b.sides, SD = S * D, C = D + 1, SC = S * C;
  const int64_t* po = bd.point_offsets;
  const int64_t n = po[1] - po[0];
  if (n < 1 || n > 256) return;
  for (int64_t c = 0; c < b.cells; ++c)
    if (po[c + 1] - po[c] != n) return;
  const double* rp = bd.ref_points;
  const int64_t P = b.total_points;
  // Normal axis: a coordinate that is exactly +-1 with one sign per side.
  int axis = -1;
  double sign[2] = {0.0, 0.0};
  for (int a = 0; a < D && axis < 0; ++a) {
    bool ok = true;
    for (int s = 0; s < S && ok; ++s) {
      const double s0 = rp[s * D + a];
      ok = std::fabs(s0) == 1.0;
      for (int64_t q = 0; q < P && ok; ++q) ok = rp[q * SD + s * D + a] == s0;
      sign[s] = s0;
    }
    if (ok) axis = a;
  }
  if (axis < 0) return;
  FacePath f;
  f.axis = axis;
  for (int a = 0, j = 0; a < D; ++a)
    if (a != axis) f.T[j++] = a;
  // Tangential tensor grid (first tangential axis fastest): each side's
  // coordinates are the same for every item up to rounding of the mapped
  // coordinates (conforming faces: both sides the Gauss points; hanging
  // interfaces batched by sub-face position: the larger side's are mapped).
  auto coord = [&](int64_t q, int s, int a) { return rp[q * SD + s * D + a]; };
  constexpr double tol = 1e-13;
  int Q0 = 1;
  if (f.T[0] >= 0) {
    Q0 = int(n);
    for (int64_t k = 1; k < n; ++k)
      if (std::fabs(coord(k, 0, f.T[0]) - coord(0, 0, f.T[0])) <= tol) {
        Q0 = int(k);
        break;
      }
  }
  if (n % Q0 != 0) return;
  const int Q1 = int(n / Q0);
  if (f.T[1] < 0 && Q1 != 1) return;
  if (Q0 > kMaxExtent || Q1 > kMaxExtent) return;
  std::vector<double> t0s[2], t1s[2];
  for (int s = 0; s < S; ++s) {
    t0s[s].assign(static_cast<size_t>(Q0), 0.0);
    t1s[s].assign(static_cast<size_t>(Q1), 0.0);
    for (int k = 0; k < Q0; ++k) t0s[s][size_t(k)] = f.T[0] >= 0 ? coord(k, s, f.T[0]) : 0.0;
    for (int k = 0; k < Q1; ++k) t1s[s][size_t(k)] = f.T[1] >= 0 ? coord(int64_t(k) * Q0, s, f.T[1]) : 0.0;
  }
  for (int64_t q = 0; q < P; ++q) {
    const int64_t k = q % n;
    for (int s = 0; s < S; ++s) {
      if (f.T[0] >= 0 && std::fabs(coord(q, s, f.T[0]) - t0s[s][size_t(k % Q0)]) > tol) return;
      if (f.T[1] >= 0 && std::fabs(coord(q, s, f.T[1]) - t1s[s][size_t(k / Q0)]) > tol) return;
    }
  }
  f.Qt[0] = Q0;
  f.Qt[1] = Q1;
  // Compressed action pattern.
  const int64_t stride = int64_t(SC) * SC;
  std::vector<char> mask(static_cast<size_t>(stride), 0);
  for (int64_t q = 0; q < P; ++q)
    for (int64_t e = 0; e < stride; ++e)
      if (bd.point_data[q * stride + e] != 0.0) mask[size_t(e)] = 1;
  std::vector<int> ent;
  for (int e = 0; e < stride; ++e)
    if (mask[size_t(e)]) ent.push_back(e);
  if (ent.empty()) ent.push_back(0);
  f.nnz = int(ent.size());
  for (int e = 0; e < f.nnz; ++e) {
    f.pr[e] = static_cast<unsigned char>(ent[size_t(e)] / SC);
    f.pc[e] = static_cast<unsigned char>(ent[size_t(e)] % SC);
  }
  std::vector<double> packed(static_cast<size_t>(P) * static_cast<size_t>(f.nnz));
  for (int64_t q = 0; q < P; ++q)
    for (int e = 0; e < f.nnz; ++e) packed[size_t(q * f.nnz + e)] = bd.point_data[q * stride + ent[size_t(e)]];
  // Tables per side: psi(+-1), psi'(+-1) along the normal, then the values and
  // derivatives of each tangential slot at its points ([mode][point]).
  std::vector<double> tab;
  for (int s = 0; s < S; ++s) {
    const int* Ns = s == 0 ? b.N : b.Np;
    const int na = Ns[axis];
    std::vector<double> v(static_cast<size_t>(na)), d(static_cast<size_t>(na));
    psi_host(sign[s], na, v.data(), d.data());
    f.toff[s][0] = int(tab.size());
    tab.insert(tab.end(), v.begin(), v.end());
    f.toff[s][1] = int(tab.size());
    tab.insert(tab.end(), d.begin(), d.end());
    for (int j = 0; j < 2; ++j) {
      const int nt = f.T[j] >= 0 ? Ns[f.T[j]] : 1;
      const std::vector<double>& pts = j == 0 ? t0s[s] : t1s[s];
      const int qn = f.Qt[j];
      std::vector<double> tv(static_cast<size_t>(nt * qn)), td(static_cast<size_t>(nt * qn));
      std::vector<double> pv(static_cast<size_t>(nt)), pd(static_cast<size_t>(nt));
      for (int q = 0; q < qn; ++q) {
        if (f.T[j] >= 0) {
          psi_host(pts[size_t(q)], nt, pv.data(), pd.data());
        } else {
          pv[0] = 1.0;
          pd[0] = 0.0;
        }
        for (int k = 0; k < nt; ++k) {
          tv[size_t(k * qn + q)] = pv[size_t(k)];
          td[size_t(k * qn + q)] = pd[size_t(k)];
        }
      }
      f.toff[s][2 + 2 * j] = int(tab.size());
      tab.insert(tab.end(), tv.begin(), tv.end());
      f.toff[s][3 + 2 * j] = int(tab.size());
      tab.insert(tab.end(), td.begin(), td.end());
    }
  }
  f.table_doubles = int(tab.size());
  f.tables = upload_array(p, tab.data(), tab.size());
  f.pdata = upload_array(p, packed.data(), packed.size());
  Bucket probe = b;
  probe.face = f;
  f.lds = face_lds(make_face_args(probe, false));
  if (f.lds > 64 * 1024) return;
  f.enabled = true;
  b.face = f;
}

void prepare_face_batch(gmf_plan& p) {
  std::vector<int> ids;
  for (size_t i = 0; i < p.buckets.size(); ++i)
    if (p.buckets[i].face.enabled && p.buckets[i].cells > 0 && p.buckets[i].cells < kFaceBatchMaxItems)
      ids.push_back(int(i));
  p.face_batch = GenericBatch{};
  if (ids.size() < 2) return;
  std::vector<FaceArgs> full, red;
  std::vector<int64_t> start(1, 0);
  bool have_red = true;
  size_t lds = 0;
  for (int i : ids) {
    const Bucket& b = p.buckets[size_t(i)];
    full.push_back(make_face_args(b, false));
    red.push_back(make_face_args(b, true));
    if (red.back().map.idx == nullptr) have_red = false;
    start.push_back(start.back() + b.cells);
    lds = std::max(lds, b.face.lds);
  }
  if (start.back() >= (int64_t(1) << 31)) return;
  GenericBatch g;
  g.threads = kFaceLanes;
  g.count = int(ids.size());
  g.buckets = ids;
  g.cells = start.back();
  g.lds = lds;
  g.args_full = upload_array(p, full.data(), full.size());
  if (have_red) g.args_red = upload_array(p, red.data(), red.size());
  g.start = upload_array(p, start.data(), start.size());
  p.face_batch = std::move(g);
}

void launch_face_batch(const gmf_plan& p, const double* x, double* y, bool reduced, hipStream_t stream,
                       char* done) {
  const GenericBatch& g = p.face_batch;
  const void* args = reduced ? g.args_red : g.args_full;
  if (g.count == 0 || args == nullptr) return;
  hipLaunchKernelGGL(face_apply_tensor_batched, dim3(unsigned(g.cells)), dim3(kFaceLanes), g.lds, stream,
                     static_cast<const FaceArgs*>(args), g.start, g.count, x, y);
  GMF_HIP(hipGetLastError());
  for (int i : g.buckets) done[size_t(i)] = 1;
}

void launch_face_rhs(const Bucket& b, const double* loads, double* y, hipStream_t stream) {
  require(b.face.enabled, GMF_EUNSUPPORTED, "device face loads cover batches on the sum-factorized face path");
  if (b.cells == 0) return;
  FaceArgs a = make_face_args(b, false);
  hipLaunchKernelGGL(face_rhs_tensor, dim3(unsigned(b.cells)), dim3(kFaceLanes), b.face.lds, stream, a, loads, y);
  GMF_HIP(hipGetLastError());
}

bool launch_face_tensor(const gmf_plan&, const Bucket& b, const double* x, double* y, bool reduced,
                        hipStream_t stream) {
  if (!b.face.enabled || b.cells == 0) return false;
  FaceArgs a = make_face_args(b, reduced);
  require(a.map.idx != nullptr, GMF_EUNSUPPORTED, "reduced index maps are not available for this plan");
  hipLaunchKernelGGL(face_apply_tensor, dim3(unsigned(b.cells)), dim3(kFaceLanes), b.face.lds, stream, a, x, y);
  GMF_HIP(hipGetLastError());
  return true;
}

}  // namespace gmf
