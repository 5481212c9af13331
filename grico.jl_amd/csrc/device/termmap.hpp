// Local-mode <-> global-dof expansions shared by the generic and point-cloud
// cell kernels (src/assembly.jl:1751-1809 _gather_local_coefficients! /
// _scatter_local_vector!): idx >= 0 is a single term (coefficient lcoef or 1),
// -1 an inactive mode, <= -2 a multi-term entry -(idx+2) of the CSR moff /
// midx / mcoef (midx < 0 skips a fixed dof in the reduced layout).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gmf {

struct TermMap {
  const int32_t* idx;
  const double* lcoef;
  const int64_t* moff;
  const int32_t* midx;
  const double* mcoef;
  int Ntot;
};

__device__ inline double gather_term(const TermMap& m, int64_t cell, int i, const double* __restrict__ x) {
  const int32_t g = m.idx[cell * m.Ntot + i];
  if (g >= 0) return m.lcoef ? m.lcoef[cell * m.Ntot + i] * x[g] : x[g];
  if (g == -1) return 0.0;
  const int64_t k = -int64_t(g) - 2;
  double v = 0.0;
  for (int64_t t = m.moff[k]; t < m.moff[k + 1]; ++t) {
    const int32_t h = m.midx[t];
    if (h >= 0) v = fma(m.mcoef[t], x[h], v);
  }
  return v;
}

__device__ inline void scatter_term(const TermMap& m, int64_t cell, int i, double v, double* __restrict__ y,
                                    bool squared) {
  const int32_t g = m.idx[cell * m.Ntot + i];
  if (g >= 0) {
    const double c = m.lcoef ? m.lcoef[cell * m.Ntot + i] : 1.0;
    unsafeAtomicAdd(&y[g], (squared ? c * c : c) * v);
    return;
  }
  if (g == -1) return;
  const int64_t k = -int64_t(g) - 2;
  for (int64_t t = m.moff[k]; t < m.moff[k + 1]; ++t) {
    const int32_t h = m.midx[t];
    if (h < 0) continue;
    const double c = m.mcoef[t];
    unsafeAtomicAdd(&y[h], (squared ? c * c : c) * v);
  }
}

}  // namespace gmf
