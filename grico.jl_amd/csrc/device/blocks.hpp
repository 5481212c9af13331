// Cell-block (additive Schwarz) smoothing of a multigrid level.
//
// The point-Jacobi smoother D^-1 cannot damp the error modes that live on
// cells whose physical part is small (finite cells cut by the boundary): their
// basis functions are nearly linearly dependent there, so the level operator
// has tiny eigenvalues that D^-1 does not lift and the coarse space cannot
// represent. With blocks, the smoother's preconditioner is
//   S = D_u^-1 + sum_b R_b^T B_b^-1 R_b,
// B_b = R_b A R_b^T the level operator restricted to the reduced dofs of block
// b (one cell), D_u^-1 the Jacobi inverse diagonal on the dofs no block
// covers (zero on covered dofs). S is symmetric positive definite when A is.
#pragma once

#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace gmf {

// Largest block (dofs) the block kernel stages in LDS.
constexpr int kMaxBlockDofs = 512;
// Blocks per 256-lane workgroup of the block kernel (one wavefront each).
constexpr int kBlocksPerGroup = 4;

// Device copy of one level's blocks; inverses stored symmetric (row = column
// major), in the level's arithmetic type R.
template <class R>
struct BlockSet {
  int64_t count = 0;
  int max_dofs = 0;
  DeviceBuffer<int64_t> offsets;  // count + 1, into dofs
  DeviceBuffer<int32_t> dofs;
  DeviceBuffer<int64_t> value_offsets;  // count + 1, into inverse (m_b^2 per block)
  DeviceBuffer<R> inverse;
};

// x += s t_b and d += s t_b (d optional) for t_b = B_b^-1 (r - q)_b (q
// optional) on every block, with atomic adds on dofs shared by blocks. One
// wavefront per block: the block's residual is staged in LDS, then lane i
// forms row i of the product (the inverse is symmetric, so lanes read a
// column: consecutive addresses).
template <class R>
__global__ void __launch_bounds__(64 * kBlocksPerGroup)
    k_block_add(R* __restrict__ x, R* __restrict__ d, const R* __restrict__ r, const R* __restrict__ q, R s,
                const int64_t* __restrict__ off, const int32_t* __restrict__ dofs, const int64_t* __restrict__ voff,
                const R* __restrict__ inv, int64_t count, int64_t ld) {
  __shared__ R rs[kBlocksPerGroup][kMaxBlockDofs];
  // column blockIdx.y of a block of columns (leading dimension ld)
  const int64_t co = int64_t(blockIdx.y) * ld;
  x += co;
  if (d) d += co;
  r += co;
  if (q) q += co;
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t b = int64_t(blockIdx.x) * kBlocksPerGroup + w;
  const bool live = b < count;
  const int64_t o = live ? off[b] : 0;
  const int m = live ? int(off[b + 1] - o) : 0;
  for (int i = lane; i < m; i += 64) {
    const int32_t g = dofs[o + i];
    rs[w][i] = q ? r[g] - q[g] : r[g];
  }
  __syncthreads();
  if (!live) return;
  const R* B = inv + voff[b];
  for (int i = lane; i < m; i += 64) {
    R t = R(0);
    for (int j = 0; j < m; ++j) t = fma(B[int64_t(j) * m + i], rs[w][j], t);
    const int32_t g = dofs[o + i];
    unsafeAtomicAdd(&x[g], s * t);
    if (d) unsafeAtomicAdd(&d[g], s * t);
  }
}

// Columns per wavefront of k_block_add_cols.
constexpr int kBlockCols = 8;

// k_block_add on up to kBlockCols columns per wavefront (column group
// blockIdx.y, leading dimension ld): each block's inverse is read once for
// the group; per column the arithmetic and summation order of k_block_add.
// Dynamic LDS: kBlocksPerGroup x kBlockCols x mmax residual values.
template <class R>
__global__ void __launch_bounds__(64 * kBlocksPerGroup)
    k_block_add_cols(R* __restrict__ x, R* __restrict__ d, const R* __restrict__ r, const R* __restrict__ q, R s,
                     const int64_t* __restrict__ off, const int32_t* __restrict__ dofs,
                     const int64_t* __restrict__ voff, const R* __restrict__ inv, int64_t count, int64_t ld, int k,
                     int mmax) {
  extern __shared__ unsigned char block_smem[];
  R* rs = reinterpret_cast<R*>(block_smem);
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t b = int64_t(blockIdx.x) * kBlocksPerGroup + w;
  const int c0 = int(blockIdx.y) * kBlockCols;
  const int kc = k - c0 < kBlockCols ? k - c0 : kBlockCols;
  const bool live = b < count;
  const int64_t o = live ? off[b] : 0;
  const int m = live ? int(off[b + 1] - o) : 0;
  R* rw = rs + int64_t(w) * kBlockCols * mmax;  // this wavefront's kc x m residuals
  for (int i = lane; i < m; i += 64) {
    const int32_t g = dofs[o + i];
    for (int c = 0; c < kc; ++c) {
      const int64_t cg = int64_t(c0 + c) * ld + g;
      rw[c * mmax + i] = q ? r[cg] - q[cg] : r[cg];
    }
  }
  __syncthreads();
  if (!live) return;
  const R* B = inv + voff[b];
  for (int i = lane; i < m; i += 64) {
    R t[kBlockCols];
#pragma unroll
    for (int c = 0; c < kBlockCols; ++c) t[c] = R(0);
    for (int j = 0; j < m; ++j) {
      const R e = B[int64_t(j) * m + i];
#pragma unroll
      for (int c = 0; c < kBlockCols; ++c) t[c] = fma(e, rw[c * mmax + j], t[c]);
    }
    const int32_t g = dofs[o + i];
    for (int c = 0; c < kc; ++c) {
      const int64_t cg = int64_t(c0 + c) * ld + g;
      unsafeAtomicAdd(&x[cg], s * t[c]);
      if (d) unsafeAtomicAdd(&d[cg], s * t[c]);
    }
  }
}

// k columns at leading dimension ld (k = 1: one vector). Several columns go
// through k_block_add_cols when its residual staging fits 64 KB of LDS.
template <class R>
inline void block_add(const BlockSet<R>& bs, R* x, R* d, const R* r, const R* q, R s, hipStream_t stream, int k = 1,
                      int64_t ld = 0) {
  if (bs.count == 0) return;
  const unsigned groups = unsigned((bs.count + kBlocksPerGroup - 1) / kBlocksPerGroup);
  const size_t cols_lds = sizeof(R) * size_t(kBlocksPerGroup) * kBlockCols * size_t(bs.max_dofs);
  if (k > 1 && cols_lds <= 64 * 1024) {
    hipLaunchKernelGGL(k_block_add_cols<R>, dim3(groups, unsigned((k + kBlockCols - 1) / kBlockCols)),
                       dim3(64 * kBlocksPerGroup), cols_lds, stream, x, d, r, q, s, bs.offsets.p, bs.dofs.p,
                       bs.value_offsets.p, bs.inverse.p, bs.count, ld, k, bs.max_dofs);
    GMF_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_block_add<R>, dim3(groups, unsigned(k)), dim3(64 * kBlocksPerGroup), 0, stream, x, d, r, q, s,
                     bs.offsets.p, bs.dofs.p, bs.value_offsets.p, bs.inverse.p, bs.count, ld);
  GMF_HIP(hipGetLastError());
}

// Host-side check of a block descriptor; returns the value offsets.
std::vector<int64_t> block_value_offsets(const gmf_block_desc& d, int64_t n);

// Upload a checked block descriptor in type R.
template <class R>
BlockSet<R> upload_blocks(const gmf_block_desc& d, int64_t n) {
  BlockSet<R> bs;
  const std::vector<int64_t> voff = block_value_offsets(d, n);
  bs.count = d.count;
  if (d.count == 0) return bs;
  for (int64_t b = 0; b < d.count; ++b) bs.max_dofs = std::max<int>(bs.max_dofs, int(d.offsets[b + 1] - d.offsets[b]));
  const size_t nd = size_t(d.offsets[d.count]), nv = size_t(voff.back());
  bs.offsets = DeviceBuffer<int64_t>(size_t(d.count) + 1);
  bs.dofs = DeviceBuffer<int32_t>(nd);
  bs.value_offsets = DeviceBuffer<int64_t>(voff.size());
  bs.inverse = DeviceBuffer<R>(nv);
  GMF_HIP(hipMemcpy(bs.offsets.p, d.offsets, (size_t(d.count) + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  GMF_HIP(hipMemcpy(bs.dofs.p, d.dofs, nd * sizeof(int32_t), hipMemcpyHostToDevice));
  GMF_HIP(hipMemcpy(bs.value_offsets.p, voff.data(), voff.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  std::vector<R> vals(nv);
  for (size_t i = 0; i < nv; ++i) vals[i] = R(d.inverse[i]);
  GMF_HIP(hipMemcpy(bs.inverse.p, vals.data(), nv * sizeof(R), hipMemcpyHostToDevice));
  return bs;
}

// invdiag[dofs[i]] = 0 for every block dof (the point part of S).
template <class R>
__global__ void k_zero_at(R* __restrict__ v, const int32_t* __restrict__ idx, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    v[idx[i]] = R(0);
}

}  // namespace gmf
