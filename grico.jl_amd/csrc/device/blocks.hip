// Cell blocks of a reduced operator for block smoothing: descriptor checks
// and the probing entry point gmf_reduced_blocks (blocks.hpp holds the block
// kernel the multigrid smoother runs).
//
// Probing: the blocks of one color touch disjoint sets of cells, so the unit
// vectors of their j-th dofs can be applied together: row i of block b reads
// (A e)_{dof_i} = A[dof_i][dof_j(b)] plus contributions of other blocks' unit
// vectors, which vanish because no cell couples dof_i to a dof of another
// block of the color. One reduced apply per color and local position.
#include <algorithm>
#include <cmath>
#include <vector>

#include "blocks.hpp"

namespace gmf {

std::vector<int64_t> block_value_offsets(const gmf_block_desc& d, int64_t n) {
  require(d.count >= 0, GMF_EINVAL, "block count must be non-negative");
  std::vector<int64_t> voff(size_t(d.count) + 1, 0);
  if (d.count == 0) return voff;
  require(d.offsets != nullptr && d.dofs != nullptr, GMF_EINVAL, "block arrays missing");
  require(d.offsets[0] == 0, GMF_EINVAL, "block offsets must start at 0");
  require(n < (int64_t(1) << 31), GMF_EUNSUPPORTED, "level too large for 32-bit block indices");
  std::vector<int64_t> seen(size_t(std::max<int64_t>(n, 1)), -1);
  for (int64_t b = 0; b < d.count; ++b) {
    const int64_t m = d.offsets[b + 1] - d.offsets[b];
    require(m >= 1 && m <= kMaxBlockDofs, GMF_EINVAL, "a block must hold 1 to 512 dofs");
    for (int64_t i = d.offsets[b]; i < d.offsets[b + 1]; ++i) {
      const int32_t g = d.dofs[i];
      require(g >= 0 && g < n, GMF_EINVAL, "block dof out of range");
      require(seen[size_t(g)] != b, GMF_EINVAL, "a block lists a dof twice");
      seen[size_t(g)] = b;
    }
    voff[size_t(b) + 1] = voff[size_t(b)] + m * m;
  }
  return voff;
}

namespace {

// e[dof_j(b)] = v for the blocks b of one color that have a j-th dof.
__global__ void k_probe_set(double* __restrict__ e, const int32_t* __restrict__ cb, int64_t nc,
                            const int64_t* __restrict__ off, const int32_t* __restrict__ dofs, int j, double v) {
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < nc; t += int64_t(gridDim.x) * blockDim.x) {
    const int32_t b = cb[t];
    if (j < off[b + 1] - off[b]) e[dofs[off[b] + j]] = v;
  }
}

// Column j of every block of one color: out[voff_b + i m + j] = y[dof_i(b)].
// One workgroup per block of the color, lanes over i.
__global__ void k_probe_gather(double* __restrict__ out, const double* __restrict__ y, const int32_t* __restrict__ cb,
                               const int64_t* __restrict__ off, const int32_t* __restrict__ dofs,
                               const int64_t* __restrict__ voff, int j) {
  const int32_t b = cb[blockIdx.x];
  const int64_t o = off[b];
  const int m = int(off[b + 1] - o);
  if (j >= m) return;
  for (int i = threadIdx.x; i < m; i += blockDim.x) out[voff[b] + int64_t(i) * m + j] = y[dofs[o + i]];
}

}  // namespace
}  // namespace gmf

using namespace gmf;

extern "C" gmf_status gmf_reduced_blocks(gmf_workspace* ws, const gmf_block_desc* blocks, const int32_t* colors,
                                         int32_t color_count, double* out) {
  return guarded_call([&] {
    require(ws != nullptr && blocks != nullptr && out != nullptr, GMF_EINVAL, "invalid argument");
    const int64_t n = ws->plan->n_red;
    const std::vector<int64_t> voff = block_value_offsets(*blocks, n);
    const int64_t nb = blocks->count;
    if (nb == 0) return;
    require(colors != nullptr && color_count >= 1, GMF_EINVAL, "block colors missing");
    // Blocks grouped by color (CSR), and each color's largest block.
    std::vector<int64_t> cptr(size_t(color_count) + 1, 0);
    for (int64_t b = 0; b < nb; ++b) {
      require(colors[b] >= 0 && colors[b] < color_count, GMF_EINVAL, "block color out of range");
      cptr[size_t(colors[b]) + 1]++;
    }
    for (int c = 0; c < color_count; ++c) cptr[size_t(c) + 1] += cptr[size_t(c)];
    std::vector<int32_t> cblocks(static_cast<size_t>(nb));
    std::vector<int> cmax(size_t(color_count), 0);
    {
      std::vector<int64_t> fillp(cptr.begin(), cptr.end() - 1);
      for (int64_t b = 0; b < nb; ++b) {
        cblocks[size_t(fillp[size_t(colors[b])]++)] = int32_t(b);
        cmax[size_t(colors[b])] =
            std::max<int>(cmax[size_t(colors[b])], int(blocks->offsets[b + 1] - blocks->offsets[b]));
      }
    }
    hipStream_t s = ws->stream;
    const size_t nd = size_t(blocks->offsets[nb]);
    DeviceBuffer<int64_t> d_off(size_t(nb) + 1), d_voff(voff.size());
    DeviceBuffer<int32_t> d_dofs(nd), d_cb(static_cast<size_t>(nb));
    DeviceBuffer<double> e(static_cast<size_t>(n)), y(static_cast<size_t>(n)), vals(static_cast<size_t>(voff.back()));
    GMF_HIP(hipMemcpy(d_off.p, blocks->offsets, (size_t(nb) + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    GMF_HIP(hipMemcpy(d_voff.p, voff.data(), voff.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    GMF_HIP(hipMemcpy(d_dofs.p, blocks->dofs, nd * sizeof(int32_t), hipMemcpyHostToDevice));
    GMF_HIP(hipMemcpy(d_cb.p, cblocks.data(), size_t(nb) * sizeof(int32_t), hipMemcpyHostToDevice));
    fill(e.p, 0.0, n, s);
    for (int c = 0; c < color_count; ++c) {
      const int64_t nc = cptr[size_t(c) + 1] - cptr[size_t(c)];
      if (nc == 0) continue;
      const int32_t* cb = d_cb.p + cptr[size_t(c)];
      for (int j = 0; j < cmax[size_t(c)]; ++j) {
        hipLaunchKernelGGL(k_probe_set, dim3(grid1(nc)), dim3(256), 0, s, e.p, cb, nc, d_off.p, d_dofs.p, j, 1.0);
        reduced_apply(ws, e.p, y.p);
        hipLaunchKernelGGL(k_probe_gather, dim3(unsigned(nc)), dim3(64), 0, s, vals.p, y.p, cb, d_off.p, d_dofs.p,
                           d_voff.p, j);
        hipLaunchKernelGGL(k_probe_set, dim3(grid1(nc)), dim3(256), 0, s, e.p, cb, nc, d_off.p, d_dofs.p, j, 0.0);
        GMF_HIP(hipGetLastError());
      }
    }
    GMF_HIP(hipMemcpyAsync(out, vals.p, size_t(voff.back()) * sizeof(double), hipMemcpyDeviceToHost, s));
    GMF_HIP(hipStreamSynchronize(s));
  });
}
