// Per-cell multilayer perceptron on BF16 matrix cores (gfx950 MFMA) with FP32
// accumulation: forward inference, backward pass and Adam training, plus the
// cell-patch gather that feeds it from reduced vectors.
//
// No reference counterpart: this is the learned block correction that plugs
// into the preconditioner seam of src/solve.jl:32 (AbstractPreconditioner),
// 641-670 (_compile_preconditioner) and 753-774 (_apply_preconditioner!);
// FGMRES accepts the resulting variable preconditioner (src/solve.jl:80-88).
//
// Network: y = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3 with W_l stored
// input-major (K_l x N_l, row-major) in BF16, biases and master weights in
// FP32. Every GEMM is one tiled kernel: 64 x 64 output tile per 256-thread
// workgroup, BK = 32 staged in LDS, four wavefronts of 2 x 2
// v_mfma_f32_16x16x16_bf16 tiles. Operand lane layout of that instruction:
// A[i][k] at lane i + 16 (k / 4), item k % 4; B[k][j] at lane j + 16 (k / 4),
// item k % 4; D[i][j] at lane j + 16 (i / 4), item i % 4.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "blas.hpp"
#include "common.hpp"
#include "internal.hpp"

typedef short gmf_bf16x4 __attribute__((ext_vector_type(4)));
typedef float gmf_f32x4 __attribute__((ext_vector_type(4)));



namespace gmf {
namespace {

#define GMF_STRIDE(i, n) \
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < (n); i += int64_t(gridDim.x) * blockDim.x)

// Round-to-nearest-even FP32 -> BF16.
__device__ inline uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

__device__ inline float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

enum Epilogue { kStoreF32 = 0, kBiasReluBF16 = 1, kBiasF32 = 2, kMaskBF16 = 3 };

struct GemmArgs {
  int M, N, K;
  const uint16_t* A;
  int64_t lda;  // A(i,k) = TA ? A[k*lda + i] : A[i*lda + k]
  const uint16_t* B;
  int64_t ldb;  // B(k,j) = TB ? B[j*ldb + k] : B[k*ldb + j]
  float* Cf;
  uint16_t* Cb;
  int64_t ldc;
  const float* bias;
  const uint16_t* mask;  // kMaskBF16: keep entries where mask(i,j) > 0 (same ldc)
};

constexpr int kBM = 64, kBN = 64, kBK = 32, kPad = 8;

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ uint16_t As[kBM][kBK + kPad];  // [m][k]
  __shared__ uint16_t Bs[kBN][kBK + kPad];  // [n][k]
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  gmf_f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = gmf_f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += kBK) {
    for (int e = t; e < kBM * kBK; e += 256) {
      int mi, ki;
      if (!TA) {
        mi = e / kBK;
        ki = e % kBK;
      } else {
        ki = e / kBM;
        mi = e % kBM;
      }
      const int gm = m0 + mi, gk = k0 + ki;
      uint16_t v = 0;
      if (gm < g.M && gk < g.K) v = TA ? g.A[int64_t(gk) * g.lda + gm] : g.A[int64_t(gm) * g.lda + gk];
      As[mi][ki] = v;
    }
    for (int e = t; e < kBN * kBK; e += 256) {
      int ni, ki;
      if (!TB) {
        ki = e / kBN;
        ni = e % kBN;
      } else {
        ni = e / kBK;
        ki = e % kBK;
      }
      const int gn = n0 + ni, gk = k0 + ki;
      uint16_t v = 0;
      if (gn < g.N && gk < g.K) v = TB ? g.B[int64_t(gn) * g.ldb + gk] : g.B[int64_t(gk) * g.ldb + gn];
      Bs[ni][ki] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 16) {
      const int kb = kk + 4 * (lane >> 4);
      gmf_bf16x4 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int row = wm + i * 16 + (lane & 15);
        a[i] = gmf_bf16x4{short(As[row][kb]), short(As[row][kb + 1]), short(As[row][kb + 2]), short(As[row][kb + 3])};
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = wn + j * 16 + (lane & 15);
        b[j] = gmf_bf16x4{short(Bs[col][kb]), short(Bs[col][kb + 1]), short(Bs[col][kb + 2]), short(Bs[col][kb + 3])};
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
        const int col = n0 + wn + j * 16 + (lane & 15);
        if (row >= g.M || col >= g.N) continue;
        float v = acc[i][j][r];
        const int64_t o = int64_t(row) * g.ldc + col;
        if (EPI == kStoreF32) {
          g.Cf[o] = v;
        } else if (EPI == kBiasReluBF16) {
          v += g.bias[col];
          g.Cb[o] = to_bf16(v > 0.f ? v : 0.f);
        } else if (EPI == kBiasF32) {
          g.Cf[o] = v + g.bias[col];
        } else {
          g.Cb[o] = (g.mask[o] != 0 && !(g.mask[o] & 0x8000u)) ? to_bf16(v) : uint16_t(0);
        }
      }
}

template <bool TA, bool TB, int EPI>
void gemm(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  dim3 grid(unsigned((g.N + kBN - 1) / kBN), unsigned((g.M + kBM - 1) / kBM));
  hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, g);
  GMF_HIP(hipGetLastError());
}

__global__ void k_to_bf16(uint16_t* __restrict__ out, const float* __restrict__ in, int64_t n) {
  GMF_STRIDE(i, n) out[i] = to_bf16(in[i]);
}

// loss partials: sum (y - t)^2 per block; dy = bf16(2 (y - t) / count).
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ tgt, uint16_t* __restrict__ dy,
                      double* __restrict__ partial, int64_t n, float scale) {
  __shared__ double red[256 / 64];
  double s = 0.0;
  GMF_STRIDE(i, n) {
    const float d = y[i] - tgt[i];
    s += double(d) * double(d);
    dy[i] = to_bf16(scale * d);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < int(blockDim.x / 64); ++w) v += red[w];
    partial[blockIdx.x] = v;
  }
}

__global__ void k_sum_partials(const double* __restrict__ partial, int count, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double v = 0.0;
    for (int i = 0; i < count; ++i) v += partial[i];
    out[0] = v;
  }
}

// Column sums of a BF16 (rows x cols) matrix into FP32 (bias gradients).
__global__ void k_colsum(const uint16_t* __restrict__ a, float* __restrict__ out, int64_t rows, int cols) {
  const int c = blockIdx.x;
  float s = 0.f;
  for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) s += from_bf16(a[r * cols + c]);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float v = 0.f;
    for (int w = 0; w < int(blockDim.x / 64); ++w) v += red[w];
    out[c] = v;
  }
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, int64_t n, float lr, float b1, float b2, float c1, float c2,
                       float eps) {
  GMF_STRIDE(i, n) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
}

// Cell patches of a reduced vector: X[c][i] = r[idx[c][i]] (0 for idx < 0) for
// the listed cells, scaled per cell to unit Euclidean norm; scale[c] keeps the
// norm. With e given, T[c][i] = e[idx] * weight[c] / scale[c].
__global__ void k_patches(const double* __restrict__ r, const double* __restrict__ e, const int32_t* __restrict__ idx,
                          const int64_t* __restrict__ cells, int64_t count, int width, const float* __restrict__ cw,
                          float* __restrict__ X, float* __restrict__ T, float* __restrict__ scale) {
  // one wavefront per patch
  const int64_t p = blockIdx.x * int64_t(blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (p >= count) return;
  const int64_t c = cells ? cells[p] : p;
  double s = 0.0;
  for (int i = lane; i < width; i += 64) {
    const int32_t j = idx[c * width + i];
    const double v = j >= 0 ? r[j] : 0.0;
    s += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  const double nrm = sqrt(s);
  const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
  for (int i = lane; i < width; i += 64) {
    const int32_t j = idx[c * width + i];
    X[p * width + i] = j >= 0 ? float(r[j] * inv) : 0.f;
    if (T) T[p * width + i] = (j >= 0 && e) ? float(e[j] * inv * double(cw[c])) : 0.f;
  }
  if (lane == 0 && scale) scale[p] = float(nrm);
}

// z[idx] += w[idx] * Y[c][i] * scale[c] / cw[c] over all cells.
__global__ void k_scatter_correction(double* __restrict__ z, const float* __restrict__ Y, const float* __restrict__ scale,
                                     const float* __restrict__ cw, const double* __restrict__ pu,
                                     const int32_t* __restrict__ idx, int64_t cells, int width) {
  GMF_STRIDE(k, cells * int64_t(width)) {
    const int64_t c = k / width;
    const int32_t j = idx[k];
    if (j < 0) continue;
    const double v = double(Y[k]) * double(scale[c]) / double(cw[c]);
    unsafeAtomicAdd(&z[j], pu[j] * v);
  }
}

// dY[c][i] = pu[idx] * g[idx] * scale[c] / cw[c] (0 where idx < 0): the output
// gradient of the per-cell network for a global gradient g of the assembled
// correction.
__global__ void k_patch_gradient(float* __restrict__ dY, const double* __restrict__ g, const float* __restrict__ scale,
                                 const float* __restrict__ cw, const double* __restrict__ pu,
                                 const int32_t* __restrict__ idx, int64_t cells, int width) {
  GMF_STRIDE(k, cells * int64_t(width)) {
    const int64_t c = k / width;
    const int32_t j = idx[k];
    dY[k] = j >= 0 ? float(pu[j] * g[j] * double(scale[c]) / double(cw[c])) : 0.f;
  }
}

__global__ void k_energy_gradient(double* __restrict__ g, const double* __restrict__ q, const double* __restrict__ r,
                                  double scale, int64_t n) {
  GMF_STRIDE(i, n) g[i] = scale * (q[i] - r[i]);
}

__global__ void k_count_cells(const int32_t* __restrict__ idx, int64_t n, double* __restrict__ count) {
  GMF_STRIDE(k, n) {
    const int32_t j = idx[k];
    if (j >= 0) unsafeAtomicAdd(&count[j], 1.0);
  }
}

__global__ void k_reciprocal(double* __restrict__ a, int64_t n) {
  GMF_STRIDE(i, n) a[i] = a[i] > 0.0 ? 1.0 / a[i] : 0.0;
}

// x_i ~ N(0, 1), counter-based: SplitMix64 of (seed, i) -> two uniforms -> Box-Muller.
__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_normal(double* __restrict__ x, int64_t n, uint64_t seed) {
  GMF_STRIDE(i, n) {
    const uint64_t a = splitmix64(seed * 0x100000001B3ull + 2 * uint64_t(i));
    const uint64_t b = splitmix64(seed * 0x100000001B3ull + 2 * uint64_t(i) + 1);
    const double u1 = (double(a >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
    const double u2 = double(b >> 11) * 0x1.0p-53;
    x[i] = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
  }
}

__global__ void k_sub(double* __restrict__ out, const double* __restrict__ a, const double* __restrict__ b, int64_t n) {
  GMF_STRIDE(i, n) out[i] = a[i] - b[i];
}

}  // namespace
}  // namespace gmf

struct gmf_mlp {
  int in = 0, h1 = 0, h2 = 0, out = 0;
  int64_t capacity = 0;  // rows of the activation buffers
  int64_t nparam = 0;
  int64_t off_w1, off_b1, off_w2, off_b2, off_w3, off_b3;
  hipStream_t stream = nullptr;
  gmf::DeviceBuffer<float> master, adam_m, adam_v, grad;
  gmf::DeviceBuffer<uint16_t> w1, w2, w3;  // BF16 copies
  gmf::DeviceBuffer<uint16_t> x16, a1, a2, dy16, dz2, dz1;
  gmf::DeviceBuffer<float> y;
  gmf::DeviceBuffer<double> partial, loss;
  int64_t adam_step = 0;

  void ensure(int64_t rows) {
    if (rows <= capacity) return;
    capacity = rows;
    x16 = gmf::DeviceBuffer<uint16_t>(size_t(rows) * size_t(in));
    a1 = gmf::DeviceBuffer<uint16_t>(size_t(rows) * size_t(h1));
    a2 = gmf::DeviceBuffer<uint16_t>(size_t(rows) * size_t(h2));
    dy16 = gmf::DeviceBuffer<uint16_t>(size_t(rows) * size_t(out));
    dz2 = gmf::DeviceBuffer<uint16_t>(size_t(rows) * size_t(h2));
    dz1 = gmf::DeviceBuffer<uint16_t>(size_t(rows) * size_t(h1));
    y = gmf::DeviceBuffer<float>(size_t(rows) * size_t(out));
  }

  void cast_weights() {
    using namespace gmf;
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1(int64_t(in) * h1)), dim3(256), 0, stream, w1.p, master.p + off_w1,
                       int64_t(in) * h1);
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1(int64_t(h1) * h2)), dim3(256), 0, stream, w2.p, master.p + off_w2,
                       int64_t(h1) * h2);
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1(int64_t(h2) * out)), dim3(256), 0, stream, w3.p, master.p + off_w3,
                       int64_t(h2) * out);
    GMF_HIP(hipGetLastError());
  }

  // y = net(x) for `rows` FP32 inputs (device).
  void forward(const float* x, int64_t rows) {
    using namespace gmf;
    ensure(rows);
    const int64_t nx = rows * in;
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1(nx)), dim3(256), 0, stream, x16.p, x, nx);
    GMF_HIP(hipGetLastError());
    GemmArgs g{};
    g = GemmArgs{int(rows), h1, in, x16.p, in, w1.p, h1, nullptr, a1.p, h1, master.p + off_b1, nullptr};
    gemm<false, false, kBiasReluBF16>(g, stream);
    g = GemmArgs{int(rows), h2, h1, a1.p, h1, w2.p, h2, nullptr, a2.p, h2, master.p + off_b2, nullptr};
    gemm<false, false, kBiasReluBF16>(g, stream);
    g = GemmArgs{int(rows), out, h2, a2.p, h2, w3.p, out, y.p, nullptr, out, master.p + off_b3, nullptr};
    gemm<false, false, kBiasF32>(g, stream);
  }

  // Loss sum((y - t)^2) / (rows * out) and its parameter gradient into grad.
  void backward(const float* tgt, int64_t rows) {
    using namespace gmf;
    const int64_t ny = rows * out;
    const unsigned blocks = grid1(ny);
    const float scale = 2.0f / float(ny);
    hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, stream, y.p, tgt, dy16.p, partial.p, ny, scale);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, stream, partial.p, int(blocks), loss.p);
    GMF_HIP(hipGetLastError());
    backprop(rows);
  }

  // Parameter gradient for an FP32 output gradient dy (rows x out) of the last forward.
  void backward_from(const float* dy, int64_t rows) {
    using namespace gmf;
    const int64_t ny = rows * out;
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1(ny)), dim3(256), 0, stream, dy16.p, dy, ny);
    GMF_HIP(hipGetLastError());
    backprop(rows);
  }

  // Backpropagation from dy16 through the activations of the last forward.
  void backprop(int64_t rows) {
    using namespace gmf;
    // layer 3: dW3 = a2^T dy, db3 = colsum(dy), dz2 = (dy W3^T) * (a2 > 0)
    GemmArgs g{};
    g = GemmArgs{h2, out, int(rows), a2.p, h2, dy16.p, out, grad.p + off_w3, nullptr, out, nullptr, nullptr};
    gemm<true, false, kStoreF32>(g, stream);
    hipLaunchKernelGGL(k_colsum, dim3(unsigned(out)), dim3(256), 0, stream, dy16.p, grad.p + off_b3, rows, out);
    g = GemmArgs{int(rows), h2, out, dy16.p, out, w3.p, out, nullptr, dz2.p, h2, nullptr, a2.p};
    gemm<false, true, kMaskBF16>(g, stream);
    // layer 2
    g = GemmArgs{h1, h2, int(rows), a1.p, h1, dz2.p, h2, grad.p + off_w2, nullptr, h2, nullptr, nullptr};
    gemm<true, false, kStoreF32>(g, stream);
    hipLaunchKernelGGL(k_colsum, dim3(unsigned(h2)), dim3(256), 0, stream, dz2.p, grad.p + off_b2, rows, h2);
    g = GemmArgs{int(rows), h1, h2, dz2.p, h2, w2.p, h2, nullptr, dz1.p, h1, nullptr, a1.p};
    gemm<false, true, kMaskBF16>(g, stream);
    // layer 1
    g = GemmArgs{in, h1, int(rows), x16.p, in, dz1.p, h1, grad.p + off_w1, nullptr, h1, nullptr, nullptr};
    gemm<true, false, kStoreF32>(g, stream);
    hipLaunchKernelGGL(k_colsum, dim3(unsigned(h1)), dim3(256), 0, stream, dz1.p, grad.p + off_b1, rows, h1);
    GMF_HIP(hipGetLastError());
  }

  void adam(float lr, float b1, float b2, float eps) {
    using namespace gmf;
    ++adam_step;
    const float c1 = 1.0f / (1.0f - std::pow(b1, float(adam_step)));
    const float c2 = 1.0f / (1.0f - std::pow(b2, float(adam_step)));
    hipLaunchKernelGGL(k_adam, dim3(grid1(nparam)), dim3(256), 0, stream, master.p, adam_m.p, adam_v.p, grad.p, nparam,
                       lr, b1, b2, c1, c2, eps);
    GMF_HIP(hipGetLastError());
    cast_weights();
  }
};

// ---- learned block correction preconditioner ---------------------------------

namespace gmf {
namespace {

struct LearnedPrecond : gmf_preconditioner {
  gmf_preconditioner* inner = nullptr;  // not owned
  gmf_mlp* net = nullptr;               // not owned
  int64_t cells = 0;
  int width = 0;
  const int32_t* idx = nullptr;  // bucket reduced index map (cells x width)
  DeviceBuffer<float> cw, X, scale;
  DeviceBuffer<double> pu, q, r1;
  double omega = 1.0;

  // c = sum_cells P^T (w * net(patches(r1)) * scale / cw) into z (added).
  void correction(const double* r1v, double* z) {
    hipStream_t s = ws->stream;
    const unsigned pb = unsigned((cells + 3) / 4);
    hipLaunchKernelGGL(k_patches, dim3(pb), dim3(256), 0, s, r1v, static_cast<const double*>(nullptr), idx,
                       static_cast<const int64_t*>(nullptr), cells, width, cw.p, X.p, static_cast<float*>(nullptr),
                       scale.p);
    GMF_HIP(hipGetLastError());
    net->stream = s;
    net->forward(X.p, cells);
    hipLaunchKernelGGL(k_scatter_correction, dim3(grid1(cells * width)), dim3(256), 0, s, z, net->y.p, scale.p, cw.p,
                       pu.p, idx, cells, width);
    GMF_HIP(hipGetLastError());
  }

  // One step of the energy-norm objective on residual rk (energy en = e'^T r'):
  // the correction c of the network, q = A c, the loss ||e' - c||_A^2 / en =
  // 1 + (c^T q - 2 c^T r') / en (returned), and its parameter gradient in
  // net->grad (no parameter update). c and gq: n doubles, dY: cells x width floats.
  double energy_step(const double* rk, double en, double* c, double* gq, float* dY) {
    const int64_t n = ws->plan->n_red;
    hipStream_t s = ws->stream;
    net->stream = s;
    fill(c, 0.0, n, s);
    correction(rk, c);
    reduced_apply(ws, c, q.p);
    dot(ws, c, q.p, n, kTmp);
    dot(ws, c, rk, n, kTmp + 1);
    read_scalars(ws, kTmp + 2);
    const double cq = ws->host_scalars[kTmp], cr = ws->host_scalars[kTmp + 1];
    hipLaunchKernelGGL(k_energy_gradient, dim3(grid1(n)), dim3(256), 0, s, gq, q.p, rk, 2.0 / en, n);
    hipLaunchKernelGGL(k_patch_gradient, dim3(grid1(cells * width)), dim3(256), 0, s, dY, gq, scale.p, cw.p, pu.p, idx,
                       cells, width);
    GMF_HIP(hipGetLastError());
    net->backward_from(dY, cells);
    return 1.0 + (cq - 2.0 * cr) / en;
  }

  // Training pairs from GMG-CG solves with the inner preconditioner V: for
  // reference solutions x* ~ N(0, 1) (seeded, one per solve) and right-hand
  // sides b = A x*, at every CG iterate x_k the pair y = x_k + V(b - A x_k),
  // r' = b - A y (stored as an FP32 row of residuals), e' = x* - y and
  // energy e'^T r'. Up to `iterations` pairs per solve; a solve also ends when
  // its residual drops below 1e-12 ||b||.
  void collect(int64_t pairs, int iterations, uint64_t seed, float* residuals, double* energies) {
    const int64_t n = ws->plan->n_red;
    hipStream_t s = ws->stream;
    require(inner != nullptr, GMF_EINVAL, "collecting training pairs needs an inner preconditioner");
    inner->ws->stream = s;
    DeviceBuffer<double> xs(static_cast<size_t>(n)), b(static_cast<size_t>(n)), x(static_cast<size_t>(n)),
        r(static_cast<size_t>(n)), z(static_cast<size_t>(n)), p(static_cast<size_t>(n)), y(static_cast<size_t>(n)),
        e1(static_cast<size_t>(n));
    const double* hs = ws->host_scalars;
    auto dot1 = [&](const double* u, const double* v) {
      dot(ws, u, v, n, kTmp);
      read_scalars(ws, kTmp + 1);
      return hs[kTmp];
    };
    int64_t got = 0;
    for (uint64_t solve = 0; got < pairs; ++solve) {
      hipLaunchKernelGGL(k_normal, dim3(grid1(n)), dim3(256), 0, s, xs.p, n, seed + solve);
      GMF_HIP(hipGetLastError());
      reduced_apply(ws, xs.p, b.p);
      fill(x.p, 0.0, n, s);
      copy(r.p, b.p, n, s);
      inner->apply(r.p, z.p);
      copy(p.p, z.p, n, s);
      double rz = dot1(r.p, z.p);
      const double bnorm = std::sqrt(dot1(b.p, b.p));
      for (int it = 0; it < iterations && got < pairs; ++it) {
        copy(y.p, x.p, n, s);
        axpby(ws, y.p, 1.0, z.p, 1.0, n);
        reduced_apply(ws, y.p, q.p);
        copy(r1.p, b.p, n, s);
        axpby(ws, r1.p, -1.0, q.p, 1.0, n);
        copy(e1.p, xs.p, n, s);
        axpby(ws, e1.p, -1.0, y.p, 1.0, n);
        narrow(residuals + got * n, r1.p, n, s);
        energies[got] = dot1(e1.p, r1.p);
        ++got;
        reduced_apply(ws, p.p, q.p);
        const double pq = dot1(p.p, q.p);
        require(pq > 0.0 && std::isfinite(pq), GMF_EBREAKDOWN, "CG encountered a non-positive operator direction");
        const double alpha = rz / pq;
        axpby(ws, x.p, alpha, p.p, 1.0, n);
        axpby(ws, r.p, -alpha, q.p, 1.0, n);
        if (std::sqrt(dot1(r.p, r.p)) <= 1e-12 * bnorm) break;
        inner->apply(r.p, z.p);
        const double rz_new = dot1(r.p, z.p);
        axpby(ws, p.p, 1.0, z.p, rz_new / rz, n);
        rz = rz_new;
      }
    }
    sync(ws);
  }

  // Energy-norm training on residuals r'_k left by the inner preconditioner
  // (FP32 rows, widened per step): minimise ||e' - c||_A^2 = e'^T A e' -
  // 2 c^T r' + c^T A c over the network, one residual per Adam step.
  // history[s] = ||e' - c||_A^2 / ||e'||_A^2 before the step (energy[k] =
  // e'^T r' = ||e'||_A^2).
  void train(const float* residuals, const double* energy, int64_t pairs, int64_t steps, float lr, float b1, float b2,
             float eps, double* history) {
    const int64_t n = ws->plan->n_red;
    DeviceBuffer<double> c(static_cast<size_t>(n)), gq(static_cast<size_t>(n)), rk(static_cast<size_t>(n));
    DeviceBuffer<float> dY(size_t(cells) * size_t(width));
    for (int64_t st = 0; st < steps; ++st) {
      const int64_t k = st % pairs;
      const double en = energy[k] > 0.0 ? energy[k] : 1.0;
      widen(rk.p, residuals + k * n, n, ws->stream);
      const double loss = energy_step(rk.p, en, c.p, gq.p, dY.p);
      if (history) history[st] = loss;
      net->adam(lr, b1, b2, eps);
    }
    sync(ws);
  }

  void apply(const double* r, double* z) override {
    const int64_t n = ws->plan->n_red;
    hipStream_t s = ws->stream;
    if (inner) {
      inner->ws->stream = s;
      inner->apply(r, z);
      reduced_apply(ws, z, q.p);
      hipLaunchKernelGGL(k_sub, dim3(grid1(n)), dim3(256), 0, s, r1.p, r, q.p, n);
    } else {
      fill(z, 0.0, n, s);
      copy(r1.p, r, n, s);
    }
    correction(r1.p, z);
  }
};

const Bucket& single_bucket(gmf_workspace* ws) {
  const gmf_plan& p = *ws->plan;
  require(p.buckets.size() == 1, GMF_EUNSUPPORTED, "the learned correction needs a single-batch plan");
  require(p.selection_map && p.buckets[0].ridx, GMF_EUNSUPPORTED, "the learned correction needs a selection reduced map");
  require(p.buckets[0].multi_count == 0, GMF_EUNSUPPORTED, "the learned correction needs single-term cell maps");
  return p.buckets[0];
}

}  // namespace
}  // namespace gmf

using namespace gmf;

extern "C" {

gmf_status gmf_mlp_create(const gmf_mlp_desc* d, gmf_mlp** out) {
  return guarded_call([&] {
    require(d && out, GMF_EINVAL, "null argument");
    require(d->input > 0 && d->hidden1 > 0 && d->hidden2 > 0 && d->output > 0, GMF_EINVAL, "layer widths must be positive");
    require(d->parameters != nullptr, GMF_EINVAL, "initial parameters are required");
    auto m = std::make_unique<gmf_mlp>();
    m->in = d->input;
    m->h1 = d->hidden1;
    m->h2 = d->hidden2;
    m->out = d->output;
    int64_t o = 0;
    m->off_w1 = o;
    o += int64_t(m->in) * m->h1;
    m->off_b1 = o;
    o += m->h1;
    m->off_w2 = o;
    o += int64_t(m->h1) * m->h2;
    m->off_b2 = o;
    o += m->h2;
    m->off_w3 = o;
    o += int64_t(m->h2) * m->out;
    m->off_b3 = o;
    o += m->out;
    m->nparam = o;
    m->master = DeviceBuffer<float>(size_t(o));
    m->adam_m = DeviceBuffer<float>(size_t(o));
    m->adam_v = DeviceBuffer<float>(size_t(o));
    m->grad = DeviceBuffer<float>(size_t(o));
    m->w1 = DeviceBuffer<uint16_t>(size_t(m->in) * size_t(m->h1));
    m->w2 = DeviceBuffer<uint16_t>(size_t(m->h1) * size_t(m->h2));
    m->w3 = DeviceBuffer<uint16_t>(size_t(m->h2) * size_t(m->out));
    m->partial = DeviceBuffer<double>(size_t(4096));
    m->loss = DeviceBuffer<double>(size_t(1));
    GMF_HIP(hipMemcpy(m->master.p, d->parameters, size_t(o) * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> zero(size_t(o), 0.f);
    GMF_HIP(hipMemcpy(m->adam_m.p, zero.data(), size_t(o) * sizeof(float), hipMemcpyHostToDevice));
    GMF_HIP(hipMemcpy(m->adam_v.p, zero.data(), size_t(o) * sizeof(float), hipMemcpyHostToDevice));
    m->cast_weights();
    GMF_HIP(hipStreamSynchronize(m->stream));
    *out = m.release();
  });
}

gmf_status gmf_mlp_destroy(gmf_mlp* m) {
  return guarded_call([&] { delete m; });
}

gmf_status gmf_mlp_parameter_count(const gmf_mlp* m, int64_t* count) {
  return guarded_call([&] {
    require(m && count, GMF_EINVAL, "null argument");
    *count = m->nparam;
  });
}

gmf_status gmf_mlp_get_parameters(gmf_mlp* m, float* host) {
  return guarded_call([&] {
    require(m && host, GMF_EINVAL, "null argument");
    GMF_HIP(hipStreamSynchronize(m->stream));
    GMF_HIP(hipMemcpy(host, m->master.p, size_t(m->nparam) * sizeof(float), hipMemcpyDeviceToHost));
  });
}

gmf_status gmf_mlp_set_stream(gmf_mlp* m, void* stream) {
  return guarded_call([&] {
    require(m, GMF_EINVAL, "null argument");
    m->stream = static_cast<hipStream_t>(stream);
  });
}

gmf_status gmf_mlp_forward(gmf_mlp* m, const float* x, float* y, int64_t rows) {
  return guarded_call([&] {
    require(m && x && y && rows >= 0, GMF_EINVAL, "invalid argument");
    require(rows < (int64_t(1) << 31), GMF_EUNSUPPORTED, "batch too large");
    if (rows == 0) return;
    require_device_pointer(x, "x");
    require_device_pointer(y, "y");
    m->forward(x, rows);
    GMF_HIP(hipMemcpyAsync(y, m->y.p, size_t(rows) * size_t(m->out) * sizeof(float), hipMemcpyDeviceToDevice,
                           m->stream));
    GMF_HIP(hipStreamSynchronize(m->stream));
  });
}

gmf_status gmf_mlp_gradient(gmf_mlp* m, const float* x, const float* t, int64_t rows, double* loss, float* grad_host) {
  return guarded_call([&] {
    require(m && x && t && rows > 0, GMF_EINVAL, "invalid argument");
    require(rows < (int64_t(1) << 31), GMF_EUNSUPPORTED, "batch too large");
    require_device_pointer(x, "x");
    require_device_pointer(t, "t");
    m->forward(x, rows);
    m->backward(t, rows);
    double l = 0.0;
    GMF_HIP(hipMemcpyAsync(&l, m->loss.p, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    GMF_HIP(hipStreamSynchronize(m->stream));
    if (loss) *loss = l / double(rows * m->out);
    if (grad_host) GMF_HIP(hipMemcpy(grad_host, m->grad.p, size_t(m->nparam) * sizeof(float), hipMemcpyDeviceToHost));
  });
}

gmf_status gmf_mlp_train(gmf_mlp* m, const float* x, const float* t, int64_t samples, int64_t batch, int64_t steps,
                         float lr, float beta1, float beta2, float eps, double* loss_history) {
  return guarded_call([&] {
    require(m && x && t && samples > 0 && batch > 0 && steps >= 0, GMF_EINVAL, "invalid argument");
    require(batch <= samples, GMF_EINVAL, "batch larger than the data set");
    require(batch < (int64_t(1) << 31), GMF_EUNSUPPORTED, "batch too large");
    require_device_pointer(x, "x");
    require_device_pointer(t, "t");
    require(std::isfinite(lr) && lr > 0 && beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps > 0, GMF_EINVAL,
            "invalid Adam parameters");
    const int64_t nb = samples / batch;
    for (int64_t s = 0; s < steps; ++s) {
      const int64_t b = s % nb;
      m->forward(x + b * batch * m->in, batch);
      m->backward(t + b * batch * m->out, batch);
      m->adam(lr, beta1, beta2, eps);
      if (loss_history) {
        double l = 0.0;
        GMF_HIP(hipMemcpyAsync(&l, m->loss.p, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        GMF_HIP(hipStreamSynchronize(m->stream));
        loss_history[s] = l / double(batch * m->out);
      }
    }
    GMF_HIP(hipStreamSynchronize(m->stream));
  });
}

gmf_status gmf_cell_patches(gmf_workspace* ws, const double* r, const double* e, const int64_t* cells, int64_t count,
                            const float* cell_weight, float* x, float* t, float* scale) {
  return guarded_call([&] {
    require(ws && r && x && cell_weight && count >= 0, GMF_EINVAL, "invalid argument");
    require(!e || t, GMF_EINVAL, "targets need an output buffer");
    require_device_pointer(r, "r");
    require_device_pointer(x, "x");
    require_device_pointer(cell_weight, "cell_weight");
    if (e) {
      require_device_pointer(e, "e");
      require_device_pointer(t, "t");
    }
    if (cells) require_device_pointer(cells, "cells");
    if (scale) require_device_pointer(scale, "scale");
    const Bucket& b = single_bucket(ws);
    if (count == 0) return;
    if (cells) {
      std::vector<int64_t> h(static_cast<size_t>(count));
      GMF_HIP(hipMemcpy(h.data(), cells, size_t(count) * sizeof(int64_t), hipMemcpyDeviceToHost));
      for (int64_t c : h) require(c >= 0 && c < b.cells, GMF_EINVAL, "cell index out of range");
    } else {
      require(count == b.cells, GMF_EINVAL, "count must equal the cell count when no cell list is given");
    }
    hipLaunchKernelGGL(k_patches, dim3(unsigned((count + 3) / 4)), dim3(256), 0, ws->stream, r, e, b.ridx, cells, count,
                       b.Ntot, cell_weight, x, e ? t : nullptr, scale);
    GMF_HIP(hipGetLastError());
    GMF_HIP(hipStreamSynchronize(ws->stream));
  });
}

gmf_status gmf_precond_create_learned(gmf_workspace* ws, gmf_preconditioner* inner, gmf_mlp* net,
                                      const float* cell_weight, gmf_preconditioner** out) {
  return guarded_call([&] {
    require(ws && net && cell_weight && out, GMF_EINVAL, "null argument");
    require(inner == nullptr || inner->ws == ws, GMF_EINVAL, "inner preconditioner belongs to a different workspace");
    const Bucket& b = single_bucket(ws);
    require(net->in == b.Ntot && net->out == b.Ntot, GMF_EINVAL, "network width must equal the cell's local modes");
    auto L = std::make_unique<LearnedPrecond>();
    L->ws = ws;
    L->inner = inner;
    L->net = net;
    L->cells = b.cells;
    L->width = b.Ntot;
    L->idx = b.ridx;
    const int64_t n = ws->plan->n_red;
    L->cw = DeviceBuffer<float>(size_t(b.cells));
    GMF_HIP(hipMemcpy(L->cw.p, cell_weight, size_t(b.cells) * sizeof(float), hipMemcpyHostToDevice));
    L->X = DeviceBuffer<float>(size_t(b.cells) * size_t(b.Ntot));
    L->scale = DeviceBuffer<float>(size_t(b.cells));
    L->pu = DeviceBuffer<double>(size_t(n));
    L->q = DeviceBuffer<double>(size_t(n));
    L->r1 = DeviceBuffer<double>(size_t(n));
    fill(L->pu.p, 0.0, n, ws->stream);
    hipLaunchKernelGGL(k_count_cells, dim3(grid1(b.cells * b.Ntot)), dim3(256), 0, ws->stream, b.ridx,
                       b.cells * int64_t(b.Ntot), L->pu.p);
    hipLaunchKernelGGL(k_reciprocal, dim3(grid1(n)), dim3(256), 0, ws->stream, L->pu.p, n);
    GMF_HIP(hipGetLastError());
    net->ensure(b.cells);
    GMF_HIP(hipStreamSynchronize(ws->stream));
    *out = L.release();
  });
}

gmf_status gmf_learned_collect(gmf_preconditioner* learned, int64_t pairs, int32_t iterations, uint64_t seed,
                               float* residuals, double* energies) {
  return guarded_call([&] {
    auto* L = dynamic_cast<LearnedPrecond*>(learned);
    require(L != nullptr, GMF_EINVAL, "not a learned-correction preconditioner");
    require(residuals && energies && pairs > 0 && iterations > 0, GMF_EINVAL, "invalid argument");
    require_device_pointer(residuals, "residuals");
    L->collect(pairs, iterations, seed, residuals, energies);
  });
}

gmf_status gmf_learned_train(gmf_preconditioner* learned, const float* residuals, const double* energies,
                             int64_t pairs, int64_t steps, float lr, float beta1, float beta2, float eps,
                             double* history) {
  return guarded_call([&] {
    auto* L = dynamic_cast<LearnedPrecond*>(learned);
    require(L != nullptr, GMF_EINVAL, "not a learned-correction preconditioner");
    require(residuals && energies && pairs > 0 && steps >= 0, GMF_EINVAL, "invalid argument");
    require(std::isfinite(lr) && lr > 0 && beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps > 0, GMF_EINVAL,
            "invalid Adam parameters");
    require_device_pointer(residuals, "residuals");
    L->train(residuals, energies, pairs, steps, lr, beta1, beta2, eps, history);
  });
}

gmf_status gmf_learned_gradient(gmf_preconditioner* learned, const double* residual, double energy, double* loss,
                                float* grad_host) {
  return guarded_call([&] {
    auto* L = dynamic_cast<LearnedPrecond*>(learned);
    require(L != nullptr, GMF_EINVAL, "not a learned-correction preconditioner");
    require(residual && std::isfinite(energy) && energy > 0.0, GMF_EINVAL, "invalid argument");
    require_device_pointer(residual, "residual");
    const int64_t n = L->ws->plan->n_red;
    DeviceBuffer<double> c(static_cast<size_t>(n)), gq(static_cast<size_t>(n));
    DeviceBuffer<float> dY(size_t(L->cells) * size_t(L->width));
    const double l = L->energy_step(residual, energy, c.p, gq.p, dY.p);
    sync(L->ws);
    if (loss) *loss = l;
    if (grad_host)
      GMF_HIP(hipMemcpy(grad_host, L->net->grad.p, size_t(L->net->nparam) * sizeof(float), hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
