// Streaming vector kernels: fills, copies, deterministic two-stage dot
// products and the fused Krylov updates (src/solve.jl:776-817 _dot,
// _dot_self, _axpy!, _update_direction!). All HBM-bound; grid-stride loops
// with 256-thread blocks, two doubles per thread per iteration.
#include "blas.hpp"

namespace gmf {

constexpr int kBlock = 256;
constexpr int kMaxPartials = 2048;

inline unsigned grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > kMaxPartials) g = kMaxPartials;
  return unsigned(g);
}

template <class T>
__global__ void k_fill(T* __restrict__ x, T v, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    x[i] = v;
}

template <class T, class S>
__global__ void k_copy(T* __restrict__ d, const S* __restrict__ s, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    d[i] = T(s[i]);
}

__device__ inline double block_sum(double v) {
  __shared__ double red[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

__global__ void __launch_bounds__(kBlock) k_dot_partial(const double* __restrict__ x, const double* __restrict__ y,
                                                        int64_t n, double* __restrict__ partials) {
  double acc = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    acc = fma(x[i], y[i], acc);
  double s = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock) k_reduce_partials(const double* __restrict__ partials, int count,
                                                            double* __restrict__ out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < count; i += blockDim.x) acc += partials[i];
  double s = block_sum(acc);
  if (threadIdx.x == 0) *out = s;
}

void fill(double* x, double v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n)), dim3(kBlock), 0, s, x, v, n);
  GMF_HIP(hipGetLastError());
}

void fill(float* x, float v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n)), dim3(kBlock), 0, s, x, v, n);
  GMF_HIP(hipGetLastError());
}

void copy(double* dst, const double* src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_copy<double, double>), dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
  GMF_HIP(hipGetLastError());
}

void copy(float* dst, const float* src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_copy<float, float>), dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
  GMF_HIP(hipGetLastError());
}

void narrow(float* dst, const double* src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_copy<float, double>), dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
  GMF_HIP(hipGetLastError());
}

void widen(double* dst, const float* src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_copy<double, float>), dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
  GMF_HIP(hipGetLastError());
}

void dot(gmf_workspace* ws, const double* x, const double* y, int64_t n, int slot) {
  unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_dot_partial, dim3(g), dim3(kBlock), 0, ws->stream, x, y, n, ws->partials);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, ws->stream, ws->partials, int(g),
                     ws->scalars + slot);
  GMF_HIP(hipGetLastError());
}

void read_scalars(gmf_workspace* ws, int count) {
  GMF_HIP(hipMemcpyAsync(ws->host_scalars, ws->scalars, sizeof(double) * size_t(count), hipMemcpyDeviceToHost,
                         ws->stream));
  GMF_HIP(hipStreamSynchronize(ws->stream));
}

void read_scalars_async(gmf_workspace* ws, int count) {
  GMF_HIP(hipMemcpyAsync(ws->pipelined_scalars, ws->scalars, sizeof(double) * size_t(count), hipMemcpyDeviceToHost,
                         ws->stream));
  GMF_HIP(hipEventRecord(ws->scalar_event, ws->stream));
}

void wait_scalars(gmf_workspace* ws) { GMF_HIP(hipEventSynchronize(ws->scalar_event)); }

// ---- Krylov building blocks ------------------------------------------------

// x += alpha p; r -= alpha q; partial sums of r.r. alpha = s[num] / s[den].
__global__ void __launch_bounds__(kBlock) k_cg_update(double* __restrict__ x, double* __restrict__ r,
                                                      const double* __restrict__ p, const double* __restrict__ q,
                                                      int64_t n, const double* __restrict__ s, int num, int den,
                                                      double* __restrict__ partials) {
  const double alpha = s[num] / s[den];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    x[i] = fma(alpha, p[i], x[i]);
    const double ri = fma(-alpha, q[i], r[i]);
    r[i] = ri;
    acc = fma(ri, ri, acc);
  }
  double t = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// z = invdiag .* r (or copy) and partial sums of r.z.
__global__ void __launch_bounds__(kBlock) k_jacobi_dot(double* __restrict__ z, const double* __restrict__ r,
                                                       const double* __restrict__ invdiag, int64_t n,
                                                       double* __restrict__ partials) {
  double acc = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const double zi = invdiag ? invdiag[i] * r[i] : r[i];
    z[i] = zi;
    acc = fma(r[i], zi, acc);
  }
  double t = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// p = z + beta p with beta = s[num]/s[den]; then s[den] = s[num] (one thread).
__global__ void __launch_bounds__(kBlock) k_direction(double* __restrict__ p, const double* __restrict__ z, int64_t n,
                                                      double* __restrict__ s, int num, int den) {
  const double beta = s[num] / s[den];
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = fma(beta, p[i], z[i]);
}

__global__ void k_set_scalar(double* s, int dst, int src) { s[dst] = s[src]; }

__global__ void k_axpby(double* __restrict__ y, double a, const double* __restrict__ x, double b, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    y[i] = fma(a, x[i], b * y[i]);
}

// y[i] -= sum_k h[k] * V[k*ld + i] for k < m (Gram-Schmidt correction).
__global__ void k_multi_axpy(double* __restrict__ y, const double* __restrict__ V, int64_t ld, const double* h, int m,
                             int64_t n, double scale) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    double acc = y[i];
    for (int k = 0; k < m; ++k) acc = fma(scale * h[k], V[int64_t(k) * ld + i], acc);
    y[i] = acc;
  }
}

// Partial dot products of w against m vectors V[k] (classical Gram-Schmidt):
// partials laid out [k][block].
__global__ void __launch_bounds__(kBlock) k_multi_dot_partial(const double* __restrict__ V, int64_t ld, int m,
                                                              const double* __restrict__ w, int64_t n,
                                                              double* __restrict__ partials, int nblocks) {
  for (int k = 0; k < m; ++k) {
    double acc = 0.0;
    const double* v = V + int64_t(k) * ld;
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
      acc = fma(v[i], w[i], acc);
    double t = block_sum(acc);
    if (threadIdx.x == 0) partials[k * nblocks + blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kBlock) k_multi_reduce(const double* __restrict__ partials, int nblocks,
                                                         double* __restrict__ out) {
  double acc = 0.0;
  const double* p = partials + blockIdx.x * nblocks;
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x) acc += p[i];
  double t = block_sum(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

void cg_update(gmf_workspace* ws, double* x, double* r, const double* p, const double* q, int64_t n, int num, int den,
               int out_slot) {
  unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_cg_update, dim3(g), dim3(kBlock), 0, ws->stream, x, r, p, q, n, ws->scalars, num, den,
                     ws->partials);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, ws->stream, ws->partials, int(g),
                     ws->scalars + out_slot);
  GMF_HIP(hipGetLastError());
}

void jacobi_dot(gmf_workspace* ws, double* z, const double* r, const double* invdiag, int64_t n, int out_slot) {
  unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_jacobi_dot, dim3(g), dim3(kBlock), 0, ws->stream, z, r, invdiag, n, ws->partials);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, ws->stream, ws->partials, int(g),
                     ws->scalars + out_slot);
  GMF_HIP(hipGetLastError());
}

void direction(gmf_workspace* ws, double* p, const double* z, int64_t n, int num, int den) {
  hipLaunchKernelGGL(k_direction, dim3(grid_for(n)), dim3(kBlock), 0, ws->stream, p, z, n, ws->scalars, num, den);
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, ws->stream, ws->scalars, den, num);
  GMF_HIP(hipGetLastError());
}

void axpby(gmf_workspace* ws, double* y, double a, const double* x, double b, int64_t n) {
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kBlock), 0, ws->stream, y, a, x, b, n);
  GMF_HIP(hipGetLastError());
}

void multi_dot(gmf_workspace* ws, const double* V, int64_t ld, int m, const double* w, int64_t n, double* out_dev) {
  unsigned g = grid_for(n);
  if (g > 256) g = 256;
  require(int64_t(m) * g <= ws->partial_capacity, GMF_EINVAL, "Krylov restart too large for the workspace");
  hipLaunchKernelGGL(k_multi_dot_partial, dim3(g), dim3(kBlock), 0, ws->stream, V, ld, m, w, n, ws->partials, int(g));
  hipLaunchKernelGGL(k_multi_reduce, dim3(unsigned(m)), dim3(kBlock), 0, ws->stream, ws->partials, int(g), out_dev);
  GMF_HIP(hipGetLastError());
}

void multi_axpy(gmf_workspace* ws, double* y, const double* V, int64_t ld, const double* h_dev, int m, int64_t n,
                double scale) {
  hipLaunchKernelGGL(k_multi_axpy, dim3(grid_for(n)), dim3(kBlock), 0, ws->stream, y, V, ld, h_dev, m, n, scale);
  GMF_HIP(hipGetLastError());
}

}  // namespace gmf
