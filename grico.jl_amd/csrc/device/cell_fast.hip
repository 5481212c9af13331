// Fast 3D tensor-product cell apply for isotropic cells with N = p+1 modes
// and N Gauss points per axis and the diffusion action form (one coefficient
// W_q c_q per point): the C2-class hot path.
//
// Same operator as the reference's tensor path (src/accumulators.jl:1955-2013,
// 24 one-dimensional passes in 3D), evaluated with 12 passes: interpolate the
// modal coefficients to the Gauss points (3 passes with B), differentiate at
// the points with the collocation matrix Dc = (B^-1 G)^T (3 passes), apply
// the per-point coefficient, and run the transposes back (3 + 3 passes). In
// exact arithmetic Dc applied to B^T u equals G^T u, so the result is the
// reference's operator; the test suite holds it to 1e-12 of the oracle. Every
// one-dimensional pass uses the even-odd split of eo.hpp.
//
// Thread layout: one cell per LPC lanes of a wavefront (LPC = N*N rounded up
// to a power of two), each active lane owning one pencil of N values in
// registers. Passes along the owned axis run in registers, the others after
// an LDS transpose. A cell never spans two wavefronts, so the transposes only
// need wavefront-level ordering (LDS operations of one wavefront complete in
// issue order) and the cell body contains no workgroup barrier.
//
// Two drivers share that cell body:
//  * brick kernel: one workgroup per brick of 2x2x2 cells that are
//    consecutive in Morton order. The brick's (2N-1)^3 lattice of modes is
//    gathered into LDS once, the eight cells read their modes from it and
//    add their results into an LDS accumulator lattice (ds_add_f64), which
//    is written back once: plain stores where this brick is the first writer
//    of the dof, plain read-modify-write otherwise. Bricks are launched one
//    color at a time (greedy DOF-disjoint coloring), so no two workgroups of
//    a launch touch the same dof and no global atomics are needed.
//  * cell kernel: independent cells with a global atomic scatter, used when
//    the cells do not form verified bricks.
// Both remap workgroup ids so that consecutive Morton-ordered work stays on
// one XCD (workgroups are dealt to the 8 XCDs round-robin).
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "eo.hpp"

namespace gmf {
namespace {

constexpr int kXcds = 8;
// Workgroups of one persistent brick launch: four per CU on 256 CUs.
constexpr int64_t kBrickGrid = 1024;

__device__ inline int64_t xcd_block(int64_t b, int64_t n) {
  const int64_t q = n / kXcds, r = n % kXcds;
  const int64_t x = b % kXcds, local = b / kXcds;
  return x * q + (x < r ? x : r) + local;
}

// Lanes per cell: N*N rounded up to a power of two (<= 64).
template <int N>
constexpr int lanes_per_cell() {
  return N * N <= 4 ? 4 : N * N <= 16 ? 16 : N * N <= 32 ? 32 : 64;
}

// Order LDS traffic of this wavefront: no compiler motion across, and the
// LDS unit completes one wavefront's operations in issue order.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-cell body. In: u = z-pencil of modes (i=t0, j=t1, k); cx, cy, cz =
// point coefficients W_q c_q of the x-pencil (q, t0, t1), the y-pencil
// (t0, q, t1) and the z-pencil (t0, t1, q); squared inverse Jacobians.
// Out: r = x-pencil of results (i, j=t0, k=t1). A/B: the cell's two N^3 LDS
// scratch arrays. Lanes with act == false hold valid (t0, t1) = (0, 0) and
// never write. Transposes rewrite an array in place only where each lane
// rewrites exactly the entries it read, so the in-order LDS of one wavefront
// makes that safe.
template <int N>
__device__ inline void cell_body(const EOTables<N>& T, const double (&u)[N], const double (&cx)[N],
                                 const double (&cy)[N], const double (&cz)[N], double jx2, double jy2, double jz2,
                                 double* A, double* B, int t0, int t1, bool act, double (&r)[N]) {
  constexpr int NN = N * N;
  double a[N], b[N];
  // z: modes -> points, A = (t0, t1, qz).
  eo_interp<N>(T, u, a);
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (act) A[t0 + N * t1 + NN * q] = a[q];
  wave_sync();
  // y: lane (i=t0, qz=t1), B = (i, qy, qz).
#pragma unroll
  for (int j = 0; j < N; ++j) a[j] = A[t0 + N * j + NN * t1];
  eo_interp<N>(T, a, b);
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (act) B[t0 + N * q + NN * t1] = b[q];
  wave_sync();
  // x: lane (qy=t0, qz=t1): point values to A; x channel f = W c J^-2 dU/dxi
  // projected back with D^T, kept in B (x layout).
#pragma unroll
  for (int i = 0; i < N; ++i) a[i] = B[i + N * t0 + NN * t1];
  eo_interp<N>(T, a, b);
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (act) A[q + N * t0 + NN * t1] = b[q];
  eo_deriv<N>(T, b, a);
#pragma unroll
  for (int q = 0; q < N; ++q) a[q] *= cx[q] * jx2;
  eo_deriv_t<N>(T, a, b);
#pragma unroll
  for (int s = 0; s < N; ++s)
    if (act) B[s + N * t0 + NN * t1] = b[s];
  wave_sync();
  // y channel on the y-pencil (qx=t0, qz=t1), added into B in place.
  {
#pragma unroll
    for (int s = 0; s < N; ++s) a[s] = A[t0 + N * s + NN * t1];
    eo_deriv<N>(T, a, b);
#pragma unroll
    for (int q = 0; q < N; ++q) b[q] *= cy[q] * jy2;
    eo_deriv_t<N>(T, b, a);
#pragma unroll
    for (int s = 0; s < N; ++s)
      if (act) B[t0 + N * s + NN * t1] += a[s];
  }
  // z channel on the z-pencil (qx=t0, qy=t1), kept in registers.
  double h[N];
  {
#pragma unroll
    for (int s = 0; s < N; ++s) a[s] = A[t0 + N * t1 + NN * s];
    eo_deriv<N>(T, a, b);
#pragma unroll
    for (int q = 0; q < N; ++q) b[q] *= cz[q] * jz2;
    eo_deriv_t<N>(T, b, h);
  }
  wave_sync();
#pragma unroll
  for (int s = 0; s < N; ++s) h[s] += B[t0 + N * t1 + NN * s];
  // Transposed z: modes k at (qx=t0, qy=t1) into A.
  eo_project<N>(T, h, a);
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (act) A[t0 + N * t1 + NN * k] = a[k];
  wave_sync();
  // Transposed y: lane (qx=t0, k=t1) -> modes j into B.
#pragma unroll
  for (int q = 0; q < N; ++q) a[q] = A[t0 + N * q + NN * t1];
  eo_project<N>(T, a, b);
#pragma unroll
  for (int j = 0; j < N; ++j)
    if (act) B[t0 + N * j + NN * t1] = b[j];
  wave_sync();
  // Transposed x: lane (j=t0, k=t1) -> modes i.
#pragma unroll
  for (int q = 0; q < N; ++q) a[q] = B[q + N * t0 + NN * t1];
  eo_project<N>(T, a, r);
}

// Lane coordinates of one cell: (t0, t1) of an active lane, (0, 0) otherwise.
template <int N>
__device__ inline void lane_coords(int lane, int& t0, int& t1, bool& act) {
  act = lane < N * N;
  const int t = act ? lane : 0;
  t0 = t % N;
  t1 = t / N;
}

// ---- per-cell driver (global atomic scatter) --------------------------------

struct CellArgs {
  int64_t cells;
  const double* jinv;   // cells x 3
  const double* coef;   // cells x N^3 (W_q c_q)
  const int32_t* idx;   // cells x N^3
  const double* lcoef;  // cells x N^3 or null
};

template <int N, bool COEF>
__global__ void __launch_bounds__(256) cell_apply_fast(CellArgs a, EOTables<N> T, const double* __restrict__ x,
                                                       double* __restrict__ y) {
  constexpr int NN = N * N, NNN = N * N * N, LPC = lanes_per_cell<N>(), CPB = 256 / LPC;
  __shared__ double S1[CPB * NNN];
  __shared__ double S2[CPB * NNN];
  const int lc = threadIdx.x / LPC;
  int t0, t1;
  bool act;
  lane_coords<N>(threadIdx.x % LPC, t0, t1, act);
  const int64_t cell = xcd_block(blockIdx.x, gridDim.x) * CPB + lc;
  const bool live = cell < a.cells;
  act = act && live;
  const int64_t c = live ? cell : 0;
  const int64_t base = c * NNN;
  double u[N], cx[N], r[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int e = t0 + N * t1 + NN * k;
    const int32_t g = a.idx[base + e];
    u[k] = g >= 0 ? (COEF ? a.lcoef[base + e] * x[g] : x[g]) : 0.0;
  }
  double cy[N], cz[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    cx[q] = a.coef[base + q + N * t0 + NN * t1];
    cy[q] = a.coef[base + t0 + N * q + NN * t1];
    cz[q] = a.coef[base + t0 + N * t1 + NN * q];
  }
  const double jx = a.jinv[c * 3], jy = a.jinv[c * 3 + 1], jz = a.jinv[c * 3 + 2];
  cell_body<N>(T, u, cx, cy, cz, jx * jx, jy * jy, jz * jz, S1 + lc * NNN, S2 + lc * NNN, t0, t1, act, r);
  if (act) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int e = i + N * t0 + NN * t1;
      const int32_t g = a.idx[base + e];
      if (g >= 0) unsafeAtomicAdd(&y[g], COEF ? a.lcoef[base + e] * r[i] : r[i]);
    }
  }
}

// ---- brick driver --------------------------------------------------------------

struct BrickArgs {
  const int64_t* order;   // bricks of this launch (one color), Morton order
  int64_t count;          // bricks in this launch
  // bricks x L lattice entries: (dof << 1) | first, -1 for no dof; `first`
  // is set where this brick is the first writer of the dof.
  const int32_t* map;
  const double* jinv;     // cells x 3
  const double* coef;     // cells x N^3
  int store_first;        // 0: always read-modify-write (y pre-zeroed)
};

// Lattice offset of local mode m along one axis of a cell: vertex 0 at 0,
// interior modes 2..N-1 at 1..N-2, vertex 1 at N-1.
template <int N>
__device__ inline int lat(int m) {
  return m == 0 ? 0 : (m == 1 ? N - 1 : m - 1);
}

// Registers of one brick in flight: its lattice entries, input values, the
// current result values of its read-modify-write points, and the x-pencil
// coefficients and Jacobian of this lane's cell.
template <int N>
struct BrickStage {
  static constexpr int M = 2 * N - 1, L = M * M * M, THREADS = 8 * lanes_per_cell<N>();
  static constexpr int PER = (L + THREADS - 1) / THREADS;
  int32_t g[PER];
  double xv[PER];
  double yold[PER];
  double cx[N];
  double jx, jy, jz;
};

template <int N>
__device__ inline void brick_map(const BrickArgs& a, int64_t brick, int tid, int32_t (&m)[BrickStage<N>::PER]) {
  using S = BrickStage<N>;
  const int32_t* map = a.map + brick * S::L;
#pragma unroll
  for (int k = 0; k < S::PER; ++k) {
    const int p = tid + k * S::THREADS;
    m[k] = p < S::L ? map[p] : -1;
  }
}

// Issue the dependent loads of one brick from its lattice entries. The y
// values of a brick are stable for the whole launch: bricks of one color are
// DOF-disjoint and earlier colors ran in earlier launches.
template <int N>
__device__ inline void brick_stage(const BrickArgs& a, int64_t brick, const int32_t (&m)[BrickStage<N>::PER], int lc,
                                   int t0, int t1, const double* __restrict__ x, const double* __restrict__ y,
                                   BrickStage<N>& s) {
  constexpr int NN = N * N, NNN = N * N * N;
  using S = BrickStage<N>;
  const int64_t cell = brick * 8 + lc;
  const int64_t base = cell * NNN;
#pragma unroll
  for (int q = 0; q < N; ++q) s.cx[q] = a.coef[base + q + N * t0 + NN * t1];
  s.jx = a.jinv[cell * 3];
  s.jy = a.jinv[cell * 3 + 1];
  s.jz = a.jinv[cell * 3 + 2];
#pragma unroll
  for (int k = 0; k < S::PER; ++k) {
    const int32_t e = m[k];
    s.g[k] = e;
    const int32_t d = e >> 1;
    s.xv[k] = e >= 0 ? x[d] : 0.0;
    s.yold[k] = (e >= 0 && !(a.store_first && (e & 1))) ? y[d] : 0.0;
  }
}

// Persistent, software-pipelined brick kernel: each workgroup walks a
// contiguous chunk of the color's bricks. While brick i computes, the loads
// of brick i+1 (inputs, results to update, coefficients) and the lattice
// entries of brick i+2 are in flight.
template <int N>
__global__ void __launch_bounds__(8 * lanes_per_cell<N>()) brick_apply(BrickArgs a, EOTables<N> T,
                                                                        const double* __restrict__ x,
                                                                        double* __restrict__ y) {
  constexpr int NN = N * N, NNN = N * N * N, LPC = lanes_per_cell<N>();
  using S = BrickStage<N>;
  constexpr int M = S::M, L = S::L, THREADS = S::THREADS, PER = S::PER;
  extern __shared__ double smem[];
  double* xl = smem;           // L input modes
  double* acc = xl + L;        // L accumulated results
  double* scratch = acc + L;   // 8 x 2 x N^3 (per cell A, B)
  const int tid = threadIdx.x;
  const int lc = tid / LPC;
  int t0, t1;
  bool act;
  lane_coords<N>(tid % LPC, t0, t1, act);
  const int64_t nblk = gridDim.x;
  const int64_t blk = xcd_block(blockIdx.x, nblk);
  const int64_t per_blk = (a.count + nblk - 1) / nblk;
  const int64_t lo = blk * per_blk;
  const int64_t hi = lo + per_blk < a.count ? lo + per_blk : a.count;
  if (lo >= hi) return;  // uniform across the workgroup
  // Cell position inside the brick (Morton child order: x fastest).
  const int ox = (lc & 1) * (N - 1), oy = ((lc >> 1) & 1) * (N - 1), oz = ((lc >> 2) & 1) * (N - 1);
  const int lx = ox + lat<N>(t0), ly = oy + lat<N>(t1);
  const int py = oy + lat<N>(t0), pz = oz + lat<N>(t1);
  double* A = scratch + lc * 2 * NNN;

  int32_t m[PER];
  S cur, nxt;
  brick_map<N>(a, a.order[lo], tid, m);
  brick_stage<N>(a, a.order[lo], m, lc, t0, t1, x, y, cur);
  if (lo + 1 < hi) brick_map<N>(a, a.order[lo + 1], tid, m);
  for (int64_t i = lo; i < hi; ++i) {
    const int64_t brick = a.order[i];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int p = tid + k * THREADS;
      if (p < L) {
        xl[p] = cur.xv[k];
        acc[p] = 0.0;
      }
    }
    __syncthreads();
    // y/z-pencil coefficients of this brick (cache hits: the x-pencil load of
    // the same lines was issued one brick earlier), then the prefetches.
    double cy[N], cz[N];
    {
      const double* cb = a.coef + (brick * 8 + lc) * NNN;
#pragma unroll
      for (int q = 0; q < N; ++q) {
        cy[q] = cb[t0 + N * q + NN * t1];
        cz[q] = cb[t0 + N * t1 + NN * q];
      }
    }
    if (i + 1 < hi) brick_stage<N>(a, a.order[i + 1], m, lc, t0, t1, x, y, nxt);
    if (i + 2 < hi) brick_map<N>(a, a.order[i + 2], tid, m);
    double u[N], r[N];
#pragma unroll
    for (int k = 0; k < N; ++k) u[k] = xl[lx + M * (ly + M * (oz + lat<N>(k)))];
    cell_body<N>(T, u, cur.cx, cy, cz, cur.jx * cur.jx, cur.jy * cur.jy, cur.jz * cur.jz, A, A + NNN, t0, t1, act,
                 r);
    if (act) {
#pragma unroll
      for (int ii = 0; ii < N; ++ii) atomicAdd(&acc[ox + lat<N>(ii) + M * (py + M * pz)], r[ii]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int p = tid + k * THREADS;
      if (cur.g[k] >= 0) y[cur.g[k] >> 1] = cur.yold[k] + acc[p];
    }
    __syncthreads();
    cur = nxt;
  }
}

template <int N>
EOTables<N> tables_of(const Bucket& b) {
  return make_eo_tables<N>(b.h_values.data(), b.h_colloc.data());
}

template <int N>
void launch_cells(const Bucket& b, const double* x, double* y, bool reduced, hipStream_t stream) {
  constexpr int CPB = 256 / lanes_per_cell<N>();
  EOTables<N> T = tables_of<N>(b);
  CellArgs a;
  a.cells = b.cells;
  a.jinv = b.jinv;
  a.coef = b.pdata;
  a.idx = reduced ? b.ridx : b.lidx;
  a.lcoef = b.lcoef;
  const int64_t blocks = (b.cells + CPB - 1) / CPB;
  if (b.lcoef)
    hipLaunchKernelGGL((cell_apply_fast<N, true>), dim3(unsigned(blocks)), dim3(256), 0, stream, a, T, x, y);
  else
    hipLaunchKernelGGL((cell_apply_fast<N, false>), dim3(unsigned(blocks)), dim3(256), 0, stream, a, T, x, y);
  GMF_HIP(hipGetLastError());
}

template <int N>
size_t brick_lds_bytes() {
  constexpr int M = 2 * N - 1;
  return sizeof(double) * size_t(2 * M * M * M + 8 * 2 * N * N * N);
}

template <int N>
void launch_bricks(const Bucket& b, const double* x, double* y, bool reduced, bool store_first, hipStream_t stream) {
  EOTables<N> T = tables_of<N>(b);
  const BrickPath& bp = b.bricks;
  for (int c = 0; c < bp.colors; ++c) {
    const int64_t lo = bp.color_offsets[size_t(c)], hi = bp.color_offsets[size_t(c) + 1];
    if (hi == lo) continue;
    BrickArgs a;
    a.order = bp.order + lo;
    a.count = hi - lo;
    a.map = reduced ? bp.map_red : bp.map_full;
    a.jinv = b.jinv;
    a.coef = b.pdata;
    a.store_first = store_first ? 1 : 0;
    const int64_t grid = std::min<int64_t>(hi - lo, kBrickGrid);
    hipLaunchKernelGGL((brick_apply<N>), dim3(unsigned(grid)), dim3(8 * lanes_per_cell<N>()), brick_lds_bytes<N>(),
                       stream, a, T, x, y);
  }
  GMF_HIP(hipGetLastError());
}

// Solve B X = G (N x N, partial pivoting) and return Dc[q*N+s] = X[s][q].
std::vector<double> collocation_derivative(const std::vector<double>& B, const std::vector<double>& G, int n) {
  std::vector<double> A(B), X(G);
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[size_t(i * n + k)]) > std::fabs(A[size_t(piv * n + k)])) piv = i;
    require(A[size_t(piv * n + k)] != 0.0, GMF_EINVAL, "singular basis collocation matrix");
    if (piv != k)
      for (int j = 0; j < n; ++j) {
        std::swap(A[size_t(k * n + j)], A[size_t(piv * n + j)]);
        std::swap(X[size_t(k * n + j)], X[size_t(piv * n + j)]);
      }
    for (int i = k + 1; i < n; ++i) {
      double f = A[size_t(i * n + k)] / A[size_t(k * n + k)];
      for (int j = k; j < n; ++j) A[size_t(i * n + j)] -= f * A[size_t(k * n + j)];
      for (int j = 0; j < n; ++j) X[size_t(i * n + j)] -= f * X[size_t(k * n + j)];
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = n - 1; i >= 0; --i) {
      double v = X[size_t(i * n + j)];
      for (int k = i + 1; k < n; ++k) v -= A[size_t(i * n + k)] * X[size_t(k * n + j)];
      X[size_t(i * n + j)] = v / A[size_t(i * n + i)];
    }
  std::vector<double> D(static_cast<size_t>(n * n));
  for (int q = 0; q < n; ++q)
    for (int s = 0; s < n; ++s) D[size_t(q * n + s)] = X[size_t(s * n + q)];
  return D;
}

// Build and verify the 2x2x2 brick decomposition of a bucket: every group of
// eight consecutive cells must place its modes on a common (2N-1)^3 lattice
// consistently (cells sharing a lattice point agree on its dof) and every
// dof may appear at most once per brick. Returns false when the cells do not
// form such bricks.
bool build_bricks(gmf_plan& p, Bucket& b, const int32_t* lidx, const std::vector<int64_t>& red_of_full) {
  const int N = b.N[0];
  const int M = 2 * N - 1, L = M * M * M;
  const int NNN = N * N * N;
  if (b.cells % 8 != 0 || b.cells == 0) return false;
  const int64_t nb = b.cells / 8;
  std::vector<int32_t> map(size_t(nb) * size_t(L));
  auto lat = [N](int m) { return m == 0 ? 0 : (m == 1 ? N - 1 : m - 1); };
  std::vector<int32_t> lattice(static_cast<size_t>(L));
  std::vector<int64_t> seen_brick(size_t(p.n_full), -1);
  for (int64_t br = 0; br < nb; ++br) {
    std::fill(lattice.begin(), lattice.end(), -2);
    for (int c = 0; c < 8; ++c) {
      const int ox = (c & 1) * (N - 1), oy = ((c >> 1) & 1) * (N - 1), oz = ((c >> 2) & 1) * (N - 1);
      const int32_t* li = lidx + (br * 8 + c) * NNN;
      for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
          for (int i = 0; i < N; ++i) {
            const int32_t g = li[i + N * j + N * N * k];
            if (g < -1) return false;
            const int pos = (ox + lat(i)) + M * ((oy + lat(j)) + M * (oz + lat(k)));
            if (lattice[size_t(pos)] == -2)
              lattice[size_t(pos)] = g;
            else if (lattice[size_t(pos)] != g)
              return false;
          }
    }
    for (int pos = 0; pos < L; ++pos) {
      const int32_t g = lattice[size_t(pos)];
      if (g == -2) return false;
      if (g >= 0) {
        if (seen_brick[size_t(g)] == br) return false;  // dof twice in one brick
        seen_brick[size_t(g)] = br;
      }
      map[size_t(br * L + pos)] = g;
    }
  }
  // Greedy DOF-disjoint coloring of bricks in Morton order.
  std::vector<uint64_t> used(size_t(p.n_full), 0);
  std::vector<int> color(static_cast<size_t>(nb));
  int ncolors = 0;
  for (int64_t br = 0; br < nb; ++br) {
    uint64_t mask = 0;
    for (int pos = 0; pos < L; ++pos) {
      const int32_t g = map[size_t(br * L + pos)];
      if (g >= 0) mask |= used[size_t(g)];
    }
    int c = 0;
    while (c < 64 && ((mask >> c) & 1u)) ++c;
    if (c == 64) return false;
    color[size_t(br)] = c;
    ncolors = std::max(ncolors, c + 1);
    for (int pos = 0; pos < L; ++pos) {
      const int32_t g = map[size_t(br * L + pos)];
      if (g >= 0) used[size_t(g)] |= uint64_t(1) << c;
    }
  }
  // First writer of each dof: the lowest color touching it.
  std::vector<uint8_t> first(size_t(nb) * size_t(L), 0);
  for (int64_t br = 0; br < nb; ++br)
    for (int pos = 0; pos < L; ++pos) {
      const int32_t g = map[size_t(br * L + pos)];
      if (g < 0) continue;
      const uint64_t m = used[size_t(g)];
      int lowest = 0;
      while (!((m >> lowest) & 1u)) ++lowest;
      first[size_t(br * L + pos)] = uint8_t(color[size_t(br)] == lowest);
    }
  BrickPath& bp = b.bricks;
  bp.colors = ncolors;
  bp.color_offsets.assign(size_t(ncolors) + 1, 0);
  for (int64_t br = 0; br < nb; ++br) bp.color_offsets[size_t(color[size_t(br)]) + 1]++;
  for (int c = 0; c < ncolors; ++c) bp.color_offsets[size_t(c) + 1] += bp.color_offsets[size_t(c)];
  std::vector<int64_t> order(static_cast<size_t>(nb));
  std::vector<int64_t> fillp(bp.color_offsets.begin(), bp.color_offsets.end() - 1);
  for (int64_t br = 0; br < nb; ++br) order[size_t(fillp[size_t(color[size_t(br)])]++)] = br;
  bp.bricks = nb;
  bp.order = upload_array(p, order.data(), order.size());
  // Lattice entries carry the first-writer flag in bit 0.
  require(p.n_full < (int64_t(1) << 30), GMF_EUNSUPPORTED, "brick maps need dof counts below 2^30");
  std::vector<int32_t> enc(map.size());
  for (size_t i = 0; i < map.size(); ++i) enc[i] = map[i] >= 0 ? (map[i] << 1) | int32_t(first[i]) : -1;
  bp.map_full = upload_array(p, enc.data(), enc.size());
  if (!red_of_full.empty()) {
    for (size_t i = 0; i < map.size(); ++i)
      enc[i] = map[i] >= 0 && red_of_full[size_t(map[i])] >= 0 ? (int32_t(red_of_full[size_t(map[i])]) << 1) | int32_t(first[i])
                                                               : -1;
    bp.map_red = upload_array(p, enc.data(), enc.size());
  }
  bp.enabled = true;
  return true;
}

}  // namespace

void prepare_fast_path(gmf_plan& p, Bucket& b, const int32_t* lidx, const std::vector<int64_t>& red_of_full) {
  b.fast = false;
  if (b.dim != 3 || b.form != GMF_FORM_DIFFUSION || b.multi_count != 0) return;
  const int n = b.N[0];
  if (n < 2 || n > 8) return;
  for (int a = 0; a < 3; ++a)
    if (b.N[a] != n || b.Q[a] != n) return;
  // The three axes must share one table (isotropic degree and rule).
  for (int a = 1; a < 3; ++a)
    for (int i = 0; i < n * n; ++i)
      if (b.h_values[size_t(b.table_offset[a] + i)] != b.h_values[size_t(i)] ||
          b.h_derivs[size_t(b.table_offset[a] + i)] != b.h_derivs[size_t(i)])
        return;
  std::vector<double> B(b.h_values.begin(), b.h_values.begin() + n * n);
  std::vector<double> G(b.h_derivs.begin(), b.h_derivs.begin() + n * n);
  b.h_colloc = collocation_derivative(B, G, n);
  b.fast = true;
  // Small batches (coarse multigrid levels) take the one-launch per-cell
  // kernel: eight colored brick launches of a few bricks each cost more than
  // the atomics they avoid. GMF_BRICK_MIN_CELLS overrides the threshold.
  int64_t min_cells = 32768;
  if (const char* env = std::getenv("GMF_BRICK_MIN_CELLS")) min_cells = std::strtoll(env, nullptr, 10);
  if (b.lcoef == nullptr && n <= 6 && b.cells >= min_cells) build_bricks(p, b, lidx, red_of_full);
}

bool launch_cell_fast(const gmf_plan& p, const Bucket& b, const double* x, double* y, bool reduced,
                      hipStream_t stream) {
  if (!b.fast || b.cells == 0) return false;
  if (reduced && !b.ridx) return false;
  if (b.bricks.enabled && (!reduced || b.bricks.map_red)) {
    const bool store_first = p.buckets.size() == 1;
    switch (b.N[0]) {
      case 2: launch_bricks<2>(b, x, y, reduced, store_first, stream); return true;
      case 3: launch_bricks<3>(b, x, y, reduced, store_first, stream); return true;
      case 4: launch_bricks<4>(b, x, y, reduced, store_first, stream); return true;
      case 5: launch_bricks<5>(b, x, y, reduced, store_first, stream); return true;
      case 6: launch_bricks<6>(b, x, y, reduced, store_first, stream); return true;
      default: break;
    }
  }
  switch (b.N[0]) {
    case 2: launch_cells<2>(b, x, y, reduced, stream); return true;
    case 3: launch_cells<3>(b, x, y, reduced, stream); return true;
    case 4: launch_cells<4>(b, x, y, reduced, stream); return true;
    case 5: launch_cells<5>(b, x, y, reduced, stream); return true;
    case 6: launch_cells<6>(b, x, y, reduced, stream); return true;
    case 7: launch_cells<7>(b, x, y, reduced, stream); return true;
    case 8: launch_cells<8>(b, x, y, reduced, stream); return true;
    default: return false;
  }
}

bool plan_writes_every_dof(const gmf_plan& p) {
  // The brick store-first path writes every dof of its lattice; with a single
  // brick bucket covering the plan no zero fill of the result is needed.
  return p.buckets.size() == 1 && p.buckets[0].bricks.enabled;
}

}  // namespace gmf
