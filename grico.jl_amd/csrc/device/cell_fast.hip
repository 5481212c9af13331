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
//    consecutive in Morton order, all bricks in one launch. The brick's
//    (2N-1)^3 lattice of modes is gathered into LDS once, the eight cells
//    read their modes from it and add their results into an LDS accumulator
//    lattice (ds_add_f64), which is written back once: dofs that no other
//    brick holds with plain stores, every occurrence of a dof shared between
//    bricks into its own halo slot, and k_brick_halo sums each shared dof's
//    slots (contiguous, in brick order) into y without global atomics. One
//    launch keeps neighbouring bricks' shared input lines in L2, which a
//    colored sequence of launches re-fetches once per color (an LRU model of
//    the 4 MB L2 per XCD on the 64^3 p=4 mesh: 1.05x instead of 2.17x the
//    input's cache lines fetched per pass).
//  * cell kernel: independent cells with a global atomic scatter, used when
//    the cells do not form verified bricks or the batch is small.
// Both remap workgroup ids so that consecutive Morton-ordered work stays on
// one XCD (workgroups are dealt to the 8 XCDs round-robin).
#include <algorithm>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "eo.hpp"

namespace gmf {
namespace {

constexpr int kXcds = 8;

__device__ inline int64_t xcd_block(int64_t b, int64_t n) {
  const int64_t q = n / kXcds, r = n % kXcds;
  const int64_t x = b % kXcds, local = b / kXcds;
  return x * q + (x < r ? x : r) + local;
}

// Lanes per cell: N*N rounded up to a power of two (<= 64).
template <int N>
constexpr int lanes_per_cell() {
  return N * N <= 4 ? 4 : N * N <= 16 ? 16 : N * N <= 32 ? 32 : N * N <= 64 ? 64 : 128;
}

// Order LDS traffic of this wavefront: no compiler motion across, and the
// LDS unit completes one wavefront's operations in issue order.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ordering of a cell's LDS transposes: wavefront order when the cell lives in
// one wavefront, a workgroup barrier when it spans two (N = 9: 81 pencils;
// every lane of the workgroup then reaches each barrier).
template <int N>
__device__ inline void cell_sync() {
  if constexpr (lanes_per_cell<N>() > 64)
    __syncthreads();
  else
    wave_sync();
}

// Per-cell body. In: u = z-pencil of modes (i=t0, j=t1, k); cx, cy, cz =
// point coefficients W_q c_q of the x-pencil (q, t0, t1), the y-pencil
// (t0, q, t1) and the z-pencil (t0, t1, q); squared inverse Jacobians.
// Out: r = x-pencil of results (i, j=t0, k=t1). A/B: the cell's two N^3 LDS
// scratch arrays. Lanes with act == false hold valid (t0, t1) = (0, 0) and
// never write. Transposes rewrite an array in place only where each lane
// rewrites exactly the entries it read, so the in-order LDS of one wavefront
// makes that safe.
template <int N, class R>
__device__ inline void cell_body(const EOTables<N, R>& T, const R (&u)[N], const R (&cx)[N], const R (&cy)[N],
                                 const R (&cz)[N], R jx2, R jy2, R jz2, R* A, R* B, int t0, int t1, bool act,
                                 R (&r)[N]) {
  constexpr int NN = N * N;
  R a[N], b[N];
  // z: modes -> points, A = (t0, t1, qz).
  eo_interp<N>(T, u, a);
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (act) A[t0 + N * t1 + NN * q] = a[q];
  cell_sync<N>();
  // y: lane (i=t0, qz=t1), B = (i, qy, qz).
#pragma unroll
  for (int j = 0; j < N; ++j) a[j] = A[t0 + N * j + NN * t1];
  eo_interp<N>(T, a, b);
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (act) B[t0 + N * q + NN * t1] = b[q];
  cell_sync<N>();
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
  cell_sync<N>();
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
  R h[N];
  {
#pragma unroll
    for (int s = 0; s < N; ++s) a[s] = A[t0 + N * t1 + NN * s];
    eo_deriv<N>(T, a, b);
#pragma unroll
    for (int q = 0; q < N; ++q) b[q] *= cz[q] * jz2;
    eo_deriv_t<N>(T, b, h);
  }
  cell_sync<N>();
#pragma unroll
  for (int s = 0; s < N; ++s) h[s] += B[t0 + N * t1 + NN * s];
  // Transposed z: modes k at (qx=t0, qy=t1) into A.
  eo_project<N>(T, h, a);
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (act) A[t0 + N * t1 + NN * k] = a[k];
  cell_sync<N>();
  // Transposed y: lane (qx=t0, k=t1) -> modes j into B.
#pragma unroll
  for (int q = 0; q < N; ++q) a[q] = A[t0 + N * q + NN * t1];
  eo_project<N>(T, a, b);
#pragma unroll
  for (int j = 0; j < N; ++j)
    if (act) B[t0 + N * j + NN * t1] = b[j];
  cell_sync<N>();
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

template <class R>
struct CellArgs {
  int64_t cells;
  const R* jinv;        // cells x 3
  const R* coef;        // cells x N^3 (W_q c_q)
  const int32_t* idx;   // cells x N^3
  const R* lcoef;       // cells x N^3 or null
};

// THREADS (64, 128 or 256 lanes per workgroup, at least one cell) is picked
// per bucket by the plan-time tuner.
template <int N, bool COEF, class R, int THREADS>
__global__ void __launch_bounds__(THREADS) cell_apply_fast(CellArgs<R> a, EOTables<N, R> T, const R* __restrict__ x,
                                                           R* __restrict__ y) {
  constexpr int NN = N * N, NNN = N * N * N, LPC = lanes_per_cell<N>(), CPB = THREADS / LPC;
  static_assert(CPB >= 1, "a workgroup holds at least one cell");
  __shared__ R S1[CPB * NNN];
  __shared__ R S2[CPB * NNN];
  const int lc = threadIdx.x / LPC;
  int t0, t1;
  bool act;
  lane_coords<N>(threadIdx.x % LPC, t0, t1, act);
  const int64_t cell = xcd_block(blockIdx.x, gridDim.x) * CPB + lc;
  const bool live = cell < a.cells;
  act = act && live;
  const int64_t c = live ? cell : 0;
  const int64_t base = c * NNN;
  R u[N], cx[N], r[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int e = t0 + N * t1 + NN * k;
    const int32_t g = a.idx[base + e];
    u[k] = g >= 0 ? (COEF ? a.lcoef[base + e] * x[g] : x[g]) : R(0);
  }
  R cy[N], cz[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    cx[q] = a.coef[base + q + N * t0 + NN * t1];
    cy[q] = a.coef[base + t0 + N * q + NN * t1];
    cz[q] = a.coef[base + t0 + N * t1 + NN * q];
  }
  const R jx = a.jinv[c * 3], jy = a.jinv[c * 3 + 1], jz = a.jinv[c * 3 + 2];
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

// ---- verification sums on fast buckets ------------------------------------------

// Verification sums of one cell (src/verification.jl:201-230): partial[2 c] =
// sum_q w_q (u_h(x_q) - e_q)^2, partial[2 c + 1] = sum_q w_q e_q^2, u_h
// interpolated from the full-layout coefficients (3 passes); the cell's lanes
// combine their sums by a fixed butterfly (deterministic).
template <int N, bool COEF>
__global__ void __launch_bounds__(256) cell_l2_fast(CellArgs<double> a, EOTables<N, double> T,
                                                    const double* __restrict__ u, const double* __restrict__ w,
                                                    const double* __restrict__ ex, double* __restrict__ partial) {
  constexpr int NN = N * N, NNN = N * N * N, LPC = lanes_per_cell<N>(), CPB = 256 / LPC;
  __shared__ double S1[CPB * NNN];
  __shared__ double S2[CPB * NNN];
  const int lc = threadIdx.x / LPC, lane_in_cell = threadIdx.x % LPC;
  int t0, t1;
  bool act;
  lane_coords<N>(lane_in_cell, t0, t1, act);
  const int64_t cell = xcd_block(blockIdx.x, gridDim.x) * CPB + lc;
  const bool live = cell < a.cells;
  act = act && live;
  const int64_t c = live ? cell : 0;
  const int64_t base = c * NNN;
  double* A = S1 + lc * NNN;
  double* B = S2 + lc * NNN;
  double v[N], t[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int e = t0 + N * t1 + NN * k;
    const int32_t gd = a.idx[base + e];
    v[k] = gd >= 0 ? (COEF ? a.lcoef[base + e] * u[gd] : u[gd]) : 0.0;
  }
  eo_interp<N>(T, v, t);
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (act) A[t0 + N * t1 + NN * q] = t[q];
  cell_sync<N>();
#pragma unroll
  for (int j = 0; j < N; ++j) v[j] = A[t0 + N * j + NN * t1];
  eo_interp<N>(T, v, t);
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (act) B[t0 + N * q + NN * t1] = t[q];
  cell_sync<N>();
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = B[i + N * t0 + NN * t1];
  eo_interp<N>(T, v, t);
  double s0 = 0.0, s1 = 0.0;
  if (act) {
#pragma unroll
    for (int q = 0; q < N; ++q) {
      const int64_t pq = base + q + N * t0 + NN * t1;
      const double ev = ex[pq], wq = w[pq];
      const double d = t[q] - ev;
      s0 = fma(wq * d, d, s0);
      s1 = fma(wq * ev, ev, s1);
    }
  }
  if constexpr (LPC <= 64) {
#pragma unroll
    for (int off = LPC / 2; off > 0; off >>= 1) {
      s0 += __shfl_xor(s0, off, LPC);
      s1 += __shfl_xor(s1, off, LPC);
    }
    if (live && lane_in_cell == 0) {
      partial[2 * c] = s0;
      partial[2 * c + 1] = s1;
    }
  } else {
    // cell over two wavefronts: butterfly in each, then the two halves in order
    __shared__ double half[CPB][2][2];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      s0 += __shfl_xor(s0, off, 64);
      s1 += __shfl_xor(s1, off, 64);
    }
    const int hw = lane_in_cell / 64;
    if (lane_in_cell % 64 == 0) {
      half[lc][hw][0] = s0;
      half[lc][hw][1] = s1;
    }
    __syncthreads();
    if (live && lane_in_cell == 0) {
      partial[2 * c] = half[lc][0][0] + half[lc][1][0];
      partial[2 * c + 1] = half[lc][0][1] + half[lc][1][1];
    }
  }
}

// ---- brick driver --------------------------------------------------------------

template <class R>
struct BrickArgs {
  // bricks x L lattice entries: (dof << 1) | exclusive; -1 for no dof, -2 for
  // no dof at a shared position (a constrained dof of the reduced layout),
  // which still takes its halo slot.
  const int32_t* map;
  // Shared positions of brick b, in ascending lattice position, write halo
  // slots halo_slot[halo_base[b] .. halo_base[b + 1]).
  const int32_t* halo_base;
  const int32_t* halo_slot;
  const R* jinv;  // cells x 3
  const R* coef;  // cells x N^3
  R* halo;        // halo values of the bucket (shared-dof major)
  // bricks x 8 cells, -1 for an empty position (zero coefficients: the
  // position adds nothing); null: brick b holds cells 8 b .. 8 b + 7.
  const int32_t* cell_of;
};

// Result stores of the plain apply: y[i] (+)= v.
template <class R>
struct StoreResult {
  R* y;
  int accumulate;  // 1: results are added to y (other batches wrote it first)
  __device__ void operator()(int64_t i, R v, R) const { y[i] = accumulate ? y[i] + v : v; }
};

// Lattice offset of local mode m along one axis of a cell: vertex 0 at 0,
// interior modes 2..N-1 at 1..N-2, vertex 1 at N-1.
template <int N>
__host__ __device__ inline int lat(int m) {
  return m == 0 ? 0 : (m == 1 ? N - 1 : m - 1);
}

// LDS placement of the brick lattice: position (x, y, z) at x + PY y + PZ z.
// Row and plane strides may exceed 2N - 1 (padding) so that the lanes of one
// wavefront, which gather rows of different cells and planes, spread over more
// LDS banks.
#ifndef GMF_BRICK_PAD_Y
#define GMF_BRICK_PAD_Y 0
#endif
#ifndef GMF_BRICK_PAD_Z
#define GMF_BRICK_PAD_Z 0
#endif
template <int N>
struct BrickLds {
  static constexpr int M = 2 * N - 1;
  static constexpr int PY = M + GMF_BRICK_PAD_Y;
  static constexpr int PZ = PY * M + GMF_BRICK_PAD_Z;
  static constexpr int SIZE = PZ * (M - 1) + PY * (M - 1) + M;  // elements spanned
  __device__ static inline int at(int x, int y, int z) { return x + PY * y + PZ * z; }
  // LDS offset of flattened lattice position p = x + M (y + M z).
  __device__ static inline int of(int p) {
    if (PY == M && PZ == M * M) return p;
    const int x = p % M, yz = p / M;
    return x + PY * (yz % M) + PZ * (yz / M);
  }
};

// One workgroup per brick, all bricks in one launch, consecutive bricks on one
// XCD so that neighbours share its L2. Phase 1 issues the lattice entries, the
// cells' coefficients (each cell's N^3 values loaded contiguously by its lanes
// and staged in the cells' A scratch) and the input gathers (staged over the
// cells' B scratch); phase 2 reads each lane's pencils, then runs the eight
// cell bodies, which add into an LDS result lattice; phase 3 stores the
// positions whose dof no other brick holds to y, and the shared positions
// (ranked in ascending lattice position by wavefront ballots) to their halo
// slots, which k_brick_halo sums per dof. The exclusive stores go
// through `store` (plain result or a fused smoother step, which may update x in
// place: x is not declared restrict).
template <int N, class R, class Store>
__global__ void __launch_bounds__(8 * lanes_per_cell<N>()) brick_apply(BrickArgs<R> a, EOTables<N, R> T,
                                                                        const R* x, Store store) {
  constexpr int NN = N * N, NNN = N * N * N, LPC = lanes_per_cell<N>();
  constexpr int M = 2 * N - 1, L = M * M * M, THREADS = 8 * LPC, PER = (L + THREADS - 1) / THREADS;
  constexpr int WAVES = (THREADS + 63) / 64, CPL = (NNN + LPC - 1) / LPC;
  extern __shared__ double smem_words[];
  __shared__ int shared_count[PER * WAVES];
  using LD = BrickLds<N>;
  static_assert(LD::SIZE <= 8 * NNN, "brick input lattice must fit the cells' B scratch");
  R* acc = reinterpret_cast<R*>(smem_words);  // lattice of accumulated results (LD layout)
  R* scratch = acc + LD::SIZE;                // cells' A (8 x N^3), then cells' B (8 x N^3)
  R* xl = scratch + 8 * NNN;                  // lattice of input modes over the B scratch, until phase 2 starts
  const int tid = threadIdx.x;
  const int lc = tid / LPC, lane_in_cell = tid % LPC;
  const int wave = tid >> 6, lane = tid & 63;
  int t0, t1;
  bool act;
  lane_coords<N>(lane_in_cell, t0, t1, act);
  const int64_t brick = xcd_block(blockIdx.x, gridDim.x);
  const int64_t slot_cell = a.cell_of ? int64_t(a.cell_of[brick * 8 + lc]) : brick * 8 + lc;
  const bool present = slot_cell >= 0;
  const int64_t cell = present ? slot_cell : 0;
  // Phase 1: entries, coefficients, slot list, inputs.
  const int32_t* map = a.map + brick * L;
  int32_t e[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int p = tid + k * THREADS;
    e[k] = p < L ? map[p] : -1;
  }
  const R* cb = a.coef + cell * NNN;
  R cv[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int i = lane_in_cell + k * LPC;
    cv[k] = i < NNN && present ? cb[i] : R(0);
  }
  const int32_t* brick_slots = a.halo_slot + a.halo_base[brick];
  const R jx = a.jinv[cell * 3], jy = a.jinv[cell * 3 + 1], jz = a.jinv[cell * 3 + 2];
  R xv[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int32_t d = e[k] >> 1;
    xv[k] = d >= 0 ? x[d] : R(0);
  }
  R* A = scratch + lc * NNN;  // coefficients, until the cell body first writes A
  R* B = scratch + 8 * NNN + lc * NNN;
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int i = lane_in_cell + k * LPC;
    if (i < NNN) A[i] = cv[k];
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int p = tid + k * THREADS;
    if (p < L) {
      const int o = LD::of(p);
      xl[o] = xv[k];
      acc[o] = R(0);
    }
  }
  __syncthreads();
  // Phase 2: cell bodies. Every lane reads its coefficient pencils (A) and
  // its input pencil (lattice over B) before any cell body writes A or B.
  R cx[N], cy[N], cz[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    cx[q] = A[q + N * t0 + NN * t1];
    cy[q] = A[t0 + N * q + NN * t1];
    cz[q] = A[t0 + N * t1 + NN * q];
  }
  const int ox = (lc & 1) * (N - 1), oy = ((lc >> 1) & 1) * (N - 1), oz = ((lc >> 2) & 1) * (N - 1);
  R u[N], r[N];
  {
    const int lx = ox + lat<N>(t0), ly = oy + lat<N>(t1);
#pragma unroll
    for (int k = 0; k < N; ++k) u[k] = xl[LD::at(lx, ly, oz + lat<N>(k))];
  }
  __syncthreads();
  cell_body<N>(T, u, cx, cy, cz, jx * jx, jy * jy, jz * jz, A, B, t0, t1, act, r);
  if (act) {
    const int py = oy + lat<N>(t0), pz = oz + lat<N>(t1);
#pragma unroll
    for (int ii = 0; ii < N; ++ii) atomicAdd(&acc[LD::at(ox + lat<N>(ii), py, pz)], r[ii]);
  }
  // Ranks of the shared positions, counted per (round, wavefront).
  uint64_t fm[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    fm[k] = __ballot(!(e[k] & 1));
    if (lane == 0) shared_count[k * WAVES + wave] = __popcll(fm[k]);
  }
  __syncthreads();
  // Phase 3.
  const uint64_t below = (uint64_t(1) << lane) - 1;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int p = tid + k * THREADS;
    const int o = LD::of(p < L ? p : 0);
    if (e[k] & 1) {
      const int32_t d = e[k] >> 1;
      if (d >= 0) store(d, acc[o], xv[k]);
    } else {
      int off = 0;
      for (int j = 0; j < k * WAVES + wave; ++j) off += shared_count[j];
      a.halo[brick_slots[off + __popcll(fm[k] & below)]] = acc[o];
    }
  }
}

// Shared dofs: store(d, sum of the dof's contiguous halo values) (the owner's
// first, then the other bricks in order: deterministic). x is the apply input
// (read only when the store needs it). Consecutive blocks stay on one XCD.
template <class R, class Store>
__global__ void k_brick_halo(const R* x, const R* __restrict__ halo, const int32_t* __restrict__ dof,
                             const int32_t* __restrict__ ptr, int64_t count, Store store, int needs_input) {
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * int64_t(blockDim.x) + threadIdx.x;
  if (i >= count) return;
  const int32_t d = dof[i];
  const int32_t lo = ptr[i], hi = ptr[i + 1];
  R s = R(0);
  for (int32_t j = lo; j < hi; ++j) s += halo[j];
  if (d >= 0) store(d, s, needs_input ? x[d] : R(0));
}

template <int N, class R>
EOTables<N, R> tables_of(const Bucket& b) {
  return make_eo_tables<N, R>(b.h_values.data(), b.h_colloc.data());
}

// Per-cell arrays of a bucket in arithmetic type R (FP64 originals or their
// FP32 copies made by prepare_fp32).
struct CellArrays64 {
  static const double* jinv(const Bucket& b) { return b.jinv; }
  static const double* coef(const Bucket& b) { return b.pdata; }
  static const double* lcoef(const Bucket& b) { return b.lcoef; }
};
struct CellArrays32 {
  static const float* jinv(const Bucket& b) { return b.jinv32; }
  static const float* coef(const Bucket& b) { return b.pdata32; }
  static const float* lcoef(const Bucket& b) { return b.lcoef32; }
};
template <class R>
using CellArrays = typename std::conditional<std::is_same<R, float>::value, CellArrays32, CellArrays64>::type;

template <int N, class R, int THREADS>
void launch_cells_shape(const Bucket& b, const R* x, R* y, bool reduced, hipStream_t stream) {
  constexpr int CPB = THREADS / lanes_per_cell<N>();
  EOTables<N, R> T = tables_of<N, R>(b);
  CellArgs<R> a;
  a.cells = b.cells;
  a.jinv = CellArrays<R>::jinv(b);
  a.coef = CellArrays<R>::coef(b);
  a.idx = reduced ? b.ridx : b.lidx;
  a.lcoef = CellArrays<R>::lcoef(b);
  const int64_t blocks = (b.cells + CPB - 1) / CPB;
  if (b.lcoef)
    hipLaunchKernelGGL((cell_apply_fast<N, true, R, THREADS>), dim3(unsigned(blocks)), dim3(THREADS), 0, stream, a,
                       T, x, y);
  else
    hipLaunchKernelGGL((cell_apply_fast<N, false, R, THREADS>), dim3(unsigned(blocks)), dim3(THREADS), 0, stream, a,
                       T, x, y);
  GMF_HIP(hipGetLastError());
}

// Per-cell launch with `threads` lanes per workgroup (64, 128, 256; raised to
// the lanes of one cell where a cell needs more).
template <int N, class R>
void launch_cells(const Bucket& b, const R* x, R* y, bool reduced, int threads, hipStream_t stream) {
  constexpr int LPC = lanes_per_cell<N>();
  if constexpr (LPC <= 64) {
    if (threads <= 64) {
      launch_cells_shape<N, R, 64>(b, x, y, reduced, stream);
      return;
    }
  }
  if (threads <= 128)
    launch_cells_shape<N, R, 128>(b, x, y, reduced, stream);
  else
    launch_cells_shape<N, R, 256>(b, x, y, reduced, stream);
}

template <int N, class R>
size_t brick_lds_bytes() {
  return sizeof(R) * size_t(BrickLds<N>::SIZE + 8 * 2 * N * N * N);
}

template <int N, class R, class Store>
void launch_bricks(const Bucket& b, const R* x, bool reduced, R* slots, hipStream_t stream, Store store,
                   bool needs_input) {
  EOTables<N, R> T = tables_of<N, R>(b);
  const BrickPath& bp = b.bricks;
  require(bp.halo_slots == 0 || slots != nullptr, GMF_EINVAL, "workspace has no brick halo slots");
  BrickArgs<R> a;
  a.map = reduced ? bp.map_red : bp.map_full;
  a.halo_base = bp.halo_base;
  a.halo_slot = bp.halo_slot;
  a.jinv = CellArrays<R>::jinv(b);
  a.coef = CellArrays<R>::coef(b);
  a.halo = slots ? slots + bp.slot_base : nullptr;
  a.cell_of = bp.cell_of;
  hipLaunchKernelGGL((brick_apply<N, R, Store>), dim3(unsigned(bp.bricks)), dim3(8 * lanes_per_cell<N>()),
                     (brick_lds_bytes<N, R>()), stream, a, T, x, store);
  if (bp.shared_count > 0)
    hipLaunchKernelGGL((k_brick_halo<R, Store>), dim3(unsigned((bp.shared_count + 255) / 256)), dim3(256), 0,
                       stream, x, a.halo, reduced ? bp.halo_red : bp.halo_full, bp.halo_ptr, bp.shared_count, store,
                       needs_input ? 1 : 0);
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
// eight consecutive cells (or the cells brick_cells assigns to a brick, with
// empty positions) must place its modes on a common (2N-1)^3 lattice
// consistently (cells sharing a lattice point agree on its dof) and every
// dof may appear at most once per brick. Lattice points no cell of a grouped
// brick covers hold no dof. Returns false when the cells do not form such
// bricks.
//
// Dofs held by one brick only are stored by brick_apply directly; every
// occurrence of a dof held by several bricks writes its own halo slot, and
// k_brick_halo sums a dof's slots (contiguous, in brick order).
bool build_bricks(gmf_plan& p, Bucket& b, const int32_t* lidx, const std::vector<int64_t>& red_of_full,
                  const int32_t* brick_cells, int64_t brick_count) {
  const int N = b.N[0];
  const int M = 2 * N - 1, L = M * M * M;
  const int NNN = N * N * N;
  if (b.cells == 0) return false;
  if (brick_cells) {
    // every cell of the batch exactly once
    require(brick_count >= 1 && brick_count <= (int64_t(1) << 27), GMF_EINVAL, "invalid brick count");
    std::vector<char> placed(size_t(b.cells), 0);
    for (int64_t i = 0; i < brick_count * 8; ++i) {
      const int32_t c = brick_cells[i];
      require(c >= -1 && c < b.cells, GMF_EINVAL, "brick cell out of range");
      if (c < 0) continue;
      require(!placed[size_t(c)], GMF_EINVAL, "a cell appears in two brick positions");
      placed[size_t(c)] = 1;
    }
    for (int64_t c = 0; c < b.cells; ++c) require(placed[size_t(c)], GMF_EINVAL, "a cell belongs to no brick");
  } else if (b.cells % 8 != 0) {
    return false;
  }
  const int64_t nb = brick_cells ? brick_count : b.cells / 8;
  if (nb > (int64_t(1) << 27) || p.n_full >= (int64_t(1) << 29)) return false;
  std::vector<int32_t> map(size_t(nb) * size_t(L));
  auto latp = [N](int m) { return m == 0 ? 0 : (m == 1 ? N - 1 : m - 1); };
  std::vector<int32_t> lattice(static_cast<size_t>(L));
  std::vector<int64_t> seen_brick(size_t(p.n_full), -1);
  for (int64_t br = 0; br < nb; ++br) {
    std::fill(lattice.begin(), lattice.end(), -2);
    for (int c = 0; c < 8; ++c) {
      const int ox = (c & 1) * (N - 1), oy = ((c >> 1) & 1) * (N - 1), oz = ((c >> 2) & 1) * (N - 1);
      const int64_t cell = brick_cells ? int64_t(brick_cells[br * 8 + c]) : br * 8 + c;
      if (cell < 0) continue;
      const int32_t* li = lidx + cell * NNN;
      for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
          for (int i = 0; i < N; ++i) {
            const int32_t g = li[i + N * j + N * N * k];
            if (g < -1) return false;
            const int pos = (ox + latp(i)) + M * ((oy + latp(j)) + M * (oz + latp(k)));
            if (lattice[size_t(pos)] == -2)
              lattice[size_t(pos)] = g;
            else if (lattice[size_t(pos)] != g)
              return false;
          }
    }
    for (int pos = 0; pos < L; ++pos) {
      int32_t g = lattice[size_t(pos)];
      if (g == -2) {
        if (!brick_cells) return false;
        g = -1;  // covered by no cell of the brick
      }
      if (g >= 0) {
        if (seen_brick[size_t(g)] == br) return false;  // dof twice in one brick
        seen_brick[size_t(g)] = br;
      }
      map[size_t(br * L + pos)] = g;
    }
  }
  // Shared dofs (held by more than one brick): their occurrences take
  // consecutive halo slots per dof, in brick order; each brick lists the slots
  // of its shared positions in ascending lattice position.
  std::vector<int32_t> occ(static_cast<size_t>(p.n_full), 0);
  for (size_t i = 0; i < map.size(); ++i)
    if (map[i] >= 0) ++occ[size_t(map[i])];
  std::vector<int32_t> shared_dofs, ptr(1, 0), next(static_cast<size_t>(p.n_full), -1);
  for (int64_t g = 0; g < p.n_full; ++g)
    if (occ[size_t(g)] > 1) {
      if (int64_t(ptr.back()) + occ[size_t(g)] >= (int64_t(1) << 31)) return false;
      next[size_t(g)] = ptr.back();
      shared_dofs.push_back(int32_t(g));
      ptr.push_back(ptr.back() + occ[size_t(g)]);
    }
  const int64_t slots = ptr.back();
  const int surface = M * M * M - (M - 2) * (M - 2) * (M - 2);
  std::vector<int32_t> base(size_t(nb) + 1, 0), slot_of;
  slot_of.reserve(size_t(slots));
  std::vector<int32_t> enc_full(map.size()), enc_red(red_of_full.empty() ? 0 : map.size());
  for (int64_t br = 0; br < nb; ++br) {
    base[size_t(br)] = int32_t(slot_of.size());
    for (int pos = 0; pos < L; ++pos) {
      const size_t i = size_t(br * L + pos);
      const int32_t g = map[i];
      const bool exclusive = g < 0 || occ[size_t(g)] == 1;
      if (!exclusive) slot_of.push_back(next[size_t(g)]++);
      enc_full[i] = g >= 0 ? (g << 1) | int32_t(exclusive) : -1;
      if (!red_of_full.empty()) {
        const int64_t rg = g >= 0 ? red_of_full[size_t(g)] : -1;
        enc_red[i] = rg >= 0 ? (int32_t(rg) << 1) | int32_t(exclusive) : (exclusive ? -1 : -2);
      }
    }
    if (int64_t(slot_of.size()) - base[size_t(br)] > surface) return false;  // the kernel's slot list bound
  }
  base[size_t(nb)] = int32_t(slot_of.size());
  BrickPath& bp = b.bricks;
  bp.bricks = nb;
  bp.map_full = upload_array(p, enc_full.data(), enc_full.size());
  if (!red_of_full.empty()) bp.map_red = upload_array(p, enc_red.data(), enc_red.size());
  bp.halo_base = upload_array(p, base.data(), base.size());
  bp.halo_slots = slots;
  bp.shared_count = int64_t(shared_dofs.size());
  if (slots > 0) bp.halo_slot = upload_array(p, slot_of.data(), slot_of.size());
  if (bp.shared_count > 0) {
    std::vector<int32_t> red(shared_dofs.size(), -1);
    if (!red_of_full.empty())
      for (size_t i = 0; i < shared_dofs.size(); ++i) red[i] = int32_t(red_of_full[size_t(shared_dofs[i])]);
    bp.halo_full = upload_array(p, shared_dofs.data(), shared_dofs.size());
    bp.halo_red = upload_array(p, red.data(), red.size());
    bp.halo_ptr = upload_array(p, ptr.data(), ptr.size());
  }
  bp.slot_base = p.brick_slots;
  p.brick_slots += slots;
  bp.cell_of = nullptr;
  if (brick_cells) {
    bool consecutive = nb * 8 == b.cells;
    for (int64_t i = 0; consecutive && i < nb * 8; ++i) consecutive = brick_cells[i] == i;
    if (!consecutive) bp.cell_of = upload_array(p, brick_cells, size_t(nb) * 8);
  }
  bp.enabled = true;
  return true;
}

}  // namespace

void prepare_fast_path(gmf_plan& p, Bucket& b, const int32_t* lidx, const std::vector<int64_t>& red_of_full,
                       const int32_t* brick_cells, int64_t brick_count) {
  b.fast = false;
  if (b.dim != 3 || b.form != GMF_FORM_DIFFUSION || b.multi_count != 0) return;
  const int n = b.N[0];
  if (n < 2 || n > 9) return;
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
  // Batches of at least 4096 cells get the brick decomposition; the plan-time
  // tuner (tune_fast) then keeps it unless the per-cell kernel is clearly
  // faster on the batch. GMF_BRICK_MIN_CELLS overrides the threshold,
  // GMF_BRICK_MIN_CELLS_N<n> for batches of n modes per axis only.
  int64_t min_cells = 4096;
  if (const char* env = std::getenv("GMF_BRICK_MIN_CELLS")) min_cells = std::strtoll(env, nullptr, 10);
  const std::string per_n = "GMF_BRICK_MIN_CELLS_N" + std::to_string(n);
  if (const char* env = std::getenv(per_n.c_str())) min_cells = std::strtoll(env, nullptr, 10);
  if (b.lcoef == nullptr && n <= 6 && b.cells >= min_cells)
    build_bricks(p, b, lidx, red_of_full, brick_cells, brick_count);
}

namespace {

template <class R>
bool launch_cells_n(const Bucket& b, const R* x, R* y, bool reduced, int threads, hipStream_t stream) {
  switch (b.N[0]) {
    case 2: launch_cells<2, R>(b, x, y, reduced, threads, stream); return true;
    case 3: launch_cells<3, R>(b, x, y, reduced, threads, stream); return true;
    case 4: launch_cells<4, R>(b, x, y, reduced, threads, stream); return true;
    case 5: launch_cells<5, R>(b, x, y, reduced, threads, stream); return true;
    case 6: launch_cells<6, R>(b, x, y, reduced, threads, stream); return true;
    case 7: launch_cells<7, R>(b, x, y, reduced, threads, stream); return true;
    case 8: launch_cells<8, R>(b, x, y, reduced, threads, stream); return true;
    case 9: launch_cells<9, R>(b, x, y, reduced, threads, stream); return true;
    default: return false;
  }
}

template <class R>
bool launch_fast(const gmf_plan& p, const Bucket& b, const R* x, R* y, bool reduced, R* slots, hipStream_t stream) {
  if (!b.fast || b.cells == 0) return false;
  if (reduced && !b.ridx) return false;
  if (b.bricks.enabled && (!reduced || b.bricks.map_red)) {
    const StoreResult<R> st{y, p.buckets.size() != 1 ? 1 : 0};
    switch (b.N[0]) {
      case 2: launch_bricks<2, R>(b, x, reduced, slots, stream, st, false); return true;
      case 3: launch_bricks<3, R>(b, x, reduced, slots, stream, st, false); return true;
      case 4: launch_bricks<4, R>(b, x, reduced, slots, stream, st, false); return true;
      case 5: launch_bricks<5, R>(b, x, reduced, slots, stream, st, false); return true;
      case 6: launch_bricks<6, R>(b, x, reduced, slots, stream, st, false); return true;
      default: break;
    }
  }
  return launch_cells_n<R>(b, x, y, reduced, b.cell_threads, stream);
}

template <class R>
bool launch_chebyshev_fused(const gmf_plan& p, const R* u, const ChebFused<R>& c, R* slots, hipStream_t stream) {
  if (!p.selection_map || p.buckets.size() != 1) return false;
  const Bucket& b = p.buckets[0];
  if (b.cloud || !b.fast || b.cells == 0 || !b.bricks.enabled || !b.bricks.map_red) return false;
  if (std::is_same<R, float>::value && !b.fp32) return false;
  switch (b.N[0]) {
    case 2: launch_bricks<2, R>(b, u, true, slots, stream, c, true); return true;
    case 3: launch_bricks<3, R>(b, u, true, slots, stream, c, true); return true;
    case 4: launch_bricks<4, R>(b, u, true, slots, stream, c, true); return true;
    case 5: launch_bricks<5, R>(b, u, true, slots, stream, c, true); return true;
    case 6: launch_bricks<6, R>(b, u, true, slots, stream, c, true); return true;
    default: return false;
  }
}

__global__ void k_round_to_float(float* __restrict__ out, const double* __restrict__ in, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    out[i] = float(in[i]);
}

float* float_copy(gmf_plan& p, const double* d, size_t n) {
  if (d == nullptr || n == 0) return nullptr;
  void* f = nullptr;
  GMF_HIP(hipMalloc(&f, n * sizeof(float)));
  p.allocations.push_back(f);
  p.device_bytes += int64_t(n * sizeof(float));
  const int64_t blocks = std::min<int64_t>((int64_t(n) + 255) / 256, 4096);
  hipLaunchKernelGGL(k_round_to_float, dim3(unsigned(blocks)), dim3(256), 0, nullptr, static_cast<float*>(f), d,
                     int64_t(n));
  GMF_HIP(hipGetLastError());
  return static_cast<float*>(f);
}

}  // namespace

// Bricks stay unless the best per-cell shape beats them by this factor: a
// brick batch also runs the multigrid smoother steps fused into its stores.
constexpr float kBrickKeepFactor = 0.9f;

void tune_fast(gmf_plan& p, TuneTimer& t) {
  int forced = 0;
  if (const char* env = std::getenv("GMF_CELL_THREADS")) forced = std::atoi(env);
  // GMF_KEEP_BRICKS=1: brick batches stay on the brick kernel.
  const char* keep_env = std::getenv("GMF_KEEP_BRICKS");
  const bool keep_bricks = keep_env && std::string(keep_env) == "1";
  for (Bucket& b : p.buckets) {
    if (!b.fast || b.cloud || b.cells == 0) continue;
    TuneInfo& ti = b.tune;
    float brick_us = -1.0f;
    if (b.bricks.enabled) {
      const StoreResult<double> st{t.y, 0};
      brick_us = t.time_us([&] {
        switch (b.N[0]) {
          case 2: launch_bricks<2, double>(b, t.x, false, t.slots, nullptr, st, false); break;
          case 3: launch_bricks<3, double>(b, t.x, false, t.slots, nullptr, st, false); break;
          case 4: launch_bricks<4, double>(b, t.x, false, t.slots, nullptr, st, false); break;
          case 5: launch_bricks<5, double>(b, t.x, false, t.slots, nullptr, st, false); break;
          default: launch_bricks<6, double>(b, t.x, false, t.slots, nullptr, st, false); break;
        }
      });
      ++ti.candidates;
    }
    float cell_us = -1.0f;
    int best_threads = 256;
    for (int threads : {64, 128, 256}) {
      if (forced && threads != forced) continue;
      const float us = t.time_us([&] { launch_cells_n<double>(b, t.x, t.y, false, threads, nullptr); });
      ++ti.candidates;
      if (threads == 256) ti.default_us = us;
      if (cell_us < 0.0f || us < cell_us) {
        cell_us = us;
        best_threads = threads;
      }
    }
    if (b.bricks.enabled) ti.default_us = brick_us;
    b.cell_threads = best_threads;
    if (b.bricks.enabled && !keep_bricks && cell_us >= 0.0f && cell_us < kBrickKeepFactor * brick_us)
      b.bricks.enabled = false;
    ti.kernel = b.bricks.enabled ? kTuneBrick : kTuneCell;
    ti.param0 = b.bricks.enabled ? 0 : b.cell_threads;
    ti.best_us = b.bricks.enabled ? brick_us : cell_us;
  }
}

bool launch_cell_fast(const gmf_plan& p, const Bucket& b, const double* x, double* y, bool reduced, double* slots,
                      hipStream_t stream) {
  return launch_fast<double>(p, b, x, y, reduced, slots, stream);
}

bool launch_cell_fast32(const gmf_plan& p, const Bucket& b, const float* x, float* y, bool reduced, float* slots,
                        hipStream_t stream) {
  if (!b.fp32) return false;
  return launch_fast<float>(p, b, x, y, reduced, slots, stream);
}

bool chebyshev_fused(const gmf_plan& p, const double* u, const ChebFused<double>& c, double* slots,
                     hipStream_t stream) {
  return launch_chebyshev_fused<double>(p, u, c, slots, stream);
}

bool chebyshev_fused(const gmf_plan& p, const float* u, const ChebFused<float>& c, float* slots,
                     hipStream_t stream) {
  return launch_chebyshev_fused<float>(p, u, c, slots, stream);
}

bool prepare_fp32(gmf_plan& p) {
  if (p.fp32_state != 0) return p.fp32_state > 0;
  bool all = p.selection_map;
  for (Bucket& b : p.buckets) {
    if (!b.fast || b.cloud) {
      all = false;
      continue;
    }
    const size_t cells = size_t(b.cells);
    b.jinv32 = float_copy(p, b.jinv, cells * 3);
    b.pdata32 = float_copy(p, b.pdata, cells * size_t(b.Qtot) * size_t(b.point_stride));
    b.lcoef32 = float_copy(p, b.lcoef, cells * size_t(b.Ntot));
    b.fp32 = true;
  }
  GMF_HIP(hipDeviceSynchronize());
  p.fp32_state = all ? 1 : -1;
  return all;
}

bool launch_cell_l2_fast(const Bucket& b, const double* u, const double* w, const double* e, double* partial,
                         hipStream_t stream) {
  if (!b.fast || b.cells == 0) return false;
  auto go = [&](auto nc) {
    constexpr int N = decltype(nc)::value;
    constexpr int CPB = 256 / lanes_per_cell<N>();
    EOTables<N, double> T = tables_of<N, double>(b);
    CellArgs<double> a{b.cells, b.jinv, nullptr, b.lidx, b.lcoef};
    const dim3 grid(unsigned((b.cells + CPB - 1) / CPB));
    if (b.lcoef)
      hipLaunchKernelGGL((cell_l2_fast<N, true>), grid, dim3(256), 0, stream, a, T, u, w, e, partial);
    else
      hipLaunchKernelGGL((cell_l2_fast<N, false>), grid, dim3(256), 0, stream, a, T, u, w, e, partial);
    GMF_HIP(hipGetLastError());
    return true;
  };
  switch (b.N[0]) {
    case 2: return go(std::integral_constant<int, 2>());
    case 3: return go(std::integral_constant<int, 3>());
    case 4: return go(std::integral_constant<int, 4>());
    case 5: return go(std::integral_constant<int, 5>());
    case 6: return go(std::integral_constant<int, 6>());
    case 7: return go(std::integral_constant<int, 7>());
    case 8: return go(std::integral_constant<int, 8>());
    case 9: return go(std::integral_constant<int, 9>());
    default: return false;
  }
}

bool plan_writes_every_dof(const gmf_plan& p) {
  // The owners of a brick bucket store every dof of its lattices; with a
  // single brick bucket covering the plan no zero fill of the result is needed.
  return p.buckets.size() == 1 && p.buckets[0].bricks.enabled;
}

}  // namespace gmf
