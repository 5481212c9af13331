// Even-odd split one-dimensional kernels for the fast tensor path.
//
// On the symmetric Gauss rule (xi_{N-1-q} = -xi_q) the integrated Legendre
// family has definite parity once the two vertex functions are recombined:
// psi_0 + psi_1 = 1 (even), psi_1 - psi_0 = xi (odd), psi_m(-xi) =
// (-1)^m psi_m(xi) for m >= 2. The collocation derivative matrix D of the
// Gauss points is centro-antisymmetric, D[N-1-q][N-1-s] = -D[q][s]. Splitting
// every pencil into its even and odd halves therefore replaces each N x N
// matrix-vector product by two products of about half the size, and halves
// the number of table entries the kernels keep in scalar registers.
//
// All helpers operate on one pencil of N values in registers and are plain
// host/device functions so that they can be checked on the CPU. R is the
// arithmetic type (double for the operator, float for the FP32 multigrid
// levels); the tables are built in double and rounded once to R.
#pragma once

#ifndef GMF_HD
#if defined(__HIPCC__)
#define GMF_HD __host__ __device__
#else
#define GMF_HD
#endif
#endif

#include <cmath>

namespace gmf {

template <int N, class R = double>
struct EOTables {
  static constexpr int H = (N + 1) / 2;  // points q < H (the middle point included for odd N)
  static constexpr int P = N / 2;        // mirrored point pairs
  static constexpr int NE = (N + 1) / 2; // even functions: 1, psi_2, psi_4, ...
  static constexpr int NO = N / 2;       // odd functions: xi, psi_3, psi_5, ...
  R be[NE * H];   // be[m*H + q] = E_m(xi_q)
  R bo[NO * P];   // bo[m*P + q] = O_m(xi_q)
  R de[P * P];    // de[q*P + s] = (D[q][s] + D[q][N-1-s]) / 2
  R dm[P + 1];    // dm[q] = D[q][mid] (odd N), q < P
  R dd[H * P];    // dd[q*P + s] = (D[q][s] - D[q][N-1-s]) / 2, q < H
};

// Build the split tables from B[i*N+q] = psi_i(xi_q) and D[q*N+s].
template <int N, class R = double>
inline EOTables<N, R> make_eo_tables(const double* B, const double* D) {
  EOTables<N, R> t{};
  constexpr int H = EOTables<N, R>::H, P = EOTables<N, R>::P;
  for (int q = 0; q < H; ++q) {
    t.be[0 * H + q] = R(B[0 * N + q] + B[1 * N + q]);  // psi_0 + psi_1 = 1
    for (int m = 1; m < EOTables<N, R>::NE; ++m) t.be[m * H + q] = R(B[(2 * m) * N + q]);
  }
  for (int q = 0; q < P; ++q) {
    t.bo[0 * P + q] = R(B[1 * N + q] - B[0 * N + q]);  // psi_1 - psi_0 = xi
    for (int m = 1; m < EOTables<N, R>::NO; ++m) t.bo[m * P + q] = R(B[(2 * m + 1) * N + q]);
  }
  for (int q = 0; q < P; ++q)
    for (int s = 0; s < P; ++s) t.de[q * P + s] = R(0.5 * (D[q * N + s] + D[q * N + (N - 1 - s)]));
  for (int q = 0; q < P; ++q) t.dm[q] = R((N % 2) ? D[q * N + P] : 0.0);
  for (int q = 0; q < H; ++q)
    for (int s = 0; s < P; ++s) t.dd[q * P + s] = R(0.5 * (D[q * N + s] - D[q * N + (N - 1 - s)]));
  return t;
}

// Modal coefficients u (psi order) -> values at the N points.
template <int N, class R>
GMF_HD inline void eo_interp(const EOTables<N, R>& t, const R (&u)[N], R (&v)[N]) {
  constexpr int H = EOTables<N, R>::H, P = EOTables<N, R>::P, NE = EOTables<N, R>::NE, NO = EOTables<N, R>::NO;
  R e[NE], o[NO];
  e[0] = R(0.5) * (u[0] + u[1]);
  o[0] = R(0.5) * (u[1] - u[0]);
#pragma unroll
  for (int m = 1; m < NE; ++m) e[m] = u[2 * m];
#pragma unroll
  for (int m = 1; m < NO; ++m) o[m] = u[2 * m + 1];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    R E = R(0), O = R(0);
#pragma unroll
    for (int m = 0; m < NE; ++m) E = fma(t.be[m * H + q], e[m], E);
#pragma unroll
    for (int m = 0; m < NO; ++m) O = fma(t.bo[m * P + q], o[m], O);
    v[q] = E + O;
    v[N - 1 - q] = E - O;
  }
  if (N % 2) {
    R E = R(0);
#pragma unroll
    for (int m = 0; m < NE; ++m) E = fma(t.be[m * H + P], e[m], E);
    v[P] = E;
  }
}

// Point values f -> r_i = sum_q psi_i(xi_q) f_q (transpose of eo_interp).
template <int N, class R>
GMF_HD inline void eo_project(const EOTables<N, R>& t, const R (&f)[N], R (&r)[N]) {
  constexpr int H = EOTables<N, R>::H, P = EOTables<N, R>::P, NE = EOTables<N, R>::NE, NO = EOTables<N, R>::NO;
  R fe[H], fo[P];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    fe[q] = f[q] + f[N - 1 - q];
    fo[q] = f[q] - f[N - 1 - q];
  }
  if (N % 2) fe[P] = f[P];
  R re[NE], ro[NO];
#pragma unroll
  for (int m = 0; m < NE; ++m) {
    R v = R(0);
#pragma unroll
    for (int q = 0; q < H; ++q) v = fma(t.be[m * H + q], fe[q], v);
    re[m] = v;
  }
#pragma unroll
  for (int m = 0; m < NO; ++m) {
    R v = R(0);
#pragma unroll
    for (int q = 0; q < P; ++q) v = fma(t.bo[m * P + q], fo[q], v);
    ro[m] = v;
  }
  r[0] = R(0.5) * (re[0] - ro[0]);
  r[1] = R(0.5) * (re[0] + ro[0]);
#pragma unroll
  for (int m = 1; m < NE; ++m) r[2 * m] = re[m];
#pragma unroll
  for (int m = 1; m < NO; ++m) r[2 * m + 1] = ro[m];
}

// Collocation derivative at the points: dv_q = sum_s D[q][s] v_s.
template <int N, class R>
GMF_HD inline void eo_deriv(const EOTables<N, R>& t, const R (&v)[N], R (&dv)[N]) {
  constexpr int P = EOTables<N, R>::P;
  R ve[P], vo[P];
#pragma unroll
  for (int s = 0; s < P; ++s) {
    ve[s] = v[s] + v[N - 1 - s];
    vo[s] = v[s] - v[N - 1 - s];
  }
#pragma unroll
  for (int q = 0; q < P; ++q) {
    R A = (N % 2) ? t.dm[q] * v[P] : R(0), Bq = R(0);
#pragma unroll
    for (int s = 0; s < P; ++s) {
      A = fma(t.de[q * P + s], ve[s], A);
      Bq = fma(t.dd[q * P + s], vo[s], Bq);
    }
    dv[q] = A + Bq;
    dv[N - 1 - q] = Bq - A;
  }
  if (N % 2) {
    R m = R(0);
#pragma unroll
    for (int s = 0; s < P; ++s) m = fma(t.dd[P * P + s], vo[s], m);
    dv[P] = m;
  }
}

// Transposed collocation derivative: h_s = sum_q D[q][s] f_q.
template <int N, class R>
GMF_HD inline void eo_deriv_t(const EOTables<N, R>& t, const R (&f)[N], R (&h)[N]) {
  constexpr int P = EOTables<N, R>::P;
  R fe[P], fo[P];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    fe[q] = f[q] + f[N - 1 - q];
    fo[q] = f[q] - f[N - 1 - q];
  }
#pragma unroll
  for (int s = 0; s < P; ++s) {
    R A = R(0), Bs = (N % 2) ? t.dd[P * P + s] * f[P] : R(0);
#pragma unroll
    for (int q = 0; q < P; ++q) {
      A = fma(t.de[q * P + s], fo[q], A);
      Bs = fma(t.dd[q * P + s], fe[q], Bs);
    }
    h[s] = A + Bs;
    h[N - 1 - s] = A - Bs;
  }
  if (N % 2) {
    R m = R(0);
#pragma unroll
    for (int q = 0; q < P; ++q) m = fma(t.dm[q], fo[q], m);
    h[P] = m;
  }
}

}  // namespace gmf
