// One-dimensional polynomial families and Gauss-Legendre rules on [-1, 1].
//
// Follows src/polynomials.jl (Legendre recurrence 294-310, integrated family
// 403-434, degree-zero finite-element constant 238-247) and
// src/quadrature.jl:342-403 (Newton from cosine guesses, mirrored roots,
// tolerance 1e3*eps, at most 32 iterations).
#pragma once
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

namespace gh {

// P_0..P_n at x (values) and their derivatives, differentiated in lockstep.
inline void legendre_values_and_derivatives(double x, int n, double* v, double* d) {
  v[0] = 1.0;
  d[0] = 0.0;
  if (n >= 1) {
    v[1] = x;
    d[1] = 1.0;
  }
  for (int k = 2; k <= n; ++k) {
    double inv_k = 1.0 / double(k);
    double a = double(2 * k - 1) * inv_k;
    double b = double(k - 1) * inv_k;
    v[k] = a * x * v[k - 1] - b * v[k - 2];
    d[k] = a * (v[k - 1] + x * d[k - 1]) - b * d[k - 2];
  }
}

inline void legendre_values(double x, int n, double* v) {
  v[0] = 1.0;
  if (n >= 1) v[1] = x;
  for (int k = 2; k <= n; ++k) {
    double inv_k = 1.0 / double(k);
    v[k] = double(2 * k - 1) * inv_k * x * v[k - 1] - double(k - 1) * inv_k * v[k - 2];
  }
}

// Finite-element basis psi_0..psi_n and derivatives: psi_0=(1-x)/2,
// psi_1=(1+x)/2, psi_k=(P_k-P_{k-2})/sqrt(4k-2), psi_k'=sqrt((2k-1)/2) P_{k-1};
// degree 0 is the cellwise constant 1 with zero derivative.
inline void fe_basis(double x, int n, double* vals, double* ders) {
  if (n == 0) {
    if (vals) vals[0] = 1.0;
    if (ders) ders[0] = 0.0;
    return;
  }
  std::vector<double> p(n + 1);
  legendre_values(x, n, p.data());
  if (ders) {
    for (int k = n; k >= 2; --k) ders[k] = std::sqrt(double(2 * k - 1) / 2.0) * p[k - 1];
    ders[0] = -0.5;
    ders[1] = 0.5;
  }
  if (vals) {
    for (int k = n; k >= 2; --k) vals[k] = (p[k] - p[k - 2]) / std::sqrt(double(4 * k - 2));
    vals[0] = 0.5 * (1.0 - x);
    vals[1] = 0.5 * (1.0 + x);
  }
}

// n-point Gauss-Legendre rule, points ascending.
inline void gauss_legendre(int n, double* pts, double* wts) {
  if (n <= 0) throw std::invalid_argument("point_count must be positive");
  std::vector<double> v(n + 1), d(n + 1);
  const double tol = 1.0e3 * std::numeric_limits<double>::epsilon();
  int half = (n + 1) / 2;
  for (int i = 1; i <= half; ++i) {
    double x = std::cos(M_PI * (double(i) - 0.25) / (double(n) + 0.5));
    bool ok = false;
    for (int it = 0; it < 32; ++it) {
      legendre_values_and_derivatives(x, n, v.data(), d.data());
      double delta = v[n] / d[n];
      if (!std::isfinite(delta)) throw std::runtime_error("Gauss-Legendre Newton update is not finite");
      x -= delta;
      if (std::fabs(delta) <= tol) {
        legendre_values_and_derivatives(x, n, v.data(), d.data());
        ok = true;
        break;
      }
    }
    if (!ok) throw std::runtime_error("Gauss-Legendre root did not converge");
    double w = 2.0 / ((1.0 - x * x) * d[n] * d[n]);
    pts[i - 1] = -x;
    wts[i - 1] = w;
    pts[n - i] = x;
    wts[n - i] = w;
  }
}

}  // namespace gh
