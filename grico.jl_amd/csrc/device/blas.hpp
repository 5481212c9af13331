// Vector / Krylov kernel launchers (blas.hip).
#pragma once
#include "common.hpp"

namespace gmf {

// Solver scalar slots in gmf_workspace::scalars.
enum Slot { kRz = 0, kPq = 1, kRr = 2, kRzNew = 3, kBb = 4, kTmp = 5, kHessian = 16 };
constexpr int kScalarCount = 16 + 512;

void cg_update(gmf_workspace* ws, double* x, double* r, const double* p, const double* q, int64_t n, int num, int den,
               int out_slot);
void jacobi_dot(gmf_workspace* ws, double* z, const double* r, const double* invdiag, int64_t n, int out_slot);
void direction(gmf_workspace* ws, double* p, const double* z, int64_t n, int num, int den);
void axpby(gmf_workspace* ws, double* y, double a, const double* x, double b, int64_t n);
void multi_dot(gmf_workspace* ws, const double* V, int64_t ld, int m, const double* w, int64_t n, double* out_dev);
void multi_axpy(gmf_workspace* ws, double* y, const double* V, int64_t ld, const double* h_dev, int m, int64_t n,
                double scale);

}  // namespace gmf
