// Internal interfaces shared by the device translation units (plan.hip, mg.hip):
// the preconditioner base, reduced operator application and the CG core.
#pragma once
#include <exception>
#include <new>
#include <string>

#include "common.hpp"

struct gmf_preconditioner {
  gmf_workspace* ws = nullptr;
  virtual ~gmf_preconditioner() = default;
  // z = M^-1 r on ws->stream.
  virtual void apply(const double* r, double* z) = 0;
  // z = M^-1 r and scalars[slot] = r.z (fused where the preconditioner allows).
  virtual void apply_dot(const double* r, double* z, int64_t n, int slot);
  // Z = M^-1 R for k columns of n rows (column-major) in one pass where the
  // preconditioner supports it; false: apply column by column.
  virtual bool apply_block(const double*, double*, int64_t, int) { return false; }
};

namespace gmf {

// Records the thread's last error message (gmf_last_error) and returns s.
gmf_status record_error(gmf_status s, const std::string& message);

template <class F>
gmf_status guarded_call(F&& f) {
  try {
    f();
    return GMF_OK;
  } catch (const Error& e) {
    return record_error(e.status, e.what());
  } catch (const std::bad_alloc&) {
    return record_error(GMF_EDEVICE, "host allocation failed");
  } catch (const std::exception& e) {
    return record_error(GMF_EINVAL, e.what());
  }
}

inline unsigned grid1(int64_t n) {
  int64_t g = (n + 255) / 256;
  return unsigned(std::max<int64_t>(1, std::min<int64_t>(g, 4096)));
}

template <class T>
T* dalloc(size_t n) {
  void* d = nullptr;
  if (n == 0) n = 1;
  GMF_HIP(hipMalloc(&d, n * sizeof(T)));
  return static_cast<T*>(d);
}

// Owning device buffer.
template <class T>
struct DeviceBuffer {
  T* p = nullptr;
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) : p(dalloc<T>(n)) {}
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p(o.p) { o.p = nullptr; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
  ~DeviceBuffer() {
    if (p) (void)hipFree(p);
  }
};

// Rejects pointers that are not device (or managed) memory of the current
// device, so a host array passed by mistake fails with GMF_EINVAL instead of
// faulting a kernel.
inline void require_device_pointer(const void* p, const char* what) {
  if (p == nullptr) throw Error(GMF_EINVAL, std::string(what) + ": null pointer");
  hipPointerAttribute_t attr{};
  const hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(GMF_EINVAL, std::string(what) + ": not a device pointer");
  }
  if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged)
    throw Error(GMF_EINVAL, std::string(what) + ": not a device pointer");
}

void sync(gmf_workspace* ws);
// GMF_PROGRESS=N (N > 0): the Krylov loops print one line to stderr every N
// iterations (long solves); 0 when unset.
int progress_interval();
// y_red = P^T A P x_red on ws->stream (src/assembly.jl:905-915).
void reduced_apply(gmf_workspace* ws, const double* xr, double* yr);
// Y = A_red X for k columns of n_red rows (column-major; block.hip).
void reduced_apply_block(gmf_workspace* ws, const double* X, double* Y, int k);
// The same operator on FP32 vectors: FP32 kernels on plans whose buckets all
// have them, otherwise the FP64 apply of the widened input, rounded back.
// The workspace must have been prepared with prepare_fp32_workspace (which
// allocates; the apply itself does not).
void reduced_apply_f32(gmf_workspace* ws, const float* xr, float* yr);
void prepare_fp32_workspace(gmf_workspace* ws);
// Exact reduced diagonal into diag (n_red) on ws->stream.
void reduced_diagonal(gmf_workspace* ws, double* diag);
// inv = 1 / d with the reference's zero / non-finite rejection (GMF_EDIAGONAL).
void invert_diagonal(gmf_workspace* ws, const double* d, double* inv, int64_t n);
void precond_apply(gmf_workspace* ws, gmf_preconditioner* M, const double* r, double* z, int64_t n);
void precond_apply_dot(gmf_workspace* ws, gmf_preconditioner* M, const double* r, double* z, int64_t n, int slot);

struct CgVectors {
  DeviceBuffer<double> r, z, p, q;
  explicit CgVectors(int64_t n)
      : r(static_cast<size_t>(n)), z(static_cast<size_t>(n)), p(static_cast<size_t>(n)), q(static_cast<size_t>(n)) {}
};

// Preconditioned CG core (src/solve.jl:376-431); throws Error on breakdown or
// non-convergence.
void cg_run(gmf_workspace* ws, gmf_preconditioner* M, const double* b, double* x, bool use_initial, double rtol,
            double atol, int64_t maxiter, gmf_solve_info& inf, CgVectors& v);

struct FgmresVectors {
  int64_t n = 0;
  int m = 0;
  DeviceBuffer<double> r, q, V, Z, ydev;
  FgmresVectors(int64_t nn, int mm)
      : n(nn),
        m(mm),
        r(static_cast<size_t>(nn)),
        q(static_cast<size_t>(nn)),
        V(static_cast<size_t>(nn) * static_cast<size_t>(mm + 1)),
        Z(static_cast<size_t>(nn) * static_cast<size_t>(mm)),
        ydev(static_cast<size_t>(mm)) {}
};

// Restarted right-preconditioned FGMRES core (src/solve.jl:433-614), restart
// m <= v.m; throws Error on breakdown or non-convergence.
void fgmres_run(gmf_workspace* ws, gmf_preconditioner* M, const double* b, double* x, bool use_initial, int m,
                double rtol, double atol, int64_t maxiter, gmf_solve_info& inf, FgmresVectors& v);

// One modified Gram-Schmidt step of q against the columns 0..col of V (column
// stride n): h_i = q.v_i, q -= h_i v_i in turn, then |q|^2. The col + 2
// values are read back into ws->host_scalars[kHessian ..].
void arnoldi_mgs(gmf_workspace* ws, double* q, const double* V, int64_t n, int col);

struct JacobiPrecond : gmf_preconditioner {
  double* invdiag = nullptr;
  bool owns = true;
  int64_t n = 0;
  ~JacobiPrecond() override {
    if (invdiag && owns) (void)hipFree(invdiag);
  }
  void apply(const double* r, double* z) override;
  void apply_dot(const double* r, double* z, int64_t nn, int slot) override;
};

}  // namespace gmf
