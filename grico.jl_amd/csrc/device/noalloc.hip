// Reduced operator apply that allocates nothing (include/grico_mf.h
// gmf_reduced_apply_noalloc): the same kernels and arithmetic as
// gmf_reduced_apply, with the workspace scratch (brick halo slots; for plans
// whose reduced map is not a selection, the full-layout input and result)
// taken from a fixed device array instead of a workspace. For callers that
// cannot allocate device memory: code under HIP graph capture, serving loops.
//
// The array lives in this file's code object, one copy per device, for the
// life of the process. Its device address is looked up once per device, with
// that device current, on the first call there; the kernels receive that
// address, never the array's host-side name. Every call on one device shares
// the array, so calls that overlap on a device race; every value a call reads
// from it was written earlier in the same call, so nothing needs clearing.
#include <mutex>

#include "common.hpp"
#include "internal.hpp"

namespace gmf {
namespace {

constexpr int64_t kScratchDoubles = int64_t(1) << 22;  // 32 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_address[kMaxDevices] = {};

// Device address of `scratch` on device `dev` (current), looked up once.
double* device_scratch(int dev) {
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (scratch_address[dev] == nullptr) {
    void* ptr = nullptr;
    GMF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    require(ptr != nullptr, GMF_EDEVICE, "hipGetSymbolAddress returned no address for the apply scratch");
    scratch_address[dev] = static_cast<double*>(ptr);
  }
  return scratch_address[dev];
}

}  // namespace
}  // namespace gmf

using namespace gmf;

extern "C" gmf_status gmf_reduced_apply_noalloc(const gmf_plan* plan, const double* x, double* y, void* stream) {
  return guarded_call([&] {
    require(plan != nullptr && x != nullptr && y != nullptr, GMF_EINVAL, "null argument");
    require(x != y, GMF_EALIAS, "result and reduced coefficient vector must not alias");
    int dev = -1;
    GMF_HIP(hipGetDevice(&dev));
    require(dev == plan->device, GMF_EINVAL, "the plan's device must be the current device");
    require(dev >= 0 && dev < kMaxDevices, GMF_EUNSUPPORTED, "device index beyond the allocation-free apply's table");
    require(plan->buckets.size() <= BucketFlags::kStackBuckets, GMF_EUNSUPPORTED,
            "the allocation-free apply supports plans of at most 64 cell batches");
    const int64_t need = plan->brick_slots + (plan->selection_map ? 0 : 2 * plan->n_full);
    require(need <= kScratchDoubles, GMF_EUNSUPPORTED,
            "plan needs " + std::to_string(need) + " doubles of apply scratch; the allocation-free apply holds " +
                std::to_string(kScratchDoubles) + " (32 MiB): use gmf_reduced_apply with a workspace");
    double* s = device_scratch(dev);
    gmf_workspace w{};
    w.plan = const_cast<gmf_plan*>(plan);
    w.stream = static_cast<hipStream_t>(stream);
    w.brick_slots = s;
    if (!plan->selection_map) {
      w.full_in = s + plan->brick_slots;
      w.full_out = w.full_in + plan->n_full;
    }
    reduced_apply(&w, x, y);
    GMF_HIP(hipGetLastError());
  });
}
