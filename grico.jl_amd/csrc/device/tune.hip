// Plan-time autotuner of the bucketed kernels' launch shapes.
//
// A plan mixes buckets of uneven degree, quadrature shape and point count
// (tensor cells on bricks or per cell, generic cells, cut cells and surface
// pieces with a few to a few thousand points each). Which launch shape is
// fastest for a bucket depends on its size and shape, so at plan creation
// every bucket's candidates are timed on scratch vectors (HIP events, one
// warm-up and three timed launches each) and the fastest is kept:
//  * fast tensor buckets: the brick kernel (when the cells form bricks)
//    against the per-cell kernel at 64, 128 and 256 lanes per workgroup;
//  * generic buckets: 64, 128 or 256 lanes per cell;
//  * point-cloud buckets: work units of 256 .. 2048 points and an LDS budget
//    of 64 or 32 KB (the chunk of points staged per pass).
// Every candidate computes the same operator (the tests run each forced
// shape against the oracle); only the work decomposition differs.
// GMF_TUNE=0 keeps the default shapes.
#include <cstdlib>
#include <cstring>
#include <string>

#include "internal.hpp"

namespace gmf {

namespace {

// Record of the shape a bucket runs with when nothing was timed.
void record_default(Bucket& b) {
  TuneInfo& ti = b.tune;
  if (b.cells == 0) {
    ti.kernel = kTuneNone;
  } else if (b.cloud) {
    ti.kernel = b.face.enabled ? kTuneFace : kTuneCloud;
    ti.param0 = b.face.enabled ? 0 : b.unit_points;
    ti.param1 = b.face.enabled ? 0 : b.lds_kb;
  } else if (b.fast) {
    ti.kernel = b.bricks.enabled ? kTuneBrick : kTuneCell;
    ti.param0 = b.bricks.enabled ? 0 : b.cell_threads;
  } else {
    ti.kernel = kTuneGeneric;
    ti.param0 = b.generic_threads;
  }
}

}  // namespace

void tune_plan(gmf_plan& p) {
  for (Bucket& b : p.buckets) record_default(b);
  const char* env = std::getenv("GMF_TUNE");
  if (env && std::string(env) == "0") return;
  bool any = false;
  for (const Bucket& b : p.buckets) any = any || (b.cells > 0 && !(b.cloud && b.face.enabled));
  if (!any) return;
  const size_t n = size_t(std::max<int64_t>(p.n_full, 1));
  DeviceBuffer<double> x(n), y(n), slots(size_t(std::max<int64_t>(p.brick_slots, 1)));
  fill(x.p, 1.0, int64_t(n), nullptr);
  fill(y.p, 0.0, int64_t(n), nullptr);
  TuneTimer t;
  t.x = x.p;
  t.y = y.p;
  t.slots = slots.p;
  GMF_HIP(hipEventCreate(&t.e0));
  GMF_HIP(hipEventCreate(&t.e1));
  try {
    tune_generic(p, t);
    tune_fast(p, t);
    tune_cloud(p, t);
    GMF_HIP(hipDeviceSynchronize());
  } catch (...) {
    (void)hipEventDestroy(t.e0);
    (void)hipEventDestroy(t.e1);
    throw;
  }
  (void)hipEventDestroy(t.e0);
  (void)hipEventDestroy(t.e1);
}

}  // namespace gmf

using namespace gmf;

extern "C" gmf_status gmf_plan_tuning(const gmf_plan* plan, int32_t bucket, gmf_bucket_tuning* out) {
  return guarded_call([&] {
    require(plan != nullptr && out != nullptr, GMF_EINVAL, "invalid argument");
    require(bucket >= 0 && size_t(bucket) < plan->buckets.size(), GMF_EINVAL, "bucket index out of range");
    const TuneInfo& ti = plan->buckets[size_t(bucket)].tune;
    std::memset(out, 0, sizeof(*out));
    out->kernel = ti.kernel;
    out->param0 = ti.param0;
    out->param1 = ti.param1;
    out->candidates = ti.candidates;
    out->best_us = double(ti.best_us);
    out->default_us = double(ti.default_us);
  });
}
