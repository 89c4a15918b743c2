// The allocation-free entry point: the iso value (capi.mojo:745-795) of grids and points
// that already live in HBM, for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops that must not call hipMalloc). It runs the kernels of
// mpr_iso_value_f64 on the caller's buffers; their per-block partial sums go to the static
// device array `scratch` below instead of a hipMalloc'd buffer, and the result goes to a
// caller-owned device double. Nothing here allocates or frees device or host memory.
#include <mutex>

#include "mpr_hip.h"
#include "mpr_internal.h"

namespace mpr {
int ensure_device();
}

using namespace mpr;

namespace {

// k_iso_partials writes two doubles (sum of w * value, sum of w) per block of kThreads
// points; the array holds them for up to kNoAllocMaxPoints points (1 MiB per device).
constexpr int64_t kNoAllocMaxPoints = int64_t(1) << 24;
constexpr size_t kScratchDoubles = 2 * (size_t)(kNoAllocMaxPoints / kThreads);
static_assert(kNoAllocMaxPoints % kThreads == 0, "whole blocks");

static __device__ __align__(256) double scratch[kScratchDoubles];

// Each device holds its own copy of `scratch` at its own address; the host name is only a
// stand-in, so kernels receive the address looked up here (once per device, with that
// device current).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

double* device_scratch() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("mpr_iso_value_f64_noalloc: no current HIP device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("mpr_iso_value_f64_noalloc: hipGetSymbolAddress(scratch): ") +
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch[dev] = static_cast<double*>(p);
  }
  return g_scratch[dev];
}

}  // namespace

extern "C" {

int mpr_iso_value_f64_noalloc(const double* field, const double* points, const double* normals,
                              int64_t point_count, int64_t n, int64_t confidence,
                              double* iso_value, void* stream) {
  if (point_count < 0 || point_count > kNoAllocMaxPoints) {
    set_error("mpr_iso_value_f64_noalloc: point_count must be in [0, 16777216] (the size of "
              "the library's static scratch)");
    return -1;
  }
  if (n < 2 || n > 1025) {
    set_error("mpr_iso_value_f64_noalloc: n must be in [2, 1025]");
    return -1;
  }
  if (!field || !iso_value || (point_count > 0 && (!points || !normals))) {
    set_error("mpr_iso_value_f64_noalloc: null device buffer");
    return -1;
  }
  if (ensure_device()) return -1;
  double* part = device_scratch();
  if (!part) return -1;
  try {
    launch_iso(natural_field(field, (int)n), points, normals, (uint32_t)point_count, (int)n,
               (int)(confidence != 0), part, iso_value, static_cast<hipStream_t>(stream));
    return 0;
  } catch (const HipError& e) {
    set_error(std::string("mpr_iso_value_f64_noalloc: ") + e.msg);
    return -1;
  }
}

}  // extern "C"
