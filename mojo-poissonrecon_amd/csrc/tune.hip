// The workgroup-size table of the batched kernels and the autotuner that fills it.
#include <algorithm>
#include <atomic>
#include <vector>

#include "mpr_batch.h"
#include "mpr_hip.h"
#include "mpr_tune.h"

namespace mpr {

int ensure_device();

namespace {

std::atomic<int> g_blocks[kTuneKernels][2][kTuneClasses];
std::atomic<int> g_groups[kTuneKernels][2][kTuneClasses];
std::atomic<int> g_planes[kTuneKernels][2][kTuneClasses];

struct TableInit {
  TableInit() {
    for (auto& k : g_blocks)
      for (auto& t : k)
        for (auto& c : t) c.store(kTuneDefault);
    for (auto& k : g_groups)
      for (auto& t : k)
        for (auto& c : t) c.store(1);
    for (auto& k : g_planes)
      for (auto& t : k)
        for (auto& c : t) c.store(1);
  }
} g_table_init;

template <class T>
struct DevArray {
  T* p = nullptr;
  explicit DevArray(size_t count) { MPR_HIP(hipMalloc((void**)&p, std::max<size_t>(1, count) * sizeof(T))); }
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  ~DevArray() {
    if (p) (void)hipFree(p);
  }
};

// milliseconds of `reps` calls of `launch()`
template <class F>
double time_launches(F&& launch, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  MPR_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) launch();
  MPR_HIP(hipEventRecord(e1, s));
  MPR_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MPR_HIP(hipEventElapsedTime(&ms, e0, e1));
  return (double)ms;
}

// Tries every candidate for (kernel, precision, class): the candidates take turns over
// kTuneRounds rounds of `reps` launches each (so clock and thermal drift hit all of them
// alike), and the median round decides. A candidate replaces the default only if it is at
// least kTuneMargin faster.
constexpr int kTuneRounds = 7;
constexpr double kTuneMargin = 0.015;

// (workgroup size, slot groups per lane, z-column planes per lane) candidates; the default
// shape is (256, 1, 1)
struct Shape {
  int block, groups, planes;
};

template <class F>
void tune_one(int kernel, bool fp64, int cls, F&& launch, int reps, hipStream_t s) {
  std::vector<Shape> cand;
  for (int g : kTuneGroupCandidates) {
    if (g != 1 && kernel != kTuneRelax) continue;
    for (int b : kTuneCandidates) cand.push_back({b, g, 1});
  }
  if (kernel == kTuneRelax)
    for (int z : kTunePlaneCandidates) {
      if (z == 1) continue;
      for (int b : kTuneCandidates) cand.push_back({b, 1, z});
    }
  const size_t nc = cand.size();
  auto apply = [&](const Shape& sh) {
    set_tuned_block(kernel, fp64, cls, sh.block);
    set_tuned_groups(kernel, fp64, cls, sh.groups);
    set_tuned_planes(kernel, fp64, cls, sh.planes);
  };
  hipEvent_t e0, e1;
  MPR_HIP(hipEventCreate(&e0));
  MPR_HIP(hipEventCreate(&e1));
  std::vector<std::vector<double>> ms(nc);
  for (size_t c = 0; c < nc; ++c) {  // warm-up: code objects loaded, caches settled
    apply(cand[c]);
    launch();
  }
  for (int round = 0; round < kTuneRounds; ++round)
    for (size_t c = 0; c < nc; ++c) {
      apply(cand[c]);
      ms[c].push_back(time_launches(launch, reps, s, e0, e1));
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::vector<double> med(nc);
  size_t best = 0, dflt = 0;
  for (size_t c = 0; c < nc; ++c) {
    std::sort(ms[c].begin(), ms[c].end());
    med[c] = ms[c][ms[c].size() / 2];
    if (med[c] < med[best]) best = c;
    if (cand[c].block == kTuneDefault && cand[c].groups == 1 && cand[c].planes == 1) dflt = c;
  }
  if (med[best] > (1.0 - kTuneMargin) * med[dflt]) best = dflt;
  apply(cand[best]);
}

// Up to kTuneInstances instances of one class with their own buffers, as a batch runs them,
// within kTuneBytes of device memory (at least one instance).
constexpr size_t kTuneInstances = 8;
constexpr size_t kTuneBytes = size_t(4) << 30;

template <class T>
void tune_class(int cls, int reps, hipStream_t s) {
  const Geom g = make_geom((1 << cls) + 1), gc = make_geom((1 << (cls - 1)) + 1);
  const size_t W = 2ull * g.C, Wc = 2ull * gc.C;
  const size_t per = sizeof(T) * (4 * W + 2 * Wc) + (sizeof(T) == 8 ? 2 * sizeof(float) * W : 0);
  const size_t K = std::max<size_t>(1, std::min(kTuneInstances, kTuneBytes / per));
  DevArray<T> x(K * W), rhs(K * W), scr(K * W), res(K * W), xc(K * Wc), rhsc(K * Wc);
  launch_fill<T>(x.p, T(0.5), K * W, s);
  launch_fill<T>(rhs.p, T(1), K * W, s);
  launch_fill<T>(scr.p, T(0.25), K * W, s);
  std::vector<StepPtrs<T>> e;
  for (size_t k = 0; k < K; ++k)
    e.push_back({x.p + k * W, rhs.p + k * W, scr.p + k * W, res.p + k * W, xc.p + k * Wc,
                 rhsc.p + k * Wc, 1});
  DevArray<StepPtrs<T>> tab(K);
  MPR_HIP(hipMemcpy(tab.p, e.data(), K * sizeof(e[0]), hipMemcpyHostToDevice));
  const bool fp64 = sizeof(T) == 8;
  tune_one(kTuneRelax, fp64, cls, [&] { launch_b_relax<T>(tab.p, K, g, 4.0, 1, s); }, reps, s);
  // the residual in the form the V-cycle launches: FP32 classes whose next-coarser class
  // also has the slot-group layout compute colour 0 only
  const bool c0 = !fp64 && group_layout(g) && group_layout(gc);
  tune_one(kTuneResidual, fp64, cls, [&] { launch_b_residual<T>(tab.p, K, g, 4.0, s, c0); }, reps,
           s);
  if (fp64) {
    DevArray<float> r32(K * W), e32(K * W);
    std::vector<TopPtrs> t;
    for (size_t k = 0; k < K; ++k)
      t.push_back({(double*)x.p + k * W, (const double*)rhs.p + k * W,
                   (const double*)scr.p + k * W, r32.p + k * W, e32.p + k * W});
    DevArray<TopPtrs> ttab(K);
    MPR_HIP(hipMemcpy(ttab.p, t.data(), K * sizeof(t[0]), hipMemcpyHostToDevice));
    tune_one(kTuneDefect, true, cls, [&] { launch_b_defect(ttab.p, K, g, 4.0, false, s); }, reps, s);
  }
  MPR_HIP(hipStreamSynchronize(s));
}

}  // namespace

bool valid_block(int block) {
  for (int b : kTuneCandidates)
    if (b == block) return true;
  return false;
}

int tuned_block(int kernel, bool fp64, int size_class) {
  if (kernel < 0 || kernel >= kTuneKernels || size_class < 0 || size_class >= kTuneClasses)
    return kTuneDefault;
  return g_blocks[kernel][fp64 ? 1 : 0][size_class].load(std::memory_order_relaxed);
}

void set_tuned_block(int kernel, bool fp64, int size_class, int block) {
  if (kernel < 0 || kernel >= kTuneKernels || size_class < 0 || size_class >= kTuneClasses ||
      !valid_block(block))
    return;
  g_blocks[kernel][fp64 ? 1 : 0][size_class].store(block, std::memory_order_relaxed);
}

int tuned_groups(int kernel, bool fp64, int size_class) {
  if (kernel < 0 || kernel >= kTuneKernels || size_class < 0 || size_class >= kTuneClasses)
    return 1;
  return g_groups[kernel][fp64 ? 1 : 0][size_class].load(std::memory_order_relaxed);
}

void set_tuned_groups(int kernel, bool fp64, int size_class, int groups) {
  if (kernel != kTuneRelax || size_class < 0 || size_class >= kTuneClasses ||
      (groups != 1 && groups != 2))
    return;
  g_groups[kernel][fp64 ? 1 : 0][size_class].store(groups, std::memory_order_relaxed);
}

bool valid_planes(int planes) {
  for (int z : kTunePlaneCandidates)
    if (z == planes) return true;
  return false;
}

int tuned_planes(int kernel, bool fp64, int size_class) {
  if (kernel < 0 || kernel >= kTuneKernels || size_class < 0 || size_class >= kTuneClasses)
    return 1;
  return g_planes[kernel][fp64 ? 1 : 0][size_class].load(std::memory_order_relaxed);
}

void set_tuned_planes(int kernel, bool fp64, int size_class, int planes) {
  if (kernel != kTuneRelax || size_class < 0 || size_class >= kTuneClasses ||
      !valid_planes(planes))
    return;
  g_planes[kernel][fp64 ? 1 : 0][size_class].store(planes, std::memory_order_relaxed);
}

namespace {
std::atomic<int> g_march{2};
}  // namespace

int march_iterations() { return g_march.load(std::memory_order_relaxed); }
void set_march_iterations(int iterations) {
  g_march.store(iterations, std::memory_order_relaxed);
}

}  // namespace mpr

using namespace mpr;

extern "C" {

int mpr_autotune(int32_t max_depth, int32_t reps, int32_t* table, int32_t* groups,
                 int32_t* planes) {
  if (max_depth < 5 || max_depth > 10 || reps < 1) {
    set_error("mpr_autotune: max_depth must be in [5, 10] and reps positive");
    return -1;
  }
  if (ensure_device()) return -1;
  hipStream_t s = nullptr;
  try {
    MPR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    // the slot-group kernels run on levels with n >= 33 (size class 5 and up)
    for (int cls = 5; cls <= max_depth; ++cls) {
      tune_class<float>(cls, reps, s);
      tune_class<double>(cls, reps, s);
    }
    (void)hipStreamDestroy(s);
  } catch (const HipError& e) {
    if (s) (void)hipStreamDestroy(s);
    set_error(e.msg);
    return -2;
  }
  for (int k = 0; k < kTuneKernels; ++k)
    for (int t = 0; t < 2; ++t)
      for (int c = 0; c < kTuneClasses; ++c) {
        if (table) table[(k * 2 + t) * kTuneClasses + c] = tuned_block(k, t == 1, c);
        if (groups) groups[(k * 2 + t) * kTuneClasses + c] = tuned_groups(k, t == 1, c);
        if (planes) planes[(k * 2 + t) * kTuneClasses + c] = tuned_planes(k, t == 1, c);
      }
  return 0;
}

int mpr_tune_get(int32_t kernel, int32_t fp64, int32_t size_class) {
  if (kernel < 0 || kernel >= kTuneKernels || size_class < 0 || size_class >= kTuneClasses) {
    set_error("mpr_tune_get: no such kernel or size class");
    return -1;
  }
  return tuned_block(kernel, fp64 != 0, size_class);
}

int mpr_tune_set(int32_t kernel, int32_t fp64, int32_t size_class, int32_t block) {
  if (kernel < 0 || kernel >= kTuneKernels || size_class < 0 || size_class >= kTuneClasses ||
      !valid_block(block)) {
    set_error("mpr_tune_set: kernel 0..2, size class 0..10, block one of 64/128/256/512/1024");
    return -1;
  }
  set_tuned_block(kernel, fp64 != 0, size_class, block);
  return 0;
}

int mpr_tune_get_groups(int32_t kernel, int32_t fp64, int32_t size_class) {
  if (kernel < 0 || kernel >= kTuneKernels || size_class < 0 || size_class >= kTuneClasses) {
    set_error("mpr_tune_get_groups: no such kernel or size class");
    return -1;
  }
  return tuned_groups(kernel, fp64 != 0, size_class);
}

int mpr_tune_set_groups(int32_t kernel, int32_t fp64, int32_t size_class, int32_t groups) {
  if (kernel != kTuneRelax || size_class < 0 || size_class >= kTuneClasses ||
      (groups != 1 && groups != 2)) {
    set_error("mpr_tune_set_groups: kernel 0 (colour sweep), size class 0..10, groups 1 or 2");
    return -1;
  }
  set_tuned_groups(kernel, fp64 != 0, size_class, groups);
  return 0;
}

int mpr_tune_get_planes(int32_t kernel, int32_t fp64, int32_t size_class) {
  if (kernel < 0 || kernel >= kTuneKernels || size_class < 0 || size_class >= kTuneClasses) {
    set_error("mpr_tune_get_planes: no such kernel or size class");
    return -1;
  }
  return tuned_planes(kernel, fp64 != 0, size_class);
}

int mpr_tune_get_march(void) { return march_iterations(); }

int mpr_fast_division_check(int32_t n, int64_t* mismatches) {
  if (!mismatches || n < 3 || n > 1025) {
    set_error("mpr_fast_division_check: n in [3, 1025] and a result pointer");
    return -1;
  }
  if (ensure_device()) return -1;
  hipStream_t s = nullptr;
  try {
    MPR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *mismatches = fast_division_mismatches(n, s);
    (void)hipStreamDestroy(s);
  } catch (const HipError& e) {
    if (s) (void)hipStreamDestroy(s);
    set_error(e.msg);
    return -2;
  }
  return 0;
}

int mpr_tune_set_march(int32_t iterations) {
  if (iterations < 0 || iterations > 2) {
    set_error("mpr_tune_set_march: iterations must be 0, 1 or 2");
    return -1;
  }
  set_march_iterations(iterations);
  return 0;
}

int mpr_tune_set_planes(int32_t kernel, int32_t fp64, int32_t size_class, int32_t planes) {
  if (kernel != kTuneRelax || size_class < 0 || size_class >= kTuneClasses ||
      !valid_planes(planes)) {
    set_error("mpr_tune_set_planes: kernel 0 (colour sweep), size class 0..10, planes 1/2/4");
    return -1;
  }
  set_tuned_planes(kernel, fp64 != 0, size_class, planes);
  return 0;
}

}  // extern "C"
