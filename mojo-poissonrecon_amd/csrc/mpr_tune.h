// Launch-shape tuning of the batched slot-group kernels: the workgroup size (and, for the
// colour sweep, the slot groups per lane) of each kernel, precision and grid size class is
// looked up in a process-wide table, which the autotuner (mpr_autotune in include/mpr_hip.h)
// fills by timing the candidates on the device. The per-node arithmetic does not depend on
// the launch shape, so results do not change.
#pragma once

#include "mpr_common.h"

namespace mpr {

enum TuneKernel : int { kTuneRelax = 0, kTuneResidual = 1, kTuneDefect = 2, kTuneKernels = 3 };
constexpr int kTuneClasses = 11;  // size classes s = 0..10, n = 2^s + 1
constexpr int kTuneCandidates[] = {64, 128, 256, 512, 1024};
constexpr int kTuneDefault = 256;

// size class of a multigrid level (n = 2^s + 1)
inline int size_class(const Geom& g) {
  int s = 0;
  while ((1 << s) + 1 < g.n && s < kTuneClasses - 1) ++s;
  return s;
}

int tuned_block(int kernel, bool fp64, int size_class);
void set_tuned_block(int kernel, bool fp64, int size_class, int block);
bool valid_block(int block);
// slot groups per lane (1 or 2; the colour sweep only, 1 for the other kernels)
constexpr int kTuneGroupCandidates[] = {1, 2};
int tuned_groups(int kernel, bool fp64, int size_class);
void set_tuned_groups(int kernel, bool fp64, int size_class, int groups);
// planes per lane of the colour sweep's z-column form (1: the plain form; 2 or 4 with one
// slot group per plane; 1 for the other kernels)
constexpr int kTunePlaneCandidates[] = {1, 2, 4};
int tuned_planes(int kernel, bool fp64, int size_class);
void set_tuned_planes(int kernel, bool fp64, int size_class, int planes);

}  // namespace mpr
