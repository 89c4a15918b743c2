// Batched multigrid: device pointer tables and the host-side V-cycle driver that runs the
// V-cycles of many instances (of any depths) with one launch per step and size class.
#pragma once

#include <vector>

#include "mpr_internal.h"

namespace mpr {

constexpr int kMaxCoarseClass = 4;                // coarse solve handles n <= 17
constexpr int kCoarseThreads = 1024;              // launch bound of the coarse kernel
constexpr size_t kCoarseLdsBytes = 160 * 1024;    // LDS per workgroup on gfx950

// one instance's buffers for one V-cycle step at a size class and the next-coarser class
template <class T>
struct StepPtrs {
  T* x;
  const T* rhs;
  const T* screen;
  T* res;
  T* xc;    // next-coarser x
  T* rhsc;  // next-coarser rhs
  int top;  // this size class is the instance's finest level
  // live slots of the finest level (LiveMask, stencil.h; null elsewhere) and whether they
  // also cover its rhs (the assembled right-hand side, not the mixed V-cycle's defect)
  int live_rhs;
  const uint32_t* live;
  // the mixed-precision loop's FP64 solution of an instance at its finest level (null
  // elsewhere): the top level's last post-smoothing sweep adds the correction into it
  double* x64;
  const int* x_initial;  // with x64: the instance's initial-guess flag (TopPtrs)
  // the level's second x buffer (LevelPtrs::xalt): the fused smoothing reads x and writes
  // xalt (null where the level has none)
  T* xalt;
};

// First sweep of a V-cycle level from a known-zero x: 0 x as stored, 1 zero for instances
// below their finest level, 2 zero for every instance
// 3 zero for every instance, and an instance with an x64 pointer (a mixed-precision top
// level) skips the sweep: the defect kernel wrote the sweep's colour-0 values
enum ZeroStart { kZeroNone = 0, kZeroBelowTop = 1, kZeroAll = 2, kZeroAllDefectTop = 3 };

// Initial guess of the solve (core.py:270)
constexpr double kInitialX = 0.5;

// top-level buffers of one instance for the FP64 defect-correction loop
struct TopPtrs {
  double* x64;
  const double* rhs64;
  const double* screen64;
  float* r32;
  float* e32;  // the FP32 correction (the top level's FP32 x)
  const uint32_t* live;  // live slots of rhs64 / screen64 (LiveMask, stencil.h), or null
  // device flag, set for a run that did not fill x64 with kInitialX: the first cycle's
  // defect reads x64 as kInitialX and its fused correction stores kInitialX + e (only for
  // instances whose correction is fused, BatchedVcycle::unfused_class)
  const int* x_initial;
};

// one instance of the LDS-resident coarse V-cycle: its top coarse class, whether that
// level starts from zero, its top x/rhs and the screening weights of every coarse class
template <class T>
struct CoarseEntry {
  int top;
  int zero_top;
  T* x;
  const T* rhs;
  const T* screen[kMaxCoarseClass + 1];
};

// zero_start: the first colour-0 sweep treats x as zero (ZeroStart); the colour-1 sweep
// after it overwrites every colour-1 node, so x needs no zero fill
// correct: the last colour-1 sweep also adds the finished correction of every entry with
// an x64 pointer into it (x64 += e for both colours) and does not store e there (FP32 only);
// first: that is the first cycle's correction (TopPtrs::x_initial)
template <class T>
// skip_first: the first colour-0 sweep is not launched (every instance of the table takes
// kZeroAllDefectTop's skip)
void launch_b_relax(const StepPtrs<T>* tab, uint32_t count, const Geom& g, double pw,
                    int iterations, hipStream_t s, int zero_start = kZeroNone,
                    bool correct = false, bool first = false, bool skip_first = false);
// colour0_only: the residual of colour 0 alone (slot-group layout only); colour1_zero: the
// restriction reads the fine colour 1 as zero (FP32 with the slot-group layout on both grids
// only). Together they serve a grid whose last sweep was a colour-1 sweep, after which the
// colour-1 residual is zero up to rounding.
template <class T>
void launch_b_residual(const StepPtrs<T>* tab, uint32_t count, const Geom& g, double pw,
                       hipStream_t s, bool colour0_only = false);
template <class T>
void launch_b_restrict(const StepPtrs<T>* tab, uint32_t count, const Geom& gf, const Geom& gc,
                       bool zero_coarse, hipStream_t s, bool colour1_zero = false);
// colour1_only: prolong only colour 1, for a level whose next step is a colour-0 sweep
// (which overwrites every colour-0 node without reading it)
template <class T>
void launch_b_prolong(const StepPtrs<T>* tab, uint32_t count, const Geom& gc, const Geom& gf,
                      bool colour1_only, hipStream_t s);
template <class T>
void launch_b_fill_x(const StepPtrs<T>* tab, uint32_t count, const Geom& g, hipStream_t s);
// first: the first cycle's defect (TopPtrs::x_initial)
void launch_b_defect(const TopPtrs* tab, uint32_t count, const Geom& g, double pw, bool first,
                     hipStream_t s, bool e0 = false);
void launch_b_correct(const TopPtrs* tab, uint32_t count, const Geom& g, hipStream_t s);
template <class T>
void launch_b_coarse(const CoarseEntry<T>* tab, uint32_t count, int top_class, double pw, int pre,
                     int post, hipStream_t s);
// uploads the coarse class table; call outside stream capture before the first launch
void init_coarse_kernels();
// Fused smoothing (march.hip): `sweeps` red-black sweeps (2, 4 or 6) of every entry in one
// pass, colours alternating from first_colour, reading x and writing xalt (or, with correct,
// adding the result of both colours into x64 for entries that have one; first: x64 holds
// kInitialX when the entry's x_initial flag is set). zero_start as launch_b_relax's: the
// colour not swept first reads +0.0. Returns false (nothing launched) where the level's shape
// or point_weight < 0 rules it out. Every entry needs an xalt buffer.
bool launch_b_march(const StepPtrs<float>* tab, uint32_t count, const Geom& g, double pw,
                    int sweeps, int first_colour, int zero_start, bool correct, bool first,
                    hipStream_t s);
bool march_supported(const Geom& g, uint32_t count, int sweeps);
// number of finite float numerators a for which the fused smoothing's division by a
// diagonal deg (n-1)^2 (deg = 3..6, a node without screening weight) differs from the IEEE
// quotient; every one of the 2^32 bit patterns is tried for each diagonal
int64_t fast_division_mismatches(int n, hipStream_t s);
// red-black iterations of each smoothing phase of the V-cycle's FP32 levels fused into one
// pass: 0 (none), 1 or 2 (process-wide; tables and graphs built later follow it)
int march_iterations();
void set_march_iterations(int iterations);
// largest coarse class whose LDS-resident hierarchy fits one workgroup
template <class T> int coarse_b_top_class();

// Host description of one instance: levels[k] is size class depth - k (levels[0] is the
// finest grid).
template <class T>
struct InstanceLevels {
  int depth;
  std::vector<LevelPtrs<T>> levels;
  const uint32_t* live = nullptr;  // live slots of levels[0] (LiveMask, stencil.h)
  bool live_rhs = false;           // ... covering levels[0].rhs too
  double* x64 = nullptr;           // FP32 correction levels: the FP64 solution it corrects
  const int* x_initial = nullptr;  // ... and its initial-guess flag (TopPtrs)
};

// highest finest class whose mixed-precision correction is not fused into the FP32 sweep
// (BatchedVcycle<float>::unfused_class)
inline int mixed_unfused_class() {
  const int c = coarse_b_top_class<float>();
  return c > 4 ? c : 4;
}

// Device tables + schedule for the V-cycles of a fixed set of instances.
template <class T>
class BatchedVcycle {
 public:
  // Builds and uploads the tables (synchronously) into `table_mem` (device memory of at
  // least table_bytes(instances) bytes).
  // stage: a pinned host buffer laid out like table_mem, which receives the tables for the
  // caller's asynchronous upload; null: the tables are copied to table_mem synchronously
  void build(const std::vector<InstanceLevels<T>>& inst, void* table_mem,
             unsigned char* stage = nullptr);
  static size_t table_bytes(const std::vector<InstanceLevels<T>>& inst);
  // One V-cycle of every instance (core.py:141-193). zero_tops: every instance's top x
  // starts at zero (the correction equation of the defect-correction loop). fuse_correct
  // (FP32, post > 0): the top level's last post-smoothing sweep adds the correction into
  // the instance's x64 for every instance whose finest class is above unfused_class(); the
  // others need DefectTables::correct(s, unfused_class()). first: this is the first cycle
  // of the defect-correction loop (TopPtrs::x_initial).
  // defect_e0 (with zero_tops, pre > 0, pw >= 0): the defect kernel already wrote each top
  // level's first colour-0 sweep (DefectTables::defect's e0), which is not launched again.
  void run(double pw, int pre, int post, bool zero_tops, hipStream_t s,
           bool fuse_correct = false, bool first = false, bool defect_e0 = false) const;
  // highest finest class without the fused correction: the LDS coarse classes and the
  // levels below 33^3, which the slot-group sweep does not cover
  int unfused_class() const { return sc > 4 ? sc : 4; }
  // the V-cycle's own steps at the finest class alone (for kernel timing): `iterations`
  // red-black sweeps (two launches each), or one residual launch
  void relax_finest(double pw, int iterations, hipStream_t s) const {
    launch_b_relax<T>(step_[smax], step_count_[smax], geom_[smax], pw, iterations, s);
  }
  // as the V-cycle launches it after pre-smoothing
  void residual_finest(double pw, hipStream_t s) const {
    launch_b_residual<T>(step_[smax], step_count_[smax], geom_[smax], pw, s,
                         colour0_residual(smax, pw, 1));
  }
  // whether class c's residual computes colour 0 only and its restriction reads the fine
  // colour 1 as zero: in FP32 after a pre-smoothing (which ends with a colour-1 sweep that
  // zeroes the colour-1 residual up to rounding, its diagonal >= 3 h^-2 > 0 for
  // point_weight >= 0), where both class c and c - 1 have the slot-group layout
  bool colour0_residual(int c, double pw, int pre) const {
    return sizeof(T) == 4 && pre > 0 && pw >= 0.0 && c > sc && group_layout(geom_[c]) &&
           group_layout(geom_[c - 1]);
  }
  // as the V-cycle launches it after pre-smoothing (colour 1 read as zero where the
  // residual computed colour 0 only)
  void restrict_finest(double pw, hipStream_t s) const {
    launch_b_restrict<T>(step_[smax], step_count_[smax], geom_[smax], geom_[smax - 1], false, s,
                         colour0_residual(smax, pw, 1));
  }
  void prolong_finest(bool colour1_only, hipStream_t s) const {
    launch_b_prolong<T>(step_[smax], step_count_[smax], geom_[smax - 1], geom_[smax],
                        colour1_only, s);
  }
  // the fused smoothing of the finest class as the V-cycle's post-smoothing launches it
  // (march_iterations() red-black iterations in one pass, x -> xalt); false where the class
  // does not march
  bool march_finest(double pw, hipStream_t s) const;
  int smax = 0, sc = 0;

 private:
  std::vector<Geom> geom_;                  // by size class
  std::vector<StepPtrs<T>*> step_;          // by class: instances with depth >= class
  std::vector<uint32_t> step_count_;
  std::vector<StepPtrs<T>*> top_;           // by class: instances whose finest class it is
  std::vector<uint32_t> top_count_;
  // by class: step_ with x and xalt exchanged (the level after a fused smoothing phase), and
  // the red-black iterations its smoothing phases fuse (0: none; FP32, every entry with an
  // xalt buffer)
  std::vector<StepPtrs<T>*> swap_;
  std::vector<int> march_;
  CoarseEntry<T>* coarse_ = nullptr;        // all instances
  CoarseEntry<T>* coarse_zero_ = nullptr;   // same, every top zeroed
  uint32_t coarse_count_ = 0;
  int coarse_class_ = 0;
};

// FP64 defect-correction tables: instances grouped by finest class.
class DefectTables {
 public:
  void build(const std::vector<int>& depths, const std::vector<TopPtrs>& tops, void* table_mem,
             unsigned char* stage = nullptr);
  static size_t table_bytes(size_t count) { return sizeof(TopPtrs) * count + 64 * 16; }
  // r32 = b - A x64; first: the first cycle of the loop (TopPtrs::x_initial). e0: also the
  // FP32 V-cycle's first colour-0 sweep of the top level from x = 0, e32 = r32 / diagonal
  // on colour 0, in that sweep's FP32 arithmetic (group layouts)
  void defect(double pw, hipStream_t s, bool first = false, bool e0 = false) const;
  // x64 += e32 for the instances of finest class <= max_class (all by default)
  void correct(hipStream_t s, int max_class = 1 << 30) const;

 private:
  std::vector<Geom> geom_;
  std::vector<TopPtrs*> tab_;
  std::vector<uint32_t> count_;
};

}  // namespace mpr
