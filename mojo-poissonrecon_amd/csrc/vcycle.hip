// Host side of the batched V-cycle: builds the per-class pointer tables for a set of
// instances and issues the schedule (one launch per step and size class).
#include <algorithm>
#include <cstring>

#include "mpr_batch.h"

namespace mpr {

namespace {

struct TableWriter {
  std::vector<unsigned char> host;
  char* device_base;
  template <class E>
  E* put(const std::vector<E>& entries) {
    if (entries.empty()) return nullptr;
    size_t off = (host.size() + 63) & ~size_t(63);
    host.resize(off + sizeof(E) * entries.size());
    std::memcpy(host.data() + off, entries.data(), sizeof(E) * entries.size());
    return reinterpret_cast<E*>(device_base + off);
  }
};

template <class T>
size_t table_words(const std::vector<InstanceLevels<T>>& inst) {
  size_t steps = 0;
  for (const auto& i : inst) steps += (size_t)i.depth + 1;
  return steps;
}

}  // namespace

template <class T>
size_t BatchedVcycle<T>::table_bytes(const std::vector<InstanceLevels<T>>& inst) {
  const size_t n = inst.size();
  return sizeof(StepPtrs<T>) * 3 * table_words(inst) + 2 * sizeof(CoarseEntry<T>) * n +
         64 * 64;
}

// the tables into pinned staging memory (the caller uploads it) or straight to the device
void put_tables(void* table_mem, unsigned char* stage, const std::vector<unsigned char>& host) {
  if (host.empty()) return;
  if (stage)
    std::memcpy(stage, host.data(), host.size());
  else
    MPR_HIP(hipMemcpy(table_mem, host.data(), host.size(), hipMemcpyHostToDevice));
}

template <class T>
void BatchedVcycle<T>::build(const std::vector<InstanceLevels<T>>& inst, void* table_mem,
                             unsigned char* stage) {
  init_coarse_kernels();
  smax = 1;
  for (const auto& i : inst) smax = std::max(smax, i.depth);
  sc = coarse_b_top_class<T>();
  geom_.assign(smax + 1, Geom());
  for (int s = 0; s <= smax; ++s) geom_[s] = make_geom((1 << s) + 1);
  TableWriter w{{}, static_cast<char*>(table_mem)};
  step_.assign(smax + 1, nullptr);
  step_count_.assign(smax + 1, 0);
  top_.assign(smax + 1, nullptr);
  top_count_.assign(smax + 1, 0);
  swap_.assign(smax + 1, nullptr);
  march_.assign(smax + 1, 0);
  for (int s = sc + 1; s <= smax; ++s) {
    std::vector<StepPtrs<T>> steps, tops, swaps;
    bool alt = true;
    for (const auto& i : inst) {
      if (i.depth < s) continue;
      const int k = i.depth - s;
      const LevelPtrs<T>& a = i.levels[k];
      const LevelPtrs<T>& b = i.levels[k + 1];
      const bool top = i.depth == s;
      StepPtrs<T> e{a.x, a.rhs, a.screen, a.res, b.x, b.rhs, top ? 1 : 0,
                    top && i.live_rhs ? 1 : 0, top ? i.live : nullptr, top ? i.x64 : nullptr,
                    top ? i.x_initial : nullptr, a.xalt};
      steps.push_back(e);
      if (i.depth == s) tops.push_back(e);
      alt = alt && a.xalt != nullptr;
      std::swap(e.x, e.xalt);
      swaps.push_back(e);
    }
    step_[s] = w.put(steps);
    step_count_[s] = (uint32_t)steps.size();
    top_[s] = w.put(tops);
    top_count_[s] = (uint32_t)tops.size();
    if constexpr (sizeof(T) == 4) {
      const int it = march_iterations();
      if (alt && it > 0 && march_supported(geom_[s], step_count_[s], 2 * it)) {
        swap_[s] = w.put(swaps);
        march_[s] = it;
      }
    }
  }
  std::vector<CoarseEntry<T>> ce, cz;
  for (const auto& i : inst) {
    CoarseEntry<T> e;
    std::memset(&e, 0, sizeof(e));
    e.top = std::min(i.depth, sc);
    const int k = i.depth - e.top;
    e.x = i.levels[k].x;
    e.rhs = i.levels[k].rhs;
    for (int s = 1; s <= e.top; ++s) e.screen[s] = i.levels[i.depth - s].screen;
    e.zero_top = i.depth > sc ? 1 : 0;
    ce.push_back(e);
    e.zero_top = 1;
    cz.push_back(e);
  }
  coarse_ = w.put(ce);
  coarse_zero_ = w.put(cz);
  coarse_count_ = (uint32_t)ce.size();
  coarse_class_ = 1;
  for (const auto& e : ce) coarse_class_ = std::max(coarse_class_, e.top);
  if (w.host.size() > table_bytes(inst)) throw HipError{"batched V-cycle table overflow"};
  put_tables(table_mem, stage, w.host);
}

template <class T>
void BatchedVcycle<T>::run(double pw, int pre, int post, bool zero_tops, hipStream_t s,
                           bool fuse_correct, bool first, bool defect_e0) const {
  // A level that starts at zero (core.py:174 solution[l+1].fill(0); every level of the
  // correction equation when zero_tops) needs no fill when pre-smoothing runs: the first
  // colour-0 sweep reads the other colour as zero, and every node of that colour has a
  // diagonal >= 3 h^-2 > 0 (point_weight >= 0), so the two colour sweeps write every node
  // before anything reads it. Otherwise the levels are zero-filled as the reference does
  // (the coarse kernel zeroes its own top level).
  const bool fill = pre <= 0 || pw < 0.0;
  if (zero_tops && fill)
    for (int c = sc + 1; c <= smax; ++c) launch_b_fill_x<T>(top_[c], top_count_[c], geom_[c], s);
  const int zero_start =
      fill ? kZeroNone
           : (zero_tops ? (defect_e0 ? kZeroAllDefectTop : kZeroAll) : kZeroBelowTop);
  // classes whose smoothing phases march: march_[c] red-black iterations of the pre-
  // smoothing in one pass from x into xalt, then the level lives in xalt (swap_) until the
  // post-smoothing's pass writes it back; the other iterations are single sweeps in place
  std::vector<char> marched(smax + 1, 0);
  for (int c = smax; c > sc; --c) {
    if (!step_count_[c]) continue;
    // every instance of the class at its top: the defect pass wrote the first sweep
    const bool skip = zero_start == kZeroAllDefectTop && top_count_[c] == step_count_[c];
    if constexpr (sizeof(T) == 4) {
      const int it = march_[c];
      // with the defect pass writing the top levels' first sweep, a class with no top level
      // starts from zero like any other (kZeroAll); one that mixes both does not march
      const int zs = zero_start == kZeroAllDefectTop && top_count_[c] == 0 ? kZeroAll : zero_start;
      if (it && !fill && post >= it && pre >= (skip ? it + 1 : it) &&
          (skip || zs != kZeroAllDefectTop) &&
          launch_b_march(step_[c], step_count_[c], geom_[c], pw, 2 * it, skip ? 1 : 0,
                         skip ? (int)kZeroNone : zs, false, false, s)) {
        marched[c] = 1;
        // the rest: colour 1 first when the defect pass wrote the first colour-0 sweep
        launch_b_relax<T>(swap_[c], step_count_[c], geom_[c], pw, pre - it, s, kZeroNone, false,
                          false, skip);
      }
    }
    if (!marched[c])
      launch_b_relax<T>(step_[c], step_count_[c], geom_[c], pw, pre, s, zero_start, false, false,
                        skip);
    const StepPtrs<T>* tab = marched[c] ? swap_[c] : step_[c];
    // FP32 after pre-smoothing: the colour-1 residual is zero up to rounding, and only
    // colour 0 is computed and restricted
    const bool c0 = colour0_residual(c, pw, pre);
    launch_b_residual<T>(tab, step_count_[c], geom_[c], pw, s, c0);
    launch_b_restrict<T>(tab, step_count_[c], geom_[c], geom_[c - 1], fill && c - 1 > sc, s, c0);
  }
  launch_b_coarse<T>(zero_tops ? coarse_zero_ : coarse_, coarse_count_, coarse_class_, pw, pre,
                     post, s);
  for (int c = sc + 1; c <= smax; ++c) {
    if (!step_count_[c]) continue;
    // The post-smoothing's first colour-0 sweep overwrites every colour-0 node without
    // reading it (its diagonal is >= 3 h^-2 > 0 for point_weight >= 0), so only colour 1
    // needs the correction.
    const StepPtrs<T>* tab = marched[c] ? swap_[c] : step_[c];
    launch_b_prolong<T>(tab, step_count_[c], geom_[c - 1], geom_[c], post > 0 && pw >= 0.0, s);
    int rest = post;
    if constexpr (sizeof(T) == 4) {
      if (marched[c]) {
        // back into x: the colour-0 sweep reads only the prolonged colour 1. With no sweep
        // after it, the pass itself adds the correction into x64.
        rest = post - march_[c];
        if (!launch_b_march(swap_[c], step_count_[c], geom_[c], pw, 2 * march_[c], 0, kZeroNone,
                            fuse_correct && rest == 0, first, s))
          throw HipError{"fused smoothing: post-smoothing launch refused"};
      }
    }
    launch_b_relax<T>(step_[c], step_count_[c], geom_[c], pw, rest, s, kZeroNone,
                      fuse_correct && post > 0, first);
  }
}

template <class T>
bool BatchedVcycle<T>::march_finest(double pw, hipStream_t s) const {
  if constexpr (sizeof(T) == 4) {
    if (march_[smax])
      return launch_b_march(step_[smax], step_count_[smax], geom_[smax], pw, 2 * march_[smax], 0,
                            kZeroNone, false, false, s);
  }
  return false;
}

void DefectTables::build(const std::vector<int>& depths, const std::vector<TopPtrs>& tops,
                         void* table_mem, unsigned char* stage) {
  int dmax = 1;
  for (int d : depths) dmax = std::max(dmax, d);
  geom_.assign(dmax + 1, Geom());
  tab_.assign(dmax + 1, nullptr);
  count_.assign(dmax + 1, 0);
  TableWriter w{{}, static_cast<char*>(table_mem)};
  for (int d = 1; d <= dmax; ++d) {
    std::vector<TopPtrs> e;
    for (size_t i = 0; i < depths.size(); ++i)
      if (depths[i] == d) e.push_back(tops[i]);
    geom_[d] = make_geom((1 << d) + 1);
    tab_[d] = w.put(e);
    count_[d] = (uint32_t)e.size();
  }
  if (w.host.size() > table_bytes(depths.size())) throw HipError{"defect table overflow"};
  put_tables(table_mem, stage, w.host);
}

void DefectTables::defect(double pw, hipStream_t s, bool first, bool e0) const {
  for (size_t d = 1; d < count_.size(); ++d)
    launch_b_defect(tab_[d], count_[d], geom_[d], pw, first, s, e0);
}

void DefectTables::correct(hipStream_t s, int max_class) const {
  for (size_t d = 1; d < count_.size() && (int)d <= max_class; ++d)
    launch_b_correct(tab_[d], count_[d], geom_[d], s);
}

template class BatchedVcycle<double>;
template class BatchedVcycle<float>;

}  // namespace mpr
