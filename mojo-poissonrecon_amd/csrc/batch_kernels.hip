// Batched multigrid kernels: one launch performs one V-cycle step for every instance that
// has a grid of the given size class (n = 2^s + 1). All instances of a class share the
// geometry, so block b of a launch maps to (instance b / nbi, local block b % nbi) with
// one division; each instance's buffers come from a pointer table in device memory.
// Per node the arithmetic is the single-instance element functions, so an instance gets
// bit-identical results alone or inside any batch.
#include <algorithm>

#include "mpr_batch.h"
#include "stencil.h"

namespace mpr {

namespace {

struct BlockMap {
  uint32_t slot, t;  // instance slot, element index inside the instance's work
};

// XCD-aware remap over the padded grid, then instance/local split
__device__ __forceinline__ BlockMap block_map(uint32_t nbi) {
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  BlockMap m;
  m.slot = lb / nbi;
  m.t = (lb - m.slot * nbi) * kThreads + threadIdx.x;
  return m;
}

template <class T>
__global__ __launch_bounds__(kThreads) void kb_relax_color(const StepPtrs<T>* tab, uint32_t count,
                                                           uint32_t nbi, Geom g, int c, T pw, T h2) {
  const BlockMap m = block_map(nbi);
  if (m.slot >= count || m.t >= g.C) return;
  int x, y, z;
  if (!split_decode(g, c, m.t, x, y, z)) return;
  const StepPtrs<T> p = tab[m.slot];
  relax_node(p.x, p.rhs, p.screen, g, (uint32_t)c * g.C + m.t, x, y, z, pw, h2);
}

template <class T>
__global__ __launch_bounds__(kThreads) void kb_residual(const StepPtrs<T>* tab, uint32_t count,
                                                        uint32_t nbi, Geom g, T pw, T h2) {
  const BlockMap m = block_map(nbi);
  if (m.slot >= count || m.t >= 2u * g.C) return;
  const StepPtrs<T> p = tab[m.slot];
  const int c = m.t >= g.C;
  int x, y, z;
  if (split_decode(g, c, m.t - (uint32_t)c * g.C, x, y, z))
    p.res[m.t] = residual_node(p.x, p.rhs, p.screen, g, m.t, x, y, z, pw, h2);
  else
    p.res[m.t] = T(0);
}

// restrict the residual into the next-coarser rhs and zero the next-coarser x
template <class T>
__global__ __launch_bounds__(kThreads) void kb_restrict(const StepPtrs<T>* tab, uint32_t count,
                                                        uint32_t nbi, Geom gf, Geom gc,
                                                        int zero_coarse) {
  const BlockMap m = block_map(nbi);
  if (m.slot >= count || m.t >= 2u * gc.C) return;
  const StepPtrs<T> p = tab[m.slot];
  const int c = m.t >= gc.C;
  int X, Y, Z;
  if (split_decode(gc, c, m.t - (uint32_t)c * gc.C, X, Y, Z))
    p.rhsc[m.t] = restrict_node(p.res, gf, X, Y, Z, T(1));
  if (zero_coarse) p.xc[m.t] = T(0);
}

template <class T>
__global__ __launch_bounds__(kThreads) void kb_prolong_add(const StepPtrs<T>* tab, uint32_t count,
                                                           uint32_t nbi, Geom gc, Geom gf) {
  const BlockMap m = block_map(nbi);
  if (m.slot >= count || m.t >= 2u * gf.C) return;
  const StepPtrs<T> p = tab[m.slot];
  const int c = m.t >= gf.C;
  int x, y, z;
  if (!split_decode(gf, c, m.t - (uint32_t)c * gf.C, x, y, z)) return;
  p.x[m.t] += prolong_node(p.xc, gc, x, y, z);
}

template <class T>
__global__ __launch_bounds__(kThreads) void kb_fill_x(const StepPtrs<T>* tab, uint32_t count,
                                                      uint32_t nbi, uint32_t words) {
  const BlockMap m = block_map(nbi);
  if (m.slot >= count || m.t >= words) return;
  tab[m.slot].x[m.t] = T(0);
}

// defect correction at the top level: r32 = b - A x in FP64, written as FP32
__global__ __launch_bounds__(kThreads) void kb_defect(const TopPtrs* tab, uint32_t count,
                                                      uint32_t nbi, Geom g, double pw, double h2) {
  const BlockMap m = block_map(nbi);
  if (m.slot >= count || m.t >= 2u * g.C) return;
  const TopPtrs p = tab[m.slot];
  const int c = m.t >= g.C;
  int x, y, z;
  if (split_decode(g, c, m.t - (uint32_t)c * g.C, x, y, z))
    p.r32[m.t] = (float)residual_node(p.x64, p.rhs64, p.screen64, g, m.t, x, y, z, pw, h2);
  else
    p.r32[m.t] = 0.0f;
}

__global__ __launch_bounds__(kThreads) void kb_correct(const TopPtrs* tab, uint32_t count,
                                                       uint32_t nbi, uint32_t words) {
  const BlockMap m = block_map(nbi);
  if (m.slot >= count || m.t >= words) return;
  const TopPtrs p = tab[m.slot];
  p.x64[m.t] += (double)p.e32[m.t];
}

template <class T>
__device__ __forceinline__ void wg_relax(T* x, const T* rhs, const T* scr, const Geom& g, T pw,
                                         int iterations) {
  const T h2 = (T)((double)(g.n - 1) * (double)(g.n - 1));
  for (int it = 0; it < iterations; ++it)
    for (int c = 0; c < 2; ++c) {
      for (uint32_t t = threadIdx.x; t < g.C; t += blockDim.x) {
        int xx, yy, zz;
        if (split_decode(g, c, t, xx, yy, zz))
          relax_node(x, rhs, scr, g, (uint32_t)c * g.C + t, xx, yy, zz, pw, h2);
      }
      __syncthreads();
    }
}

// Size class S of the per-instance LDS V-cycle (core.py:141-193). Classes are template
// parameters so geometries and LDS offsets are compile-time-indexed kernel arguments.
template <class T, int S>
__device__ void wg_class(const CoarseArgsB<T>& a, const CoarseEntry<T>& e, T* sm) {
  if constexpr (S >= 1 && S <= kMaxCoarseClass) {
    if (S > e.top) {  // this instance's hierarchy starts below S
      wg_class<T, S - 1>(a, e, sm);
      return;
    }
    const Geom g = a.geom[S];
    const uint32_t m = 2u * g.C;
    T* x = sm + a.off[S];
    T* rhs = x + m;
    T* scr = rhs + m;
    T* res = scr + m;
    if (S == 1) {  // coarsest (n = 3): 40 iterations (core.py:152-156)
      wg_relax(x, rhs, scr, g, a.pw, 40);
      return;
    }
    wg_relax(x, rhs, scr, g, a.pw, a.pre);
    const T h2 = (T)((double)(g.n - 1) * (double)(g.n - 1));
    for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
      const int c = t >= g.C;
      int xx, yy, zz;
      if (split_decode(g, c, t - (uint32_t)c * g.C, xx, yy, zz))
        res[t] = residual_node(x, rhs, scr, g, t, xx, yy, zz, a.pw, h2);
    }
    __syncthreads();
    if constexpr (S > 1) {
      const Geom gc = a.geom[S - 1];
      const uint32_t mc = 2u * gc.C;
      T* xc = sm + a.off[S - 1];
      T* rhsc = xc + mc;
      for (uint32_t t = threadIdx.x; t < mc; t += blockDim.x) {
        const int c = t >= gc.C;
        int X, Y, Z;
        if (split_decode(gc, c, t - (uint32_t)c * gc.C, X, Y, Z))
          rhsc[t] = restrict_node(res, g, X, Y, Z, T(1));
        xc[t] = T(0);
      }
      __syncthreads();
      wg_class<T, S - 1>(a, e, sm);
      for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) {
        const int c = t >= g.C;
        int xx, yy, zz;
        if (split_decode(g, c, t - (uint32_t)c * g.C, xx, yy, zz))
          x[t] += prolong_node(xc, gc, xx, yy, zz);
      }
      __syncthreads();
      wg_relax(x, rhs, scr, g, a.pw, a.post);
    }
  }
}

// one workgroup per instance: the whole V-cycle from the instance's coarse top class down,
// with every level in LDS; only the top x is read (unless zeroed) and written back
template <class T>
__global__ __launch_bounds__(kCoarseThreads) void kb_coarse(CoarseArgsB<T> a,
                                                            const CoarseEntry<T>* tab) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  const CoarseEntry<T> e = tab[blockIdx.x];
  const Geom g0 = a.geom[e.top];
  const uint32_t m0 = 2u * g0.C;
  T* x0 = sm + a.off[e.top];
  for (uint32_t i = threadIdx.x; i < m0; i += blockDim.x) {
    x0[i] = e.zero_top ? T(0) : e.x[i];
    x0[m0 + i] = e.rhs[i];
  }
  for (int s = 1; s <= e.top; ++s) {
    const uint32_t m = 2u * a.geom[s].C;
    T* scr = sm + a.off[s] + 2 * m;
    const T* src = e.screen[s];
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) scr[i] = src[i];
  }
  __syncthreads();
  wg_class<T, kMaxCoarseClass>(a, e, sm);
  for (uint32_t i = threadIdx.x; i < m0; i += blockDim.x) e.x[i] = x0[i];
}

inline unsigned padded8(uint64_t blocks) { return (unsigned)((blocks + 7u) & ~7ull); }

}  // namespace

template <class T>
void launch_b_relax(const StepPtrs<T>* tab, uint32_t count, const Geom& g, double pw,
                    int iterations, hipStream_t s) {
  if (!count || !iterations) return;
  const uint32_t nbi = blocks_for(g.C, kThreads);
  const T h2 = (T)((double)(g.n - 1) * (double)(g.n - 1));
  for (int it = 0; it < iterations; ++it)
    for (int c = 0; c < 2; ++c) {
      hipLaunchKernelGGL(kb_relax_color<T>, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads),
                         0, s, tab, count, nbi, g, c, (T)pw, h2);
      MPR_LAUNCH_CHECK();
    }
}

template <class T>
void launch_b_residual(const StepPtrs<T>* tab, uint32_t count, const Geom& g, double pw,
                       hipStream_t s) {
  if (!count) return;
  const uint32_t nbi = blocks_for(2ull * g.C, kThreads);
  const T h2 = (T)((double)(g.n - 1) * (double)(g.n - 1));
  hipLaunchKernelGGL(kb_residual<T>, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads), 0, s,
                     tab, count, nbi, g, (T)pw, h2);
  MPR_LAUNCH_CHECK();
}

template <class T>
void launch_b_restrict(const StepPtrs<T>* tab, uint32_t count, const Geom& gf, const Geom& gc,
                       bool zero_coarse, hipStream_t s) {
  if (!count) return;
  const uint32_t nbi = blocks_for(2ull * gc.C, kThreads);
  hipLaunchKernelGGL(kb_restrict<T>, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads), 0, s,
                     tab, count, nbi, gf, gc, zero_coarse ? 1 : 0);
  MPR_LAUNCH_CHECK();
}

template <class T>
void launch_b_prolong(const StepPtrs<T>* tab, uint32_t count, const Geom& gc, const Geom& gf,
                      hipStream_t s) {
  if (!count) return;
  const uint32_t nbi = blocks_for(2ull * gf.C, kThreads);
  hipLaunchKernelGGL(kb_prolong_add<T>, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads), 0,
                     s, tab, count, nbi, gc, gf);
  MPR_LAUNCH_CHECK();
}

template <class T>
void launch_b_fill_x(const StepPtrs<T>* tab, uint32_t count, const Geom& g, hipStream_t s) {
  if (!count) return;
  const uint32_t nbi = blocks_for(2ull * g.C, kThreads);
  hipLaunchKernelGGL(kb_fill_x<T>, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads), 0, s,
                     tab, count, nbi, 2u * g.C);
  MPR_LAUNCH_CHECK();
}

void launch_b_defect(const TopPtrs* tab, uint32_t count, const Geom& g, double pw, hipStream_t s) {
  if (!count) return;
  const uint32_t nbi = blocks_for(2ull * g.C, kThreads);
  const double h2 = (double)(g.n - 1) * (double)(g.n - 1);
  hipLaunchKernelGGL(kb_defect, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads), 0, s, tab,
                     count, nbi, g, pw, h2);
  MPR_LAUNCH_CHECK();
}

void launch_b_correct(const TopPtrs* tab, uint32_t count, const Geom& g, hipStream_t s) {
  if (!count) return;
  const uint32_t nbi = blocks_for(2ull * g.C, kThreads);
  hipLaunchKernelGGL(kb_correct, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads), 0, s, tab,
                     count, nbi, 2u * g.C);
  MPR_LAUNCH_CHECK();
}

template <class T>
size_t coarse_b_lds_bytes(int top_class) {
  size_t words = 0;
  for (int s = 1; s <= top_class; ++s) {
    const Geom g = make_geom((1 << s) + 1);
    words += 4ull * 2ull * g.C;
  }
  return words * sizeof(T);
}

template <class T>
int coarse_b_top_class() {
  int s = 1;
  while (s < kMaxCoarseClass && coarse_b_lds_bytes<T>(s + 1) <= kCoarseLdsBytes) ++s;
  return s;
}

template <class T>
void launch_b_coarse(const CoarseEntry<T>* tab, uint32_t count, int top_class, double pw, int pre,
                     int post, hipStream_t s) {
  if (!count) return;
  static bool attr_set = false;
  if (!attr_set) {
    MPR_HIP(hipFuncSetAttribute((const void*)kb_coarse<T>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, kCoarseLdsBytes));
    attr_set = true;
  }
  CoarseArgsB<T> a;
  uint32_t off = 0;
  for (int c = 0; c <= kMaxCoarseClass; ++c) {
    a.geom[c] = make_geom((1 << c) + 1);
    a.off[c] = 0;
  }
  for (int c = 1; c <= top_class; ++c) {
    a.off[c] = off;
    off += 4u * 2u * a.geom[c].C;
  }
  a.pw = (T)pw;
  a.pre = pre;
  a.post = post;
  const size_t bytes = coarse_b_lds_bytes<T>(top_class);
  hipLaunchKernelGGL(kb_coarse<T>, dim3(count), dim3(kCoarseThreads), bytes, s, a, tab);
  MPR_LAUNCH_CHECK();
}

#define MPR_BINST(T)                                                                           \
  template void launch_b_relax<T>(const StepPtrs<T>*, uint32_t, const Geom&, double, int,      \
                                  hipStream_t);                                                \
  template void launch_b_residual<T>(const StepPtrs<T>*, uint32_t, const Geom&, double,        \
                                     hipStream_t);                                             \
  template void launch_b_restrict<T>(const StepPtrs<T>*, uint32_t, const Geom&, const Geom&,   \
                                     bool, hipStream_t);                                       \
  template void launch_b_prolong<T>(const StepPtrs<T>*, uint32_t, const Geom&, const Geom&,    \
                                    hipStream_t);                                              \
  template void launch_b_fill_x<T>(const StepPtrs<T>*, uint32_t, const Geom&, hipStream_t);    \
  template void launch_b_coarse<T>(const CoarseEntry<T>*, uint32_t, int, double, int, int,     \
                                   hipStream_t);                                               \
  template int coarse_b_top_class<T>();
MPR_BINST(double)
MPR_BINST(float)
#undef MPR_BINST

}  // namespace mpr
