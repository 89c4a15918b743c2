// Batched multigrid kernels: one launch performs one V-cycle step for every instance that
// has a grid of the given size class (n = 2^s + 1). All instances of a class share the
// geometry, so block b of a launch maps to (instance b / nbi, local block b % nbi) with
// one division; each instance's buffers come from a pointer table in device memory.
// Per node the arithmetic is the single-instance element functions, so an instance gets
// bit-identical results alone or inside any batch.
#include <algorithm>
#include <mutex>
#include <string>

#include "mpr_batch.h"
#include "mpr_tune.h"
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

// Streaming kernels handle kNodesPerThread nodes per lane (t, t + 256, ...) with every load
// issued before any store (NodeLoads in stencil.h).
constexpr uint32_t kSpan = kThreads * kNodesPerThread;

// whether an instance's x counts as zero in this launch (ZeroStart)
template <class T>
__device__ __forceinline__ bool start_zero(const StepPtrs<T>& p, int zero_start) {
  return zero_start == kZeroAll || zero_start == kZeroAllDefectTop ||
         (zero_start == kZeroBelowTop && !p.top);
}

template <class T>
__global__ __launch_bounds__(kThreads) void kb_relax_color(const StepPtrs<T>* tab, uint32_t count,
                                                           uint32_t nbi, Geom g, int c, T pw, T h2,
                                                           int zero_start) {
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const uint32_t base = (lb - slot * nbi) * kSpan + threadIdx.x;
  const StepPtrs<T> p = tab[slot];
  const gptr<T> X = as_global(p.x);
  const gptr<const T> R = as_global(p.rhs), S = as_global(p.screen);
  const bool zero = start_zero(p, zero_start);
  NodeLoads<T> L[kNodesPerThread];
#pragma unroll
  for (int k = 0; k < kNodesPerThread; ++k)
    load_node(X, R, S, g, c, base + k * kThreads, false, L[k], zero);
#pragma unroll
  for (int k = 0; k < kNodesPerThread; ++k) {
    T v;
    if (loaded_relax(L[k], pw, h2, v)) X[L[k].idx] = v;
  }
}

template <class T, class PO>
__device__ __forceinline__ void residual_span(gptr<const T> X, gptr<const T> R, gptr<const T> S,
                                              PO out, const Geom& g, uint32_t base, T pw, T h2) {
  NodeLoads<T> L[kNodesPerThread];
#pragma unroll
  for (int k = 0; k < kNodesPerThread; ++k) {
    const uint32_t t = base + k * kThreads;
    const int c = t >= g.C;
    load_node(X, R, S, g, c, t - (uint32_t)c * g.C, true, L[k]);
  }
#pragma unroll
  for (int k = 0; k < kNodesPerThread; ++k) {
    const uint32_t t = base + k * kThreads;
    if (t >= 2u * g.C) continue;
    out[t] = L[k].ok ? loaded_residual(L[k], pw, h2) : T(0);
  }
}

template <class T>
__global__ __launch_bounds__(kThreads) void kb_residual(const StepPtrs<T>* tab, uint32_t count,
                                                        uint32_t nbi, Geom g, T pw, T h2) {
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const StepPtrs<T> p = tab[slot];
  residual_span<T>(as_global((const T*)p.x), as_global(p.rhs), as_global(p.screen),
                   as_global(p.res), g, (lb - slot * nbi) * kSpan + threadIdx.x, pw, h2);
}

// Slot-group forms of the sweep, residual and defect for levels with 16-byte aligned rows
// (Geom::H a multiple of 4): one lane per 16 bytes of a colour row (stencil.h).
template <class T>
constexpr int kGroup = 16 / (int)sizeof(T);

// x64 += e for the group's colour-1 slots (e1: this sweep's new values, written slots in
// mask1) and the colour-0 slots at the same positions (their finished values, loaded as
// the sweep's same-row neighbours), as kb_correct does for every slot. x_initial: x64 holds
// the initial guess kInitialX implicitly (it was never filled), so x64 = kInitialX + e is
// stored without a load: the sum x64 += e would give.
template <int V>
__device__ __forceinline__ void add_correction(double* x64, const SlotGroup<float, V>& G,
                                               const Geom& g, const float (&e1)[V],
                                               uint32_t mask1, bool x_initial) {
  if (!G.in) return;
  using D2 = double __attribute__((ext_vector_type(2)));
  gptr<double> X = as_global(x64);
  const uint32_t own1 = G.own, own0 = G.own - g.C;
  // colour-0 node j of this row is at x = x0 + 1 - 2p + 2j (its parity is 1 - p)
  const int x00 = G.x0 + 1 - 2 * G.p;
  uint32_t mask0 = 0;
#pragma unroll
  for (int j = 0; j < V; ++j)
    if (x00 + 2 * j < g.n) mask0 |= 1u << j;
  float e0[V];
#pragma unroll
  for (int j = 0; j < V; ++j) e0[j] = G.row[j / 4][j % 4];
  constexpr uint32_t full = (1u << V) - 1u;
  if (mask0 == full && mask1 == full) {
    D2 a[V / 2], b[V / 2];
#pragma unroll
    for (int i = 0; i < V / 2; ++i) {
      if (x_initial) {
        a[i] = b[i] = D2(kInitialX);
      } else {
        a[i] = *(const __attribute__((address_space(1))) D2*)(X + own0 + 2 * i);
        b[i] = *(const __attribute__((address_space(1))) D2*)(X + own1 + 2 * i);
      }
    }
#pragma unroll
    for (int i = 0; i < V / 2; ++i) {
      a[i][0] += (double)e0[2 * i];
      a[i][1] += (double)e0[2 * i + 1];
      b[i][0] += (double)e1[2 * i];
      b[i][1] += (double)e1[2 * i + 1];
      *(__attribute__((address_space(1))) D2*)(X + own0 + 2 * i) = a[i];
      *(__attribute__((address_space(1))) D2*)(X + own1 + 2 * i) = b[i];
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < V; ++j) {
    if ((mask0 >> j) & 1u) X[own0 + j] = (x_initial ? kInitialX : X[own0 + j]) + (double)e0[j];
    if ((mask1 >> j) & 1u) X[own1 + j] = (x_initial ? kInitialX : X[own1 + j]) + (double)e1[j];
  }
}

// B: workgroup size; NG: consecutive slot groups per lane (all of a lane's loads are issued
// before its arithmetic, so NG = 2 doubles the loads in flight). Both are chosen per
// precision and size class by the tuner (mpr_tune.h). correct (the last colour-1 sweep of
// a mixed-precision top level): entries with an x64 pointer get x64 += e for both colours
// instead of storing e (the next V-cycle starts from zero and never reads it). first: this
// is the first cycle's correction, so an entry whose x_initial flag is set holds kInitialX.
template <class T, int B, int NG>
__global__ __launch_bounds__(B) void kb_relax_group(const StepPtrs<T>* tab, uint32_t count,
                                                    uint32_t nbi, Geom g, uint32_t lanes, int c,
                                                    T pw, T h2, int zero_start, int correct,
                                                    int first) {
  constexpr int V = kGroup<T>;
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const StepPtrs<T> p = tab[slot];
  if (zero_start == kZeroAllDefectTop && p.x64) return;  // written by the defect kernel
  const uint32_t l0 = ((lb - slot * nbi) * B + threadIdx.x) * NG;
  const LiveMask live{p.live, p.live_rhs != 0};
  // the edge slot is loaded, not exchanged between lanes (group_edges): the shuffle waits
  // for the row chunk and measured slower here, though faster in the residual
  SlotGroup<T, V> G[NG];
#pragma unroll
  for (int k = 0; k < NG; ++k)
    load_group<T, V>(as_global((const T*)p.x), as_global(p.rhs), as_global(p.screen), g, c,
                     l0 + (uint32_t)k, lanes, false, G[k], start_zero(p, zero_start), live);
#pragma unroll
  for (int k = 0; k < NG; ++k) {
    T out[V];
    const uint32_t mask = group_relax<T, V>(G[k], g.n, pw, h2, out);
    if constexpr (sizeof(T) == 4) {
      if (correct && p.x64) {
        add_correction<V>(p.x64, G[k], g, out, mask, first && p.x_initial && *p.x_initial != 0);
        continue;
      }
    }
    store_group<T, T, V>(as_global(p.x) + G[k].own, out, mask);
  }
}

// z-column form of the colour sweep (no zero start, no fused correction): each lane relaxes
// the slot groups at one (k0, y) position of NZ consecutive planes z0 .. z0 + NZ - 1. The
// other colour's row chunk of plane z holds the same slots as this colour's nodes in planes
// z - 1 and z + 1 (the x parity alternates with z in both colours), so it is the row
// operand of plane z, the z+1 operand of plane z - 1 and the z-1 operand of plane z + 1:
// a lane loads NZ + 2 such rows instead of 3 NZ. lanes = nH * ceil(n / NZ) / V; lane * V
// decodes as a slot of a grid whose planes are the columns. Measured on 8 instances of
// 257^3: the FP64 sweep is 4-5 % faster with NZ = 4, the FP32 sweep no faster with NZ = 2
// and slower with 4 (half the occupancy); the tuner chooses per precision and class.
template <class T, int B, int NZ>
__global__ __launch_bounds__(B) void kb_relax_column(const StepPtrs<T>* tab, uint32_t count,
                                                     uint32_t nbi, Geom g, uint32_t lanes, int c,
                                                     T pw, T h2) {
  constexpr int V = kGroup<T>;
  using SG = SlotGroup<T, V>;
  constexpr int NC = SG::NC;
  using CT = typename SG::CT;
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const StepPtrs<T> p = tab[slot];
  uint32_t lane = (lb - slot * nbi) * B + threadIdx.x;
  const bool in = lane < lanes;
  lane = in ? lane : lanes - 1;
  const uint32_t s0 = lane * (uint32_t)V;
  const uint32_t zb = g.div_nH.div(s0);
  const uint32_t rr = s0 - zb * g.nH;
  const uint32_t y = g.div_H.div(rr);
  const uint32_t k0 = rr - y * (uint32_t)g.H;
  const int n = g.n;
  const int z0 = (int)zb * NZ;
  const uint32_t base = rr + (uint32_t)z0 * g.nH;  // slot of plane z0 in a colour block
  const uint32_t o = (uint32_t)(1 - c) * g.C + base, own = (uint32_t)c * g.C + base;
  gptr<const T> X = as_global((const T*)p.x);
  gptr<const T> R = as_global(p.rhs), S = as_global(p.screen);
  // planes outside the grid read plane z0 (their values are never used)
  auto plane = [&](int j) -> uint32_t {
    const int z = z0 + j;
    return z >= 0 && z < n ? (uint32_t)((int)g.nH * j) : 0u;
  };
  CT rows[NZ + 2][NC];
#pragma unroll
  for (int j = 0; j < NZ + 2; ++j) load_chunks<T>(rows[j], X + (o + plane(j - 1)));
  SG G[NZ];
  const LiveMask live{p.live, p.live_rhs != 0};
#pragma unroll
  for (int j = 0; j < NZ; ++j) {
    const uint32_t oj = o + plane(j);
    const int pj = (c ^ (int)y ^ (z0 + j)) & 1;
    G[j].own = own + plane(j);
    G[j].x0 = 2 * (int)k0 + pj;
    G[j].y = (int)y;
    G[j].z = z0 + j;
    G[j].p = pj;
    G[j].in = in && z0 + j < n;
    const uint32_t eo = pj == 0 ? (k0 > 0 ? oj - 1 : oj) : (k0 + V < (uint32_t)g.H ? oj + V : oj);
    G[j].edge = X[eo];
    load_chunks<T>(G[j].ym, X + (y > 0 ? oj - (uint32_t)g.H : oj));
    load_chunks<T>(G[j].yp, X + ((int)y + 1 < n ? oj + (uint32_t)g.H : oj));
    const uint32_t ow = G[j].own;
    const bool dead = live.bits && !((live.bits[ow >> 5] >> (ow & 31u)) & ((1u << V) - 1u));
    if (dead && live.rhs) {
#pragma unroll
      for (int i = 0; i < NC; ++i) G[j].r[i] = CT(T(0));
    } else {
      load_chunks<T>(G[j].r, R + ow);
    }
    if (dead) {
#pragma unroll
      for (int i = 0; i < NC; ++i) G[j].s[i] = CT(T(0));
    } else {
      load_chunks<T>(G[j].s, S + ow);
    }
  }
#pragma unroll
  for (int j = 0; j < NZ; ++j) {
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      G[j].zm[i] = rows[j][i];
      G[j].row[i] = rows[j + 1][i];
      G[j].zp[i] = rows[j + 2][i];
    }
    T out[V];
    const uint32_t mask = group_relax<T, V>(G[j], n, pw, h2, out);
    store_group<T, T, V>(as_global(p.x) + G[j].own, out, mask);
  }
}

// residual of slot group `lane` in both colours. One lane takes the same slots of both
// colour blocks, so the words each colour reads as neighbours it also reads as its own
// nodes at the same time (from cache), and x streams from memory once.
// KNOWN: every x word holds xv (load_group); a compile-time choice, so the loading form
// keeps its load schedule.
// C0: colour 0 only; the colour-1 residual words are left as they are (for a grid whose
// last sweep was a colour-1 sweep, where the colour-1 residual is zero up to rounding and
// the restriction reads it as zero).
// The FP32 V-cycle's first colour-0 sweep of a top level from x = 0, for colour-0 slot j of
// a group at row (y, z) with first node x0: e = (r + 0 h2) / (deg h2 + (pw s) h2) with the
// sweep's FP32 operands (r the FP32 defect, s the FP32 screening weight, h2 = (float) h^-2),
// the values group_relax_one computes when every x word is +0.0
__device__ __forceinline__ float first_sweep_value(float r, float s, int x, int y, int z, int n,
                                                   float pw, float h2) {
  const int deg = (int)(x > 0) + (int)(x + 1 < n) + (int)(y > 0) + (int)(y + 1 < n) +
                  (int)(z > 0) + (int)(z + 1 < n);
  const float screening = pw * s * h2;
  const float diag = (float)deg * h2 + screening;
  return (r + 0.0f * h2) / diag;
}

template <class T, class TO, bool KNOWN = false, bool C0 = false, bool E0 = false>
__device__ __forceinline__ void residual_group(gptr<const T> X, gptr<const T> R, gptr<const T> S,
                                               gptr<TO> out, const Geom& g, uint32_t lane,
                                               uint32_t lanes, T pw, T h2, LiveMask live,
                                               T xv = T(0), gptr<float> e0 = nullptr,
                                               float pwf = 0.f, float h2f = 0.f) {
  constexpr int V = kGroup<T>;
  SlotGroup<T, V> G0[1], G1[1];
  load_group<T, V>(X, R, S, g, 0, lane, lanes, true, G0[0], KNOWN, live, false, nullptr, xv);
  // colour 1 at the same slots: its row and own chunks are colour 0's own and row chunks
  if (!C0) load_group<T, V>(X, R, S, g, 1, lane, lanes, true, G1[0], KNOWN, live, false, G0, xv);
  group_edges<T, V, 1>(X, g, 0, G0, KNOWN, xv);
  if (!C0) group_edges<T, V, 1>(X, g, 1, G1, KNOWN, xv);
  T v[V];
  group_residual<T, V>(G0[0], g.n, pw, h2, v);
  store_group<T, TO, V>(out + G0[0].own, v, G0[0].in ? (1u << V) - 1 : 0u);
  if constexpr (E0) {
    float e[V];
    uint32_t mask = 0;
#pragma unroll
    for (int j = 0; j < V; ++j) {
      const int x = G0[0].x0 + 2 * j;
      const float s32 = (float)G0[0].s[j / SlotGroup<T, V>::W][j % SlotGroup<T, V>::W];
      e[j] = first_sweep_value((float)v[j], s32, x, G0[0].y, G0[0].z, g.n, pwf, h2f);
      if (G0[0].in && x < g.n) mask |= 1u << j;
    }
    static_assert(V == 2, "FP64 slot groups: two slots, one 8-byte FP32 store");
    using F2 = float __attribute__((ext_vector_type(2)));
    if (mask == 3u)
      *(__attribute__((address_space(1))) F2*)(e0 + G0[0].own) = F2{e[0], e[1]};
    else if (mask == 1u)
      e0[G0[0].own] = e[0];
  }
  if (C0) return;
  group_residual<T, V>(G1[0], g.n, pw, h2, v);
  store_group<T, TO, V>(out + G1[0].own, v, G1[0].in ? (1u << V) - 1 : 0u);
}

template <class T, int B, bool C0>
__global__ __launch_bounds__(B) void kb_residual_group(const StepPtrs<T>* tab, uint32_t count,
                                                       uint32_t nbi, Geom g, uint32_t lanes, T pw,
                                                       T h2) {
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const StepPtrs<T> p = tab[slot];
  residual_group<T, T, false, C0>(as_global((const T*)p.x), as_global(p.rhs),
                                  as_global(p.screen), as_global(p.res), g,
                                  (lb - slot * nbi) * B + threadIdx.x, lanes, pw, h2,
                                  LiveMask{p.live, p.live_rhs != 0});
}

// restrict the residual into the next-coarser rhs and zero the next-coarser x
template <class T>
__global__ __launch_bounds__(kThreads) void kb_restrict(const StepPtrs<T>* tab, uint32_t count,
                                                        uint32_t nbi, Geom gf, Geom gc,
                                                        int zero_coarse) {
  const BlockMap m = block_map(nbi);
  if (m.slot >= count || m.t >= 2u * gc.C) return;
  const StepPtrs<T> p = tab[m.slot];
  // colours interleaved (neighbouring threads restrict neighbouring coarse nodes, which
  // read overlapping fine nodes of both colours)
  const int c = (int)(m.t & 1u);
  const uint32_t i = (uint32_t)c * gc.C + (m.t >> 1);
  int X, Y, Z;
  if (split_decode(gc, c, m.t >> 1, X, Y, Z)) p.rhsc[i] = restrict_node(p.res, gf, X, Y, Z, T(1));
  if (zero_coarse) p.xc[i] = T(0);
}

// C1Z: the fine colour-1 residual is read as zero (residual_group's C0)
template <class T, bool C1Z>
__global__ __launch_bounds__(kThreads) void kb_restrict_group(const StepPtrs<T>* tab,
                                                              uint32_t count, uint32_t nbi, Geom gf,
                                                              Geom gc, uint32_t lanes,
                                                              int zero_coarse) {
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const StepPtrs<T> p = tab[slot];
  const uint32_t lane = (lb - slot * nbi) * kThreads + threadIdx.x;
  restrict_group<T, kGroup<T>, C1Z>(as_global((const T*)p.res), gf, as_global(p.rhsc),
                                    zero_coarse ? as_global(p.xc) : gptr<T>(nullptr), gc, lane,
                                    lanes, T(1));
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
  p.x[m.t] += prolong_node<T>(p.xc, gc, x, y, z);
}

// slot-group prolong-add of colours c_first .. c_first + ncol - 1 (lanes per colour each)
template <class T>
__global__ __launch_bounds__(kThreads) void kb_prolong_group(const StepPtrs<T>* tab,
                                                             uint32_t count, uint32_t nbi, Geom gc,
                                                             Geom gf, uint32_t lanes, int c_first,
                                                             int ncol, int vec_coarse) {
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const StepPtrs<T> p = tab[slot];
  const uint32_t lane = (lb - slot * nbi) * kThreads + threadIdx.x;
  // lanes past the end keep a valid colour and load clamped (prolong_group stores nothing)
  const int k = (ncol == 2 && lane >= lanes) ? 1 : 0;
  const uint32_t l = lane < (uint32_t)ncol * lanes ? lane - (uint32_t)k * lanes : lanes;
  prolong_group<T, kGroup<T>>(as_global((const T*)p.xc), gc, as_global(p.x), gf, c_first + k,
                              l, lanes, vec_coarse != 0);
}

// FP32 prolong-add of colour 1 (the form before a post-smoothing) by blocks of four fine rows:
// a lane takes the 4-slot chunk k of fine rows (2Y + a, 2Z + b), a, b in {0, 1}. All four
// rows interpolate from the same coarse rows (Y | min(Y+1, nc-1), Z | min(Z+1, nc-1)) over
// the same coarse x = 4k .. 4k + 4 (only tx differs, by the row's x parity), so the coarse
// words are loaded once for 16 fine nodes instead of once per 4. Per node the arithmetic and
// the clamping are prolong_group's (vec_coarse form), in the same order.
__global__ __launch_bounds__(kThreads) void kb_prolong_c1_block(const StepPtrs<float>* tab,
                                                                uint32_t count, uint32_t nbi,
                                                                Geom gc, Geom gf, uint32_t nk,
                                                                uint32_t nb2, uint32_t lanes) {
  using F2 = float __attribute__((ext_vector_type(2)));
  using F4 = float __attribute__((ext_vector_type(4)));
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const StepPtrs<float> p = tab[slot];
  const uint32_t lane = (lb - slot * nbi) * kThreads + threadIdx.x;
  if (lane >= lanes) return;
  const uint32_t k = lane % nk, rest = lane / nk;
  const int Y = (int)(rest % nb2), Z = (int)(rest / nb2);
  const int n = gf.n, nc = gc.n;
  const uint32_t k0 = 4u * k;
  const gptr<const float> XC = as_global((const float*)p.xc);
  const gptr<float> F = as_global(p.x);
  const int y1 = Y + 1 < nc - 1 ? Y + 1 : nc - 1;
  const int z1 = Z + 1 < nc - 1 ? Z + 1 : nc - 1;
  // coarse rows (Y,Z), (y1,Z), (Y,z1), (y1,z1) at coarse x = k0 .. k0 + 4
  float cv[4][5];
  const uint32_t h = k0 >> 1;
  const int m = nc - 1 - (int)k0;  // x = k0 + m is the last coarse node
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int yr = (r & 1) ? y1 : Y, zr = (r & 2) ? z1 : Z;
    const uint32_t row = ((uint32_t)zr * (uint32_t)nc + (uint32_t)yr) * (uint32_t)gc.H;
    const uint32_t ca = (uint32_t)((yr + zr) & 1);  // colour of the even x = k0
    const uint32_t oa = ca * gc.C + row + h, ob = (1u - ca) * gc.C + row + h;
    const F2 a = *(const __attribute__((address_space(1))) F2*)(XC + oa);
    const F2 b = *(const __attribute__((address_space(1))) F2*)(XC + ob);
    const float a2 = XC[oa + 2];
    const float raw[5] = {a[0], b[0], a[1], b[1], a2};
#pragma unroll
    for (int i = 0; i <= 4; ++i) {
      float v = raw[i];
#pragma unroll
      for (int q = 0; q < i; ++q)
        if (i > m && q == m) v = raw[q];
      cv[r][i] = v;
    }
  }
  F4 xf[4];
  uint32_t own[4];
  bool in[4];
#pragma unroll
  for (int ab = 0; ab < 4; ++ab) {
    const int y = 2 * Y + (ab & 1), z = 2 * Z + (ab >> 1);
    in[ab] = y < n && z < n;
    own[ab] = gf.C + ((uint32_t)(in[ab] ? z : 0) * (uint32_t)n + (uint32_t)(in[ab] ? y : 0)) *
                         (uint32_t)gf.H + k0;
    xf[ab] = *(const __attribute__((address_space(1))) F4*)(F + own[ab]);
  }
#pragma unroll
  for (int ab = 0; ab < 4; ++ab) {
    if (!in[ab]) continue;
    const int y = 2 * Y + (ab & 1), z = 2 * Z + (ab >> 1);
    const int px = (1 ^ y ^ z) & 1;
    const float tx = px ? 0.5f : 0.0f, ty = (y & 1) ? 0.5f : 0.0f, tz = (z & 1) ? 0.5f : 0.0f;
    float out[4];
    uint32_t mask = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float c00 = cv[0][j] * (1.0f - tx) + cv[0][j + 1] * tx;
      const float c10 = cv[1][j] * (1.0f - tx) + cv[1][j + 1] * tx;
      const float c01 = cv[2][j] * (1.0f - tx) + cv[2][j + 1] * tx;
      const float c11 = cv[3][j] * (1.0f - tx) + cv[3][j + 1] * tx;
      const float c0 = c00 * (1.0f - ty) + c10 * ty;
      const float c1 = c01 * (1.0f - ty) + c11 * ty;
      out[j] = xf[ab][j] + (c0 * (1.0f - tz) + c1 * tz);
      if (2 * ((int)k0 + j) + px < n) mask |= 1u << j;
    }
    store_group<float, float, 4>(F + own[ab], out, mask);
  }
}

template <class T>
__global__ __launch_bounds__(kThreads) void kb_fill_x(const StepPtrs<T>* tab, uint32_t count,
                                                      uint32_t nbi, uint32_t words) {
  const BlockMap m = block_map(nbi);
  if (m.slot >= count || m.t >= words) return;
  tab[m.slot].x[m.t] = T(0);
}

// defect correction at the top level: r32 = b - A x in FP64, written as FP32
struct ToFloat {  // stores a double as float through a global pointer
  gptr<float> p;
  struct Ref {
    gptr<float> q;
    __device__ void operator=(double v) const { *q = (float)v; }
  };
  __device__ Ref operator[](uint32_t i) const { return Ref{p + i}; }
};

__global__ __launch_bounds__(kThreads) void kb_defect(const TopPtrs* tab, uint32_t count,
                                                      uint32_t nbi, Geom g, double pw, double h2) {
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const TopPtrs p = tab[slot];
  residual_span<double>(as_global((const double*)p.x64), as_global(p.rhs64),
                        as_global(p.screen64), ToFloat{as_global(p.r32)}, g,
                        (lb - slot * nbi) * kSpan + threadIdx.x, pw, h2);
}

// The first defect of an entry whose x_initial flag is set: x holds kInitialX everywhere, so
// a node needs only its own rhs and screening words. The arithmetic is residual_group's with
// x known, in the same order: the neighbour sum of deg copies of kInitialX (0.5) is
// deg * kInitialX exactly, and r = rhs - ((deg h2 + pw s h2) x - sum h2). Lanes as
// residual_group<double>: the same 2 slots of both colours.
__device__ __forceinline__ void initial_defect_group(gptr<const double> R, gptr<const double> S,
                                                     gptr<float> out, const Geom& g,
                                                     uint32_t lane, uint32_t lanes, double pw,
                                                     double h2, const uint32_t* live,
                                                     gptr<float> e0, float pwf, float h2f) {
  constexpr int V = kGroup<double>;
  static_assert(V == 2, "one double2 per colour and lane");
  if (lane >= lanes) return;
  const uint32_t s0 = lane * (uint32_t)V;
  const uint32_t z = g.div_nH.div(s0);
  const uint32_t rr = s0 - z * g.nH;
  const uint32_t y = g.div_H.div(rr);
  const uint32_t k0 = rr - y * (uint32_t)g.H;
  const int n = g.n;
  const int dyz = (int)(y > 0) + (int)((int)y + 1 < n) + (int)(z > 0) + (int)((int)z + 1 < n);
  using D2 = typename Chunk16<double>::type;
  using F2 = float __attribute__((ext_vector_type(2)));
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const uint32_t own = (uint32_t)c * g.C + s0;
    const int p = (c ^ (int)y ^ (int)z) & 1;
    // a group without a live slot has rhs and screen exactly +0.0 (LiveMask)
    const bool dead = live && !((live[own >> 5] >> (own & 31u)) & ((1u << V) - 1u));
    D2 rv = D2(0.0), sv = D2(0.0);
    if (!dead) {
      rv = *(const __attribute__((address_space(1))) D2*)(R + own);
      sv = *(const __attribute__((address_space(1))) D2*)(S + own);
    }
    float v[V];
#pragma unroll
    for (int j = 0; j < V; ++j) {
      const int x = 2 * ((int)k0 + j) + p;
      const int deg = (int)(x > 0) + (int)(x + 1 < n) + dyz;
      const double screening = pw * (j ? sv.y : sv.x) * h2;
      const double applied =
          ((double)deg * h2 + screening) * kInitialX - ((double)deg * kInitialX) * h2;
      v[j] = x < n ? (float)((j ? rv.y : rv.x) - applied) : 0.f;
    }
    *(__attribute__((address_space(1))) F2*)(out + own) = F2{v[0], v[1]};
    if (e0 && c == 0) {  // the first colour-0 sweep (residual_group's E0)
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const int x = 2 * ((int)k0 + j) + p;
        if (x < n)
          e0[own + j] = first_sweep_value(v[j], (float)(j ? sv.y : sv.x), x, (int)y, (int)z, n,
                                          pwf, h2f);
      }
    }
  }
}

// The first cycle's defect of the entries whose x_initial flag is set (x64 was never filled
// and holds kInitialX everywhere); a kernel of its own, so that its few registers allow the
// occupancy that hides its two dependent loads
__global__ __launch_bounds__(kThreads) void kb_defect_initial(const TopPtrs* tab, uint32_t count,
                                                              uint32_t nbi, Geom g, uint32_t lanes,
                                                              double pw, double h2, int e0,
                                                              float pwf, float h2f) {
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const TopPtrs p = tab[slot];
  if (!p.x_initial || *p.x_initial == 0) return;
  initial_defect_group(as_global(p.rhs64), as_global(p.screen64), as_global(p.r32), g,
                       (lb - slot * nbi) * kThreads + threadIdx.x, lanes, pw, h2, p.live,
                       e0 ? as_global(p.e32) : gptr<float>(nullptr), pwf, h2f);
}

// FIRST: the first cycle's defect; entries whose x_initial flag is set are left to
// kb_defect_initial
template <int B, bool FIRST, bool E0>
__global__ __launch_bounds__(B) void kb_defect_group(const TopPtrs* tab, uint32_t count,
                                                     uint32_t nbi, Geom g, uint32_t lanes,
                                                     double pw, double h2, float pwf,
                                                     float h2f) {
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t slot = lb / nbi;
  if (slot >= count) return;
  const TopPtrs p = tab[slot];
  const uint32_t lane = (lb - slot * nbi) * B + threadIdx.x;
  if (FIRST && p.x_initial && *p.x_initial != 0) return;
  residual_group<double, float, false, false, E0>(
      as_global((const double*)p.x64), as_global(p.rhs64), as_global(p.screen64),
      as_global(p.r32), g, lane, lanes, pw, h2, LiveMask{p.live, true}, 0.0,
      as_global(p.e32), pwf, h2f);
}

__global__ __launch_bounds__(kThreads) void kb_correct(const TopPtrs* tab, uint32_t count,
                                                       uint32_t nbi, uint32_t words) {
  const BlockMap m = block_map(nbi);
  if (m.slot >= count || m.t >= words) return;
  const TopPtrs p = tab[m.slot];
  p.x64[m.t] += (double)p.e32[m.t];
}

// Geometry and LDS word offset of every coarse size class (n = 2^s + 1). Offsets count
// elements from class 1 upwards, so they are the same for float and double and for every
// top class; the table lives in constant memory and is read with uniform (scalar) loads.
struct CoarseClassInfo {
  Geom g;
  uint32_t off = 0;
};
__constant__ CoarseClassInfo g_coarse_info[kMaxCoarseClass + 1];

// Per-node constants of one smoothing call on an LDS-resident level: the LDS indices of the
// six neighbours (the node's own slot where one is absent), which of them exist, the rhs
// and the diagonal deg h^-2 + (alpha screen) h^-2 -- all fixed while x is iterated, so
// they are computed once per call instead of once per sweep.
template <class T>
struct CoarseNode {
  uint32_t idx, nb[6], flags;
  T r, diag;
  bool live;  // a real node whose update is applied (diag > 1e-30, as relax_node)
};

template <class T>
__device__ __forceinline__ void coarse_node(const T* rhs, const T* scr, const Geom& g, int c,
                                            uint32_t t, T pw, T h2, CoarseNode<T>& N) {
  int x = 0, y = 0, z = 0;
  const bool in = t < g.C;
  const bool ok = in && split_decode(g, c, in ? t : 0u, x, y, z);
  const int n = g.n;
  const bool bx0 = x > 0, bx1 = x + 1 < n, by0 = y > 0, by1 = y + 1 < n, bz0 = z > 0,
             bz1 = z + 1 < n;
  const uint32_t self = (uint32_t)c * g.C + (in ? t : 0u);
  N.idx = self;
  N.nb[0] = bx0 ? split_index(g, x - 1, y, z) : self;
  N.nb[1] = bx1 ? split_index(g, x + 1, y, z) : self;
  N.nb[2] = by0 ? split_index(g, x, y - 1, z) : self;
  N.nb[3] = by1 ? split_index(g, x, y + 1, z) : self;
  N.nb[4] = bz0 ? split_index(g, x, y, z - 1) : self;
  N.nb[5] = bz1 ? split_index(g, x, y, z + 1) : self;
  N.flags = (uint32_t)bx0 | ((uint32_t)bx1 << 1) | ((uint32_t)by0 << 2) | ((uint32_t)by1 << 3) |
            ((uint32_t)bz0 << 4) | ((uint32_t)bz1 << 5);
  const int deg = (int)bx0 + (int)bx1 + (int)by0 + (int)by1 + (int)bz0 + (int)bz1;
  const T screening = pw * scr[self] * h2;
  N.diag = (T)deg * h2 + screening;
  N.r = rhs[self];
  N.live = ok && N.diag > T(1.0e-30);
}

// relax_node's update of one node from its constants
template <class T>
__device__ __forceinline__ void coarse_update(T* x, const CoarseNode<T>& N, T h2) {
  if (!N.live) return;
  T v[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) v[j] = x[N.nb[j]];
  T s = T(0);
#pragma unroll
  for (int j = 0; j < 6; ++j)
    if ((N.flags >> j) & 1u) s += v[j];
  x[N.idx] = (N.r + s * h2) / N.diag;
}

// Red-black sweeps of an LDS-resident level by the workgroup (capi.mojo:309-410). Levels
// with at most 128 slots per colour (n <= 5) are swept by the first wavefront alone, with
// wavefront-level ordering between the colours instead of workgroup barriers; larger ones
// by all 1024 lanes, at most three nodes per lane and colour (n <= 17).
template <class T>
__device__ __forceinline__ void wg_relax(T* x, const T* rhs, const T* scr, const Geom& g, T pw,
                                         int iterations) {
  constexpr int kPer = 3;
  const T h2 = (T)((double)(g.n - 1) * (double)(g.n - 1));
  if (g.C <= 2u * 64u) {
    if (threadIdx.x < 64) {
      CoarseNode<T> N[2][2];
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < 2; ++k) coarse_node(rhs, scr, g, c, threadIdx.x + 64u * k, pw, h2, N[c][k]);
#pragma unroll 1
      for (int it = 0; it < iterations; ++it)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
#pragma unroll
          for (int k = 0; k < 2; ++k) coarse_update(x, N[c][k], h2);
          // one wavefront: its LDS accesses complete in program order
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    __syncthreads();
    return;
  }
  {
    // n <= 17 (kMaxCoarseClass 4): C <= 17 * 17 * 9 = 2601 <= kPer * kCoarseThreads
    static_assert(kMaxCoarseClass <= 4, "coarse levels above n = 17 need more nodes per lane");
    CoarseNode<T> N[2][kPer];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int k = 0; k < kPer; ++k)
        coarse_node(rhs, scr, g, c, threadIdx.x + (uint32_t)kCoarseThreads * k, pw, h2, N[c][k]);
#pragma unroll 1
    for (int it = 0; it < iterations; ++it)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) coarse_update(x, N[c][k], h2);
        __syncthreads();
      }
  }
}

// One workgroup per instance: the instance's V-cycle from its top coarse class down
// (core.py:141-193 as a down loop and an up loop), every level resident in LDS. Only the
// top x (unless it starts at zero), the top rhs and the screening weights are read from
// HBM, and only the top x is written back.
template <class T>
__global__ __launch_bounds__(kCoarseThreads) void kb_coarse(const CoarseEntry<T>* tab, T pw,
                                                            int pre, int post) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  const CoarseEntry<T> e = tab[blockIdx.x];
  const int top = e.top;
  {
    const uint32_t m0 = 2u * g_coarse_info[top].g.C;
    T* x0 = sm + g_coarse_info[top].off;
    for (uint32_t i = threadIdx.x; i < m0; i += kCoarseThreads) {
      x0[i] = e.zero_top ? T(0) : e.x[i];
      x0[m0 + i] = e.rhs[i];
    }
#pragma unroll 1
    for (int s = 1; s <= top; ++s) {
      const uint32_t m = 2u * g_coarse_info[s].g.C;
      T* scr = sm + g_coarse_info[s].off + 2 * m;
      const T* src = e.screen[s];
      for (uint32_t i = threadIdx.x; i < m; i += kCoarseThreads) scr[i] = src[i];
    }
  }
  __syncthreads();
#pragma unroll 1
  for (int s = top; s >= 2; --s) {  // pre-smooth, residual, restrict, zero the coarser x
    const Geom g = g_coarse_info[s].g;
    const uint32_t m = 2u * g.C;
    T* x = sm + g_coarse_info[s].off;
    T* rhs = x + m;
    T* scr = rhs + m;
    T* res = scr + m;
    wg_relax(x, rhs, scr, g, pw, pre);
    const T h2 = (T)((double)(g.n - 1) * (double)(g.n - 1));
#pragma unroll 1
    for (uint32_t t = threadIdx.x; t < m; t += kCoarseThreads) {
      const int c = t >= g.C;
      int xx, yy, zz;
      if (split_decode(g, c, t - (uint32_t)c * g.C, xx, yy, zz))
        res[t] = residual_node(x, rhs, scr, g, t, xx, yy, zz, pw, h2);
    }
    __syncthreads();
    const Geom gc = g_coarse_info[s - 1].g;
    const uint32_t mc = 2u * gc.C;
    T* xc = sm + g_coarse_info[s - 1].off;
    T* rhsc = xc + mc;
#pragma unroll 1
    for (uint32_t t = threadIdx.x; t < mc; t += kCoarseThreads) {
      const int c = t >= gc.C;
      int X, Y, Z;
      if (split_decode(gc, c, t - (uint32_t)c * gc.C, X, Y, Z))
        rhsc[t] = restrict_node(res, g, X, Y, Z, T(1));
      xc[t] = T(0);
    }
    __syncthreads();
  }
  {  // coarsest (n = 3): 40 iterations (core.py:152-156)
    const Geom g = g_coarse_info[1].g;
    const uint32_t m = 2u * g.C;
    T* x = sm + g_coarse_info[1].off;
    wg_relax(x, x + m, x + 2 * m, g, pw, 40);
  }
#pragma unroll 1
  for (int s = 2; s <= top; ++s) {  // prolong-add, post-smooth
    const Geom g = g_coarse_info[s].g;
    const Geom gc = g_coarse_info[s - 1].g;
    const uint32_t m = 2u * g.C;
    T* x = sm + g_coarse_info[s].off;
    const T* xc = sm + g_coarse_info[s - 1].off;
#pragma unroll 1
    for (uint32_t t = threadIdx.x; t < m; t += kCoarseThreads) {
      const int c = t >= g.C;
      int xx, yy, zz;
      if (split_decode(g, c, t - (uint32_t)c * g.C, xx, yy, zz))
        x[t] += prolong_node<T>(xc, gc, xx, yy, zz);
    }
    __syncthreads();
    wg_relax(x, x + m, x + 2 * m, g, pw, post);
  }
  const uint32_t m0 = 2u * g_coarse_info[top].g.C;
  const T* x0 = sm + g_coarse_info[top].off;
  for (uint32_t i = threadIdx.x; i < m0; i += kCoarseThreads) e.x[i] = x0[i];
}

inline unsigned padded8(uint64_t blocks) { return (unsigned)((blocks + 7u) & ~7ull); }

// launch KERNEL<..., B> with the tuned workgroup size B (one of kTuneCandidates)
#define MPR_WITH_BLOCK(BLOCK, LAUNCH) \
  switch (BLOCK) {                    \
    case 64: LAUNCH(64); break;       \
    case 128: LAUNCH(128); break;     \
    case 512: LAUNCH(512); break;     \
    case 1024: LAUNCH(1024); break;   \
    default: LAUNCH(256); break;      \
  }

}  // namespace

template <class T>
void launch_b_relax(const StepPtrs<T>* tab, uint32_t count, const Geom& g, double pw,
                    int iterations, hipStream_t s, int zero_start, bool correct, bool first,
                    bool skip_first) {
  if (!count || !iterations) return;
  const T h2 = (T)((double)(g.n - 1) * (double)(g.n - 1));
  if (group_layout(g)) {
    const uint32_t lanes = g.C / (uint32_t)kGroup<T>;
    const int block = tuned_block(kTuneRelax, sizeof(T) == 8, size_class(g));
    const int ng = tuned_groups(kTuneRelax, sizeof(T) == 8, size_class(g));
    const int nz = tuned_planes(kTuneRelax, sizeof(T) == 8, size_class(g));
    const uint32_t nbi = blocks_for((lanes + ng - 1) / ng, (unsigned)block);
    const uint32_t col_lanes =
        g.nH * (((uint32_t)g.n + (uint32_t)nz - 1) / (uint32_t)nz) / (uint32_t)kGroup<T>;
    const uint32_t col_nbi = blocks_for(col_lanes, (unsigned)block);
    for (int it = 0; it < iterations; ++it)
      for (int c = 0; c < 2; ++c) {
        if (skip_first && it == 0 && c == 0) continue;
        const int zs = it == 0 && c == 0 ? zero_start : (int)kZeroNone;
        const int corr = correct && sizeof(T) == 4 && it == iterations - 1 && c == 1 ? 1 : 0;
        if (nz > 1 && zs == kZeroNone && !corr) {
#define MPR_LAUNCH_COLUMN(B, NZ)                                                               \
  hipLaunchKernelGGL((kb_relax_column<T, B, NZ>), dim3(padded8((uint64_t)col_nbi * count)),    \
                     dim3(B), 0, s, tab, count, col_nbi, g, col_lanes, c, (T)pw, h2)
#define MPR_LAUNCH_COLUMN2(B) MPR_LAUNCH_COLUMN(B, 2)
#define MPR_LAUNCH_COLUMN4(B) MPR_LAUNCH_COLUMN(B, 4)
          if (nz == 2) {
            MPR_WITH_BLOCK(block, MPR_LAUNCH_COLUMN2)
          } else {
            MPR_WITH_BLOCK(block, MPR_LAUNCH_COLUMN4)
          }
#undef MPR_LAUNCH_COLUMN4
#undef MPR_LAUNCH_COLUMN2
#undef MPR_LAUNCH_COLUMN
          MPR_LAUNCH_CHECK();
          continue;
        }
#define MPR_LAUNCH_RELAX_NG(B, NG)                                                 \
  hipLaunchKernelGGL((kb_relax_group<T, B, NG>), dim3(padded8((uint64_t)nbi * count)), \
                     dim3(B), 0, s, tab, count, nbi, g, lanes, c, (T)pw, h2, zs, corr, \
                     first ? 1 : 0)
#define MPR_LAUNCH_RELAX1(B) MPR_LAUNCH_RELAX_NG(B, 1)
#define MPR_LAUNCH_RELAX2(B) MPR_LAUNCH_RELAX_NG(B, 2)
        if (ng == 2) {
          MPR_WITH_BLOCK(block, MPR_LAUNCH_RELAX2)
        } else {
          MPR_WITH_BLOCK(block, MPR_LAUNCH_RELAX1)
        }
#undef MPR_LAUNCH_RELAX1
#undef MPR_LAUNCH_RELAX2
#undef MPR_LAUNCH_RELAX_NG
        MPR_LAUNCH_CHECK();
      }
    return;
  }
  const uint32_t nbi = blocks_for(g.C, kSpan);
  for (int it = 0; it < iterations; ++it)
    for (int c = 0; c < 2; ++c) {
      if (skip_first && it == 0 && c == 0) continue;
      hipLaunchKernelGGL(kb_relax_color<T>, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads),
                         0, s, tab, count, nbi, g, c, (T)pw, h2,
                         it == 0 && c == 0 ? zero_start : (int)kZeroNone);
      MPR_LAUNCH_CHECK();
    }
}

template <class T>
void launch_b_residual(const StepPtrs<T>* tab, uint32_t count, const Geom& g, double pw,
                       hipStream_t s, bool colour0_only) {
  if (!count) return;
  const T h2 = (T)((double)(g.n - 1) * (double)(g.n - 1));
  if (group_layout(g)) {
    const uint32_t lanes = g.C / (uint32_t)kGroup<T>;
    const int block = tuned_block(kTuneResidual, sizeof(T) == 8, size_class(g));
    const uint32_t nbi = blocks_for(lanes, (unsigned)block);  // both colours per lane
#define MPR_LAUNCH_RESIDUAL(B)                                                               \
  if (colour0_only)                                                                          \
    hipLaunchKernelGGL((kb_residual_group<T, B, true>), dim3(padded8((uint64_t)nbi * count)),  \
                       dim3(B), 0, s, tab, count, nbi, g, lanes, (T)pw, h2);                   \
  else                                                                                       \
    hipLaunchKernelGGL((kb_residual_group<T, B, false>), dim3(padded8((uint64_t)nbi * count)), \
                       dim3(B), 0, s, tab, count, nbi, g, lanes, (T)pw, h2)
    MPR_WITH_BLOCK(block, MPR_LAUNCH_RESIDUAL)
#undef MPR_LAUNCH_RESIDUAL
    MPR_LAUNCH_CHECK();
    return;
  }
  const uint32_t nbi = blocks_for(2ull * g.C, kSpan);
  hipLaunchKernelGGL(kb_residual<T>, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads), 0, s,
                     tab, count, nbi, g, (T)pw, h2);
  MPR_LAUNCH_CHECK();
}

template <class T>
void launch_b_restrict(const StepPtrs<T>* tab, uint32_t count, const Geom& gf, const Geom& gc,
                       bool zero_coarse, hipStream_t s, bool colour1_zero) {
  if (!count) return;
  if (sizeof(T) == 4 && group_layout(gc) && group_layout(gf)) {  // as launch_restrict
    const uint32_t lanes = gc.C / (uint32_t)kGroup<T>;
    const uint32_t nbi = blocks_for(lanes, kThreads);  // both coarse colours per lane
    if (colour1_zero)
      hipLaunchKernelGGL((kb_restrict_group<T, true>), dim3(padded8((uint64_t)nbi * count)),
                         dim3(kThreads), 0, s, tab, count, nbi, gf, gc, lanes, zero_coarse ? 1 : 0);
    else
      hipLaunchKernelGGL((kb_restrict_group<T, false>), dim3(padded8((uint64_t)nbi * count)),
                         dim3(kThreads), 0, s, tab, count, nbi, gf, gc, lanes, zero_coarse ? 1 : 0);
    MPR_LAUNCH_CHECK();
    return;
  }
  if (colour1_zero) throw HipError{"colour-1-zero restriction needs the slot-group layout"};
  const uint32_t nbi = blocks_for(2ull * gc.C, kThreads);
  hipLaunchKernelGGL(kb_restrict<T>, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads), 0, s,
                     tab, count, nbi, gf, gc, zero_coarse ? 1 : 0);
  MPR_LAUNCH_CHECK();
}

template <class T>
void launch_b_prolong(const StepPtrs<T>* tab, uint32_t count, const Geom& gc, const Geom& gf,
                      bool colour1_only, hipStream_t s) {
  if (!count) return;
  if constexpr (sizeof(T) == 4) {
    if (colour1_only && group_layout(gf) && group_layout(gc)) {
      const uint32_t nk = (uint32_t)gf.H / 4u, nb2 = (uint32_t)gc.n;
      const uint32_t lanes = nk * nb2 * nb2;
      const uint32_t nbi = blocks_for(lanes, kThreads);
      hipLaunchKernelGGL(kb_prolong_c1_block, dim3(padded8((uint64_t)nbi * count)),
                         dim3(kThreads), 0, s, tab, count, nbi, gc, gf, nk, nb2, lanes);
      MPR_LAUNCH_CHECK();
      return;
    }
  }
  if (group_layout(gf)) {
    const uint32_t lanes = gf.C / (uint32_t)kGroup<T>;
    const int ncol = colour1_only ? 1 : 2;
    const uint32_t nbi = blocks_for((uint64_t)ncol * lanes, kThreads);
    hipLaunchKernelGGL(kb_prolong_group<T>, dim3(padded8((uint64_t)nbi * count)), dim3(kThreads),
                       0, s, tab, count, nbi, gc, gf, lanes, colour1_only ? 1 : 0, ncol,
                       group_layout(gc) ? 1 : 0);
    MPR_LAUNCH_CHECK();
    return;
  }
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

void launch_b_defect(const TopPtrs* tab, uint32_t count, const Geom& g, double pw, bool first,
                     hipStream_t s, bool e0) {
  if (!count) return;
  const double h2 = (double)(g.n - 1) * (double)(g.n - 1);
  // the FP32 sweep's own operands (launch_b_relax<float>)
  const float pwf = (float)pw, h2f = (float)h2;
  if (group_layout(g)) {
    const uint32_t lanes = g.C / (uint32_t)kGroup<double>;
    const int block = tuned_block(kTuneDefect, true, size_class(g));
    const uint32_t nbi = blocks_for(lanes, (unsigned)block);  // both colours per lane
#define MPR_LAUNCH_DEFECT_F(B, F, E)                                                            \
  hipLaunchKernelGGL((kb_defect_group<B, F, E>), dim3(padded8((uint64_t)nbi * count)), dim3(B), \
                     0, s, tab, count, nbi, g, lanes, pw, h2, pwf, h2f)
#define MPR_LAUNCH_DEFECT1E(B) MPR_LAUNCH_DEFECT_F(B, true, true)
#define MPR_LAUNCH_DEFECT0E(B) MPR_LAUNCH_DEFECT_F(B, false, true)
#define MPR_LAUNCH_DEFECT1(B) MPR_LAUNCH_DEFECT_F(B, true, false)
#define MPR_LAUNCH_DEFECT0(B) MPR_LAUNCH_DEFECT_F(B, false, false)
    if (first) {
      hipLaunchKernelGGL(kb_defect_initial, dim3(padded8((uint64_t)blocks_for(lanes, kThreads) *
                                                         count)),
                         dim3(kThreads), 0, s, tab, count, blocks_for(lanes, kThreads), g, lanes,
                         pw, h2, e0 ? 1 : 0, pwf, h2f);
      MPR_LAUNCH_CHECK();
      if (e0) {
        MPR_WITH_BLOCK(block, MPR_LAUNCH_DEFECT1E)
      } else {
        MPR_WITH_BLOCK(block, MPR_LAUNCH_DEFECT1)
      }
    } else {
      if (e0) {
        MPR_WITH_BLOCK(block, MPR_LAUNCH_DEFECT0E)
      } else {
        MPR_WITH_BLOCK(block, MPR_LAUNCH_DEFECT0)
      }
    }
#undef MPR_LAUNCH_DEFECT0
#undef MPR_LAUNCH_DEFECT1
#undef MPR_LAUNCH_DEFECT0E
#undef MPR_LAUNCH_DEFECT1E
#undef MPR_LAUNCH_DEFECT_F
    MPR_LAUNCH_CHECK();
    return;
  }
  const uint32_t nbi = blocks_for(2ull * g.C, kSpan);
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
  const size_t bytes = coarse_b_lds_bytes<T>(top_class);
  hipLaunchKernelGGL(kb_coarse<T>, dim3(count), dim3(kCoarseThreads), bytes, s, tab, (T)pw, pre,
                     post);
  MPR_LAUNCH_CHECK();
}

// Uploads the coarse class table and raises the coarse kernels' LDS limit. Runs once per
// process, outside any stream capture (from the table builders).
void init_coarse_kernels() {
  static std::once_flag once;
  static std::string failure;
  std::call_once(once, [] {
    CoarseClassInfo info[kMaxCoarseClass + 1];
    uint32_t off = 0;
    for (int c = 0; c <= kMaxCoarseClass; ++c) {
      info[c].g = make_geom((1 << c) + 1);
      info[c].off = c >= 1 ? off : 0;
      if (c >= 1) off += 4u * 2u * info[c].g.C;
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_coarse_info), info, sizeof(info)) != hipSuccess ||
        hipFuncSetAttribute((const void*)kb_coarse<double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kCoarseLdsBytes) != hipSuccess ||
        hipFuncSetAttribute((const void*)kb_coarse<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kCoarseLdsBytes) != hipSuccess)
      failure = "coarse kernel initialisation failed";
  });
  if (!failure.empty()) throw HipError{failure};
}

#define MPR_BINST(T)                                                                           \
  template void launch_b_relax<T>(const StepPtrs<T>*, uint32_t, const Geom&, double, int,      \
                                  hipStream_t, int, bool, bool, bool);                               \
  template void launch_b_residual<T>(const StepPtrs<T>*, uint32_t, const Geom&, double,        \
                                     hipStream_t, bool);                                       \
  template void launch_b_restrict<T>(const StepPtrs<T>*, uint32_t, const Geom&, const Geom&,   \
                                     bool, hipStream_t, bool);                                 \
  template void launch_b_prolong<T>(const StepPtrs<T>*, uint32_t, const Geom&, const Geom&,    \
                                    bool, hipStream_t);                                        \
  template void launch_b_fill_x<T>(const StepPtrs<T>*, uint32_t, const Geom&, hipStream_t);    \
  template void launch_b_coarse<T>(const CoarseEntry<T>*, uint32_t, int, double, int, int,     \
                                   hipStream_t);                                               \
  template int coarse_b_top_class<T>();
MPR_BINST(double)
MPR_BINST(float)
#undef MPR_BINST

}  // namespace mpr
