// Per-node element functions of the multigrid solver on the colour-split layout.
//
// Each function computes one node exactly as the reference kernel does (same operand
// order, no contraction: the library is compiled with -ffp-contract=off), so the FP64
// instantiation is bit-identical to src/capi.mojo on the same inputs. The multi-block
// kernels and the single-workgroup coarse-level kernel both call these.
#pragma once

#include <utility>

#include "mpr_common.h"

namespace mpr {

// Neumann 7-point neighbour sum in the order x-1, x+1, y-1, y+1, z-1, z+1 and the number
// of in-grid neighbours (capi.mojo:282-306). All neighbours of (x,y,z) have the other
// colour, so with the split layout they live in colour block 1-c.
//
// All six loads are issued unconditionally from in-grid addresses (a missing neighbour
// reads the node's own slot instead) and only the additions are guarded, so a lane keeps
// six loads in flight instead of waiting on each in turn; the guarded sum is the same
// sequence of roundings as skipping the term.
template <class T>
MPR_HD T neighbour_sum(const T* X, const Geom& g, int x, int y, int z, int& deg) {
  const int n = g.n;
  const bool bx0 = x > 0, bx1 = x + 1 < n, by0 = y > 0, by1 = y + 1 < n, bz0 = z > 0,
             bz1 = z + 1 < n;
  const uint32_t self = split_index(g, x, y, z);
  const T v0 = X[bx0 ? split_index(g, x - 1, y, z) : self];
  const T v1 = X[bx1 ? split_index(g, x + 1, y, z) : self];
  const T v2 = X[by0 ? split_index(g, x, y - 1, z) : self];
  const T v3 = X[by1 ? split_index(g, x, y + 1, z) : self];
  const T v4 = X[bz0 ? split_index(g, x, y, z - 1) : self];
  const T v5 = X[bz1 ? split_index(g, x, y, z + 1) : self];
  T s = T(0);
  if (bx0) s += v0;
  if (bx1) s += v1;
  if (by0) s += v2;
  if (by1) s += v3;
  if (bz0) s += v4;
  if (bz1) s += v5;
  deg = (int)bx0 + (int)bx1 + (int)by0 + (int)by1 + (int)bz0 + (int)bz1;
  return s;
}

// Red-black Gauss-Seidel update of node idx (capi.mojo:343-401):
// x = (rhs + h^-2 * sum) / (deg h^-2 + (alpha * screen) h^-2), skipped if the diagonal
// is not positive (the reference's scalar-path guard, applied at every node).
template <class T>
MPR_HD void relax_node(T* X, const T* R, const T* S, const Geom& g, uint32_t idx, int x, int y,
                       int z, T pw, T h2) {
  int deg;
  T s = neighbour_sum(X, g, x, y, z, deg);
  T screening = pw * S[idx] * h2;
  T diag = (T)deg * h2 + screening;
  if (diag > T(1.0e-30)) X[idx] = (R[idx] + s * h2) / diag;
}

// r = rhs - A x (capi.mojo:413-470)
template <class T>
MPR_HD T residual_node(const T* X, const T* R, const T* S, const Geom& g, uint32_t idx, int x,
                       int y, int z, T pw, T h2) {
  int deg;
  T s = neighbour_sum(X, g, x, y, z, deg);
  T screening = pw * S[idx] * h2;
  T applied = ((T)deg * h2 + screening) * X[idx] - s * h2;
  return R[idx] - applied;
}

// ---- two-phase form for streaming kernels ---------------------------------------------
// A lane first issues every load of its NPT nodes (neighbours from the other colour, rhs,
// screen, own value), then computes and stores. Within a colour sweep no node reads a node
// that is written, so hoisting all loads above the stores is exact, and it doubles the
// bytes each lane keeps in flight. The arithmetic is relax_node / residual_node's.
template <class T>
struct NodeLoads {
  T v[6];       // neighbours x-1, x+1, y-1, y+1, z-1, z+1 (own value where absent)
  T r, s, self;
  uint32_t idx;
  uint32_t flags;  // bit j: neighbour j exists
  bool ok;         // a real node (not a pad slot or past the end)
};

// t: index inside colour block c (values >= g.C are clamped and marked not ok)
template <class T, class PX, class PC>
MPR_HD void load_node(PX X, PC R, PC S, const Geom& g, int c, uint32_t t, bool want_self,
                      NodeLoads<T>& L, bool x_zero = false) {
  const bool in = t < g.C;
  t = in ? t : g.C - 1;
  int x, y, z;
  L.ok = split_decode(g, c, t, x, y, z) && in;
  const int n = g.n;
  const bool bx0 = x > 0, bx1 = x + 1 < n, by0 = y > 0, by1 = y + 1 < n, bz0 = z > 0,
             bz1 = z + 1 < n;
  const uint32_t self = (uint32_t)c * g.C + t;
  L.idx = self;
  L.flags = (uint32_t)bx0 | ((uint32_t)bx1 << 1) | ((uint32_t)by0 << 2) | ((uint32_t)by1 << 3) |
            ((uint32_t)bz0 << 4) | ((uint32_t)bz1 << 5);
  if (x_zero) {  // as load_group: the other colour is known to be zero
#pragma unroll
    for (int j = 0; j < 6; ++j) L.v[j] = T(0);
  } else {
    L.v[0] = X[bx0 ? split_index(g, x - 1, y, z) : self];
    L.v[1] = X[bx1 ? split_index(g, x + 1, y, z) : self];
    L.v[2] = X[by0 ? split_index(g, x, y - 1, z) : self];
    L.v[3] = X[by1 ? split_index(g, x, y + 1, z) : self];
    L.v[4] = X[bz0 ? split_index(g, x, y, z - 1) : self];
    L.v[5] = X[bz1 ? split_index(g, x, y, z + 1) : self];
  }
  L.r = R[self];
  L.s = S[self];
  L.self = want_self ? X[self] : T(0);
}

template <class T>
MPR_HD T loaded_sum(const NodeLoads<T>& L, int& deg) {
  T s = T(0);
  int d = 0;
#pragma unroll
  for (int j = 0; j < 6; ++j)
    if ((L.flags >> j) & 1u) {
      s += L.v[j];
      ++d;
    }
  deg = d;
  return s;
}

// relax_node on loaded values; returns whether the node is written
template <class T>
MPR_HD bool loaded_relax(const NodeLoads<T>& L, T pw, T h2, T& out) {
  int deg;
  const T s = loaded_sum(L, deg);
  const T screening = pw * L.s * h2;
  const T diag = (T)deg * h2 + screening;
  out = (L.r + s * h2) / diag;
  return L.ok && diag > T(1.0e-30);
}

// residual_node on loaded values
template <class T>
MPR_HD T loaded_residual(const NodeLoads<T>& L, T pw, T h2) {
  int deg;
  const T s = loaded_sum(L, deg);
  const T screening = pw * L.s * h2;
  const T applied = ((T)deg * h2 + screening) * L.self - s * h2;
  return L.r - applied;
}

// ---- slot-group form: V consecutive slots of one colour row per lane -------------------
// On levels whose row length H is a multiple of 4 (make_geom pads rows of n >= 33), a lane
// takes slots k0..k0+V-1 of row (y,z) in colour c and reads every operand with 16-byte
// loads: the same-row neighbours (x-1, x+1) are the other colour's slots k0..k0+V-1 plus
// one edge slot (k0-1 for rows whose first colour-c node is at even x, k0+V otherwise),
// and the y/z neighbours are the same V slots of the other colour's rows (y+-1, z+-1).
// Per node the arithmetic is relax_node / residual_node's, in the same order.
template <class T>
struct Chunk16;
template <>
struct Chunk16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Chunk16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

template <class T, int V>
struct SlotGroup {
  static constexpr int W = 16 / (int)sizeof(T);  // elements per 16-byte chunk
  static constexpr int NC = V / W;               // chunks per operand
  static_assert(V % W == 0, "a slot group is a whole number of 16-byte chunks");
  using CT = typename Chunk16<T>::type;
  CT row[NC], ym[NC], yp[NC], zm[NC], zp[NC], r[NC], s[NC], self[NC];
  T edge;
  uint32_t own;  // index of the group's first slot (colour block offset included)
  int x0, y, z, p;
  bool in;  // lane inside the level (lanes past the end load clamped and store nothing)
};

template <class T, class CT, int NC>
__device__ __forceinline__ void load_chunks(CT (&d)[NC], gptr<const T> p) {
  constexpr int W = 16 / (int)sizeof(T);
#pragma unroll
  for (int i = 0; i < NC; ++i) d[i] = *(const __attribute__((address_space(1))) CT*)(p + i * W);
}

// Live slots of a level whose right-hand side and screening weights are exactly +0.0 away
// from the splat: bit k of `bits` (k = colour * C + slot, the split index) is set where the
// assembled rhs or screen of slot k is non-zero. A group whose V bits are all clear reads
// the screening weights (and, with `rhs`, the right-hand side) as +0.0 without loading them:
// the values it would load. bits == nullptr: load everything.
struct LiveMask {
  const uint32_t* bits = nullptr;
  bool rhs = false;
};

// lane: index of the group inside colour block c; lanes = C / V. x_known: every x word is
// known to hold xv (+0 for the first sweep from a zero start, the initial guess before the
// first defect), so the x loads are skipped and the values are exactly what they would load.
template <class T, int V>
__device__ __forceinline__ void load_group(gptr<const T> X, gptr<const T> R, gptr<const T> S,
                                           const Geom& g, int c, uint32_t lane, uint32_t lanes,
                                           bool want_self, SlotGroup<T, V>& G,
                                           bool x_known = false, LiveMask live = LiveMask(),
                                           bool load_edge = true,
                                           const SlotGroup<T, V>* partner = nullptr,
                                           T xv = T(0)) {
  G.in = lane < lanes;
  lane = G.in ? lane : lanes - 1;
  const uint32_t s0 = lane * (uint32_t)V;
  const uint32_t z = g.div_nH.div(s0);
  const uint32_t rr = s0 - z * g.nH;
  const uint32_t y = g.div_H.div(rr);
  const uint32_t k0 = rr - y * (uint32_t)g.H;
  const int p = (c ^ (int)y ^ (int)z) & 1;
  const uint32_t o = (uint32_t)(1 - c) * g.C + s0;
  G.own = (uint32_t)c * g.C + s0;
  G.x0 = 2 * (int)k0 + p;
  G.y = (int)y;
  G.z = (int)z;
  G.p = p;
  const int n = g.n;
  if (x_known) {
#pragma unroll
    for (int i = 0; i < SlotGroup<T, V>::NC; ++i)
      G.row[i] = G.ym[i] = G.yp[i] = G.zm[i] = G.zp[i] = typename SlotGroup<T, V>::CT(xv);
    G.edge = xv;
  } else {
    // partner: the other colour's group at the same slots (loaded with want_self), whose own
    // chunk is this group's row and whose row is this group's own chunk
    if (partner) {
#pragma unroll
      for (int i = 0; i < SlotGroup<T, V>::NC; ++i) G.row[i] = partner->self[i];
    } else {
      load_chunks<T>(G.row, X + o);
    }
    if (load_edge) {
      const uint32_t eo = p == 0 ? (k0 > 0 ? o - 1 : o) : (k0 + V < (uint32_t)g.H ? o + V : o);
      G.edge = X[eo];
    }
    load_chunks<T>(G.ym, X + (y > 0 ? o - (uint32_t)g.H : o));
    load_chunks<T>(G.yp, X + ((int)y + 1 < n ? o + (uint32_t)g.H : o));
    load_chunks<T>(G.zm, X + (z > 0 ? o - g.nH : o));
    load_chunks<T>(G.zp, X + ((int)z + 1 < n ? o + g.nH : o));
  }
  // own is a multiple of V (C is a multiple of 4), so the group's V bits share one word
  const bool dead =
      live.bits && !((live.bits[G.own >> 5] >> (G.own & 31u)) & ((1u << V) - 1u));
  using CT = typename SlotGroup<T, V>::CT;
  if (dead && live.rhs) {
#pragma unroll
    for (int i = 0; i < SlotGroup<T, V>::NC; ++i) G.r[i] = CT(T(0));
  } else {
    load_chunks<T>(G.r, R + G.own);
  }
  if (dead) {
#pragma unroll
    for (int i = 0; i < SlotGroup<T, V>::NC; ++i) G.s[i] = CT(T(0));
  } else {
    load_chunks<T>(G.s, S + G.own);
  }
  if (want_self) {
    if (x_known) {
#pragma unroll
      for (int i = 0; i < SlotGroup<T, V>::NC; ++i) G.self[i] = CT(xv);
    } else if (partner) {
#pragma unroll
      for (int i = 0; i < SlotGroup<T, V>::NC; ++i) G.self[i] = partner->row[i];
    } else {
      load_chunks<T>(G.self, X + G.own);
    }
  }
}

// Edge slots of a lane's NG consecutive slot groups (group index lane * NG + k, the lanes of
// a wavefront consecutive, all in colour c, loaded with load_edge = false). A group's edge
// is slot k0 - 1 (x parity 0) or k0 + V (parity 1) of the other colour's row: the last slot
// of the previous group's or the first slot of the next group's row chunk whenever it is
// inside the row, taken from this lane or a neighbouring lane; only the first and last
// lanes of a wavefront load it. Every lane must call this (the exchange is wavefront-wide).
// x_known: x is the constant xv (load_group).
template <class T, int V, int NG>
__device__ __forceinline__ void group_edges(gptr<const T> X, const Geom& g, int c,
                                            SlotGroup<T, V> (&G)[NG], bool x_known,
                                            T xv = T(0)) {
  constexpr int W = SlotGroup<T, V>::W;
  constexpr int NC = SlotGroup<T, V>::NC;
  const T prev_lane = __shfl_up(G[NG - 1].row[NC - 1][W - 1], 1, 64);
  const T next_lane = __shfl_down(G[0].row[0][0], 1, 64);
  const int lid = (int)(threadIdx.x & 63u);
#pragma unroll
  for (int k = 0; k < NG; ++k) {
    if (x_known) {
      G[k].edge = xv;
      continue;
    }
    const uint32_t k0 = (uint32_t)(G[k].x0 - G[k].p) >> 1;
    const uint32_t o = (uint32_t)(1 - c) * g.C + (G[k].own - (uint32_t)c * g.C);
    if (G[k].p == 0) {
      const bool in_row = k0 > 0;
      if (k > 0)
        G[k].edge = G[k - 1].row[NC - 1][W - 1];
      else if (lid > 0 || !in_row)
        G[k].edge = prev_lane;
      else
        G[k].edge = X[o - 1];
    } else {
      const bool in_row = k0 + V < (uint32_t)g.H;
      if (k < NG - 1)
        G[k].edge = G[k + 1].row[0][0];
      else if (lid < 63 || !in_row)
        G[k].edge = next_lane;
      else
        G[k].edge = X[o + V];
    }
  }
}

// element J of an operand (J is a compile-time constant, so every access is static)
template <int J, class T, class CT, int NC>
__device__ __forceinline__ T elem(const CT (&d)[NC]) {
  constexpr int W = 16 / (int)sizeof(T);
  return d[J / W][J % W];
}

// neighbour sum and degree of slot J, terms in the order of neighbour_sum
template <int J, class T, int V>
__device__ __forceinline__ T group_sum(const SlotGroup<T, V>& G, int n, int& deg) {
  const T cur = elem<J, T>(G.row);
  T prev, next;
  if constexpr (J == 0)
    prev = G.edge;
  else
    prev = elem<J - 1, T>(G.row);
  if constexpr (J == V - 1)
    next = G.edge;
  else
    next = elem<J + 1, T>(G.row);
  const T v0 = G.p ? cur : prev;  // x - 1
  const T v1 = G.p ? next : cur;  // x + 1
  const int x = G.x0 + 2 * J;
  const bool bx0 = x > 0, bx1 = x + 1 < n, by0 = G.y > 0, by1 = G.y + 1 < n, bz0 = G.z > 0,
             bz1 = G.z + 1 < n;
  T s = T(0);
  if (bx0) s += v0;
  if (bx1) s += v1;
  if (by0) s += elem<J, T>(G.ym);
  if (by1) s += elem<J, T>(G.yp);
  if (bz0) s += elem<J, T>(G.zm);
  if (bz1) s += elem<J, T>(G.zp);
  deg = (int)bx0 + (int)bx1 + (int)by0 + (int)by1 + (int)bz0 + (int)bz1;
  return s;
}

template <int J, class T, int V>
__device__ __forceinline__ void group_relax_one(const SlotGroup<T, V>& G, int n, T pw, T h2,
                                                T (&out)[V], uint32_t& mask) {
  int deg;
  const T s = group_sum<J, T, V>(G, n, deg);
  const T screening = pw * elem<J, T>(G.s) * h2;
  const T diag = (T)deg * h2 + screening;
  out[J] = (elem<J, T>(G.r) + s * h2) / diag;
  if (G.in && G.x0 + 2 * J < n && diag > T(1.0e-30)) mask |= 1u << J;
}

template <int J, class T, int V>
__device__ __forceinline__ T group_residual_one(const SlotGroup<T, V>& G, int n, T pw, T h2) {
  int deg;
  const T s = group_sum<J, T, V>(G, n, deg);
  const T screening = pw * elem<J, T>(G.s) * h2;
  const T applied = ((T)deg * h2 + screening) * elem<J, T>(G.self) - s * h2;
  return G.x0 + 2 * J < n ? elem<J, T>(G.r) - applied : T(0);
}

template <class T, int V, int... J>
__device__ __forceinline__ uint32_t group_relax_all(const SlotGroup<T, V>& G, int n, T pw, T h2,
                                                    T (&out)[V],
                                                    std::integer_sequence<int, J...>) {
  uint32_t mask = 0;
  (group_relax_one<J, T, V>(G, n, pw, h2, out, mask), ...);
  return mask;
}

// Relaxed values of the group's V slots; returns the mask of slots to store.
template <class T, int V>
__device__ __forceinline__ uint32_t group_relax(const SlotGroup<T, V>& G, int n, T pw, T h2,
                                                T (&out)[V]) {
  return group_relax_all<T, V>(G, n, pw, h2, out, std::make_integer_sequence<int, V>{});
}

template <class T, int V, int... J>
__device__ __forceinline__ void group_residual_all(const SlotGroup<T, V>& G, int n, T pw, T h2,
                                                   T (&out)[V],
                                                   std::integer_sequence<int, J...>) {
  ((out[J] = group_residual_one<J, T, V>(G, n, pw, h2)), ...);
}

// Residuals of the group's V slots (0 in pad slots, as the scalar kernels write).
template <class T, int V>
__device__ __forceinline__ void group_residual(const SlotGroup<T, V>& G, int n, T pw, T h2,
                                               T (&out)[V]) {
  group_residual_all<T, V>(G, n, pw, h2, out, std::make_integer_sequence<int, V>{});
}

// Stores V values at p (16-byte aligned): whole chunks when every slot is stored, else
// slot by slot. Values narrower than T (the FP64 defect written as FP32) go slot by slot,
// which measured faster than one converted vector store.
template <class T, class U, int V>
__device__ __forceinline__ void store_group(__attribute__((address_space(1))) U* p,
                                            const T (&v)[V], uint32_t mask) {
  if constexpr (sizeof(U) == sizeof(T)) {
    using CT = typename SlotGroup<T, V>::CT;
    constexpr int W = SlotGroup<T, V>::W;
    if (mask == (1u << V) - 1) {
#pragma unroll
      for (int i = 0; i < V / W; ++i) {
        CT c;
#pragma unroll
        for (int j = 0; j < W; ++j) c[j] = v[i * W + j];
        *(__attribute__((address_space(1))) CT*)(p + i * W) = c;
      }
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < V; ++j)
    if ((mask >> j) & 1u) p[j] = (U)v[j];
}

// Prolong-add of a slot group: fine slots k0..k0+V-1 of row (y,z) in colour c. All V nodes
// have the same x parity p, so they share tx, ty, tz and the four coarse rows
// (y0|y1, z0|z1); node j needs coarse x = k0+j and min(k0+j+1, nc-1), so each row is read
// once at x = k0..k0+V (coarse nodes alternate colour along x). Per node the arithmetic is
// prolong_node's followed by fine += value. vec_coarse (FP32, coarse rows 16-byte aligned):
// coarse x = k0..k0+4 are slots k0/2..k0/2+2 of one colour and k0/2..k0/2+1 of the other
// (k0 is a multiple of 4), read as two 8-byte loads and one scalar per row instead of five
// scalars; x past the last node takes the last node's value in registers, as the clamped
// index does.
template <class T, int V>
__device__ __forceinline__ void prolong_group(gptr<const T> XC, const Geom& gc, gptr<T> F,
                                              const Geom& gf, int c, uint32_t lane,
                                              uint32_t lanes, bool vec_coarse = false) {
  const bool in = lane < lanes;
  lane = in ? lane : lanes - 1;
  const uint32_t s0 = lane * (uint32_t)V;
  const uint32_t z = gf.div_nH.div(s0);
  const uint32_t rr = s0 - z * gf.nH;
  const uint32_t y = gf.div_H.div(rr);
  const uint32_t k0 = rr - y * (uint32_t)gf.H;
  const int p = (c ^ (int)y ^ (int)z) & 1;
  const uint32_t own = (uint32_t)c * gf.C + s0;
  typename SlotGroup<T, V>::CT xf[SlotGroup<T, V>::NC];
  load_chunks<T>(xf, (gptr<const T>)F + own);
  const int nc = gc.n;
  const int y0 = (int)y >> 1, z0 = (int)z >> 1;
  const int y1 = y0 + 1 < nc - 1 ? y0 + 1 : nc - 1;
  const int z1 = z0 + 1 < nc - 1 ? z0 + 1 : nc - 1;
  const T tx = p ? T(0.5) : T(0), ty = (y & 1) ? T(0.5) : T(0), tz = (z & 1) ? T(0.5) : T(0);
  T cv[4][V + 1];  // rows (y0,z0), (y1,z0), (y0,z1), (y1,z1) at coarse x = k0..k0+V
  bool loaded = false;
  if constexpr (sizeof(T) == 4 && V == 4) {
    if (vec_coarse) {
      using F2 = float __attribute__((ext_vector_type(2)));
      const uint32_t h = k0 >> 1;
      const int m = nc - 1 - (int)k0;  // x = k0 + m is the last coarse node
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int yr = (r & 1) ? y1 : y0, zr = (r & 2) ? z1 : z0;
        const uint32_t row = ((uint32_t)zr * (uint32_t)nc + (uint32_t)yr) * (uint32_t)gc.H;
        const uint32_t ca = (uint32_t)((yr + zr) & 1);  // colour of the even x = k0
        const uint32_t oa = ca * gc.C + row + h, ob = (1u - ca) * gc.C + row + h;
        const F2 a = *(const __attribute__((address_space(1))) F2*)(XC + oa);
        const F2 b = *(const __attribute__((address_space(1))) F2*)(XC + ob);
        const T a2 = XC[oa + 2];
        const T raw[V + 1] = {a[0], b[0], a[1], b[1], a2};
#pragma unroll
        for (int i = 0; i <= V; ++i) {
          T v = raw[i];
#pragma unroll
          for (int q = 0; q < i; ++q)
            if (i > m && q == m) v = raw[q];
          cv[r][i] = v;
        }
      }
      loaded = true;
    }
  }
  if (!loaded) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int yr = (r & 1) ? y1 : y0, zr = (r & 2) ? z1 : z0;
      const uint32_t row = ((uint32_t)zr * (uint32_t)nc + (uint32_t)yr) * (uint32_t)gc.H;
#pragma unroll
      for (int i = 0; i <= V; ++i) {
        const int X = (int)k0 + i < nc - 1 ? (int)k0 + i : nc - 1;
        const uint32_t cc = (uint32_t)((X + yr + zr) & 1);
        cv[r][i] = XC[cc * gc.C + row + (uint32_t)(X >> 1)];
      }
    }
  }
  T out[V];
  uint32_t mask = 0;
#pragma unroll
  for (int j = 0; j < V; ++j) {
    const T c00 = cv[0][j] * (T(1) - tx) + cv[0][j + 1] * tx;
    const T c10 = cv[1][j] * (T(1) - tx) + cv[1][j + 1] * tx;
    const T c01 = cv[2][j] * (T(1) - tx) + cv[2][j + 1] * tx;
    const T c11 = cv[3][j] * (T(1) - tx) + cv[3][j + 1] * tx;
    const T c0 = c00 * (T(1) - ty) + c10 * ty;
    const T c1 = c01 * (T(1) - ty) + c11 * ty;
    out[j] = xf[j / SlotGroup<T, V>::W][j % SlotGroup<T, V>::W] + (c0 * (T(1) - tz) + c1 * tz);
    if (in && 2 * ((int)k0 + j) + p < gf.n) mask |= 1u << j;
  }
  store_group<T, T, V>(F + own, out, mask);
}

// Full-weighting restriction of a slot group of coarse nodes: coarse slots K0..K0+V-1 of
// row (Y,Z) in colour cc, nodes X_j = 2(K0+j) + P. Coarse node j is centred at fine
// x = 2 X_j, whose colour in fine row (yy,zz) is q = (yy+zz)&1 at slot 2(K0+j)+P; its x-1
// and x+1 neighbours are the other colour's slots 2(K0+j)+P-1 and 2(K0+j)+P. So each of the
// nine fine rows is read as two 16-byte chunks of both colours starting at slot 2 K0, plus
// one slot before them. Terms are accumulated per coarse node in restrict_node's
// (oz, oy, ox) order with its weights, so the result is restrict_node's. Needs 16-byte
// aligned rows on both levels. Both coarse colours' groups at slots K0.. of row (Y,Z) read
// the same fine chunks, so one lane restricts both from one set of loads.
//
// The nine rows are taken one z-plane (three rows) at a time: the per-node sums still run
// over r = 0..8 in order, and a lane holds one plane's chunks instead of all nine rows'
// (about a third of the registers, so more waves hide the load latency). C1Z: the fine
// colour-1 values are known to be zero (the residual computed colour 0 only) and only the
// colour-0 chunks are loaded.
template <class T, int V, bool C1Z = false>
__device__ __forceinline__ void restrict_group(gptr<const T> F, const Geom& gf, gptr<T> RC,
                                               gptr<T> XC, const Geom& gc, uint32_t lane,
                                               uint32_t lanes, T scale) {
  using CT = typename SlotGroup<T, V>::CT;
  constexpr int W = SlotGroup<T, V>::W;
  constexpr int NR = 2 * V / W;  // chunks per fine row and colour
  const bool in = lane < lanes;
  lane = in ? lane : lanes - 1;
  const uint32_t s0 = lane * (uint32_t)V;
  const uint32_t Z = gc.div_nH.div(s0);
  const uint32_t rr = s0 - Z * gc.nH;
  const uint32_t Y = gc.div_H.div(rr);
  const uint32_t K0 = rr - Y * (uint32_t)gc.H;
  const int nf = gf.n;
  const uint32_t base = 2u * K0;
  // chunks past the row end only hold pad slots (never used): read the row start
  uint32_t slot[NR];
#pragma unroll
  for (int i = 0; i < NR; ++i)
    slot[i] = base + (uint32_t)(i * W) + W <= (uint32_t)gf.H ? base + i * W : 0u;
  const uint32_t bslot = base > 0 ? base - 1 : 0;
  T sum[2][V], wsum[2][V];
#pragma unroll
  for (int cc = 0; cc < 2; ++cc)
#pragma unroll
    for (int j = 0; j < V; ++j) sum[cc][j] = wsum[cc][j] = T(0);
#pragma unroll
  for (int oz = -1; oz <= 1; ++oz) {
    // this plane's three rows, every load first; the scheduling barrier keeps the machine
    // scheduler from hoisting the next plane's loads above this plane's sums
    __builtin_amdgcn_sched_barrier(0);
    if (C1Z) {
      // colour 0 only. In a fine row of parity q the colour-0 nodes sit at x = q mod 2, so
      // of coarse node j's three terms (x-1, x, x+1) the centre is colour 0 when q = 0 and
      // the two sides when q = 1; the others are the +0.0 the full form would load. The
      // window a[] holds colour-0 slots 2 K0 - 1 .. 2 K0 + 2V - 1 of the row.
      T a[3][2 * V + 1];
      uint32_t qk[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int zz = 2 * (int)Z + oz, yy = 2 * (int)Y + k - 1;
        const int zc = zz < 0 ? 0 : (zz >= nf ? nf - 1 : zz);
        const int yc = yy < 0 ? 0 : (yy >= nf ? nf - 1 : yy);
        qk[k] = (uint32_t)((yc + zc) & 1);
        const uint32_t row = ((uint32_t)zc * (uint32_t)nf + (uint32_t)yc) * (uint32_t)gf.H;
        a[k][0] = F[row + bslot];
#pragma unroll
        for (int i = 0; i < NR; ++i) {
          const CT c = *(const __attribute__((address_space(1))) CT*)(F + row + slot[i]);
#pragma unroll
          for (int w = 0; w < W; ++w) a[k][1 + i * W + w] = c[w];
        }
      }
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const int P = (cc ^ (int)Y ^ (int)Z) & 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int oy = k - 1;
          const int zz = 2 * (int)Z + oz, yy = 2 * (int)Y + oy;
          const bool iny = zz >= 0 && zz < nf && yy >= 0 && yy < nf;
          const T wz = oz == 0 ? T(2) : T(1), wy = oy == 0 ? T(2) : T(1);
          const bool q = qk[k] != 0u;
#pragma unroll
          for (int j = 0; j < V; ++j) {
            // slot 2(K0+j)+P is window index 2j+P+1, slot 2(K0+j)+P-1 index 2j+P
            const T hi = P ? a[k][2 * j + 2] : a[k][2 * j + 1];
            const T lo = P ? a[k][2 * j + 1] : a[k][2 * j];  // unused at x = 0 (inm false)
            const T vm = q ? lo : T(0);  // x - 1
            const T v0 = q ? T(0) : hi;  // x
            const T vp = q ? hi : T(0);  // x + 1
            const int xc = 2 * (2 * ((int)K0 + j) + P);
            const bool inm = iny && xc - 1 >= 0, in0 = iny && xc < nf, inp = iny && xc + 1 < nf;
            const T wm = T(1) * wy * wz, w0 = T(2) * wy * wz, wp = T(1) * wy * wz;
            if (inm) {
              sum[cc][j] += vm * wm;
              wsum[cc][j] += wm;
            }
            if (in0) {
              sum[cc][j] += v0 * w0;
              wsum[cc][j] += w0;
            }
            if (inp) {
              sum[cc][j] += vp * wp;
              wsum[cc][j] += wp;
            }
          }
        }
      }
      // pin this plane's partial sums here, so the compiler cannot sink a colour's whole
      // chain of additions below the last plane (which would keep every plane's words live)
#pragma unroll
      for (int cc = 0; cc < 2; ++cc)
#pragma unroll
        for (int j = 0; j < V; ++j) asm volatile("" : "+v"(sum[cc][j]), "+v"(wsum[cc][j]));
      continue;
    }
    CT aq[3][NR], ao[3][NR];
    T before[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int oy = k - 1;
      const int zz = 2 * (int)Z + oz, yy = 2 * (int)Y + oy;
      const int zc = zz < 0 ? 0 : (zz >= nf ? nf - 1 : zz);
      const int yc = yy < 0 ? 0 : (yy >= nf ? nf - 1 : yy);
      const uint32_t q = (uint32_t)((yc + zc) & 1);
      const uint32_t row = ((uint32_t)zc * (uint32_t)nf + (uint32_t)yc) * (uint32_t)gf.H;
      const uint32_t rq = q * gf.C + row, ro = (1u - q) * gf.C + row;
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        aq[k][i] = *(const __attribute__((address_space(1))) CT*)(F + rq + slot[i]);
        ao[k][i] = *(const __attribute__((address_space(1))) CT*)(F + ro + slot[i]);
      }
      before[k] = F[ro + bslot];
    }
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const int P = (cc ^ (int)Y ^ (int)Z) & 1;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int oy = k - 1;
        const int zz = 2 * (int)Z + oz, yy = 2 * (int)Y + oy;
        const bool iny = zz >= 0 && zz < nf && yy >= 0 && yy < nf;
        const T wz = oz == 0 ? T(2) : T(1), wy = oy == 0 ? T(2) : T(1);
#pragma unroll
        for (int j = 0; j < V; ++j) {
          const T q0 = aq[k][(2 * j) / W][(2 * j) % W];
          const T q1 = aq[k][(2 * j + 1) / W][(2 * j + 1) % W];
          const T o0 = ao[k][(2 * j) / W][(2 * j) % W];
          const T o1 = ao[k][(2 * j + 1) / W][(2 * j + 1) % W];
          T om;
          if (j == 0)
            om = before[k];
          else
            om = ao[k][(2 * j - 1) / W][(2 * j - 1) % W];
          const T vm = P ? o0 : om;  // x - 1
          const T v0 = P ? q1 : q0;  // x
          const T vp = P ? o1 : o0;  // x + 1
          const int xc = 2 * (2 * ((int)K0 + j) + P);
          const bool inm = iny && xc - 1 >= 0, in0 = iny && xc < nf, inp = iny && xc + 1 < nf;
          const T wm = T(1) * wy * wz, w0 = T(2) * wy * wz, wp = T(1) * wy * wz;
          if (inm) {
            sum[cc][j] += vm * wm;
            wsum[cc][j] += wm;
          }
          if (in0) {
            sum[cc][j] += v0 * w0;
            wsum[cc][j] += w0;
          }
          if (inp) {
            sum[cc][j] += vp * wp;
            wsum[cc][j] += wp;
          }
        }
      }
    }
#pragma unroll
    for (int cc = 0; cc < 2; ++cc)
#pragma unroll
      for (int j = 0; j < V; ++j) asm volatile("" : "+v"(sum[cc][j]), "+v"(wsum[cc][j]));
  }
#pragma unroll
  for (int cc = 0; cc < 2; ++cc) {
    const int P = (cc ^ (int)Y ^ (int)Z) & 1;
    T out[V];
    uint32_t mask = 0;
#pragma unroll
    for (int j = 0; j < V; ++j) {
      out[j] = sum[cc][j] / wsum[cc][j] * scale;
      if (in && 2 * ((int)K0 + j) + P < gc.n) mask |= 1u << j;
    }
    const uint32_t own = (uint32_t)cc * gc.C + s0;
    store_group<T, T, V>(RC + own, out, mask);
    if (XC) {
      T zero[V];
#pragma unroll
      for (int j = 0; j < V; ++j) zero[j] = T(0);
      store_group<T, T, V>(XC + own, zero, in ? (1u << V) - 1 : 0u);
    }
  }
}

// Full-weighting restriction of coarse node (X,Y,Z) (capi.mojo:556-673): weights 2 at the
// centre and 1 at the sides per axis, terms in (oz, oy, ox) order, divided by the sum of
// the in-grid weights, then multiplied by `scale` (1 for residuals, 4 for the screening
// pyramid of core.py:267-269; both scalings are exact).
//
// As in neighbour_sum, the 27 loads use clamped (in-grid) coordinates and only the
// accumulation is guarded, so all loads can be in flight at once.
template <class T, class PF>
MPR_HD T restrict_node(PF F, const Geom& gf, int X, int Y, int Z, T scale) {
  const int nf = gf.n;
  T v[27];
#pragma unroll
  for (int oz = -1; oz <= 1; ++oz) {
    int zz = 2 * Z + oz;
    zz = zz < 0 ? 0 : (zz >= nf ? nf - 1 : zz);
#pragma unroll
    for (int oy = -1; oy <= 1; ++oy) {
      int yy = 2 * Y + oy;
      yy = yy < 0 ? 0 : (yy >= nf ? nf - 1 : yy);
#pragma unroll
      for (int ox = -1; ox <= 1; ++ox) {
        int xx = 2 * X + ox;
        xx = xx < 0 ? 0 : (xx >= nf ? nf - 1 : xx);
        v[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)] = F[split_index(gf, xx, yy, zz)];
      }
    }
  }
  T sum = T(0), wsum = T(0);
#pragma unroll
  for (int oz = -1; oz <= 1; ++oz) {
    const int zz = 2 * Z + oz;
    const bool inz = zz >= 0 && zz < nf;
    const T wz = oz == 0 ? T(2) : T(1);
#pragma unroll
    for (int oy = -1; oy <= 1; ++oy) {
      const int yy = 2 * Y + oy;
      const bool iny = inz && yy >= 0 && yy < nf;
      const T wy = oy == 0 ? T(2) : T(1);
#pragma unroll
      for (int ox = -1; ox <= 1; ++ox) {
        const int xx = 2 * X + ox;
        const bool in = iny && xx >= 0 && xx < nf;
        const T wx = ox == 0 ? T(2) : T(1);
        const T w = wx * wy * wz;
        if (in) {
          sum += v[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)] * w;
          wsum += w;
        }
      }
    }
  }
  return sum / wsum * scale;
}

// Trilinear prolongation value at fine node (x,y,z) (capi.mojo:676-742).
template <class T, class PC>
MPR_HD T prolong_node(PC Cg, const Geom& gc, int x, int y, int z) {
  const int nc = gc.n;
  int x0 = x >> 1, y0 = y >> 1, z0 = z >> 1;
  int x1 = x0 + 1 < nc - 1 ? x0 + 1 : nc - 1;
  int y1 = y0 + 1 < nc - 1 ? y0 + 1 : nc - 1;
  int z1 = z0 + 1 < nc - 1 ? z0 + 1 : nc - 1;
  T tx = (x & 1) ? T(0.5) : T(0), ty = (y & 1) ? T(0.5) : T(0), tz = (z & 1) ? T(0.5) : T(0);
  T c000 = Cg[split_index(gc, x0, y0, z0)], c100 = Cg[split_index(gc, x1, y0, z0)];
  T c010 = Cg[split_index(gc, x0, y1, z0)], c110 = Cg[split_index(gc, x1, y1, z0)];
  T c001 = Cg[split_index(gc, x0, y0, z1)], c101 = Cg[split_index(gc, x1, y0, z1)];
  T c011 = Cg[split_index(gc, x0, y1, z1)], c111 = Cg[split_index(gc, x1, y1, z1)];
  T c00 = c000 * (T(1) - tx) + c100 * tx;
  T c10 = c010 * (T(1) - tx) + c110 * tx;
  T c01 = c001 * (T(1) - tx) + c101 * tx;
  T c11 = c011 * (T(1) - tx) + c111 * tx;
  T c0 = c00 * (T(1) - ty) + c10 * ty;
  T c1 = c01 * (T(1) - ty) + c11 * ty;
  return c0 * (T(1) - tz) + c1 * tz;
}

}  // namespace mpr
