// Multigrid grid kernels for gfx950: red-black smoothing, residual (+ squared norm),
// full-weighting restriction, trilinear prolongation, layout conversions and RHS assembly.
//
// All of them are HBM-bound streaming stencils (well under 1 flop per byte), so they are
// written for bandwidth: one node per lane, flattened indices so every wavefront touches
// contiguous 64-word runs, the colour-split layout so a colour sweep never drags the
// other colour's words through the cache, and an XCD-aware block order (blocks are
// dispatched round-robin over the 8 XCDs, so block b is remapped to give each XCD one
// contiguous slab of planes and keep the z-1/z+1 neighbour planes in that XCD's L2).
#include <algorithm>

#include "block_scan.h"
#include "mpr_internal.h"
#include "multi.h"
#include "stencil.h"

namespace mpr {

namespace {

__device__ __forceinline__ uint32_t logical_block() {
  const uint32_t per = gridDim.x >> 3;  // launch grids are padded to a multiple of 8
  return (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
}

inline unsigned xcd_grid(uint64_t work, unsigned threads) {
  unsigned nb = blocks_for(work, threads);
  return (nb + 7u) & ~7u;
}

template <class T>
__device__ __forceinline__ void fill_body(Blk blk_, T* p, T v, uint64_t count) {
  uint64_t i = (uint64_t)blk_.id * kThreads + threadIdx.x;
  uint64_t stride = (uint64_t)blk_.count * kThreads;
  for (; i < count; i += stride) p[i] = v;
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_relax_color(T* X, const T* R, const T* S, Geom g,
                                                          int c, T pw, T h2) {
  const uint32_t base = logical_block() * (kThreads * kNodesPerThread) + threadIdx.x;
  NodeLoads<T> L[kNodesPerThread];
#pragma unroll
  for (int k = 0; k < kNodesPerThread; ++k) load_node(X, R, S, g, c, base + k * kThreads, false, L[k]);
#pragma unroll
  for (int k = 0; k < kNodesPerThread; ++k) {
    T v;
    if (loaded_relax(L[k], pw, h2, v)) X[L[k].idx] = v;
  }
}

// slot-group form (one lane per 16 bytes of a colour row) for levels with aligned rows
template <class T>
__global__ __launch_bounds__(kThreads) void k_relax_group(T* X, const T* R, const T* S, Geom g,
                                                          uint32_t lanes, int c, T pw, T h2) {
  constexpr int V = 16 / (int)sizeof(T);
  SlotGroup<T, V> G;
  load_group<T, V>(as_global((const T*)X), as_global(R), as_global(S), g, c,
                   logical_block() * kThreads + threadIdx.x, lanes, false, G);
  T out[V];
  const uint32_t mask = group_relax<T, V>(G, g.n, pw, h2, out);
  store_group<T, T, V>(as_global(X) + G.own, out, mask);
}

__device__ __forceinline__ double block_sum(double v, double* lds) {
  // fixed-order reduction: wavefront butterfly, then the 4 wave totals in order
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double total = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; ++w) total += lds[w];
  return total;
}

template <class T, class TR>
__global__ __launch_bounds__(kThreads) void k_residual(const T* X, const T* R, const T* S,
                                                       TR* out, Geom g, T pw, T h2,
                                                       double* partials) {
  __shared__ double lds[kThreads / 64];
  const uint32_t lb = logical_block();
  const uint32_t base = lb * (kThreads * kNodesPerThread) + threadIdx.x;
  NodeLoads<T> L[kNodesPerThread];
  uint32_t tt[kNodesPerThread];
#pragma unroll
  for (int k = 0; k < kNodesPerThread; ++k) {
    const uint32_t t = base + k * kThreads;
    tt[k] = t;
    const int c = t >= g.C;
    load_node(X, R, S, g, c, t - (uint32_t)c * g.C, true, L[k]);
    if (t >= 2u * g.C) L[k].ok = false;
  }
  double sq = 0.0;
#pragma unroll
  for (int k = 0; k < kNodesPerThread; ++k) {
    if (tt[k] >= 2u * g.C) continue;
    if (L[k].ok) {
      const T r = loaded_residual(L[k], pw, h2);
      if (out) out[tt[k]] = (TR)r;
      sq += (double)r * (double)r;
    } else if (out) {
      out[tt[k]] = TR(0);
    }
  }
  if (partials) {
    double total = block_sum(sq, lds);
    if (threadIdx.x == 0 && lb * (kThreads * kNodesPerThread) < 2u * g.C) partials[lb] = total;
  }
}

// residual of slot group `lane` in both colours (one lane takes the same slots of both
// colour blocks, so x streams from memory once); block partials of the squared norm as
// k_residual (colour 0 then colour 1 per lane)
// With SIGN (FP64 solutions): also the sign bits of the extraction (node x of row (z, y) is
// set when its value is below *iso, natural layout, bw words per row, launch_bit_extract's
// k_sign_bits), from the x words the residual loads anyway. A lane's slots cover 4 nodes
// x = 2 k0 .. 2 k0 + 3 of both colours; the 8 lanes of a word OR their nibbles with shuffles
// and the first lane of each run adds them to the (zeroed) word.
template <class T, class TR, bool SIGN = false>
__device__ __forceinline__ void residual_group_body(Blk blk_, const T* X, const T* R, const T* S,
                                                    TR* out, Geom g, uint32_t lanes, T pw, T h2,
                                                    double* partials, const uint32_t* live,
                                                    const double* iso = nullptr,
                                                    uint32_t* bits = nullptr, uint32_t bw = 0) {
  constexpr int V = 16 / (int)sizeof(T);
  __shared__ double lds[kThreads / 64];
  const uint32_t lb = xcd_block(blk_);
  const uint32_t lane = lb * kThreads + threadIdx.x;
  double sq = 0.0;
  SlotGroup<T, V> G[2][1];
  load_group<T, V>(as_global(X), as_global(R), as_global(S), g, 0, lane, lanes, true, G[0][0],
                   false, LiveMask{live, true}, false);
  // colour 1 at the same slots: its row and own chunks are colour 0's own and row chunks
  load_group<T, V>(as_global(X), as_global(R), as_global(S), g, 1, lane, lanes, true, G[1][0],
                   false, LiveMask{live, true}, false, G[0]);
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    group_edges<T, V, 1>(as_global(X), g, c, G[c], false);
    T v[V];
    group_residual<T, V>(G[c][0], g.n, pw, h2, v);
    if (out)
      store_group<T, TR, V>(as_global(out) + G[c][0].own, v, G[c][0].in ? (1u << V) - 1 : 0u);
    if (G[c][0].in)
#pragma unroll
      for (int j = 0; j < V; ++j) sq += (double)v[j] * (double)v[j];
  }
  if constexpr (SIGN) {
    const double level = *iso;
    const int x_first = G[0][0].x0 - G[0][0].p;  // 2 k0, a multiple of 4
    uint32_t nib = 0;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int J = 0; J < V; ++J) {
        const int x = G[c][0].x0 + 2 * J;
        const T val = G[c][0].self[J / SlotGroup<T, V>::W][J % SlotGroup<T, V>::W];
        if (G[c][0].in && x < g.n && (double)val < level) nib |= 1u << (x - x_first);
      }
    const uint32_t wid = G[0][0].in ? ((uint32_t)G[0][0].z * (uint32_t)g.n + (uint32_t)G[0][0].y) *
                                              bw + ((uint32_t)x_first >> 5)
                                        : 0xffffffffu;
    uint32_t v = nib << ((uint32_t)x_first & 31u);
    const uint32_t wl = threadIdx.x & 63u;
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
      const uint32_t ov = (uint32_t)__shfl_down((int)v, off, 64);
      const uint32_t ow = (uint32_t)__shfl_down((int)wid, off, 64);
      if (wl + (uint32_t)off < 64u && ow == wid) v |= ov;
    }
    const uint32_t pw_id = (uint32_t)__shfl_up((int)wid, 1, 64);
    const bool head = wl == 0u || pw_id != wid;
    if (head && v && wid != 0xffffffffu) atomicOr(&bits[wid], v);
  }
  if (partials) {
    const double total = block_sum(sq, lds);
    if (threadIdx.x == 0 && lb * kThreads < lanes) partials[lb] = total;
  }
}

__device__ __forceinline__ void norm_finish_body(Blk blk_, const double* partials, int count,
                                                      double* out) {
  __shared__ double lds[1024];
  double v = 0.0;
  for (int i = threadIdx.x; i < count; i += 1024) v += partials[i];
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sqrt(lds[0]);
}

template <class T>
__device__ __forceinline__ void restrict_nodes_body(Blk blk_, const T* F, Geom gf, T* Cc, Geom gc,
                                                       T scale, float* copy32) {
  const uint32_t t = xcd_block(blk_) * kThreads + threadIdx.x;
  if (t >= 2u * gc.C) return;
  // colours interleaved (neighbouring threads restrict neighbouring coarse nodes, which
  // read overlapping fine nodes of both colours)
  const int c = (int)(t & 1u);
  const uint32_t i = (uint32_t)c * gc.C + (t >> 1);
  int X, Y, Z;
  if (!split_decode(gc, c, t >> 1, X, Y, Z)) return;
  const T v = restrict_node(F, gf, X, Y, Z, scale);
  Cc[i] = v;
  if (copy32) copy32[i] = (float)v;
}

// slot-group restriction, both coarse colours of slot group `lane` per lane
template <class T>
__global__ __launch_bounds__(kThreads) void k_restrict_group(const T* F, Geom gf, T* Cc, Geom gc,
                                                             uint32_t lanes, T scale) {
  constexpr int V = 16 / (int)sizeof(T);
  const uint32_t lane = logical_block() * kThreads + threadIdx.x;
  restrict_group<T, V>(as_global(F), gf, as_global(Cc), gptr<T>(nullptr), gc, lane, lanes, scale);
}

// slot-group prolong-add over both colours (lanes [0, lanes) colour 0, then colour 1)
template <class T>
__global__ __launch_bounds__(kThreads) void k_prolong_group(const T* Cc, Geom gc, T* F, Geom gf,
                                                            uint32_t lanes, int vec_coarse) {
  constexpr int V = 16 / (int)sizeof(T);
  const uint32_t lane = logical_block() * kThreads + threadIdx.x;
  const int c = lane >= lanes;
  const uint32_t l = lane - (uint32_t)c * lanes;
  prolong_group<T, V>(as_global(Cc), gc, as_global(F), gf, c, lane < 2u * lanes ? l : lanes,
                      lanes, vec_coarse != 0);
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_prolong_add(const T* Cc, Geom gc, T* F, Geom gf) {
  const uint32_t t = logical_block() * kThreads + threadIdx.x;
  if (t >= 2u * gf.C) return;
  const int c = t >= gf.C;
  int x, y, z;
  if (!split_decode(gf, c, t - (uint32_t)c * gf.C, x, y, z)) return;
  F[t] += prolong_node<T>(Cc, gc, x, y, z);
}

template <class TI, class TO>
__global__ __launch_bounds__(kThreads) void k_nat_to_split(const TI* nat, TO* split, Geom g) {
  const uint32_t t = logical_block() * kThreads + threadIdx.x;
  if (t >= 2u * g.C) return;
  const int c = t >= g.C;
  int x, y, z;
  if (split_decode(g, c, t - (uint32_t)c * g.C, x, y, z))
    split[t] = (TO)nat[nat_index(x, y, z, g.n)];
  else
    split[t] = TO(0);
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_split_to_nat(const T* split, double* nat, Geom g,
                                                           FastDiv div_n) {
  const uint32_t t = logical_block() * kThreads + threadIdx.x;
  const uint32_t n = (uint32_t)g.n;
  if (t >= n * n * n) return;
  uint32_t zy = div_n.div(t);
  uint32_t x = t - zy * n;
  uint32_t z = div_n.div(zy);
  uint32_t y = zy - z * n;
  nat[t] = (double)split[split_index(g, (int)x, (int)y, (int)z)];
}

template <class TI, class TO>
__global__ __launch_bounds__(kThreads) void k_convert(const TI* in, TO* out, uint64_t count) {
  uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (; i < count; i += stride) out[i] = (TO)in[i];
}

__global__ __launch_bounds__(kThreads) void k_add_correction(double* X, const float* E,
                                                             uint64_t count) {
  uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (; i < count; i += stride) X[i] += (double)E[i];
}

// value of a splat grid at node i: unset `touched` bits mean the node received nothing.
// The word is loaded whatever the bit (a node that received nothing holds a stale value,
// which the bit then replaces by +0.0), so the two loads are in flight together.
__device__ __forceinline__ double splat_value(const double* v, const uint32_t* touched,
                                              uint64_t i) {
  const double val = v[i];
  if (!touched) return val;
  return ((touched[i >> 5] >> (i & 31u)) & 1u) ? val : 0.0;
}

// 32 node bits starting at node b of the union of two bit sets (nodes outside [0, 32 words)
// read as unset)
__device__ __forceinline__ uint32_t bits_at(const uint32_t* a, const uint32_t* b, int64_t bit,
                                            int64_t words) {
  const int64_t q = bit >> 5;  // arithmetic shift: floor for negative offsets
  const uint32_t r = (uint32_t)(bit & 31);
  const uint32_t lo = (q >= 0 && q < words) ? (a[q] | b[q]) : 0u;
  if (!r) return lo;
  const uint32_t hi = (q + 1 >= 0 && q + 1 < words) ? (a[q + 1] | b[q + 1]) : 0u;
  return (lo >> r) | (hi << (32u - r));
}

// any of bits [b, b + len) set (len <= 64, all inside the mask)
__device__ __forceinline__ bool any_bit(const uint32_t* w, uint32_t b, uint32_t len) {
  const uint32_t e = b + len, q0 = b >> 5, qe = (e - 1) >> 5;
  uint32_t acc = 0;
  for (uint32_t q = q0; q <= qe; ++q) {
    uint32_t m = w[q];
    if (q == q0) m &= ~0u << (b & 31u);
    if (q == qe && (e & 31u)) m &= (1u << (e & 31u)) - 1u;
    acc |= m;
  }
  return acc != 0;
}

// need = nodes whose 7-point stencil (linear offsets 0, +-1, +-n, +-n^2) meets a node of
// cur or prev. Linear offsets also reach across row and plane ends, which only adds nodes
// (a superset is harmless: an added node is written with the value the assembly computes).
__device__ __forceinline__ void need_mask_body(Blk blk_, const uint32_t* cur, const uint32_t* prev,
                                                        int64_t words, int64_t sy, int64_t sz,
                                                        uint32_t* need) {
  const int64_t w = (int64_t)blk_.id * kThreads + threadIdx.x;
  if (w >= words) return;
  const int64_t b = 32 * w;
  need[w] = bits_at(cur, prev, b, words) | bits_at(cur, prev, b + 1, words) |
            bits_at(cur, prev, b - 1, words) | bits_at(cur, prev, b + sy, words) |
            bits_at(cur, prev, b - sy, words) | bits_at(cur, prev, b + sz, words) |
            bits_at(cur, prev, b - sz, words);
}

// outputs and parameters of the assembly (capi.mojo:227-279); null outputs are not written
struct AssembleArgs {
  const double *vx, *vy, *vz, *scr;
  double *rhs_nat, *rhs_split, *scr_split, *density, *scr_nat_out;
  float* scr_split32;
  const uint32_t* touched;
  uint32_t* live;
  double spn, pw, target;
};

// assembly of node t = (x, y, z); returns whether its split slot (si) is live (rhs or
// screen not +0.0, by bit pattern, so -0.0 counts as live); the callers set the live bits
__device__ __forceinline__ bool assemble_node(const AssembleArgs& a, const Geom& g, uint32_t t,
                                              int x, int y, int z, uint32_t& si_out) {
  const int n = g.n;
  const double h_inv = (double)(n - 1);
  const double h_inv2 = h_inv * h_inv;
  const int64_t sy = n, sz = (int64_t)n * n;
  const uint32_t* tb = a.touched;
  double div = 0.0;
  if (x == 0) div += (splat_value(a.vx, tb, t + 1) - splat_value(a.vx, tb, t)) * h_inv;
  else if (x == n - 1) div += (splat_value(a.vx, tb, t) - splat_value(a.vx, tb, t - 1)) * h_inv;
  else div += (splat_value(a.vx, tb, t + 1) - splat_value(a.vx, tb, t - 1)) * (0.5 * h_inv);
  if (y == 0) div += (splat_value(a.vy, tb, t + sy) - splat_value(a.vy, tb, t)) * h_inv;
  else if (y == n - 1) div += (splat_value(a.vy, tb, t) - splat_value(a.vy, tb, t - sy)) * h_inv;
  else div += (splat_value(a.vy, tb, t + sy) - splat_value(a.vy, tb, t - sy)) * (0.5 * h_inv);
  if (z == 0) div += (splat_value(a.vz, tb, t + sz) - splat_value(a.vz, tb, t)) * h_inv;
  else if (z == n - 1) div += (splat_value(a.vz, tb, t) - splat_value(a.vz, tb, t - sz)) * h_inv;
  else div += (splat_value(a.vz, tb, t + sz) - splat_value(a.vz, tb, t - sz)) * (0.5 * h_inv);
  const double s = splat_value(a.scr, tb, t);
  const double screening = a.pw * s * h_inv2;
  const double r = -div + screening * a.target;
  if (a.rhs_nat) a.rhs_nat[t] = r;
  if (a.scr_nat_out) a.scr_nat_out[t] = s;
  bool live = false;
  si_out = 0u;
  if (a.rhs_split) {
    const uint32_t si = split_index(g, x, y, z);
    a.rhs_split[si] = r;
    a.scr_split[si] = s;
    if (a.scr_split32) a.scr_split32[si] = (float)s;
    live = a.live && (__double_as_longlong(r) | __double_as_longlong(s)) != 0;
    si_out = si;
  }
  if (a.density) a.density[t] = splat_value(a.density, tb, t) / a.spn;
  return live;
}

// dense assembly: every node; with `prev`, also records the touched bits for the next run
__device__ __forceinline__ void assemble_dense_body(Blk blk_, AssembleArgs a, Geom g, FastDiv div_n,
                                                       uint32_t* prev) {
  const uint32_t t = xcd_block(blk_) * kThreads + threadIdx.x;
  const int n = g.n;
  const bool act = t < (uint32_t)n * n * n;
  bool live = false;
  uint32_t si = 0u;
  if (act) {
    if (prev && (t & 31u) == 0) prev[t >> 5] = a.touched[t >> 5];
    const uint32_t zy = div_n.div(t);
    const int x = (int)(t - zy * (uint32_t)n);
    const int z = (int)div_n.div(zy);
    const int y = (int)(zy - (uint32_t)z * (uint32_t)n);
    live = assemble_node(a, g, t, x, y, z, si);
  }
  if (a.live) wave_set_bits(a.live, live, si);
}

// Sparse assembly, step 1: lists the words of the need mask that hold a node (list order is
// free: nodes are independent), and records this run's touched bits for the next run's need
// mask (the need mask was computed from the old ones by the previous kernel). A block takes
// kNeedWordsPerThread words per lane and reserves its list entries with one atomic: every
// reservation of a job goes to one counter, whose atomics serialise in the L2, so one per
// wavefront (about 8k per 257^3 job) cost about 20 us per job and one per block 64x fewer.
constexpr uint32_t kNeedWordsPerThread = 8;
__device__ __forceinline__ void need_words_body(Blk blk_, const uint32_t* need,
                                                const uint32_t* touched, uint32_t* prev,
                                                uint32_t words, uint32_t* list,
                                                uint32_t* count) {
  __shared__ uint32_t lds4[4];
  __shared__ uint32_t block_base;
  const uint32_t w0 = blk_.id * (kThreads * kNeedWordsPerThread) + threadIdx.x;
  uint32_t hit = 0u;  // bit i: word w0 + i kThreads is in the need mask
#pragma unroll
  for (uint32_t i = 0; i < kNeedWordsPerThread; ++i) {
    const uint32_t w = w0 + i * kThreads;
    if (w < words) {
      prev[w] = touched[w];
      if (need[w]) hit |= 1u << i;
    }
  }
  uint32_t total;
  const uint32_t at = block_exclusive256((uint32_t)__popc(hit), lds4, total);
  if (threadIdx.x == 0 && total) block_base = atomicAdd(count, total);
  __syncthreads();
  if (!total) return;
  uint32_t* out = list + block_base + at;
#pragma unroll
  for (uint32_t i = 0; i < kNeedWordsPerThread; ++i)
    if (hit & (1u << i)) *out++ = w0 + i * kThreads;
}

// Sparse assembly, step 2: a fixed grid walks the listed need words, 32 lanes per word
// (node t = 32 w + lane, natural layout); a node outside the need mask has no touched
// neighbour now or in the previous run, so its assembled words are the +0.0 they already
// hold and it is skipped.
__device__ __forceinline__ void assemble_words_body(Blk blk_, AssembleArgs a, Geom g,
                                                    FastDiv div_n, const uint32_t* list,
                                                    const uint32_t* count,
                                                    const uint32_t* need) {
  const uint32_t sub = threadIdx.x & 31u;
  const uint32_t slots = blk_.count * (kThreads / 32);
  const uint32_t total = *count;
  const uint64_t m = (uint64_t)g.n * g.n * g.n;
  // k is uniform over each half-wavefront, and wave_set_bits needs every lane of the
  // wavefront: the loop runs to the bound of the wavefront's first half-wave slot
  const uint32_t first = blk_.id * (kThreads / 32) + ((threadIdx.x >> 6) << 1);
  for (uint32_t base = first; base < total; base += slots) {
    const uint32_t k = base + ((threadIdx.x >> 5) & 1u);
    bool live = false;
    uint32_t si = 0u;
    if (k < total) {
      const uint32_t w = list[k];
      const uint32_t t = 32u * w + sub;
      if (t < m && ((need[w] >> sub) & 1u)) {
        const uint32_t zy = div_n.div(t);
        const int x = (int)(t - zy * (uint32_t)g.n);
        const int z = (int)div_n.div(zy);
        const int y = (int)(zy - (uint32_t)z * (uint32_t)g.n);
        live = assemble_node(a, g, t, x, y, z, si);
      }
    }
    if (a.live) wave_set_bits(a.live, live, si);
  }
}

// 64 bits of a bit set starting at bit `pos`, those at or past `end` cleared
__device__ __forceinline__ uint64_t bits64_at(const uint32_t* b, uint64_t pos, uint64_t end,
                                              uint64_t words) {
  const uint64_t q = pos >> 5;
  const uint32_t sh = (uint32_t)(pos & 31u);
  const uint64_t w0 = b[q], w1 = q + 1 < words ? b[q + 1] : 0u, w2 = q + 2 < words ? b[q + 2] : 0u;
  uint64_t v = (w0 | (w1 << 32)) >> sh;
  if (sh) v |= w2 << (64u - sh);
  if (end <= pos) return 0u;
  if (end - pos < 64u) v &= (1ull << (end - pos)) - 1ull;
  return v;
}

// the even-position bits of x, packed into the low 32 bits
__device__ __forceinline__ uint32_t even_bits(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x >> 4)) & 0x00ff00ff00ff00ffull;
  x = (x | (x >> 8)) & 0x0000ffff0000ffffull;
  x = (x | (x >> 16)) & 0x00000000ffffffffull;
  return (uint32_t)x;
}

// Coarse "support" bits: coarse node (X,Y,Z) is set when a fine node of its 27-node
// restriction support (2X+ox, 2Y+oy, 2Z+oz, in the grid) is set in `fine` (natural layout,
// nf^3 bits). One lane per word of the coarse bit set, which holds cw = ceil(nc / 32) words
// per coarse row: coarse x 32 w .. 32 w + 31 of row (Z, Y) take fine x 64 w - 1 .. 64 w + 63.
__device__ __forceinline__ void coarse_support_body(Blk blk_, const uint32_t* fine, int nf,
                                                    int nc, uint32_t cw, uint32_t* out,
                                                    uint32_t* list, uint32_t* count) {
  __shared__ uint32_t lds4[4];
  __shared__ uint32_t block_base;
  const uint64_t q = (uint64_t)blk_.id * kThreads + threadIdx.x;
  const uint64_t total = (uint64_t)nc * nc * cw;
  uint32_t acc = 0;
  if (q < total) {
    const uint32_t w = (uint32_t)(q % cw);
    const uint64_t zy = q / cw;
    const int Z = (int)(zy / (uint64_t)nc), Y = (int)(zy % (uint64_t)nc);
    const uint64_t words = ((uint64_t)nf * nf * nf + 31) / 32;
    for (int oz = -1; oz <= 1; ++oz) {
      const int zz = 2 * Z + oz;
      if (zz < 0 || zz >= nf) continue;
      for (int oy = -1; oy <= 1; ++oy) {
        const int yy = 2 * Y + oy;
        if (yy < 0 || yy >= nf) continue;
        const uint64_t base = ((uint64_t)zz * nf + yy) * nf, x0 = 64ull * w;
        if (x0 >= (uint64_t)nf) continue;
        const uint64_t v = bits64_at(fine, base + x0, base + nf, words);
        uint32_t before = 0;  // fine x = 64 w - 1, the x - 1 neighbour of coarse x = 32 w
        if (w > 0) before = (fine[(base + x0 - 1) >> 5] >> ((base + x0 - 1) & 31u)) & 1u;
        const uint32_t e = even_bits(v), o = even_bits(v >> 1);
        acc |= e | o | (o << 1) | before;
      }
    }
    const uint32_t x_first = 32u * w;
    if ((uint32_t)nc - x_first < 32u) acc &= (1u << ((uint32_t)nc - x_first)) - 1u;
    out[q] = acc;
  }
  // list order is free (nodes are independent): one reservation per block, since every
  // reservation of a job goes to one counter and its atomics serialise in the L2
  uint32_t listed;
  const uint32_t at = block_prefix256(acc != 0u, lds4, listed);
  if (threadIdx.x == 0 && listed) block_base = atomicAdd(count, listed);
  __syncthreads();
  if (acc) list[block_base + at] = (uint32_t)q;
}

// restrict_nodes_body over the coarse nodes set in `support` (coarse_support_body of the
// fine nodes that this or the previous run may have written); every other coarse node keeps
// the +0.0 it holds: the restriction of +0.0 words. A fixed grid walks the list of non-zero
// support words, 32 lanes per word (coarse x = 32 w + lane of row (Z, Y)).
template <class T>
__device__ __forceinline__ void restrict_sparse_body(Blk blk_, const T* F, Geom gf, T* Cc, Geom gc,
                                                     T scale, float* copy32,
                                                     const uint32_t* support,
                                                     const uint32_t* list, const uint32_t* count) {
  const uint32_t cw = ((uint32_t)gc.n + 31u) / 32u;
  const uint32_t total = *count;
  const uint32_t sub = threadIdx.x & 31u;
  const uint32_t slots = blk_.count * (kThreads / 32);
  for (uint32_t k = blk_.id * (kThreads / 32) + (threadIdx.x >> 5); k < total; k += slots) {
    const uint32_t q = list[k];
    const uint32_t word = support[q];
    if (!((word >> sub) & 1u)) continue;
    const uint32_t zy = q / cw, w = q - zy * cw;
    const int Z = (int)(zy / (uint32_t)gc.n), Y = (int)(zy - (uint32_t)Z * (uint32_t)gc.n);
    const int X = (int)(32u * w + sub);
    const uint32_t i = split_index(gc, X, Y, Z);
    const T v = restrict_node(F, gf, X, Y, Z, scale);
    Cc[i] = v;
    if (copy32) copy32[i] = (float)v;
  }
}

// ---- multi-instance kernel bodies (multi.h) ---------------------------------------------
template <class T, class TR>
struct ResidualGroup {
  struct Args {
    const T *X, *R, *S;
    TR* out;
    Geom g;
    uint32_t lanes;
    T pw, h2;
    double* partials;
    const uint32_t* live;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    residual_group_body<T, TR>(b, a.X, a.R, a.S, a.out, a.g, a.lanes, a.pw, a.h2, a.partials,
                               a.live);
  }
};

struct ResidualSignBits {
  struct Args {
    ResidualGroup<double, double>::Args r;
    const double* iso;
    uint32_t* bits;
    uint32_t bw;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    residual_group_body<double, double, true>(b, a.r.X, a.r.R, a.r.S, a.r.out, a.r.g, a.r.lanes,
                                              a.r.pw, a.r.h2, a.r.partials, a.r.live, a.iso,
                                              a.bits, a.bw);
  }
};

struct NormFinish {
  struct Args {
    const double* partials;
    int count;
    double* out;
  };
  static constexpr int kThreads = 1024;
  static __device__ void run(const Args& a, Blk b) {
    norm_finish_body(b, a.partials, a.count, a.out);
  }
};

template <class T>
struct Fill {
  struct Args {
    T* p;
    T v;
    uint64_t count;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) { fill_body<T>(b, a.p, a.v, a.count); }
};

template <class T>
struct RestrictNodes {
  struct Args {
    const T* F;
    Geom gf;
    T* Cc;
    Geom gc;
    T scale;
    float* copy32;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    restrict_nodes_body<T>(b, a.F, a.gf, a.Cc, a.gc, a.scale, a.copy32);
  }
};

template <class T>
struct RestrictSparse {
  struct Args {
    const T* F;
    Geom gf;
    T* Cc;
    Geom gc;
    T scale;
    float* copy32;
    const uint32_t *support, *list, *count;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    restrict_sparse_body<T>(b, a.F, a.gf, a.Cc, a.gc, a.scale, a.copy32, a.support, a.list,
                            a.count);
  }
};

struct CoarseSupport {
  struct Args {
    const uint32_t* fine;
    int nf, nc;
    uint32_t cw;
    uint32_t *out, *list, *count;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    coarse_support_body(b, a.fine, a.nf, a.nc, a.cw, a.out, a.list, a.count);
  }
};

struct NeedMask {
  struct Args {
    const uint32_t *cur, *prev;
    int64_t words, sy, sz;
    uint32_t* need;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    need_mask_body(b, a.cur, a.prev, a.words, a.sy, a.sz, a.need);
  }
};

struct NeedWords {
  struct Args {
    const uint32_t *need, *touched;
    uint32_t* prev;
    uint32_t words;
    uint32_t *list, *count;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    need_words_body(b, a.need, a.touched, a.prev, a.words, a.list, a.count);
  }
};

struct AssembleWords {
  struct Args {
    AssembleArgs a;
    Geom g;
    FastDiv div_n;
    const uint32_t *list, *count, *need;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    assemble_words_body(b, a.a, a.g, a.div_n, a.list, a.count, a.need);
  }
};

struct AssembleDense {
  struct Args {
    AssembleArgs a;
    Geom g;
    FastDiv div_n;
    uint32_t* prev;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    assemble_dense_body(b, a.a, a.g, a.div_n, a.prev);
  }
};

}  // namespace

inline uint32_t fill_blocks(uint64_t count) {
  return (uint32_t)std::min<uint64_t>(blocks_for(count, kThreads), 8192);
}

template <class T>
void launch_fill_multi(const FillJob<T>* jobs, size_t count, hipStream_t s) {
  std::vector<typename Fill<T>::Args> a(count);
  std::vector<uint32_t> nb(count);
  for (size_t i = 0; i < count; ++i) {
    a[i] = {jobs[i].p, jobs[i].value, jobs[i].count};
    nb[i] = jobs[i].count ? fill_blocks(jobs[i].count) : 0u;
  }
  launch_multi<Fill<T>>(a, nb, s);
}

template <class T>
void launch_fill(T* p, T value, uint64_t count, hipStream_t s) {
  if (!count) return;
  launch_one<Fill<T>>(typename Fill<T>::Args{p, value, count}, fill_blocks(count), s);
}

template <class T>
void launch_relax_color(T* X, const T* R, const T* S, const Geom& g, double pw, int c,
                        hipStream_t s) {
  const T h2 = (T)((double)(g.n - 1) * (double)(g.n - 1));
  if (group_layout(g)) {
    const uint32_t lanes = g.C / (uint32_t)(16 / sizeof(T));
    hipLaunchKernelGGL(k_relax_group<T>, dim3(xcd_grid(lanes, kThreads)), dim3(kThreads), 0, s, X,
                       R, S, g, lanes, c, (T)pw, h2);
    MPR_LAUNCH_CHECK();
    return;
  }
  const unsigned nb = xcd_grid(g.C, kThreads * kNodesPerThread);
  hipLaunchKernelGGL(k_relax_color<T>, dim3(nb), dim3(kThreads), 0, s, X, R, S, g, c, (T)pw, h2);
  MPR_LAUNCH_CHECK();
}

template <class T>
void launch_relax(T* X, const T* R, const T* S, const Geom& g, double pw, int iterations,
                  hipStream_t s) {
  for (int it = 0; it < iterations; ++it)
    for (int c = 0; c < 2; ++c) launch_relax_color<T>(X, R, S, g, pw, c, s);
}

template <class T, class TR>
int launch_residual(const T* X, const T* R, const T* S, TR* out, const Geom& g, double pw,
                    double* partials, hipStream_t s, const uint32_t* live) {
  const T h2 = (T)((double)(g.n - 1) * (double)(g.n - 1));
  if (group_layout(g)) {
    const uint32_t lanes = g.C / (uint32_t)(16 / sizeof(T));
    launch_one<ResidualGroup<T, TR>>({X, R, S, out, g, lanes, (T)pw, h2, partials, live},
                                     xcd_grid(lanes, kThreads), s);
    return (int)blocks_for(lanes, kThreads);
  }
  const unsigned nb = xcd_grid(2ull * g.C, kThreads * kNodesPerThread);
  hipLaunchKernelGGL((k_residual<T, TR>), dim3(nb), dim3(kThreads), 0, s, X, R, S, out, g,
                     (T)pw, h2, partials);
  MPR_LAUNCH_CHECK();
  return (int)blocks_for(2ull * g.C, kThreads * kNodesPerThread);
}

int residual_partial_count(const Geom& g, size_t word_bytes) {
  if (group_layout(g)) return (int)blocks_for(g.C / (uint32_t)(16 / word_bytes), kThreads);
  return (int)blocks_for(2ull * g.C, kThreads * kNodesPerThread);
}

void launch_norm_finish(const double* partials, int count, double* out, hipStream_t s) {
  launch_one<NormFinish>({partials, count, out}, 1, s);
}

bool residual_norm_sets_sign_bits(const Geom& g) { return group_layout(g); }

void launch_residual_norm_multi(const NormJob* jobs, size_t count, hipStream_t s) {
  std::vector<ResidualGroup<double, double>::Args> rg(count);
  std::vector<ResidualSignBits::Args> rs(count);
  std::vector<FillJob<uint32_t>> zero(count);
  std::vector<NormFinish::Args> nf(count);
  std::vector<uint32_t> nb(count, 0u), nbs(count, 0u), one(count, 1u);
  bool any_sign = false;
  for (size_t i = 0; i < count; ++i) {
    const NormJob& j = jobs[i];
    nf[i] = {j.partials, j.nparts, j.out};
    zero[i] = {nullptr, 0u, 0};
    if (!group_layout(j.g)) {  // small grids: the node-by-node kernel, launched alone
      launch_residual<double, double>(j.X, j.R, j.S, (double*)nullptr, j.g, j.pw, j.partials, s,
                                      j.live);
      continue;
    }
    const uint32_t lanes = j.g.C / 2u;
    const double h2 = (double)(j.g.n - 1) * (double)(j.g.n - 1);
    const ResidualGroup<double, double>::Args r{j.X, j.R, j.S, nullptr, j.g, lanes, j.pw, h2,
                                                j.partials, j.live};
    if (j.sign_bits) {
      const uint32_t bw = ((uint32_t)j.g.n + 31u) / 32u;
      zero[i] = {j.sign_bits, 0u, (uint64_t)j.g.n * j.g.n * bw};
      rs[i] = {r, j.iso, j.sign_bits, bw};
      nbs[i] = xcd_grid(lanes, kThreads);
      any_sign = true;
    } else {
      rg[i] = r;
      nb[i] = xcd_grid(lanes, kThreads);
    }
  }
  if (any_sign) launch_fill_multi(zero.data(), count, s);
  launch_multi<ResidualGroup<double, double>>(rg, nb, s);
  launch_multi<ResidualSignBits>(rs, nbs, s);
  launch_multi<NormFinish>(nf, one, s);
}

template <class T>
void launch_restrict(const T* F, const Geom& gf, T* Cc, const Geom& gc, double scale,
                     hipStream_t s, float* copy32) {
  // FP32 only: with 8-byte words the grouped form reads 22.5 loads per coarse node against
  // the scalar form's 27 and measured slower
  if (sizeof(T) == 4 && !copy32 && group_layout(gc) && group_layout(gf)) {
    const uint32_t lanes = gc.C / (uint32_t)(16 / sizeof(T));
    hipLaunchKernelGGL(k_restrict_group<T>, dim3(xcd_grid(lanes, kThreads)),
                       dim3(kThreads), 0, s, F, gf, Cc, gc, lanes, (T)scale);
    MPR_LAUNCH_CHECK();
    return;
  }
  const RestrictJob<T> j{F, gf, Cc, gc, scale, copy32, nullptr, nullptr};
  launch_restrict_nodes_multi(&j, 1, s);
}

template <class T>
void launch_restrict_nodes_multi(const RestrictJob<T>* jobs, size_t count, hipStream_t s) {
  std::vector<typename RestrictNodes<T>::Args> a(count);
  std::vector<typename RestrictSparse<T>::Args> sp(count);
  std::vector<CoarseSupport::Args> cs(count);
  std::vector<FillJob<uint32_t>> zero(count);
  std::vector<uint32_t> nb(count, 0u), nbs(count, 0u), nbc(count, 0u);
  bool any_sparse = false;
  for (size_t i = 0; i < count; ++i) {
    const RestrictJob<T>& j = jobs[i];
    zero[i] = {nullptr, 0u, 0};
    if (!j.F) continue;
    const uint32_t blocks = xcd_grid(2ull * j.gc.C, kThreads);
    if (j.need) {
      const uint32_t cw = ((uint32_t)j.gc.n + 31u) / 32u;
      const uint64_t words = (uint64_t)j.gc.n * j.gc.n * cw;
      // the support words, then the list of the non-zero ones and its length
      uint32_t* list = j.support + words;
      uint32_t* len = list + words;
      zero[i] = {len, 0u, 1};
      cs[i] = {j.need, j.gf.n, j.gc.n, cw, j.support, list, len};
      nbc[i] = blocks_for(words, kThreads);
      sp[i] = {j.F, j.gf, j.Cc, j.gc, (T)j.scale, j.copy32, j.support, list, len};
      // a fixed grid (graph-capturable) walking the list, 8 words per block and round
      nbs[i] = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(1, (words + 7) / 8), 2048);
      any_sparse = true;
    } else {
      a[i] = {j.F, j.gf, j.Cc, j.gc, (T)j.scale, j.copy32};
      nb[i] = blocks;
    }
  }
  launch_multi<RestrictNodes<T>>(a, nb, s);
  if (!any_sparse) return;
  launch_fill_multi(zero.data(), count, s);
  launch_multi<CoarseSupport>(cs, nbc, s);
  launch_multi<RestrictSparse<T>>(sp, nbs, s);
}

template <class T>
void launch_prolong_add(const T* Cc, const Geom& gc, T* F, const Geom& gf, hipStream_t s) {
  if (group_layout(gf)) {
    const uint32_t lanes = gf.C / (uint32_t)(16 / sizeof(T));
    hipLaunchKernelGGL(k_prolong_group<T>, dim3(xcd_grid(2ull * lanes, kThreads)), dim3(kThreads),
                       0, s, Cc, gc, F, gf, lanes, group_layout(gc) ? 1 : 0);
    MPR_LAUNCH_CHECK();
    return;
  }
  const unsigned nb = xcd_grid(2ull * gf.C, kThreads);
  hipLaunchKernelGGL(k_prolong_add<T>, dim3(nb), dim3(kThreads), 0, s, Cc, gc, F, gf);
  MPR_LAUNCH_CHECK();
}

template <class TI, class TO>
void launch_nat_to_split(const TI* nat, TO* split, const Geom& g, hipStream_t s) {
  const unsigned nb = xcd_grid(2ull * g.C, kThreads);
  hipLaunchKernelGGL((k_nat_to_split<TI, TO>), dim3(nb), dim3(kThreads), 0, s, nat, split, g);
  MPR_LAUNCH_CHECK();
}

template <class T>
void launch_split_to_nat(const T* split, double* nat, const Geom& g, hipStream_t s) {
  const uint64_t m = (uint64_t)g.n * g.n * g.n;
  const unsigned nb = xcd_grid(m, kThreads);
  hipLaunchKernelGGL(k_split_to_nat<T>, dim3(nb), dim3(kThreads), 0, s, split, nat, g,
                     FastDiv((uint32_t)g.n));
  MPR_LAUNCH_CHECK();
}

template <class TI, class TO>
void launch_convert(const TI* in, TO* out, uint64_t count, hipStream_t s) {
  if (!count) return;
  unsigned nb = (unsigned)std::min<uint64_t>(blocks_for(count, kThreads), 8192);
  hipLaunchKernelGGL((k_convert<TI, TO>), dim3(nb), dim3(kThreads), 0, s, in, out, count);
  MPR_LAUNCH_CHECK();
}

void launch_add_correction(double* X, const float* E, uint64_t count, hipStream_t s) {
  unsigned nb = (unsigned)std::min<uint64_t>(blocks_for(count, kThreads), 8192);
  hipLaunchKernelGGL(k_add_correction, dim3(nb), dim3(kThreads), 0, s, X, E, count);
  MPR_LAUNCH_CHECK();
}

void launch_assemble_multi(const AssembleJob* jobs, size_t count, hipStream_t s) {
  std::vector<AssembleDense::Args> dense(count);
  std::vector<NeedMask::Args> nm(count);
  std::vector<FillJob<uint32_t>> zero(count);
  std::vector<NeedWords::Args> nc(count);
  std::vector<AssembleWords::Args> ac(count);
  std::vector<uint32_t> nb_dense(count, 0u), nb_mask(count, 0u), nb_need(count, 0u),
      nb_asm(count, 0u);
  bool any_dense = false, any_sparse = false;
  for (size_t i = 0; i < count; ++i) {
    const AssembleJob& j = jobs[i];
    const Geom& g = j.g;
    const uint64_t m = (uint64_t)g.n * g.n * g.n;
    const AssembleArgs a{j.vx, j.vy, j.vz, j.screen_nat, j.rhs_nat, j.rhs_split, j.screen_split,
                         j.density, j.screen_nat_out, j.screen_split32, j.touched, j.live, j.spn,
                         j.pw, j.target};
    const TouchedHistory* hist = j.hist;
    if (hist && (!j.touched || !j.rhs_split))
      throw HipError{"assembly history needs the touched bits"};
    zero[i] = {nullptr, 0u, 0};
    if (!hist || hist->write_all) {
      dense[i] = {a, g, FastDiv((uint32_t)g.n), hist ? hist->prev : nullptr};
      nb_dense[i] = xcd_grid(m, kThreads);
      any_dense = true;
      continue;
    }
    // only the split outputs keep the zero-input value at untouched nodes; the
    // natural-layout outputs are written by the dense assembly
    if (j.rhs_nat || j.density || j.screen_nat_out)
      throw HipError{"sparse assembly writes the split outputs only"};
    const uint32_t words = (uint32_t)((m + 31) / 32);
    nm[i] = {j.touched, hist->prev, (int64_t)words, (int64_t)g.n, (int64_t)g.n * g.n, hist->need};
    nb_mask[i] = blocks_for(words, kThreads);
    zero[i] = {hist->count, 0u, 1};
    nc[i] = {hist->need, j.touched, hist->prev, words, hist->list, hist->count};
    nb_need[i] = blocks_for(words, kThreads * kNeedWordsPerThread);
    ac[i] = {a, g, FastDiv((uint32_t)g.n), hist->list, hist->count, hist->need};
    // a fixed grid (graph-capturable) walking the word list, 8 words per block and round
    nb_asm[i] = std::min<unsigned>(blocks_for(words, 8), 4096u);
    any_sparse = true;
  }
  if (any_dense) launch_multi<AssembleDense>(dense, nb_dense, s);
  if (!any_sparse) return;
  launch_multi<NeedMask>(nm, nb_mask, s);
  launch_fill_multi(zero.data(), count, s);
  launch_multi<NeedWords>(nc, nb_need, s);
  launch_multi<AssembleWords>(ac, nb_asm, s);
}

void launch_assemble(const double* vx, const double* vy, const double* vz,
                     const double* screen_nat, double* rhs_nat, double* rhs_split,
                     double* screen_split, double* density, double spn, const Geom& g,
                     double pw, double target, const uint32_t* touched, double* screen_nat_out,
                     float* screen_split32, const TouchedHistory* hist, uint32_t* live,
                     hipStream_t s) {
  const AssembleJob j{vx, vy, vz, screen_nat, rhs_nat, rhs_split, screen_split, density, spn, g,
                      pw, target, touched, screen_nat_out, screen_split32, hist, live};
  launch_assemble_multi(&j, 1, s);
}

#define MPR_INST(T)                                                                          \
  template void launch_fill<T>(T*, T, uint64_t, hipStream_t);                                \
  template void launch_relax<T>(T*, const T*, const T*, const Geom&, double, int, hipStream_t); \
  template void launch_relax_color<T>(T*, const T*, const T*, const Geom&, double, int,         \
                                      hipStream_t);                                            \
  template void launch_restrict<T>(const T*, const Geom&, T*, const Geom&, double, hipStream_t,  \
                                   float*);                                                     \
  template void launch_prolong_add<T>(const T*, const Geom&, T*, const Geom&, hipStream_t);  \
  template void launch_split_to_nat<T>(const T*, double*, const Geom&, hipStream_t);
MPR_INST(double)
MPR_INST(float)
#undef MPR_INST
template void launch_fill<int32_t>(int32_t*, int32_t, uint64_t, hipStream_t);
template void launch_fill_multi<int32_t>(const FillJob<int32_t>*, size_t, hipStream_t);
template void launch_fill_multi<uint32_t>(const FillJob<uint32_t>*, size_t, hipStream_t);
template void launch_fill_multi<double>(const FillJob<double>*, size_t, hipStream_t);
template void launch_fill_multi<unsigned long long>(const FillJob<unsigned long long>*, size_t,
                                                    hipStream_t);
template void launch_restrict_nodes_multi<double>(const RestrictJob<double>*, size_t,
                                                   hipStream_t);
template void launch_fill<uint32_t>(uint32_t*, uint32_t, uint64_t, hipStream_t);
template void launch_fill<int64_t>(int64_t*, int64_t, uint64_t, hipStream_t);
template void launch_fill<unsigned long long>(unsigned long long*, unsigned long long, uint64_t,
                                              hipStream_t);
template int launch_residual<double, double>(const double*, const double*, const double*,
                                             double*, const Geom&, double, double*, hipStream_t,
                                             const uint32_t*);
template int launch_residual<double, float>(const double*, const double*, const double*, float*,
                                            const Geom&, double, double*, hipStream_t,
                                            const uint32_t*);
template int launch_residual<float, float>(const float*, const float*, const float*, float*,
                                           const Geom&, double, double*, hipStream_t,
                                           const uint32_t*);
template void launch_nat_to_split<double, double>(const double*, double*, const Geom&,
                                                  hipStream_t);
template void launch_nat_to_split<double, float>(const double*, float*, const Geom&,
                                                 hipStream_t);
template void launch_convert<double, float>(const double*, float*, uint64_t, hipStream_t);
template void launch_convert<float, double>(const float*, double*, uint64_t, hipStream_t);

}  // namespace mpr
