// Dual-grid (surface-nets) iso-surface extraction for gfx950 (capi.mojo:798-1129).
//
// The reference assigns vertex indices in raster order with a serial loop and emits faces
// in three serial edge passes. Here both become order-preserving stream compactions:
// a count kernel (one 4096-item tile per workgroup, wave64 ballot + popcount), a one-
// workgroup exclusive scan of the tile counts, and an emit kernel that re-derives each
// flag and its rank inside the tile with ballots, so indices and face order are exactly
// the reference's.
#include <algorithm>

#include "block_scan.h"
#include "mpr_internal.h"
#include "multi.h"

namespace mpr {

namespace {

constexpr int kTileItems = 16;
constexpr uint32_t kTile = kThreads * kTileItems;

__device__ __forceinline__ bool cell_active(const double* f, int x, int y, int z, int n,
                                            double iso) {
  int below = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    below += f[nat_index(x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1), n)] < iso;
  return below != 0 && below != 8;
}

__device__ __forceinline__ void cell_decode(uint32_t i, uint32_t cells, const FastDiv& dc, int& x,
                                            int& y, int& z) {
  const uint32_t q = dc.div(i);
  x = (int)(i - q * cells);
  const uint32_t zz = dc.div(q);
  y = (int)(q - zz * cells);
  z = (int)zz;
}

__global__ __launch_bounds__(kThreads) void k_cells_count(const double* f, int n, const double* isop,
                                                          FastDiv dc, uint32_t* counts) {
  __shared__ uint32_t lds4[4];
  const double iso = *isop;
  const uint32_t cells = (uint32_t)(n - 1), total = cells * cells * cells;
  uint32_t c = 0;
  for (int r = 0; r < kTileItems; ++r) {
    const uint32_t i = blockIdx.x * kTile + r * kThreads + threadIdx.x;
    if (i < total) {
      int x, y, z;
      cell_decode(i, cells, dc, x, y, z);
      c += cell_active(f, x, y, z, n, iso);
    }
  }
  c = block_count256(c, lds4);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// count mode of the compatibility ABI: one partial count per 4-plane z chunk
__global__ __launch_bounds__(kThreads) void k_cells_chunk_counts(const double* f, int n,
                                                                 const double* isop, FastDiv dc,
                                                                 unsigned long long* chunks) {
  const double iso = *isop;
  const uint32_t cells = (uint32_t)(n - 1);
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= cells * cells * cells) return;
  int x, y, z;
  cell_decode(i, cells, dc, x, y, z);
  if (cell_active(f, x, y, z, n, iso)) atomicAdd(&chunks[z >> 2], 1ull);
}

// vertex of an active cell: mean of the edge roots in x-, y-, z-axis order (capi.mojo:958-1038)
__device__ __forceinline__ void dual_vertex(const FieldView& f, const DensityView& dens, int x, int y,
                                            int z, int n, double iso, double out[3],
                                            double& vdens) {
  double v[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) v[c] = f.at(x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1));
  double sx = 0.0, sy = 0.0, sz = 0.0;
  int roots = 0;
#pragma unroll
  for (int axis = 0; axis < 3; ++axis)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        int c0, c1;
        double x0, y0, z0, x1, y1, z1;
        if (axis == 0) {
          c0 = 2 * a + 4 * b; c1 = c0 + 1;
          x0 = 0.0; y0 = (double)a; z0 = (double)b; x1 = 1.0; y1 = y0; z1 = z0;
        } else if (axis == 1) {
          c0 = a + 4 * b; c1 = c0 + 2;
          x0 = (double)a; y0 = 0.0; z0 = (double)b; x1 = x0; y1 = 1.0; z1 = z0;
        } else {
          c0 = a + 2 * b; c1 = c0 + 4;
          x0 = (double)a; y0 = (double)b; z0 = 0.0; x1 = x0; y1 = y0; z1 = 1.0;
        }
        const double v0 = v[c0], v1 = v[c1];
        if ((v0 < iso) == (v1 < iso)) continue;
        double t = (iso - v0) / (v1 - v0);
        t = t < 0.0 ? 0.0 : t;
        t = t > 1.0 ? 1.0 : t;
        sx = sx + x0 + (x1 - x0) * t;
        sy = sy + y0 + (y1 - y0) * t;
        sz = sz + z0 + (z1 - z0) * t;
        roots += 1;
      }
  const double inv = 1.0 / (double)roots;
  const double cells = (double)(n - 1);
  out[0] = ((double)x + sx * inv) / cells;
  out[1] = ((double)y + sy * inv) / cells;
  out[2] = ((double)z + sz * inv) / cells;
  double ds = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    ds += dens.at(nat_index(x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1), n));
  vdens = ds * 0.125;
}

// Fill mode: active cells get consecutive raster-order indices. Cells whose index reaches
// `capacity` keep the mark 0 and emit nothing (capi.mojo:905-963). If `xform` is non-null
// ([origin x, y, z, width]) vertices are written as v * width + origin (core.py:425).
template <class CM>
__global__ __launch_bounds__(kThreads) void k_cells_emit(const double* f, const double* dens,
                                                         int n, const double* isop,
                                                         FastDiv dc,
                                                         const uint32_t* tile_offsets,
                                                         CM* cell_map, double* verts,
                                                         double* vdens, uint64_t capacity,
                                                         const double* xform) {
  __shared__ uint32_t lds4[4];
  const double iso = *isop;
  const uint32_t cells = (uint32_t)(n - 1), total = cells * cells * cells;
  uint32_t running = tile_offsets[blockIdx.x];
  for (int r = 0; r < kTileItems; ++r) {
    const uint32_t i = blockIdx.x * kTile + r * kThreads + threadIdx.x;
    int x = 0, y = 0, z = 0;
    bool act = false;
    if (i < total) {
      cell_decode(i, cells, dc, x, y, z);
      act = cell_active(f, x, y, z, n, iso);
    }
    uint32_t cnt;
    const uint32_t pre = block_prefix256(act, lds4, cnt);
    if (i < total) {
      if (act) {
        const uint64_t idx = (uint64_t)running + pre;
        if (idx < capacity) {
          double v[3], d;
          FieldView fv;
          fv.nat = f;
          fv.g.n = n;
          dual_vertex(fv, DensityView{dens, nullptr, 1.0}, x, y, z, n, iso, v, d);
          if (xform) {
            for (int a = 0; a < 3; ++a) v[a] = v[a] * xform[3] + xform[a];
          }
          verts[3 * idx] = v[0];
          verts[3 * idx + 1] = v[1];
          verts[3 * idx + 2] = v[2];
          vdens[idx] = d;
          cell_map[i] = (CM)idx;
        } else {
          cell_map[i] = (CM)0;
        }
      } else {
        cell_map[i] = (CM)-1;
      }
    }
    running += cnt;
  }
}

struct EdgeGeom {
  uint32_t cells, per_pass;  // per_pass = (cells-1)^2 * cells
  FastDiv d_pass, d_cells, d_cm1;
};

// Decodes edge item e of the three passes (capi.mojo:1090-1128) into the edge's lower node
// and the four incident cells (a, b, c, d) in the reference's quad order.
__device__ __forceinline__ void edge_decode(uint32_t e, const EdgeGeom& G, int& axis, int& x,
                                            int& y, int& z, int q[4][3]) {
  const uint32_t cells = G.cells, cm1 = cells - 1;
  axis = (int)G.d_pass.div(e);
  const uint32_t j = e - (uint32_t)axis * G.per_pass;
  if (axis == 0) {
    const uint32_t r = G.d_cells.div(j);
    x = (int)(j - r * cells);
    const uint32_t zz = G.d_cm1.div(r);
    y = 1 + (int)(r - zz * cm1);
    z = 1 + (int)zz;
    const int t[4][3] = {{x, y - 1, z - 1}, {x, y, z - 1}, {x, y, z}, {x, y - 1, z}};
    for (int k = 0; k < 4; ++k) for (int a = 0; a < 3; ++a) q[k][a] = t[k][a];
  } else if (axis == 1) {
    const uint32_t r = G.d_cm1.div(j);
    x = 1 + (int)(j - r * cm1);
    const uint32_t zz = G.d_cells.div(r);
    y = (int)(r - zz * cells);
    z = 1 + (int)zz;
    const int t[4][3] = {{x - 1, y, z - 1}, {x - 1, y, z}, {x, y, z}, {x, y, z - 1}};
    for (int k = 0; k < 4; ++k) for (int a = 0; a < 3; ++a) q[k][a] = t[k][a];
  } else {
    const uint32_t r = G.d_cm1.div(j);
    x = 1 + (int)(j - r * cm1);
    const uint32_t zz = G.d_cm1.div(r);
    y = 1 + (int)(r - zz * cm1);
    z = (int)zz;
    const int t[4][3] = {{x - 1, y - 1, z}, {x, y - 1, z}, {x, y, z}, {x - 1, y, z}};
    for (int k = 0; k < 4; ++k) for (int a = 0; a < 3; ++a) q[k][a] = t[k][a];
  }
}

template <class CM>
__device__ __forceinline__ bool edge_quad(const double* f, int n, double iso, const CM* cm,
                                          const EdgeGeom& G, uint32_t e, int64_t idx[4],
                                          bool& reverse) {
  int axis, x, y, z, q[4][3];
  edge_decode(e, G, axis, x, y, z, q);
  const double v0 = f[nat_index(x, y, z, n)];
  const double v1 = f[nat_index(x + (axis == 0), y + (axis == 1), z + (axis == 2), n)];
  reverse = v0 < iso;
  if ((v0 < iso) == (v1 < iso)) return false;
  const uint32_t cells = G.cells;
  bool ok = true;
  for (int k = 0; k < 4; ++k) {
    idx[k] = (int64_t)cm[((uint64_t)q[k][2] * cells + q[k][1]) * cells + q[k][0]];
    ok = ok && idx[k] >= 0;
  }
  return ok;
}

template <class CM>
__global__ __launch_bounds__(kThreads) void k_edges_count(const double* f, int n,
                                                          const double* isop, const CM* cm,
                                                          EdgeGeom G, uint32_t* counts) {
  __shared__ uint32_t lds4[4];
  const double iso = *isop;
  const uint32_t total = 3u * G.per_pass;
  uint32_t c = 0;
  for (int r = 0; r < kTileItems; ++r) {
    const uint32_t e = blockIdx.x * kTile + r * kThreads + threadIdx.x;
    if (e < total) {
      int64_t idx[4];
      bool rev;
      c += edge_quad(f, n, iso, cm, G, e, idx, rev);
    }
  }
  c = block_count256(c, lds4);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// capi.mojo:1044-1072: quad (a,b,c,d) -> [a,b,c],[a,c,d], or [a,d,c],[a,c,b] when v0 < iso
template <class CM>
__global__ __launch_bounds__(kThreads) void k_edges_emit(const double* f, int n,
                                                         const double* isop, const CM* cm,
                                                         EdgeGeom G,
                                                         const uint32_t* tile_offsets,
                                                         int64_t* faces, uint64_t capacity) {
  __shared__ uint32_t lds4[4];
  const double iso = *isop;
  const uint32_t total = 3u * G.per_pass;
  uint32_t running = tile_offsets[blockIdx.x];
  for (int r = 0; r < kTileItems; ++r) {
    const uint32_t e = blockIdx.x * kTile + r * kThreads + threadIdx.x;
    int64_t q[4] = {0, 0, 0, 0};
    bool rev = false, on = false;
    if (e < total) on = edge_quad(f, n, iso, cm, G, e, q, rev);
    uint32_t cnt;
    const uint32_t pre = block_prefix256(on, lds4, cnt);
    if (on) {
      const uint64_t fc = 2ull * ((uint64_t)running + pre);
      if (fc + 2 <= capacity) {
        int64_t* o = faces + 3 * fc;
        if (rev) {
          o[0] = q[0]; o[1] = q[3]; o[2] = q[2]; o[3] = q[0]; o[4] = q[2]; o[5] = q[1];
        } else {
          o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = q[0]; o[4] = q[2]; o[5] = q[3];
        }
      }
    }
    running += cnt;
  }
}

// ---- sign-bit extraction (the device pipeline) -----------------------------------------
// One pass packs s(x,y,z) = field < iso into 32-bit words along x (word w of node row
// (y,z) holds x = 32w .. 32w+31). Cell activity (some but not all of the 8 corner bits
// set) and edge sign changes are then whole-word operations, 32 cells or edges per lane, in
// the reference's raster and pass order. A cell's vertex index is the exclusive count of
// active cells before its 32-cell run plus the active cells before it in the run, so no
// per-cell index map is stored; the field and density are read only at active cells.
struct BitGeom {
  uint32_t n, cells, bw, cw;  // node words and cell runs per row
  uint32_t r0, r1, r2;        // rows of the three edge passes
  FastDiv d_bw, d_cw, d_cells, d_cm1, d_n;
};

__device__ __forceinline__ uint32_t node_word(const uint32_t* bits, const BitGeom& G, uint32_t y,
                                              uint32_t z, uint32_t w) {
  return w < G.bw ? bits[(z * G.n + y) * G.bw + w] : 0u;
}

// bits of nodes x = 32w+k (lo) and x = 32w+k+1 (hi) of node row (y,z)
__device__ __forceinline__ void node_pair(const uint32_t* bits, const BitGeom& G, uint32_t y,
                                          uint32_t z, uint32_t w, uint32_t& lo, uint32_t& hi) {
  lo = node_word(bits, G, y, z, w);
  hi = (lo >> 1) | (node_word(bits, G, y, z, w + 1) << 31);
}

__device__ __forceinline__ uint32_t low_bits(uint32_t count) {
  return count >= 32u ? 0xffffffffu : ((1u << count) - 1u);
}

__device__ __forceinline__ void sign_bits_body(Blk blk_, FieldView f, const double* isop,
                                                        BitGeom G, uint32_t* bits) {
  const uint64_t i = (uint64_t)blk_.id * kThreads + threadIdx.x;
  const uint32_t wi = (uint32_t)(i >> 5), k = (uint32_t)(i & 31u);
  const uint32_t words = G.n * G.n * G.bw;
  const uint32_t row = G.d_bw.div(wi < words ? wi : 0u);
  const uint32_t x = 32u * (wi - row * G.bw) + k;
  const bool in = wi < words && x < G.n;
  const uint32_t z = G.d_n.div(row), y = row - z * G.n;
  const bool below = in && f.at(in ? (int)x : 0, (int)y, (int)z) < *isop;
  const uint64_t bal = __ballot(below);
  if (k == 0 && wi < words) bits[wi] = (uint32_t)(bal >> (threadIdx.x & 32u));
}

// active-cell mask of cell run r (cells x = 32w .. 32w+31 of cell row (y,z))
__device__ __forceinline__ uint32_t cell_mask(const uint32_t* bits, const BitGeom& G, uint32_t r,
                                              uint32_t& y, uint32_t& z, uint32_t& w) {
  const uint32_t row = G.d_cw.div(r);
  w = r - row * G.cw;
  z = G.d_cells.div(row);
  y = row - z * G.cells;
  uint32_t all = 0xffffffffu, any = 0u;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    uint32_t lo, hi;
    node_pair(bits, G, y + (c & 1), z + (c >> 1), w, lo, hi);
    all &= lo & hi;
    any |= lo | hi;
  }
  return (any & ~all) & low_bits(G.cells - 32u * w);
}

__device__ __forceinline__ void bit_cells_count_body(Blk blk_, const uint32_t* bits, BitGeom G,
                                                              uint32_t* cmask, uint32_t* counts) {
  __shared__ uint32_t lds4[4];
  const uint32_t r = blk_.id * kThreads + threadIdx.x;
  uint32_t m = 0;
  if (r < G.cells * G.cells * G.cw) {
    uint32_t y, z, w;
    m = cell_mask(bits, G, r, y, z, w);
    cmask[r] = m;
  }
  const uint32_t c = block_count256((uint32_t)__popc(m), lds4);
  if (threadIdx.x == 0) counts[blk_.id] = c;
}

// Raster-order indices of the active cells: coff[r] = index of run r's first vertex, and
// vcell[idx] = the cell (x, run) of vertex idx for the vertices below the capacity.
__device__ __forceinline__ void bit_cells_index_body(Blk blk_, BitGeom G, const uint32_t* cmask,
                                                              const uint32_t* tile_offsets,
                                                              uint32_t* coff, uint32_t* vcell,
                                                              uint64_t capacity) {
  __shared__ uint32_t lds4[4];
  const uint32_t r = blk_.id * kThreads + threadIdx.x;
  const bool live = r < G.cells * G.cells * G.cw;
  uint32_t m = live ? cmask[r] : 0u;
  uint32_t total;
  const uint32_t pre = block_exclusive256((uint32_t)__popc(m), lds4, total);
  if (!live) return;
  const uint32_t base = tile_offsets[blk_.id] + pre;
  coff[r] = base;
  uint64_t idx = base;
  while (m && idx < capacity) {
    const uint32_t k = (uint32_t)__ffs(m) - 1u;
    m &= m - 1u;
    vcell[idx++] = r * 32u + k;
  }
}

// one lane per vertex: the dual vertex of its cell (capi.mojo:958-1038)
__device__ __forceinline__ void bit_vertices_body(Blk blk_, FieldView f, DensityView dens,
                                                           const double* isop, BitGeom G,
                                                           const uint32_t* vcell,
                                                           const uint32_t* count, double* verts,
                                                           double* vdens, uint64_t capacity,
                                                           const double* xform) {
  const uint64_t idx = (uint64_t)blk_.id * kThreads + threadIdx.x;
  if (idx >= capacity || idx >= *count) return;
  const uint32_t c = vcell[idx];
  const uint32_t r = c >> 5, k = c & 31u;
  const uint32_t row = G.d_cw.div(r);
  const uint32_t w = r - row * G.cw;
  const uint32_t z = G.d_cells.div(row);
  const uint32_t y = row - z * G.cells;
  double v[3], d;
  dual_vertex(f, dens, (int)(32u * w + k), (int)y, (int)z, (int)G.n, *isop, v, d);
  if (xform) {
    for (int a = 0; a < 3; ++a) v[a] = v[a] * xform[3] + xform[a];
  }
  verts[3 * idx] = v[0];
  verts[3 * idx + 1] = v[1];
  verts[3 * idx + 2] = v[2];
  vdens[idx] = d;
}

// Sign-change mask of edge run e over the three passes (capi.mojo:1090-1128): pass 0 edges
// along x at y,z in [1, cells-1], x in [0, cells-1]; pass 1 along y at z in [1, cells-1],
// y in [0, cells-1], x in [1, cells-1]; pass 2 along z at z in [0, cells-1], y, x in
// [1, cells-1]; rows in raster (z, y) order. `below` receives the lower node's bits (the
// face orientation, capi.mojo:1044-1072).
__device__ __forceinline__ uint32_t edge_mask(const uint32_t* bits, const BitGeom& G, uint32_t e,
                                              int& axis, uint32_t& y, uint32_t& z, uint32_t& w,
                                              uint32_t& below) {
  const uint32_t c = G.cells, cm1 = c - 1u;
  uint32_t row;
  if (e < G.r0 * G.cw) {
    axis = 0;
    row = G.d_cw.div(e);
    w = e - row * G.cw;
    const uint32_t zz = G.d_cm1.div(row);
    y = 1u + (row - zz * cm1);
    z = 1u + zz;
    uint32_t lo, hi;
    node_pair(bits, G, y, z, w, lo, hi);
    below = lo;
    return (lo ^ hi) & low_bits(c - 32u * w);
  }
  e -= G.r0 * G.cw;
  uint32_t other;
  if (e < G.r1 * G.cw) {
    axis = 1;
    row = G.d_cw.div(e);
    w = e - row * G.cw;
    const uint32_t zz = G.d_cells.div(row);
    y = row - zz * c;
    z = 1u + zz;
    other = node_word(bits, G, y + 1u, z, w);
  } else {
    e -= G.r1 * G.cw;
    axis = 2;
    row = G.d_cw.div(e);
    w = e - row * G.cw;
    const uint32_t zz = G.d_cm1.div(row);
    y = 1u + (row - zz * cm1);
    z = zz;
    other = node_word(bits, G, y, z + 1u, w);
  }
  below = node_word(bits, G, y, z, w);
  const uint32_t valid = low_bits(c - 32u * w) & (w == 0 ? ~1u : 0xffffffffu);  // x >= 1
  return (below ^ other) & valid;
}

__device__ __forceinline__ uint32_t edge_runs(const BitGeom& G) {
  return (G.r0 + G.r1 + G.r2) * G.cw;
}

__device__ __forceinline__ void bit_edges_count_body(Blk blk_, const uint32_t* bits, BitGeom G,
                                                              uint32_t* counts) {
  __shared__ uint32_t lds4[4];
  const uint32_t e = blk_.id * kThreads + threadIdx.x;
  uint32_t m = 0;
  if (e < edge_runs(G)) {
    int axis;
    uint32_t y, z, w, below;
    m = edge_mask(bits, G, e, axis, y, z, w, below);
  }
  const uint32_t c = block_count256((uint32_t)__popc(m), lds4);
  if (threadIdx.x == 0) counts[blk_.id] = c;
}

__device__ __forceinline__ int64_t vertex_of(const BitGeom& G, const uint32_t* cmask,
                                             const uint32_t* coff, uint32_t x, uint32_t y,
                                             uint32_t z) {
  const uint32_t r = (z * G.cells + y) * G.cw + (x >> 5);
  return (int64_t)coff[r] + __popc(cmask[r] & low_bits(x & 31u));
}

__device__ __forceinline__ void bit_edges_emit_body(Blk blk_, const uint32_t* bits, BitGeom G,
                                                             const uint32_t* cmask,
                                                             const uint32_t* coff,
                                                             const uint32_t* tile_offsets,
                                                             int64_t* faces, uint64_t capacity) {
  __shared__ uint32_t lds4[4];
  const uint32_t e = blk_.id * kThreads + threadIdx.x;
  uint32_t m = 0, y = 0, z = 0, w = 0, below = 0;
  int axis = 0;
  if (e < edge_runs(G)) m = edge_mask(bits, G, e, axis, y, z, w, below);
  uint32_t total;
  const uint32_t pre = block_exclusive256((uint32_t)__popc(m), lds4, total);
  uint64_t quad = (uint64_t)tile_offsets[blk_.id] + pre;
  while (m) {
    const uint32_t k = (uint32_t)__ffs(m) - 1u;
    m &= m - 1u;
    const uint64_t fc = 2ull * quad++;
    if (fc + 2 > capacity) continue;
    const uint32_t x = 32u * w + k;
    uint32_t cx[4], cy[4], cz[4];
    if (axis == 0) {
      const uint32_t t[4][3] = {{x, y - 1, z - 1}, {x, y, z - 1}, {x, y, z}, {x, y - 1, z}};
      for (int j = 0; j < 4; ++j) cx[j] = t[j][0], cy[j] = t[j][1], cz[j] = t[j][2];
    } else if (axis == 1) {
      const uint32_t t[4][3] = {{x - 1, y, z - 1}, {x - 1, y, z}, {x, y, z}, {x, y, z - 1}};
      for (int j = 0; j < 4; ++j) cx[j] = t[j][0], cy[j] = t[j][1], cz[j] = t[j][2];
    } else {
      const uint32_t t[4][3] = {{x - 1, y - 1, z}, {x, y - 1, z}, {x, y, z}, {x - 1, y, z}};
      for (int j = 0; j < 4; ++j) cx[j] = t[j][0], cy[j] = t[j][1], cz[j] = t[j][2];
    }
    int64_t q[4];
    for (int j = 0; j < 4; ++j) q[j] = vertex_of(G, cmask, coff, cx[j], cy[j], cz[j]);
    int64_t* o = faces + 3 * fc;
    if ((below >> k) & 1u) {
      o[0] = q[0]; o[1] = q[3]; o[2] = q[2]; o[3] = q[0]; o[4] = q[2]; o[5] = q[1];
    } else {
      o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = q[0]; o[4] = q[2]; o[5] = q[3];
    }
  }
}

BitGeom make_bit_geom(int n) {
  BitGeom G;
  G.n = (uint32_t)n;
  G.cells = (uint32_t)(n - 1);
  G.bw = (G.n + 31u) / 32u;
  G.cw = std::max<uint32_t>(1u, (G.cells + 31u) / 32u);
  const uint32_t c = G.cells, cm1 = c ? c - 1u : 0u;
  G.r0 = cm1 * cm1;
  G.r1 = cm1 * c;
  G.r2 = c * cm1;
  G.d_bw = FastDiv(G.bw);
  G.d_cw = FastDiv(G.cw);
  G.d_cells = FastDiv(std::max<uint32_t>(1u, c));
  G.d_cm1 = FastDiv(std::max<uint32_t>(1u, cm1));
  G.d_n = FastDiv(G.n);
  return G;
}

EdgeGeom make_edge_geom(int n) {
  EdgeGeom G;
  G.cells = (uint32_t)(n - 1);
  G.per_pass = (G.cells - 1) * (G.cells - 1) * G.cells;
  G.d_pass = FastDiv(std::max<uint32_t>(1u, G.per_pass));
  G.d_cells = FastDiv(std::max<uint32_t>(1u, G.cells));
  G.d_cm1 = FastDiv(std::max<uint32_t>(1u, G.cells - 1));
  return G;
}

// ---- multi-instance kernel bodies (multi.h) ---------------------------------------------
struct SignBits {
  struct Args {
    FieldView f;
    const double* iso;
    BitGeom G;
    uint32_t* bits;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) { sign_bits_body(b, a.f, a.iso, a.G, a.bits); }
};

struct BitCellsCount {
  struct Args {
    const uint32_t* bits;
    BitGeom G;
    uint32_t *cmask, *counts;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    bit_cells_count_body(b, a.bits, a.G, a.cmask, a.counts);
  }
};

struct BitCellsIndex {
  struct Args {
    BitGeom G;
    const uint32_t *cmask, *tile_offsets;
    uint32_t *coff, *vcell;
    uint64_t capacity;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    bit_cells_index_body(b, a.G, a.cmask, a.tile_offsets, a.coff, a.vcell, a.capacity);
  }
};

struct BitVertices {
  struct Args {
    FieldView f;
    DensityView dens;
    const double* iso;
    BitGeom G;
    const uint32_t *vcell, *count;
    double *verts, *vdens;
    uint64_t capacity;
    const double* xform;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    bit_vertices_body(b, a.f, a.dens, a.iso, a.G, a.vcell, a.count, a.verts, a.vdens, a.capacity,
                      a.xform);
  }
};

struct BitEdgesCount {
  struct Args {
    const uint32_t* bits;
    BitGeom G;
    uint32_t* counts;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    bit_edges_count_body(b, a.bits, a.G, a.counts);
  }
};

struct BitEdgesEmit {
  struct Args {
    const uint32_t* bits;
    BitGeom G;
    const uint32_t *cmask, *coff, *tile_offsets;
    int64_t* faces;
    uint64_t capacity;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    bit_edges_emit_body(b, a.bits, a.G, a.cmask, a.coff, a.tile_offsets, a.faces, a.capacity);
  }
};

}  // namespace

uint32_t cell_tiles(int n) {
  const uint64_t cells = (uint64_t)(n - 1);
  return (uint32_t)std::max<uint64_t>(1, (cells * cells * cells + kTile - 1) / kTile);
}

uint32_t edge_tiles(int n) {
  const uint64_t cells = (uint64_t)(n - 1);
  const uint64_t e = cells ? 3ull * (cells - 1) * (cells - 1) * cells : 0;
  return (uint32_t)std::max<uint64_t>(1, (e + kTile - 1) / kTile);
}

void launch_cells_count(const double* f, int n, const double* iso, uint32_t* counts,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_cells_count, dim3(cell_tiles(n)), dim3(kThreads), 0, s, f, n, iso,
                     FastDiv((uint32_t)(n - 1)), counts);
  MPR_LAUNCH_CHECK();
}

void launch_cells_chunk_counts(const double* f, int n, const double* iso, int64_t* chunks,
                               hipStream_t s) {
  const uint64_t cells = (uint64_t)(n - 1);
  const uint64_t total = cells * cells * cells;
  if (!total) return;
  hipLaunchKernelGGL(k_cells_chunk_counts, dim3(blocks_for(total, kThreads)), dim3(kThreads), 0,
                     s, f, n, iso, FastDiv((uint32_t)cells), (unsigned long long*)chunks);
  MPR_LAUNCH_CHECK();
}

template <class CM>
void launch_cells_emit(const double* f, const double* dens, int n, const double* iso,
                       const uint32_t* tile_offsets, CM* cell_map, double* verts, double* vdens,
                       uint64_t capacity, const double* xform, hipStream_t s) {
  hipLaunchKernelGGL(k_cells_emit<CM>, dim3(cell_tiles(n)), dim3(kThreads), 0, s, f, dens, n, iso,
                     FastDiv((uint32_t)(n - 1)), tile_offsets, cell_map, verts, vdens, capacity,
                     xform);
  MPR_LAUNCH_CHECK();
}

template <class CM>
void launch_edges_count(const double* f, int n, const double* iso, const CM* cm,
                        uint32_t* counts,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_edges_count<CM>, dim3(edge_tiles(n)), dim3(kThreads), 0, s, f, n, iso, cm,
                     make_edge_geom(n), counts);
  MPR_LAUNCH_CHECK();
}

template <class CM>
void launch_edges_emit(const double* f, int n, const double* iso, const CM* cm,
                       const uint32_t* tile_offsets, int64_t* faces, uint64_t capacity,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_edges_emit<CM>, dim3(edge_tiles(n)), dim3(kThreads), 0, s, f, n, iso, cm,
                     make_edge_geom(n), tile_offsets, faces, capacity);
  MPR_LAUNCH_CHECK();
}

BitSurface bit_surface_sizes(int n) {
  const BitGeom G = make_bit_geom(n);
  BitSurface b;
  b.bit_words = (uint64_t)G.n * G.n * G.bw;
  b.runs = (uint64_t)G.cells * G.cells * G.cw;
  b.cell_tiles = (uint32_t)std::max<uint64_t>(1, (b.runs + kThreads - 1) / kThreads);
  b.edge_tiles = (uint32_t)std::max<uint64_t>(
      1, ((uint64_t)(G.r0 + G.r1 + G.r2) * G.cw + kThreads - 1) / kThreads);
  return b;
}

void launch_bit_extract_multi(const ExtractJob* jobs, size_t count, hipStream_t s) {
  std::vector<SignBits::Args> sb(count);
  std::vector<BitCellsCount::Args> cc(count);
  std::vector<BitCellsIndex::Args> ci(count);
  std::vector<BitVertices::Args> bv(count);
  std::vector<BitEdgesCount::Args> ec(count);
  std::vector<BitEdgesEmit::Args> ee(count);
  std::vector<ScanJob> scan_cells(count), scan_edges(count);
  std::vector<uint32_t> nb_sign(count), nb_cells(count), nb_vert(count), nb_edges(count);
  for (size_t i = 0; i < count; ++i) {
    const ExtractJob& j = jobs[i];
    const BitSurface& b = *j.b;
    const BitGeom G = make_bit_geom(j.n);
    const uint64_t lanes = b.bit_words * 32ull;
    const uint64_t vlanes = std::min<uint64_t>(j.vcap, b.vcell_cap);
    sb[i] = {j.f, j.iso, G, b.bits};
    nb_sign[i] = j.bits_ready ? 0u : (uint32_t)((lanes + kThreads - 1) / kThreads);
    cc[i] = {b.bits, G, b.cmask, b.cell_tc};
    nb_cells[i] = b.cell_tiles;
    scan_cells[i] = {b.cell_tc, b.cell_to, b.cell_tiles, &j.counts[0], nullptr};
    ci[i] = {G, b.cmask, b.cell_to, b.coff, b.vcell, vlanes};
    bv[i] = {j.f, j.dens, j.iso, G, b.vcell, &j.counts[0], j.verts, j.vdens, vlanes, j.xform};
    nb_vert[i] = (uint32_t)((vlanes + kThreads - 1) / kThreads);
    ec[i] = {b.bits, G, b.edge_tc};
    nb_edges[i] = b.edge_tiles;
    scan_edges[i] = {b.edge_tc, b.edge_to, b.edge_tiles, &j.counts[1], nullptr};
    ee[i] = {b.bits, G, b.cmask, b.coff, b.edge_to, j.faces, j.fcap};
  }
  launch_multi<SignBits>(sb, nb_sign, s);
  launch_multi<BitCellsCount>(cc, nb_cells, s);
  launch_scan_exclusive_multi(scan_cells.data(), count, s);
  launch_multi<BitCellsIndex>(ci, nb_cells, s);
  launch_multi<BitVertices>(bv, nb_vert, s);
  launch_multi<BitEdgesCount>(ec, nb_edges, s);
  launch_scan_exclusive_multi(scan_edges.data(), count, s);
  launch_multi<BitEdgesEmit>(ee, nb_edges, s);
}

void launch_bit_extract(const FieldView& f, const DensityView& dens, int n, const double* iso,
                        const BitSurface& b, uint32_t* counts, double* verts, double* vdens,
                        uint64_t vcap, int64_t* faces, uint64_t fcap, const double* xform,
                        hipStream_t s) {
  const ExtractJob j{f, dens, n, iso, &b, counts, verts, vdens, vcap, faces, fcap, xform, false};
  launch_bit_extract_multi(&j, 1, s);
}

// Host copy of the mesh: the first counts[0] vertices (3 doubles each) and vertex densities
// and 2 * counts[1] faces (3 int64 each), clamped to the capacities, copied word for word to
// (pinned) host buffers of the same layout. The counts live on the device, so the copy is
// sized there: a fixed grid strides over 16-byte word pairs of the three arrays in turn.
__device__ __forceinline__ void mesh_to_host_body(Blk blk_, const uint64_t* v, const uint64_t* d,
                                                           const uint64_t* f,
                                                           const uint32_t* counts, uint64_t vcap,
                                                           uint64_t fcap, uint64_t* hv,
                                                           uint64_t* hd, uint64_t* hf) {
  using U2 = unsigned long long __attribute__((ext_vector_type(2)));
  const uint64_t nv = counts[0] < vcap ? counts[0] : vcap;
  const uint64_t nf = 2ull * counts[1] < fcap ? 2ull * counts[1] : fcap;
  const uint64_t w[3] = {3 * nv, nv, 3 * nf};
  const uint64_t p0 = (w[0] + 1) / 2, p1 = (w[1] + 1) / 2, p2 = (w[2] + 1) / 2;
  const uint64_t total = p0 + p1 + p2;
  for (uint64_t i = (uint64_t)blk_.id * kThreads + threadIdx.x; i < total;
       i += (uint64_t)blk_.count * kThreads) {
    const uint64_t* src;
    uint64_t* dst;
    uint64_t j, words;
    if (i < p0) {
      src = v, dst = hv, j = i, words = w[0];
    } else if (i < p0 + p1) {
      src = d, dst = hd, j = i - p0, words = w[1];
    } else {
      src = f, dst = hf, j = i - p0 - p1, words = w[2];
    }
    if (2 * j + 1 < words)
      *(U2*)(dst + 2 * j) = *(const U2*)(src + 2 * j);
    else
      dst[2 * j] = src[2 * j];
  }
}

namespace {
struct MeshToHost {
  struct Args {
    const uint64_t *v, *d, *f;
    const uint32_t* counts;
    uint64_t vcap, fcap;
    uint64_t *hv, *hd, *hf;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    mesh_to_host_body(b, a.v, a.d, a.f, a.counts, a.vcap, a.fcap, a.hv, a.hd, a.hf);
  }
};
}  // namespace

void launch_mesh_to_host_multi(const MeshCopyJob* jobs, size_t count, hipStream_t s) {
  std::vector<MeshToHost::Args> a(count);
  std::vector<uint32_t> nb(count);
  for (size_t i = 0; i < count; ++i) {
    const MeshCopyJob& j = jobs[i];
    const uint64_t pairs = (3 * j.vcap + 1) / 2 + (j.vcap + 1) / 2 + (3 * j.fcap + 1) / 2;
    nb[i] = (uint32_t)std::min<uint64_t>(2048, std::max<uint64_t>(1, (pairs + kThreads - 1) / kThreads));
    a[i] = {(const uint64_t*)j.verts, (const uint64_t*)j.vdens, (const uint64_t*)j.faces, j.counts,
            j.vcap, j.fcap, (uint64_t*)j.h_verts, (uint64_t*)j.h_vdens, (uint64_t*)j.h_faces};
  }
  launch_multi<MeshToHost>(a, nb, s);
}

void launch_mesh_to_host(const double* verts, const double* vdens, const int64_t* faces,
                         const uint32_t* counts, uint64_t vcap, uint64_t fcap, double* h_verts,
                         double* h_vdens, int64_t* h_faces, hipStream_t s) {
  const MeshCopyJob j{verts, vdens, faces, counts, vcap, fcap, h_verts, h_vdens, h_faces};
  launch_mesh_to_host_multi(&j, 1, s);
}

template void launch_cells_emit<int32_t>(const double*, const double*, int, const double*, const uint32_t*,
                                         int32_t*, double*, double*, uint64_t, const double*,
                                         hipStream_t);
template void launch_cells_emit<int64_t>(const double*, const double*, int, const double*, const uint32_t*,
                                         int64_t*, double*, double*, uint64_t, const double*,
                                         hipStream_t);
template void launch_edges_count<int32_t>(const double*, int, const double*, const int32_t*, uint32_t*,
                                          hipStream_t);
template void launch_edges_count<int64_t>(const double*, int, const double*, const int64_t*, uint32_t*,
                                          hipStream_t);
template void launch_edges_emit<int32_t>(const double*, int, const double*, const int32_t*,
                                         const uint32_t*, int64_t*, uint64_t, hipStream_t);
template void launch_edges_emit<int64_t>(const double*, int, const double*, const int64_t*,
                                         const uint32_t*, int64_t*, uint64_t, hipStream_t);

}  // namespace mpr
