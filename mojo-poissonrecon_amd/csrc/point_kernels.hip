// Point-stage kernels for gfx950: unit-cube normalisation, Morton codes, a stable LSB
// radix sort, the breadth-first octree arrays, the two-pass trilinear splat and the
// iso-value reduction. All are byte-bound scatters/gathers; the splat accumulates with
// the native FP64 atomic add of the MI355X L2 (global_atomic_add_f64).
#include <algorithm>

#include "block_scan.h"
#include "mpr_internal.h"
#include "multi.h"

namespace mpr {

namespace {

// capi.mojo:51-59: scaled Euclidean norm
__device__ __forceinline__ double norm3(double x, double y, double z) {
  double ax = fabs(x), ay = fabs(y), az = fabs(z);
  double m = ay > az ? ay : az;
  double scale = ax > m ? ax : m;
  if (scale == 0.0) return 0.0;
  double sx = x / scale, sy = y / scale, sz = z / scale;
  return scale * sqrt(sx * sx + sy * sy + sz * sz);
}

__device__ __forceinline__ double block_sum_d(double v, double* lds4) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds4[wave] = v;
  __syncthreads();
  double tot = ((lds4[0] + lds4[1]) + lds4[2]) + lds4[3];
  __syncthreads();
  return tot;
}

// ---- normalisation (core.py:377-385) -------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_bbox_partials(const double* pts, uint32_t P,
                                                            double* part) {
  __shared__ double lds[6][kThreads];
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < P; i += gridDim.x * kThreads)
    for (int a = 0; a < 3; ++a) {
      double v = pts[3ull * i + a];
      mn[a] = v < mn[a] ? v : mn[a];
      mx[a] = v > mx[a] ? v : mx[a];
    }
  for (int a = 0; a < 3; ++a) {
    lds[a][threadIdx.x] = mn[a];
    lds[3 + a][threadIdx.x] = mx[a];
  }
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        double o = lds[a][threadIdx.x + w];
        if (o < lds[a][threadIdx.x]) lds[a][threadIdx.x] = o;
        o = lds[3 + a][threadIdx.x + w];
        if (o > lds[3 + a][threadIdx.x]) lds[3 + a][threadIdx.x] = o;
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = lds[threadIdx.x][0];
}

// params: [origin x, y, z, width]
__device__ __forceinline__ void bbox_finish_body(Blk blk_, const double* part, int nparts,
                                                          double scale, double* params) {
  __shared__ double lds[6][kThreads];
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nparts; b += kThreads)
    for (int a = 0; a < 3; ++a) {
      double lo = part[6 * b + a], hi = part[6 * b + 3 + a];
      mn[a] = lo < mn[a] ? lo : mn[a];
      mx[a] = hi > mx[a] ? hi : mx[a];
    }
  for (int a = 0; a < 3; ++a) {
    lds[a][threadIdx.x] = mn[a];
    lds[3 + a][threadIdx.x] = mx[a];
  }
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        double o = lds[a][threadIdx.x + w];
        if (o < lds[a][threadIdx.x]) lds[a][threadIdx.x] = o;
        o = lds[3 + a][threadIdx.x + w];
        if (o > lds[3 + a][threadIdx.x]) lds[3 + a][threadIdx.x] = o;
      }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  for (int a = 0; a < 3; ++a) {
    mn[a] = lds[a][0];
    mx[a] = lds[3 + a][0];
  }
  if (!(mn[0] <= mx[0])) {  // no points (every row dropped): the identity transform
    params[0] = params[1] = params[2] = 0.0;
    params[3] = 1.0;
    return;
  }
  double ext = mx[0] - mn[0];
  for (int a = 1; a < 3; ++a) ext = (mx[a] - mn[a]) > ext ? (mx[a] - mn[a]) : ext;
  double width = ext * scale;
  if (!isfinite(width) || width <= 0.0) width = 1.0;
  for (int a = 0; a < 3; ++a) {
    double center = 0.5 * (mn[a] + mx[a]);
    params[a] = center - 0.5 * width;
  }
  params[3] = width;
}

// ---- row validation (core.py:73-83) fused with the normalisation ------------------------
// A row is kept when its six values are finite and its normal has a non-zero component.
// The reference also asks that the normal scaled by its largest |component| s have a finite,
// positive squared norm; for a finite normal with s > 0 that is a sum of three squares of
// values in [-1, 1], one of them exactly 1 (s / s), so it lies in [1, 3] and both tests keep
// the same rows on every input.
__device__ __forceinline__ bool row_valid(const double* pts, const double* nrm, uint32_t i) {
  const double p0 = pts[3ull * i], p1 = pts[3ull * i + 1], p2 = pts[3ull * i + 2];
  const double n0 = nrm[3ull * i], n1 = nrm[3ull * i + 1], n2 = nrm[3ull * i + 2];
  return isfinite(p0) && isfinite(p1) && isfinite(p2) && isfinite(n0) && isfinite(n1) &&
         isfinite(n2) && (n0 != 0.0 || n1 != 0.0 || n2 != 0.0);
}

// per block of kThreads rows: the number of kept rows and their bounding box
__device__ __forceinline__ void filter_count_body(Blk blk_, const double* pts, const double* nrm,
                                                           uint32_t P, uint32_t* counts,
                                                           double* part) {
  __shared__ double lds[6][kThreads];
  __shared__ uint32_t lds4[4];
  const uint32_t i = blk_.id * kThreads + threadIdx.x;
  const bool ok = i < P && row_valid(pts, nrm, i);
  const uint32_t c = block_count256(ok ? 1u : 0u, lds4);
  for (int a = 0; a < 3; ++a) {
    const double v = ok ? pts[3ull * i + a] : 0.0;
    lds[a][threadIdx.x] = ok ? v : INFINITY;
    lds[3 + a][threadIdx.x] = ok ? v : -INFINITY;
  }
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        double o = lds[a][threadIdx.x + w];
        if (o < lds[a][threadIdx.x]) lds[a][threadIdx.x] = o;
        o = lds[3 + a][threadIdx.x + w];
        if (o > lds[3 + a][threadIdx.x]) lds[3 + a][threadIdx.x] = o;
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blk_.id + threadIdx.x] = lds[threadIdx.x][0];
  if (threadIdx.x == 0) counts[blk_.id] = c;
}

// Kept rows in their order to the front (unit-cube points (p - origin) / width, normals as
// given), dropped rows behind them as zero points with zero normals, which the splat rejects
// and the iso value skips (rows past *total).
__device__ __forceinline__ void filter_compact_body(Blk blk_, const double* pts, const double* nrm,
                                                             uint32_t P, const uint32_t* offsets,
                                                             const uint32_t* total,
                                                             const double* params, double* unit,
                                                             double* nrm_out) {
  __shared__ uint32_t lds4[4];
  const uint32_t i = blk_.id * kThreads + threadIdx.x;
  const bool ok = i < P && row_valid(pts, nrm, i);
  uint32_t block_total;
  const uint32_t rank = block_prefix256(ok, lds4, block_total);
  if (i >= P) return;
  const uint32_t before = offsets[blk_.id] + rank;  // kept rows before row i
  const uint32_t dst = ok ? before : *total + (i - before);
  for (int a = 0; a < 3; ++a) {
    unit[3ull * dst + a] = ok ? (pts[3ull * i + a] - params[a]) / params[3] : 0.0;
    nrm_out[3ull * dst + a] = ok ? nrm[3ull * i + a] : 0.0;
  }
}

__global__ __launch_bounds__(kThreads) void k_normalize(const double* pts, uint32_t P,
                                                        const double* params, double* unit) {
  const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= 3ull * P) return;
  const int a = (int)(i % 3);
  unit[i] = (pts[i] - params[a]) / params[3];
}

// ---- Morton codes (capi.mojo:62-99) ---------------------------------------------------
__device__ __forceinline__ int64_t level_child(double v, int64_t scale) {
  int64_t q = (int64_t)(v * (double)scale);
  if (q < 0) q = 0;
  if (q >= scale) q = scale - 1;
  return q & 1;
}

__global__ __launch_bounds__(kThreads) void k_octree_codes(const double* pts, uint32_t P,
                                                           int depth, int64_t* codes) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= P) return;
  const double x = pts[3ull * i], y = pts[3ull * i + 1], z = pts[3ull * i + 2];
  int64_t code = 0;
  int64_t* row = codes + (uint64_t)i * (depth + 1);
  row[0] = 0;
  for (int d = 0; d < depth; ++d) {
    const int64_t scale = (int64_t)1 << (d + 1);
    const int64_t child = level_child(x, scale) | (level_child(y, scale) << 1) |
                          (level_child(z, scale) << 2);
    code = (code << 3) | child;
    row[d + 1] = code;
  }
}

// leaf (level-depth) code and the identity permutation, the sort's input
__global__ __launch_bounds__(kThreads) void k_leaf_keys(const double* pts, uint32_t P, int depth,
                                                        uint64_t* keys, uint32_t* perm) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= P) return;
  const double x = pts[3ull * i], y = pts[3ull * i + 1], z = pts[3ull * i + 2];
  int64_t code = 0;
  for (int d = 0; d < depth; ++d) {
    const int64_t scale = (int64_t)1 << (d + 1);
    code = (code << 3) | level_child(x, scale) | (level_child(y, scale) << 1) |
           (level_child(z, scale) << 2);
  }
  keys[i] = (uint64_t)code;
  perm[i] = i;
}

// ---- stable LSB radix sort, 8- or 9-bit digits -----------------------------------------
constexpr int kRadixItems = 8;                        // items per thread per tile
constexpr int kRadixTile = kThreads * kRadixItems;    // 2048 keys per block
constexpr int kRadixMaxBins = 512;                    // 9-bit digits

// Tile of a workgroup: the 8 XCDs take workgroups round-robin, so tile t goes to workgroup
// 8 (t mod per) + t / per, giving each XCD a contiguous run of tiles. Neighbouring tiles
// write neighbouring positions of every digit, which then meet in one XCD's L2.
// nblocks (the grid) is a multiple of 8.
__device__ __forceinline__ uint32_t radix_tile() {
  return (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
}

// per-tile digit counts, digit-major: hist[d * nblocks + tile]
template <class K, int DB>
__device__ __forceinline__ void radix_hist_body(Blk blk_, const K* keys, uint32_t n, int shift,
                                                         uint32_t* hist, uint32_t nblocks) {
  constexpr uint32_t kBins = 1u << DB;
  __shared__ uint32_t h[kBins];
  for (uint32_t b = threadIdx.x; b < kBins; b += kThreads) h[b] = 0;
  __syncthreads();
  const uint32_t tile = xcd_block(blk_);
  const uint32_t base = tile * kRadixTile;
  for (int r = 0; r < kRadixItems; ++r) {
    const uint32_t i = base + r * kThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kBins - 1u)], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < kBins; b += kThreads) hist[b * nblocks + tile] = h[b];
}

// Moves each key to offsets[digit][block] + its rank among the block's keys of that digit,
// in input order (stable). Three phases with two workgroup barriers: (1) each wavefront
// ranks its keys of every round by ballots on the DB digit bits and records its count per
// (round, wavefront, digit); (2) one lane per digit turns the counts into running offsets
// in input order (round-major, wavefront-minor); (3) every key is written. gdst (optional):
// gdst[pos] = gsrc[val] beside each moved pair, a payload gathered into the sorted order.
template <class K, int DB>
__device__ __forceinline__ void radix_scatter_body(Blk blk_, const K* kin, const uint32_t* vin,
                                                            K* kout, uint32_t* vout, uint32_t n,
                                                            int shift, const uint32_t* offsets,
                                                            uint32_t nblocks, const double* gsrc,
                                                            double* gdst) {
  constexpr uint32_t kBins = 1u << DB;
  constexpr int kWaves = kThreads / 64;
  __shared__ uint32_t cnt[kRadixItems][kWaves][kBins];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t* flat = &cnt[0][0][0];
  for (uint32_t i = threadIdx.x; i < (uint32_t)(kRadixItems * kWaves) * kBins; i += kThreads)
    flat[i] = 0;
  const uint32_t tile = xcd_block(blk_);
  const uint32_t base = tile * kRadixTile;
  K keys[kRadixItems];
  uint32_t vals[kRadixItems], digit[kRadixItems], rank[kRadixItems];
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const uint32_t i = base + r * kThreads + threadIdx.x;
    keys[r] = i < n ? kin[i] : K(0);
    vals[r] = i < n ? vin[i] : 0u;
  }
  __syncthreads();  // counts cleared
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const bool valid = base + r * kThreads + threadIdx.x < n;
    digit[r] = valid ? (uint32_t)(keys[r] >> shift) & (kBins - 1u) : 0u;
    uint64_t match = __ballot(valid);
    for (int b = 0; b < DB; ++b) {
      const bool bit = (digit[r] >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      match &= bit ? bal : ~bal;
    }
    rank[r] = (uint32_t)__popcll(match & ((1ull << lane) - 1ull));
    if (valid && rank[r] == 0) cnt[r][wave][digit[r]] = (uint32_t)__popcll(match);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < kBins; d += kThreads) {
    uint32_t run = offsets[d * nblocks + tile];
    for (int r = 0; r < kRadixItems; ++r)
      for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = cnt[r][w][d];
        cnt[r][w][d] = run;
        run += c;
      }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    if (base + r * kThreads + threadIdx.x >= n) continue;
    const uint32_t pos = cnt[r][wave][digit[r]] + rank[r];
    kout[pos] = keys[r];
    vout[pos] = vals[r];
    if (gdst) gdst[pos] = gsrc[vals[r]];
  }
}

// ---- octree levels (core.py:99-127) -----------------------------------------------------
constexpr int kTileItems = 16;
constexpr uint32_t kTile = kThreads * kTileItems;  // 4096 items per block for compaction

__device__ __forceinline__ bool level_head(const uint64_t* keys, uint32_t i, int sh) {
  return i == 0 || (keys[i] >> sh) != (keys[i - 1] >> sh);
}

__global__ __launch_bounds__(kThreads) void k_level_count(const uint64_t* keys, uint32_t n,
                                                          int sh, uint32_t* counts) {
  __shared__ uint32_t lds4[4];
  uint32_t c = 0;
  for (int r = 0; r < kTileItems; ++r) {
    const uint32_t i = blockIdx.x * kTile + r * kThreads + threadIdx.x;
    if (i < n) c += level_head(keys, i, sh);
  }
  c = block_count256(c, lds4);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// level offsets: offsets[d] = sum of level sizes before d; offsets[levels] = node count
__global__ void k_level_offsets(const uint32_t* level_count, int levels, int64_t* offsets) {
  if (threadIdx.x != 0) return;
  int64_t acc = 0;
  for (int d = 0; d < levels; ++d) {
    offsets[d] = acc;
    acc += level_count[d];
  }
  offsets[levels] = acc;
}

// Writes the level-d nodes (first occurrence of each level-d code in sorted order):
// morton, depth, parent, first-child and child count of the parent; records for every
// sorted point the level-d local node index (uidx_cur) used by level d+1.
__global__ __launch_bounds__(kThreads) void k_level_emit(
    const uint64_t* keys, uint32_t n, int d, int depth, const int64_t* offsets,
    const uint32_t* tile_offsets, const uint32_t* uidx_prev, uint32_t* uidx_cur,
    int64_t* morton, int64_t* depth_out, int64_t* parent, int64_t* child_start,
    unsigned long long* child_count) {
  __shared__ uint32_t lds4[4];
  const int sh = 3 * (depth - d);
  const int64_t off_d = offsets[d];
  const int64_t off_p = d > 0 ? offsets[d - 1] : 0;
  uint32_t running = tile_offsets[blockIdx.x];
  for (int r = 0; r < kTileItems; ++r) {
    const uint32_t i = blockIdx.x * kTile + r * kThreads + threadIdx.x;
    const bool valid = i < n;
    const bool head = valid && level_head(keys, i, sh);
    uint32_t total;
    const uint32_t pre = block_prefix256(head, lds4, total);
    if (valid) {
      // local index of this point's level-d node = number of heads up to and including i, -1
      const uint32_t u = running + pre + (head ? 1u : 0u) - 1u;
      if (uidx_cur) uidx_cur[i] = u;
      if (head) {
        const int64_t node = off_d + u;
        morton[node] = (int64_t)(keys[i] >> sh);
        depth_out[node] = d;
        if (d > 0) {
          const int64_t p = off_p + uidx_prev[i];
          parent[node] = p;
          if (level_head(keys, i, sh + 3)) child_start[p] = node;
          atomicAdd(&child_count[p], 1ull);
        } else {
          parent[node] = -1;
        }
      }
    }
    running += total;
  }
}

__global__ __launch_bounds__(kThreads) void k_point_leaf(const uint32_t* perm,
                                                         const uint32_t* uidx, uint32_t n,
                                                         const int64_t* offsets, int depth,
                                                         int64_t* point_leaf) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) point_leaf[perm[i]] = offsets[depth] + uidx[i];
}

// ---- device-wide exclusive scan of tile counts (one workgroup) -------------------------
__device__ __forceinline__ void scan_exclusive_body(Blk blk_, const uint32_t* in, uint32_t* out,
                                                         uint32_t count, uint32_t* total) {
  __shared__ uint32_t part[1024];
  const uint32_t per = (count + 1023u) / 1024u;
  const uint32_t b = threadIdx.x * per, e = min(count, b + per);
  uint32_t s = 0;
  for (uint32_t i = b; i < e; ++i) s += in[i];
  part[threadIdx.x] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the 1024 chunk sums
  for (uint32_t w = 1; w < 1024; w <<= 1) {
    uint32_t v = threadIdx.x >= w ? part[threadIdx.x - w] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
  for (uint32_t i = b; i < e; ++i) {
    const uint32_t v = in[i];
    out[i] = run;
    run += v;
  }
  if (threadIdx.x == 1023 && total) *total = part[1023];
}

// Multi-block exclusive scan for long inputs: per-chunk sums, a one-block scan of the sums,
// then each chunk scanned locally from its offset.
constexpr uint32_t kScanChunk = 4096;  // 1024 lanes x 4 values

__device__ __forceinline__ uint32_t block_exclusive1024(uint32_t v, uint32_t* lds16,
                                                        uint32_t& total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, 64);
    if (lane >= (uint32_t)off) inc += t;
  }
  if (lane == 63) lds16[wave] = inc;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (uint32_t w = 0; w < 16; ++w) {
    const uint32_t c = lds16[w];
    base += w < wave ? c : 0u;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return base + inc - v;
}

__device__ __forceinline__ void scan_chunk_sums_body(Blk blk_, const uint32_t* in, uint32_t count,
                                                          uint32_t* sums) {
  __shared__ uint32_t lds16[16];
  const uint32_t base = blk_.id * kScanChunk + threadIdx.x * 4u;
  uint32_t v = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k)
    if (base + k < count) v += in[base + k];
  uint32_t total;
  block_exclusive1024(v, lds16, total);
  if (threadIdx.x == 0) sums[blk_.id] = total;
}

__device__ __forceinline__ void scan_chunks_body(Blk blk_, const uint32_t* in, uint32_t* out,
                                                      uint32_t count, const uint32_t* offsets) {
  __shared__ uint32_t lds16[16];
  const uint32_t base = blk_.id * kScanChunk + threadIdx.x * 4u;
  uint32_t v[4], sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    v[k] = base + k < count ? in[base + k] : 0u;
    sum += v[k];
  }
  uint32_t total;
  uint32_t run = offsets[blk_.id] + block_exclusive1024(sum, lds16, total);
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    if (base + k < count) out[base + k] = run;
    run += v[k];
  }
}

// ---- splat (capi.mojo:102-224) ---------------------------------------------------------
// Deterministic, in the reference's order. Every accepted point adds to its 8 corner nodes
// (dz, dy, dx order) and the reference applies those additions point by point. Here each
// (node, point) contribution becomes a pair (key = node, id = 8 p + corner) in point order;
// a stable radix sort by node keeps each node's contributions in point order, and one lane
// per node then adds them to the grid value sequentially -- the same additions in the same
// order as the serial loop, so the grids are bit-identical to it and no atomics are used.
// Pass 2 needs the finished density, so its contributions are computed after pass 1's sums
// and summed along the same sorted pairs.

// per point: accepted? and the trilinear cell of pass 1 (capi.mojo:127-173)
__device__ __forceinline__ void splat_pairs_body(Blk blk_, const double* pts, const double* nrm,
                                                          uint32_t P, int n, int confidence,
                                                          uint32_t* keys, uint32_t* ids,
                                                          double* w1, uint32_t sentinel,
                                                          unsigned long long* accepted) {
  __shared__ uint32_t lds4[4];
  const uint32_t p = blk_.id * kThreads + threadIdx.x;
  uint32_t ok = 0;
  if (p < P) {
    const uint64_t base = 8ull * p;
    const double length = norm3(nrm[3ull * p], nrm[3ull * p + 1], nrm[3ull * p + 2]);
    if (length > 0.0) {
      ok = 1;
      const int res = n - 1;
      const double sw = confidence ? length : 1.0;
      double gx = pts[3ull * p] * (double)res, gy = pts[3ull * p + 1] * (double)res,
             gz = pts[3ull * p + 2] * (double)res;
      int64_t ix = (int64_t)gx, iy = (int64_t)gy, iz = (int64_t)gz;
      if (ix < 0) ix = 0;
      if (iy < 0) iy = 0;
      if (iz < 0) iz = 0;
      if (ix >= res) { ix = res - 1; gx = (double)res; }
      if (iy >= res) { iy = res - 1; gy = (double)res; }
      if (iz >= res) { iz = res - 1; gz = (double)res; }
      const double fx = gx - (double)ix, fy = gy - (double)iy, fz = gz - (double)iz;
      int k = 0;
      for (int dz = 0; dz < 2; ++dz) {
        const double wz = dz ? fz : 1.0 - fz;
        for (int dy = 0; dy < 2; ++dy) {
          const double wy = dy ? fy : 1.0 - fy;
          for (int dx = 0; dx < 2; ++dx, ++k) {
            const double wx = dx ? fx : 1.0 - fx;
            keys[base + k] = (uint32_t)nat_index((int)ix + dx, (int)iy + dy, (int)iz + dz, n);
            ids[base + k] = (uint32_t)(base + k);
            w1[base + k] = wx * wy * wz * sw;
          }
        }
      }
    } else {
      for (int k = 0; k < 8; ++k) {
        keys[base + k] = sentinel;  // sorts after every node and is skipped
        ids[base + k] = (uint32_t)(base + k);
      }
    }
  }
  const uint32_t c = block_count256(ok, lds4);
  if (threadIdx.x == 0 && c) atomicAdd(accepted, (unsigned long long)c);
}

// Runs of equal keys in the sorted pairs: position j starts a run when its key differs from
// position j - 1's. The lane of a starting position learns from a wavefront ballot of the run
// starts where the run's part in this wavefront ends, so it can issue those loads together;
// only a run that reaches the wavefront's end is followed on key by key. Every lane of the
// wavefront must call this (it exchanges keys).
struct RunStart {
  bool head;     // position j starts the run of a node (key < M)
  bool open;     // the run may continue past `end`, the wavefront's end
  uint32_t key;
  uint32_t end;  // first position past the run's part in this wavefront (at most count)
};

__device__ __forceinline__ RunStart run_start(const uint32_t* keys, uint32_t j, uint32_t count,
                                              uint32_t M) {
  const uint32_t lane = threadIdx.x & 63u;
  const bool valid = j < count;
  const uint32_t key = valid ? keys[j] : 0xffffffffu;
  uint32_t prev = __shfl_up(key, 1, 64);
  if (lane == 0) prev = valid && j > 0 ? keys[j - 1] : ~key;
  const bool starts = !valid || key != prev;
  const uint64_t sm = __ballot(starts);
  const uint64_t above = lane == 63u ? 0ull : sm & (~0ull << (lane + 1u));
  RunStart r;
  r.key = key;
  r.head = valid && starts && key < M;
  r.open = above == 0ull;
  r.end = (j - lane) + (r.open ? 64u : (uint32_t)__builtin_ctzll(above));
  if (r.end > count) r.end = count;
  return r;
}

// density += 2 w and screen += w over each node's contributions, in point order (ws: pass
// 1's weights in the sorted order). Every lane loads its own position's weight; the lane
// that starts a run then adds the run's weights from LDS in order (a run that continues
// past the wavefront goes on with loads). With `touched`, the grids count as zero-filled
// (the sums start at +0.0, as on the reference's fresh grids) and the node's bit is set
// instead of reading the grid.
__device__ __forceinline__ void splat_sum1_body(Blk blk_, const uint32_t* keys, const double* ws,
                                                uint32_t count, uint32_t M, double* density,
                                                double* screen, uint32_t* touched) {
  __shared__ double lw[kThreads];
  const uint32_t j = blk_.id * kThreads + threadIdx.x;
  const RunStart r = run_start(keys, j, count, M);
  lw[threadIdx.x] = j < count ? ws[j] : 0.0;
  if (touched) wave_set_bits(touched, r.head, r.key);
  __syncthreads();
  if (!r.head) return;
  const uint32_t key = r.key;
  double d = 0.0, sc = 0.0;
  if (!touched) {
    d = density[key];
    sc = screen[key];
  }
  uint32_t i = j;
  for (; i < r.end; ++i) {
    const double w = lw[threadIdx.x + (i - j)];
    d += 2.0 * w;
    sc += w;
  }
  if (r.open)
    for (; i < count && keys[i] == key; ++i) {
      const double w = ws[i];
      d += 2.0 * w;
      sc += w;
    }
  density[key] = d;
  screen[key] = sc;
}

// pass 2 (capi.mojo:176-223), per point: the density factor and what each corner's
// contribution n^ * (w (n-1) / factor) is made of
__device__ __forceinline__ void splat_pass2_body(Blk blk_, const double* pts,
                                                                const double* nrm, uint32_t P,
                                                                int n, int confidence,
                                                                const double* density,
                                                                double* c2, double spn) {
  const uint32_t p = blk_.id * kThreads + threadIdx.x;
  if (p >= P) return;
  double nx = nrm[3ull * p], ny = nrm[3ull * p + 1], nz = nrm[3ull * p + 2];
  const double length = norm3(nx, ny, nz);
  if (!(length > 0.0)) return;
  const double sw = confidence ? length : 1.0;
  nx /= length;
  ny /= length;
  nz /= length;
  const int64_t res = n - 1;
  const double gx = pts[3ull * p] * (double)res, gy = pts[3ull * p + 1] * (double)res,
               gz = pts[3ull * p + 2] * (double)res;
  int64_t ix = (int64_t)gx, iy = (int64_t)gy, iz = (int64_t)gz;
  ix = ix < 0 ? 0 : ix; ix = ix > res - 1 ? res - 1 : ix;
  iy = iy < 0 ? 0 : iy; iy = iy > res - 1 ? res - 1 : iy;
  iz = iz < 0 ? 0 : iz; iz = iz > res - 1 ? res - 1 : iz;
  double fx = gx - (double)ix, fy = gy - (double)iy, fz = gz - (double)iz;
  fx = fx < 0.0 ? 0.0 : fx; fx = fx > 1.0 ? 1.0 : fx;
  fy = fy < 0.0 ? 0.0 : fy; fy = fy > 1.0 ? 1.0 : fy;
  fz = fz < 0.0 ? 0.0 : fz; fz = fz > 1.0 ? 1.0 : fz;
  double local = 0.0;
  for (int dz = 0; dz < 2; ++dz) {
    const double wz = dz ? fz : 1.0 - fz;
    for (int dy = 0; dy < 2; ++dy) {
      const double wy = dy ? fy : 1.0 - fy;
      for (int dx = 0; dx < 2; ++dx) {
        const double wx = dx ? fx : 1.0 - fx;
        local += density[nat_index((int)ix + dx, (int)iy + dy, (int)iz + dz, n)] * wx * wy * wz;
      }
    }
  }
  const double q = local / spn;
  const double factor = q > 1.0 ? q : 1.0;
  // the point's record: pass 2's clamped fractions, weight, unit normal and density factor,
  // from which each of its 8 contributions is formed where it is summed (splat_sum2_body)
  double* rec = c2 + 8ull * p;
  rec[0] = fx;
  rec[1] = fy;
  rec[2] = fz;
  rec[3] = sw;
  rec[4] = nx;
  rec[5] = ny;
  rec[6] = nz;
  rec[7] = factor;
}

// v -= n^ vw over each node's contributions in point order (from +0.0 when the grids count
// as zero-filled, as splat_sum1_body). Every lane forms its own position's contribution
// from its point's record (pair id 8 p + k, k = 4 dz + 2 dy + dx, splat_pairs_body's corner
// order; pass 2's w = wx wy wz sw, vw = w (n-1) / factor); the lane that starts a run then
// subtracts the run's contributions from LDS in order.
__device__ __forceinline__ void splat_contribution(const double* rec, uint32_t id, double res,
                                                   double (&c)[3]) {
  using D2 = double __attribute__((ext_vector_type(2)));
  const D2* q = reinterpret_cast<const D2*>(rec + 8ull * (id >> 3));
  const D2 f01 = q[0], f2s = q[1], n01 = q[2], n2f = q[3];  // fx fy, fz sw, nx ny, nz factor
  const uint32_t k = id & 7u;
  const double wz = (k & 4u) ? f2s[0] : 1.0 - f2s[0];
  const double wy = (k & 2u) ? f01[1] : 1.0 - f01[1];
  const double wx = (k & 1u) ? f01[0] : 1.0 - f01[0];
  const double w = wx * wy * wz * f2s[1];
  const double vw = w * res / n2f[1];
  c[0] = n01[0] * vw;
  c[1] = n01[1] * vw;
  c[2] = n2f[0] * vw;
}

__device__ __forceinline__ void splat_sum2_body(Blk blk_, const uint32_t* keys,
                                                const uint32_t* ids, const double* rec,
                                                uint32_t count, uint32_t M, int n, double* vx,
                                                double* vy, double* vz, int from_zero) {
  __shared__ double lc[3][kThreads];
  const uint32_t j = blk_.id * kThreads + threadIdx.x;
  const RunStart r = run_start(keys, j, count, M);
  const double res = (double)(n - 1);
  double c[3] = {0.0, 0.0, 0.0};
  if (j < count && r.key < M) splat_contribution(rec, ids[j], res, c);
  lc[0][threadIdx.x] = c[0];
  lc[1][threadIdx.x] = c[1];
  lc[2][threadIdx.x] = c[2];
  __syncthreads();
  if (!r.head) return;
  const uint32_t key = r.key;
  double a = 0.0, b = 0.0, cz = 0.0;
  if (!from_zero) {
    a = vx[key];
    b = vy[key];
    cz = vz[key];
  }
  uint32_t i = j;
  for (; i < r.end; ++i) {
    const uint32_t t = threadIdx.x + (i - j);
    a -= lc[0][t];
    b -= lc[1][t];
    cz -= lc[2][t];
  }
  if (r.open)
    for (; i < count && keys[i] == key; ++i) {
      splat_contribution(rec, ids[i], res, c);
      a -= c[0];
      b -= c[1];
      cz -= c[2];
    }
  vx[key] = a;
  vy[key] = b;
  vz[key] = cz;
}

// ---- iso value (capi.mojo:745-795) -----------------------------------------------------
// rows at or past *count (when given) are absent
__device__ __forceinline__ void iso_partials_body(Blk blk_, FieldView field,
                                                           const double* pts, const double* nrm,
                                                           uint32_t P, int n, int confidence,
                                                           double* part, const uint32_t* count) {
  __shared__ double lds4[4];
  const uint32_t p = blk_.id * kThreads + threadIdx.x;
  double vw = 0.0, wt = 0.0;
  if (p < P && (!count || p < *count)) {
    const double weight =
        confidence ? norm3(nrm[3ull * p], nrm[3ull * p + 1], nrm[3ull * p + 2]) : 1.0;
    const int64_t res = n - 1;
    const double gx = pts[3ull * p] * (double)res, gy = pts[3ull * p + 1] * (double)res,
                 gz = pts[3ull * p + 2] * (double)res;
    int64_t ix = (int64_t)gx, iy = (int64_t)gy, iz = (int64_t)gz;
    ix = ix < 0 ? 0 : ix; ix = ix > res - 1 ? res - 1 : ix;
    iy = iy < 0 ? 0 : iy; iy = iy > res - 1 ? res - 1 : iy;
    iz = iz < 0 ? 0 : iz; iz = iz > res - 1 ? res - 1 : iz;
    double fx = gx - (double)ix, fy = gy - (double)iy, fz = gz - (double)iz;
    fx = fx < 0.0 ? 0.0 : fx; fx = fx > 1.0 ? 1.0 : fx;
    fy = fy < 0.0 ? 0.0 : fy; fy = fy > 1.0 ? 1.0 : fy;
    fz = fz < 0.0 ? 0.0 : fz; fz = fz > 1.0 ? 1.0 : fz;
    double value = 0.0;
    for (int dz = 0; dz < 2; ++dz) {
      const double wz = dz ? fz : 1.0 - fz;
      for (int dy = 0; dy < 2; ++dy) {
        const double wy = dy ? fy : 1.0 - fy;
        for (int dx = 0; dx < 2; ++dx) {
          const double wx = dx ? fx : 1.0 - fx;
          value += field.at((int)ix + dx, (int)iy + dy, (int)iz + dz) * wx * wy * wz;
        }
      }
    }
    vw = value * weight;
    wt = weight;
  }
  const double a = block_sum_d(vw, lds4);
  const double b = block_sum_d(wt, lds4);
  if (threadIdx.x == 0) {
    part[2 * blk_.id] = a;
    part[2 * blk_.id + 1] = b;
  }
}

// fixed summation order: thread t sums partials t, t+256, ...; then a fixed tree
__device__ __forceinline__ void iso_finish_body(Blk blk_, const double* part, int nparts,
                                                         double* iso) {
  __shared__ double la[kThreads], lb[kThreads];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kThreads) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  la[threadIdx.x] = a;
  lb[threadIdx.x] = b;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      la[threadIdx.x] += la[threadIdx.x + w];
      lb[threadIdx.x] += lb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) iso[0] = lb[0] > 0.0 ? la[0] / lb[0] : 0.0;
}

// ---- multi-instance kernel bodies (multi.h) ---------------------------------------------
struct BboxFinish {
  struct Args {
    const double* part;
    int nparts;
    double scale;
    double* params;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    bbox_finish_body(b, a.part, a.nparts, a.scale, a.params);
  }
};

struct FilterCount {
  struct Args {
    const double *pts, *nrm;
    uint32_t P;
    uint32_t* counts;
    double* part;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    filter_count_body(b, a.pts, a.nrm, a.P, a.counts, a.part);
  }
};

struct FilterCompact {
  struct Args {
    const double *pts, *nrm;
    uint32_t P;
    const uint32_t *offsets, *total;
    const double* params;
    double *unit, *nrm_out;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    filter_compact_body(b, a.pts, a.nrm, a.P, a.offsets, a.total, a.params, a.unit, a.nrm_out);
  }
};

struct ScanExclusive {
  struct Args {
    const uint32_t* in;
    uint32_t* out;
    uint32_t count;
    uint32_t* total;
  };
  static constexpr int kThreads = 1024;
  static __device__ void run(const Args& a, Blk b) {
    scan_exclusive_body(b, a.in, a.out, a.count, a.total);
  }
};

struct ScanChunkSums {
  struct Args {
    const uint32_t* in;
    uint32_t count;
    uint32_t* sums;
  };
  static constexpr int kThreads = 1024;
  static __device__ void run(const Args& a, Blk b) { scan_chunk_sums_body(b, a.in, a.count, a.sums); }
};

struct ScanChunks {
  struct Args {
    const uint32_t* in;
    uint32_t* out;
    uint32_t count;
    const uint32_t* offsets;
  };
  static constexpr int kThreads = 1024;
  static __device__ void run(const Args& a, Blk b) {
    scan_chunks_body(b, a.in, a.out, a.count, a.offsets);
  }
};

template <class K>
struct RadixHistArgs {
  const K* keys;
  uint32_t n;
  int shift;
  uint32_t* hist;
};

template <class K>
struct RadixScatterArgs {
  const K* kin;
  const uint32_t* vin;
  K* kout;
  uint32_t* vout;
  uint32_t n;
  int shift;
  const uint32_t* offsets;
  const double* gsrc;
  double* gdst;
};

template <class K, int DB>
struct RadixHist {
  using Args = RadixHistArgs<K>;
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    radix_hist_body<K, DB>(b, a.keys, a.n, a.shift, a.hist, b.count);
  }
};

template <class K, int DB>
struct RadixScatter {
  using Args = RadixScatterArgs<K>;
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    radix_scatter_body<K, DB>(b, a.kin, a.vin, a.kout, a.vout, a.n, a.shift, a.offsets, b.count,
                              a.gsrc, a.gdst);
  }
};

struct IsoPartials {
  struct Args {
    FieldView field;
    const double *pts, *nrm;
    uint32_t P;
    int n, confidence;
    double* part;
    const uint32_t* count;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    iso_partials_body(b, a.field, a.pts, a.nrm, a.P, a.n, a.confidence, a.part, a.count);
  }
};

struct IsoFinish {
  struct Args {
    const double* part;
    int nparts;
    double* iso;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) { iso_finish_body(b, a.part, a.nparts, a.iso); }
};

struct SplatPairs {
  struct Args {
    const double *pts, *nrm;
    uint32_t P;
    int n, confidence;
    uint32_t *keys, *ids;
    double* w1;
    uint32_t sentinel;
    unsigned long long* accepted;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    splat_pairs_body(b, a.pts, a.nrm, a.P, a.n, a.confidence, a.keys, a.ids, a.w1, a.sentinel,
                     a.accepted);
  }
};

struct SplatSum1 {
  struct Args {
    const uint32_t* keys;
    const double* ws;
    uint32_t count, M;
    double *density, *screen;
    uint32_t* touched;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    splat_sum1_body(b, a.keys, a.ws, a.count, a.M, a.density, a.screen, a.touched);
  }
};

struct SplatPass2 {
  struct Args {
    const double *pts, *nrm;
    uint32_t P;
    int n, confidence;
    const double* density;
    double* c2;
    double spn;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    splat_pass2_body(b, a.pts, a.nrm, a.P, a.n, a.confidence, a.density, a.c2, a.spn);
  }
};

struct SplatSum2 {
  struct Args {
    const uint32_t *keys, *ids;
    const double* rec;
    uint32_t count, M;
    int n;
    double *vx, *vy, *vz;
    int from_zero;
  };
  static constexpr int kThreads = mpr::kThreads;
  static __device__ void run(const Args& a, Blk b) {
    splat_sum2_body(b, a.keys, a.ids, a.rec, a.count, a.M, a.n, a.vx, a.vy, a.vz, a.from_zero);
  }
};

}  // namespace

// ------------------------------------------------------------------------------------------
int bbox_partial_count(uint32_t P) {
  return (int)std::min<uint32_t>(std::max<uint32_t>(1u, blocks_for(P, kThreads)), 1024u);
}

void launch_normalize(const double* pts, uint32_t P, double scale, double* part, double* params,
                      double* unit, hipStream_t s) {
  const int nb = bbox_partial_count(P);
  hipLaunchKernelGGL(k_bbox_partials, dim3(nb), dim3(kThreads), 0, s, pts, P, part);
  MPR_LAUNCH_CHECK();
  launch_one<BboxFinish>(BboxFinish::Args{part, nb, scale, params}, 1, s);
  if (P) {
    hipLaunchKernelGGL(k_normalize, dim3(blocks_for(3ull * P, kThreads)), dim3(kThreads), 0, s,
                       pts, P, params, unit);
    MPR_LAUNCH_CHECK();
  }
}

uint32_t filter_blocks(uint32_t P) { return std::max<uint32_t>(1u, blocks_for(P, kThreads)); }

void launch_filter_normalize_multi(const FilterJob* jobs, size_t count, hipStream_t s) {
  std::vector<FilterCount::Args> fc(count);
  std::vector<BboxFinish::Args> bf(count);
  std::vector<FilterCompact::Args> cp(count);
  std::vector<ScanJob> sc(count);
  std::vector<uint32_t> nb(count), one(count, 1u), ncp(count);
  for (size_t i = 0; i < count; ++i) {
    const FilterJob& j = jobs[i];
    nb[i] = filter_blocks(j.P);
    fc[i] = {j.pts, j.nrm, j.P, j.f.counts, j.f.part};
    bf[i] = {j.f.part, (int)nb[i], j.scale, j.params};
    sc[i] = {j.f.counts, j.f.offsets, nb[i], j.total, j.f.scan};
    cp[i] = {j.pts, j.nrm, j.P, j.f.offsets, j.total, j.params, j.unit, j.nrm_out};
    ncp[i] = j.P ? nb[i] : 0u;
  }
  launch_multi<FilterCount>(fc, nb, s);
  launch_multi<BboxFinish>(bf, one, s);
  launch_scan_exclusive_multi(sc.data(), count, s);
  launch_multi<FilterCompact>(cp, ncp, s);
}

void launch_filter_normalize(const double* pts, const double* nrm, uint32_t P, double scale,
                             const FilterScratch& f, uint32_t* total, double* params, double* unit,
                             double* nrm_out, hipStream_t s) {
  const FilterJob j{pts, nrm, P, scale, f, total, params, unit, nrm_out};
  launch_filter_normalize_multi(&j, 1, s);
}

void launch_octree_codes(const double* pts, uint32_t P, int depth, int64_t* codes,
                         hipStream_t s) {
  if (!P) return;
  hipLaunchKernelGGL(k_octree_codes, dim3(blocks_for(P, kThreads)), dim3(kThreads), 0, s, pts, P,
                     depth, codes);
  MPR_LAUNCH_CHECK();
}

void launch_scan_exclusive_multi(const ScanJob* jobs, size_t count, hipStream_t s) {
  // jobs of more than one chunk (with scratch) scan their chunk sums first
  std::vector<ScanChunkSums::Args> cs(count);
  std::vector<ScanExclusive::Args> ex(count);
  std::vector<ScanChunks::Args> ch(count);
  std::vector<uint32_t> nchunk(count, 0u), one(count, 1u);
  bool chunked = false;
  for (size_t i = 0; i < count; ++i) {
    const ScanJob& j = jobs[i];
    if (j.count > kScanChunk && j.scratch) {
      const uint32_t chunks = (j.count + kScanChunk - 1) / kScanChunk;
      nchunk[i] = chunks;
      chunked = true;
      cs[i] = {j.in, j.count, j.scratch};
      ex[i] = {j.scratch, j.scratch, chunks, j.total};
      ch[i] = {j.in, j.out, j.count, j.scratch};
    } else {
      ex[i] = {j.in, j.out, j.count, j.total};
    }
  }
  if (chunked) launch_multi<ScanChunkSums>(cs, nchunk, s);
  launch_multi<ScanExclusive>(ex, one, s);
  if (chunked) launch_multi<ScanChunks>(ch, nchunk, s);
}

void launch_scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t count, uint32_t* total,
                           hipStream_t s, uint32_t* scratch) {
  const ScanJob j{in, out, count, total, scratch};
  launch_scan_exclusive_multi(&j, 1, s);
}

uint32_t scan_scratch_words(uint32_t count) { return (count + kScanChunk - 1) / kScanChunk + 1; }

// tiles of a sort of n keys, a multiple of 8 (radix_tile)
inline uint32_t radix_tiles(uint32_t n) {
  return ((std::max<uint32_t>(1u, (n + kRadixTile - 1) / kRadixTile) + 7u) / 8u) * 8u;
}

size_t radix_scratch_words(uint32_t n) {
  const uint32_t nb = radix_tiles(n);
  // histogram, then the scan's sums
  return (size_t)kRadixMaxBins * nb + scan_scratch_words((uint32_t)kRadixMaxBins * nb);
}

// Sorts each job's (keys, vals) by the low `bits` key bits, stably, with 9-bit digits where
// that saves a pass (25..27 bits: depth-8 node keys, depth-9 Morton codes) and 8-bit digits
// otherwise; the jobs of one call share the digit width and pass count of the widest key
// (a pass over all-zero digits moves nothing: a stable sort's order is the same for any
// digit schedule). (k_tmp, v_tmp) are ping-pong buffers; hist and offs hold
// radix_scratch_words(n) words. The result ends in (keys, vals), or, with result_keys /
// result_vals, in whichever pair of buffers the last pass wrote (no copy back), which they
// then point at. gather_dst (optional, n words): gather_dst[i] = gather_src[vals[i]] in the
// sorted order, written by the last pass.
template <class K>
void launch_radix_sort_multi(const SortJob<K>* jobs, size_t count, hipStream_t s) {
  int bits = 0;
  std::vector<char> active(count, 0);
  for (size_t i = 0; i < count; ++i) {
    const SortJob<K>& j = jobs[i];
    if (j.result_keys) *j.result_keys = j.keys;
    if (j.result_vals) *j.result_vals = j.vals;
    if (j.n == 0 || (j.n == 1 && !j.gather_dst)) continue;
    if (j.bits > 8 * (int)sizeof(K)) throw HipError{"radix sort: more key bits than the key holds"};
    active[i] = 1;
    bits = std::max(bits, j.bits);
  }
  if (!bits) return;
  const int wide = (bits + 8) / 9 < (bits + 7) / 8;
  const int db = wide ? 9 : 8;
  const int passes = (bits + db - 1) / db;
  const uint32_t bins = 1u << db;
  std::vector<K*> ka(count), kb(count);
  std::vector<uint32_t*> va(count), vb(count);
  std::vector<uint32_t> nb(count, 0u);
  for (size_t i = 0; i < count; ++i) {
    ka[i] = jobs[i].keys;
    kb[i] = jobs[i].k_tmp;
    va[i] = jobs[i].vals;
    vb[i] = jobs[i].v_tmp;
    if (active[i]) nb[i] = radix_tiles(jobs[i].n);
  }
  std::vector<RadixHistArgs<K>> h(count);
  std::vector<RadixScatterArgs<K>> sc(count);
  std::vector<ScanJob> scan(count);
  for (int p = 0; p < passes; ++p) {
    const int shift = db * p;
    const bool last = p == passes - 1;
    size_t n_active = 0;
    for (size_t i = 0; i < count; ++i) {
      const SortJob<K>& j = jobs[i];
      h[i] = {ka[i], j.n, shift, j.hist};
      sc[i] = {ka[i], va[i], kb[i], vb[i], j.n, shift, j.offs, last ? j.gather_src : nullptr,
               last ? j.gather_dst : nullptr};
      // inactive jobs scan nothing
      scan[n_active] = {j.hist, j.offs, bins * nb[i], nullptr, j.hist + (size_t)bins * nb[i]};
      n_active += active[i] ? 1 : 0;
    }
    if (wide)
      launch_multi<RadixHist<K, 9>>(h, nb, s);
    else
      launch_multi<RadixHist<K, 8>>(h, nb, s);
    launch_scan_exclusive_multi(scan.data(), n_active, s);
    if (wide)
      launch_multi<RadixScatter<K, 9>>(sc, nb, s);
    else
      launch_multi<RadixScatter<K, 8>>(sc, nb, s);
    for (size_t i = 0; i < count; ++i)
      if (active[i]) {
        std::swap(ka[i], kb[i]);
        std::swap(va[i], vb[i]);
      }
  }
  for (size_t i = 0; i < count; ++i) {
    const SortJob<K>& j = jobs[i];
    if (!active[i]) continue;
    if (j.result_keys && j.result_vals) {
      *j.result_keys = ka[i];
      *j.result_vals = va[i];
    } else if (ka[i] != j.keys) {
      MPR_HIP(hipMemcpyAsync(j.keys, ka[i], sizeof(K) * j.n, hipMemcpyDeviceToDevice, s));
      MPR_HIP(hipMemcpyAsync(j.vals, va[i], sizeof(uint32_t) * j.n, hipMemcpyDeviceToDevice, s));
    }
  }
}

template <class K>
void launch_radix_sort(K* keys, uint32_t* vals, K* k_tmp, uint32_t* v_tmp, uint32_t n, int bits,
                       uint32_t* hist, uint32_t* offs, hipStream_t s, const double* gather_src,
                       double* gather_dst, K** result_keys, uint32_t** result_vals) {
  const SortJob<K> j{keys, vals, k_tmp, v_tmp, n, bits, hist, offs, gather_src, gather_dst,
                     result_keys, result_vals};
  launch_radix_sort_multi(&j, 1, s);
}
template void launch_radix_sort<uint32_t>(uint32_t*, uint32_t*, uint32_t*, uint32_t*, uint32_t,
                                          int, uint32_t*, uint32_t*, hipStream_t, const double*,
                                          double*, uint32_t**, uint32_t**);
template void launch_radix_sort<uint64_t>(uint64_t*, uint32_t*, uint64_t*, uint32_t*, uint32_t,
                                          int, uint32_t*, uint32_t*, hipStream_t, const double*,
                                          double*, uint64_t**, uint32_t**);

void launch_leaf_keys(const double* pts, uint32_t P, int depth, uint64_t* keys, uint32_t* perm,
                      hipStream_t s) {
  if (!P) return;
  hipLaunchKernelGGL(k_leaf_keys, dim3(blocks_for(P, kThreads)), dim3(kThreads), 0, s, pts, P,
                     depth, keys, perm);
  MPR_LAUNCH_CHECK();
}

uint32_t compaction_tiles(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + kTile - 1) / kTile); }

void launch_level_counts(const uint64_t* keys, uint32_t n, int d, int depth, uint32_t* counts,
                         uint32_t* offsets_out, uint32_t* level_total, hipStream_t s) {
  const uint32_t nb = compaction_tiles(n);
  hipLaunchKernelGGL(k_level_count, dim3(nb), dim3(kThreads), 0, s, keys, n, 3 * (depth - d),
                     counts);
  MPR_LAUNCH_CHECK();
  launch_scan_exclusive(counts, offsets_out, nb, level_total, s);
}

void launch_level_offsets(const uint32_t* level_count, int levels, int64_t* offsets,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_level_offsets, dim3(1), dim3(64), 0, s, level_count, levels, offsets);
  MPR_LAUNCH_CHECK();
}

void launch_level_emit(const uint64_t* keys, uint32_t n, int d, int depth, const int64_t* offsets,
                       const uint32_t* tile_offsets, const uint32_t* uidx_prev, uint32_t* uidx_cur,
                       int64_t* morton, int64_t* depth_out, int64_t* parent, int64_t* child_start,
                       int64_t* child_count, hipStream_t s) {
  const uint32_t nb = compaction_tiles(n);
  hipLaunchKernelGGL(k_level_emit, dim3(nb), dim3(kThreads), 0, s, keys, n, d, depth, offsets,
                     tile_offsets, uidx_prev, uidx_cur, morton, depth_out, parent, child_start,
                     (unsigned long long*)child_count);
  MPR_LAUNCH_CHECK();
}

void launch_point_leaf(const uint32_t* perm, const uint32_t* uidx, uint32_t n,
                       const int64_t* offsets, int depth, int64_t* point_leaf, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_point_leaf, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, perm, uidx,
                     n, offsets, depth, point_leaf);
  MPR_LAUNCH_CHECK();
}

SplatScratch splat_scratch_sizes(uint32_t P) {
  SplatScratch b;
  b.pairs = 8u * std::max<uint32_t>(1u, P);
  b.radix_words = radix_scratch_words(b.pairs);
  return b;
}

void launch_splat_multi(const SplatJob* jobs, size_t count, hipStream_t s) {
  std::vector<SplatPairs::Args> pa(count);
  std::vector<SortJob<uint32_t>> so(count);
  std::vector<SplatSum1::Args> s1(count);
  std::vector<SplatPass2::Args> p2(count);
  std::vector<SplatSum2::Args> s2(count);
  std::vector<uint32_t> nb(count, 0u), np(count, 0u);
  std::vector<uint32_t*> keys(count, nullptr), ids(count, nullptr);
  for (size_t i = 0; i < count; ++i) {
    const SplatJob& j = jobs[i];
    const uint32_t pairs = 8u * j.P;
    if (pairs > j.b.pairs) throw HipError{"splat scratch smaller than the point count"};
    // node keys and the sentinel M (sorts after every node) fit 32 bits: M <= 1025^3 < 2^31
    const uint32_t M = (uint32_t)((uint64_t)j.n * j.n * j.n);
    int bits = 1;
    while ((1ull << bits) <= M) ++bits;
    if (j.P) {
      nb[i] = blocks_for(j.P, kThreads);
      np[i] = blocks_for(pairs, kThreads);
    }
    pa[i] = {j.pts, j.nrm, j.P, j.n, j.confidence, j.b.keys, j.b.ids, j.b.w1, M, j.accepted};
    // the sort's last pass also lays pass 1's weights out in the sorted order, in the space
    // of pass 2's contributions (written after k_splat_sum1 has read them); the sorted
    // pairs stay in whichever buffers the last pass wrote
    so[i] = {j.b.keys, j.b.ids, j.b.ktmp, j.b.itmp, j.P ? pairs : 0u, bits, j.b.hist, j.b.offs,
             j.b.w1, j.b.c2, &keys[i], &ids[i]};
  }
  launch_multi<SplatPairs>(pa, nb, s);
  launch_radix_sort_multi(so.data(), count, s);
  for (size_t i = 0; i < count; ++i) {
    const SplatJob& j = jobs[i];
    const uint32_t pairs = 8u * j.P, M = (uint32_t)((uint64_t)j.n * j.n * j.n);
    s1[i] = {keys[i], j.b.c2, pairs, M, j.density, j.screen, j.touched};
    // per-point records of pass 2 behind the sorted pass-1 weights (8 P <= 2 pairs words of
    // the 3 pairs c2 holds)
    double* rec = j.b.c2 + pairs;
    p2[i] = {j.pts, j.nrm, j.P, j.n, j.confidence, j.density, rec, j.spn};
    s2[i] = {keys[i], ids[i], rec, pairs, M, j.n, j.vx, j.vy, j.vz, j.touched ? 1 : 0};
  }
  launch_multi<SplatSum1>(s1, np, s);
  launch_multi<SplatPass2>(p2, nb, s);
  launch_multi<SplatSum2>(s2, np, s);
}

void launch_splat(const double* pts, const double* nrm, uint32_t P, int n, int confidence,
                  double* density, double* vx, double* vy, double* vz, double* screen,
                  double spn, unsigned long long* accepted, const SplatScratch& b,
                  uint32_t* touched, hipStream_t s) {
  const SplatJob j{pts, nrm, P, n, confidence, density, vx, vy, vz, screen, spn, accepted, b,
                   touched};
  launch_splat_multi(&j, 1, s);
}

int iso_partial_count(uint32_t P) { return (int)std::max<uint32_t>(1u, blocks_for(P, kThreads)); }

void launch_iso_multi(const IsoJob* jobs, size_t count, hipStream_t s) {
  std::vector<IsoPartials::Args> pa(count);
  std::vector<IsoFinish::Args> fi(count);
  std::vector<uint32_t> nb(count), one(count, 1u);
  for (size_t i = 0; i < count; ++i) {
    const IsoJob& j = jobs[i];
    nb[i] = (uint32_t)iso_partial_count(j.P);
    pa[i] = {j.field, j.pts, j.nrm, j.P, j.n, j.confidence, j.part, j.count};
    fi[i] = {j.part, (int)nb[i], j.iso};
  }
  launch_multi<IsoPartials>(pa, nb, s);
  launch_multi<IsoFinish>(fi, one, s);
}

void launch_iso(const FieldView& field, const double* pts, const double* nrm, uint32_t P, int n,
                int confidence, double* part, double* iso, hipStream_t s, const uint32_t* count) {
  const IsoJob j{field, pts, nrm, P, n, confidence, part, iso, count};
  launch_iso_multi(&j, 1, s);
}

}  // namespace mpr
