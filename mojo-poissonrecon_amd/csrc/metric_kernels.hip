// Mesh-to-mesh distances for the accuracy checks of the reconstruction (the reference's
// tolerance metric is a symmetric distance between meshes normalised by the bounding-box
// diagonal, tests/test_reconstruction.py:50-62 of the reference). For every vertex of one
// mesh the distance to the nearest point of the other mesh's triangles is computed exactly
// in FP64; the Hausdorff distance is the larger of the two directions' maxima and the
// vertex chamfer the mean of the two directions' means.
//
// Search structure: the target mesh's triangles are binned into a uniform grid of cubic
// cells of side h (each triangle into every cell its bounding box overlaps). A query point
// tests the triangles of the 27 cells around its own cell: the ball of radius h around the
// point lies inside those cells, so a nearest distance <= h found there is exact (a closer
// triangle would meet the ball and therefore be binned in one of them; the test keeps a
// margin of 0.1 % of h for the rounding of cell coordinates). Queries whose best distance
// exceeds that, or that found no triangle, are answered again on grids 8 times coarser,
// down to a grid of at most 2 cells per axis, where the 27 cells are every cell. Minima do
// not depend on the order of the tests, and the reductions run in a fixed order, so the
// results are deterministic.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mpr_hip.h"
#include "mpr_internal.h"

namespace mpr {
int ensure_device();
}

using namespace mpr;

namespace {

struct CellGrid {
  double lo[3];
  double inv_h;
  int32_t nc[3];
  double h2;  // (0.999 h)^2: a squared distance up to this is exact from the 27 cells
};

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 axpy(V3 a, double t, V3 d) {
  return {a.x + t * d.x, a.y + t * d.y, a.z + t * d.z};
}
__device__ __forceinline__ double dist2(V3 a, V3 b) {
  const V3 d = sub(a, b);
  return dot(d, d);
}
__device__ __forceinline__ V3 load3(const double* v, int64_t i) {
  return {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
}

// squared distance from p to the segment [a, b]
__device__ double seg_dist2(V3 p, V3 a, V3 b) {
  const V3 ab = sub(b, a);
  const double len2 = dot(ab, ab);
  double t = len2 > 0.0 ? dot(sub(p, a), ab) / len2 : 0.0;
  t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  return dist2(p, axpy(a, t, ab));
}

// Squared distance from p to the triangle (a, b, c): the closest point by the Voronoi
// regions of the vertices, edges and face (Ericson, Real-Time Collision Detection, 2005,
// section 5.1.5). A degenerate triangle (no interior) falls back to its three edges.
__device__ double tri_dist2(V3 p, V3 a, V3 b, V3 c) {
  const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return dot(ap, ap);
  const V3 bp = sub(p, b);
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return dot(bp, bp);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0 && d1 - d3 > 0.0)
    return dist2(p, axpy(a, d1 / (d1 - d3), ab));
  const V3 cp = sub(p, c);
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return dot(cp, cp);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0 && d2 - d6 > 0.0)
    return dist2(p, axpy(a, d2 / (d2 - d6), ac));
  const double va = d3 * d6 - d5 * d4;
  const double e43 = d4 - d3, e56 = d5 - d6;
  if (va <= 0.0 && e43 >= 0.0 && e56 >= 0.0 && e43 + e56 > 0.0)
    return dist2(p, axpy(b, e43 / (e43 + e56), sub(c, b)));
  const double sum = va + vb + vc;
  if (!(sum > 0.0) || !std::isfinite(sum)) {
    const double e0 = seg_dist2(p, a, b), e1 = seg_dist2(p, b, c), e2 = seg_dist2(p, c, a);
    return fmin(e0, fmin(e1, e2));
  }
  const double v = vb / sum, w = vc / sum;
  const V3 q = {a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w};
  return dist2(p, q);
}

__device__ __forceinline__ int cell_coord(double x, const CellGrid& g, int a) {
  int i = (int)floor((x - g.lo[a]) * g.inv_h);
  return i < 0 ? 0 : (i >= g.nc[a] ? g.nc[a] - 1 : i);
}

__device__ __forceinline__ void tri_cells(const double* verts, const int32_t* faces, uint32_t t,
                                          const CellGrid& g, int* lo, int* hi) {
  const V3 a = load3(verts, faces[3 * t]), b = load3(verts, faces[3 * t + 1]),
           c = load3(verts, faces[3 * t + 2]);
  const double mn[3] = {fmin(a.x, fmin(b.x, c.x)), fmin(a.y, fmin(b.y, c.y)),
                        fmin(a.z, fmin(b.z, c.z))};
  const double mx[3] = {fmax(a.x, fmax(b.x, c.x)), fmax(a.y, fmax(b.y, c.y)),
                        fmax(a.z, fmax(b.z, c.z))};
  for (int k = 0; k < 3; ++k) {
    lo[k] = cell_coord(mn[k], g, k);
    hi[k] = cell_coord(mx[k], g, k);
  }
}

__global__ __launch_bounds__(kThreads) void k_bin_count(const double* verts, const int32_t* faces,
                                                        uint32_t F, CellGrid g, uint32_t* counts) {
  const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= F) return;
  int lo[3], hi[3];
  tri_cells(verts, faces, t, g, lo, hi);
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x)
        atomicAdd(&counts[((uint64_t)z * g.nc[1] + y) * g.nc[0] + x], 1u);
}

__global__ __launch_bounds__(kThreads) void k_bin_fill(const double* verts, const int32_t* faces,
                                                       uint32_t F, CellGrid g,
                                                       const uint32_t* starts, uint32_t* cursor,
                                                       uint32_t* items) {
  const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= F) return;
  int lo[3], hi[3];
  tri_cells(verts, faces, t, g, lo, hi);
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const uint64_t c = ((uint64_t)z * g.nc[1] + y) * g.nc[0] + x;
        items[starts[c] + atomicAdd(&cursor[c], 1u)] = t;
      }
}

// Nearest squared distance from query points (all of them, or those listed in `ids`) to the
// binned triangles of the 27 cells around each; a query whose best distance exceeds h (it
// might have a closer triangle further out) is appended to `far_out` for a coarser grid.
__global__ __launch_bounds__(kThreads) void k_near(const double* q, const uint32_t* ids, uint32_t Q,
                                                   const double* verts, const int32_t* faces,
                                                   CellGrid g, const uint32_t* starts,
                                                   const uint32_t* counts, const uint32_t* items,
                                                   double* d2_out, uint32_t* far_out,
                                                   uint32_t* far_count) {
  const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= Q) return;
  const uint32_t i = ids ? ids[j] : j;
  const V3 p = load3(q, i);
  const int cx = cell_coord(p.x, g, 0), cy = cell_coord(p.y, g, 1), cz = cell_coord(p.z, g, 2);
  double best = INFINITY;
  for (int z = max(cz - 1, 0); z <= min(cz + 1, g.nc[2] - 1); ++z)
    for (int y = max(cy - 1, 0); y <= min(cy + 1, g.nc[1] - 1); ++y)
      for (int x = max(cx - 1, 0); x <= min(cx + 1, g.nc[0] - 1); ++x) {
        const uint64_t c = ((uint64_t)z * g.nc[1] + y) * g.nc[0] + x;
        const uint32_t s = starts[c], e = s + counts[c];
        for (uint32_t k = s; k < e; ++k) {
          const uint32_t t = items[k];
          const double d = tri_dist2(p, load3(verts, faces[3 * t]), load3(verts, faces[3 * t + 1]),
                                     load3(verts, faces[3 * t + 2]));
          best = d < best ? d : best;
        }
      }
  d2_out[i] = best;
  if (!(best <= g.h2)) far_out[atomicAdd(far_count, 1u)] = i;
}

// per-block maximum and sum of the distances (sqrt of the squared minima)
__global__ __launch_bounds__(kThreads) void k_dist_partials(const double* d2, uint32_t Q,
                                                            double* part) {
  __shared__ double lmax[kThreads], lsum[kThreads];
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  const double d = i < Q ? sqrt(d2[i]) : 0.0;
  lmax[threadIdx.x] = d;
  lsum[threadIdx.x] = d;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      lmax[threadIdx.x] = fmax(lmax[threadIdx.x], lmax[threadIdx.x + w]);
      lsum[threadIdx.x] += lsum[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = lmax[0];
    part[2 * blockIdx.x + 1] = lsum[0];
  }
}

__global__ __launch_bounds__(kThreads) void k_dist_finish(const double* part, uint32_t nparts,
                                                          uint32_t Q, double* out) {
  __shared__ double lmax[kThreads], lsum[kThreads];
  double m = 0.0, s = 0.0;
  for (uint32_t k = threadIdx.x; k < nparts; k += kThreads) {
    m = fmax(m, part[2 * k]);
    s += part[2 * k + 1];
  }
  lmax[threadIdx.x] = m;
  lsum[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      lmax[threadIdx.x] = fmax(lmax[threadIdx.x], lmax[threadIdx.x + w]);
      lsum[threadIdx.x] += lsum[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = lmax[0];
    out[1] = Q ? lsum[0] / (double)Q : 0.0;
  }
}

template <class T>
struct Dev {
  T* p = nullptr;
  explicit Dev(size_t count) { MPR_HIP(hipMalloc((void**)&p, std::max<size_t>(1, count) * sizeof(T))); }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};

constexpr uint64_t kMaxCells = uint64_t(1) << 24;

// Sampled mean edge length of a mesh (0 for none).
double mean_edge(const double* verts, const int64_t* faces, int64_t F) {
  double sum = 0.0;
  int64_t used = 0;
  const int64_t step = std::max<int64_t>(1, F / 65536);
  for (int64_t t = 0; t < F; t += step, ++used) {
    const double* a = verts + 3 * faces[3 * t];
    const double* b = verts + 3 * faces[3 * t + 1];
    sum += std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                     (a[2] - b[2]) * (a[2] - b[2]));
  }
  return used ? sum / (double)used : 0.0;
}

// Grid over lo..hi with cells of side at least h, coarsened until it has at most kMaxCells
// cells. When every axis has at most 2 cells, the 27 cells around any cell are the whole
// grid, so every distance found is exact (h2 = infinity).
CellGrid make_grid(const double* lo, const double* hi, double& h) {
  CellGrid g;
  for (;;) {
    uint64_t total = 1;
    for (int k = 0; k < 3; ++k) {
      g.nc[k] = (int32_t)std::min<double>(1 << 20, std::floor((hi[k] - lo[k]) / h) + 1.0);
      total *= (uint64_t)g.nc[k];
    }
    if (total <= kMaxCells) break;
    h *= 1.25;
  }
  for (int k = 0; k < 3; ++k) g.lo[k] = lo[k];
  g.inv_h = 1.0 / h;
  g.h2 = (g.nc[0] <= 2 && g.nc[1] <= 2 && g.nc[2] <= 2) ? INFINITY : (0.999 * h) * (0.999 * h);
  return g;
}

// out[0] = max, out[1] = mean over the query points of the distance to the target mesh.
// Queries are first answered on a grid of cells twice the target's mean edge; those it
// cannot settle are answered again on grids 8 times coarser, until the grid is at most 2
// cells per axis (at most five rounds from the finest cell of extent / 4096).
void directed(const double* qv, int64_t Q, const double* tv_host, const double* tv,
              const int64_t* tf_host, const int32_t* tf, int64_t F, const double* lo,
              const double* hi, double* out) {
  if (Q == 0) {
    out[0] = out[1] = 0.0;
    return;
  }
  if (F == 0) {
    out[0] = out[1] = INFINITY;
    return;
  }
  double ext = 0.0;
  for (int k = 0; k < 3; ++k) ext = std::max(ext, hi[k] - lo[k]);
  if (!(ext > 0.0)) ext = 1.0;
  double h = 2.0 * mean_edge(tv_host, tf_host, F);
  if (!(h > ext / 4096.0)) h = ext / 4096.0;
  const hipStream_t s = nullptr;
  Dev<double> d2((size_t)Q);
  Dev<uint32_t> far_a((size_t)Q), far_b((size_t)Q), far_count(1), total(1);
  const unsigned nbF = blocks_for((uint64_t)F, kThreads);
  const uint32_t* ids = nullptr;  // round 0: every query
  uint32_t nq = (uint32_t)Q;
  uint32_t* far_out = far_a.p;
  for (;;) {
    const CellGrid g = make_grid(lo, hi, h);
    const uint64_t cells = (uint64_t)g.nc[0] * g.nc[1] * g.nc[2];
    Dev<uint32_t> counts(cells), starts(cells), cursor(cells);
    Dev<uint32_t> scan_scratch(scan_scratch_words((uint32_t)cells));
    launch_fill<uint32_t>(counts.p, 0u, cells, s);
    launch_fill<uint32_t>(cursor.p, 0u, cells, s);
    hipLaunchKernelGGL(k_bin_count, dim3(nbF), dim3(kThreads), 0, s, tv, tf, (uint32_t)F, g,
                       counts.p);
    MPR_LAUNCH_CHECK();
    launch_scan_exclusive(counts.p, starts.p, (uint32_t)cells, total.p, s, scan_scratch.p);
    uint32_t entries = 0;
    MPR_HIP(hipMemcpy(&entries, total.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    Dev<uint32_t> items(entries);
    hipLaunchKernelGGL(k_bin_fill, dim3(nbF), dim3(kThreads), 0, s, tv, tf, (uint32_t)F, g,
                       starts.p, cursor.p, items.p);
    MPR_LAUNCH_CHECK();
    launch_fill<uint32_t>(far_count.p, 0u, 1, s);
    hipLaunchKernelGGL(k_near, dim3(blocks_for(nq, kThreads)), dim3(kThreads), 0, s, qv, ids, nq,
                       tv, tf, g, starts.p, counts.p, items.p, d2.p, far_out, far_count.p);
    MPR_LAUNCH_CHECK();
    uint32_t nfar = 0;
    MPR_HIP(hipMemcpy(&nfar, far_count.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (nfar == 0) break;
    ids = far_out;
    nq = nfar;
    far_out = far_out == far_a.p ? far_b.p : far_a.p;
    h *= 8.0;
  }
  const unsigned nbQ = blocks_for((uint64_t)Q, kThreads);
  Dev<double> part(2ull * nbQ), res(2);
  hipLaunchKernelGGL(k_dist_partials, dim3(nbQ), dim3(kThreads), 0, s, d2.p, (uint32_t)Q, part.p);
  MPR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dist_finish, dim3(1), dim3(kThreads), 0, s, part.p, nbQ, (uint32_t)Q,
                     res.p);
  MPR_LAUNCH_CHECK();
  MPR_HIP(hipMemcpy(out, res.p, 2 * sizeof(double), hipMemcpyDeviceToHost));
}

bool faces_ok(const int64_t* f, int64_t F, int64_t nv) {
  for (int64_t i = 0; i < 3 * F; ++i)
    if (f[i] < 0 || f[i] >= nv) return false;
  return true;
}

}  // namespace

extern "C" {

int mpr_mesh_distance(const double* va, int64_t nva, const int64_t* fa, int64_t nfa,
                      const double* vb, int64_t nvb, const int64_t* fb, int64_t nfb, double* out) {
  const int64_t kMax = int64_t(1) << 31;
  if (!out || nva < 0 || nvb < 0 || nfa < 0 || nfb < 0 || nva >= kMax || nvb >= kMax ||
      nfa >= kMax / 3 || nfb >= kMax / 3 || (nva && !va) || (nvb && !vb) || (nfa && !fa) ||
      (nfb && !fb)) {
    set_error("mpr_mesh_distance: invalid sizes or buffers");
    return -1;
  }
  if (!faces_ok(fa, nfa, nva) || !faces_ok(fb, nfb, nvb)) {
    set_error("mpr_mesh_distance: face index out of range");
    return -1;
  }
  for (int64_t i = 0; i < 3 * nva; ++i)
    if (!std::isfinite(va[i])) {
      set_error("mpr_mesh_distance: vertices must be finite");
      return -1;
    }
  for (int64_t i = 0; i < 3 * nvb; ++i)
    if (!std::isfinite(vb[i])) {
      set_error("mpr_mesh_distance: vertices must be finite");
      return -1;
    }
  if (ensure_device()) return -1;
  try {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const auto& m : {std::make_pair(va, nva), std::make_pair(vb, nvb)})
      for (int64_t i = 0; i < m.second; ++i)
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::min(lo[k], m.first[3 * i + k]);
          hi[k] = std::max(hi[k], m.first[3 * i + k]);
        }
    Dev<double> dva(3 * (size_t)nva), dvb(3 * (size_t)nvb);
    Dev<int32_t> dfa(3 * (size_t)nfa), dfb(3 * (size_t)nfb);
    std::vector<int32_t> tmp;
    auto up_faces = [&](const int64_t* f, int64_t F, int32_t* d) {
      tmp.assign(f, f + 3 * F);
      if (F) MPR_HIP(hipMemcpy(d, tmp.data(), 12 * (size_t)F, hipMemcpyHostToDevice));
    };
    if (nva) MPR_HIP(hipMemcpy(dva.p, va, 24 * (size_t)nva, hipMemcpyHostToDevice));
    if (nvb) MPR_HIP(hipMemcpy(dvb.p, vb, 24 * (size_t)nvb, hipMemcpyHostToDevice));
    up_faces(fa, nfa, dfa.p);
    up_faces(fb, nfb, dfb.p);
    double ab[2], ba[2];
    directed(dva.p, nva, vb, dvb.p, fb, dfb.p, nfb, lo, hi, ab);
    directed(dvb.p, nvb, va, dva.p, fa, dfa.p, nfa, lo, hi, ba);
    out[0] = ab[0];
    out[1] = ba[0];
    out[2] = ab[1];
    out[3] = ba[1];
    return 0;
  } catch (const HipError& e) {
    set_error(std::string("mpr_mesh_distance: ") + e.msg);
    return -1;
  }
}

}  // extern "C"
