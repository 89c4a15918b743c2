// Oriented points of a triangle mesh (core.py:429-450): every referenced vertex with a
// finite, non-zero sum of the (unnormalised) cross products of its faces, normalised.
//
// The reference adds each face's cross product to its three corners in face order, skipping
// faces whose cross product is not finite or has a zero squared norm. Here each kept face
// emits (vertex, 3 f + corner) pairs in face order; a stable radix sort by vertex keeps each
// vertex's faces in that order, and one lane per vertex adds them sequentially from +0.0:
// the same additions in the same order, so the normals are bit-identical to the per-face
// loop (cross product a1 b2 - a2 b1, ..., as numpy.cross; norm sqrt((x^2 + y^2) + z^2), as
// numpy.linalg.norm over the last axis of three). Kept vertices are then compacted in order.
#include <algorithm>
#include <cmath>

#include "block_scan.h"
#include "mpr_hip.h"
#include "mpr_internal.h"

namespace mpr {
int ensure_device();
}

using namespace mpr;

namespace {

__device__ __forceinline__ void face_cross(const double* v, const int64_t* f, uint32_t t,
                                           double (&c)[3], bool& ok) {
  const int64_t a = f[3ull * t], b = f[3ull * t + 1], d = f[3ull * t + 2];
  const double e1[3] = {v[3 * b] - v[3 * a], v[3 * b + 1] - v[3 * a + 1], v[3 * b + 2] - v[3 * a + 2]};
  const double e2[3] = {v[3 * d] - v[3 * a], v[3 * d + 1] - v[3 * a + 1], v[3 * d + 2] - v[3 * a + 2]};
  c[0] = e1[1] * e2[2] - e1[2] * e2[1];
  c[1] = e1[2] * e2[0] - e1[0] * e2[2];
  c[2] = e1[0] * e2[1] - e1[1] * e2[0];
  // np.isfinite(cross).all() and np.dot(cross, cross) > 0: a sum of squares is positive
  // exactly when one of its (rounded) squares is
  ok = isfinite(c[0]) && isfinite(c[1]) && isfinite(c[2]) &&
       (c[0] * c[0] > 0.0 || c[1] * c[1] > 0.0 || c[2] * c[2] > 0.0);
}

// pairs (vertex, 3 t + corner) of kept faces, (sentinel, ...) for skipped ones
__global__ __launch_bounds__(kThreads) void k_face_pairs(const double* v, const int64_t* f,
                                                         uint32_t F, uint32_t sentinel,
                                                         uint32_t* keys, uint32_t* ids) {
  const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= F) return;
  double c[3];
  bool ok;
  face_cross(v, f, t, c, ok);
  for (int k = 0; k < 3; ++k) {
    keys[3ull * t + k] = ok ? (uint32_t)f[3ull * t + k] : sentinel;
    ids[3ull * t + k] = 3u * t + (uint32_t)k;
  }
}

// One lane per run of equal keys (a vertex): the sum of its faces' cross products in face
// order, then kept / normalised as core.py:447-450.
__global__ __launch_bounds__(kThreads) void k_vertex_normals(const uint32_t* keys,
                                                             const uint32_t* ids, uint32_t n,
                                                             uint32_t sentinel, const double* v,
                                                             const int64_t* f, double* normals,
                                                             uint32_t* keep) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint32_t key = keys[i];
  if (key == sentinel || (i > 0 && keys[i - 1] == key)) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (uint32_t j = i; j < n && keys[j] == key; ++j) {
    double c[3];
    bool ok;
    face_cross(v, f, ids[j] / 3u, c, ok);
    s[0] += c[0];
    s[1] += c[1];
    s[2] += c[2];
  }
  const double len = sqrt((s[0] * s[0] + s[1] * s[1]) + s[2] * s[2]);
  const bool k = len > 0.0 && isfinite(len);
  keep[key] = k ? 1u : 0u;
  if (k)
    for (int a = 0; a < 3; ++a) normals[3ull * key + a] = s[a] / len;
}

__global__ __launch_bounds__(kThreads) void k_keep_counts(const uint32_t* keep, uint32_t V,
                                                          uint32_t* counts) {
  __shared__ uint32_t lds4[4];
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  const uint32_t c = block_count256(i < V ? keep[i] : 0u, lds4);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

__global__ __launch_bounds__(kThreads) void k_keep_compact(const uint32_t* keep, uint32_t V,
                                                           const uint32_t* offsets,
                                                           const double* v, const double* nrm,
                                                           double* v_out, double* n_out) {
  __shared__ uint32_t lds4[4];
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  const bool k = i < V && keep[i] != 0u;
  uint32_t tot;
  const uint32_t rank = block_prefix256(k, lds4, tot);
  if (!k) return;
  const uint64_t d = offsets[blockIdx.x] + rank;
  for (int a = 0; a < 3; ++a) {
    v_out[3 * d + a] = v[3ull * i + a];
    n_out[3 * d + a] = nrm[3ull * i + a];
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

}  // namespace

extern "C" {

int mpr_oriented_points_from_mesh(const double* vertices, int64_t vertex_count,
                                  const int64_t* faces, int64_t face_count, double* points_out,
                                  double* normals_out, int64_t* count_out) {
  if (!count_out || vertex_count < 0 || face_count < 0 || vertex_count >= (int64_t(1) << 31) ||
      face_count >= (int64_t(1) << 30) || (vertex_count && (!vertices || !points_out || !normals_out)) ||
      (face_count && !faces)) {
    set_error("mpr_oriented_points_from_mesh: invalid sizes or buffers");
    return -1;
  }
  for (int64_t i = 0; i < 3 * face_count; ++i)
    if (faces[i] < 0 || faces[i] >= vertex_count) {
      set_error("mpr_oriented_points_from_mesh: face index out of range");
      return -1;
    }
  *count_out = 0;
  if (vertex_count == 0 || face_count == 0) return 0;
  if (ensure_device()) return -1;
  try {
    const hipStream_t s = nullptr;
    const uint32_t V = (uint32_t)vertex_count, F = (uint32_t)face_count, n = 3u * F;
    int bits = 1;
    while (bits < 32 && (uint64_t(1) << bits) <= (uint64_t)V) ++bits;  // V is the sentinel
    Dev<double> v(3ull * V), nrm(3ull * V), vo(3ull * V), no(3ull * V);
    Dev<int64_t> f(3ull * F);
    Dev<uint32_t> keys(n), ids(n), ktmp(n), itmp(n), keep(V);
    const size_t rw = radix_scratch_words(n);
    Dev<uint32_t> hist(rw), offs(rw);
    const uint32_t nb = std::max<uint32_t>(1u, blocks_for(V, kThreads));
    Dev<uint32_t> counts(nb), offsets(nb), total(1), scan(scan_scratch_words(nb));
    MPR_HIP(hipMemcpy(v.p, vertices, 24ull * V, hipMemcpyHostToDevice));
    MPR_HIP(hipMemcpy(f.p, faces, 24ull * F, hipMemcpyHostToDevice));
    launch_fill<uint32_t>(keep.p, 0u, V, s);
    hipLaunchKernelGGL(k_face_pairs, dim3(blocks_for(F, kThreads)), dim3(kThreads), 0, s, v.p,
                       f.p, F, V, keys.p, ids.p);
    MPR_LAUNCH_CHECK();
    launch_radix_sort<uint32_t>(keys.p, ids.p, ktmp.p, itmp.p, n, bits, hist.p, offs.p, s);
    hipLaunchKernelGGL(k_vertex_normals, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s,
                       keys.p, ids.p, n, V, v.p, f.p, nrm.p, keep.p);
    MPR_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_keep_counts, dim3(nb), dim3(kThreads), 0, s, keep.p, V, counts.p);
    MPR_LAUNCH_CHECK();
    launch_scan_exclusive(counts.p, offsets.p, nb, total.p, s, scan.p);
    hipLaunchKernelGGL(k_keep_compact, dim3(nb), dim3(kThreads), 0, s, keep.p, V, offsets.p, v.p,
                       nrm.p, vo.p, no.p);
    MPR_LAUNCH_CHECK();
    uint32_t kept = 0;
    MPR_HIP(hipMemcpy(&kept, total.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (kept) {
      MPR_HIP(hipMemcpy(points_out, vo.p, 24ull * kept, hipMemcpyDeviceToHost));
      MPR_HIP(hipMemcpy(normals_out, no.p, 24ull * kept, hipMemcpyDeviceToHost));
    }
    *count_out = kept;
    return 0;
  } catch (const HipError& e) {
    set_error(std::string("mpr_oriented_points_from_mesh: ") + e.msg);
    return -1;
  }
}

}  // extern "C"
