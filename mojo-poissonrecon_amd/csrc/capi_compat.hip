// Drop-in kernel ABI: the reference's eleven ctypes symbols (python/mojo_poissonrecon/_lib.py:
// 20-32) on caller-owned host buffers. Each call stages its buffers in HBM, runs the same
// HIP kernels the device pipeline uses (grids converted to and from the colour-split layout
// on the device) and copies results back before returning. The reference's native side
// cannot raise; neither do these: on a device error they report on stderr, set
// mpr_last_error() and return -1 / NaN.
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "mpr_batch.h"
#include "mpr_hip.h"
#include "mpr_internal.h"

namespace mpr {
int ensure_device();
}

using namespace mpr;

namespace {

template <class T>
struct DevBuf {
  T* p = nullptr;
  explicit DevBuf(size_t count) {
    MPR_HIP(hipMalloc((void**)&p, std::max<size_t>(1, count) * sizeof(T)));
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

template <class T>
void h2d(T* dst, const T* src, size_t count) {
  if (count) MPR_HIP(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyHostToDevice));
}

template <class T>
void d2h(T* dst, const T* src, size_t count) {
  if (count) MPR_HIP(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyDeviceToHost));
}

void report(const char* fn, const std::string& msg) {
  set_error(std::string(fn) + ": " + msg);
  std::fprintf(stderr, "libmpr_hip %s: %s\n", fn, msg.c_str());
}

const double kNaN = std::numeric_limits<double>::quiet_NaN();
const hipStream_t kNull = nullptr;

void upload_f32_as_split(const float* host, float* split, const Geom& g, DevBuf<double>& nat,
                         size_t M) {
  std::vector<double> wide(host, host + M);
  h2d(nat.p, wide.data(), M);
  launch_nat_to_split<double, float>(nat.p, split, g, kNull);
}
void download_split_f32(const float* split, float* host, const Geom& g, DevBuf<double>& nat,
                        size_t M) {
  launch_split_to_nat<float>(split, nat.p, g, kNull);
  std::vector<double> wide(M);
  d2h(wide.data(), nat.p, M);
  for (size_t i = 0; i < M; ++i) host[i] = (float)wide[i];
}

}  // namespace

extern "C" {

int64_t mpr_parallel_init(void) { return ensure_device() == 0 ? 0 : -1; }

void mpr_octree_codes_f64(const double* points, int64_t point_count, int64_t depth,
                          int64_t* codes) {
  if (point_count <= 0 || ensure_device()) return;
  try {
    DevBuf<double> p(3 * point_count);
    DevBuf<int64_t> c(point_count * (depth + 1));
    h2d(p.p, points, 3 * point_count);
    launch_octree_codes(p.p, (uint32_t)point_count, (int)depth, c.p, kNull);
    d2h(codes, c.p, point_count * (depth + 1));
  } catch (const HipError& e) {
    report(__func__, e.msg);
  }
}

int64_t mpr_splat_f64(const double* points, const double* normals, int64_t point_count, int64_t n,
                      int64_t confidence, double* density, double* vx, double* vy, double* vz,
                      double* screen, double samples_per_node) {
  if (point_count < 0 || point_count > kMaxPoints || n < 2) {
    report(__func__, "point_count must be in [0, 2^28] and n >= 2");
    return -1;
  }
  if (point_count == 0) return 0;
  if (ensure_device()) return -1;
  try {
    const size_t M = (size_t)n * n * n;
    DevBuf<double> p(3 * point_count), q(3 * point_count), d(M), x(M), y(M), z(M), s(M);
    DevBuf<unsigned long long> acc(1);
    h2d(p.p, points, 3 * point_count);
    h2d(q.p, normals, 3 * point_count);
    h2d(d.p, density, M);
    h2d(x.p, vx, M);
    h2d(y.p, vy, M);
    h2d(z.p, vz, M);
    h2d(s.p, screen, M);
    launch_fill<unsigned long long>(acc.p, 0ull, 1, kNull);
    SplatScratch sc = splat_scratch_sizes((uint32_t)point_count);
    DevBuf<uint32_t> keys(sc.pairs), ktmp(sc.pairs);
    DevBuf<uint32_t> ids(sc.pairs), itmp(sc.pairs), hist(sc.radix_words), offs(sc.radix_words);
    DevBuf<double> w1(sc.pairs), c2(3ull * sc.pairs);
    sc.keys = keys.p;
    sc.ktmp = ktmp.p;
    sc.ids = ids.p;
    sc.itmp = itmp.p;
    sc.hist = hist.p;
    sc.offs = offs.p;
    sc.w1 = w1.p;
    sc.c2 = c2.p;
    launch_splat(p.p, q.p, (uint32_t)point_count, (int)n, (int)(confidence != 0), d.p, x.p, y.p,
                 z.p, s.p, samples_per_node, acc.p, sc, nullptr, kNull);
    unsigned long long accepted = 0;
    d2h(&accepted, acc.p, 1);
    d2h(density, d.p, M);
    d2h(vx, x.p, M);
    d2h(vy, y.p, M);
    d2h(vz, z.p, M);
    d2h(screen, s.p, M);
    return (int64_t)accepted;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return -1;
  }
}

void mpr_assemble_f64(const double* vx, const double* vy, const double* vz, const double* screen,
                      double* rhs, int64_t n, double point_weight, double target_value) {
  if (ensure_device()) return;
  try {
    const size_t M = (size_t)n * n * n;
    DevBuf<double> x(M), y(M), z(M), s(M), r(M);
    h2d(x.p, vx, M);
    h2d(y.p, vy, M);
    h2d(z.p, vz, M);
    h2d(s.p, screen, M);
    launch_assemble(x.p, y.p, z.p, s.p, r.p, nullptr, nullptr, nullptr, 1.0,
                    make_geom((int32_t)n), point_weight, target_value, nullptr, nullptr,
                    nullptr, nullptr, nullptr, kNull);
    d2h(rhs, r.p, M);
  } catch (const HipError& e) {
    report(__func__, e.msg);
  }
}

void mpr_relax_f64(double* solution, const double* rhs, const double* screen, int64_t n,
                   double point_weight, int64_t iterations) {
  if (n < 1 || ensure_device()) return;
  try {
    const Geom g = make_geom((int32_t)n);
    const size_t M = (size_t)n * n * n, W = 2ull * g.C;
    DevBuf<double> nat(M), x(W), r(W), s(W);
    h2d(nat.p, solution, M);
    launch_nat_to_split<double, double>(nat.p, x.p, g, kNull);
    h2d(nat.p, rhs, M);
    launch_nat_to_split<double, double>(nat.p, r.p, g, kNull);
    h2d(nat.p, screen, M);
    launch_nat_to_split<double, double>(nat.p, s.p, g, kNull);
    launch_relax<double>(x.p, r.p, s.p, g, point_weight, (int)iterations, kNull);
    launch_split_to_nat<double>(x.p, nat.p, g, kNull);
    d2h(solution, nat.p, M);
  } catch (const HipError& e) {
    report(__func__, e.msg);
  }
}

double mpr_residual_f64(const double* solution, const double* rhs, const double* screen,
                        double* residual, int64_t n, double point_weight) {
  if (n < 1 || ensure_device()) return kNaN;
  try {
    const Geom g = make_geom((int32_t)n);
    const size_t M = (size_t)n * n * n, W = 2ull * g.C;
    DevBuf<double> nat(M), x(W), r(W), s(W), out(W), norm(1);
    DevBuf<double> part((size_t)residual_partial_count(g, sizeof(double)));
    h2d(nat.p, solution, M);
    launch_nat_to_split<double, double>(nat.p, x.p, g, kNull);
    h2d(nat.p, rhs, M);
    launch_nat_to_split<double, double>(nat.p, r.p, g, kNull);
    h2d(nat.p, screen, M);
    launch_nat_to_split<double, double>(nat.p, s.p, g, kNull);
    const int parts = launch_residual<double, double>(x.p, r.p, s.p, out.p, g, point_weight,
                                                      part.p, kNull);
    launch_norm_finish(part.p, parts, norm.p, kNull);
    launch_split_to_nat<double>(out.p, nat.p, g, kNull);
    d2h(residual, nat.p, M);
    double value = kNaN;
    d2h(&value, norm.p, 1);
    return value;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return kNaN;
  }
}

// FP32 instantiations of the smoother and residual (the mixed-precision V-cycle's
// kernels) on host float32 grids: staged through the same layout conversions.
int mpr_relax_f32(float* solution, const float* rhs, const float* screen, int64_t n,
                  double point_weight, int64_t iterations) {
  if (n < 1 || n > 1025 || iterations < 0) {
    set_error("mpr_relax_f32: n must be in [1, 1025] and iterations non-negative");
    return -1;
  }
  if (ensure_device()) return -1;
  try {
    const Geom g = make_geom((int32_t)n);
    const size_t M = (size_t)n * n * n, W = 2ull * g.C;
    DevBuf<double> nat(M);
    DevBuf<float> x(W), r(W), s(W);
    upload_f32_as_split(solution, x.p, g, nat, M);
    upload_f32_as_split(rhs, r.p, g, nat, M);
    upload_f32_as_split(screen, s.p, g, nat, M);
    launch_relax<float>(x.p, r.p, s.p, g, point_weight, (int)iterations, kNull);
    download_split_f32(x.p, solution, g, nat, M);
    return 0;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return -1;
  }
}

int mpr_residual_f32(const float* solution, const float* rhs, const float* screen,
                     float* residual, int64_t n, double point_weight) {
  if (n < 1 || n > 1025) {
    set_error("mpr_residual_f32: n must be in [1, 1025]");
    return -1;
  }
  if (ensure_device()) return -1;
  try {
    const Geom g = make_geom((int32_t)n);
    const size_t M = (size_t)n * n * n, W = 2ull * g.C;
    DevBuf<double> nat(M);
    DevBuf<float> x(W), r(W), s(W), out(W);
    upload_f32_as_split(solution, x.p, g, nat, M);
    upload_f32_as_split(rhs, r.p, g, nat, M);
    upload_f32_as_split(screen, s.p, g, nat, M);
    launch_residual<float, float>(x.p, r.p, s.p, out.p, g, point_weight, nullptr, kNull);
    download_split_f32(out.p, residual, g, nat, M);
    return 0;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return -1;
  }
}

int mpr_restrict_f32(const float* fine, float* coarse, int64_t fine_n) {
  if (fine_n < 1 || fine_n > 1025) {
    set_error("mpr_restrict_f32: fine_n must be in [1, 1025]");
    return -1;
  }
  if (ensure_device()) return -1;
  try {
    const int64_t cn = (fine_n + 1) / 2;
    const Geom gf = make_geom((int32_t)fine_n), gc = make_geom((int32_t)cn);
    const size_t Mf = (size_t)fine_n * fine_n * fine_n, Mc = (size_t)cn * cn * cn;
    DevBuf<double> nat(Mf);
    DevBuf<float> f(2ull * gf.C), c(2ull * gc.C);
    upload_f32_as_split(fine, f.p, gf, nat, Mf);
    launch_restrict<float>(f.p, gf, c.p, gc, 1.0, kNull);
    download_split_f32(c.p, coarse, gc, nat, Mc);
    return 0;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return -1;
  }
}

int mpr_prolong_add_f32(const float* coarse, float* fine, int64_t fine_n) {
  if (fine_n < 2 || fine_n > 1025) {
    set_error("mpr_prolong_add_f32: fine_n must be in [2, 1025]");
    return -1;
  }
  if (ensure_device()) return -1;
  try {
    const int64_t cn = (fine_n + 1) / 2;
    const Geom gf = make_geom((int32_t)fine_n), gc = make_geom((int32_t)cn);
    const size_t Mf = (size_t)fine_n * fine_n * fine_n, Mc = (size_t)cn * cn * cn;
    DevBuf<double> nat(Mf);
    DevBuf<float> f(2ull * gf.C), c(2ull * gc.C);
    upload_f32_as_split(coarse, c.p, gc, nat, Mc);
    upload_f32_as_split(fine, f.p, gf, nat, Mf);
    launch_prolong_add<float>(c.p, gc, f.p, gf, kNull);
    download_split_f32(f.p, fine, gf, nat, Mf);
    return 0;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return -1;
  }
}

int mpr_vcycle_transfer_f32(const float* solution, const float* rhs, const float* screen,
                            float* residual, float* coarse_rhs, int64_t n, double point_weight) {
  if (n < 65 || n > 1025 || ((n - 1) & (n - 2)) != 0 || !(point_weight >= 0.0)) {
    set_error("mpr_vcycle_transfer_f32: n must be 2^s + 1 in [65, 1025], point_weight >= 0");
    return -1;
  }
  if (ensure_device()) return -1;
  try {
    const int64_t cn = (n + 1) / 2;
    const Geom g = make_geom((int32_t)n), gc = make_geom((int32_t)cn);
    const size_t M = (size_t)n * n * n, Mc = (size_t)cn * cn * cn, W = 2ull * g.C;
    DevBuf<double> nat(M);
    DevBuf<float> x(W), r(W), sc(W), res(W), rc(2ull * gc.C);
    upload_f32_as_split(solution, x.p, g, nat, M);
    upload_f32_as_split(rhs, r.p, g, nat, M);
    upload_f32_as_split(screen, sc.p, g, nat, M);
    launch_fill<float>(res.p, 0.0f, W, kNull);
    launch_fill<float>(rc.p, 0.0f, 2ull * gc.C, kNull);
    StepPtrs<float> step{};
    step.x = x.p;
    step.rhs = r.p;
    step.screen = sc.p;
    step.res = res.p;
    step.rhsc = rc.p;
    step.top = 1;
    DevBuf<StepPtrs<float>> tab(1);
    h2d(tab.p, &step, 1);
    launch_b_residual<float>(tab.p, 1, g, point_weight, kNull, true);
    launch_b_restrict<float>(tab.p, 1, g, gc, false, kNull, true);
    download_split_f32(res.p, residual, g, nat, M);
    download_split_f32(rc.p, coarse_rhs, gc, nat, Mc);
    return 0;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return -1;
  }
}

int mpr_prolong_c1_f32(const float* coarse, float* fine, int64_t fine_n) {
  if (fine_n < 65 || fine_n > 1025 || ((fine_n - 1) & (fine_n - 2)) != 0) {
    set_error("mpr_prolong_c1_f32: fine_n must be 2^s + 1 in [65, 1025]");
    return -1;
  }
  if (ensure_device()) return -1;
  try {
    const int64_t cn = (fine_n + 1) / 2;
    const Geom gf = make_geom((int32_t)fine_n), gc = make_geom((int32_t)cn);
    const size_t Mf = (size_t)fine_n * fine_n * fine_n, Mc = (size_t)cn * cn * cn;
    DevBuf<double> nat(Mf);
    DevBuf<float> f(2ull * gf.C), c(2ull * gc.C);
    upload_f32_as_split(coarse, c.p, gc, nat, Mc);
    upload_f32_as_split(fine, f.p, gf, nat, Mf);
    StepPtrs<float> step{};
    step.x = f.p;
    step.xc = c.p;
    DevBuf<StepPtrs<float>> tab(1);
    h2d(tab.p, &step, 1);
    launch_b_prolong<float>(tab.p, 1, gc, gf, true, kNull);
    download_split_f32(f.p, fine, gf, nat, Mf);
    return 0;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return -1;
  }
}

void mpr_restrict_f64(const double* fine, double* coarse, int64_t fine_n) {
  if (fine_n < 1 || ensure_device()) return;
  try {
    const int64_t cn = (fine_n + 1) / 2;
    const Geom gf = make_geom((int32_t)fine_n), gc = make_geom((int32_t)cn);
    DevBuf<double> nf((size_t)fine_n * fine_n * fine_n), f(2ull * gf.C), c(2ull * gc.C),
        nc((size_t)cn * cn * cn);
    h2d(nf.p, fine, (size_t)fine_n * fine_n * fine_n);
    launch_nat_to_split<double, double>(nf.p, f.p, gf, kNull);
    launch_restrict<double>(f.p, gf, c.p, gc, 1.0, kNull);
    launch_split_to_nat<double>(c.p, nc.p, gc, kNull);
    d2h(coarse, nc.p, (size_t)cn * cn * cn);
  } catch (const HipError& e) {
    report(__func__, e.msg);
  }
}

void mpr_prolong_add_f64(const double* coarse, double* fine, int64_t fine_n) {
  if (fine_n < 1 || ensure_device()) return;
  try {
    const int64_t cn = (fine_n + 1) / 2;
    const Geom gf = make_geom((int32_t)fine_n), gc = make_geom((int32_t)cn);
    const size_t Mf = (size_t)fine_n * fine_n * fine_n, Mc = (size_t)cn * cn * cn;
    DevBuf<double> nf(Mf), f(2ull * gf.C), c(2ull * gc.C), nc(Mc);
    h2d(nc.p, coarse, Mc);
    launch_nat_to_split<double, double>(nc.p, c.p, gc, kNull);
    h2d(nf.p, fine, Mf);
    launch_nat_to_split<double, double>(nf.p, f.p, gf, kNull);
    launch_prolong_add<double>(c.p, gc, f.p, gf, kNull);
    launch_split_to_nat<double>(f.p, nf.p, gf, kNull);
    d2h(fine, nf.p, Mf);
  } catch (const HipError& e) {
    report(__func__, e.msg);
  }
}

double mpr_iso_value_f64(const double* field, const double* points, const double* normals,
                         int64_t point_count, int64_t n, int64_t confidence) {
  if (ensure_device()) return kNaN;
  try {
    const size_t M = (size_t)n * n * n;
    DevBuf<double> f(M), p(3 * point_count), q(3 * point_count), iso(1);
    DevBuf<double> part(2ull * iso_partial_count((uint32_t)std::max<int64_t>(1, point_count)));
    h2d(f.p, field, M);
    h2d(p.p, points, 3 * point_count);
    h2d(q.p, normals, 3 * point_count);
    launch_iso(natural_field(f.p, (int)n), p.p, q.p, (uint32_t)point_count, (int)n, (int)(confidence != 0), part.p, iso.p,
               kNull);
    double value = kNaN;
    d2h(&value, iso.p, 1);
    return value;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return kNaN;
  }
}

int64_t mpr_dual_vertices_f64(const double* field, const double* density, int64_t n, double iso,
                              int64_t* cell_map, double* vertices, double* vertex_density,
                              int64_t capacity) {
  if (n < 2) return 0;
  if (ensure_device()) return -1;
  try {
    const size_t M = (size_t)n * n * n, cells = (size_t)(n - 1), C3 = cells * cells * cells;
    DevBuf<double> f(M), isod(1);
    h2d(f.p, field, M);
    h2d(isod.p, &iso, 1);
    if (capacity <= 0) {  // count mode: per-4-plane partial counts (capi.mojo:855-903)
      const size_t chunks = (cells + 3) / 4;
      DevBuf<int64_t> ch(chunks);
      launch_fill<int64_t>(ch.p, 0, chunks, kNull);
      launch_cells_chunk_counts(f.p, (int)n, isod.p, ch.p, kNull);
      std::vector<int64_t> h(chunks);
      d2h(h.data(), ch.p, chunks);
      int64_t total = 0;
      for (size_t i = 0; i < chunks; ++i) {
        cell_map[i] = h[i];
        total += h[i];
      }
      return total;
    }
    DevBuf<double> d(M);
    h2d(d.p, density, M);
    const uint32_t tiles = cell_tiles((int)n);
    DevBuf<uint32_t> tc(tiles), to(tiles), total(1);
    DevBuf<int64_t> cm(C3);
    launch_cells_count(f.p, (int)n, isod.p, tc.p, kNull);
    launch_scan_exclusive(tc.p, to.p, tiles, total.p, kNull);
    uint32_t count = 0;
    d2h(&count, total.p, 1);
    const size_t written = std::min<size_t>(count, (size_t)capacity);
    DevBuf<double> v(3 * written), vd(written);
    launch_cells_emit<int64_t>(f.p, d.p, (int)n, isod.p, to.p, cm.p, v.p, vd.p, (uint64_t)capacity,
                               nullptr, kNull);
    d2h(cell_map, cm.p, C3);
    d2h(vertices, v.p, 3 * written);
    d2h(vertex_density, vd.p, written);
    return (int64_t)count;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return -1;
  }
}

int64_t mpr_dual_faces_f64(const double* field, int64_t n, double iso, const int64_t* cell_map,
                           int64_t* faces, int64_t capacity) {
  if (n < 2) return 0;
  if (ensure_device()) return -1;
  try {
    const size_t M = (size_t)n * n * n, cells = (size_t)(n - 1), C3 = cells * cells * cells;
    DevBuf<double> f(M), isod(1);
    DevBuf<int64_t> cm(C3);
    h2d(f.p, field, M);
    h2d(isod.p, &iso, 1);
    h2d(cm.p, cell_map, C3);
    const uint32_t tiles = edge_tiles((int)n);
    DevBuf<uint32_t> tc(tiles), to(tiles), total(1);
    launch_edges_count<int64_t>(f.p, (int)n, isod.p, cm.p, tc.p, kNull);
    launch_scan_exclusive(tc.p, to.p, tiles, total.p, kNull);
    uint32_t quads = 0;
    d2h(&quads, total.p, 1);
    const int64_t face_count = 2 * (int64_t)quads;
    if (capacity > 0 && quads) {
      const size_t written = 2 * std::min<size_t>(quads, (size_t)capacity / 2);
      DevBuf<int64_t> out(3 * std::max<size_t>(1, written));
      launch_edges_emit<int64_t>(f.p, (int)n, isod.p, cm.p, to.p, out.p, (uint64_t)capacity, kNull);
      d2h(faces, out.p, 3 * written);
    }
    return face_count;
  } catch (const HipError& e) {
    report(__func__, e.msg);
    return -1;
  }
}

}  // extern "C"
