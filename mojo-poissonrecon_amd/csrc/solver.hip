// Host-side driver of the device-resident pipeline: jobs (one reconstruction slot each),
// the multigrid V-cycle schedule, the FP64 defect-correction loop around FP32 V-cycles,
// extraction with capacity growth, the octree build, HIP graph capture, and the C ABI.
//
// The V-cycle schedule is the reference's (core.py:141-193, 274-285): pre 3 / post 3
// red-black iterations, correction scheme, x = 0 on coarse levels before each descent,
// 40 iterations on the coarsest (n <= 3) grid. Levels small enough to live in LDS are
// handled by one single-workgroup kernel per cycle instead of ~100 tiny launches.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "mpr_hip.h"
#include "mpr_internal.h"

namespace mpr {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

int ensure_device() {
  static std::once_flag once;
  static int status = 0;
  std::call_once(once, [] {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
      status = -1;
      return;
    }
    if (hipSetDevice(0) != hipSuccess) status = -1;
  });
  if (status != 0) set_error("no HIP device available");
  return status;
}

namespace {

struct DevArena {
  std::vector<void*> ptrs;
  size_t bytes = 0;
  bool dry = false;  // footprint accounting only
  template <class T>
  T* alloc(size_t count) {
    const size_t b = std::max<size_t>(256, ((count * sizeof(T)) + 255) & ~size_t(255));
    bytes += b;
    if (dry) return nullptr;
    void* p = nullptr;
    MPR_HIP(hipMalloc(&p, b));
    ptrs.push_back(p);
    return static_cast<T*>(p);
  }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
    bytes = 0;
  }
  DevArena() = default;
  DevArena(const DevArena&) = delete;
  DevArena& operator=(const DevArena&) = delete;
  ~DevArena() { release(); }
};

struct HostScalars {
  unsigned long long accepted;
  uint32_t counts[2];  // active cells, quads
  double iso;
  double resnorm;
  double params[4];  // origin xyz, width
};

template <class T>
int coarse_start(const std::vector<LevelPtrs<T>>& L) {
  const int nl = (int)L.size();
  int from = nl - 1;
  std::vector<Geom> g(nl);
  for (int l = 0; l < nl; ++l) g[l] = L[l].g;
  while (from > 0 && nl - (from - 1) <= kMaxCoarseLevels &&
         coarse_lds_bytes<T>(&g[from - 1], nl - (from - 1)) <= 160 * 1024)
    --from;
  return from;
}

// core.py:141-193 for levels [top, L.size()); multi-block kernels above `cfrom`, one
// LDS-resident workgroup from `cfrom` down. `zero_top`: the top level's x starts at 0.
template <class T>
void vcycle(std::vector<LevelPtrs<T>>& L, int top, int cfrom, bool zero_top, double pw, int pre,
            int post, hipStream_t s) {
  int l = top;
  if (zero_top && top < cfrom) launch_fill<T>(L[top].x, T(0), 2ull * L[top].g.C, s);
  for (; l < cfrom; ++l) {
    launch_relax<T>(L[l].x, L[l].rhs, L[l].screen, L[l].g, pw, pre, s);
    launch_residual<T, T>(L[l].x, L[l].rhs, L[l].screen, L[l].res, L[l].g, pw, nullptr, s);
    launch_restrict<T>(L[l].res, L[l].g, L[l + 1].rhs, L[l + 1].g, 1.0, s);
    // solution[l+1].fill(0) (core.py:174); the coarse kernel zeroes its own top level
    if (l + 1 < cfrom) launch_fill<T>(L[l + 1].x, T(0), 2ull * L[l + 1].g.C, s);
  }
  const bool zero_c = l > top || zero_top;
  if (!launch_coarse_vcycle<T>(&L[l], (int)L.size() - l, zero_c, pw, pre, post, s))
    throw HipError{"coarse levels do not fit in LDS"};
  for (l = l - 1; l >= top; --l) {
    launch_prolong_add<T>(L[l + 1].x, L[l + 1].g, L[l].x, L[l].g, s);
    launch_relax<T>(L[l].x, L[l].rhs, L[l].screen, L[l].g, pw, post, s);
  }
}

}  // namespace

struct Job {
  int depth = 0, n = 0, precision = 0;
  uint32_t max_points = 0, P = 0;
  hipStream_t stream = nullptr;
  DevArena arena;
  std::vector<Geom> geoms;
  // inputs
  double *pts = nullptr, *nrm = nullptr, *unit = nullptr;
  // natural-layout grids
  double *density = nullptr, *screen = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr,
         *rhs = nullptr, *field = nullptr;
  std::vector<LevelPtrs<double>> L64;
  std::vector<LevelPtrs<float>> L32;
  int cfrom64 = 0, cfrom32 = 0;
  // scalars and partial sums
  unsigned long long* d_accepted = nullptr;
  double *d_iso = nullptr, *d_resnorm = nullptr, *d_params = nullptr;
  uint32_t* d_counts = nullptr;
  double *norm_part = nullptr, *bbox_part = nullptr, *iso_part = nullptr;
  int norm_nparts = 0;
  // extraction
  uint32_t *cell_tc = nullptr, *cell_to = nullptr, *edge_tc = nullptr, *edge_to = nullptr;
  int32_t* cell_map = nullptr;
  double *verts = nullptr, *vdens = nullptr;
  int64_t* faces = nullptr;
  uint64_t vcap = 0, fcap = 0;
  std::vector<void*> mesh_bufs;  // grown extraction outputs (owned separately)
  // host results
  HostScalars* host = nullptr;
  mpr_params last{};
  bool ran = false, grids_kept = false;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // octree (allocated on demand)
  DevArena oct;
  int64_t oct_nodes = 0;
  int oct_depth = -1;
  int64_t *o_depth = nullptr, *o_morton = nullptr, *o_parent = nullptr, *o_cstart = nullptr,
          *o_ccount = nullptr, *o_leaf = nullptr;

  void plan(DevArena& a);
  void enqueue(const mpr_params& p);
  void enqueue_extract(bool world);
  ~Job();
};

void Job::plan(DevArena& a) {
  n = (1 << depth) + 1;
  const uint64_t M = (uint64_t)n * n * n;
  const uint64_t P3 = 3ull * std::max<uint32_t>(1u, max_points);
  pts = a.alloc<double>(P3);
  nrm = a.alloc<double>(P3);
  unit = a.alloc<double>(P3);
  density = a.alloc<double>(M);
  screen = a.alloc<double>(M);
  vx = a.alloc<double>(M);
  vy = a.alloc<double>(M);
  vz = a.alloc<double>(M);
  rhs = a.alloc<double>(M);
  field = a.alloc<double>(M);
  geoms.clear();
  for (int m = n;; m = (m + 1) / 2) {
    geoms.push_back(make_geom(m));
    if (m <= 3) break;
  }
  L64.clear();
  L32.clear();
  for (const Geom& g : geoms) {
    LevelPtrs<double> l;
    const uint64_t w = 2ull * g.C;
    l.g = g;
    l.x = a.alloc<double>(w);
    l.rhs = a.alloc<double>(w);
    l.screen = a.alloc<double>(w);
    l.res = a.alloc<double>(w);
    L64.push_back(l);
    if (precision == MPR_PRECISION_MIXED) {
      LevelPtrs<float> f;
      f.g = g;
      f.x = a.alloc<float>(w);
      f.rhs = a.alloc<float>(w);
      f.screen = a.alloc<float>(w);
      f.res = a.alloc<float>(w);
      L32.push_back(f);
    }
  }
  d_accepted = a.alloc<unsigned long long>(1);
  d_iso = a.alloc<double>(1);
  d_resnorm = a.alloc<double>(1);
  d_params = a.alloc<double>(4);
  d_counts = a.alloc<uint32_t>(2);
  norm_nparts = (int)blocks_for(2ull * geoms[0].C, kThreads);
  norm_part = a.alloc<double>(norm_nparts);
  bbox_part = a.alloc<double>(6ull * bbox_partial_count(std::max<uint32_t>(1u, max_points)));
  iso_part = a.alloc<double>(2ull * iso_partial_count(std::max<uint32_t>(1u, max_points)));
  cell_tc = a.alloc<uint32_t>(cell_tiles(n));
  cell_to = a.alloc<uint32_t>(cell_tiles(n));
  edge_tc = a.alloc<uint32_t>(edge_tiles(n));
  edge_to = a.alloc<uint32_t>(edge_tiles(n));
  const uint64_t cells = (uint64_t)(n - 1);
  cell_map = a.alloc<int32_t>(cells * cells * cells);
  vcap = std::min<uint64_t>(cells * cells * cells, std::max<uint64_t>(1u << 16, 12ull * n * n));
  fcap = 5 * vcap / 2 + 16;
  verts = a.alloc<double>(3 * vcap);
  vdens = a.alloc<double>(vcap);
  faces = a.alloc<int64_t>(3 * fcap);
  cfrom64 = coarse_start(L64);
  if (!L32.empty()) cfrom32 = coarse_start(L32);
}

Job::~Job() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (void* p : mesh_bufs) (void)hipFree(p);
  oct.release();
  arena.release();
  if (host) (void)hipHostFree(host);
  if (stream) (void)hipStreamDestroy(stream);
}

void Job::enqueue_extract(bool world) {
  hipStream_t s = stream;
  launch_cells_count(field, n, d_iso, cell_tc, s);
  launch_scan_exclusive(cell_tc, cell_to, cell_tiles(n), &d_counts[0], s);
  launch_cells_emit<int32_t>(field, density, n, d_iso, cell_to, cell_map, verts, vdens, vcap,
                             world ? d_params : nullptr, s);
  launch_edges_count<int32_t>(field, n, d_iso, cell_map, edge_tc, s);
  launch_scan_exclusive(edge_tc, edge_to, edge_tiles(n), &d_counts[1], s);
  launch_edges_emit<int32_t>(field, n, d_iso, cell_map, edge_to, faces, fcap, s);
}

void Job::enqueue(const mpr_params& p) {
  hipStream_t s = stream;
  const uint64_t M = (uint64_t)n * n * n;
  const double* upts = pts;
  if (p.stages & MPR_STAGE_NORMALIZE) {
    launch_normalize(pts, P, p.scale, bbox_part, d_params, unit, s);
    upts = unit;
  }
  if (p.stages & MPR_STAGE_SOLVE) {
    const double pw = p.point_weight;
    launch_fill<double>(density, 0.0, M, s);
    launch_fill<double>(screen, 0.0, M, s);
    launch_fill<double>(vx, 0.0, M, s);
    launch_fill<double>(vy, 0.0, M, s);
    launch_fill<double>(vz, 0.0, M, s);
    launch_fill<unsigned long long>(d_accepted, 0ull, 1, s);
    launch_splat(upts, nrm, P, n, p.confidence, density, vx, vy, vz, screen, p.samples_per_node,
                 d_accepted, s);
    launch_assemble(vx, vy, vz, screen, (p.stages & MPR_STAGE_GRIDS) ? rhs : nullptr, L64[0].rhs,
                    L64[0].screen, density, p.samples_per_node, geoms[0], pw, 0.5, s);
    for (size_t l = 1; l < L64.size(); ++l)
      launch_restrict<double>(L64[l - 1].screen, geoms[l - 1], L64[l].screen, geoms[l], 4.0, s);
    if (P) {
      launch_fill<double>(L64[0].x, 0.5, 2ull * geoms[0].C, s);
      if (precision == MPR_PRECISION_MIXED)
        for (size_t l = 0; l < L32.size(); ++l)
          launch_convert<double, float>(L64[l].screen, L32[l].screen, 2ull * geoms[l].C, s);
      for (int c = 0; c < p.cycles; ++c) {
        if (precision == MPR_PRECISION_MIXED) {
          // defect correction: r = b - A x in FP64, e = V32(r) from e = 0, x += e
          launch_residual<double, float>(L64[0].x, L64[0].rhs, L64[0].screen, L32[0].rhs,
                                         geoms[0], pw, nullptr, s);
          vcycle<float>(L32, 0, cfrom32, true, pw, p.pre_smooth, p.post_smooth, s);
          launch_add_correction(L64[0].x, L32[0].x, 2ull * geoms[0].C, s);
        } else {
          vcycle<double>(L64, 0, cfrom64, false, pw, p.pre_smooth, p.post_smooth, s);
        }
      }
    } else {
      launch_fill<double>(L64[0].x, 0.0, 2ull * geoms[0].C, s);  // core.py:286-287
    }
    launch_residual<double, double>(L64[0].x, L64[0].rhs, L64[0].screen, L64[0].res, geoms[0], pw,
                                    norm_part, s);
    launch_norm_finish(norm_part, norm_nparts, d_resnorm, s);
    launch_split_to_nat<double>(L64[0].x, field, geoms[0], s);
    if (P)
      launch_iso(field, upts, nrm, P, n, p.confidence, iso_part, d_iso, s);
    else
      launch_fill<double>(d_iso, 0.0, 1, s);
  }
  if (p.stages & MPR_STAGE_EXTRACT) enqueue_extract((p.stages & MPR_STAGE_WORLD) != 0);
  MPR_HIP(hipMemcpyAsync(&host->accepted, d_accepted, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, s));
  MPR_HIP(hipMemcpyAsync(host->counts, d_counts, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MPR_HIP(hipMemcpyAsync(&host->iso, d_iso, sizeof(double), hipMemcpyDeviceToHost, s));
  MPR_HIP(hipMemcpyAsync(&host->resnorm, d_resnorm, sizeof(double), hipMemcpyDeviceToHost, s));
  MPR_HIP(hipMemcpyAsync(host->params, d_params, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
}

}  // namespace mpr

struct mpr_job : mpr::Job {};

struct mpr_mesh {
  int n = 0;
  hipStream_t stream = nullptr;
  mpr::DevArena arena;
  uint64_t nv = 0, nq = 0;
  double *verts = nullptr, *vdens = nullptr;
  int64_t* faces = nullptr;
  ~mpr_mesh() {
    arena.release();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

using namespace mpr;

#define MPR_GUARD_BEGIN try {
#define MPR_GUARD_END                      \
  }                                        \
  catch (const HipError& e) {              \
    set_error(e.msg);                      \
    return -2;                             \
  }                                        \
  catch (const std::exception& e) {        \
    set_error(e.what());                   \
    return -3;                             \
  }

static bool valid_params(const mpr_params* p) {
  if (!p) return false;
  return p->cycles >= 0 && p->pre_smooth >= 0 && p->post_smooth >= 0 && p->samples_per_node > 0;
}

extern "C" {

const char* mpr_last_error(void) { return mpr::last_error(); }

int mpr_device_info(char* name, int64_t name_len, int64_t* total_bytes, int64_t* cu_count) {
  if (ensure_device()) return -1;
  MPR_GUARD_BEGIN
  hipDeviceProp_t prop;
  MPR_HIP(hipGetDeviceProperties(&prop, 0));
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (total_bytes) *total_bytes = (int64_t)prop.totalGlobalMem;
  if (cu_count) *cu_count = prop.multiProcessorCount;
  return 0;
  MPR_GUARD_END
}

int64_t mpr_job_footprint(int32_t depth, int64_t max_points, int32_t precision) {
  if (depth < 1 || depth > 10 || max_points < 0 || max_points > 0xffffffffll) return -1;
  Job j;
  j.depth = depth;
  j.max_points = (uint32_t)max_points;
  j.precision = precision;
  DevArena dry;
  dry.dry = true;
  try {
    j.plan(dry);
  } catch (...) {
    return -1;
  }
  return (int64_t)dry.bytes;
}

int mpr_job_create(int32_t depth, int64_t max_points, int32_t precision, mpr_job** out) {
  if (!out) return -1;
  *out = nullptr;
  if (depth < 1 || depth > 10) {
    set_error("depth must be between 1 and 10");
    return -1;
  }
  if (max_points < 0 || max_points > 0x7fffffffll) {
    set_error("max_points out of range");
    return -1;
  }
  if (precision != MPR_PRECISION_FP64 && precision != MPR_PRECISION_MIXED) {
    set_error("precision must be MPR_PRECISION_FP64 or MPR_PRECISION_MIXED");
    return -1;
  }
  if (ensure_device()) return -1;
  mpr_job* j = new mpr_job();
  MPR_GUARD_BEGIN
  j->depth = depth;
  j->max_points = (uint32_t)max_points;
  j->precision = precision;
  MPR_HIP(hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking));
  MPR_HIP(hipHostMalloc((void**)&j->host, sizeof(HostScalars), hipHostMallocDefault));
  std::memset(j->host, 0, sizeof(HostScalars));
  j->plan(j->arena);
  const double ident[4] = {0.0, 0.0, 0.0, 1.0};
  MPR_HIP(hipMemcpyAsync(j->d_params, ident, sizeof(ident), hipMemcpyHostToDevice, j->stream));
  MPR_HIP(hipMemsetAsync(j->d_counts, 0, 2 * sizeof(uint32_t), j->stream));
  MPR_HIP(hipMemsetAsync(j->d_accepted, 0, sizeof(unsigned long long), j->stream));
  MPR_HIP(hipStreamSynchronize(j->stream));
  *out = j;
  return 0;
  }
  catch (const HipError& e) {
    set_error(e.msg);
    delete j;
    return -2;
  }
}

int mpr_job_destroy(mpr_job* job) {
  if (!job) return 0;
  if (job->stream) (void)hipStreamSynchronize(job->stream);
  delete job;
  return 0;
}

int mpr_job_set_points(mpr_job* job, const double* points, const double* normals, int64_t P) {
  if (!job || P < 0 || (uint64_t)P > job->max_points || (P && (!points || !normals))) {
    set_error("invalid points for job (null buffer or more than max_points)");
    return -1;
  }
  MPR_GUARD_BEGIN
  job->P = (uint32_t)P;
  if (P) {
    MPR_HIP(hipMemcpyAsync(job->pts, points, 24ull * P, hipMemcpyHostToDevice, job->stream));
    MPR_HIP(hipMemcpyAsync(job->nrm, normals, 24ull * P, hipMemcpyHostToDevice, job->stream));
  }
  job->oct_depth = -1;
  return 0;
  MPR_GUARD_END
}

int mpr_job_device_points(mpr_job* job, double** points, double** normals) {
  if (!job) return -1;
  if (points) *points = job->pts;
  if (normals) *normals = job->nrm;
  return 0;
}

int mpr_job_set_count(mpr_job* job, int64_t P) {
  if (!job || P < 0 || (uint64_t)P > job->max_points) return -1;
  job->P = (uint32_t)P;
  job->oct_depth = -1;
  return 0;
}

int mpr_job_run(mpr_job* job, const mpr_params* params) {
  if (!job || !valid_params(params)) {
    set_error("invalid job or parameters");
    return -1;
  }
  MPR_GUARD_BEGIN
  job->last = *params;
  job->grids_kept = (params->stages & MPR_STAGE_GRIDS) != 0;
  job->enqueue(*params);
  job->ran = true;
  return 0;
  MPR_GUARD_END
}

int mpr_job_capture(mpr_job* job, const mpr_params* params) {
  if (!job || !valid_params(params)) {
    set_error("invalid job or parameters");
    return -1;
  }
  MPR_GUARD_BEGIN
  if (job->exec) {
    MPR_HIP(hipGraphExecDestroy(job->exec));
    job->exec = nullptr;
  }
  if (job->graph) {
    MPR_HIP(hipGraphDestroy(job->graph));
    job->graph = nullptr;
  }
  MPR_HIP(hipStreamSynchronize(job->stream));
  MPR_HIP(hipStreamBeginCapture(job->stream, hipStreamCaptureModeThreadLocal));
  try {
    job->enqueue(*params);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(job->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  MPR_HIP(hipStreamEndCapture(job->stream, &job->graph));
  MPR_HIP(hipGraphInstantiate(&job->exec, job->graph, nullptr, nullptr, 0));
  job->last = *params;
  job->grids_kept = (params->stages & MPR_STAGE_GRIDS) != 0;
  return 0;
  MPR_GUARD_END
}

int mpr_job_launch(mpr_job* job) {
  if (!job || !job->exec) {
    set_error("job has no captured graph");
    return -1;
  }
  MPR_GUARD_BEGIN
  MPR_HIP(hipGraphLaunch(job->exec, job->stream));
  job->ran = true;
  return 0;
  MPR_GUARD_END
}

int mpr_job_query(mpr_job* job) {
  if (!job) return -1;
  hipError_t e = hipStreamQuery(job->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  set_error(hipGetErrorString(e));
  return -2;
}

int mpr_job_wait(mpr_job* job, mpr_result* r) {
  if (!job) return -1;
  MPR_GUARD_BEGIN
  MPR_HIP(hipStreamSynchronize(job->stream));
  if (job->ran && (job->last.stages & MPR_STAGE_EXTRACT)) {
    const uint64_t nv = job->host->counts[0], nf = 2ull * job->host->counts[1];
    if (nv > job->vcap || nf > job->fcap) {
      // grow the mesh buffers to the exact counts and re-run the extraction
      const uint64_t vcap = std::max<uint64_t>(nv, job->vcap), fcap = std::max<uint64_t>(nf, job->fcap);
      void *v = nullptr, *d = nullptr, *f = nullptr;
      MPR_HIP(hipMalloc(&v, 24 * vcap));
      MPR_HIP(hipMalloc(&d, 8 * vcap));
      MPR_HIP(hipMalloc(&f, 24 * fcap));
      job->mesh_bufs.push_back(v);
      job->mesh_bufs.push_back(d);
      job->mesh_bufs.push_back(f);
      job->verts = (double*)v;
      job->vdens = (double*)d;
      job->faces = (int64_t*)f;
      job->vcap = vcap;
      job->fcap = fcap;
      if (job->exec) {  // the captured graph points at the old buffers
        MPR_HIP(hipGraphExecDestroy(job->exec));
        MPR_HIP(hipGraphDestroy(job->graph));
        job->exec = nullptr;
        job->graph = nullptr;
      }
      job->enqueue_extract((job->last.stages & MPR_STAGE_WORLD) != 0);
      MPR_HIP(hipMemcpyAsync(job->host->counts, job->d_counts, 2 * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, job->stream));
      MPR_HIP(hipStreamSynchronize(job->stream));
    }
  }
  if (r) {
    r->accepted = (int64_t)job->host->accepted;
    r->vertex_count = job->host->counts[0];
    r->face_count = 2ll * job->host->counts[1];
    r->iso_value = job->host->iso;
    r->residual_norm = job->host->resnorm;
    for (int a = 0; a < 3; ++a) r->origin[a] = job->host->params[a];
    r->width = job->host->params[3];
  }
  return 0;
  MPR_GUARD_END
}

int mpr_job_fetch_mesh(mpr_job* job, double* vertices, double* vertex_density, int64_t* faces) {
  if (!job) return -1;
  MPR_GUARD_BEGIN
  const uint64_t nv = job->host->counts[0], nf = 2ull * job->host->counts[1];
  if (nv > job->vcap || nf > job->fcap) {
    set_error("mesh exceeds capacity; call mpr_job_wait first");
    return -1;
  }
  if (vertices && nv)
    MPR_HIP(hipMemcpyAsync(vertices, job->verts, 24 * nv, hipMemcpyDeviceToHost, job->stream));
  if (vertex_density && nv)
    MPR_HIP(hipMemcpyAsync(vertex_density, job->vdens, 8 * nv, hipMemcpyDeviceToHost, job->stream));
  if (faces && nf)
    MPR_HIP(hipMemcpyAsync(faces, job->faces, 24 * nf, hipMemcpyDeviceToHost, job->stream));
  MPR_HIP(hipStreamSynchronize(job->stream));
  return 0;
  MPR_GUARD_END
}

int mpr_job_fetch_grids(mpr_job* job, double* field, double* density, double* screen, double* rhs) {
  if (!job) return -1;
  if (rhs && !job->grids_kept) {
    set_error("rhs was not kept: run with MPR_STAGE_GRIDS");
    return -1;
  }
  MPR_GUARD_BEGIN
  const size_t b = 8ull * job->n * job->n * job->n;
  if (field) MPR_HIP(hipMemcpyAsync(field, job->field, b, hipMemcpyDeviceToHost, job->stream));
  if (density) MPR_HIP(hipMemcpyAsync(density, job->density, b, hipMemcpyDeviceToHost, job->stream));
  if (screen) MPR_HIP(hipMemcpyAsync(screen, job->screen, b, hipMemcpyDeviceToHost, job->stream));
  if (rhs) MPR_HIP(hipMemcpyAsync(rhs, job->rhs, b, hipMemcpyDeviceToHost, job->stream));
  MPR_HIP(hipStreamSynchronize(job->stream));
  return 0;
  MPR_GUARD_END
}

int mpr_job_build_octree(mpr_job* job, int32_t depth, int64_t* node_count) {
  if (!job || depth < 0 || depth > 19) {
    set_error("octree depth must be between 0 and 19");
    return -1;
  }
  MPR_GUARD_BEGIN
  hipStream_t s = job->stream;
  const uint32_t P = job->P;
  const double* upts = (job->ran && (job->last.stages & MPR_STAGE_NORMALIZE)) ? job->unit : job->pts;
  job->oct.release();
  DevArena& a = job->oct;
  if (P == 0) {
    job->oct_nodes = 1;
    job->oct_depth = depth;
    if (node_count) *node_count = 1;
    return 0;
  }
  uint64_t* keys = a.alloc<uint64_t>(P);
  uint64_t* ktmp = a.alloc<uint64_t>(P);
  uint32_t* perm = a.alloc<uint32_t>(P);
  uint32_t* vtmp = a.alloc<uint32_t>(P);
  const size_t rw = radix_scratch_words(P);
  uint32_t* hist = a.alloc<uint32_t>(rw);
  uint32_t* offs = a.alloc<uint32_t>(rw);
  launch_leaf_keys(upts, P, depth, keys, perm, s);
  launch_radix_sort(keys, perm, ktmp, vtmp, P, 3 * depth, hist, offs, s);
  const int levels = depth + 1;
  const uint32_t tiles = compaction_tiles(P);
  uint32_t* tcount = a.alloc<uint32_t>((size_t)tiles * levels);
  uint32_t* toffs = a.alloc<uint32_t>((size_t)tiles * levels);
  uint32_t* lcount = a.alloc<uint32_t>(levels);
  int64_t* offsets = a.alloc<int64_t>(levels + 1);
  for (int d = 0; d < levels; ++d)
    launch_level_counts(keys, P, d, depth, tcount + (size_t)d * tiles, toffs + (size_t)d * tiles,
                        lcount + d, s);
  launch_level_offsets(lcount, levels, offsets, s);
  int64_t total = 0;
  MPR_HIP(hipMemcpyAsync(&total, offsets + levels, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  MPR_HIP(hipStreamSynchronize(s));
  job->o_depth = a.alloc<int64_t>(total);
  job->o_morton = a.alloc<int64_t>(total);
  job->o_parent = a.alloc<int64_t>(total);
  job->o_cstart = a.alloc<int64_t>(total);
  job->o_ccount = a.alloc<int64_t>(total);
  job->o_leaf = a.alloc<int64_t>(P);
  launch_fill<int64_t>(job->o_cstart, -1, total, s);
  launch_fill<int64_t>(job->o_ccount, 0, total, s);
  uint32_t* ua = a.alloc<uint32_t>(P);
  uint32_t* ub = a.alloc<uint32_t>(P);
  for (int d = 0; d < levels; ++d) {
    launch_level_emit(keys, P, d, depth, offsets, toffs + (size_t)d * tiles, ua, ub, job->o_morton,
                      job->o_depth, job->o_parent, job->o_cstart, job->o_ccount, s);
    std::swap(ua, ub);
  }
  launch_point_leaf(perm, ua, P, offsets, depth, job->o_leaf, s);
  MPR_HIP(hipStreamSynchronize(s));
  job->oct_nodes = total;
  job->oct_depth = depth;
  if (node_count) *node_count = total;
  return 0;
  MPR_GUARD_END
}

int mpr_job_fetch_octree(mpr_job* job, int64_t* depth, int64_t* morton, int64_t* parent,
                         int64_t* child_start, int64_t* child_count, int64_t* point_leaf) {
  if (!job || job->oct_depth < 0) {
    set_error("no octree built");
    return -1;
  }
  MPR_GUARD_BEGIN
  if (job->P == 0) {  // core.py:94-97
    if (depth) depth[0] = 0;
    if (morton) morton[0] = 0;
    if (parent) parent[0] = -1;
    if (child_start) child_start[0] = -1;
    if (child_count) child_count[0] = 0;
    return 0;
  }
  const size_t b = 8ull * job->oct_nodes;
  hipStream_t s = job->stream;
  if (depth) MPR_HIP(hipMemcpyAsync(depth, job->o_depth, b, hipMemcpyDeviceToHost, s));
  if (morton) MPR_HIP(hipMemcpyAsync(morton, job->o_morton, b, hipMemcpyDeviceToHost, s));
  if (parent) MPR_HIP(hipMemcpyAsync(parent, job->o_parent, b, hipMemcpyDeviceToHost, s));
  if (child_start) MPR_HIP(hipMemcpyAsync(child_start, job->o_cstart, b, hipMemcpyDeviceToHost, s));
  if (child_count) MPR_HIP(hipMemcpyAsync(child_count, job->o_ccount, b, hipMemcpyDeviceToHost, s));
  if (point_leaf)
    MPR_HIP(hipMemcpyAsync(point_leaf, job->o_leaf, 8ull * job->P, hipMemcpyDeviceToHost, s));
  MPR_HIP(hipStreamSynchronize(s));
  return 0;
  MPR_GUARD_END
}

int mpr_job_time_kernel(mpr_job* job, int32_t kernel, int32_t reps, double* ms_per_launch,
                        double* bytes_per_launch) {
  if (!job || reps < 1 || kernel < 0 || kernel > 6) {
    set_error("invalid kernel timing request");
    return -1;
  }
  if ((kernel == 1 || kernel == 5 || kernel == 6) && job->L32.empty()) {
    set_error("FP32 kernels need a mixed-precision job");
    return -1;
  }
  MPR_GUARD_BEGIN
  hipStream_t s = job->stream;
  const Geom& g0 = job->geoms[0];
  const Geom& g1 = job->geoms[1];
  const double M0 = (double)g0.n * g0.n * g0.n, M1 = (double)g1.n * g1.n * g1.n;
  const double pw = job->ran ? job->last.point_weight : 4.0;
  auto& a = job->L64[0];
  double bytes = 0;
  hipEvent_t e0, e1;
  MPR_HIP(hipEventCreate(&e0));
  MPR_HIP(hipEventCreate(&e1));
  MPR_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) {
    switch (kernel) {
      case 0:  // one colour: read other colour x, this colour rhs + screen, write x
        launch_relax_color<double>(a.x, a.rhs, a.screen, g0, pw, r & 1, s);
        bytes = 4.0 * 8.0 * M0 / 2.0;
        break;
      case 1: {
        auto& f = job->L32[0];
        launch_relax_color<float>(f.x, f.rhs, f.screen, g0, pw, r & 1, s);
        bytes = 4.0 * 4.0 * M0 / 2.0;
        break;
      }
      case 2:
        launch_residual<double, double>(a.x, a.rhs, a.screen, a.res, g0, pw, nullptr, s);
        bytes = 4.0 * 8.0 * M0;
        break;
      case 3:
        launch_restrict<double>(a.res, g0, job->L64[1].rhs, g1, 1.0, s);
        bytes = 8.0 * (M0 + M1);
        break;
      case 4:
        launch_prolong_add<double>(job->L64[1].x, g1, a.x, g0, s);
        bytes = 8.0 * (M1 + 2.0 * M0);
        break;
      case 5: {
        auto& f = job->L32[0];
        launch_residual<float, float>(f.x, f.rhs, f.screen, f.res, g0, pw, nullptr, s);
        bytes = 4.0 * 4.0 * M0;
        break;
      }
      case 6:
        launch_residual<double, float>(a.x, a.rhs, a.screen, job->L32[0].rhs, g0, pw, nullptr, s);
        bytes = (3.0 * 8.0 + 4.0) * M0;
        break;
    }
  }
  MPR_HIP(hipEventRecord(e1, s));
  MPR_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MPR_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_per_launch) *ms_per_launch = (double)ms / reps;
  if (bytes_per_launch) *bytes_per_launch = bytes;
  return 0;
  MPR_GUARD_END
}

int mpr_extract(const double* field, const double* density, int64_t n, double iso, mpr_mesh** out,
                int64_t* vertex_count, int64_t* face_count) {
  if (!out || !field || n < 2 || n > 1025) {
    set_error("field must be a cubic grid with 2 <= n <= 1025");
    return -1;
  }
  *out = nullptr;
  if (ensure_device()) return -1;
  mpr_mesh* m = new mpr_mesh();
  try {
    m->n = (int)n;
    MPR_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    hipStream_t s = m->stream;
    DevArena scratch;
    const uint64_t M = (uint64_t)n * n * n, cells = (uint64_t)(n - 1);
    double* f = scratch.alloc<double>(M);
    double* d = scratch.alloc<double>(M);
    double* isod = scratch.alloc<double>(1);
    uint32_t* ctc = scratch.alloc<uint32_t>(cell_tiles((int)n));
    uint32_t* cto = scratch.alloc<uint32_t>(cell_tiles((int)n));
    uint32_t* etc_ = scratch.alloc<uint32_t>(edge_tiles((int)n));
    uint32_t* eto = scratch.alloc<uint32_t>(edge_tiles((int)n));
    uint32_t* counts = scratch.alloc<uint32_t>(2);
    int32_t* cm = scratch.alloc<int32_t>(cells * cells * cells);
    MPR_HIP(hipMemcpyAsync(f, field, 8 * M, hipMemcpyHostToDevice, s));
    if (density)
      MPR_HIP(hipMemcpyAsync(d, density, 8 * M, hipMemcpyHostToDevice, s));
    else
      launch_fill<double>(d, 0.0, M, s);
    MPR_HIP(hipMemcpyAsync(isod, &iso, 8, hipMemcpyHostToDevice, s));
    launch_cells_count(f, (int)n, isod, ctc, s);
    launch_scan_exclusive(ctc, cto, cell_tiles((int)n), counts, s);
    uint32_t hc[2] = {0, 0};
    MPR_HIP(hipMemcpyAsync(hc, counts, 4, hipMemcpyDeviceToHost, s));
    MPR_HIP(hipStreamSynchronize(s));
    m->nv = hc[0];
    m->verts = m->arena.alloc<double>(3 * std::max<uint64_t>(1, m->nv));
    m->vdens = m->arena.alloc<double>(std::max<uint64_t>(1, m->nv));
    launch_cells_emit<int32_t>(f, d, (int)n, isod, cto, cm, m->verts, m->vdens, m->nv, nullptr, s);
    launch_edges_count<int32_t>(f, (int)n, isod, cm, etc_, s);
    launch_scan_exclusive(etc_, eto, edge_tiles((int)n), counts + 1, s);
    MPR_HIP(hipMemcpyAsync(hc + 1, counts + 1, 4, hipMemcpyDeviceToHost, s));
    MPR_HIP(hipStreamSynchronize(s));
    m->nq = hc[1];
    m->faces = m->arena.alloc<int64_t>(6 * std::max<uint64_t>(1, m->nq));
    launch_edges_emit<int32_t>(f, (int)n, isod, cm, eto, m->faces, 2 * m->nq, s);
    MPR_HIP(hipStreamSynchronize(s));
    scratch.release();
    if (vertex_count) *vertex_count = (int64_t)m->nv;
    if (face_count) *face_count = 2 * (int64_t)m->nq;
    *out = m;
    return 0;
  } catch (const HipError& e) {
    set_error(e.msg);
    delete m;
    return -2;
  }
}

int mpr_mesh_fetch(mpr_mesh* m, double* vertices, double* vertex_density, int64_t* faces) {
  if (!m) return -1;
  MPR_GUARD_BEGIN
  if (vertices && m->nv)
    MPR_HIP(hipMemcpyAsync(vertices, m->verts, 24 * m->nv, hipMemcpyDeviceToHost, m->stream));
  if (vertex_density && m->nv)
    MPR_HIP(hipMemcpyAsync(vertex_density, m->vdens, 8 * m->nv, hipMemcpyDeviceToHost, m->stream));
  if (faces && m->nq)
    MPR_HIP(hipMemcpyAsync(faces, m->faces, 48 * m->nq, hipMemcpyDeviceToHost, m->stream));
  MPR_HIP(hipStreamSynchronize(m->stream));
  return 0;
  MPR_GUARD_END
}

int mpr_mesh_destroy(mpr_mesh* m) {
  delete m;
  return 0;
}

}  // extern "C"
