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
#include <map>
#include <mutex>
#include <vector>

#include "mpr_hip.h"
#include "mpr_batch.h"
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

// Device buffers of one owner. Each buffer takes its size rounded up to 256 bytes, and a
// buffer of 1 MiB or more starts on a 2 MiB boundary (as a separate allocation of that size
// would). After reserve(total) the buffers are carved in order out of one allocation of
// `total` bytes (the sum, padding included, that a dry run of the same requests counts), so
// the owner's HBM footprint is that sum; otherwise each buffer is its own allocation.
struct DevArena {
  static constexpr size_t kBig = size_t(1) << 20, kPage = size_t(2) << 20;
  std::vector<void*> ptrs;
  size_t bytes = 0;
  bool dry = false;  // footprint accounting only
  char* base = nullptr;
  size_t cap = 0;
  // one allocation of `total` bytes, which must include kPage for aligning its start
  void reserve(size_t total) {
    char* p = nullptr;
    MPR_HIP(hipMalloc((void**)&p, total));
    ptrs.push_back(p);
    base = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(p) + kPage - 1) & ~(kPage - 1));
    cap = total - (size_t)(base - p);
  }
  template <class T>
  T* alloc(size_t count) {
    const size_t b = std::max<size_t>(256, ((count * sizeof(T)) + 255) & ~size_t(255));
    if (b >= kBig) bytes = (bytes + kPage - 1) & ~(kPage - 1);
    const size_t off = bytes;
    bytes += b;
    if (dry) return nullptr;
    if (base) {
      if (bytes > cap) throw HipError{"device arena smaller than its buffers"};
      return reinterpret_cast<T*>(base + off);
    }
    void* p = nullptr;
    MPR_HIP(hipMalloc(&p, b));
    ptrs.push_back(p);
    return static_cast<T*>(p);
  }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
    bytes = 0;
    base = nullptr;
    cap = 0;
  }
  DevArena() = default;
  DevArena(const DevArena&) = delete;
  DevArena& operator=(const DevArena&) = delete;
  ~DevArena() { release(); }
};

// the scalars a run reports, laid out alike on the device and in pinned host memory so that
// one copy reads them back
struct HostScalars {
  unsigned long long accepted;
  uint32_t counts[2];  // active cells, quads
  double iso;
  double resnorm;
  double params[4];  // origin xyz, width
  uint32_t valid;    // rows kept by MPR_STAGE_FILTER
  uint32_t pad;
};

}  // namespace

struct Job {
  int depth = 0, n = 0, precision = 0;
  uint32_t max_points = 0, P = 0;
  hipStream_t stream = nullptr;
  hipEvent_t queued = nullptr;  // recorded on stream when a batch takes the job's work over
  DevArena arena;
  std::vector<Geom> geoms;
  // inputs; with MPR_STAGE_FILTER the kept rows' normals are compacted into nrm_kept
  double *pts = nullptr, *nrm = nullptr, *unit = nullptr, *nrm_kept = nullptr;
  FilterScratch filt;
  // natural-layout grids
  double *density = nullptr, *screen = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr,
         *rhs = nullptr, *field = nullptr;
  std::vector<LevelPtrs<double>> L64;
  std::vector<LevelPtrs<float>> L32;
  // this job's own V-cycle schedule (a batch of one) and defect-correction tables
  BatchedVcycle<double> vc64;
  BatchedVcycle<float> vc32;
  DefectTables dtab;
  // scalars (fields of one device HostScalars block) and partial sums
  HostScalars* d_scalars = nullptr;
  unsigned long long* d_accepted = nullptr;
  double *d_iso = nullptr, *d_resnorm = nullptr, *d_params = nullptr;
  uint32_t* d_counts = nullptr;
  double *norm_part = nullptr, *bbox_part = nullptr, *iso_part = nullptr;
  int norm_nparts = 0;
  // extraction
  BitSurface surf;
  SplatScratch splat;
  uint32_t* touched = nullptr;  // splat node bits, ceil(n^3 / 32) words
  // the previous assembly's touched bits and the sparse assembly's need mask (same size)
  uint32_t *prev_touched = nullptr, *need_bits = nullptr;
  uint32_t* support_bits = nullptr;  // coarse support of need_bits (the first pyramid step)
  uint32_t *chunk_list = nullptr, *chunk_count = nullptr;  // the sparse assembly's words
  // live slots of the finest level (LiveMask, stencil.h): rhs or screen non-zero
  uint32_t* live_bits = nullptr;
  int* x_initial = nullptr;  // the FP64 solution was not filled with kInitialX (TopPtrs)
  // how the density grid holds the last solve's values: raw splat sums read through
  // `touched` and divided by dens_spn, or normalised in place (runs with MPR_STAGE_GRIDS)
  bool dens_raw = false;
  double dens_spn = 1.0;
  double *verts = nullptr, *vdens = nullptr;
  int64_t* faces = nullptr;
  uint64_t vcap = 0, fcap = 0;
  std::vector<void*> mesh_bufs;  // grown extraction outputs (owned separately, latest set only)
  // host mesh: each extraction also copies the mesh into these pinned host buffers on the
  // job's stream (capacities vcap / fcap), so fetching it is a host copy
  bool host_mesh = false;
  double *h_verts = nullptr, *h_vdens = nullptr;
  int64_t* h_faces = nullptr;
  void free_host_mesh() {
    for (void* p : {(void*)h_verts, (void*)h_vdens, (void*)h_faces})
      if (p) (void)hipHostFree(p);
    h_verts = h_vdens = nullptr;
    h_faces = nullptr;
  }
  void alloc_host_mesh() {
    free_host_mesh();
    MPR_HIP(hipHostMalloc((void**)&h_verts, 24 * std::max<uint64_t>(1, vcap), hipHostMallocDefault));
    MPR_HIP(hipHostMalloc((void**)&h_vdens, 8 * std::max<uint64_t>(1, vcap), hipHostMallocDefault));
    MPR_HIP(hipHostMalloc((void**)&h_faces, 24 * std::max<uint64_t>(1, fcap), hipHostMallocDefault));
  }
  // host results
  HostScalars* host = nullptr;
  mpr_params last{}, captured{};
  bool has_capture = false;  // `captured` was recorded (the graph may since have been dropped)
  bool ran = false, grids_kept = false;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // octree (allocated on demand)
  DevArena oct;
  int64_t oct_nodes = 0;
  uint32_t oct_points = 0;  // points the tree was built over (point_leaf entries)
  int oct_depth = -1;
  int64_t *o_depth = nullptr, *o_morton = nullptr, *o_parent = nullptr, *o_cstart = nullptr,
          *o_ccount = nullptr, *o_leaf = nullptr;

  void plan(DevArena& a);
  // the finest level's live slots cover the assembled rhs and screen (FP64) and the FP32
  // screen; the mixed V-cycle's FP32 rhs is the (dense) defect
  InstanceLevels<double> levels64() const { return {depth, L64, live_bits, true}; }
  InstanceLevels<float> levels32() const {
    return {depth, L32, live_bits, false, L64[0].x, x_initial};
  }
  TopPtrs top_ptrs() const {
    return {L64[0].x, L64[0].rhs, L64[0].screen, L32.empty() ? nullptr : L32[0].rhs,
            L32.empty() ? nullptr : L32[0].x, live_bits, x_initial};
  }
  uint64_t live_words() const { return (2ull * geoms[0].C + 31) / 32; }
  const double* unit_points(const mpr_params& p) const {
    return (p.stages & (MPR_STAGE_NORMALIZE | MPR_STAGE_FILTER)) ? unit : pts;
  }
  const double* normals(const mpr_params& p) const {
    return (p.stages & MPR_STAGE_FILTER) ? nrm_kept : nrm;
  }
  // rows the splat and the iso value see: all P, of which the first host->valid are kept
  // (the rest zero rows) with MPR_STAGE_FILTER
  const uint32_t* kept_count(const mpr_params& p) const {
    return (p.stages & MPR_STAGE_FILTER) ? &d_scalars->valid : nullptr;
  }
  // pipeline stages; all enqueue on `s` without host synchronisation
  void setup(const mpr_params& p, hipStream_t s);
  // the setups of several jobs, one launch per step for all of them (multi.h); each job's
  // results are those of its own setup
  static void setup_all(Job* const* jobs, const mpr_params* const* params, size_t count,
                        hipStream_t s);
  // finish, extraction and readback of several jobs, one launch per step for all of them
  static void finish_all(Job* const* jobs, const mpr_params* const* params, size_t count,
                         hipStream_t s);
  void enqueue_extract(bool world, hipStream_t s);
  // the parameters of the work last enqueued or launched on this job (eagerly or as a graph)
  void adopt(const mpr_params& p) {
    last = p;
    grids_kept = (p.stages & MPR_STAGE_GRIDS) != 0;
    if (p.stages & MPR_STAGE_SOLVE) {
      dens_raw = !grids_kept;
      dens_spn = p.samples_per_node;
    }
  }
  struct State {
    mpr_params last;
    bool grids_kept, dens_raw;
    double dens_spn;
  };
  State state() const { return {last, grids_kept, dens_raw, dens_spn}; }
  void restore(const State& st) {
    last = st.last;
    grids_kept = st.grids_kept;
    dens_raw = st.dens_raw;
    dens_spn = st.dens_spn;
  }
  DensityView density_view() const {
    return dens_raw ? DensityView{density, touched, dens_spn}
                    : DensityView{density, nullptr, 1.0};
  }
  void readback(hipStream_t s);
  void enqueue(const mpr_params& p);
  // after the stream finished: if the mesh outgrew its buffers, grow them to the exact
  // counts and re-run the extraction on `s` (synchronously)
  void resolve_overflow(hipStream_t s);
  void result(mpr_result* r) const;
  uint64_t mesh_generation = 0;  // bumped when the mesh buffers move (graphs must re-record)
  ~Job();
};

void run_cycles(const BatchedVcycle<double>& v64, const BatchedVcycle<float>* v32,
                const DefectTables* dt, const mpr_params& p, hipStream_t s);

void Job::resolve_overflow(hipStream_t s) {
  if (!ran || !(last.stages & MPR_STAGE_EXTRACT)) return;
  const uint64_t nv = host->counts[0], nf = 2ull * host->counts[1];
  if (nv <= vcap && nf <= fcap) return;
  const uint64_t nvc = std::max<uint64_t>(nv, vcap), nfc = std::max<uint64_t>(nf, fcap);
  // the stream is idle (the caller waited for it) and the job's graph is dropped below, so
  // the previously grown set is no longer referenced (a batch graph holding it records again
  // when mesh_generation changes): free it, keeping only the latest set
  for (void* p : mesh_bufs) (void)hipFree(p);
  mesh_bufs.clear();
  void *v = nullptr, *d = nullptr, *f = nullptr, *c = nullptr;
  MPR_HIP(hipMalloc(&v, 24 * nvc));
  mesh_bufs.push_back(v);
  MPR_HIP(hipMalloc(&d, 8 * nvc));
  mesh_bufs.push_back(d);
  MPR_HIP(hipMalloc(&f, 24 * nfc));
  mesh_bufs.push_back(f);
  MPR_HIP(hipMalloc(&c, 4 * nvc));
  mesh_bufs.push_back(c);
  verts = (double*)v;
  vdens = (double*)d;
  faces = (int64_t*)f;
  surf.vcell = (uint32_t*)c;
  surf.vcell_cap = nvc;
  vcap = nvc;
  fcap = nfc;
  if (host_mesh) alloc_host_mesh();
  ++mesh_generation;
  if (exec) {  // the captured graph points at the old buffers
    MPR_HIP(hipGraphExecDestroy(exec));
    MPR_HIP(hipGraphDestroy(graph));
    exec = nullptr;
    graph = nullptr;
  }
  enqueue_extract((last.stages & MPR_STAGE_WORLD) != 0, s);
  MPR_HIP(hipMemcpyAsync(host->counts, d_counts, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MPR_HIP(hipStreamSynchronize(s));
}

void Job::result(mpr_result* r) const {
  if (!r) return;
  r->accepted = (int64_t)host->accepted;
  r->vertex_count = host->counts[0];
  r->face_count = 2ll * host->counts[1];
  r->iso_value = host->iso;
  r->residual_norm = host->resnorm;
  for (int a = 0; a < 3; ++a) r->origin[a] = host->params[a];
  r->width = host->params[3];
  r->valid_rows = (last.stages & MPR_STAGE_FILTER) ? (int64_t)host->valid : (int64_t)P;
}

// A fixed set of jobs whose V-cycles run together: one launch per V-cycle step and grid
// size class covers every job (mpr_batch in include/mpr_hip.h).
// Streams and table buffers reused across batches: forming and closing a batch in a
// serving loop then allocates and frees no device memory after warm-up (a device free can
// wait for work queued on other streams).
class BatchResources {
 public:
  static BatchResources& get() {
    static BatchResources r;
    return r;
  }
  hipStream_t take_stream() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      if (!streams_.empty()) {
        hipStream_t s = streams_.back();
        streams_.pop_back();
        return s;
      }
    }
    hipStream_t s = nullptr;
    MPR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
  }
  void give_stream(hipStream_t s) {
    std::lock_guard<std::mutex> lock(mu_);
    streams_.push_back(s);
  }
  // a buffer of at least `bytes` (capacity a power of two, at least 64 KiB)
  void* take_buffer(size_t bytes, size_t& cap) {
    cap = 64 * 1024;
    while (cap < bytes) cap <<= 1;
    {
      std::lock_guard<std::mutex> lock(mu_);
      auto it = buffers_.find(cap);
      if (it != buffers_.end()) {
        void* p = it->second;
        buffers_.erase(it);
        return p;
      }
    }
    void* p = nullptr;
    MPR_HIP(hipMalloc(&p, cap));
    return p;
  }
  void give_buffer(void* p, size_t cap) {
    std::lock_guard<std::mutex> lock(mu_);
    buffers_.emplace(cap, p);
  }
  // pinned host staging of `cap` bytes (a capacity take_buffer returned)
  unsigned char* take_stage(size_t cap) {
    {
      std::lock_guard<std::mutex> lock(mu_);
      auto it = stages_.find(cap);
      if (it != stages_.end()) {
        unsigned char* p = it->second;
        stages_.erase(it);
        return p;
      }
    }
    void* p = nullptr;
    MPR_HIP(hipHostMalloc(&p, cap, hipHostMallocDefault));
    return static_cast<unsigned char*>(p);
  }
  void give_stage(unsigned char* p, size_t cap) {
    std::lock_guard<std::mutex> lock(mu_);
    stages_.emplace(cap, p);
  }

 private:
  std::mutex mu_;
  std::vector<hipStream_t> streams_;
  std::multimap<size_t, void*> buffers_;
  std::multimap<size_t, unsigned char*> stages_;
};

// A batch runs everything on its own stream: the jobs' own stages (normalise .. pyramid,
// and residual norm .. readback, chains of mostly small kernels) as multi-instance launches
// that take up to 8 jobs each (Job::setup_all, Job::finish_all), the V-cycles one launch per
// step for all jobs.
struct Batch {
  std::vector<Job*> jobs;
  int precision = 0;
  hipStream_t stream = nullptr;
  void* tables = nullptr;  // pooled (BatchResources)
  unsigned char* stage = nullptr;  // its pinned host staging, same capacity
  size_t tables_cap = 0;
  BatchedVcycle<double> vc64;
  BatchedVcycle<float> vc32;
  DefectTables dtab;
  std::vector<mpr_params> last;
  std::vector<uint64_t> generations;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void build();
  void enqueue(const mpr_params* p);
  // the batch's stream waits (on the device) for the work queued so far on its jobs' own
  // streams, such as their point uploads
  void after_job_streams() {
    for (Job* j : jobs) {
      MPR_HIP(hipEventRecord(j->queued, j->stream));
      MPR_HIP(hipStreamWaitEvent(stream, j->queued, 0));
    }
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~Batch() {
    drop_graph();
    if (tables) BatchResources::get().give_buffer(tables, tables_cap);
    if (stage) BatchResources::get().give_stage(stage, tables_cap);
    if (stream) BatchResources::get().give_stream(stream);
  }
};

void Batch::build() {
  std::vector<InstanceLevels<double>> i64;
  std::vector<InstanceLevels<float>> i32;
  std::vector<int> depths;
  std::vector<TopPtrs> tops;
  for (Job* j : jobs) {
    i64.push_back(j->levels64());
    if (precision == MPR_PRECISION_MIXED) {
      i32.push_back(j->levels32());
      depths.push_back(j->depth);
      tops.push_back(j->top_ptrs());
    }
  }
  const bool mixed = precision == MPR_PRECISION_MIXED;
  auto round = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t b64 = round(BatchedVcycle<double>::table_bytes(i64));
  const size_t b32 = mixed ? round(BatchedVcycle<float>::table_bytes(i32)) : 0;
  const size_t bd = mixed ? round(DefectTables::table_bytes(tops.size())) : 0;
  tables = BatchResources::get().take_buffer(b64 + b32 + bd, tables_cap);
  stage = BatchResources::get().take_stage(tables_cap);
  unsigned char* t = static_cast<unsigned char*>(tables);
  vc64.build(i64, t, stage);
  if (mixed) {
    vc32.build(i32, t + b64, stage + b64);
    dtab.build(depths, tops, t + b64 + b32, stage + b64 + b32);
  }
  // one asynchronous upload on the batch's stream (a synchronous copy would wait behind the
  // work of other batches in flight); the staging stays the batch's until it is destroyed
  MPR_HIP(hipMemcpyAsync(tables, stage, b64 + b32 + bd, hipMemcpyHostToDevice, stream));
}

void Batch::enqueue(const mpr_params* p) {
  {
    std::vector<const mpr_params*> pp(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) pp[i] = &p[i];
    Job::setup_all(jobs.data(), pp.data(), jobs.size(), stream);
  }
  if (p[0].stages & MPR_STAGE_SOLVE)
    run_cycles(vc64, precision == MPR_PRECISION_MIXED ? &vc32 : nullptr, &dtab, p[0], stream);
  {
    std::vector<const mpr_params*> pp(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) pp[i] = &p[i];
    Job::finish_all(jobs.data(), pp.data(), jobs.size(), stream);
  }
}

// V-cycles of a set of jobs (core.py:274-285): FP64 as the reference, or FP32 V-cycles on
// the correction equation inside an FP64 defect-correction loop, inner_cycles FP32 V-cycles
// per FP64 defect.
void run_cycles(const BatchedVcycle<double>& v64, const BatchedVcycle<float>* v32,
                const DefectTables* dt, const mpr_params& p, hipStream_t s) {
  if (!v32) {
    for (int c = 0; c < p.cycles; ++c)
      v64.run(p.point_weight, p.pre_smooth, p.post_smooth, false, s);
    return;
  }
  const int inner = p.inner_cycles > 0 ? p.inner_cycles : 1;
  for (int c = 0; c < p.cycles; c += inner) {
    const int k = std::min(inner, p.cycles - c);
    // r = b - A x (FP64), stored as FP32; with pre-smoothing from e = 0 (point_weight >= 0)
    // the same pass writes the top level's first colour-0 sweep, e = r / diagonal
    const bool e0 = p.pre_smooth > 0 && p.point_weight >= 0.0;
    dt->defect(p.point_weight, s, c == 0, e0);
    // e = V32(r) from e = 0 (k V-cycles, the later ones from the current e), and x += e: with
    // post-smoothing the top level's last sweep adds e into x for instances deeper than the
    // coarse classes, the rest use kb_correct
    const bool fuse = p.post_smooth > 0;
    for (int i = 0; i < k; ++i)
      v32->run(p.point_weight, p.pre_smooth, p.post_smooth, i == 0, s, fuse && i == k - 1,
               c == 0, e0 && i == 0);
    dt->correct(s, fuse ? v32->unfused_class() : 1 << 30);
  }
}

void Job::plan(DevArena& a) {
  n = (1 << depth) + 1;
  const uint64_t M = (uint64_t)n * n * n;
  const uint64_t P3 = 3ull * std::max<uint32_t>(1u, max_points);
  pts = a.alloc<double>(P3);
  nrm = a.alloc<double>(P3);
  unit = a.alloc<double>(P3);
  nrm_kept = a.alloc<double>(P3);
  {
    const uint32_t fb = filter_blocks(std::max<uint32_t>(1u, max_points));
    filt.counts = a.alloc<uint32_t>(fb);
    filt.offsets = a.alloc<uint32_t>(fb);
    filt.scan = a.alloc<uint32_t>(scan_scratch_words(fb));
    filt.part = a.alloc<double>(6ull * fb);
  }
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
      // the fused smoothing's second x buffer (march.hip) on the levels it can run on
      f.xalt = march_supported(g, 1, 4) ? a.alloc<float>(w) : nullptr;
      L32.push_back(f);
    }
  }
  d_scalars = a.alloc<HostScalars>(1);
  if (d_scalars) {
    d_accepted = &d_scalars->accepted;
    d_counts = d_scalars->counts;
    d_iso = &d_scalars->iso;
    d_resnorm = &d_scalars->resnorm;
    d_params = d_scalars->params;
  }
  norm_nparts = residual_partial_count(geoms[0], sizeof(double));
  norm_part = a.alloc<double>(norm_nparts);
  bbox_part = a.alloc<double>(6ull * bbox_partial_count(std::max<uint32_t>(1u, max_points)));
  iso_part = a.alloc<double>(2ull * iso_partial_count(std::max<uint32_t>(1u, max_points)));
  touched = a.alloc<uint32_t>((M + 31) / 32);
  prev_touched = a.alloc<uint32_t>((M + 31) / 32);
  need_bits = a.alloc<uint32_t>((M + 31) / 32);
  support_bits = geoms.size() > 1 ? a.alloc<uint32_t>(coarse_support_words(geoms[1].n)) : nullptr;
  chunk_list = a.alloc<uint32_t>((M + 31) / 32);
  chunk_count = a.alloc<uint32_t>(1);
  live_bits = a.alloc<uint32_t>(live_words());
  x_initial = a.alloc<int>(1);
  splat = splat_scratch_sizes(std::max<uint32_t>(1u, max_points));
  splat.keys = a.alloc<uint32_t>(splat.pairs);
  splat.ktmp = a.alloc<uint32_t>(splat.pairs);
  splat.ids = a.alloc<uint32_t>(splat.pairs);
  splat.itmp = a.alloc<uint32_t>(splat.pairs);
  splat.hist = a.alloc<uint32_t>(splat.radix_words);
  splat.offs = a.alloc<uint32_t>(splat.radix_words);
  splat.w1 = a.alloc<double>(splat.pairs);
  splat.c2 = a.alloc<double>(3ull * splat.pairs);
  surf = bit_surface_sizes(n);
  surf.bits = a.alloc<uint32_t>(surf.bit_words);
  surf.cmask = a.alloc<uint32_t>(surf.runs);
  surf.coff = a.alloc<uint32_t>(surf.runs);
  surf.cell_tc = a.alloc<uint32_t>(surf.cell_tiles);
  surf.cell_to = a.alloc<uint32_t>(surf.cell_tiles);
  surf.edge_tc = a.alloc<uint32_t>(surf.edge_tiles);
  surf.edge_to = a.alloc<uint32_t>(surf.edge_tiles);
  const uint64_t cells = (uint64_t)(n - 1);
  vcap = std::min<uint64_t>(cells * cells * cells, std::max<uint64_t>(1u << 16, 12ull * n * n));
  fcap = 5 * vcap / 2 + 16;
  surf.vcell_cap = vcap;
  surf.vcell = a.alloc<uint32_t>(vcap);
  verts = a.alloc<double>(3 * vcap);
  vdens = a.alloc<double>(vcap);
  faces = a.alloc<int64_t>(3 * fcap);
  // single-instance V-cycle tables
  const std::vector<InstanceLevels<double>> i64{levels64()};
  void* t64 = a.alloc<unsigned char>(BatchedVcycle<double>::table_bytes(i64));
  void* t32 = nullptr;
  void* td = nullptr;
  std::vector<InstanceLevels<float>> i32;
  if (precision == MPR_PRECISION_MIXED) {
    i32.push_back(levels32());
    t32 = a.alloc<unsigned char>(BatchedVcycle<float>::table_bytes(i32));
    td = a.alloc<unsigned char>(DefectTables::table_bytes(1));
  }
  if (!a.dry) {
    vc64.build(i64, t64);
    if (precision == MPR_PRECISION_MIXED) {
      vc32.build(i32, t32);
      dtab.build({depth}, {top_ptrs()}, td);
    }
  }
}

Job::~Job() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (void* p : mesh_bufs) (void)hipFree(p);
  free_host_mesh();
  oct.release();
  arena.release();
  if (host) (void)hipHostFree(host);
  if (queued) (void)hipEventDestroy(queued);
  if (stream) (void)hipStreamDestroy(stream);
}

void Job::enqueue_extract(bool world, hipStream_t s) {
  launch_bit_extract(split_field(L64[0].x, geoms[0]), density_view(), n, d_iso, surf, d_counts,
                     verts, vdens, vcap, faces, fcap, world ? d_params : nullptr, s);
  if (host_mesh)
    launch_mesh_to_host(verts, vdens, faces, d_counts, vcap, fcap, h_verts, h_vdens, h_faces, s);
}

// normalise, splat, assemble, screening pyramid and initial guess (core.py:229-273)
void Job::setup(const mpr_params& p, hipStream_t s) {
  Job* self = this;
  const mpr_params* pp = &p;
  setup_all(&self, &pp, 1, s);
}

void Job::setup_all(Job* const* jobs, const mpr_params* const* params, size_t count,
                    hipStream_t s) {
  std::vector<FilterJob> filter;
  std::vector<Job*> solve;
  std::vector<const mpr_params*> sp;
  for (size_t i = 0; i < count; ++i) {
    Job& j = *jobs[i];
    const mpr_params& p = *params[i];
    if (p.stages & MPR_STAGE_FILTER)
      filter.push_back({j.pts, j.nrm, j.P, p.scale, j.filt, &j.d_scalars->valid, j.d_params,
                        j.unit, j.nrm_kept});
    else if (p.stages & MPR_STAGE_NORMALIZE)
      launch_normalize(j.pts, j.P, p.scale, j.bbox_part, j.d_params, j.unit, s);
    if (p.stages & MPR_STAGE_SOLVE) {
      solve.push_back(&j);
      sp.push_back(&p);
    }
  }
  if (!filter.empty()) launch_filter_normalize_multi(filter.data(), filter.size(), s);
  const size_t m = solve.size();
  if (!m) return;
  // the splat grids are not zero-filled: the splat marks the nodes it writes and the
  // assembly and the extraction read every other node as zero
  std::vector<FillJob<uint32_t>> touched(m), live(m);
  std::vector<FillJob<unsigned long long>> accepted(m);
  std::vector<SplatJob> splat(m);
  std::vector<AssembleJob> assemble(m);
  std::vector<TouchedHistory> hist(m);
  size_t levels = 0;
  for (size_t i = 0; i < m; ++i) {
    Job& j = *solve[i];
    const mpr_params& p = *sp[i];
    const uint64_t M = (uint64_t)j.n * j.n * j.n;
    const bool grids = (p.stages & MPR_STAGE_GRIDS) != 0;
    const bool mixed = j.precision == MPR_PRECISION_MIXED;
    touched[i] = {j.touched, 0u, (M + 31) / 32};
    accepted[i] = {j.d_accepted, 0ull, 1};
    splat[i] = {j.unit_points(p), j.normals(p), j.P, j.n, p.confidence, j.density, j.vx, j.vy,
                j.vz, j.screen, p.samples_per_node, j.d_accepted, j.splat, j.touched};
    // Without MPR_STAGE_GRIDS the assembly writes only the nodes near this run's and the
    // previous run's splat (the rest of the right-hand side and screening weights stay
    // +0.0) and leaves the density raw; with it, every grid is written in natural layout
    // too. The FP32 screening weights of the mixed V-cycle are written beside the FP64
    // ones. The live-slot bits let the finest level's kernels skip loading the +0.0 rhs and
    // screen words away from the splat.
    hist[i] = {j.prev_touched, j.need_bits, grids, j.chunk_list, j.chunk_count};
    live[i] = {j.live_bits, 0u, j.live_words()};
    assemble[i] = {j.vx, j.vy, j.vz, j.screen, grids ? j.rhs : nullptr, j.L64[0].rhs,
                   j.L64[0].screen, grids ? j.density : nullptr, p.samples_per_node, j.geoms[0],
                   p.point_weight, 0.5, j.touched, grids ? j.screen : nullptr,
                   mixed ? j.L32[0].screen : nullptr, &hist[i], j.live_bits};
    levels = std::max(levels, j.L64.size());
  }
  launch_fill_multi(touched.data(), m, s);
  launch_fill_multi(accepted.data(), m, s);
  launch_splat_multi(splat.data(), m, s);
  launch_fill_multi(live.data(), m, s);
  launch_assemble_multi(assemble.data(), m, s);
  for (size_t i = 0; i < m; ++i)
    solve[i]->adopt(*sp[i]);  // the extraction enqueued next reads the density as assembled
  // screening pyramid: level l of every job that has one, in one launch
  std::vector<RestrictJob<double>> pyr(m);
  for (size_t l = 1; l < levels; ++l) {
    for (size_t i = 0; i < m; ++i) {
      Job& j = *solve[i];
      const bool mixed = j.precision == MPR_PRECISION_MIXED;
      // The first restriction reads the finest screen, which is +0.0 outside the need mask
      // of a sparse assembly (the nodes with a touched neighbour now or in the previous run,
      // all nodes on a job's first run); a coarse node whose support misses the mask is
      // +0.0 and was already +0.0 after the previous run, so only the others are written.
      const bool sparse = l == 1 && !(sp[i]->stages & MPR_STAGE_GRIDS);
      if (l < j.L64.size())
        pyr[i] = {j.L64[l - 1].screen, j.geoms[l - 1], j.L64[l].screen, j.geoms[l], 4.0,
                  mixed ? j.L32[l].screen : nullptr, sparse ? j.need_bits : nullptr,
                  j.support_bits};
      else
        pyr[i] = {nullptr, Geom{}, nullptr, Geom{}, 4.0, nullptr, nullptr, nullptr};
    }
    launch_restrict_nodes_multi(pyr.data(), m, s);
  }
  // initial guess 0.5 (core.py:270); no samples: the solution is 0 (core.py:286-287), which
  // the V-cycles then leave at 0 (zero right-hand side and screening). When the first
  // mixed-precision cycle's defect and fused correction cover this job, they read the
  // guess as the constant and write every node of x, so x is not filled.
  std::vector<FillJob<int32_t>> flag(m);
  std::vector<FillJob<double>> x(m);
  for (size_t i = 0; i < m; ++i) {
    Job& j = *solve[i];
    const mpr_params& p = *sp[i];
    const bool mixed = j.precision == MPR_PRECISION_MIXED;
    const bool implicit = mixed && j.P && p.cycles > 0 && p.post_smooth > 0 &&
                          j.depth > mixed_unfused_class();
    flag[i] = {j.x_initial, implicit ? 1 : 0, 1};
    x[i] = {j.L64[0].x, j.P ? kInitialX : 0.0, implicit ? 0 : 2ull * j.geoms[0].C};
  }
  launch_fill_multi(flag.data(), m, s);
  launch_fill_multi(x.data(), m, s);
}

// Final FP64 residual norm, natural-layout field and iso value (core.py:289-303), then the
// extraction and the readback, of several jobs. Only the norm of the final residual is
// reported; the iso value and the extraction read the solution in the solver's layout, so
// the natural field is written for fetching only.
void Job::finish_all(Job* const* jobs, const mpr_params* const* params, size_t count,
                     hipStream_t s) {
  std::vector<NormJob> norm;
  std::vector<IsoJob> iso;
  std::vector<FillJob<double>> iso_zero;
  std::vector<ExtractJob> ext;
  std::vector<MeshCopyJob> copy;
  for (size_t i = 0; i < count; ++i) {
    Job& j = *jobs[i];
    const mpr_params& p = *params[i];
    const bool extract = (p.stages & MPR_STAGE_EXTRACT) != 0;
    bool fused = false;
    if (p.stages & MPR_STAGE_SOLVE) {
      // the iso value first: the residual pass then also sets the extraction's sign bits
      fused = extract && residual_norm_sets_sign_bits(j.geoms[0]);
      norm.push_back({j.L64[0].x, j.L64[0].rhs, j.L64[0].screen, j.geoms[0], p.point_weight,
                      j.norm_part, j.norm_nparts, j.d_resnorm, j.live_bits,
                      fused ? j.d_iso : nullptr, fused ? j.surf.bits : nullptr});
      if (p.stages & MPR_STAGE_GRIDS)
        launch_split_to_nat<double>(j.L64[0].x, j.field, j.geoms[0], s);
      if (j.P)
        iso.push_back({split_field(j.L64[0].x, j.geoms[0]), j.unit_points(p), j.normals(p), j.P,
                       j.n, p.confidence, j.iso_part, j.d_iso, j.kept_count(p)});
      else
        iso_zero.push_back({j.d_iso, 0.0, 1});
    }
    if (extract) {
      ext.push_back({split_field(j.L64[0].x, j.geoms[0]), j.density_view(), j.n, j.d_iso,
                     &j.surf, j.d_counts, j.verts, j.vdens, j.vcap, j.faces, j.fcap,
                     (p.stages & MPR_STAGE_WORLD) ? j.d_params : nullptr, fused});
      if (j.host_mesh)
        copy.push_back({j.verts, j.vdens, j.faces, j.d_counts, j.vcap, j.fcap, j.h_verts,
                        j.h_vdens, j.h_faces});
    }
  }
  if (!iso.empty()) launch_iso_multi(iso.data(), iso.size(), s);
  if (!iso_zero.empty()) launch_fill_multi(iso_zero.data(), iso_zero.size(), s);
  if (!norm.empty()) launch_residual_norm_multi(norm.data(), norm.size(), s);
  if (!ext.empty()) launch_bit_extract_multi(ext.data(), ext.size(), s);
  if (!copy.empty()) launch_mesh_to_host_multi(copy.data(), copy.size(), s);
  for (size_t i = 0; i < count; ++i) jobs[i]->readback(s);
}

void Job::readback(hipStream_t s) {
  MPR_HIP(hipMemcpyAsync(host, d_scalars, sizeof(HostScalars), hipMemcpyDeviceToHost, s));
}

void Job::enqueue(const mpr_params& p) {
  hipStream_t s = stream;
  setup(p, s);
  if ((p.stages & MPR_STAGE_SOLVE) && P)
    run_cycles(vc64, precision == MPR_PRECISION_MIXED ? &vc32 : nullptr, &dtab, p, s);
  Job* self = this;
  const mpr_params* pp = &p;
  finish_all(&self, &pp, 1, s);
}

}  // namespace mpr

struct mpr_job : mpr::Job {};
struct mpr_batch : mpr::Batch {};

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
  // core.py:213-220: finite, point_weight >= 0, samples_per_node > 0
  return p->cycles >= 0 && p->pre_smooth >= 0 && p->post_smooth >= 0 && p->inner_cycles >= 0 &&
         std::isfinite(p->samples_per_node) && p->samples_per_node > 0 &&
         std::isfinite(p->point_weight) && p->point_weight >= 0.0;
}

// A point weight of -0.0 runs as +0.0 (only the sign of zero right-hand-side values could
// differ), so that every assembly gives the same +0.0 at nodes away from the splat.
static mpr_params canonical(const mpr_params& p) {
  mpr_params q = p;
  q.point_weight = p.point_weight + 0.0;
  return q;
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

int mpr_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
  if (ensure_device()) return -1;
  MPR_GUARD_BEGIN
  size_t f = 0, t = 0;
  MPR_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = (int64_t)f;
  if (total_bytes) *total_bytes = (int64_t)t;
  return 0;
  MPR_GUARD_END
}

int64_t mpr_job_footprint(int32_t depth, int64_t max_points, int32_t precision) {
  if (depth < 1 || depth > 10 || max_points < 0 || max_points > kMaxPoints) return -1;
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
  // the job's one allocation: its buffers and room to start them on a 2 MiB boundary
  return (int64_t)(dry.bytes + DevArena::kPage);
}

int mpr_job_create(int32_t depth, int64_t max_points, int32_t precision, mpr_job** out) {
  if (!out) return -1;
  *out = nullptr;
  if (depth < 1 || depth > 10) {
    set_error("depth must be between 1 and 10");
    return -1;
  }
  if (max_points < 0 || max_points > kMaxPoints) {
    set_error("max_points must be between 0 and 2^28");
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
  MPR_HIP(hipEventCreateWithFlags(&j->queued, hipEventDisableTiming));
  MPR_HIP(hipHostMalloc((void**)&j->host, sizeof(HostScalars), hipHostMallocDefault));
  std::memset(j->host, 0, sizeof(HostScalars));
  // one allocation for every buffer of the job: its HBM footprint is mpr_job_footprint
  const int64_t total = mpr_job_footprint(depth, max_points, precision);
  if (total < 0) throw HipError{"job footprint"};
  j->arena.reserve((size_t)total);
  j->plan(j->arena);
  const double ident[4] = {0.0, 0.0, 0.0, 1.0};
  MPR_HIP(hipMemcpyAsync(j->d_params, ident, sizeof(ident), hipMemcpyHostToDevice, j->stream));
  MPR_HIP(hipMemsetAsync(j->d_counts, 0, 2 * sizeof(uint32_t), j->stream));
  MPR_HIP(hipMemsetAsync(j->d_accepted, 0, sizeof(unsigned long long), j->stream));
  // no assembly yet: the first one writes every node
  const uint64_t words = ((uint64_t)j->n * j->n * j->n + 31) / 32;
  launch_fill<uint32_t>(j->prev_touched, 0xffffffffu, words, j->stream);
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

int mpr_job_fetch_filtered(mpr_job* job, int64_t* count, double* unit_points, double* normals) {
  if (!job || !job->ran || !(job->last.stages & MPR_STAGE_FILTER)) {
    set_error("the job's last run had no MPR_STAGE_FILTER");
    return -1;
  }
  MPR_GUARD_BEGIN
  MPR_HIP(hipStreamSynchronize(job->stream));
  const uint64_t v = job->host->valid;
  if (count) *count = (int64_t)v;
  if (unit_points && v)
    MPR_HIP(hipMemcpyAsync(unit_points, job->unit, 24 * v, hipMemcpyDeviceToHost, job->stream));
  if (normals && v)
    MPR_HIP(hipMemcpyAsync(normals, job->nrm_kept, 24 * v, hipMemcpyDeviceToHost, job->stream));
  MPR_HIP(hipStreamSynchronize(job->stream));
  return 0;
  MPR_GUARD_END
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
  const mpr_params p = canonical(*params);
  job->adopt(p);
  job->enqueue(p);
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
  const mpr_params p = canonical(*params);
  const Job::State before = job->state();
  MPR_HIP(hipStreamBeginCapture(job->stream, hipStreamCaptureModeThreadLocal));
  try {
    job->enqueue(p);
  } catch (...) {
    job->restore(before);
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(job->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  MPR_HIP(hipStreamEndCapture(job->stream, &job->graph));
  MPR_HIP(hipGraphInstantiate(&job->exec, job->graph, nullptr, nullptr, 0));
  // recording ran nothing: the job's state is still that of its last executed work
  job->captured = p;
  job->has_capture = true;
  job->restore(before);
  return 0;
  MPR_GUARD_END
}

int mpr_job_launch(mpr_job* job) {
  if (!job || (!job->exec && !job->has_capture)) {
    set_error("job has no captured graph");
    return -1;
  }
  if (!job->exec) {  // dropped when the mesh buffers moved: record the same run again
    const mpr_params p = job->captured;
    if (int rc = mpr_job_capture(job, &p)) return rc;
  }
  MPR_GUARD_BEGIN
  MPR_HIP(hipGraphLaunch(job->exec, job->stream));
  job->adopt(job->captured);
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
  job->resolve_overflow(job->stream);
  job->result(r);
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
  if (job->host_mesh) {  // the run's extraction copied the mesh to pinned host memory
    MPR_HIP(hipStreamSynchronize(job->stream));
    if (vertices && nv) std::memcpy(vertices, job->h_verts, 24 * nv);
    if (vertex_density && nv) std::memcpy(vertex_density, job->h_vdens, 8 * nv);
    if (faces && nf) std::memcpy(faces, job->h_faces, 24 * nf);
    return 0;
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

int mpr_job_set_host_mesh(mpr_job* job, int32_t enable) {
  if (!job) return -1;
  MPR_GUARD_BEGIN
  const bool on = enable != 0;
  if (on == job->host_mesh) return 0;
  MPR_HIP(hipStreamSynchronize(job->stream));
  if (on)
    job->alloc_host_mesh();
  else
    job->free_host_mesh();
  job->host_mesh = on;
  // recorded graphs of this job or of batches holding it have the other extraction
  if (job->exec) {
    MPR_HIP(hipGraphExecDestroy(job->exec));
    MPR_HIP(hipGraphDestroy(job->graph));
    job->exec = nullptr;
    job->graph = nullptr;
  }
  ++job->mesh_generation;
  return 0;
  MPR_GUARD_END
}

int mpr_job_host_mesh(mpr_job* job, const double** vertices, const double** vertex_density,
                      const int64_t** faces) {
  if (!job || !job->host_mesh) {
    set_error("the job does not copy its mesh to host memory (mpr_job_set_host_mesh)");
    return -1;
  }
  if (vertices) *vertices = job->h_verts;
  if (vertex_density) *vertex_density = job->h_vdens;
  if (faces) *faces = job->h_faces;
  return 0;
}

int mpr_job_fetch_grids(mpr_job* job, double* field, double* density, double* screen, double* rhs) {
  if (!job) return -1;
  if ((field || density || rhs || screen) && !job->grids_kept) {
    set_error("field, density, screen and rhs are kept only by runs with MPR_STAGE_GRIDS");
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
  // after a filtered run the tree covers the kept rows (their count is read back with the
  // run's scalars, so the run must have finished)
  const bool filtered = job->ran && (job->last.stages & MPR_STAGE_FILTER);
  if (filtered) MPR_HIP(hipStreamSynchronize(s));
  const uint32_t P = filtered ? job->host->valid : job->P;
  const double* upts =
      (job->ran && (job->last.stages & (MPR_STAGE_NORMALIZE | MPR_STAGE_FILTER))) ? job->unit
                                                                                   : job->pts;
  job->oct_points = P;
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
  if (job->oct_points == 0) {  // core.py:94-97
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
    MPR_HIP(hipMemcpyAsync(point_leaf, job->o_leaf, 8ull * job->oct_points, hipMemcpyDeviceToHost,
                           s));
  MPR_HIP(hipStreamSynchronize(s));
  return 0;
  MPR_GUARD_END
}

static bool batch_params_ok(const mpr_batch* b, const mpr_params* p) {
  for (size_t i = 0; i < b->jobs.size(); ++i) {
    if (!valid_params(&p[i])) return false;
    if (p[i].cycles != p[0].cycles || p[i].pre_smooth != p[0].pre_smooth ||
        p[i].post_smooth != p[0].post_smooth || p[i].point_weight != p[0].point_weight ||
        std::max(p[i].inner_cycles, 1) != std::max(p[0].inner_cycles, 1) ||
        (p[i].stages & MPR_STAGE_SOLVE) != (p[0].stages & MPR_STAGE_SOLVE))
      return false;
  }
  return true;
}

int mpr_batch_create(mpr_job* const* jobs, int32_t count, mpr_batch** out) {
  if (!out || !jobs || count < 1) {
    set_error("a batch needs at least one job");
    return -1;
  }
  *out = nullptr;
  for (int i = 0; i < count; ++i)
    if (!jobs[i] || jobs[i]->precision != jobs[0]->precision) {
      set_error("batch jobs must be non-null and share one precision");
      return -1;
    }
  mpr_batch* b = new mpr_batch();
  try {
    b->precision = jobs[0]->precision;
    for (int i = 0; i < count; ++i) b->jobs.push_back(jobs[i]);
    b->stream = BatchResources::get().take_stream();
    b->build();  // host work and a table upload into an idle buffer: no wait on the jobs
    *out = b;
    return 0;
  } catch (const HipError& e) {
    set_error(e.msg);
    delete b;
    return -2;
  }
}

int mpr_batch_destroy(mpr_batch* b) {
  if (!b) return 0;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  delete b;
  return 0;
}

int mpr_batch_run(mpr_batch* b, const mpr_params* params) {
  if (!b || !params || !batch_params_ok(b, params)) {
    set_error("batch parameters must be valid and share cycles, inner cycles, smoothing and "
              "point weight");
    return -1;
  }
  MPR_GUARD_BEGIN
  b->after_job_streams();  // after the jobs' uploads, without waiting on the host
  b->last.clear();
  for (size_t i = 0; i < b->jobs.size(); ++i) b->last.push_back(canonical(params[i]));
  for (size_t i = 0; i < b->jobs.size(); ++i) {
    b->jobs[i]->adopt(b->last[i]);
    b->jobs[i]->ran = true;
  }
  b->enqueue(b->last.data());
  return 0;
  MPR_GUARD_END
}

int mpr_batch_capture(mpr_batch* b, const mpr_params* params) {
  if (!b || !params || !batch_params_ok(b, params)) {
    set_error("batch parameters must be valid and share cycles, inner cycles, smoothing and "
              "point weight");
    return -1;
  }
  MPR_GUARD_BEGIN
  b->drop_graph();
  for (Job* j : b->jobs) MPR_HIP(hipStreamSynchronize(j->stream));
  MPR_HIP(hipStreamSynchronize(b->stream));
  std::vector<mpr_params> cp;
  std::vector<Job::State> before;
  for (size_t i = 0; i < b->jobs.size(); ++i) {
    cp.push_back(canonical(params[i]));
    before.push_back(b->jobs[i]->state());
  }
  auto restore = [&] {  // recording ran nothing: each job keeps its executed work's state
    for (size_t i = 0; i < b->jobs.size(); ++i) b->jobs[i]->restore(before[i]);
  };
  MPR_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  try {
    b->enqueue(cp.data());
  } catch (...) {
    restore();
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(b->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  restore();
  MPR_HIP(hipStreamEndCapture(b->stream, &b->graph));
  MPR_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->last = cp;
  b->generations.clear();
  for (Job* j : b->jobs) b->generations.push_back(j->mesh_generation);
  return 0;
  MPR_GUARD_END
}

int mpr_batch_launch(mpr_batch* b) {
  if (!b || !b->exec) {
    set_error("batch has no captured graph");
    return -1;
  }
  MPR_GUARD_BEGIN
  bool stale = false;
  for (size_t i = 0; i < b->jobs.size(); ++i)
    stale = stale || b->jobs[i]->mesh_generation != b->generations[i];
  if (stale) {  // a job's mesh buffers grew since the capture: record again
    std::vector<mpr_params> p = b->last;
    if (int rc = mpr_batch_capture(b, p.data())) return rc;
  }
  b->after_job_streams();
  for (size_t i = 0; i < b->jobs.size(); ++i) {
    b->jobs[i]->adopt(b->last[i]);
    b->jobs[i]->ran = true;
  }
  MPR_HIP(hipGraphLaunch(b->exec, b->stream));
  return 0;
  MPR_GUARD_END
}

int mpr_batch_wait(mpr_batch* b, mpr_result* results) {
  if (!b) return -1;
  MPR_GUARD_BEGIN
  MPR_HIP(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < b->jobs.size(); ++i) {
    b->jobs[i]->resolve_overflow(b->stream);
    b->jobs[i]->result(results ? &results[i] : nullptr);
  }
  return 0;
  MPR_GUARD_END
}

int mpr_batch_query(mpr_batch* b) {
  if (!b) return -1;
  hipError_t e = hipStreamQuery(b->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  set_error(hipGetErrorString(e));
  return -2;
}

// fraction of the finest level's V-slot groups with a live slot (the groups whose rhs and
// screen words the pipeline's kernels load), for the kernels' algorithmic byte counts
static double live_group_fraction(const mpr_job* job, int V) {
  std::vector<uint32_t> bits(job->live_words());
  MPR_HIP(hipMemcpyAsync(bits.data(), job->live_bits, 4 * bits.size(), hipMemcpyDeviceToHost,
                         job->stream));
  MPR_HIP(hipStreamSynchronize(job->stream));
  const uint64_t slots = 2ull * job->geoms[0].C, groups = slots / V;
  const uint32_t gm = (1u << V) - 1u;
  uint64_t live = 0;
  for (uint64_t k = 0; k < groups; ++k) {
    const uint64_t b = k * V;
    live += ((bits[b >> 5] >> (b & 31)) & gm) != 0;
  }
  return groups ? (double)live / (double)groups : 1.0;
}

// algorithmic bytes of one launch of pipeline kernel 7, 8, 9 or 13 (mpr_job_time_kernel) on
// the job's finest level: words each kernel must move, with the rhs / screen words of
// groups without a live slot left out (the kernels skip them) and one mask bit per slot
static double pipeline_kernel_bytes(const mpr_job* job, int kernel, double pw) {
  const Geom& g0 = job->geoms[0];
  const double M0 = (double)g0.n * g0.n * g0.n;
  const bool skip = group_layout(g0);
  // slots per lane: 4 in FP32, 2 in FP64
  const double live = skip ? live_group_fraction(job, kernel == 8 || kernel == 13 ? 2 : 4) : 1.0;
  const double mask = skip ? M0 / 8.0 : 0.0;
  switch (kernel) {
    case 7:  // one colour: other colour x, rhs (the dense defect), x written; live screen
      return (3.0 + live) * 4.0 * M0 / 2.0 + mask / 2.0;
    case 8:  // one colour: other colour x, x written; live rhs and screen
      return (2.0 + 2.0 * live) * 8.0 * M0 / 2.0 + mask / 2.0;
    case 9:
      // colour 0 only (FP32 after pre-smoothing, BatchedVcycle::colour0_residual): all of x,
      // colour-0 rhs and residual written; live colour-0 screen
      if (job->vc32.colour0_residual(job->vc32.smax, pw, 1))
        return (4.0 + live) * 4.0 * M0 / 2.0 + mask / 2.0;
      return (3.0 + live) * 4.0 * M0 + mask;  // x, rhs, residual written; live screen
    case 13:  // x64 read, r32 written; live rhs64 and screen64
      return (8.0 + 4.0 + 16.0 * live) * M0 + mask;
    case 14:  // fused smoothing, two iterations: the colour read first, the rhs (the dense
              // defect), both colours written once; live screen words and the live bits
      return (2.5 + live) * 4.0 * M0 + mask;
    case 10: {  // fine residual read (colour 0 only in the C1Z form), coarse rhs written
      const Geom& g1 = job->geoms[1];
      const double M1 = (double)g1.n * g1.n * g1.n;
      return job->vc32.colour0_residual(job->vc32.smax, pw, 1) ? 4.0 * (M0 / 2.0 + M1)
                                                               : 4.0 * (M0 + M1);
    }
    case 11: {  // colour 1 only: fine colour-1 x read and written, coarse x read
      const Geom& g1 = job->geoms[1];
      const double M1 = (double)g1.n * g1.n * g1.n;
      return 4.0 * (M0 + M1);
    }
  }
  return 0.0;
}

int mpr_job_time_kernel(mpr_job* job, int32_t kernel, int32_t reps, double* ms_per_launch,
                        double* bytes_per_launch) {
  if (!job || reps < 1 || kernel < 0 || kernel > 14) {
    set_error("invalid kernel timing request");
    return -1;
  }
  if ((kernel == 1 || kernel == 5 || kernel == 6 || kernel == 7 || kernel == 9 ||
       kernel == 10 || kernel == 11 || kernel == 13 || kernel == 14) &&
      job->L32.empty()) {
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
  // the pipeline kernels' byte counts read the live-slot bits back: before the timed region
  double bytes = (kernel == 7 || kernel == 8 || kernel == 9 || kernel == 13 || kernel == 14)
                     ? pipeline_kernel_bytes(job, kernel, pw)
                     : 0.0;
  int launches = reps;
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
  // the V-cycle's batched kernels at the finest class, as the pipeline launches them
  switch (kernel) {
    case 7:
      job->vc32.relax_finest(pw, reps, s);
      launches = 2 * reps;
      break;
    case 8:
      job->vc64.relax_finest(pw, reps, s);
      launches = 2 * reps;
      break;
    case 9:
      for (int r = 0; r < reps; ++r) job->vc32.residual_finest(pw, s);
      break;
    case 10:  // fine colour 0 (colour 1 read as zero after pre-smoothing), coarse written
      for (int r = 0; r < reps; ++r) job->vc32.restrict_finest(pw, s);
      bytes = job->vc32.colour0_residual(job->vc32.smax, pw, 1) ? 4.0 * (M0 / 2.0 + M1)
                                                                 : 4.0 * (M0 + M1);
      break;
    case 11:  // colour 1 only: read coarse and fine colour 1, write fine colour 1
      for (int r = 0; r < reps; ++r) job->vc32.prolong_finest(true, s);
      bytes = 4.0 * (M1 + M0);
      break;
    case 12:
      for (int r = 0; r < reps; ++r) job->vc64.restrict_finest(pw, s);
      bytes = 8.0 * (M0 + M1);
      break;
    case 13:
      for (int r = 0; r < reps; ++r) job->dtab.defect(pw, s);
      break;
    case 14:
      for (int r = 0; r < reps; ++r)
        if (!job->vc32.march_finest(pw, s)) throw HipError{"the finest level does not march"};
      break;
  }
  MPR_HIP(hipEventRecord(e1, s));
  MPR_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MPR_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_per_launch) *ms_per_launch = (double)ms / launches;
  if (bytes_per_launch) *bytes_per_launch = bytes;
  return 0;
  MPR_GUARD_END
}

int mpr_batch_time_kernel(mpr_batch* b, int32_t kernel, int32_t reps, double* ms_per_launch,
                          double* bytes_per_launch) {
  if (!b || reps < 1 ||
      (kernel != 7 && kernel != 8 && kernel != 9 && kernel != 10 && kernel != 11 &&
       kernel != 13 && kernel != 14)) {
    set_error("batch kernel timing: kernel 7, 8, 9, 10, 11, 13 or 14 and reps >= 1");
    return -1;
  }
  const bool mixed = b->precision == MPR_PRECISION_MIXED;
  if ((kernel == 8) == mixed) {
    set_error("batch kernel timing: kernel 8 needs an FP64 batch, 7, 9, 10, 11, 13 and 14 a mixed one");
    return -1;
  }
  for (Job* j : b->jobs)
    if (kernel == 13 && j->depth != b->jobs[0]->depth) {
      set_error("batch kernel timing: kernel 13 needs jobs of one depth");
      return -1;
    }
  MPR_GUARD_BEGIN
  hipStream_t s = b->stream;
  for (Job* j : b->jobs) MPR_HIP(hipStreamSynchronize(j->stream));
  MPR_HIP(hipStreamSynchronize(s));
  const double pw = b->last.empty() ? 4.0 : b->last[0].point_weight;
  const int smax = mixed ? b->vc32.smax : b->vc64.smax;
  // one launch covers every instance of the batch's finest class
  double bytes = 0.0;
  for (Job* j : b->jobs)
    if (j->depth == smax) bytes += pipeline_kernel_bytes(static_cast<const mpr_job*>(j), kernel, pw);
  hipEvent_t e0, e1;
  MPR_HIP(hipEventCreate(&e0));
  MPR_HIP(hipEventCreate(&e1));
  MPR_HIP(hipEventRecord(e0, s));
  int launches = reps;
  switch (kernel) {
    case 7:
      b->vc32.relax_finest(pw, reps, s);
      launches = 2 * reps;
      break;
    case 8:
      b->vc64.relax_finest(pw, reps, s);
      launches = 2 * reps;
      break;
    case 9:
      for (int r = 0; r < reps; ++r) b->vc32.residual_finest(pw, s);
      break;
    case 10:
      for (int r = 0; r < reps; ++r) b->vc32.restrict_finest(pw, s);
      break;
    case 11:
      for (int r = 0; r < reps; ++r) b->vc32.prolong_finest(true, s);
      break;
    case 13:
      for (int r = 0; r < reps; ++r) b->dtab.defect(pw, s);
      break;
    case 14:
      for (int r = 0; r < reps; ++r)
        if (!b->vc32.march_finest(pw, s)) throw HipError{"the finest class does not march"};
      break;
  }
  MPR_HIP(hipEventRecord(e1, s));
  MPR_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MPR_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_per_launch) *ms_per_launch = (double)ms / launches;
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
    uint32_t* counts = scratch.alloc<uint32_t>(2);
    BitSurface surf = bit_surface_sizes((int)n);
    surf.bits = scratch.alloc<uint32_t>(surf.bit_words);
    surf.cmask = scratch.alloc<uint32_t>(surf.runs);
    surf.coff = scratch.alloc<uint32_t>(surf.runs);
    surf.cell_tc = scratch.alloc<uint32_t>(surf.cell_tiles);
    surf.cell_to = scratch.alloc<uint32_t>(surf.cell_tiles);
    surf.edge_tc = scratch.alloc<uint32_t>(surf.edge_tiles);
    surf.edge_to = scratch.alloc<uint32_t>(surf.edge_tiles);
    MPR_HIP(hipMemcpyAsync(f, field, 8 * M, hipMemcpyHostToDevice, s));
    if (density)
      MPR_HIP(hipMemcpyAsync(d, density, 8 * M, hipMemcpyHostToDevice, s));
    else
      launch_fill<double>(d, 0.0, M, s);
    MPR_HIP(hipMemcpyAsync(isod, &iso, 8, hipMemcpyHostToDevice, s));
    // count (nothing fits a zero capacity), then emit into exactly sized buffers
    launch_bit_extract(natural_field(f, (int)n), DensityView{d, nullptr, 1.0}, (int)n, isod,
                       surf, counts, nullptr, nullptr, 0, nullptr, 0, nullptr, s);
    uint32_t hc[2] = {0, 0};
    MPR_HIP(hipMemcpyAsync(hc, counts, 8, hipMemcpyDeviceToHost, s));
    MPR_HIP(hipStreamSynchronize(s));
    m->nv = hc[0];
    m->nq = hc[1];
    surf.vcell_cap = m->nv;
    surf.vcell = m->arena.alloc<uint32_t>(std::max<uint64_t>(1, m->nv));
    m->verts = m->arena.alloc<double>(3 * std::max<uint64_t>(1, m->nv));
    m->vdens = m->arena.alloc<double>(std::max<uint64_t>(1, m->nv));
    m->faces = m->arena.alloc<int64_t>(6 * std::max<uint64_t>(1, m->nq));
    launch_bit_extract(natural_field(f, (int)n), DensityView{d, nullptr, 1.0}, (int)n, isod,
                       surf, counts, m->verts, m->vdens, m->nv, m->faces, 2 * m->nq, nullptr, s);
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
