/*
 * C ABI of the MI355X (gfx950) screened Poisson reconstruction library
 * (libmpr_hip.so). Plain pointers and sizes only; no exceptions cross this boundary.
 *
 * Two groups of entry points:
 *
 * 1. Drop-in kernel ABI. The same symbol names, argument lists, argument meaning and
 *    return values as the reference's ctypes table (python/mojo_poissonrecon/_lib.py:20-32),
 *    operating on caller-owned HOST buffers (C-contiguous float64/int64, grids (n,n,n)
 *    row-major [z][y][x]). Each call copies its inputs to the GPU, runs the HIP kernels and
 *    copies the results back before returning. The reference passes buffers as int64
 *    addresses; these take the same addresses typed as pointers. Like the reference they
 *    cannot raise: on a device error they print to stderr, record mpr_last_error() and
 *    return -1 (integer results) or NaN (float results).
 *
 * 2. Device-resident pipeline ("jobs"). A job owns every device buffer of one
 *    reconstruction at a fixed depth and point capacity; points are uploaded once and the
 *    whole solve (normalise, splat, assemble, multigrid, iso value, extraction) runs on the
 *    job's stream without host round trips. These return 0 on success and a negative
 *    status on failure, with the message in mpr_last_error().
 */
#ifndef MPR_HIP_H
#define MPR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. drop-in kernel ABI ------------------------------------------------------------ */

/* Replaces mpr_parallel_init (src/capi.mojo:36-38; _lib.py:21,85): selects and initialises
 * the GPU; returns 0, or -1 if no usable device is present. */
int64_t mpr_parallel_init(void);

/* Replaces mpr_octree_codes_f64 (src/capi.mojo:62-99; _lib.py:22): codes (P, depth+1). */
void mpr_octree_codes_f64(const double *points, int64_t point_count, int64_t depth,
                          int64_t *codes);

/* Replaces mpr_splat_f64 (src/capi.mojo:102-224; _lib.py:23): accumulates into the five
 * caller-zeroed grids; returns the number of accepted samples. */
int64_t mpr_splat_f64(const double *points, const double *normals, int64_t point_count,
                      int64_t n, int64_t confidence, double *density, double *vx, double *vy,
                      double *vz, double *screen, double samples_per_node);

/* Replaces mpr_assemble_f64 (src/capi.mojo:227-279; _lib.py:24). */
void mpr_assemble_f64(const double *vx, const double *vy, const double *vz,
                      const double *screen, double *rhs, int64_t n, double point_weight,
                      double target_value);

/* Replaces mpr_relax_f64 (src/capi.mojo:309-410; _lib.py:25): in place. */
void mpr_relax_f64(double *solution, const double *rhs, const double *screen, int64_t n,
                   double point_weight, int64_t iterations);

/* Replaces mpr_residual_f64 (src/capi.mojo:473-553; _lib.py:26): returns ||r||_2. */
double mpr_residual_f64(const double *solution, const double *rhs, const double *screen,
                        double *residual, int64_t n, double point_weight);

/* FP32 instantiations of mpr_relax_f64 / mpr_residual_f64 (the smoother and residual of
 * the mixed-precision V-cycle, run by the same kernels) on host float32 grids. No reference
 * counterpart (the reference is FP64 only); they let the FP32 kernels be checked against a
 * float32 restatement. Return 0, or -1 with mpr_last_error() set. */
int mpr_relax_f32(float *solution, const float *rhs, const float *screen, int64_t n,
                  double point_weight, int64_t iterations);
int mpr_residual_f32(const float *solution, const float *rhs, const float *screen,
                     float *residual, int64_t n, double point_weight);
/* FP32 instantiations of mpr_restrict_f64 / mpr_prolong_add_f64 (the mixed-precision
 * V-cycle's grid transfers), same contract. No reference counterpart. */
int mpr_restrict_f32(const float *fine, float *coarse, int64_t fine_n);
int mpr_prolong_add_f32(const float *coarse, float *fine, int64_t fine_n);

/* The mixed V-cycle's step after a level's pre-smoothing, on host float32 grids (n = 2^s + 1
 * in [65, 1025]): the batched colour-0-only residual (colour-1 residual words left at zero;
 * the sweep that ends pre-smoothing leaves them at rounding level) and the restriction that
 * reads the fine colour 1 as zero, into coarse_rhs ((n+1)/2)^3. No reference counterpart: a
 * check of those two kernels against a float32 restatement. 0, or -1 with mpr_last_error(). */
int mpr_vcycle_transfer_f32(const float *solution, const float *rhs, const float *screen,
                            float *residual, float *coarse_rhs, int64_t n, double point_weight);

/* The mixed V-cycle's prolongation before a post-smoothing, on host float32 grids
 * (fine_n = 2^s + 1 in [65, 1025]): fine += trilinear(coarse) on the fine colour-1 nodes
 * ((x + y + z) odd) only, by the batched kernel the V-cycle launches (colour 0 is
 * overwritten by the next sweep without being read). No reference counterpart: a check of
 * that kernel against a float32 restatement. 0, or -1 with mpr_last_error(). */
int mpr_prolong_c1_f32(const float *coarse, float *fine, int64_t fine_n);

/* Replaces mpr_restrict_f64 (src/capi.mojo:556-673; _lib.py:27). */
void mpr_restrict_f64(const double *fine, double *coarse, int64_t fine_n);

/* Replaces mpr_prolong_add_f64 (src/capi.mojo:676-742; _lib.py:28): fine += P coarse. */
void mpr_prolong_add_f64(const double *coarse, double *fine, int64_t fine_n);

/* Replaces mpr_iso_value_f64 (src/capi.mojo:745-795; _lib.py:29). */
double mpr_iso_value_f64(const double *field, const double *points, const double *normals,
                         int64_t point_count, int64_t n, int64_t confidence);

/* Replaces mpr_dual_vertices_f64 (src/capi.mojo:839-1041; _lib.py:30). capacity <= 0:
 * count mode, writes ceil((n-1)/4) partial counts into cell_map. */
int64_t mpr_dual_vertices_f64(const double *field, const double *density, int64_t n, double iso,
                              int64_t *cell_map, double *vertices, double *vertex_density,
                              int64_t capacity);

/* Replaces mpr_dual_faces_f64 (src/capi.mojo:1075-1129; _lib.py:31). */
int64_t mpr_dual_faces_f64(const double *field, int64_t n, double iso, const int64_t *cell_map,
                           int64_t *faces, int64_t capacity);

/* ---- status ---------------------------------------------------------------------------- */

/* Message of the last failure on the calling thread ("" if none). No reference
 * counterpart: the reference's native side cannot fail (MOJO_NOTES.md:41-42). */
const char *mpr_last_error(void);

/* Device name, total HBM bytes and compute-unit count of the current GPU. */
int mpr_device_info(char *name, int64_t name_len, int64_t *total_bytes, int64_t *cu_count);

/* Free and total HBM bytes of the current GPU (hipMemGetInfo); either may be NULL. No
 * reference counterpart (memory accounting for admission by footprint). */
int mpr_device_memory(int64_t *free_bytes, int64_t *total_bytes);

/* ---- 2. device-resident pipeline --------------------------------------------------------
 * The stages replace core.py:196-306 (solve_screened_poisson), core.py:309-374
 * (extract_isosurface) and core.py:388-426 (reconstruct) as one device pass. */

enum {
  MPR_PRECISION_FP64 = 0,  /* every stage in float64, as the reference */
  MPR_PRECISION_MIXED = 1  /* float32 V-cycles inside a float64 defect-correction loop */
};

enum {
  MPR_STAGE_NORMALIZE = 1,  /* points are world space: bbox normalise on device (core.py:377-385) */
  MPR_STAGE_SOLVE = 2,      /* splat, assemble, multigrid, residual, iso value */
  MPR_STAGE_EXTRACT = 4,    /* dual-grid mesh of the solution at the iso value */
  MPR_STAGE_GRIDS = 8,      /* also write natural-layout field, density, screen, rhs for fetching */
  MPR_STAGE_WORLD = 16,     /* write mesh vertices in world space (core.py:425) */
  MPR_STAGE_FILTER = 32     /* drop rows with a non-finite value or a zero normal on device,
                               keeping the order (core.py:73-83, 407-409), then normalise the
                               kept rows as MPR_STAGE_NORMALIZE does */
};

typedef struct mpr_params {
  double point_weight;     /* alpha, default 4.0; finite and >= 0 (-0.0 runs as +0.0) */
  double samples_per_node; /* default 1.5 */
  double scale;            /* bbox scale for MPR_STAGE_NORMALIZE, default 1.1 */
  int32_t cycles;          /* V-cycles, default 4 */
  int32_t confidence;      /* 1: weight samples by normal length */
  int32_t pre_smooth;      /* default 3 */
  int32_t post_smooth;     /* default 3 */
  int32_t stages;          /* MPR_STAGE_* bits */
  int32_t inner_cycles;    /* MPR_PRECISION_MIXED: FP32 V-cycles per FP64 defect (>= 0; 0 and 1
                              give one, the default). The cycles V-cycles run in outer steps
                              of this many: one FP64 defect r = b - A x, that many V-cycles on
                              A e = r (the first from e = 0, the rest from the current e), one
                              x += e. One per defect makes the mixed iterate track the FP64
                              iterate cycle for cycle; two save two FP64 defects and
                              corrections of a 4-cycle solve at a larger distance to the FP64
                              mesh (depth-8 batch of 8: 17.1 -> 15.4 ms; mesh Hausdorff 2e-11
                              -> 2e-10 of the diagonal on the torus, worst of 112 uneven
                              requests 2e-8 -> 5e-6). Ignored in FP64. */
} mpr_params;

typedef struct mpr_result {
  int64_t accepted;       /* splat-accepted samples */
  int64_t vertex_count;
  int64_t face_count;
  double iso_value;
  double residual_norm;   /* float64 ||b - A x|| on the finest grid */
  double origin[3];       /* normalisation transform (MPR_STAGE_NORMALIZE) */
  double width;
  int64_t valid_rows;     /* rows kept by MPR_STAGE_FILTER (all P without it); the splat
                             accepts exactly these */
} mpr_result;

typedef struct mpr_job mpr_job;

/* Device bytes a job of this shape holds (used for admission by HBM footprint): its
 * buffers are carved out of one device allocation of exactly this size, made by
 * mpr_job_create. A run whose mesh outgrows the preallocated mesh buffers adds larger ones. */
int64_t mpr_job_footprint(int32_t depth, int64_t max_points, int32_t precision);
int mpr_job_create(int32_t depth, int64_t max_points, int32_t precision, mpr_job **out);
int mpr_job_destroy(mpr_job *job);
/* Copies P oriented points (host, (P,3) float64 each) to the job; asynchronous w.r.t. host
 * only in that it is ordered on the job's stream. */
int mpr_job_set_points(mpr_job *job, const double *points, const double *normals, int64_t P);
/* Enqueues the requested stages on the job's stream and returns without waiting. */
int mpr_job_run(mpr_job *job, const mpr_params *params);
/* Waits for the job, resolves extraction capacity overflow (re-emitting into larger
 * buffers) and returns the scalar results. */
int mpr_job_wait(mpr_job *job, mpr_result *result);
/* Non-blocking completion test: 1 done, 0 running, <0 error. */
int mpr_job_query(mpr_job *job);
int mpr_job_fetch_mesh(mpr_job *job, double *vertices, double *vertex_density, int64_t *faces);
/* Host mesh output (the mesh half of core.py:422-426 without a separate device-to-host
 * copy): with enable != 0 every extraction of the job also copies its mesh (world space with
 * MPR_STAGE_WORLD) into pinned host buffers owned by the job, on the job's stream (sized on
 * the device from the counts, so it is part of a recorded graph too); mpr_job_fetch_mesh then
 * copies from host memory, and mpr_job_host_mesh gives the pinned buffers themselves (valid
 * until the next run, overflow growth or mpr_job_set_host_mesh). Default off. No reference
 * counterpart. */
int mpr_job_set_host_mesh(mpr_job *job, int32_t enable);
int mpr_job_host_mesh(mpr_job *job, const double **vertices, const double **vertex_density,
                      const int64_t **faces);
/* Any pointer may be NULL. Each grid requires MPR_STAGE_GRIDS on the run (without it the
 * assembly writes only the solver-layout grids near the splat and leaves the density as
 * splatted). */
int mpr_job_fetch_grids(mpr_job *job, double *field, double *density, double *screen,
                        double *rhs);
/* Breadth-first octree of the job's (unit-cube) points (core.py:86-127); synchronises and
 * returns the node count; then fetch the six arrays (point_leaf has P entries). */
int mpr_job_build_octree(mpr_job *job, int32_t depth, int64_t *node_count);
int mpr_job_fetch_octree(mpr_job *job, int64_t *depth, int64_t *morton, int64_t *parent,
                         int64_t *child_start, int64_t *child_count, int64_t *point_leaf);
/* Records the job's run (for fixed params and point count) into a HIP graph (recording
 * runs nothing); mpr_job_launch replays it without re-recording, except that a graph dropped
 * because the job's mesh buffers moved (overflow growth, mpr_job_set_host_mesh) is first
 * recorded again with the same parameters. */
int mpr_job_capture(mpr_job *job, const mpr_params *params);
int mpr_job_launch(mpr_job *job);
/* Device pointers of the job's input buffers (for callers whose points are already in HBM). */
int mpr_job_device_points(mpr_job *job, double **points, double **normals);
/* After a run with MPR_STAGE_FILTER: the number of kept rows and (either may be NULL) their
 * unit-cube points and normals in order ((count,3) float64 each; the caller sizes them for
 * all P rows or first asks with NULL buffers). A check of the device filter and
 * normalisation against core.py:73-83,377-385; no reference counterpart. */
int mpr_job_fetch_filtered(mpr_job *job, int64_t *count, double *unit_points, double *normals);
int mpr_job_set_count(mpr_job *job, int64_t P);

/* ---- batches: many jobs, one launch per V-cycle step ----------------------------------
 * No reference counterpart (the reference solves one request per call). The jobs keep
 * their own buffers; the batch owns a stream and device pointer tables so that every
 * V-cycle step is one launch over all jobs that have a grid of that size (jobs may have
 * different depths). Each job's arithmetic is exactly its solo run's. All jobs of a batch
 * share a precision; per-job params may differ except cycles, pre/post smoothing and
 * point_weight, which must be equal. While a batch runs, its jobs must not run alone. */
typedef struct mpr_batch mpr_batch;
int mpr_batch_create(mpr_job *const *jobs, int32_t count, mpr_batch **out);
int mpr_batch_destroy(mpr_batch *batch);
/* params: one per job */
int mpr_batch_run(mpr_batch *batch, const mpr_params *params);
/* records the whole batched run (setup, V-cycles, extraction, read-back of the scalars)
 * as one HIP graph; mpr_batch_launch replays it in a single launch */
int mpr_batch_capture(mpr_batch *batch, const mpr_params *params);
int mpr_batch_launch(mpr_batch *batch);
int mpr_batch_query(mpr_batch *batch);
/* waits, resolves extraction overflow per job; results: one per job (may be NULL) */
int mpr_batch_wait(mpr_batch *batch, mpr_result *results);
/* Kernel timing of a batch on its own stream with HIP events (as mpr_job_time_kernel, after
 * a run): one launch covers every job of the batch's finest depth. Kernels 7 (FP32 colour
 * sweep), 9 (FP32 residual), 10 (FP32 restriction of the finest residual), 11 (FP32
 * colour-1 prolongation into the finest level), 13 (FP64->FP32 defect; jobs of one
 * depth) and 14 (fused smoothing, two red-black iterations in one pass) for a mixed batch,
 * 8 (FP64 colour sweep) for an FP64 one. bytes_per_launch sums the jobs' algorithmic
 * bytes. No reference counterpart (measurement hook). */
int mpr_batch_time_kernel(mpr_batch *batch, int32_t kernel, int32_t reps, double *ms_per_launch,
                          double *bytes_per_launch);

/* Kernel timing on the job's own stream with HIP events: launches kernel `kernel` `reps`
 * times on the job's finest level (after a run) and returns the mean milliseconds per
 * launch and the algorithmic HBM bytes one launch moves. Kernels: 0 FP64 red-black colour
 * sweep, 1 FP32 colour sweep (mixed jobs), 2 FP64 residual, 3 FP64 restriction,
 * 4 FP64 prolongation, 5 FP32 residual, 6 FP64->FP32 defect residual; the batched V-cycle
 * kernels the pipeline itself launches: 7 FP32 colour sweep, 8 FP64 colour sweep, 9 FP32
 * residual, 10 FP32 restriction, 11 FP32 prolongation of colour 1, 12 FP64 restriction,
 * 13 FP64->FP32 defect residual, 14 FP32 fused smoothing (two red-black iterations of the
 * finest level in one pass, mpr_tune_set_march). Kernels 7, 8, 9, 13 and 14 skip the
 * right-hand-side and screening words that are +0.0 away from the splat (live-slot bits),
 * and their byte counts include only the words they load. No reference counterpart (measurement hook). */
int mpr_job_time_kernel(mpr_job *job, int32_t kernel, int32_t reps, double *ms_per_launch,
                        double *bytes_per_launch);

/* Launch-shape autotuner of the batched slot-group kernels (colour sweep = 0, residual = 1,
 * FP64->FP32 defect residual = 2) on levels n = 2^s + 1, s = 5..max_depth (max_depth in
 * [5, 10]): times the workgroup sizes 64, 128, 256, 512 and 1024 (for the colour sweep also
 * with 2 slot groups per lane, or with one slot group in each of 2 or 4 consecutive planes
 * per lane, the z-column form) per (kernel, precision, s) on up to 8 synthetic
 * instances (within 4 GiB, at least one instance; it allocates and frees them), 7
 * interleaved rounds of `reps` launches each, and keeps the shape with the fastest median
 * round for later launches in this process if it beats the default (256, 1, 1) by 1.5%.
 * Results do not depend on the launch shape. If `table` / `groups` / `planes` are not NULL
 * they receive the chosen workgroup sizes / groups per lane / planes per lane,
 * int32[3][2][11] indexed [kernel][fp64][s]. Run it before capturing graphs (captured
 * launches keep the shape they were recorded with). No reference counterpart (launch
 * tuning). */
int mpr_autotune(int32_t max_depth, int32_t reps, int32_t *table, int32_t *groups,
                 int32_t *planes);
/* current workgroup size of (kernel, fp64, s), or -1 */
int mpr_tune_get(int32_t kernel, int32_t fp64, int32_t size_class);
/* sets it (block one of 64, 128, 256, 512, 1024); 0 or -1 */
int mpr_tune_set(int32_t kernel, int32_t fp64, int32_t size_class, int32_t block);
/* slot groups per lane of (kernel, fp64, s) (1 or 2; only the colour sweep, kernel 0, has
 * a choice), or -1; and its setter (0 or -1) */
int mpr_tune_get_groups(int32_t kernel, int32_t fp64, int32_t size_class);
int mpr_tune_set_groups(int32_t kernel, int32_t fp64, int32_t size_class, int32_t groups);
/* planes per lane of (kernel, fp64, s) (1, 2 or 4; only the colour sweep has a choice;
 * above 1 the sweep runs in z-column form with one slot group per plane, whatever the
 * groups setting), or -1; and its setter (0 or -1) */
int mpr_tune_get_planes(int32_t kernel, int32_t fp64, int32_t size_class);
int mpr_tune_set_planes(int32_t kernel, int32_t fp64, int32_t size_class, int32_t planes);
/* Fused smoothing of the mixed-precision V-cycle: on FP32 levels with 16-byte aligned rows
 * (n >= 33), `iterations` red-black iterations of each smoothing phase run as one pass over
 * the level (a z-marching kernel holding a ring of planes on chip) instead of 2 x iterations
 * single colour sweeps: 0 (every sweep on its own), 1, or 2 (the default). Results are
 * bit-identical for every setting. It applies to jobs and batches created afterwards. No
 * reference counterpart (a schedule of relax_f64's arithmetic, capi.mojo:309-410). get
 * returns the setting, set 0 or -1. */
int mpr_tune_get_march(void);
int mpr_tune_set_march(int32_t iterations);
/* Test hook of the fused smoothing's division: on a node without screening weight the
 * diagonal is d = deg (n-1)^2 (deg = 3..6) and the pass divides by it with one
 * multiplication and one FMA (the reciprocal split into two floats; numerators of magnitude
 * below 2^-60 divide the IEEE way). Tries all 2^32 float bit patterns as the
 * numerator for each of the four d of level n (skipping inf and NaN) and stores in
 * *mismatches how many quotients differ in any bit from the IEEE quotient (0 means the fast
 * division is exact for every finite numerator). Takes about a second per level. No
 * reference counterpart. 0 or a negative status. */
int mpr_fast_division_check(int32_t n, int64_t *mismatches);

/* Allocation-free variant of mpr_iso_value_f64. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops
 * that must not call hipMalloc). field ((n,n,n) float64), points and normals ((P,3) float64)
 * and iso_value (one double, the result) are DEVICE pointers; the work is enqueued on
 * `stream` (a hipStream_t; NULL is the default stream) and the call returns without
 * waiting. The kernels are mpr_iso_value_f64's, so the value is the same bit for bit. Their
 * partial sums live in a static device array of the library (1 MiB per GPU, held until the
 * process ends), which bounds point_count to 2^24; n must be in [2, 1025]. Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (on different streams or
 * host threads) race. Allocates nothing (no hipMalloc, hipFree or allocating library
 * call). Returns 0, or -1 with mpr_last_error() set. */
int mpr_iso_value_f64_noalloc(const double *field, const double *points, const double *normals,
                              int64_t point_count, int64_t n, int64_t confidence,
                              double *iso_value, void *stream);

/* Referenced vertices and area-weighted unit normals of a triangle mesh (core.py:429-450,
 * oriented_points_from_mesh), on the GPU with the arithmetic of the reference's per-face
 * loop (results bit-identical): host vertices (V,3) float64 and faces (F,3) int64 in;
 * points_out and normals_out ((V,3) float64 each, caller-sized for all V) receive the kept
 * vertices in order and *count_out their number. Returns 0, or -1 (face index out of range,
 * device error) with mpr_last_error() set. */
int mpr_oriented_points_from_mesh(const double *vertices, int64_t vertex_count,
                                  const int64_t *faces, int64_t face_count, double *points_out,
                                  double *normals_out, int64_t *count_out);

/* Distances between two triangle meshes (host buffers: vertices (V,3) float64, faces (F,3)
 * int64), the accuracy metric of the reference's end-to-end tests
 * (tests/test_reconstruction.py:50-62), computed exactly (point to triangle) on the GPU:
 * out[0] = max over A's vertices of the distance to B's surface, out[1] = the same from B to
 * A, out[2] and out[3] the corresponding means. The symmetric Hausdorff distance is
 * max(out[0], out[1]); the vertex chamfer (out[2] + out[3]) / 2. A direction with no query
 * vertex gives 0; one with vertices but no target triangle gives +inf. Returns 0, or -1 with
 * mpr_last_error() set (face index out of range, non-finite vertex, device error). */
int mpr_mesh_distance(const double *vertices_a, int64_t vertex_count_a, const int64_t *faces_a,
                      int64_t face_count_a, const double *vertices_b, int64_t vertex_count_b,
                      const int64_t *faces_b, int64_t face_count_b, double *out);

/* Extraction of a caller grid (core.py:309-374): field and density host (n,n,n) float64
 * (density may be NULL); returns counts; then fetch and destroy. */
typedef struct mpr_mesh mpr_mesh;
int mpr_extract(const double *field, const double *density, int64_t n, double iso,
                mpr_mesh **out, int64_t *vertex_count, int64_t *face_count);
int mpr_mesh_fetch(mpr_mesh *mesh, double *vertices, double *vertex_density, int64_t *faces);
int mpr_mesh_destroy(mpr_mesh *mesh);

#ifdef __cplusplus
}
#endif
#endif
