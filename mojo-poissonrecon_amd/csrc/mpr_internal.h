// Internal launch interface between the kernel files and the host-side driver.
// Every launcher enqueues on the given stream and never synchronises.
#pragma once

#include "mpr_common.h"

namespace mpr {

constexpr int kThreads = 256;

template <class T>
struct LevelPtrs {
  T* x;
  T* rhs;
  T* screen;
  T* res;
  Geom g;
  // second x buffer of an FP32 level that the fused smoothing (march.hip) writes out of
  // place, or null
  T* xalt = nullptr;
};

// ---- grid kernels (grid_kernels.hip) ----------------------------------------------------
template <class T> void launch_fill(T* p, T value, uint64_t count, hipStream_t s);
// fills of several buffers in one launch (multi.h); count 0 skips a job
template <class T>
struct FillJob {
  T* p;
  T value;
  uint64_t count;
};
template <class T> void launch_fill_multi(const FillJob<T>* jobs, size_t count, hipStream_t s);
// node-by-node restriction of several grids in one launch (launch_restrict's scalar form);
// a job with F null is skipped. With `need` (a natural-layout bit set of the fine level)
// only the coarse nodes whose restriction support holds one of its nodes are written: the
// caller guarantees that the fine words outside it are +0.0 and that the coarse words
// outside its support already hold +0.0 (the restriction of +0.0 words). `support` then
// holds coarse_support_words(gc.n) words of scratch.
template <class T>
struct RestrictJob {
  const T* F;
  Geom gf;
  T* Cc;
  Geom gc;
  double scale;
  float* copy32;
  const uint32_t* need;
  uint32_t* support;
};
// the support words, a list of as many word indices and the list's length
inline uint64_t coarse_support_words(int nc) {
  return 2 * (uint64_t)nc * nc * (((uint64_t)nc + 31) / 32) + 1;
}
template <class T>
void launch_restrict_nodes_multi(const RestrictJob<T>* jobs, size_t count, hipStream_t s);
template <class T>
void launch_relax(T* X, const T* R, const T* S, const Geom& g, double pw, int iterations,
                  hipStream_t s);
// one red-black half sweep (colour c)
template <class T>
void launch_relax_color(T* X, const T* R, const T* S, const Geom& g, double pw, int c,
                        hipStream_t s);
// r = rhs - A x written as TR into `out` (split layout). When `partials` is non-null each
// block also writes the sum of squares of its residuals there; returns the block count.
// live: the level's live-slot bits covering R and S (LiveMask, stencil.h), or null.
template <class T, class TR>
int launch_residual(const T* X, const T* R, const T* S, TR* out, const Geom& g, double pw,
                    double* partials, hipStream_t s, const uint32_t* live = nullptr);
// number of squared-norm partials launch_residual writes for a level of word_bytes words
int residual_partial_count(const Geom& g, size_t word_bytes);
// out = sqrt(sum partials) with a fixed summation order
void launch_norm_finish(const double* partials, int count, double* out, hipStream_t s);
// the final FP64 residual norm of several solutions in one launch per step: out = sqrt of
// the sum of squared residuals of (X, R, S), nparts = residual_partial_count(g, 8)
struct NormJob {
  const double *X, *R, *S;
  Geom g;
  double pw;
  double* partials;
  int nparts;
  double* out;
  const uint32_t* live;
  // optional, where residual_norm_sets_sign_bits(g): the extraction's sign bits of X against
  // *iso (n * n * ceil(n / 32) words, launch_bit_extract's layout), set in the same pass
  const double* iso;
  uint32_t* sign_bits;
};
bool residual_norm_sets_sign_bits(const Geom& g);
void launch_residual_norm_multi(const NormJob* jobs, size_t count, hipStream_t s);
template <class T>
void launch_restrict(const T* F, const Geom& gf, T* Cc, const Geom& gc, double scale,
                     hipStream_t s, float* copy32 = nullptr);  // copy32: also store as float
template <class T>
void launch_prolong_add(const T* Cc, const Geom& gc, T* F, const Geom& gf, hipStream_t s);
template <class TI, class TO>
void launch_nat_to_split(const TI* nat, TO* split, const Geom& g, hipStream_t s);
template <class T> void launch_split_to_nat(const T* split, double* nat, const Geom& g, hipStream_t s);
template <class TI, class TO> void launch_convert(const TI* in, TO* out, uint64_t count, hipStream_t s);
void launch_add_correction(double* X, const float* E, uint64_t count, hipStream_t s);
// rhs = -div(v) + alpha * screen * h^-2 * target (capi.mojo:227-279), written in the natural
// layout (rhs_nat, optional) and the split layout (rhs_split); also copies the screening
// weights into the split layout and, if density is non-null, divides it by spn in place
// (core.py:255).
// RHS assembly (capi.mojo:227-279) from the natural splat grids; rhs_nat / rhs_split /
// screen_split (+ its float copy) / screen_nat_out / density (/= spn, core.py:255) are
// written when non-null.
// touched: the splat's node bits; unset nodes read as +0.0 (their grids were not filled).
//
// Sparse assembly (hist != null, write_all false; split outputs only, density left as
// splatted): a node whose 7-point stencil holds no touched node gets
// rhs = -(+0) + pw * (+0) * h^-2 * target = +0.0 and screen = +0.0 (pw >= +0, target >= 0),
// the same in every run. So only the 64-node row chunks holding a node whose stencil meets
// this run's or the previous run's touched nodes are written, one wavefront per listed
// chunk; every other node keeps the +0.0 an earlier run wrote.
// `prev` holds the previous assembly's touched bits (all ones before the first, so that one
// writes every node) and is updated in place by every assembly given `hist` (write_all
// true: a dense assembly that still records its touched bits); `need` is scratch of the
// same size. live (nullable, zeroed by the caller): the split layout's live-slot bits
// (LiveMask, stencil.h) are set where rhs or screen is written non-zero.
struct TouchedHistory {
  uint32_t* prev;
  uint32_t* need;
  bool write_all;
  uint32_t* list;   // ceil(n^3 / 32) words: the need-mask words holding a node
  uint32_t* count;  // one word: their number
};
inline uint32_t assemble_chunks_per_row(int n) { return (uint32_t)(n + 63) / 64u; }
void launch_assemble(const double* vx, const double* vy, const double* vz,
                     const double* screen_nat, double* rhs_nat, double* rhs_split,
                     double* screen_split, double* density, double spn, const Geom& g,
                     double pw, double target, const uint32_t* touched, double* screen_nat_out,
                     float* screen_split32, const TouchedHistory* hist, uint32_t* live,
                     hipStream_t s);
// one instance of launch_assemble_multi (launch_assemble's arguments)
struct AssembleJob {
  const double *vx, *vy, *vz, *screen_nat;
  double *rhs_nat, *rhs_split, *screen_split, *density;
  double spn;
  Geom g;
  double pw, target;
  const uint32_t* touched;
  double* screen_nat_out;
  float* screen_split32;
  const TouchedHistory* hist;
  uint32_t* live;
};
void launch_assemble_multi(const AssembleJob* jobs, size_t count, hipStream_t s);
// ---- point kernels (point_kernels.hip) --------------------------------------------------
int bbox_partial_count(uint32_t P);
// core.py:377-385 on the device: params = [origin x, y, z, width], unit = (p - origin)/width
void launch_normalize(const double* pts, uint32_t P, double scale, double* part, double* params,
                      double* unit, hipStream_t s);
// Row filter of core.py:73-83 fused with the normalisation of core.py:377-385: the rows with
// six finite values and a non-zero normal, in order, go to unit[0..V) (normalised by their
// bounding box) and nrm_out[0..V); the other P - V rows follow as zero points and normals.
// *total (device) receives V. Scratch: filter_blocks(P) counts and offsets, 6 * that many
// doubles of bbox partials, scan_scratch_words(filter_blocks(P)) scan words.
struct FilterScratch {
  uint32_t *counts = nullptr, *offsets = nullptr, *scan = nullptr;
  double* part = nullptr;
};
uint32_t filter_blocks(uint32_t P);
// one instance of the row filter + normalisation (launch_filter_normalize's arguments)
struct FilterJob {
  const double *pts, *nrm;
  uint32_t P;
  double scale;
  FilterScratch f;
  uint32_t* total;
  double *params, *unit, *nrm_out;
};
// every job's filter in one launch per step (multi.h)
void launch_filter_normalize_multi(const FilterJob* jobs, size_t count, hipStream_t s);
struct ScanJob {
  const uint32_t* in;
  uint32_t* out;
  uint32_t count;
  uint32_t* total;
  uint32_t* scratch;
};
void launch_scan_exclusive_multi(const ScanJob* jobs, size_t count, hipStream_t s);
void launch_filter_normalize(const double* pts, const double* nrm, uint32_t P, double scale,
                             const FilterScratch& f, uint32_t* total, double* params, double* unit,
                             double* nrm_out, hipStream_t s);
void launch_octree_codes(const double* pts, uint32_t P, int depth, int64_t* codes, hipStream_t s);
void launch_scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t count, uint32_t* total,
                           hipStream_t s, uint32_t* scratch = nullptr);
// words of `scratch` that let launch_scan_exclusive use many workgroups for `count` values
uint32_t scan_scratch_words(uint32_t count);
size_t radix_scratch_words(uint32_t n);
// gather_dst (optional, n doubles): gather_src[vals[i]] in the sorted order. With
// result_keys and result_vals the sorted pairs stay in whichever buffers the last pass
// wrote (keys/vals or k_tmp/v_tmp), which they are set to; without, they end in keys/vals.
template <class K>  // uint32_t (splat node keys) or uint64_t (Morton keys)
void launch_radix_sort(K* keys, uint32_t* vals, K* k_tmp, uint32_t* v_tmp, uint32_t n, int bits,
                       uint32_t* hist, uint32_t* offs, hipStream_t s,
                       const double* gather_src = nullptr, double* gather_dst = nullptr,
                       K** result_keys = nullptr, uint32_t** result_vals = nullptr);
// one sort of launch_radix_sort_multi (launch_radix_sort's arguments)
template <class K>
struct SortJob {
  K* keys;
  uint32_t* vals;
  K* k_tmp;
  uint32_t* v_tmp;
  uint32_t n;
  int bits;
  uint32_t *hist, *offs;
  const double* gather_src;
  double* gather_dst;
  K** result_keys;
  uint32_t** result_vals;
};
// every job's sort in one launch per step (multi.h); the jobs share the digit schedule of
// the widest key, which leaves each job's sorted order what its own sort gives
template <class K>
void launch_radix_sort_multi(const SortJob<K>* jobs, size_t count, hipStream_t s);
void launch_leaf_keys(const double* pts, uint32_t P, int depth, uint64_t* keys, uint32_t* perm,
                      hipStream_t s);
uint32_t compaction_tiles(uint64_t n);
void launch_level_counts(const uint64_t* keys, uint32_t n, int d, int depth, uint32_t* counts,
                         uint32_t* offsets_out, uint32_t* level_total, hipStream_t s);
void launch_level_offsets(const uint32_t* level_count, int levels, int64_t* offsets,
                          hipStream_t s);
void launch_level_emit(const uint64_t* keys, uint32_t n, int d, int depth, const int64_t* offsets,
                       const uint32_t* tile_offsets, const uint32_t* uidx_prev, uint32_t* uidx_cur,
                       int64_t* morton, int64_t* depth_out, int64_t* parent, int64_t* child_start,
                       int64_t* child_count, hipStream_t s);
void launch_point_leaf(const uint32_t* perm, const uint32_t* uidx, uint32_t n,
                       const int64_t* offsets, int depth, int64_t* point_leaf, hipStream_t s);
// capi.mojo:102-224; density/screen/vx/vy/vz accumulate (callers zero them)
// Largest point count of one job or splat call (8 splat contributions per point are indexed
// with 32 bits).
constexpr int64_t kMaxPoints = int64_t(1) << 28;

// Device scratch of the deterministic splat for up to pairs / 8 points.
struct SplatScratch {
  uint32_t pairs = 0;        // 8 per point
  size_t radix_words = 0;    // words of hist and of offs
  uint32_t *keys = nullptr, *ktmp = nullptr;  // node index of each (node, point) pair
  uint32_t *ids = nullptr, *itmp = nullptr, *hist = nullptr, *offs = nullptr;
  double* w1 = nullptr;  // pairs
  double* c2 = nullptr;  // 3 * pairs
};
SplatScratch splat_scratch_sizes(uint32_t P);
// Both splat passes (capi.mojo:102-224), accumulating into the five grids in the reference's
// point order; `accepted` (device) is incremented by the accepted point count. With
// `touched` (ceil(n^3 / 32) zeroed words) the grids are taken as zero-filled without being
// filled: only the nodes that receive contributions are written, and their bits are set
// (readers treat an unset node as +0.0).
void launch_splat(const double* pts, const double* nrm, uint32_t P, int n, int confidence,
                  double* density, double* vx, double* vy, double* vz, double* screen,
                  double spn, unsigned long long* accepted, const SplatScratch& b,
                  uint32_t* touched, hipStream_t s);
// one instance of launch_splat_multi (launch_splat's arguments)
struct SplatJob {
  const double *pts, *nrm;
  uint32_t P;
  int n, confidence;
  double *density, *vx, *vy, *vz, *screen;
  double spn;
  unsigned long long* accepted;
  SplatScratch b;
  uint32_t* touched;
};
// every job's splat in one launch per step (multi.h), each job's results those of its own
void launch_splat_multi(const SplatJob* jobs, size_t count, hipStream_t s);
int iso_partial_count(uint32_t P);
// one instance of launch_iso_multi (launch_iso's arguments)
struct IsoJob {
  FieldView field;
  const double *pts, *nrm;
  uint32_t P;
  int n, confidence;
  double *part, *iso;
  const uint32_t* count;
};
void launch_iso_multi(const IsoJob* jobs, size_t count, hipStream_t s);
// count (device, nullable): only rows below *count take part
void launch_iso(const FieldView& field, const double* pts, const double* nrm, uint32_t P, int n,
                int confidence, double* part, double* iso, hipStream_t s,
                const uint32_t* count = nullptr);

// ---- extraction kernels (extract_kernels.hip) --------------------------------------------
uint32_t cell_tiles(int n);
uint32_t edge_tiles(int n);
// iso is a device pointer: the pipeline computes it on the GPU
// Device buffers of the sign-bit extraction (launch_bit_extract) for an n^3 field.
struct BitSurface {
  uint64_t bit_words = 0, runs = 0;  // sizes of bits and of cmask / coff
  uint32_t cell_tiles = 0, edge_tiles = 0;
  uint32_t *bits = nullptr, *cmask = nullptr, *coff = nullptr;
  uint32_t* vcell = nullptr;  // cell of each vertex, vcell_cap entries (>= the vertex capacity)
  uint64_t vcell_cap = 0;
  uint32_t *cell_tc = nullptr, *cell_to = nullptr, *edge_tc = nullptr, *edge_to = nullptr;
};
BitSurface bit_surface_sizes(int n);
// Dual-grid extraction of the device pipeline (capi.mojo:839-1129, same vertex indices and
// face order): counts[0] = vertices, counts[1] = quads; vertices past vcap and faces past
// fcap are counted but not written. xform: optional [origin x, y, z, width] world transform.
void launch_bit_extract(const FieldView& f, const DensityView& dens, int n, const double* iso,
                        const BitSurface& b, uint32_t* counts, double* verts, double* vdens,
                        uint64_t vcap, int64_t* faces, uint64_t fcap, const double* xform,
                        hipStream_t s);

// one instance of launch_bit_extract_multi (launch_bit_extract's arguments)
struct ExtractJob {
  FieldView f;
  DensityView dens;
  int n;
  const double* iso;
  const BitSurface* b;
  uint32_t* counts;
  double *verts, *vdens;
  uint64_t vcap;
  int64_t* faces;
  uint64_t fcap;
  const double* xform;
  bool bits_ready;  // b->bits already hold the sign bits (launch_residual_norm_multi)
};
void launch_bit_extract_multi(const ExtractJob* jobs, size_t count, hipStream_t s);

// copies the mesh (counts[0] vertices, 2 * counts[1] faces, clamped to the capacities, sizes
// read on the device) to host buffers of the same layout, e.g. pinned host memory
void launch_mesh_to_host(const double* verts, const double* vdens, const int64_t* faces,
                         const uint32_t* counts, uint64_t vcap, uint64_t fcap, double* h_verts,
                         double* h_vdens, int64_t* h_faces, hipStream_t s);
struct MeshCopyJob {
  const double *verts, *vdens;
  const int64_t* faces;
  const uint32_t* counts;
  uint64_t vcap, fcap;
  double *h_verts, *h_vdens;
  int64_t* h_faces;
};
void launch_mesh_to_host_multi(const MeshCopyJob* jobs, size_t count, hipStream_t s);

void launch_cells_count(const double* f, int n, const double* iso, uint32_t* counts, hipStream_t s);
void launch_cells_chunk_counts(const double* f, int n, const double* iso, int64_t* chunks,
                               hipStream_t s);
template <class CM>
void launch_cells_emit(const double* f, const double* dens, int n, const double* iso,
                       const uint32_t* tile_offsets, CM* cell_map, double* verts, double* vdens,
                       uint64_t capacity, const double* xform, hipStream_t s);
template <class CM>
void launch_edges_count(const double* f, int n, const double* iso, const CM* cm, uint32_t* counts,
                        hipStream_t s);
template <class CM>
void launch_edges_emit(const double* f, int n, const double* iso, const CM* cm,
                       const uint32_t* tile_offsets, int64_t* faces, uint64_t capacity,
                       hipStream_t s);

}  // namespace mpr
