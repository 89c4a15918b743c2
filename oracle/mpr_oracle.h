/*
 * TEST INFRASTRUCTURE ONLY. This is the CPU oracle: an op-for-op FP64
 * restatement of the reference kernels in src/capi.mojo. Only tests/, the
 * smoke check and bench.py's cpu_baseline leg may load it, and only as the
 * checker (or the timed CPU baseline). The product library never links it.
 *
 * Every function cites the reference lines it restates. Grids are (n,n,n)
 * row-major [z][y][x] (capi.mojo:41-48).
 */
#ifndef MPR_ORACLE_H
#define MPR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

double oracle_norm3(double x, double y, double z);
/* threads the grid kernels use (1 unless built with OpenMP) */
int oracle_threads(void);
void oracle_octree_codes(const double *points, int64_t point_count, int64_t depth,
                         int64_t *codes);
int64_t oracle_splat(const double *points, const double *normals, int64_t point_count,
                     int64_t n, int64_t confidence, double *density, double *vx, double *vy,
                     double *vz, double *screen, double samples_per_node);
void oracle_assemble(const double *vx, const double *vy, const double *vz,
                     const double *screen, double *rhs, int64_t n, double point_weight,
                     double target_value);
void oracle_relax(double *solution, const double *rhs, const double *screen, int64_t n,
                  double point_weight, int64_t iterations);
double oracle_residual(const double *solution, const double *rhs, const double *screen,
                       double *residual, int64_t n, double point_weight);
void oracle_restrict(const double *fine, double *coarse, int64_t fine_n);
void oracle_prolong_add(const double *coarse, double *fine, int64_t fine_n);
double oracle_iso_value(const double *field, const double *points, const double *normals,
                        int64_t point_count, int64_t n, int64_t confidence);
int64_t oracle_dual_vertices(const double *field, const double *density, int64_t n,
                             double iso, int64_t *cell_map, double *vertices,
                             double *vertex_density, int64_t capacity);
int64_t oracle_dual_faces(const double *field, int64_t n, double iso, const int64_t *cell_map,
                          int64_t *faces, int64_t capacity);

/* core.py:429-450: referenced vertices and area-weighted normals (the per-face loop). */
int64_t oracle_oriented_points(const double *vertices, int64_t vertex_count, const int64_t *faces,
                               int64_t face_count, double *points_out, double *normals_out);

/* core.py:196-306 orchestration after validation and the octree: splat, density/spn,
 * assemble, screen pyramid, `cycles` V-cycles, final residual and iso value.
 * Outputs are (n,n,n) with n = 2^depth + 1; returns the splat's accepted count. */
int64_t oracle_solve(const double *points, const double *normals, int64_t point_count,
                     int64_t depth, double point_weight, int64_t confidence, int64_t cycles,
                     double samples_per_node, double *field, double *density, double *screen,
                     double *rhs, double *iso_value, double *residual_norm);

#ifdef __cplusplus
}
#endif
#endif
