/*
 * TEST INFRASTRUCTURE ONLY: the CPU oracle (see mpr_oracle.h). A plain-C, serial,
 * FP64 restatement of the reference kernels in src/capi.mojo and of the solver
 * orchestration in python/mojo_poissonrecon/core.py. Built with
 * -ffp-contract=off so every product and sum rounds exactly as written.
 *
 * Where the reference has a SIMD body plus a scalar tail that evaluate the same
 * expression, this file writes the expression once; the per-element operation
 * order is kept (neighbour sums x-1, x+1, y-1, y+1, z-1, z+1; restriction terms in
 * (oz, oy, ox) order; trilinear corners in (dz, dy, dx) order).
 */
#include "mpr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Test-only slips in the solver orchestration: tests/test_oracle_solve.py builds copies of
 * this file with -DORACLE_MUTATION=k (k = 1..9) and shows that its checks reject each one.
 * The oracle proper is built with the default, 0, which is the restatement itself. */
#ifndef ORACLE_MUTATION
#define ORACLE_MUTATION 0
#endif
#define MUTATED(k) (ORACLE_MUTATION == (k))

/* Built with -fopenmp (oracle.build_openmp), the plane loops of the grid kernels run on
 * several threads: within one colour sweep, restriction, prolongation, assembly or residual
 * every output depends only on inputs no other iteration writes, so the results equal the
 * serial build's bit for bit (the residual norm, the splat and the extraction stay serial).
 * It exists to time the reference algorithm on all host cores. */
#ifdef _OPENMP
#include <omp.h>
#define ORACLE_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#else
#define ORACLE_PARALLEL_FOR
#endif

int oracle_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static inline int64_t gidx(int64_t x, int64_t y, int64_t z, int64_t n) {
    return (z * n + y) * n + x; /* capi.mojo:41-43 */
}

/* capi.mojo:51-59 */
double oracle_norm3(double x, double y, double z) {
    double ax = fabs(x), ay = fabs(y), az = fabs(z);
    double m = ay > az ? ay : az;
    double scale = ax > m ? ax : m;
    if (scale == 0.0) return 0.0;
    double sx = x / scale, sy = y / scale, sz = z / scale;
    return scale * sqrt(sx * sx + sy * sy + sz * sz);
}

/* capi.mojo:62-99 */
void oracle_octree_codes(const double *points, int64_t point_count, int64_t depth,
                         int64_t *codes) {
    for (int64_t i = 0; i < point_count; ++i) {
        double x = points[3 * i], y = points[3 * i + 1], z = points[3 * i + 2];
        int64_t code = 0;
        codes[i * (depth + 1)] = 0;
        for (int64_t d = 0; d < depth; ++d) {
            int64_t scale = (int64_t)1 << (d + 1);
            int64_t ix = (int64_t)(x * (double)scale);
            int64_t iy = (int64_t)(y * (double)scale);
            int64_t iz = (int64_t)(z * (double)scale);
            if (ix < 0) ix = 0;
            if (iy < 0) iy = 0;
            if (iz < 0) iz = 0;
            if (ix >= scale) ix = scale - 1;
            if (iy >= scale) iy = scale - 1;
            if (iz >= scale) iz = scale - 1;
            int64_t child = (ix & 1) | ((iy & 1) << 1) | ((iz & 1) << 2);
            code = (code << 3) | child;
            codes[i * (depth + 1) + d + 1] = code;
        }
    }
}

/* capi.mojo:102-224 (two passes: density/screen, then density-normalised vector field) */
int64_t oracle_splat(const double *points, const double *normals, int64_t point_count,
                     int64_t n, int64_t confidence, double *density, double *vx, double *vy,
                     double *vz, double *screen, double samples_per_node) {
    int64_t res = n - 1;
    int64_t accepted = 0;
    for (int64_t p = 0; p < point_count; ++p) {
        double nx = normals[3 * p], ny = normals[3 * p + 1], nz = normals[3 * p + 2];
        double length = oracle_norm3(nx, ny, nz);
        if (length <= 0.0) continue;
        double sw = confidence != 0 ? length : 1.0;
        double gx = points[3 * p] * (double)res;
        double gy = points[3 * p + 1] * (double)res;
        double gz = points[3 * p + 2] * (double)res;
        int64_t ix = (int64_t)gx, iy = (int64_t)gy, iz = (int64_t)gz;
        if (ix < 0) ix = 0;
        if (iy < 0) iy = 0;
        if (iz < 0) iz = 0;
        if (ix >= res) { ix = res - 1; gx = (double)res; }
        if (iy >= res) { iy = res - 1; gy = (double)res; }
        if (iz >= res) { iz = res - 1; gz = (double)res; }
        double fx = gx - (double)ix, fy = gy - (double)iy, fz = gz - (double)iz;
        for (int dz = 0; dz < 2; ++dz) {
            double wz = dz == 1 ? fz : 1.0 - fz;
            for (int dy = 0; dy < 2; ++dy) {
                double wy = dy == 1 ? fy : 1.0 - fy;
                for (int dx = 0; dx < 2; ++dx) {
                    double wx = dx == 1 ? fx : 1.0 - fx;
                    double w = wx * wy * wz * sw;
                    int64_t idx = gidx(ix + dx, iy + dy, iz + dz, n);
                    density[idx] += 2.0 * w;
                    screen[idx] += w;
                }
            }
        }
        accepted += 1;
    }
    for (int64_t p = 0; p < point_count; ++p) {
        double nx = normals[3 * p], ny = normals[3 * p + 1], nz = normals[3 * p + 2];
        double length = oracle_norm3(nx, ny, nz);
        if (length <= 0.0) continue;
        double sw = confidence != 0 ? length : 1.0;
        nx /= length;
        ny /= length;
        nz /= length;
        double gx = points[3 * p] * (double)res;
        double gy = points[3 * p + 1] * (double)res;
        double gz = points[3 * p + 2] * (double)res;
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
            double wz = dz == 1 ? fz : 1.0 - fz;
            for (int dy = 0; dy < 2; ++dy) {
                double wy = dy == 1 ? fy : 1.0 - fy;
                for (int dx = 0; dx < 2; ++dx) {
                    double wx = dx == 1 ? fx : 1.0 - fx;
                    local += density[gidx(ix + dx, iy + dy, iz + dz, n)] * wx * wy * wz;
                }
            }
        }
        double q = local / samples_per_node;
        double factor = q > 1.0 ? q : 1.0;
        for (int dz = 0; dz < 2; ++dz) {
            double wz = dz == 1 ? fz : 1.0 - fz;
            for (int dy = 0; dy < 2; ++dy) {
                double wy = dy == 1 ? fy : 1.0 - fy;
                for (int dx = 0; dx < 2; ++dx) {
                    double wx = dx == 1 ? fx : 1.0 - fx;
                    double w = wx * wy * wz * sw;
                    int64_t idx = gidx(ix + dx, iy + dy, iz + dz, n);
                    double vw = w * (double)res / factor;
                    vx[idx] -= nx * vw;
                    vy[idx] -= ny * vw;
                    vz[idx] -= nz * vw;
                }
            }
        }
    }
    return accepted;
}

/* capi.mojo:227-279 */
void oracle_assemble(const double *vx, const double *vy, const double *vz,
                     const double *screen, double *rhs, int64_t n, double point_weight,
                     double target_value) {
    double h_inv = (double)(n - 1);
    double h_inv2 = h_inv * h_inv;
    ORACLE_PARALLEL_FOR
    for (int64_t z = 0; z < n; ++z)
        for (int64_t y = 0; y < n; ++y)
            for (int64_t x = 0; x < n; ++x) {
                int64_t idx = gidx(x, y, z, n);
                double div = 0.0;
                if (x == 0) div += (vx[gidx(1, y, z, n)] - vx[idx]) * h_inv;
                else if (x == n - 1) div += (vx[idx] - vx[gidx(x - 1, y, z, n)]) * h_inv;
                else div += (vx[gidx(x + 1, y, z, n)] - vx[gidx(x - 1, y, z, n)]) * (0.5 * h_inv);
                if (y == 0) div += (vy[gidx(x, 1, z, n)] - vy[idx]) * h_inv;
                else if (y == n - 1) div += (vy[idx] - vy[gidx(x, y - 1, z, n)]) * h_inv;
                else div += (vy[gidx(x, y + 1, z, n)] - vy[gidx(x, y - 1, z, n)]) * (0.5 * h_inv);
                if (z == 0) div += (vz[gidx(x, y, 1, n)] - vz[idx]) * h_inv;
                else if (z == n - 1) div += (vz[idx] - vz[gidx(x, y, z - 1, n)]) * h_inv;
                else div += (vz[gidx(x, y, z + 1, n)] - vz[gidx(x, y, z - 1, n)]) * (0.5 * h_inv);
                double screening = point_weight * screen[idx] * h_inv2;
                rhs[idx] = -div + screening * target_value;
            }
}

/* capi.mojo:282-306: Neumann 7-point neighbour sum (order x-1, x+1, y-1, y+1, z-1, z+1)
 * and the number of in-grid neighbours. */
static inline double stencil_sum(const double *v, int64_t x, int64_t y, int64_t z, int64_t n,
                                 int *degree) {
    double s = 0.0;
    int d = 0;
    if (x > 0) { s += v[gidx(x - 1, y, z, n)]; d++; }
    if (x + 1 < n) { s += v[gidx(x + 1, y, z, n)]; d++; }
    if (y > 0) { s += v[gidx(x, y - 1, z, n)]; d++; }
    if (y + 1 < n) { s += v[gidx(x, y + 1, z, n)]; d++; }
    if (z > 0) { s += v[gidx(x, y, z - 1, n)]; d++; }
    if (z + 1 < n) { s += v[gidx(x, y, z + 1, n)]; d++; }
    *degree = d;
    return s;
}

/* capi.mojo:309-410. Within one colour no updated node reads another updated node,
 * so the sweep order inside a colour does not change any result. The diagonal guard
 * (capi.mojo:355,398) is applied at every node, as in tests/test_kernels.py:158. */
void oracle_relax(double *solution, const double *rhs, const double *screen, int64_t n,
                  double point_weight, int64_t iterations) {
    double h_inv = (double)(n - 1);
    double h_inv2 = h_inv * h_inv;
    for (int64_t it = 0; it < iterations; ++it)
        for (int64_t color = 0; color < 2; ++color) {
            ORACLE_PARALLEL_FOR
            for (int64_t z = 0; z < n; ++z)
                for (int64_t y = 0; y < n; ++y)
                    for (int64_t x = 0; x < n; ++x) {
                        if (((x + y + z) & 1) != color) continue;
                        int64_t idx = gidx(x, y, z, n);
                        int deg;
                        double nsum = stencil_sum(solution, x, y, z, n, &deg);
                        double screening = point_weight * screen[idx] * h_inv2;
                        double diag = (double)deg * h_inv2 + screening;
                        if (diag > 1.0e-30) solution[idx] = (rhs[idx] + nsum * h_inv2) / diag;
                    }
        }
}

/* capi.mojo:413-553. The norm sums squares in index order (the reference uses
 * SIMD-lane partial sums, so the norm agrees to rounding, not bit for bit). */
double oracle_residual(const double *solution, const double *rhs, const double *screen,
                       double *residual, int64_t n, double point_weight) {
    double h_inv = (double)(n - 1);
    double h_inv2 = h_inv * h_inv;
    ORACLE_PARALLEL_FOR
    for (int64_t z = 0; z < n; ++z)
        for (int64_t y = 0; y < n; ++y)
            for (int64_t x = 0; x < n; ++x) {
                int64_t idx = gidx(x, y, z, n);
                int deg;
                double nsum = stencil_sum(solution, x, y, z, n, &deg);
                double screening = point_weight * screen[idx] * h_inv2;
                double applied = ((double)deg * h_inv2 + screening) * solution[idx] - nsum * h_inv2;
                residual[idx] = rhs[idx] - applied;
            }
    /* capi.mojo:543-553: W lane partial sums, halving lane reduction, serial tail.
     * W = 4 (the float64 SIMD width of the AVX2 machine the reference's numbers were
     * taken on); the norm is compared with a tolerance everywhere. */
    int64_t size = n * n * n;
    int64_t vend = size / 4 * 4;
    double lane[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = 0; i < vend; i += 4)
        for (int l = 0; l < 4; ++l) lane[l] += residual[i + l] * residual[i + l];
    double sq = (lane[0] + lane[2]) + (lane[1] + lane[3]);
    for (int64_t i = vend; i < size; ++i) sq += residual[i] * residual[i];
    return sqrt(sq);
}

/* capi.mojo:556-673: full weighting, weights 2 (centre) / 1 per axis, renormalised by
 * the in-grid weight sum; terms accumulate in (oz, oy, ox) order. */
void oracle_restrict(const double *fine, double *coarse, int64_t fine_n) {
    int64_t cn = (fine_n + 1) / 2;
    ORACLE_PARALLEL_FOR
    for (int64_t z = 0; z < cn; ++z)
        for (int64_t y = 0; y < cn; ++y)
            for (int64_t x = 0; x < cn; ++x) {
                double sum = 0.0, wsum = 0.0;
                for (int oz = -1; oz <= 1; ++oz) {
                    int64_t zz = 2 * z + oz;
                    if (zz < 0 || zz >= fine_n) continue;
                    double wz = oz == 0 ? 2.0 : 1.0;
                    for (int oy = -1; oy <= 1; ++oy) {
                        int64_t yy = 2 * y + oy;
                        if (yy < 0 || yy >= fine_n) continue;
                        double wy = oy == 0 ? 2.0 : 1.0;
                        for (int ox = -1; ox <= 1; ++ox) {
                            int64_t xx = 2 * x + ox;
                            if (xx < 0 || xx >= fine_n) continue;
                            double wx = ox == 0 ? 2.0 : 1.0;
                            double w = wx * wy * wz;
                            sum += fine[gidx(xx, yy, zz, fine_n)] * w;
                            wsum += w;
                        }
                    }
                }
                coarse[gidx(x, y, z, cn)] = sum / wsum;
            }
}

/* capi.mojo:676-742 */
void oracle_prolong_add(const double *coarse, double *fine, int64_t fine_n) {
    int64_t cn = (fine_n + 1) / 2;
    ORACLE_PARALLEL_FOR
    for (int64_t z = 0; z < fine_n; ++z) {
        int64_t z0 = z / 2, z1 = z0 + 1 < cn - 1 ? z0 + 1 : cn - 1;
        double tz = (z & 1) ? 0.5 : 0.0;
        for (int64_t y = 0; y < fine_n; ++y) {
            int64_t y0 = y / 2, y1 = y0 + 1 < cn - 1 ? y0 + 1 : cn - 1;
            double ty = (y & 1) ? 0.5 : 0.0;
            for (int64_t x = 0; x < fine_n; ++x) {
                int64_t x0 = x / 2, x1 = x0 + 1 < cn - 1 ? x0 + 1 : cn - 1;
                double tx = (x & 1) ? 0.5 : 0.0;
                double c000 = coarse[gidx(x0, y0, z0, cn)], c100 = coarse[gidx(x1, y0, z0, cn)];
                double c010 = coarse[gidx(x0, y1, z0, cn)], c110 = coarse[gidx(x1, y1, z0, cn)];
                double c001 = coarse[gidx(x0, y0, z1, cn)], c101 = coarse[gidx(x1, y0, z1, cn)];
                double c011 = coarse[gidx(x0, y1, z1, cn)], c111 = coarse[gidx(x1, y1, z1, cn)];
                double c00 = c000 * (1.0 - tx) + c100 * tx;
                double c10 = c010 * (1.0 - tx) + c110 * tx;
                double c01 = c001 * (1.0 - tx) + c101 * tx;
                double c11 = c011 * (1.0 - tx) + c111 * tx;
                double c0 = c00 * (1.0 - ty) + c10 * ty;
                double c1 = c01 * (1.0 - ty) + c11 * ty;
                fine[gidx(x, y, z, fine_n)] += c0 * (1.0 - tz) + c1 * tz;
            }
        }
    }
}

/* capi.mojo:745-795 */
double oracle_iso_value(const double *field, const double *points, const double *normals,
                        int64_t point_count, int64_t n, int64_t confidence) {
    int64_t res = n - 1;
    double vsum = 0.0, wsum = 0.0;
    for (int64_t p = 0; p < point_count; ++p) {
        double weight = confidence != 0
                            ? oracle_norm3(normals[3 * p], normals[3 * p + 1], normals[3 * p + 2])
                            : 1.0;
        double gx = points[3 * p] * (double)res;
        double gy = points[3 * p + 1] * (double)res;
        double gz = points[3 * p + 2] * (double)res;
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
            double wz = dz == 1 ? fz : 1.0 - fz;
            for (int dy = 0; dy < 2; ++dy) {
                double wy = dy == 1 ? fy : 1.0 - fy;
                for (int dx = 0; dx < 2; ++dx) {
                    double wx = dx == 1 ? fx : 1.0 - fx;
                    value += field[gidx(ix + dx, iy + dy, iz + dz, n)] * wx * wy * wz;
                }
            }
        }
        vsum += value * weight;
        wsum += weight;
    }
    return wsum > 0.0 ? vsum / wsum : 0.0;
}

static inline double corner_value(const double *f, int64_t x, int64_t y, int64_t z, int64_t n,
                                  int c) {
    return f[gidx(x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1), n)]; /* capi.mojo:798-809 */
}

static inline int cell_active(const double *f, int64_t x, int64_t y, int64_t z, int64_t n,
                              double iso) {
    int below = 0;
    for (int c = 0; c < 8; ++c) below += corner_value(f, x, y, z, n, c) < iso;
    return below != 0 && below != 8;
}

/* capi.mojo:839-1041. Count mode (capacity <= 0) writes one partial count per
 * 4-plane z chunk into cell_map[0 .. ceil(cells/4)) (capi.mojo:855-903). */
int64_t oracle_dual_vertices(const double *field, const double *density, int64_t n,
                             double iso, int64_t *cell_map, double *vertices,
                             double *vertex_density, int64_t capacity) {
    int64_t cells = n - 1;
    int64_t count = 0;
    if (capacity <= 0) {
        int64_t tasks = (cells + 3) / 4;
        for (int64_t t = 0; t < tasks; ++t) {
            int64_t local = 0;
            int64_t z1 = t * 4 + 4 < cells ? t * 4 + 4 : cells;
            for (int64_t z = t * 4; z < z1; ++z)
                for (int64_t y = 0; y < cells; ++y)
                    for (int64_t x = 0; x < cells; ++x) local += cell_active(field, x, y, z, n, iso);
            cell_map[t] = local;
            count += local;
        }
        return count;
    }
    for (int64_t z = 0; z < cells; ++z)
        for (int64_t y = 0; y < cells; ++y)
            for (int64_t x = 0; x < cells; ++x)
                cell_map[(z * cells + y) * cells + x] = cell_active(field, x, y, z, n, iso) ? 0 : -1;
    for (int64_t z = 0; z < cells; ++z)
        for (int64_t y = 0; y < cells; ++y)
            for (int64_t x = 0; x < cells; ++x) {
                int64_t ci = (z * cells + y) * cells + x;
                if (cell_map[ci] < 0) continue;
                double v[8];
                for (int c = 0; c < 8; ++c) v[c] = corner_value(field, x, y, z, n, c);
                if (count >= capacity) { count += 1; continue; }
                double sx = 0.0, sy = 0.0, sz = 0.0;
                int64_t roots = 0;
                for (int axis = 0; axis < 3; ++axis)
                    for (int a = 0; a < 2; ++a)
                        for (int b = 0; b < 2; ++b) {
                            int c0, c1;
                            double x0, y0, z0, x1, y1, z1;
                            if (axis == 0) {
                                c0 = 2 * a + 4 * b; c1 = c0 + 1;
                                x0 = 0.0; y0 = (double)a; z0 = (double)b;
                                x1 = 1.0; y1 = y0; z1 = z0;
                            } else if (axis == 1) {
                                c0 = a + 4 * b; c1 = c0 + 2;
                                x0 = (double)a; y0 = 0.0; z0 = (double)b;
                                x1 = x0; y1 = 1.0; z1 = z0;
                            } else {
                                c0 = a + 2 * b; c1 = c0 + 4;
                                x0 = (double)a; y0 = (double)b; z0 = 0.0;
                                x1 = x0; y1 = y0; z1 = 1.0;
                            }
                            /* capi.mojo:812-836 */
                            double v0 = v[c0], v1 = v[c1];
                            if ((v0 < iso) == (v1 < iso)) continue;
                            double t = (iso - v0) / (v1 - v0);
                            t = t < 0.0 ? 0.0 : t;
                            t = t > 1.0 ? 1.0 : t;
                            sx = sx + x0 + (x1 - x0) * t;
                            sy = sy + y0 + (y1 - y0) * t;
                            sz = sz + z0 + (z1 - z0) * t;
                            roots += 1;
                        }
                double inv = 1.0 / (double)roots;
                vertices[3 * count] = ((double)x + sx * inv) / (double)cells;
                vertices[3 * count + 1] = ((double)y + sy * inv) / (double)cells;
                vertices[3 * count + 2] = ((double)z + sz * inv) / (double)cells;
                double dsum = 0.0;
                for (int c = 0; c < 8; ++c) dsum += corner_value(density, x, y, z, n, c);
                vertex_density[count] = dsum * 0.125;
                cell_map[ci] = count;
                count += 1;
            }
    return count;
}

/* capi.mojo:1044-1072 */
static inline int64_t emit_quad(int64_t *faces, int64_t fc, int64_t cap, int64_t a, int64_t b,
                                int64_t c, int64_t d, int reverse) {
    if (a < 0 || b < 0 || c < 0 || d < 0) return fc;
    if (fc + 2 <= cap) {
        int64_t *f = faces + 3 * fc;
        if (reverse) { f[0] = a; f[1] = d; f[2] = c; f[3] = a; f[4] = c; f[5] = b; }
        else { f[0] = a; f[1] = b; f[2] = c; f[3] = a; f[4] = c; f[5] = d; }
    }
    return fc + 2;
}

/* capi.mojo:1075-1129 */
int64_t oracle_dual_faces(const double *field, int64_t n, double iso, const int64_t *cell_map,
                          int64_t *faces, int64_t capacity) {
    int64_t cells = n - 1, fc = 0;
#define CM(x, y, z) cell_map[((z) * cells + (y)) * cells + (x)]
    for (int64_t z = 1; z < cells; ++z)
        for (int64_t y = 1; y < cells; ++y)
            for (int64_t x = 0; x < cells; ++x) {
                double v0 = field[gidx(x, y, z, n)], v1 = field[gidx(x + 1, y, z, n)];
                if ((v0 < iso) != (v1 < iso))
                    fc = emit_quad(faces, fc, capacity, CM(x, y - 1, z - 1), CM(x, y, z - 1),
                                   CM(x, y, z), CM(x, y - 1, z), v0 < iso);
            }
    for (int64_t z = 1; z < cells; ++z)
        for (int64_t y = 0; y < cells; ++y)
            for (int64_t x = 1; x < cells; ++x) {
                double v0 = field[gidx(x, y, z, n)], v1 = field[gidx(x, y + 1, z, n)];
                if ((v0 < iso) != (v1 < iso))
                    fc = emit_quad(faces, fc, capacity, CM(x - 1, y, z - 1), CM(x - 1, y, z),
                                   CM(x, y, z), CM(x, y, z - 1), v0 < iso);
            }
    for (int64_t z = 0; z < cells; ++z)
        for (int64_t y = 1; y < cells; ++y)
            for (int64_t x = 1; x < cells; ++x) {
                double v0 = field[gidx(x, y, z, n)], v1 = field[gidx(x, y, z + 1, n)];
                if ((v0 < iso) != (v1 < iso))
                    fc = emit_quad(faces, fc, capacity, CM(x - 1, y - 1, z), CM(x, y - 1, z),
                                   CM(x, y, z), CM(x - 1, y, z), v0 < iso);
            }
#undef CM
    return fc;
}

/* core.py:429-450: per-face loop. cross = (b - a) x (c - a) as numpy.cross computes it;
 * a face counts when its cross product is finite with a positive squared norm (np.dot);
 * its cross product is added to its three corners in face order; a vertex is kept when it is
 * referenced and the norm of its sum (numpy.linalg.norm: sqrt((x^2 + y^2) + z^2)) is finite
 * and positive, and its normal is the sum divided by that norm. Kept vertices keep their
 * order. Returns their count. */
int64_t oracle_oriented_points(const double *vertices, int64_t vertex_count, const int64_t *faces,
                               int64_t face_count, double *points_out, double *normals_out) {
    double *sum = calloc((size_t)(3 * vertex_count + 1), sizeof(double));
    unsigned char *referenced = calloc((size_t)(vertex_count + 1), 1);
    for (int64_t f = 0; f < face_count; ++f) {
        int64_t a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
        double e1[3], e2[3], cr[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = vertices[3 * b + k] - vertices[3 * a + k];
            e2[k] = vertices[3 * c + k] - vertices[3 * a + k];
        }
        cr[0] = e1[1] * e2[2] - e1[2] * e2[1];
        cr[1] = e1[2] * e2[0] - e1[0] * e2[2];
        cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
        if (!(isfinite(cr[0]) && isfinite(cr[1]) && isfinite(cr[2]))) continue;
        if (!(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2] > 0.0)) continue;
        int64_t corner[3] = {a, b, c};
        for (int j = 0; j < 3; ++j) {
            for (int k = 0; k < 3; ++k) sum[3 * corner[j] + k] += cr[k];
            referenced[corner[j]] = 1;
        }
    }
    int64_t kept = 0;
    for (int64_t i = 0; i < vertex_count; ++i) {
        double x = sum[3 * i], y = sum[3 * i + 1], z = sum[3 * i + 2];
        double len = sqrt((x * x + y * y) + z * z);
        if (!referenced[i] || !(len > 0.0) || !isfinite(len)) continue;
        for (int k = 0; k < 3; ++k) {
            points_out[3 * kept + k] = vertices[3 * i + k];
            normals_out[3 * kept + k] = sum[3 * i + k] / len;
        }
        kept += 1;
    }
    free(sum);
    free(referenced);
    return kept;
}

/* core.py:141-193: correction-scheme V-cycle, coarsest (n <= 3) gets 40 iterations. */
static void v_cycle(int level, int levels, double **sol, double **rhs, double **scr, double **res,
                    const int64_t *ns, double pw, int pre, int post) {
    int64_t n = ns[level];
    if (n <= 3) {
        oracle_relax(sol[level], rhs[level], scr[level], n, pw, MUTATED(7) ? 20 : 40);
        return;
    }
    oracle_relax(sol[level], rhs[level], scr[level], n, pw, MUTATED(6) ? pre - 1 : pre);
    oracle_residual(sol[level], rhs[level], scr[level], res[level], n, pw);
    oracle_restrict(MUTATED(9) ? sol[level] : res[level], rhs[level + 1], n);
    int64_t cn = ns[level + 1];
    if (!MUTATED(3)) memset(sol[level + 1], 0, sizeof(double) * (size_t)(cn * cn * cn));
    v_cycle(level + 1, levels, sol, rhs, scr, res, ns, pw, pre, post);
    if (MUTATED(4)) {
        oracle_relax(sol[level], rhs[level], scr[level], n, pw, post);
        oracle_prolong_add(sol[level + 1], sol[level], n);
    } else {
        oracle_prolong_add(sol[level + 1], sol[level], n);
        oracle_relax(sol[level], rhs[level], scr[level], n, pw, post);
    }
    (void)levels;
}

/* core.py:229-306 (from the grid allocation on; validation and the octree live in the
 * Python side of the oracle). */
int64_t oracle_solve(const double *points, const double *normals, int64_t point_count,
                     int64_t depth, double point_weight, int64_t confidence, int64_t cycles,
                     double samples_per_node, double *field, double *density, double *screen,
                     double *rhs, double *iso_value, double *residual_norm) {
    int64_t n = ((int64_t)1 << depth) + 1;
    size_t m = (size_t)(n * n * n);
    double *vx = calloc(m, sizeof(double)), *vy = calloc(m, sizeof(double)),
           *vz = calloc(m, sizeof(double));
    memset(density, 0, m * sizeof(double));
    memset(screen, 0, m * sizeof(double));
    memset(rhs, 0, m * sizeof(double));
    int64_t accepted = 0;
    if (point_count) {
        accepted = oracle_splat(points, normals, point_count, n, confidence, density, vx, vy, vz,
                                screen, samples_per_node);
        if (!MUTATED(5))
            for (size_t i = 0; i < m; ++i) density[i] /= samples_per_node;
        oracle_assemble(vx, vy, vz, screen, rhs, n, point_weight, MUTATED(8) ? 1.0 : 0.5);
    }
    free(vx); free(vy); free(vz);

    int levels = 1;
    int64_t ns[32];
    ns[0] = n;
    while (ns[levels - 1] > 3) { ns[levels] = (ns[levels - 1] + 1) / 2; levels++; }
    double *sol[32], *rh[32], *sc[32], *rs[32];
    for (int l = 0; l < levels; ++l) {
        size_t ml = (size_t)(ns[l] * ns[l] * ns[l]);
        sol[l] = malloc(ml * sizeof(double));
        rh[l] = calloc(ml, sizeof(double));
        rs[l] = malloc(ml * sizeof(double));
        sc[l] = l == 0 ? screen : malloc(ml * sizeof(double));
        for (size_t i = 0; i < ml; ++i) sol[l][i] = MUTATED(2) ? 0.0 : 0.5;
        if (l > 0) {
            oracle_restrict(sc[l - 1], sc[l], ns[l - 1]);
            for (size_t i = 0; i < ml; ++i) sc[l][i] *= MUTATED(1) ? 2.0 : 4.0; /* core.py:267-269 */
        }
    }
    memcpy(rh[0], rhs, m * sizeof(double));
    if (point_count) {
        for (int64_t c = 0; c < cycles; ++c)
            v_cycle(0, levels, sol, rh, sc, rs, ns, point_weight, 3, 3);
    } else {
        memset(sol[0], 0, m * sizeof(double));
    }
    *residual_norm = oracle_residual(sol[0], rhs, screen, rs[0], n, point_weight);
    *iso_value = point_count ? oracle_iso_value(sol[0], points, normals, point_count, n, confidence)
                             : 0.0;
    memcpy(field, sol[0], m * sizeof(double));
    for (int l = 0; l < levels; ++l) {
        free(sol[l]); free(rh[l]); free(rs[l]);
        if (l > 0) free(sc[l]);
    }
    return accepted;
}
