/*
 * TEST INFRASTRUCTURE ONLY. A driver that runs every oracle function on small inputs, for a
 * build with AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_oracle_sanitize.py
 * compiles it together with mpr_oracle.c and runs it). Odd and even grid sizes, the count
 * and fill protocols of the dual contouring and the whole solve are exercised; the process
 * exits non-zero on any sanitizer report.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "mpr_oracle.h"

static uint64_t rng = 0x9e3779b97f4a7c15ull;

static double uniform(void) {  /* xorshift64*, in [0, 1) */
  rng ^= rng >> 12;
  rng ^= rng << 25;
  rng ^= rng >> 27;
  return (double)((rng * 0x2545f4914f6cdd1dull) >> 11) * (1.0 / 9007199254740992.0);
}

static double *grid(int64_t n, int random) {
  double *g = calloc((size_t)(n * n * n), sizeof(double));
  if (!g) exit(2);
  if (random)
    for (int64_t i = 0; i < n * n * n; ++i) g[i] = uniform() - 0.5;
  return g;
}

int main(void) {
  /* points on a sphere of radius 0.3 about the cube centre, unit normals; one zero normal */
  const int64_t P = 400;
  double *pts = malloc(3 * P * sizeof(double)), *nrm = malloc(3 * P * sizeof(double));
  for (int64_t i = 0; i < P; ++i) {
    const double z = 2.0 * uniform() - 1.0, t = 6.283185307179586 * uniform(), r = sqrt(1.0 - z * z);
    const double d[3] = {r * cos(t), r * sin(t), z};
    for (int a = 0; a < 3; ++a) {
      pts[3 * i + a] = 0.5 + 0.3 * d[a];
      nrm[3 * i + a] = i == 7 ? 0.0 : d[a];
    }
  }
  int64_t *codes = malloc((size_t)(P * 6) * sizeof(int64_t));
  oracle_octree_codes(pts, P, 5, codes);

  const int64_t depth = 4, n = 17;
  double *field = grid(n, 0), *density = grid(n, 0), *screen = grid(n, 0), *rhs = grid(n, 0);
  double iso = 0.0, res = 0.0;
  const int64_t accepted = oracle_solve(pts, nrm, P, depth, 4.0, 1, 2, 1.5, field, density,
                                        screen, rhs, &iso, &res);
  const double iso2 = oracle_iso_value(field, pts, nrm, P, n, 1);

  /* dual contouring: count mode, then fill with the counted capacity */
  const int64_t cells = n - 1, cells3 = cells * cells * cells;
  int64_t *cell_map = malloc((size_t)cells3 * sizeof(int64_t));
  const int64_t nv = oracle_dual_vertices(field, density, n, iso, cell_map, NULL, NULL, 0);
  double *verts = malloc((size_t)(3 * (nv > 0 ? nv : 1)) * sizeof(double));
  double *vdens = malloc((size_t)(nv > 0 ? nv : 1) * sizeof(double));
  const int64_t nv2 = oracle_dual_vertices(field, density, n, iso, cell_map, verts, vdens, nv);
  const int64_t nf = oracle_dual_faces(field, n, iso, cell_map, NULL, 0);
  int64_t *faces = malloc((size_t)(3 * (nf > 0 ? nf : 1)) * sizeof(int64_t));
  const int64_t nf2 = oracle_dual_faces(field, n, iso, cell_map, faces, nf);

  /* grid kernels on odd and even sizes */
  double checksum = 0.0;
  const int64_t sizes[] = {3, 4, 9, 10};
  for (int k = 0; k < 4; ++k) {
    const int64_t m = sizes[k], mc = (m + 1) / 2;
    double *x = grid(m, 1), *b = grid(m, 1), *s = grid(m, 1), *r = grid(m, 0);
    for (int64_t i = 0; i < m * m * m; ++i) s[i] = fabs(s[i]);
    double *c = grid(mc, 0);
    oracle_relax(x, b, s, m, 4.0, 2);
    checksum += oracle_residual(x, b, s, r, m, 4.0);
    oracle_restrict(r, c, m);
    oracle_prolong_add(c, x, m);
    double *vx = grid(m, 1), *vy = grid(m, 1), *vz = grid(m, 1);
    oracle_assemble(vx, vy, vz, s, b, m, 4.0, 0.5);
    for (int64_t i = 0; i < m * m * m; ++i) checksum += x[i] + b[i];
    free(x); free(b); free(s); free(r); free(c); free(vx); free(vy); free(vz);
  }

  /* mesh normals of the extracted mesh (core.py:429-450) */
  double *opts = malloc((size_t)(3 * (nv > 0 ? nv : 1)) * sizeof(double));
  double *onrm = malloc((size_t)(3 * (nv > 0 ? nv : 1)) * sizeof(double));
  const int64_t kept = oracle_oriented_points(verts, nv, faces, nf, opts, onrm);
  checksum += (double)kept;
  free(opts); free(onrm);

  printf("sanitize ok: accepted %lld, iso %.6f (%.6f), residual %.3f, vertices %lld/%lld, "
         "faces %lld/%lld, checksum %.6g\n",
         (long long)accepted, iso, iso2, res, (long long)nv, (long long)nv2, (long long)nf,
         (long long)nf2, checksum);
  free(pts); free(nrm); free(codes); free(field); free(density); free(screen); free(rhs);
  free(cell_map); free(verts); free(vdens); free(faces);
  return nv == nv2 && nf == nf2 && accepted == P - 1 ? 0 : 1;
}
