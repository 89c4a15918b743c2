// Per-node element functions of the multigrid solver on the colour-split layout.
//
// Each function computes one node exactly as the reference kernel does (same operand
// order, no contraction: the library is compiled with -ffp-contract=off), so the FP64
// instantiation is bit-identical to src/capi.mojo on the same inputs. The multi-block
// kernels and the single-workgroup coarse-level kernel both call these.
#pragma once

#include "mpr_common.h"

namespace mpr {

// Neumann 7-point neighbour sum in the order x-1, x+1, y-1, y+1, z-1, z+1 and the number
// of in-grid neighbours (capi.mojo:282-306). All neighbours of (x,y,z) have the other
// colour, so with the split layout they live in colour block 1-c.
template <class T>
MPR_HD T neighbour_sum(const T* X, const Geom& g, int x, int y, int z, int& deg) {
  const int n = g.n;
  T s = T(0);
  int d = 0;
  if (x > 0) { s += X[split_index(g, x - 1, y, z)]; ++d; }
  if (x + 1 < n) { s += X[split_index(g, x + 1, y, z)]; ++d; }
  if (y > 0) { s += X[split_index(g, x, y - 1, z)]; ++d; }
  if (y + 1 < n) { s += X[split_index(g, x, y + 1, z)]; ++d; }
  if (z > 0) { s += X[split_index(g, x, y, z - 1)]; ++d; }
  if (z + 1 < n) { s += X[split_index(g, x, y, z + 1)]; ++d; }
  deg = d;
  return s;
}

// Red-black Gauss-Seidel update of node idx (capi.mojo:343-401):
// x = (rhs + h^-2 * sum) / (deg h^-2 + (alpha * screen) h^-2), skipped if the diagonal
// is not positive (the reference's scalar-path guard, applied at every node).
template <class T>
MPR_HD void relax_node(T* X, const T* R, const T* S, const Geom& g, uint32_t idx, int x, int y,
                       int z, T pw, T h2) {
  int deg;
  T s = neighbour_sum(X, g, x, y, z, deg);
  T screening = pw * S[idx] * h2;
  T diag = (T)deg * h2 + screening;
  if (diag > T(1.0e-30)) X[idx] = (R[idx] + s * h2) / diag;
}

// r = rhs - A x (capi.mojo:413-470)
template <class T>
MPR_HD T residual_node(const T* X, const T* R, const T* S, const Geom& g, uint32_t idx, int x,
                       int y, int z, T pw, T h2) {
  int deg;
  T s = neighbour_sum(X, g, x, y, z, deg);
  T screening = pw * S[idx] * h2;
  T applied = ((T)deg * h2 + screening) * X[idx] - s * h2;
  return R[idx] - applied;
}

// Full-weighting restriction of coarse node (X,Y,Z) (capi.mojo:556-673): weights 2 at the
// centre and 1 at the sides per axis, terms in (oz, oy, ox) order, divided by the sum of
// the in-grid weights, then multiplied by `scale` (1 for residuals, 4 for the screening
// pyramid of core.py:267-269; both scalings are exact).
template <class T>
MPR_HD T restrict_node(const T* F, const Geom& gf, int X, int Y, int Z, T scale) {
  const int nf = gf.n;
  T sum = T(0), wsum = T(0);
  for (int oz = -1; oz <= 1; ++oz) {
    int zz = 2 * Z + oz;
    if (zz < 0 || zz >= nf) continue;
    T wz = oz == 0 ? T(2) : T(1);
    for (int oy = -1; oy <= 1; ++oy) {
      int yy = 2 * Y + oy;
      if (yy < 0 || yy >= nf) continue;
      T wy = oy == 0 ? T(2) : T(1);
      for (int ox = -1; ox <= 1; ++ox) {
        int xx = 2 * X + ox;
        if (xx < 0 || xx >= nf) continue;
        T wx = ox == 0 ? T(2) : T(1);
        T w = wx * wy * wz;
        sum += F[split_index(gf, xx, yy, zz)] * w;
        wsum += w;
      }
    }
  }
  return sum / wsum * scale;
}

// Trilinear prolongation value at fine node (x,y,z) (capi.mojo:676-742).
template <class T>
MPR_HD T prolong_node(const T* Cg, const Geom& gc, int x, int y, int z) {
  const int nc = gc.n;
  int x0 = x >> 1, y0 = y >> 1, z0 = z >> 1;
  int x1 = x0 + 1 < nc - 1 ? x0 + 1 : nc - 1;
  int y1 = y0 + 1 < nc - 1 ? y0 + 1 : nc - 1;
  int z1 = z0 + 1 < nc - 1 ? z0 + 1 : nc - 1;
  T tx = (x & 1) ? T(0.5) : T(0), ty = (y & 1) ? T(0.5) : T(0), tz = (z & 1) ? T(0.5) : T(0);
  T c000 = Cg[split_index(gc, x0, y0, z0)], c100 = Cg[split_index(gc, x1, y0, z0)];
  T c010 = Cg[split_index(gc, x0, y1, z0)], c110 = Cg[split_index(gc, x1, y1, z0)];
  T c001 = Cg[split_index(gc, x0, y0, z1)], c101 = Cg[split_index(gc, x1, y0, z1)];
  T c011 = Cg[split_index(gc, x0, y1, z1)], c111 = Cg[split_index(gc, x1, y1, z1)];
  T c00 = c000 * (T(1) - tx) + c100 * tx;
  T c10 = c010 * (T(1) - tx) + c110 * tx;
  T c01 = c001 * (T(1) - tx) + c101 * tx;
  T c11 = c011 * (T(1) - tx) + c111 * tx;
  T c0 = c00 * (T(1) - ty) + c10 * ty;
  T c1 = c01 * (T(1) - ty) + c11 * ty;
  return c0 * (T(1) - tz) + c1 * tz;
}

}  // namespace mpr
