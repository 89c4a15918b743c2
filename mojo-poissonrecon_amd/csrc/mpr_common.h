// Shared definitions for the MI355X (gfx950) screened Poisson library.
//
// Grid layouts in HBM
// -------------------
// * natural: (n,n,n) row-major [z][y][x], index (z*n + y)*n + x — the layout of every grid
//   that crosses the API (capi.mojo:41-48).
// * colour-split: the solver's layout. Node (x,y,z) has colour c = (x+y+z)&1. Each colour
//   owns a block of C = n*n*H words, H >= (n+1)/2, holding row (z,y) of that colour as
//   H consecutive words, node x at slot x>>1; the slots past the row's last node are
//   unused pads (one for rows whose first node of colour c is at x = 1, plus the
//   alignment padding of make_geom). A red-black sweep of colour c then reads
//   and writes colour c contiguously and reads colour 1-c contiguously: no half-used
//   cache lines, so a full Gauss-Seidel iteration moves 4 words per node.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#define MPR_HD __host__ __device__ __forceinline__

namespace mpr {

// ---------------------------------------------------------------------------------------
// error plumbing: no C++ exception crosses the C ABI; every entry point returns a status
// and records a message retrievable through mpr_last_error().
void set_error(const std::string& msg);
const char* last_error();

struct HipError {
  std::string msg;
};

#define MPR_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t _e = (call);                                                            \
    if (_e != hipSuccess)                                                              \
      throw ::mpr::HipError{std::string(#call) + ": " + hipGetErrorString(_e) + " (" + \
                            __FILE__ + ":" + std::to_string(__LINE__) + ")"};          \
  } while (0)

#define MPR_LAUNCH_CHECK() MPR_HIP(hipGetLastError())

// ---------------------------------------------------------------------------------------
// unsigned 32-bit division by a run-time invariant divisor (multiply-high + shift,
// Granlund & Montgomery 1994), used to decode flattened thread indices.
struct FastDiv {
  uint32_t d, m, s;
  MPR_HD constexpr FastDiv() : d(1), m(1), s(0) {}
  explicit FastDiv(uint32_t div) : d(div) {
    uint32_t l = 0;
    while ((1ull << l) < div) ++l;
    s = l;
    m = (uint32_t)((((1ull << 32) * ((1ull << l) - div)) / div) + 1);
  }
  MPR_HD uint32_t div(uint32_t t) const {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t hi = __umulhi(t, m);
#else
    uint32_t hi = (uint32_t)(((uint64_t)t * m) >> 32);
#endif
    return (uint32_t)(((uint64_t)hi + t) >> s);
  }
};

// geometry of one grid level in the colour-split layout
struct Geom {
  int32_t n = 0;     // nodes per side
  int32_t H = 0;     // slots per colour row, >= (n+1)/2 (make_geom)
  uint32_t C = 0;    // words per colour block, n*n*H
  uint32_t nH = 0;   // n*H
  FastDiv div_nH, div_H, div_C;
};

// Rows of levels with n >= 33 are padded to a multiple of 4 slots, so every colour row
// starts on a 16-byte boundary (FP32 and FP64) and the streaming kernels read whole 16-byte
// chunks (group_layout). Smaller levels, including every level the LDS coarse solve holds,
// keep H = (n+1)/2.
constexpr int32_t kGroupMinN = 33;

inline Geom make_geom(int32_t n) {
  Geom g;
  g.n = n;
  g.H = (n + 1) / 2;
  if (n >= kGroupMinN) g.H = (g.H + 3) & ~3;
  g.nH = (uint32_t)n * (uint32_t)g.H;
  g.C = g.nH * (uint32_t)n;
  g.div_nH = FastDiv(g.nH);
  g.div_H = FastDiv((uint32_t)g.H);
  g.div_C = FastDiv(g.C);
  return g;
}

// whether a level's rows are 16-byte aligned chunks (the slot-group kernels' precondition)
MPR_HD bool group_layout(const Geom& g) { return (g.H & 3) == 0; }

MPR_HD int64_t nat_index(int x, int y, int z, int n) { return ((int64_t)z * n + y) * n + x; }

MPR_HD uint32_t split_index(const Geom& g, int x, int y, int z) {
  uint32_t c = (uint32_t)((x + y + z) & 1);
  return c * g.C + ((uint32_t)z * (uint32_t)g.n + (uint32_t)y) * (uint32_t)g.H + (uint32_t)(x >> 1);
}

// Decode an index inside colour block c into (x,y,z); returns false for a pad slot.
MPR_HD bool split_decode(const Geom& g, int c, uint32_t t, int& x, int& y, int& z) {
  uint32_t zz = g.div_nH.div(t);
  uint32_t r = t - zz * g.nH;
  uint32_t yy = g.div_H.div(r);
  uint32_t k = r - yy * (uint32_t)g.H;
  int p = (c ^ (int)yy ^ (int)zz) & 1;
  x = 2 * (int)k + p;
  y = (int)yy;
  z = (int)zz;
  return x < g.n;
}

// A scalar n^3 grid in either layout: natural [z][y][x] (split == nullptr) or colour-split.
struct FieldView {
  const double* nat = nullptr;
  const double* split = nullptr;
  Geom g;  // g.n is the side in both layouts; the rest describes the split layout
  MPR_HD double at(int x, int y, int z) const {
    return split ? split[split_index(g, x, y, z)] : nat[nat_index(x, y, z, g.n)];
  }
};

// The density grid (natural layout) as the extraction reads it. With `touched` it is the
// splat's raw grid: nodes whose bit is unset read as zero (the grid was not zero-filled) and
// every value is divided by samples_per_node, the exact value of core.py:255's in-place
// `density /= samples_per_node`. Without it, the values are read as stored.
struct DensityView {
  const double* p = nullptr;
  const uint32_t* touched = nullptr;
  double spn = 1.0;
  MPR_HD double at(uint64_t i) const {
    if (!touched) return p[i];
    const double raw = ((touched[i >> 5] >> (i & 31u)) & 1u) ? p[i] : 0.0;
    return raw / spn;
  }
};

inline FieldView natural_field(const double* f, int n) {
  FieldView v;
  v.nat = f;
  v.g.n = n;
  return v;
}

inline FieldView split_field(const double* f, const Geom& g) {
  FieldView v;
  v.split = f;
  v.g = g;
  return v;
}

inline unsigned blocks_for(uint64_t work, unsigned threads) {
  return (unsigned)((work + threads - 1) / threads);
}

// Pointer into global memory (address space 1). Buffers reached through a device pointer
// table are generic pointers to the compiler; re-typing them lets it emit global_load/store
// instead of the slower generic flat_load/store.
template <class T>
using gptr = __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ gptr<T> as_global(T* p) {
  return (gptr<T>)p;
}

constexpr int kNodesPerThread = 2;  // streaming kernels: nodes per lane (loads hoisted)

}  // namespace mpr
