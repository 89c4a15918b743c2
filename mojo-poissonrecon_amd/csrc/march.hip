// Fused red-black smoothing: S colour sweeps of one FP32 level in one pass over the grid
// (temporal blocking of capi.mojo:309-410's sweeps).
//
// A workgroup owns a tile of whole x-rows (NR rows of one z-range of one instance) and
// marches along z. Sweep j (colour c_j, alternating from c_1) at step P relaxes plane
// P - j + 1: it needs sweep j-1's values at planes P-j .. P-j+2, which sweep j-1 produced at
// steps P-2 .. P (sweeps run in order inside a step). Every value of a plane stays on chip
// from the step its input is read to the step its last sweep is done, in an LDS ring of
// W planes per colour (W = S + 1 for even S, S + 2 for odd S), updated in place:
//   * own-column operands (z-1, z, z+1) are the lane's own ring entries;
//   * the y+-1 rows and the x edge slot are other lanes' entries of the same plane, written
//     in an earlier step, so one workgroup barrier per step orders every exchange;
//   * the halo: sweep j is valid on the rows whose distance from the tile edge exceeds its
//     distance from the start, so a tile of NR rows yields NR - 2S output rows (the grid's
//     own edges excepted), and a z-range [z0, z1) starts S planes early and ends S late.
// The result is written out of place (x_in -> x_out: neighbouring tiles read x_in as their
// halo), once per node, or added into the FP64 solution for the mixed loop's last sweep.
//
// Missing neighbours read -0.0 (the pads of the ring and of the grid): x + (-0.0) == x for
// every x, so the unguarded sum of six terms is the guarded neighbour_sum's sequence of
// roundings. Per node the arithmetic is relax_node's, so results are bit-identical to S
// single sweeps. Precondition (checked by the launcher): point_weight >= 0 and non-negative
// screening weights, so every diagonal is >= 3 h^-2 > 0 and relax_node's diagonal guard
// never skips a node (the first sweep's colour is not read, only overwritten).
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "mpr_batch.h"
#include "stencil.h"

namespace mpr {

namespace {

using F4 = float __attribute__((ext_vector_type(4)));

// the live-bit word read for every chunk of a level without live bits
__device__ const uint32_t g_all_live = 0xffffffffu;
// lanes per workgroup: one workgroup per CU (its LDS ring), 4 waves per SIMD (the march's
// registers fit 128 VGPRs for S <= 4)
template <int S>
constexpr int march_threads() {
  return 1024;
}
constexpr size_t kMarchLds = 160 * 1024;
constexpr int kMarchMinN = 257;

__device__ __forceinline__ F4 ld4(gptr<const float> p) {
  return *(const __attribute__((address_space(1))) F4*)p;
}
__device__ __forceinline__ void st4(gptr<float> p, F4 v) {
  *(__attribute__((address_space(1))) F4*)p = v;
}

template <int S>
constexpr int ring_planes() {
  return S % 2 == 0 ? S + 1 : S + 2;
}

// positive residue of a compile-time offset
template <int W>
constexpr int wrap(int v) {
  return ((v % W) + W) % W;
}

struct MarchArgs {
  int32_t ny, ty, nr, zt, lr;  // tiles in y, output rows per tile, rows per tile, planes
                               // per z-range, lanes per row (H / 4)
  int32_t cA;                  // colour of sweep 1
  int32_t zero_start;          // ZeroStart: the input colour reads +0.0 (not loaded)
  int32_t correct, first;      // last phase of the mixed loop (x64 += e), first cycle
  float pw, h2;
  // the reciprocal of deg h^-2 for deg = 0..6 split into a float and the float nearest its
  // remainder (recip_split; deg = 3..6 are used)
  float zh[7], zl[7];
};

// Division by a diagonal known in advance, d = deg h^-2 (a node without screening weight),
// with one multiplication and one FMA: q = RN(a zh + RN(a zl)), where zh = RN(1 / d) and
// zl = RN(1 / d - zh) (Brisebarre, Muller and Raina, "Accelerating correctly rounded
// floating-point division when the divisor is known in advance", IEEE Trans. Computers
// 53(8), 2004). Whether it rounds correctly for every numerator depends on the divisor's
// significand only, here that of 3, 5 (4: zl = 0, an exact scaling) or 6: checked
// exhaustively on the device for every finite float numerator and the four diagonals of a
// level (mpr_fast_division_check). A zero quotient takes the sign of a (with zl < 0 the FMA
// would add a zero of the other sign). Once |a zl| is subnormal the product loses bits, so
// the sweep divides nonzero |a| < kTinyNumerator the IEEE way: fast_quotient.
constexpr float kTinyNumerator = 0x1p-60f;

__device__ __forceinline__ float div_known(float a, float zh, float zl) {
  const float q = __builtin_fmaf(a, zh, a * zl);
  return __builtin_bit_cast(float, (__builtin_bit_cast(uint32_t, q) & 0x7fffffffu) |
                                       (__builtin_bit_cast(uint32_t, a) & 0x80000000u));
}

// a nonzero numerator whose quotient may be subnormal
__device__ __forceinline__ bool tiny_numerator(float a) {
  return __builtin_fabsf(a) < kTinyNumerator && a != 0.0f;
}

// the sweep's division of a by an unscreened diagonal d = 1 / (zh + zl), as a function of one
// value
__device__ __forceinline__ float fast_quotient(float a, float d, float zh, float zl) {
  return tiny_numerator(a) ? a / d : div_known(a, zh, zl);
}

// zh = RN(1 / d) and zl = RN(1 / d - zh) for d > 0 (0 and 0 for d = 0). d zh is exact in
// double (24 x 24 bits), so 1 - d zh is the exact remainder and zl its quotient by d, within
// one rounding of the double quotient.
void recip_split(float d, float& zh, float& zl) {
  if (!(d > 0.0f)) {
    zh = zl = 0.0f;
    return;
  }
  zh = (float)(1.0 / (double)d);
  zl = (float)((1.0 - (double)d * (double)zh) / (double)d);
}

// Per-lane state of the march. Sweep j of step P reads the rhs of plane P - j + 1 in its
// colour; sweeps 1 and 2 are the first to read theirs (F, loaded a step ahead, Fn the next
// step's), sweep j > 2 the second (the same words sweep j - 2 read two steps earlier, reloaded
// at the start of the step from L2). lbs: the live-bit nibbles of planes P+2, P+1, ...
// (colour slot 0 = c_1 in the low nibble, 1 the other colour in the high), 8 bits per plane,
// newest lowest. Sc: the screening weights of the current step's sweeps (loaded a step
// ahead, only for chunks with a live slot: the others are +0.0).
template <int S, int W>
struct MarchRegs {
  F4 F[2], Fn[2];
  uint64_t lbs;
  F4 Sc[S];
};

template <int S, int LRT>
struct March {
  static constexpr int W = ring_planes<S>();
  // LDS, row-major: NR + 2 rows (pad rows 0 and NR + 1), each a segment of LR + 1 16-byte
  // chunks per colour slot and ring plane (slot-major), chunk 0 of every segment a pad; every
  // pad holds -0.0 for the whole march. With the row length a compile-time constant every
  // ring access is the lane's one base address plus an immediate offset.
  static constexpr int RSEG = LRT + 1;
  static constexpr int ROW = 2 * W * RSEG;
  F4* ring;
  F4* base;  // the lane's chunk in colour slot 0, ring plane 0 of its row
  const StepPtrs<float>* pp;
  const float *X, *RH, *SC;
  float* XO;
  const uint32_t* live;
  Geom g;
  MarchArgs a;
  int n, LR, r, k, y, z0, z1, yo0, yo1, Pb, Pe, cB, by;
  uint32_t lane_off;  // (y H + 4 k): the lane's chunk inside a plane of one colour
  bool act, zin, corr, init64, live_rhs, xedge, klast;
  float zhb, zlb, zhm, zlm;  // the diagonal's reciprocal pair without / with an x boundary

  // real slots of the lane's chunk in colour c's row of plane q
  __device__ uint32_t vmask(int c, int q) const {
    const int par = (c ^ y ^ q) & 1;
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (8 * k + 2 * i + par < n) m |= 1u << i;
    return m;
  }
  // the lane's chunk of colour slot cs, ring plane sl, dr rows away
  __device__ F4* L(int cs, int sl, int dr = 0) const {
    return base + ((cs * W + sl) * RSEG + dr * ROW);
  }
  // offset of colour c's plane q (uniform over the workgroup)
  __device__ uint32_t plane(int c, int q) const {
    return (uint32_t)c * g.C + (uint32_t)q * g.nH;
  }
  __device__ int colour(int cs) const { return cs ? cB : a.cA; }
  // input colour (slot 1) of plane q: loaded, +0.0 (zero start) or the -0.0 pad. Pad slots
  // past a row's last node keep whatever they hold: the one real read of such a slot (the
  // x+1 neighbour of x = n-1) is replaced by -0.0 in the sweep.
  // Loads at the start of a step are issued on every lane and for every plane, from an
  // in-grid address (clamped plane; inactive lanes read their workgroup's first row), and
  // their values replaced afterwards where needed: with no branch around them the compiler
  // counts the loads in flight exactly and waits only for the one a value comes from, never
  // for the prefetches behind it.
  __device__ uint32_t clamp_plane(int q) const { return (uint32_t)min(max(q, 0), n - 1); }
  // input colour (slot 1) of plane q: as loaded, +0.0 (zero start) or the -0.0 pad. Pad slots
  // past a row's last node keep whatever they hold: the one real read of such a slot (the
  // x+1 neighbour of x = n-1) is replaced by -0.0 in the sweep.
  __device__ F4 load_input(int q) const {
    const F4 v = ld4(as_global(X + plane(cB, (int)clamp_plane(q))) + lane_off);
    if (q < 0 || q >= n) return F4(-0.0f);
    return zin ? F4(0.0f) : v;
  }
  // the live-bit words of both colours' chunks of plane q (all ones without a live mask)
  __device__ void live_words(int q, uint32_t& wa, uint32_t& wb) const {
    const gptr<const uint32_t> lp = as_global(live ? live : &g_all_live);
    const uint32_t ia = plane(a.cA, (int)clamp_plane(q)) + lane_off;
    const uint32_t ib = plane(cB, (int)clamp_plane(q)) + lane_off;
    wa = lp[live ? ia >> 5 : 0u];
    wb = lp[live ? ib >> 5 : 0u];
  }
  // the nibbles of the words live_words(q) read (colour slot 0 low, slot 1 high)
  __device__ uint32_t live_nibbles(int q, uint32_t wa, uint32_t wb) const {
    if (!live || q < 0 || q >= n) return 0xffu;
    const uint32_t ia = plane(a.cA, q) + lane_off, ib = plane(cB, q) + lane_off;
    return ((wa >> (ia & 31u)) & 15u) | (((wb >> (ib & 31u)) & 15u) << 4);
  }
  // live nibble of colour slot cs of the plane d planes below the newest in lbs
  template <int D, int CS>
  __device__ static bool dead_at(uint64_t lbs) {
    return ((lbs >> (8 * D + 4 * CS)) & 15u) == 0u;
  }
  // rhs of colour slot cs, plane q (+0.0 outside the grid). Rhs words the live bits show to
  // be +0.0 are loaded all the same (they are +0.0).
  __device__ F4 rhs(int cs, int q) const {
    const F4 v = ld4(as_global(RH + plane(colour(cs), (int)clamp_plane(q))) + lane_off);
    return q >= 0 && q < n ? v : F4(0.0f);
  }
  // screening weights of sweep J at step P (plane P - J + 1, D planes below the newest in
  // lbs), +0.0 without a live slot. These loads are predicated and issued last in a step.
  template <int J, int D>
  __device__ void load_screen(MarchRegs<S, W>& s, int P) const {
    constexpr int cs = (J & 1) ? 0 : 1;
    const int q = P - J + 1;
    const bool in = q >= 0 && q < n;
    s.Sc[J - 1] = !in || dead_at<D, cs>(s.lbs)
                      ? F4(0.0f)
                      : ld4(as_global(SC + plane(colour(cs), in ? q : 0)) + lane_off);
  }
  // the screening weights of step P's sweeps, when lbs's newest plane is P + 1
  template <int... J>
  __device__ void load_screens(MarchRegs<S, W>& s, int P, std::integer_sequence<int, J...>) const {
    (load_screen<J + 1, J + 1>(s, P), ...);
  }

  // sweep j of step P (U: P's position in the unrolled period)
  template <int U, int J>
  __device__ void sweep(MarchRegs<S, W>& s, const F4 (&T)[S > 2 ? S - 2 : 1], int P) const {
    constexpr int cs = (J & 1) ? 0 : 1;
    constexpr int os = 1 - cs;
    constexpr int sl = wrap<W>(U - J + 1), slm = wrap<W>(U - J), slp = wrap<W>(U - J + 2);
    const int q = P - J + 1;
    const int lo = max(0, z0 - (S - J)), hi = min(n, z1 + (S - J));
    if (q < lo || q >= hi || !act) return;
    const int c = colour(cs);
    const int pr = (c ^ y ^ q) & 1;
    // the other colour's chunk of this row (one aligned 16-byte read) and its edge slot: the
    // last slot of chunk k - 1 for parity 0, the first of chunk k + 1 for parity 1 (the
    // segment's pad chunk, or the next segment's, supplies the -0.0 beyond the row)
    const F4 row = *L(os, sl);
    const float* Of = reinterpret_cast<const float*>(L(os, sl));
    float edge = pr ? Of[4] : Of[-1];
    float w[5];  // node i's x - 1 and x + 1 neighbours are w[i] and w[i + 1]
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const float lo = i == 0 ? edge : row[i - 1];
      const float hi = i == 4 ? edge : row[i];
      w[i] = pr ? hi : lo;
    }
    if (klast && pr == 0) w[1] = -0.0f;  // x = n - 1 has no x + 1 neighbour
    const F4 ym = *L(os, sl, -1), yp = *L(os, sl, 1);
    const F4 zm = *L(os, slm), zp = *L(os, slp);
    const F4 rh = J <= 2 ? s.F[J - 1] : T[J > 2 ? J - 3 : 0];
    const float h2 = a.h2;
    float num[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float sum = 0.0f;
      sum += w[i];
      sum += w[i + 1];
      sum += ym[i];
      sum += yp[i];
      sum += zm[i];
      sum += zp[i];
      num[i] = rh[i] + sum * h2;
    }
    const bool xb = xedge && pr == 0;  // node 0 is x = 0 or x = n - 1
    const bool bq = q == 0 || q == n - 1;
    // lbs's newest plane is P + 1 during the sweeps: this sweep's plane is J below it
    const bool dead = dead_at<J, cs>(s.lbs);
    F4 out;
    bool tiny = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float zh = i == 0 && xb ? zhm : zhb;
      const float zl = i == 0 && xb ? zlm : zlb;
      out[i] = div_known(num[i], zh, zl);
      tiny = tiny || tiny_numerator(num[i]);
    }
    // screened nodes, z-boundary planes and tiny numerators (fast_quotient's IEEE branch)
    if (!dead || bq || tiny) {
      // relax_node's diagonal: deg h^-2 + (alpha screen) h^-2, and an IEEE division
      const F4 sc = s.Sc[J - 1];
      const int dyz = 4 - by - (int)bq;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int deg = dyz + 2 - (int)(i == 0 && xb);
        const float screening = a.pw * sc[i] * h2;
        const float diag = (float)deg * h2 + screening;
        out[i] = num[i] / diag;
      }
    }
    *L(cs, sl) = out;
  }

  // the sweeps of one step in order; the scheduling barriers keep each sweep's operands
  // from being hoisted above the previous sweep (which would hold S sweeps' operands live)
  template <int U, int... J>
  __device__ void sweeps(MarchRegs<S, W>& s, const F4 (&T)[S > 2 ? S - 2 : 1], int P,
                         std::integer_sequence<int, J...>) const {
    ((sweep<U, J + 1>(s, T, P), __builtin_amdgcn_sched_barrier(0)), ...);
  }
  // second reads of the rhs (sweeps 3 .. S of step P), from L2
  template <int... J>
  __device__ void reload_rhs(F4 (&T)[S > 2 ? S - 2 : 1], int P,
                             std::integer_sequence<int, J...>) const {
    ((T[J] = rhs(((J + 3) & 1) ? 0 : 1, P - (J + 3) + 1)), ...);
  }

  // both colours of the finished plane q (slot SL) into x_out, or added into x64
  template <int SL>
  __device__ void output(int q) const {
    if (!act || q < z0 || q >= z1 || y < yo0 || y >= yo1) return;
    const F4 va = *L(0, SL), vb = *L(1, SL);
    if (!corr) {
      st4(as_global(XO + plane(a.cA, q)) + lane_off, va);
      st4(as_global(XO + plane(cB, q)) + lane_off, vb);
      return;
    }
    using D2 = double __attribute__((ext_vector_type(2)));
    gptr<double> X64 = as_global(pp->x64);
#pragma unroll
    for (int cs = 0; cs < 2; ++cs) {
      const int c = colour(cs);
      const F4 e = cs ? vb : va;
      const uint32_t o = plane(c, q) + lane_off;
      const uint32_t m = vmask(c, q);
      if (m == 15u) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          D2 v = init64 ? D2(kInitialX)
                        : *(const __attribute__((address_space(1))) D2*)(X64 + o + 2 * h);
          v[0] += (double)e[2 * h];
          v[1] += (double)e[2 * h + 1];
          *(__attribute__((address_space(1))) D2*)(X64 + o + 2 * h) = v;
        }
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if ((m >> i) & 1u) X64[o + i] = (init64 ? kInitialX : X64[o + i]) + (double)e[i];
      }
    }
  }

  template <int U>
  __device__ void step(MarchRegs<S, W>& s, int P) const {
    if (P >= Pe) return;  // uniform over the workgroup
    constexpr int s1 = wrap<W>(U + 1), s2 = wrap<W>(U + 2), so = wrap<W>(U - S + 1);
    F4 T[S > 2 ? S - 2 : 1];
    // loads first, on every lane: the input colour of plane P+2 (stored at the end of this
    // step), the live bits of plane P+2, the rhs of this step's later sweeps and of the next
    // step's first two
    const F4 xin = load_input(P + 2);
    uint32_t wa, wb;
    live_words(P + 2, wa, wb);
    if constexpr (S > 2) reload_rhs(T, P, std::make_integer_sequence<int, S - 2>{});
    s.Fn[0] = rhs(0, P + 1);
    s.Fn[1] = rhs(1, P);
    sweeps<U>(s, T, P, std::make_integer_sequence<int, S>{});
    s.F[0] = s.Fn[0];
    s.F[1] = s.Fn[1];
    output<so>(P - S + 1);
    if (act) {
      if (P + 1 < 0 || P + 1 >= n) *L(0, s1) = F4(-0.0f);
      *L(1, s2) = xin;
    }
    // then the next step's screening weights, predicated on the live bits (lbs's newest
    // plane is now P + 2), as the last loads of the step
    s.lbs = (s.lbs << 8) | live_nibbles(P + 2, wa, wb);
    if (act) load_screens(s, P + 1, std::make_integer_sequence<int, S>{});
    __syncthreads();
  }

  template <int... U>
  __device__ void period(MarchRegs<S, W>& s, int P0, std::integer_sequence<int, U...>) const {
    (step<U>(s, P0 + U), ...);
  }
};

template <int S, int LRT>
__global__ __launch_bounds__(march_threads<S>()) void kb_march(const StepPtrs<float>* tab,
                                                          uint32_t count, uint32_t per_inst,
                                                          Geom g, MarchArgs args) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using M = March<S, LRT>;
  constexpr int W = M::W;
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t inst = lb / per_inst;
  if (inst >= count) return;  // the whole workgroup
  const StepPtrs<float>& p = tab[inst];
  M m;
  m.ring = reinterpret_cast<F4*>(smem_raw);
  m.pp = &p;
  m.g = g;
  m.a = args;
  m.n = g.n;
  m.LR = LRT;
  const uint32_t w = lb - inst * per_inst;
  const int ty = (int)(w % (uint32_t)args.ny), tz = (int)(w / (uint32_t)args.ny);
  const int n = g.n, NR = args.nr, LR = LRT;
  m.yo0 = ty * args.ty;
  m.yo1 = min(n, m.yo0 + args.ty);
  const int ybase = max(0, min(m.yo0 - S, n - NR));
  m.z0 = tz * args.zt;
  m.z1 = min(n, m.z0 + args.zt);
  m.Pb = max(0, m.z0 - S + 1);
  m.Pe = m.z1 + S - 1;
  const int t = (int)threadIdx.x;
  m.r = t / LR;
  m.k = t - m.r * LR;
  m.act = m.r < NR;
  m.y = ybase + (m.act ? m.r : 0);
  m.cB = 1 - args.cA;
  m.base = m.ring + ((m.act ? m.r + 1 : 0) * M::ROW + 1 + m.k);
  m.lane_off = (uint32_t)m.y * (uint32_t)g.H + 4u * (uint32_t)m.k;
  m.X = p.x;
  m.XO = p.xalt;
  m.RH = p.rhs;
  m.SC = p.screen;
  m.live = p.live;
  m.live_rhs = p.live_rhs != 0;
  m.zin = args.zero_start == kZeroAll || (args.zero_start == kZeroBelowTop && !p.top);
  m.corr = args.correct && p.x64 != nullptr;
  m.init64 = m.corr && args.first && p.x_initial && *p.x_initial != 0;
  // n = 2^s + 1: x = n - 1 = 8 (LR - 1) is node 0 of the last lane (row parity 0), x = 0 node
  // 0 of the first
  m.klast = m.k == LR - 1;
  m.xedge = m.k == 0 || m.klast;
  m.by = (int)(m.y == 0) + (int)(m.y == n - 1);
  m.zhb = args.zh[6 - m.by];
  m.zlb = args.zl[6 - m.by];
  m.zhm = args.zh[5 - m.by];
  m.zlm = args.zl[5 - m.by];
  // pads: chunk 0 of every segment and the whole of rows 0 and NR + 1
  F4* ring = m.ring;
  const uint32_t segs = (uint32_t)(NR + 2) * 2u * (uint32_t)W;
  const uint32_t pads = segs + 2u * (uint32_t)M::ROW;
  for (uint32_t i = (uint32_t)t; i < pads; i += (uint32_t)march_threads<S>()) {
    uint32_t at;
    if (i < segs)
      at = i * (uint32_t)M::RSEG;  // chunk 0 of segment i
    else if (i < segs + (uint32_t)M::ROW)
      at = i - segs;  // row 0
    else
      at = (uint32_t)(NR + 1) * (uint32_t)M::ROW + (i - segs - (uint32_t)M::ROW);  // row NR + 1
    ring[at] = F4(-0.0f);
  }
  MarchRegs<S, W> s;
  s.lbs = 0;
  // prologue: the input colour of planes Pb-1 .. Pb+1, the live bits of Pb-S .. Pb+1, the
  // first step's first-read rhs and screening weights; slot(q) = (q - Pb) mod W
  const F4 a0 = m.load_input(m.Pb - 1), a1 = m.load_input(m.Pb), a2 = m.load_input(m.Pb + 1);
  {
#pragma unroll
    for (int d = -S; d <= 1; ++d) {
      uint32_t wa, wb;
      m.live_words(m.Pb + d, wa, wb);
      s.lbs = (s.lbs << 8) | m.live_nibbles(m.Pb + d, wa, wb);
    }
    s.F[0] = m.rhs(0, m.Pb);
    s.F[1] = m.rhs(1, m.Pb - 1);
    m.load_screens(s, m.Pb, std::make_integer_sequence<int, S>{});
  }
  if (m.act) {
    *m.L(1, W - 1) = a0;
    *m.L(1, 0) = a1;
    *m.L(1, 1) = a2;
    if (m.Pb - 1 < 0) *m.L(0, W - 1) = F4(-0.0f);
  }
  __syncthreads();
  for (int P0 = m.Pb; P0 < m.Pe; P0 += W)
    m.period(s, P0, std::make_integer_sequence<int, W>{});
}

// rows per tile and tiles per level for S fused sweeps on a level of lr lanes per row
struct MarchPlan {
  bool ok = false;
  MarchArgs a{};
  uint32_t per_inst = 0;
  size_t lds = 0;
};

int device_cus() {
  static int cus = 0;
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cus = prop.multiProcessorCount;
    if (cus <= 0) cus = 256;
  });
  return cus;
}

template <int S>
MarchPlan march_plan(const Geom& g, uint32_t count) {
  MarchPlan pl;
  if (!group_layout(g) || count == 0) return pl;
  constexpr int W = ring_planes<S>();
  const int n = g.n, lr = g.H / 4;
  // the kernel's row widths, and levels of 257^3: on smaller ones single sweeps are
  // latency-bound and the march's pipeline fill and halos cost as much or more (measured,
  // 8 instances per launch: 129^3 98 us per fused pass against 4 x 24 us, 65^3 42.5 us
  // against 4 x 7 us, 33^3 29.5 us against 4 x 5 us)
  if (n < kMarchMinN || (lr != 5 && lr != 9 && lr != 17 && lr != 33)) return pl;
  const int rows_lds = (int)(kMarchLds / (2ull * W * (size_t)(lr + 1) * 16ull)) - 2;
  const int nr_max = std::min({march_threads<S>() / lr, rows_lds, n});
  if (nr_max < 2 * S + 1 && nr_max < n) return pl;
  const int cus = device_cus();
  double best = 0;
  for (int ny = 1; ny <= n; ++ny) {
    const int ty = (n + ny - 1) / ny;
    const int nr = std::min(n, ty + 2 * S);
    if (nr > nr_max) continue;
    if ((ny - 1) * ty >= n) continue;  // an empty tile
    for (int nz = 1; nz <= n; ++nz) {
      const int zt = (n + nz - 1) / nz;
      if ((nz - 1) * zt >= n) continue;
      const uint64_t wgs = (uint64_t)count * ny * nz;
      const double rounds = (double)((wgs + cus - 1) / cus);
      const double cost = rounds * (double)(zt + 2 * S - 2) * (double)nr;
      if (!pl.ok || cost < best - 1e-9) {
        pl.ok = true;
        best = cost;
        pl.a.ny = ny;
        pl.a.ty = ty;
        pl.a.nr = nr;
        pl.a.zt = zt;
        pl.per_inst = (uint32_t)(ny * nz);
      }
    }
  }
  pl.a.lr = lr;
  pl.lds = 2ull * W * (size_t)(pl.a.nr + 2) * (size_t)(lr + 1) * 16ull;
  return pl;
}

// march_plan's search over tilings (a few hundred thousand steps on 257^3 and larger levels)
// is done once per level shape, instance count and sweep count, then looked up: the
// V-cycle asks for a plan at every launch
template <int S>
MarchPlan cached_march_plan(const Geom& g, uint32_t count) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, uint32_t>, MarchPlan> plans;
  const auto key = std::make_tuple(g.n, g.H, count);
  {
    std::lock_guard<std::mutex> lock(mu);
    const auto it = plans.find(key);
    if (it != plans.end()) return it->second;
  }
  const MarchPlan pl = march_plan<S>(g, count);
  std::lock_guard<std::mutex> lock(mu);
  plans.emplace(key, pl);
  return pl;
}

template <int S, int LRT>
void set_march_lds() {
  static std::once_flag once;
  static bool ok = true;
  std::call_once(once, [] {
    ok = hipFuncSetAttribute((const void*)kb_march<S, LRT>,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kMarchLds) == hipSuccess;
  });
  if (!ok) throw HipError{"fused smoothing kernel initialisation failed"};
}

template <int S, int LRT>
void launch_march_kernel(const StepPtrs<float>* tab, uint32_t count, const MarchPlan& pl,
                         const Geom& g, const MarchArgs& a, hipStream_t s) {
  set_march_lds<S, LRT>();
  const uint64_t blocks = (uint64_t)pl.per_inst * count;
  hipLaunchKernelGGL((kb_march<S, LRT>), dim3((unsigned)((blocks + 7u) & ~7ull)),
                     dim3(march_threads<S>()), pl.lds, s, tab, count, pl.per_inst, g, a);
  MPR_LAUNCH_CHECK();
}

template <int S>
bool launch_march_s(const StepPtrs<float>* tab, uint32_t count, const Geom& g, double pw,
                    int first_colour, int zero_start, bool correct, bool first, hipStream_t s) {
  const MarchPlan pl = cached_march_plan<S>(g, count);
  if (!pl.ok) return false;
  MarchArgs a = pl.a;
  a.cA = first_colour;
  a.zero_start = zero_start;
  a.correct = correct ? 1 : 0;
  a.first = first ? 1 : 0;
  a.pw = (float)pw;
  a.h2 = (float)((double)(g.n - 1) * (double)(g.n - 1));
  for (int d = 0; d <= 6; ++d) recip_split((float)d * a.h2, a.zh[d], a.zl[d]);
  switch (pl.a.lr) {
    case 5: launch_march_kernel<S, 5>(tab, count, pl, g, a, s); break;
    case 9: launch_march_kernel<S, 9>(tab, count, pl, g, a, s); break;
    case 17: launch_march_kernel<S, 17>(tab, count, pl, g, a, s); break;
    default: launch_march_kernel<S, 33>(tab, count, pl, g, a, s); break;
  }
  return true;
}

// counts the finite floats a (bit patterns start .. start + count - 1) whose fast_quotient
// differs from the IEEE quotient a / d in any bit
__global__ __launch_bounds__(256) void k_fast_div_check(float d, float zh, float zl,
                                                         uint64_t start, uint64_t count,
                                                         unsigned long long* bad) {
  unsigned long long local = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint32_t bits = (uint32_t)(start + i);
    if ((bits & 0x7f800000u) == 0x7f800000u) continue;  // inf and NaN
    const float a = __builtin_bit_cast(float, bits);
    const float want = a / d;
    if (__builtin_bit_cast(uint32_t, fast_quotient(a, d, zh, zl)) !=
        __builtin_bit_cast(uint32_t, want))
      ++local;
  }
  if (local) atomicAdd(bad, local);
}

}  // namespace

int64_t fast_division_mismatches(int n, hipStream_t s) {
  unsigned long long* bad = nullptr;
  MPR_HIP(hipMalloc((void**)&bad, sizeof(unsigned long long)));
  unsigned long long total = 0;
  try {
    MPR_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned long long), s));
    const float h2 = (float)((double)(n - 1) * (double)(n - 1));
    for (int deg = 3; deg <= 6; ++deg) {
      const float d = (float)deg * h2;
      float zh, zl;
      recip_split(d, zh, zl);
      hipLaunchKernelGGL(k_fast_div_check, dim3(8192), dim3(256), 0, s, d, zh, zl, (uint64_t)0,
                         (uint64_t)1 << 32, bad);
      MPR_LAUNCH_CHECK();
    }
    MPR_HIP(hipMemcpyAsync(&total, bad, sizeof(total), hipMemcpyDeviceToHost, s));
    MPR_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(bad);
    throw;
  }
  MPR_HIP(hipFree(bad));
  return (int64_t)total;
}

bool march_supported(const Geom& g, uint32_t count, int sweeps) {
  switch (sweeps) {
    case 2: return cached_march_plan<2>(g, count).ok;
    case 4: return cached_march_plan<4>(g, count).ok;
    default: return false;
  }
}

bool launch_b_march(const StepPtrs<float>* tab, uint32_t count, const Geom& g, double pw,
                    int sweeps, int first_colour, int zero_start, bool correct, bool first,
                    hipStream_t s) {
  if (!count) return true;
  if (!(pw >= 0.0)) return false;
  switch (sweeps) {
    case 2: return launch_march_s<2>(tab, count, g, pw, first_colour, zero_start, correct, first, s);
    case 4: return launch_march_s<4>(tab, count, g, pw, first_colour, zero_start, correct, first, s);
    default: return false;
  }
}

}  // namespace mpr
