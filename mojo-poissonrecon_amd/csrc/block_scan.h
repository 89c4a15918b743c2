// Workgroup-level helpers for order-preserving compaction on wave64 hardware.
#pragma once

#include "mpr_common.h"

namespace mpr {

// Sets bit index & 31 of words[index >> 5] for the lanes with `set`, one atomicOr per
// distinct word of the wavefront (one atomic per lane would serialise the lanes that share
// a word in the L2). Every lane of the wavefront calls it.
__device__ __forceinline__ void wave_set_bits(uint32_t* words, bool set, uint32_t index) {
  const uint32_t word = index >> 5, bit = 1u << (index & 31u);
  uint64_t pending = __ballot(set);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t w = (uint32_t)__shfl((int)word, leader);
    const bool mine = set && word == w;
    uint32_t v = mine ? bit : 0u;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v |= (uint32_t)__shfl_xor((int)v, o);
    if ((int)(threadIdx.x & 63u) == leader) atomicOr(&words[w], v);
    pending &= ~__ballot(mine);
  }
}

// Exclusive prefix of `flag` over the 256 threads of a block in thread order, using one
// ballot per wavefront and the 4 wave totals in LDS. `total` receives the block's count.
// Every thread of the block must call it (it synchronises twice).
__device__ __forceinline__ uint32_t block_prefix256(bool flag, uint32_t* lds4, uint32_t& total) {
  const uint64_t bal = __ballot(flag);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t in_wave = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
  if (lane == 0) lds4[wave] = (uint32_t)__popcll(bal);
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (uint32_t w = 0; w < 4; ++w) {
    const uint32_t c = lds4[w];
    base += w < wave ? c : 0u;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return base + in_wave;
}

// Exclusive prefix sum of `v` over the 256 threads of a block in thread order (wave64
// shuffle scan, then the 4 wave totals in LDS); `total` receives the block's sum. Every
// thread of the block must call it.
__device__ __forceinline__ uint32_t block_exclusive256(uint32_t v, uint32_t* lds4,
                                                       uint32_t& total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, 64);
    if (lane >= (uint32_t)off) inc += t;
  }
  if (lane == 63) lds4[wave] = inc;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (uint32_t w = 0; w < 4; ++w) {
    const uint32_t c = lds4[w];
    base += w < wave ? c : 0u;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return base + inc - v;
}

// Sum of a per-thread count over a 256-thread block (result valid in every thread).
__device__ __forceinline__ uint32_t block_count256(uint32_t v, uint32_t* lds4) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) lds4[wave] = v;
  __syncthreads();
  const uint32_t tot = lds4[0] + lds4[1] + lds4[2] + lds4[3];
  __syncthreads();
  return tot;
}

}  // namespace mpr
