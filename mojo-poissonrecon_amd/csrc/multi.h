// Several instances of one kernel in one launch.
//
// A job's setup and finish are chains of small kernels (a few hundred thousand points, a
// grid's worth of bits): launched job by job, each is too small to fill the GPU and the
// chains of a batch run one after another. Here one launch runs the same kernel for up to
// kMultiMax instances (jobs): instance j owns row blockIdx.y = j of the grid, its first
// blocks[j] workgroups do its work (the rest return at once), and its arguments args[j]
// travel by value in the kernel's argument segment. A kernel is a body type:
//
//   struct Body {
//     struct Args {...};                        // one instance's arguments
//     static constexpr int kThreads = 256;      // workgroup size
//     static __device__ void run(const Args&, Blk);
//   };
//
// Blk gives the instance's own block index and block count, so bodies that remap blocks
// for the XCDs use the instance's count; the launch pads the grid's x extent to a multiple
// of 8, so a block's XCD is its x index mod 8 as in a launch of its own.
#ifndef MPR_MULTI_H
#define MPR_MULTI_H

#include <algorithm>
#include <vector>

#include "mpr_common.h"

namespace mpr {

constexpr int kMultiMax = 8;

struct Blk {
  uint32_t id;     // this workgroup's block index within its instance
  uint32_t count;  // the instance's blocks
};

template <class A>
struct MultiArgs {
  A args[kMultiMax];
  uint32_t blocks[kMultiMax];
};

template <class Body>
__global__ __launch_bounds__(Body::kThreads) void k_multi(const MultiArgs<typename Body::Args> m) {
  const uint32_t j = blockIdx.y;
  const uint32_t nb = m.blocks[j];
  if (blockIdx.x >= nb) return;
  Body::run(m.args[j], Blk{blockIdx.x, nb});
}

// XCD-aware block order inside an instance (count a multiple of 8): the 8 XCDs take
// workgroups round-robin, so block b is remapped to give each XCD one contiguous range.
__device__ __forceinline__ uint32_t xcd_block(Blk b) {
  return (b.id & 7u) * (b.count >> 3) + (b.id >> 3);
}

inline uint32_t pad8(uint32_t nb) { return (nb + 7u) & ~7u; }

// One launch per kMultiMax instances. blocks[i] = 0 skips instance i.
template <class Body>
void launch_multi(const typename Body::Args* args, const uint32_t* blocks, size_t count,
                  hipStream_t s) {
  for (size_t first = 0; first < count; first += kMultiMax) {
    const size_t k = std::min<size_t>(kMultiMax, count - first);
    MultiArgs<typename Body::Args> m{};
    uint32_t gx = 0;
    for (size_t i = 0; i < k; ++i) {
      m.args[i] = args[first + i];
      m.blocks[i] = blocks[first + i];
      gx = std::max(gx, blocks[first + i]);
    }
    if (!gx) continue;
    hipLaunchKernelGGL(k_multi<Body>, dim3(pad8(gx), (unsigned)k), dim3(Body::kThreads), 0, s,
                       m);
    MPR_LAUNCH_CHECK();
  }
}

template <class Body>
void launch_multi(const std::vector<typename Body::Args>& args,
                  const std::vector<uint32_t>& blocks, hipStream_t s) {
  launch_multi<Body>(args.data(), blocks.data(), args.size(), s);
}

template <class Body>
void launch_one(const typename Body::Args& a, uint32_t blocks, hipStream_t s) {
  launch_multi<Body>(&a, &blocks, 1, s);
}

}  // namespace mpr

#endif
