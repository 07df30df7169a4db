// Kernel launch with per-family timing, for the .hip translation units.
//
// Every launch adds its algorithmic bytes to its family's total.  One launch
// in prof_every (and the family's first) is timed: the dispatch packet itself
// carries the start/stop events (hipExtLaunchKernel), so the measured
// duration is the kernel's execution alone, without the queueing latency that
// a pair of hipEventRecord markers around the launch would add.
#pragma once

#include <hip/hip_ext.h>

#include "gm_internal.h"

// Workgroups of a 1D grid are dispatched round-robin to the 8 XCDs, each with
// its own L2.  Element kernels remap blockIdx.x so that every XCD works
// through one contiguous range of an element list: neighbouring elements,
// which share nodes, then share an L2 (measured: a few per cent on the
// element-block smoother; neutral to slower on the row kernels, which keep
// the plain order).  A bijection on [0, gridDim.x).
__device__ inline int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// sum_k a[k] x[col[k]] over k = k0 + lane, k0 + lane + G, ... < k1, with two
// independent accumulators (two loads in flight per lane).
template <typename AT, typename VT, typename XT, int G>
__device__ inline AT row_dot(const VT *__restrict__ a, const int *__restrict__ col,
                             const XT *__restrict__ x, int k0, int k1, int lane) {
  AT s0 = 0, s1 = 0;
  int k = k0 + lane;
  for (; k + G < k1; k += 2 * G) {
    s0 += (AT)a[k] * (AT)x[col[k]];
    s1 += (AT)a[k + G] * (AT)x[col[k + G]];
  }
  if (k < k1) s0 += (AT)a[k] * (AT)x[col[k]];
  return s0 + s1;
}

template <typename... P, typename... A>
static inline void gm_launch(gm_ctx *pc, int id, double bytes, void (*kern)(P...), dim3 grid,
                             dim3 block, size_t shm, hipStream_t s, A... args) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  hipEvent_t eb = nullptr, ee = nullptr;
  if (pc && gm_prof_take(pc, id, bytes, &eb, &ee))
    hipExtLaunchKernelGGL(kern, grid, block, (std::uint32_t)shm, s, eb, ee, 0u, ((P)args)...);
  else
    kern<<<grid, block, shm, s>>>(((P)args)...);
}
