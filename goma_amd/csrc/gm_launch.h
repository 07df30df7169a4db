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
