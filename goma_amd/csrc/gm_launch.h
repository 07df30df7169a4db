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
// element-block smoother, about 1 % on the HEX27 node-row kernel; neutral to
// slower on the per-row kernels, which keep the plain order).  A bijection on
// [0, gridDim.x).
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

// v from the lane R places along in the same row of 16 lanes (DPP row_ror)
template <int R>
__device__ inline int row_ror_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x120 + R, 0xf, 0xf, false);
}
template <int R>
__device__ inline float row_ror_f(float v) {
  return __builtin_bit_cast(float, row_ror_i<R>(__builtin_bit_cast(int, v)));
}
template <int R>
__device__ inline double row_ror_f(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)row_ror_i<R>((int)(unsigned)b);
  const unsigned hi = (unsigned)row_ror_i<R>((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// The rows of the MSR matrix one node at a time, without reading a column
// index per nonzero.  Row u of node n holds, for every neighbour node m of n
// in ascending order (n itself included), the unknowns of m in ascending
// order, minus the P columns in an energy row (the interaction mask) and
// minus u itself (the diagonal, stored apart): the order
// find_MSR_problem_graph builds (src/mm_fill_util.c:3229-3445).  One wave per
// node writes the node's column pattern to LDS from its neighbour list (one
// lane per neighbour, offsets by a wave prefix sum of the neighbours'
// unknown counts): PA for rows that take every column, PT for the energy
// row.  It then takes the node's rows four at a time, 16 lanes (one DPP row)
// per row, with coalesced value loads (two entries per lane and step) and a
// reduction by DPP row rotations; entry e of row r is pattern position e, or
// e + 1 from the diagonal's position on.  The node graph costs 4 bytes per
// neighbour for all the node's rows instead of 4 bytes per nonzero.
// Measured at 48^3 HEX27: 2.05 ms per product (4.05 TB/s algorithmic), the
// multigrid residual 0.596 ms averaged over the levels (4.0 TB/s).  Slower
// variants: the node's rows in turn with 64 lanes each (2.49 ms: one row's
// loads in flight at a time), 64 lanes per row with the column list (2.85),
// four accumulators per lane (2.10), a workgroup per node with wave 0
// building the pattern and a row per wave (3.39), the node's rows walked as
// one block with per-entry row selection (3.55).  `cap` is the longest
// pattern of the mesh (LDS: 2 cap ints per wave).  RESID: y = b - A x;
// otherwise y = A x.
template <typename AT, typename VT, typename XT, typename YT, bool RESID>
__global__ __launch_bounds__(256) void k_msr_nodes(int num_nodes, int dim, int nvn, int cap,
                                                   const int *__restrict__ fu,
                                                   const int *__restrict__ nn_ptr,
                                                   const int *__restrict__ nn_list,
                                                   const int *__restrict__ bindx,
                                                   const VT *__restrict__ a,
                                                   const XT *__restrict__ x,
                                                   const XT *__restrict__ b, YT *__restrict__ y) {
  extern __shared__ int s_pat[];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = xcd_block() * (blockDim.x >> 6) + w;
  const bool valid = n < num_nodes;
  int *PA = s_pat + (size_t)w * 2 * cap, *PT = PA + cap;
  int f0 = 0, nu = 0, selfA = 0, selfT = 0;
  if (valid) {
    f0 = fu[n];
    nu = fu[n + 1] - f0;
    const int j0 = nn_ptr[n], nb = nn_ptr[n + 1] - j0;
    int carryA = 0, carryT = 0;
    for (int c0 = 0; c0 < nb; c0 += 64) {
      const int jj = c0 + lane;
      int m = -1, fm = 0, lm = 0;
      if (jj < nb) {
        m = nn_list[j0 + jj];
        fm = fu[m];
        lm = fu[m + 1] - fm;
      }
      const int lenT = lm < nvn ? lm : nvn;
      int incA = lm, incT = lenT;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int tA = __shfl_up(incA, o), tT = __shfl_up(incT, o);
        if (lane >= o) {
          incA += tA;
          incT += tT;
        }
      }
      const int offA = carryA + incA - lm, offT = carryT + incT - lenT;
      for (int q = 0; q < lm; q++) PA[offA + q] = fm + q;
      for (int q = 0; q < lenT; q++) PT[offT + q] = fm + q;
      const unsigned long long sb = __ballot(m == n);
      if (sb) {
        const int sl = __builtin_ctzll(sb);
        selfA = __shfl(offA, sl);
        selfT = __shfl(offT, sl);
      }
      carryA += __shfl(incA, 63);
      carryT += __shfl(incT, 63);
    }
  }
  __syncthreads();
  if (!valid) return;
  const bool energy = nvn > dim;
  // four rows at a time, 16 lanes (one DPP row) per row
  const int g = lane >> 4, l = lane & 15;
  for (int r0 = 0; r0 < nu; r0 += 4) {
    const int r = r0 + g;
    AT s0 = 0, s1 = 0;
    int u = 0;
    if (r < nu) {
      u = f0 + r;
      const int k0 = bindx[u], len = bindx[u + 1] - k0;
      const bool tr = energy && r == dim;
      const int *P = tr ? PT : PA;
      const int hole = (tr ? selfT : selfA) + r;
      int e = l;
      for (; e + 16 < len; e += 32) {
        const int e1 = e + 16;
        s0 += (AT)a[k0 + e] * (AT)x[P[e + (e >= hole ? 1 : 0)]];
        s1 += (AT)a[k0 + e1] * (AT)x[P[e1 + (e1 >= hole ? 1 : 0)]];
      }
      if (e < len) s0 += (AT)a[k0 + e] * (AT)x[P[e + (e >= hole ? 1 : 0)]];
    }
    AT v = s0 + s1;
    v += row_ror_f<8>(v);
    v += row_ror_f<4>(v);
    v += row_ror_f<2>(v);
    v += row_ror_f<1>(v);
    if (l == 0 && r < nu) {
      const AT d = (AT)a[u] * (AT)x[u];
      y[u] = RESID ? (YT)((AT)b[u] - v - d) : (YT)(v + d);
    }
  }
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
