// Geometric multigrid preconditioner with element-block (Vanka) smoothing.
//
// Levels: level 0 is the context's own mesh and (row-sum-scaled) Jacobian;
// coarser levels are contexts on coarsened meshes whose Jacobians are
// re-assembled by the same fill kernels at the injected state.  Transfers are
// CSR matrices P (coarse -> fine) and R = P^T supplied by the caller.
//
// Smoother: for each element colour, r = b - A x on every row, then every
// element of the colour solves its local system (all velocity, temperature
// and pressure unknowns of the element, entries taken from the assembled
// matrix through the element-to-nonzero map) and adds omega * A_e^-1 r_e to x.
// Elements of one colour share no unknowns, and the residual is formed before
// the colour's updates, so the sweep is deterministic (block Gauss-Seidel
// across colours, block Jacobi within a colour).  The local inverses are
// formed once per Newton step by Gauss-Jordan elimination with partial
// pivoting in LDS.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gm_launch.h"

template <int DIM, int EN>
struct Loc {
  static constexpr int NVE = DIM == 2 ? 9 : 27;
  static constexpr int NP = DIM + 1;
  static constexpr int NV = DIM + EN;
  static constexpr int N = NV * NVE + NP;
};

struct LevelDev {
  const int *conn, *fu, *centroid, *e2nz, *bindx;
  const signed char *dbc;
  long long E, off_mp, off_cv, off_cp;
};

template <int DIM, int EN>
__device__ inline int local_dof(const LevelDev &L, int e, int li) {
  using Lc = Loc<DIM, EN>;
  if (li < Lc::NV * Lc::NVE) {
    const int a = li / Lc::NVE, i = li - a * Lc::NVE;
    return L.fu[L.conn[(size_t)e * Lc::NVE + i]] + a;
  }
  return L.fu[L.centroid[e]] + DIM + EN + (li - Lc::NV * Lc::NVE);
}

// e2nz position of local entry (li, lj), -1 when the pair is not in the graph
template <int DIM, int EN>
__device__ inline long long local_pos(const LevelDev &L, int li, int lj) {
  using Lc = Loc<DIM, EN>;
  constexpr int NVE = Lc::NVE, NP = Lc::NP, NV = Lc::NV, NN = NV * NVE;
  if (li < NN) {
    const int a = li / NVE, i = li - a * NVE;
    if (lj < NN) {
      const int b = lj / NVE, j = lj - b * NVE;
      return (long long)(a * NV + b) * NVE * NVE + i * NVE + j;
    }
    if (a >= DIM) return -1;  // energy row x pressure column: not in the graph
    return L.off_mp + (long long)a * NVE * NP + i * NP + (lj - NN);
  }
  const int k = li - NN;
  if (lj < NN) {
    const int b = lj / NVE, j = lj - b * NVE;
    if (b >= DIM) return -1;  // continuity x temperature: identically zero
    return L.off_cv + (long long)b * NP * NVE + k * NVE + j;
  }
  return L.off_cp + (long long)k * NP + (lj - NN);
}

// local matrix -> inverse (stored transposed: ainv[e][j][i] = (A_e^-1)[i][j])
// Element-block inverse by Gauss-Jordan with partial pivoting in LDS, one
// workgroup per element, in compute type CT (FP64, or FP32 for the mixed
// precision).  Per pivot step: wave 0 finds the pivot while all threads copy
// column k; then the pivot row is copied and the rows are interchanged; then
// every entry is updated from the two copies.  The row interchanges are undone
// as column interchanges at the end, one row per thread.  The inverse is
// stored transposed (column-major) for the update kernel.
template <int DIM, int EN, typename VT, typename AT, typename CT, int NT>
__global__ __launch_bounds__(NT) void k_vanka_invert(LevelDev L, const VT *__restrict__ val,
                                                     AT *__restrict__ ainv) {
  using Lc = Loc<DIM, EN>;
  constexpr int n = Lc::N;
  static_assert(n <= 32 && NT % 32 == 0, "2D element blocks: at most 32 unknowns, rows on 32 lanes");
  __shared__ CT A[n][n + 1];
  __shared__ CT s_col[n], s_row[n];
  __shared__ int s_piv[n];
  __shared__ int s_p;
  const int e = blockIdx.x, tid = threadIdx.x;
  const int *map = L.e2nz + (size_t)e * L.E;
  for (int idx = tid; idx < n * n; idx += NT) {
    const int i = idx / n, j = idx - i * n;
    const long long p = local_pos<DIM, EN>(L, i, j);
    A[i][j] = p >= 0 ? (CT)val[map[p]] : (CT)0;
  }
  __syncthreads();
  for (int k = 0; k < n; k++) {
    for (int i = tid; i < n; i += NT) s_col[i] = A[i][k];
    if (tid < 64) {
      CT best = (CT)-1;
      int bi = n;
      for (int i = k + tid; i < n; i += 64) {
        const CT v = fabs(A[i][k]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const CT ov = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > best || (ov == best && oi < bi)) {
          best = ov;
          bi = oi;
        }
      }
      if (tid == 0) {
        if (bi >= n) bi = k;  // no comparable candidate (NaN column): keep the row
        s_p = bi;
        s_piv[k] = bi;
      }
    }
    __syncthreads();
    const int p = s_p;
    for (int j = tid; j < n; j += NT) {
      s_row[j] = A[p][j];
      if (p != k) A[p][j] = A[k][j];
    }
    __syncthreads();
    const CT piv = s_row[k];
    // a structurally empty column (a local pressure mode with no velocity
    // coupling inside the element) leaves that unknown untouched
    const CT d = fabs(piv) > (CT)1e-300 ? (CT)1 / piv : (CT)0;
    // thread (row ti, column group tg): no index division in the update
    const int ti = tid & 31, tg = tid >> 5;
    if (ti < n) {
      const int i = ti;
      if (i == k) {
        for (int j = tg; j < n; j += NT / 32) A[k][j] = j == k ? d : s_row[j] * d;
      } else {
        const CT f = i == p ? s_col[k] : s_col[i];
        for (int j = tg; j < n; j += NT / 32) A[i][j] = j == k ? -f * d : A[i][j] - f * (s_row[j] * d);
      }
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, last step first
  for (int i = tid; i < n; i += NT)
    for (int k = n - 1; k >= 0; k--) {
      const int p = s_piv[k];
      if (p != k) {
        const CT t = A[i][k];
        A[i][k] = A[i][p];
        A[i][p] = t;
      }
    }
  __syncthreads();
  AT *out = ainv + (size_t)e * n * n;
  for (int idx = tid; idx < n * n; idx += NT) {
    const int j = idx / n, i = idx - j * n;
    out[idx] = (AT)A[i][j];
  }
}

// the larger |entry| of (best, bi) and the lane R places along; lower row on ties
template <int R, typename CT>
__device__ inline void pivot_combine(CT &best, int &bi) {
  const CT ov = row_ror_f<R>(best);
  const int oi = row_ror_i<R>(bi);
  if (ov > best || (ov == best && oi < bi)) {
    best = ov;
    bi = oi;
  }
}

// Pivot of one column: the TR lanes of one DPP row hold its entries cv[a]
// (rows tr + TR a).  They stage the column in col, pick the row not yet used
// (bit a of `used`) with the largest |entry| (a NaN counts as zero, ties go to
// the lower row, so an unused row is always found), and the lane holding it
// posts the row to *pp, its reciprocal to *dv (0 for a structurally empty
// column: a local pressure mode with no velocity coupling inside the element,
// which leaves that unknown untouched) and clears the row's factor in col.
template <int n, int RA, int TR, typename CT>
__device__ inline void post_pivot(const CT (&cv)[RA], unsigned used, int tr, CT *col, int *pp,
                                  CT *dv) {
  CT best = (CT)-1;
  int bi = n;
#pragma unroll
  for (int a = 0; a < RA; a++) {
    const int i = tr + TR * a;
    const CT v = cv[a];
    col[i] = v;
    CT m = fabs(v);
    m = m == m ? m : (CT)0;
    if ((n % TR == 0 || i < n) && !((used >> a) & 1u) && m > best) {
      best = m;
      bi = i;
    }
  }
  pivot_combine<8>(best, bi);
  pivot_combine<4>(best, bi);
  pivot_combine<2>(best, bi);
  pivot_combine<1>(best, bi);
  if (tr == (bi & (TR - 1))) {
    const CT pv = col[bi];
    *dv = fabs(pv) > (CT)1e-300 ? (CT)1 / pv : (CT)0;
    col[bi] = (CT)0;
    *pp = bi;
  }
}

// The same inverse with the block held in registers: thread (tr, tc) of a
// TR x TC grid owns rows tr + TR a and columns tc + TC b.  Pivoting is
// implicit: step k takes its pivot row p_k among the rows not yet used and
// leaves every row where it is.  Scaling is lazy: the pivot row keeps its
// entries (column k set to 1) and is multiplied by its pivot's reciprocal d_k
// when the block is written out; the other rows subtract f (d_k row) as usual.
// The stored result M then holds the inverse with both permutations applied,
// inv[k][p_c] = d_k M[p_k][c], undone at the write.  So no register is
// written under a runtime row index: the loop over column blocks is unrolled
// (column k's registers are addressed at compile time) and the pivot row is
// read through a scalar switch on its block index, which is uniform over the
// workgroup.  One barrier a step: the lanes owning column k post its pivot,
// the pivot's reciprocal and the column in LDS (double-buffered by step
// parity); the pivot row's entries that a thread needs are held by lanes of
// its own wave (row rp, same tc) and pass through LDS within the wave.
template <int DIM, int EN, typename VT, typename AT, typename CT, int TR, int TC>
__global__ __launch_bounds__(TR *TC) void k_vanka_invert_reg(LevelDev L, const VT *__restrict__ val,
                                                             AT *__restrict__ ainv) {
  using Lc = Loc<DIM, EN>;
  constexpr int n = Lc::N;
  constexpr int RA = (n + TR - 1) / TR, CB = (n + TC - 1) / TC;
  static_assert(RA <= 8, "pivot-row switch covers eight row blocks");
  static_assert(TR == 16, "a column's TR lanes form one DPP row");
  // every row / column index of the thread grid is inside the block
  constexpr bool RFULL = n % TR == 0, CFULL = n % TC == 0;
  __shared__ CT s_col[2][RA * TR], s_row[CB * TC], s_dv[2], s_d[n];
  __shared__ int s_prow[n], s_kinv[n], s_pp[2];
  const int e = blockIdx.x, tid = threadIdx.x;
  const int tr = tid % TR, tc = tid / TR;
  const int *map = L.e2nz + (size_t)e * L.E;
  CT A[RA][CB];
#pragma unroll
  for (int a = 0; a < RA; a++)
#pragma unroll
    for (int b = 0; b < CB; b++) {
      const int i = tr + TR * a, j = tc + TC * b;
      const long long p = ((RFULL || i < n) && (CFULL || j < n)) ? local_pos<DIM, EN>(L, i, j) : -1;
      A[a][b] = p >= 0 ? (CT)val[map[p]] : (CT)0;
    }
  unsigned used = 0;  // bit a: row tr + TR a has been a pivot row
#pragma unroll
  for (int K = 0; K < CB; K++) {
    for (int r = 0; r < TC; r++) {
      const int k = TC * K + r;
      if (k >= n) break;
      const int q = k & 1;
      if (tc == r) {
        CT cv[RA];
#pragma unroll
        for (int a = 0; a < RA; a++) cv[a] = A[a][K];
        post_pivot<n, RA, TR>(cv, used, tr, s_col[q], &s_pp[q], &s_dv[q]);
      }
      __syncthreads();
      const int p = __builtin_amdgcn_readfirstlane(s_pp[q]);
      const CT d = s_dv[q];
      if (tid == 0) {
        s_prow[k] = p;
        s_kinv[p] = k;
        s_d[k] = d;
      }
      const int ap = p / TR, rp = p - ap * TR;
      // This wave's columns of the pivot row pass through s_row: the lanes of
      // row p write them and the same wave reads them back (LDS keeps one
      // wave's operations in order; the fences keep the compiler from moving
      // the reads above the writes), so no workgroup barrier is needed.
      if (tr == rp) {
        switch (ap) {
#define GM_PIVOT_ROW(c)                                            \
  case c:                                                          \
    if constexpr (c < RA) {                                        \
      _Pragma("unroll") for (int b = 0; b < CB; b++) {             \
        const int j = tc + TC * b;                                 \
        if (CFULL || j < n) s_row[j] = A[c][b];                    \
      }                                                            \
    }                                                              \
    break;
          GM_PIVOT_ROW(0) GM_PIVOT_ROW(1) GM_PIVOT_ROW(2) GM_PIVOT_ROW(3)
          GM_PIVOT_ROW(4) GM_PIVOT_ROW(5) GM_PIVOT_ROW(6) GM_PIVOT_ROW(7)
#undef GM_PIVOT_ROW
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // Step k of Gauss-Jordan on pivot row p: every other row i subtracts f_i
      // times d row p (f_i its column-k entry) and takes -f_i d in column k;
      // row p keeps its entries, with 1 in column k.  One FMA per entry after
      // column k is set to 1 (row p) and 0 (the others); f_p is 0.
      CT rb[CB], fa[RA];
#pragma unroll
      for (int b = 0; b < CB; b++) {
        const int j = tc + TC * b;
        rb[b] = (CFULL || j < n) ? s_row[j] * d : (CT)0;
      }
      rb[K] = tc == r ? d : rb[K];
#pragma unroll
      for (int a = 0; a < RA; a++) fa[a] = s_col[q][tr + TR * a];
      if (tc == r) {
#pragma unroll
        for (int a = 0; a < RA; a++) A[a][K] = tr + TR * a == p ? (CT)1 : (CT)0;
      }
#pragma unroll
      for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < CB; b++) A[a][b] -= fa[a] * rb[b];
      if (tr == rp) used |= 1u << ap;
    }
  }
  __syncthreads();
  AT *out = ainv + (size_t)e * n * n;
#pragma unroll
  for (int a = 0; a < RA; a++) {
    const int i = tr + TR * a;
    if (!RFULL && i >= n) continue;
    const int ki = s_kinv[i];
    const CT di = s_d[ki];
#pragma unroll
    for (int b = 0; b < CB; b++) {
      const int c = tc + TC * b;
      if (CFULL || c < n) out[(size_t)s_prow[c] * n + ki] = (AT)(A[a][b] * di);
    }
  }
}

// r = b - A x over all rows (MSR), G lanes per row
template <typename VT, typename BT, int G>
__global__ __launch_bounds__(256) void k_mg_resid(int N, const VT *__restrict__ a,
                                                  const int *__restrict__ bindx,
                                                  const BT *__restrict__ x,
                                                  const BT *__restrict__ b, BT *__restrict__ r) {
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  if (row >= N) return;
  using AT = typename std::conditional<sizeof(BT) == 8, double, float>::type;
  const int k0 = bindx[row], k1 = bindx[row + 1];
  AT s = row_dot<AT, VT, BT, G>(a, bindx, x, k0, k1, lane);
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (lane == 0) r[row] = (BT)((AT)b[row] - s - (AT)a[row] * (AT)x[row]);
}

// r = b - A x on a list of rows (the unknowns of one element colour)
template <typename VT, typename BT, int G>
__global__ __launch_bounds__(256) void k_mg_resid_rows(int nrows, const int *__restrict__ rows,
                                                       const VT *__restrict__ a,
                                                       const int *__restrict__ bindx,
                                                       const BT *__restrict__ x,
                                                       const BT *__restrict__ b, BT *__restrict__ r) {
  const int k = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  if (k >= nrows) return;
  const int row = rows[k];
  using AT = typename std::conditional<sizeof(BT) == 8, double, float>::type;
  const int k0 = bindx[row], k1 = bindx[row + 1];
  AT s = row_dot<AT, VT, BT, G>(a, bindx, x, k0, k1, lane);
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (lane == 0) r[row] = (BT)((AT)b[row] - s - (AT)a[row] * (AT)x[row]);
}

// ---- exact coarse solve: dense inverse by in-place Gauss-Jordan with
// partial pivoting (two launches per pivot), applied as a dense GEMV.
template <typename VT>
__global__ void k_msr_to_dense(int n, const VT *__restrict__ a, const int *__restrict__ bindx,
                               double *__restrict__ D) {
  const int row = blockIdx.x;
  for (int k = bindx[row] + threadIdx.x; k < bindx[row + 1]; k += blockDim.x)
    D[(size_t)row * n + bindx[k]] = (double)a[k];
  if (threadIdx.x == 0) D[(size_t)row * n + row] = (double)a[row];
}

// Dense Gauss-Jordan inverse of the coarsest level with partial pivoting,
// blocked in panels of at most GJ_NB columns (same pivots and interchanges as
// the column-at-a-time algorithm).  For panel K:
//   k_gjb_panel    one workgroup runs the panel's column steps on the panel
//                  columns only, the panel in registers (pivot: largest
//                  |entry| at or below the diagonal, first on ties; whole
//                  panel rows interchanged);
//   k_gjb_swap     the panel's interchanges on every other column, in order,
//                  and a copy of the pivot rows B_K of those columns;
//   k_gjb_update   every other column: B <- B' + Pan B_K, B' = B with the
//                  pivot rows zeroed (the panel now holds the K columns of
//                  the accumulated elimination, so pivot rows become
//                  A_KK^-1 B_K and the others B_R - A_RK A_KK^-1 B_K).
// D is row-major n x n.  perm[k] = the row interchanged with k at step k.
constexpr int GJ_NB = 32;

// The panel lives in registers for its column steps: thread t of 512 owns
// rows t + 512 r (r < RPT) of the NB panel columns (RPT NB <= 64 doubles),
// so D is read and written once per panel.  Per step: the pivot candidate of
// each wave (DPP row rotations, then the wave's four rows by readlane) goes to
// LDS; after one barrier every thread merges the eight candidates itself; the
// pivot row and row k pass through LDS; a second barrier.  The LDS buffers
// alternate by step parity, so no third barrier guards their reuse.
template <int NB, int RPT>
__global__ __launch_bounds__(512) void k_gjb_panel(int n, int kb0, int nb, double *__restrict__ D,
                                                    int *__restrict__ perm, int *__restrict__ flag) {
  __shared__ double s_rk[2][NB], s_rp[2][NB];
  __shared__ double sv[2][8];
  __shared__ int si[2][8];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double P[RPT][NB];
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const int i = tid + 512 * r;
#pragma unroll
    for (int j = 0; j < NB; j++) P[r][j] = (i < n && j < nb) ? D[(size_t)i * n + kb0 + j] : 0.0;
  }
#pragma unroll
  for (int kk = 0; kk < NB; kk++) {
    if (kk < nb) {
      const int k = kb0 + kk, q = kk & 1;
      double best = -1.0;
      int bi = n;
#pragma unroll
      for (int r = 0; r < RPT; r++) {
        const int i = tid + 512 * r;
        const double v = fabs(P[r][kk]);
        if (i >= k && i < n && v > best) {
          best = v;
          bi = i;
        }
      }
      pivot_combine<8>(best, bi);
      pivot_combine<4>(best, bi);
      pivot_combine<2>(best, bi);
      pivot_combine<1>(best, bi);
      // lanes 0, 16, 32, 48 hold their rows' candidates
#pragma unroll
      for (int g = 1; g < 4; g++) {
        const unsigned long long bb = __builtin_bit_cast(unsigned long long, best);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)bb, 16 * g);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(bb >> 32), 16 * g);
        const double ov = __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
        const int oi = __builtin_amdgcn_readlane(bi, 16 * g);
        if (lane == 0 && (ov > best || (ov == best && oi < bi))) {
          best = ov;
          bi = oi;
        }
      }
      if (lane == 0) {
        sv[q][w] = best;
        si[q][w] = bi;
      }
      __syncthreads();
      double bv = sv[q][0];
      int p = si[q][0];
#pragma unroll
      for (int g = 1; g < 8; g++) {
        const double ov = sv[q][g];
        const int oi = si[q][g];
        if (ov > bv || (ov == bv && oi < p)) {
          bv = ov;
          p = oi;
        }
      }
      if (p >= n) p = k;
      if (tid == 0) {
        if (!(bv > 0.0)) flag[3] = 1;  // singular coarse matrix
        perm[k] = p;
      }
      // rows k and p to LDS (row p is the new pivot row), then interchanged
#pragma unroll
      for (int r = 0; r < RPT; r++) {
        const int i = tid + 512 * r;
        if (i == p)
#pragma unroll
          for (int j = 0; j < NB; j++) s_rp[q][j] = P[r][j];
        if (i == k && p != k)
#pragma unroll
          for (int j = 0; j < NB; j++) s_rk[q][j] = P[r][j];
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RPT; r++) {
        const int i = tid + 512 * r;
        if (p != k && i == k)
#pragma unroll
          for (int j = 0; j < NB; j++) P[r][j] = s_rp[q][j];
        if (p != k && i == p)
#pragma unroll
          for (int j = 0; j < NB; j++) P[r][j] = s_rk[q][j];
      }
      const double piv = s_rp[q][kk];
      const double inv = piv != 0.0 ? 1.0 / piv : 0.0;
#pragma unroll
      for (int r = 0; r < RPT; r++) {
        const int i = tid + 512 * r;
        const double f = P[r][kk];
        if (i == k) {
#pragma unroll
          for (int j = 0; j < NB; j++) P[r][j] = j == kk ? inv : s_rp[q][j] * inv;
        } else {
          const double g = f * inv;
#pragma unroll
          for (int j = 0; j < NB; j++) P[r][j] = j == kk ? -g : P[r][j] - g * s_rp[q][j];
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const int i = tid + 512 * r;
    if (i < n)
#pragma unroll
      for (int j = 0; j < NB; j++)
        if (j < nb) D[(size_t)i * n + kb0 + j] = P[r][j];
  }
}

__global__ __launch_bounds__(256) void k_gjb_swap(int n, int kb0, int nb, double *__restrict__ D,
                                                  const int *__restrict__ perm,
                                                  double *__restrict__ BK) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || (j >= kb0 && j < kb0 + nb)) return;
  for (int kk = 0; kk < nb; kk++) {
    const int k = kb0 + kk, p = perm[k];
    if (p != k) {
      const double t = D[(size_t)k * n + j];
      D[(size_t)k * n + j] = D[(size_t)p * n + j];
      D[(size_t)p * n + j] = t;
    }
  }
  for (int kk = 0; kk < nb; kk++) BK[(size_t)kk * n + j] = D[(size_t)(kb0 + kk) * n + j];
}

// 64 x 64 output tile per workgroup, 4 x 4 per thread; columns of the panel
// are left alone
__global__ __launch_bounds__(256) void k_gjb_update(int n, int kb0, int nb, double *__restrict__ D,
                                                    const double *__restrict__ BK) {
  __shared__ double sP[64][GJ_NB + 1];
  __shared__ double sB[GJ_NB][64];
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64, tid = threadIdx.x;
  for (int idx = tid; idx < 64 * GJ_NB; idx += 256) {
    const int r = idx / GJ_NB, t = idx - r * GJ_NB, i = i0 + r;
    sP[r][t] = (i < n && t < nb) ? D[(size_t)i * n + kb0 + t] : 0.0;
  }
  for (int idx = tid; idx < GJ_NB * 64; idx += 256) {
    const int t = idx >> 6, c = idx & 63, j = j0 + c;
    sB[t][c] = (t < nb && j < n) ? BK[(size_t)t * n + j] : 0.0;
  }
  __syncthreads();
  const int tr = tid >> 4, tc = tid & 15;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[a][c] = 0.0;
  for (int t = 0; t < nb; t++) {
    double pa[4], pb[4];
#pragma unroll
    for (int a = 0; a < 4; a++) pa[a] = sP[tr + 16 * a][t];
#pragma unroll
    for (int c = 0; c < 4; c++) pb[c] = sB[t][tc + 16 * c];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[a][c] += pa[a] * pb[c];
  }
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int i = i0 + tr + 16 * a;
    if (i >= n) continue;
    const bool krow = i >= kb0 && i < kb0 + nb;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int j = j0 + tc + 16 * c;
      if (j >= n || (j >= kb0 && j < kb0 + nb)) continue;
      const size_t at = (size_t)i * n + j;
      D[at] = (krow ? 0.0 : D[at]) + acc[a][c];
    }
  }
}

__global__ void k_gj_unpermute(int n, const int *__restrict__ perm, double *__restrict__ D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double *r = D + (size_t)i * n;
  for (int k = n - 1; k >= 0; k--) {
    const int p = perm[k];
    if (p != k) {
      const double t = r[k];
      r[k] = r[p];
      r[p] = t;
    }
  }
}

// x = D b, one wave per row
template <typename BT>
__global__ __launch_bounds__(256) void k_dense_gemv(int n, const double *__restrict__ D,
                                                    const BT *__restrict__ b, BT *__restrict__ x) {
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (row >= n) return;
  const double *r = D + (size_t)row * n;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += r[j] * (double)b[j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) x[row] = (BT)s;
}

// x_e += omega * A_e^-1 r_e for the nelem elements of one colour
template <int DIM, int EN, typename AT, typename BT, int NT>
__global__ __launch_bounds__(NT) void k_vanka_update(LevelDev L, int nelem, const int *__restrict__ elems,
                                                     const AT *__restrict__ ainv,
                                                     const BT *__restrict__ r, BT *__restrict__ x,
                                                     double omega, const double *__restrict__ wmul) {
  using Lc = Loc<DIM, EN>;
  constexpr int n = Lc::N;
  // two elements per workgroup when a block fits in half of it (2D: 30 of 64 lanes)
  constexpr int EPB = (2 * n <= NT) ? 2 : 1;
  constexpr int HALF = NT / EPB;
  __shared__ double s_r2[EPB][n];
  __shared__ int s_g2[EPB][n];
  const int tid = threadIdx.x, sub = tid / HALF, t = tid - sub * HALF;
  const int eidx = xcd_block() * EPB + sub;
  const bool act = eidx < nelem;
  const int e = act ? elems[eidx] : 0;
  if (act)
    for (int i = t; i < n; i += HALF) {
      const int g = local_dof<DIM, EN>(L, e, i);
      s_g2[sub][i] = g;
      s_r2[sub][i] = (double)r[g];
    }
  __syncthreads();
  if (act) {
    const AT *M = ainv + (size_t)e * n * n;
    for (int i = t; i < n; i += HALF) {
      double s = 0.0;
      for (int j = 0; j < n; j++) s += (double)M[j * n + i] * s_r2[sub][j];
      const int g = s_g2[sub][i];
      x[g] += (BT)(omega * (wmul ? wmul[g] : 1.0) * s);
    }
  }
}

// Additive smoother, first half: every element's correction c_e = A_e^-1 r_e
// (all elements in one launch; nothing is written to x).
template <int DIM, int EN, typename AT, typename BT, int NT>
__global__ __launch_bounds__(NT) void k_vanka_corr(LevelDev L, int nelem, const AT *__restrict__ ainv,
                                                   const BT *__restrict__ r, double *__restrict__ corr) {
  using Lc = Loc<DIM, EN>;
  constexpr int n = Lc::N;
  constexpr int EPB = (2 * n <= NT) ? 2 : 1;
  constexpr int HALF = NT / EPB;
  __shared__ double s_r2[EPB][n];
  const int tid = threadIdx.x, sub = tid / HALF, t = tid - sub * HALF;
  const int e = xcd_block() * EPB + sub;
  const bool act = e < nelem;
  if (act)
    for (int i = t; i < n; i += HALF) s_r2[sub][i] = (double)r[local_dof<DIM, EN>(L, e, i)];
  __syncthreads();
  if (act) {
    const AT *M = ainv + (size_t)e * n * n;
    for (int i = t; i < n; i += HALF) {
      double s = 0.0;
      for (int j = 0; j < n; j++) s += (double)M[j * n + i] * s_r2[sub][j];
      corr[(size_t)e * n + i] = s;
    }
  }
}

// Second half: x_u += omega / (elements sharing u) * sum of u's corrections, in
// a fixed order (deterministic); assign: x_u = that (x is zero on entry).
template <typename BT>
__global__ __launch_bounds__(256) void k_vanka_gather(int N, const int *__restrict__ gptr,
                                                      const int *__restrict__ gent,
                                                      const double *__restrict__ corr,
                                                      const double *__restrict__ wmul, double omega,
                                                      BT *__restrict__ x, int assign) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= N) return;
  double s = 0.0;
  for (int k = gptr[u]; k < gptr[u + 1]; k++) s += corr[gent[k]];
  const BT d = (BT)(omega * wmul[u] * s);
  x[u] = assign ? d : x[u] + d;
}

// y = M v (CSR), optionally scaling v by sc and masking Dirichlet rows; G
// lanes per row (consecutive lanes read consecutive entries of the row)
template <typename BT, int G>
__global__ __launch_bounds__(256) void k_csr_apply(int nrows, const int *__restrict__ rp,
                                                   const int *__restrict__ ci,
                                                   const double *__restrict__ cv,
                                                   const BT *__restrict__ v,
                                                   const double *__restrict__ sc,
                                                   const signed char *__restrict__ mask,
                                                   BT *__restrict__ y, int accumulate) {
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  if (i >= nrows) return;  // whole groups of G lanes leave together
  double s = 0.0;
  for (int k = rp[i] + lane; k < rp[i + 1]; k += G) {
    const int j = ci[k];
    s += cv[k] * (double)v[j] * (sc ? sc[j] : 1.0);
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (lane) return;
  if (mask && mask[i]) s = 0.0;
  if (accumulate) y[i] += (BT)s;
  else y[i] = (BT)s;
}

// coarse state: nodal values injected from the finer level, pressure 0
__global__ void k_inject(int ncn, const int *__restrict__ inj, const int *__restrict__ fu_f,
                         const int *__restrict__ fu_c, int nv, const double *__restrict__ xf,
                         double *__restrict__ xc) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < ncn; n += gridDim.x * blockDim.x) {
    const int f = fu_f[inj[n]], cc = fu_c[n], cnt = fu_c[n + 1] - cc;
    for (int a = 0; a < cnt; a++) xc[cc + a] = a < nv ? xf[f + a] : 0.0;
  }
}

template <typename T>
__global__ void k_set(T *p, long long n, T v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------- host side

static LevelDev level_dev(const gm_ctx *c) {
  LevelDev L;
  L.conn = c->d_conn;
  L.fu = c->d_fu;
  L.centroid = c->d_centroid;
  L.e2nz = c->d_e2nz;
  L.bindx = c->d_bindx;
  L.dbc = c->d_dbc;
  L.E = c->lay.total;
  L.off_mp = c->lay.off_mp;
  L.off_cv = c->lay.off_cv;
  L.off_cp = c->lay.off_cp;
  return L;
}

static int nloc_of(const gm_ctx *c) { return c->nv * c->nve + c->np; }

template <typename T>
static int upv(T **d, const T *h, size_t n) {
  GM_HIP(hipMalloc(d, std::max<size_t>(n, 1) * sizeof(T)));
  if (n) GM_HIP(hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return GM_OK;
}


// Largest number of nonzeros of one child's nodal interpolation W that the
// Galerkin kernel keeps in LDS (a Q2 child of a Q2 parent: 125 in 3D, 25 in 2D)
template <int DIM>
constexpr int gal_ws_max() { return DIM == 2 ? 32 : 128; }

// Galerkin coarse operators.  When every finer element lies inside one
// coarser element (its parent) and P maps the parent's unknowns to the
// element's unknowns, R A P = sum over finer elements of P_e^T K_e P_e, with K_e
// the element matrix before Dirichlet rows and P_e the element-local block of
// P.  The coarse element matrix of element E is then the sum over its 2^dim
// children, and the coarse level assembles it like any element matrix
// (the same element-to-nonzero map, its own Dirichlet rows).  P_e is one
// nodal weight matrix W (the same for every nodal variable) and one P1 matrix Q.
static std::string build_galerkin_maps(const gm_ctx *f, const gm_ctx *c, const int *Prow,
                                       const int *Pcol, const double *Pval, MgLevel &L) {
  const int nve = f->nve, np = f->np, nv = f->nv, nch = 1 << f->dim;
  std::vector<int> node_of(c->N, -1);
  for (int n = 0; n < c->num_nodes; n++)
    for (int k = c->first_unknown[n]; k < c->first_unknown[n + 1]; k++) node_of[k] = n;
  std::vector<int> elem_of_centroid(c->num_nodes, -1);
  for (int e = 0; e < c->num_elems; e++) elem_of_centroid[c->elem_centroid[e]] = e;
  std::vector<int> child((size_t)c->num_elems * nch, -1), nchild(c->num_elems, 0);
  std::vector<double> W((size_t)f->num_elems * nve * nve, 0.0);
  std::vector<double> Q((size_t)f->num_elems * np * np, 0.0);
  for (int e = 0; e < f->num_elems; e++) {
    // parent: the coarse element whose centroid node enters the interpolation
    // of this element's centroid
    const int ur = f->first_unknown[f->elem_centroid[e]];
    int parent = -1;
    for (int k = Prow[ur]; k < Prow[ur + 1]; k++) {
      const int pe = elem_of_centroid[node_of[Pcol[k]]];
      if (pe < 0 || Pval[k] == 0.0) continue;
      if (parent >= 0 && parent != pe) return "an element's centroid interpolates from two coarse elements";
      parent = pe;
    }
    if (parent < 0) return "an element has no parent coarse element";
    if (nchild[parent] >= nch) return "a coarse element has more than 2^dim children";
    child[(size_t)parent * nch + nchild[parent]++] = e;
    const int *cn = &c->conn[(size_t)parent * nve];
    for (int i = 0; i < nve; i++) {
      const int u = f->first_unknown[f->conn[(size_t)e * nve + i]];  // first nodal variable
      for (int k = Prow[u]; k < Prow[u + 1]; k++) {
        const int n = node_of[Pcol[k]];
        if (Pcol[k] != c->first_unknown[n]) return "P couples different variables";
        int m = 0;
        while (m < nve && cn[m] != n) m++;
        if (m == nve) return "P couples an element to unknowns outside its parent";
        W[(size_t)e * nve * nve + i * nve + m] = Pval[k];
      }
    }
    const int pf = f->first_unknown[f->elem_centroid[e]] + nv;
    const int pc = c->first_unknown[c->elem_centroid[parent]] + nv;
    for (int kk = 0; kk < np; kk++)
      for (int k = Prow[pf + kk]; k < Prow[pf + kk + 1]; k++) {
        const int l = Pcol[k] - pc;
        if (l < 0 || l >= np) return "P couples pressure outside the parent element";
        Q[(size_t)e * np * np + kk * np + l] = Pval[k];
      }
  }
  for (int e = 0; e < c->num_elems; e++)
    if (nchild[e] != nch) return "a coarse element without 2^dim children";
  // W by columns, nonzeros only (a Q2 child of a Q2 parent: 125 of 729 in 3D)
  int stride = 1;
  for (int e = 0; e < f->num_elems; e++) {
    int nz = 0;
    for (int t = 0; t < nve * nve; t++) nz += W[(size_t)e * nve * nve + t] != 0.0;
    stride = std::max(stride, nz);
  }
  if (stride > (f->dim == 2 ? gal_ws_max<2>() : gal_ws_max<3>()))
    return "a child's nodal interpolation has more nonzeros than the Galerkin kernel keeps";
  std::vector<int> Wp((size_t)f->num_elems * (nve + 1)), Wi((size_t)f->num_elems * stride, 0);
  std::vector<double> Wv((size_t)f->num_elems * stride, 0.0);
  for (int e = 0; e < f->num_elems; e++) {
    const double *We = &W[(size_t)e * nve * nve];
    int k = 0;
    for (int m = 0; m < nve; m++) {
      Wp[(size_t)e * (nve + 1) + m] = k;
      for (int i = 0; i < nve; i++)
        if (We[i * nve + m] != 0.0) {
          Wi[(size_t)e * stride + k] = i;
          Wv[(size_t)e * stride + k] = We[i * nve + m];
          k++;
        }
    }
    Wp[(size_t)e * (nve + 1) + nve] = k;
  }
  L.W_stride = stride;
  if (upv(&L.d_child, child.data(), child.size()) || upv(&L.d_Wp, Wp.data(), Wp.size()) ||
      upv(&L.d_Wi, Wi.data(), Wi.size()) || upv(&L.d_Wv, Wv.data(), Wv.size()) ||
      upv(&L.d_Q, Q.data(), Q.size()))
    return "device allocation failed";
  return "";
}

// Coarse element matrices P_e^T K_e P_e summed over the children of each coarse
// element (one workgroup per coarse element), in the e2nz layout: nodal blocks
// (a, b) as [i][j], then momentum x P, continuity x velocity, P x P.
template <int DIM, int EN, typename CT>
__global__ __launch_bounds__(256) void k_galerkin(const double *__restrict__ kf,
                                                  const int *__restrict__ child,
                                                  const int *__restrict__ Wp,
                                                  const int *__restrict__ Wi,
                                                  const double *__restrict__ Wv, int ws,
                                                  const double *__restrict__ Q, long long E,
                                                  long long off_mp, double *__restrict__ kc) {
  constexpr int NVE = DIM == 2 ? 9 : 27, NP = DIM + 1, NV = DIM + EN, NCH = 1 << DIM, NT = 256;
  constexpr int NVE2 = NVE * NVE;
  constexpr int RN = (NVE2 + NT - 1) / NT;
  constexpr int NMP = DIM * NVE * NP;          // momentum x P (= continuity x velocity) entries
  constexpr int NPC = 2 * NMP + NP * NP;       // pressure couplings, contiguous from off_mp
  constexpr int RP = (NPC + NT - 1) / NT;
  constexpr int WSM = gal_ws_max<DIM>();
  // W of every child by columns, loaded once: s_wp[c][m] .. s_wp[c][m + 1]
  // index the finer nodes s_wi (ascending) interpolating from coarse node m
  // and their weights s_wv; only these terms of the dense sums are nonzero,
  // taken in the same order
  __shared__ CT sK[NVE2], sT[NVE2], sQ[NP * NP], sKp[NPC], s_wv[NCH * WSM];
  __shared__ int s_wp[NCH][NVE + 1], s_wi[NCH * WSM];
  const int ec = blockIdx.x, tid = threadIdx.x;
  const int *ch = child + (size_t)ec * NCH;
  double *__restrict__ out = kc + (size_t)ec * E;
  for (int t = tid; t < NCH * (NVE + 1); t += NT) {
    const int c = t / (NVE + 1), m = t - c * (NVE + 1);
    s_wp[c][m] = Wp[(size_t)ch[c] * (NVE + 1) + m];
  }
  for (int t = tid; t < NCH * WSM; t += NT) {
    const int c = t / WSM, k = t - c * WSM;
    if (k < ws) {
      s_wi[t] = Wi[(size_t)ch[c] * ws + k];
      s_wv[t] = (CT)Wv[(size_t)ch[c] * ws + k];
    }
  }
  // The finer element matrices' blocks, (variable block, child) in turn: the
  // next one is loaded into registers while the current one is reduced.
  double kr[RN];
  auto fetch = [&](int it) {
    const int blk = it / NCH, f = ch[it - blk * NCH];
    const double *Kf = kf + (size_t)f * E + (size_t)blk * NVE2;
#pragma unroll
    for (int r = 0; r < RN; r++) {
      const int t = tid + r * NT;
      kr[r] = t < NVE2 ? Kf[t] : 0.0;
    }
  };
  fetch(0);
  for (int blk = 0; blk < NV * NV; blk++) {
    CT acc[RN];
#pragma unroll
    for (int r = 0; r < RN; r++) acc[r] = 0.0;
    for (int sidx = 0; sidx < NCH; sidx++) {
#pragma unroll
      for (int r = 0; r < RN; r++) {
        const int t = tid + r * NT;
        if (t < NVE2) sK[t] = (CT)kr[r];
      }
      const int it = blk * NCH + sidx;
      if (it + 1 < NV * NV * NCH) fetch(it + 1);
      const int *wp = s_wp[sidx], *wi = s_wi + sidx * WSM;
      const CT *wv = s_wv + sidx * WSM;
      __syncthreads();
      for (int t = tid; t < NVE2; t += NT) {  // T = K W, column m of W
        const int m = t / NVE, i = t - m * NVE;
        CT v = 0;
        for (int k = wp[m]; k < wp[m + 1]; k++) v += sK[i * NVE + wi[k]] * wv[k];
        sT[i * NVE + m] = v;
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RN; r++) {  // acc += W^T T
        const int t = tid + r * NT;
        if (t < NVE2) {
          const int m = t / NVE, n = t - m * NVE;
          CT v = 0;
          for (int k = wp[m]; k < wp[m + 1]; k++) v += wv[k] * sT[wi[k] * NVE + n];
          acc[r] += v;
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < RN; r++) {
      const int t = tid + r * NT;
      if (t < NVE2) out[(size_t)blk * NVE2 + t] = acc[r];
    }
  }
  CT accp[RP];
#pragma unroll
  for (int r = 0; r < RP; r++) accp[r] = 0.0;
  for (int sidx = 0; sidx < NCH; sidx++) {
    const int f = ch[sidx];
    const double *Kf = kf + (size_t)f * E + off_mp;
    const int *wp = s_wp[sidx], *wi = s_wi + sidx * WSM;
    const CT *wv = s_wv + sidx * WSM;
    if (tid < NP * NP) sQ[tid] = (CT)Q[(size_t)f * NP * NP + tid];
    for (int t = tid; t < NPC; t += NT) sKp[t] = (CT)Kf[t];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RP; r++) {
      const int t = tid + r * NT;
      if (t >= NPC) continue;
      CT v = 0;
      if (t < NMP) {  // row (a, m) momentum, column l pressure
        const int a = t / (NVE * NP), rem = t - a * NVE * NP, m = rem / NP, l = rem - m * NP;
        for (int k2 = wp[m]; k2 < wp[m + 1]; k2++) {
          const int i = wi[k2];
          CT q = 0;
#pragma unroll
          for (int k = 0; k < NP; k++) q += sKp[a * NVE * NP + i * NP + k] * sQ[k * NP + l];
          v += wv[k2] * q;
        }
      } else if (t < 2 * NMP) {  // row l continuity, column (b, n) velocity
        const int t2 = t - NMP, b = t2 / (NP * NVE), rem = t2 - b * NP * NVE;
        const int l = rem / NVE, n = rem - l * NVE;
#pragma unroll
        for (int k = 0; k < NP; k++) {
          CT q = 0;
          for (int k2 = wp[n]; k2 < wp[n + 1]; k2++)
            q += sKp[NMP + b * NP * NVE + k * NVE + wi[k2]] * wv[k2];
          v += sQ[k * NP + l] * q;
        }
      } else {  // P x P
        const int t3 = t - 2 * NMP, l = t3 / NP, l2 = t3 - l * NP;
#pragma unroll
        for (int k = 0; k < NP; k++)
#pragma unroll
          for (int k2 = 0; k2 < NP; k2++)
            v += sQ[k * NP + l] * sKp[2 * NMP + k * NP + k2] * sQ[k2 * NP + l2];
      }
      accp[r] += v;
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < RP; r++) {
    const int t = tid + r * NT;
    if (t < NPC) out[off_mp + t] = accp[r];
  }
}

// Adds the element matrices of one element colour into an MSR matrix with
// Goma's per-element Dirichlet rows (a Dirichlet row keeps only a unit
// diagonal, src/bc_dirich.c:44-153), as the assembly kernel does.
template <int DIM, typename JT>
__global__ __launch_bounds__(256) void k_elem_scatter(const double *__restrict__ kel,
                                                      const int *__restrict__ elems, int nv,
                                                      const int *__restrict__ conn,
                                                      const int *__restrict__ fu,
                                                      const int *__restrict__ centroid,
                                                      const signed char *__restrict__ dbc,
                                                      const int *__restrict__ e2nz, long long E,
                                                      long long off_mp, long long off_cv,
                                                      long long off_cp, JT *__restrict__ val) {
  constexpr int NVE = DIM == 2 ? 9 : 27, NP = DIM + 1, NVE2 = NVE * NVE;
  const int e = elems[blockIdx.x];
  __shared__ int s_f[NVE];
  __shared__ int s_pc;
  if (threadIdx.x < NVE) s_f[threadIdx.x] = fu[conn[(size_t)e * NVE + threadIdx.x]];
  if (threadIdx.x == 0) s_pc = fu[centroid[e]] + nv;
  __syncthreads();
  const double *__restrict__ K = kel + (size_t)e * E;
  const int *__restrict__ map = e2nz + (size_t)e * E;
  for (long long idx = threadIdx.x; idx < E; idx += blockDim.x) {
    int row;
    bool diag;
    if (idx < off_mp) {
      const int blk = (int)(idx / NVE2), t = (int)(idx - (long long)blk * NVE2);
      const int a = blk / nv, b = blk - a * nv, i = t / NVE, j = t - i * NVE;
      row = s_f[i] + a;
      diag = i == j && a == b;
    } else if (idx < off_cv) {
      const int r = (int)(idx - off_mp);
      const int a = r / (NVE * NP), t = r - a * NVE * NP;
      row = s_f[t / NP] + a;
      diag = false;
    } else if (idx < off_cp) {
      const int t = (int)((idx - off_cv) % (NP * NVE));
      row = s_pc + t / NVE;
      diag = false;
    } else {
      const int r = (int)(idx - off_cp);
      row = s_pc + r / NP;
      diag = r / NP == r % NP;
    }
    double v = K[idx];
    if (dbc[row]) {
      if (!diag) continue;
      v = 1.0;
    }
    val[map[idx]] += (JT)v;
  }
}

template <typename JT>
static void galerkin_level(gm_ctx *prof, const gm_ctx *above, MgLevel &L, JT *val) {
  gm_ctx *lc = L.ctx;
  const ElemLayout &ly = lc->lay;
  const int nch = 1 << lc->dim;
  const double gbytes = ((double)above->num_elems * (ly.total + lc->np * lc->np) +
                         (double)lc->num_elems * ly.total) * 8.0 + 4.0 * nch * lc->num_elems +
                        (double)above->num_elems * (L.W_stride * 12.0 + (lc->nve + 1) * 4.0);
  // products in the precision of the coarse matrices (FP32 under mixed precision)
  using CT = JT;
  auto kern = lc->dim == 2 ? (lc->energy ? k_galerkin<2, 1, CT> : k_galerkin<2, 0, CT>)
                           : (lc->energy ? k_galerkin<3, 1, CT> : k_galerkin<3, 0, CT>);
  gm_launch(prof, GM_K_GALERKIN, gbytes, kern, dim3(lc->num_elems), dim3(256), 0, lc->stream,
            above->d_kel, L.d_child, L.d_Wp, L.d_Wi, L.d_Wv, L.W_stride, L.d_Q, ly.total, ly.off_mp,
            lc->d_kel);
  for (int col = 0; col < lc->n_ecol; col++) {
    const int cnt = lc->ecol_ptr[col + 1] - lc->ecol_ptr[col];
    if (!cnt) continue;
    const double sbytes = (double)cnt * ly.total * (8.0 + 4.0 + 2.0 * sizeof(JT));
    auto kern = lc->dim == 2 ? k_elem_scatter<2, JT> : k_elem_scatter<3, JT>;
    gm_launch(prof, GM_K_ELEM_SCATTER, sbytes, kern, dim3(cnt), dim3(256), 0, lc->stream,
              (const double *)lc->d_kel, (const int *)(lc->d_ecol_elems + lc->ecol_ptr[col]), lc->nv,
              (const int *)lc->d_conn, (const int *)lc->d_fu, (const int *)lc->d_centroid,
              (const signed char *)lc->d_dbc, (const int *)lc->d_e2nz, ly.total, ly.off_mp,
              ly.off_cv, ly.off_cp, val);
  }
}

int gm_mg_kel_alloc(gm_ctx *c) {
  if (!c->d_kel) GM_HIP(hipMalloc(&c->d_kel, (size_t)c->num_elems * c->lay.total * sizeof(double)));
  return GM_OK;
}

extern "C" int gm_mg_copy_level_matrix(const gm_ctx *c, int level, double *val) {
  if (!c || !val || level < 1 || level > (int)c->mg.size()) return GM_ERR_ARG;
  const MgLevel &L = c->mg[level - 1];
  if (!L.d_val) {
    gm_set_error("multigrid level %d has no matrix yet", level);
    return GM_ERR_ARG;
  }
  const size_t n = L.ctx->nval;
  if (L.d_val == L.d_val_p[1]) {
    std::vector<float> h(n);
    GM_HIP(hipMemcpy(h.data(), L.d_val, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) val[i] = h[i];
  } else {
    GM_HIP(hipMemcpy(val, L.d_val, n * sizeof(double), hipMemcpyDeviceToHost));
  }
  return GM_OK;
}

extern "C" int gm_mg_copy_block_inverse(const gm_ctx *c, int level, int elem, double *out) {
  if (!c || !out || level < 0 || level > (int)c->mg.size()) return GM_ERR_ARG;
  const MgLevel &L = level == 0 ? c->mg0 : c->mg[level - 1];
  const gm_ctx *lc = level == 0 ? c : L.ctx;
  if (elem < 0 || elem >= lc->num_elems) return GM_ERR_ARG;
  if (!L.d_ainv) {
    gm_set_error("multigrid level %d has no element-block inverses yet", level);
    return GM_ERR_ARG;
  }
  const size_t n = nloc_of(lc), off = (size_t)elem * n * n;
  if (L.d_ainv == L.d_ainv_p[1]) {
    std::vector<float> h(n * n);
    GM_HIP(hipMemcpy(h.data(), (const float *)L.d_ainv + off, n * n * sizeof(float),
                     hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n * n; i++) out[i] = h[i];
  } else {
    GM_HIP(hipMemcpy(out, (const double *)L.d_ainv + off, n * n * sizeof(double),
                     hipMemcpyDeviceToHost));
  }
  return (int)n;
}

extern "C" int gm_mg_copy_coarse_inverse(const gm_ctx *c, double *out) {
  if (!c || !out || c->mg.empty()) return GM_ERR_ARG;
  const MgLevel &L = c->mg.back();
  if (!L.d_dense || L.dense_n != L.ctx->N) {
    gm_set_error("the coarsest multigrid level has no dense inverse yet (no set-up with an "
                 "exact coarse solve)");
    return GM_ERR_ARG;
  }
  const size_t n = (size_t)L.dense_n;
  GM_HIP(hipMemcpy(out, L.d_dense, n * n * sizeof(double), hipMemcpyDeviceToHost));
  return (int)n;
}

extern "C" int gm_mg_add_level(gm_ctx *fine, gm_ctx *coarse, const int *P_rowptr, const int *P_col,
                               const double *P_val, const int *inject) {
  if (!fine || !coarse || !P_rowptr || !P_col || !P_val || !inject) return GM_ERR_ARG;
  gm_ctx *above = fine->mg.empty() ? fine : fine->mg.back().ctx;
  if (coarse->dim != fine->dim || coarse->energy != fine->energy) {
    gm_set_error("coarse level has a different dimension or variable set");
    return GM_ERR_ARG;
  }
  MgLevel L;
  L.ctx = coarse;
  const int Nf = above->N, Nc = coarse->N;
  const int nnz = P_rowptr[Nf];
  for (int k = 0; k < nnz; k++)
    if (P_col[k] < 0 || P_col[k] >= Nc) {
      gm_set_error("prolongation column %d outside the coarse level", P_col[k]);
      return GM_ERR_ARG;
    }
  for (int n = 0; n < coarse->num_nodes; n++)
    if (inject[n] < 0 || inject[n] >= above->num_nodes) {
      gm_set_error("injection map entry outside the finer level");
      return GM_ERR_ARG;
    }
  // R = P^T
  std::vector<int> rrow(Nc + 1, 0), rcol(nnz);
  std::vector<double> rval(nnz);
  for (int k = 0; k < nnz; k++) rrow[P_col[k] + 1]++;
  for (int i = 0; i < Nc; i++) rrow[i + 1] += rrow[i];
  std::vector<int> fp(rrow.begin(), rrow.end() - 1);
  for (int i = 0; i < Nf; i++)
    for (int k = P_rowptr[i]; k < P_rowptr[i + 1]; k++) {
      const int q = fp[P_col[k]]++;
      rcol[q] = i;
      rval[q] = P_val[k];
    }
  int rc;
  if ((rc = upv(&L.d_Prow, P_rowptr, Nf + 1)) || (rc = upv(&L.d_Pcol, P_col, nnz)) ||
      (rc = upv(&L.d_Pval, P_val, nnz)) || (rc = upv(&L.d_Rrow, rrow.data(), Nc + 1)) ||
      (rc = upv(&L.d_Rcol, rcol.data(), nnz)) || (rc = upv(&L.d_Rval, rval.data(), nnz)) ||
      (rc = upv(&L.d_inject, inject, coarse->num_nodes)))
    return rc;
  L.nnz = nnz;
  L.gal_err = build_galerkin_maps(above, coarse, P_rowptr, P_col, P_val, L);
  L.gal_ok = L.gal_err.empty();
  GM_HIP(hipMalloc(&L.d_state, Nc * sizeof(double)));
  GM_HIP(hipMalloc(&L.d_rdummy, Nc * sizeof(double)));
  fine->mg.push_back(L);
  return GM_OK;
}

extern "C" int gm_mg_num_levels(const gm_ctx *c) { return c ? (int)c->mg.size() + 1 : GM_ERR_ARG; }

// Level vectors are FP64 in both precisions; matrices and element-block
// inverses are FP32 in the mixed precision.
static int alloc_level_vectors(MgLevel &L, const gm_ctx *lc, int prec, bool coarse) {
  const size_t N = lc->N;
  const size_t es = prec == GM_PREC_MIXED ? sizeof(float) : sizeof(double);
  if (!L.d_x) {
    GM_HIP(hipMalloc(&L.d_x, N * sizeof(double)));
    GM_HIP(hipMalloc(&L.d_b, N * sizeof(double)));
    GM_HIP(hipMalloc(&L.d_r, N * sizeof(double)));
    L.vec_prec = 0;
  }
  const size_t nl = nloc_of(lc);
  const int pi = prec == GM_PREC_MIXED ? 1 : 0;
  if (!L.d_ainv_p[pi]) GM_HIP(hipMalloc(&L.d_ainv_p[pi], (size_t)lc->num_elems * nl * nl * es));
  L.d_ainv = L.d_ainv_p[pi];
  if (coarse) {
    if (!L.d_val_p[pi]) GM_HIP(hipMalloc(&L.d_val_p[pi], (size_t)lc->nval * es));
    L.d_val = L.d_val_p[pi];
  }
  return GM_OK;
}

template <int DIM, int EN, typename VT, typename AT>
static void launch_invert(gm_ctx *prof, gm_ctx *lc, hipStream_t s, const VT *val, AT *ainv) {
  constexpr int n = Loc<DIM, EN>::N;
  // gathered entries + map + inverses written
  const double bytes = (double)lc->num_elems * n * n * (sizeof(VT) + 4 + sizeof(AT));
  if constexpr (DIM == 3) {
    // 3D blocks (n = 85 / 112) in registers, 16 x 16 threads: three
    // workgroups per CU (164 VGPRs) overlap their pivot steps' barriers
    gm_launch(prof, GM_K_VANKA_INV, bytes, k_vanka_invert_reg<DIM, EN, VT, AT, AT, 16, 16>,
              dim3(lc->num_elems), dim3(256), 0, s, level_dev(lc), val, ainv);
  } else {
    gm_launch(prof, GM_K_VANKA_INV, bytes, k_vanka_invert<DIM, EN, VT, AT, AT, 128>,
              dim3(lc->num_elems), dim3(128), 0, s, level_dev(lc), val, ainv);
  }
}

template <typename VT>
static void invert(gm_ctx *prof, gm_ctx *lc, hipStream_t s, const VT *val, VT *ainv) {
  if (lc->dim == 2) {
    if (lc->energy) launch_invert<2, 1, VT, VT>(prof, lc, s, val, ainv);
    else launch_invert<2, 0, VT, VT>(prof, lc, s, val, ainv);
  } else {
    if (lc->energy) launch_invert<3, 1, VT, VT>(prof, lc, s, val, ainv);
    else launch_invert<3, 0, VT, VT>(prof, lc, s, val, ainv);
  }
}

// Rows touched by each element colour: the unknowns of its elements.
static int build_colour_rows(gm_ctx *lc) {
  if (lc->d_mg_rows) return GM_OK;
  const int nve = lc->nve, cl = lc->dim == 2 ? 8 : 20;
  std::vector<int> rows;
  std::vector<char> seen(lc->N, 0);
  lc->mg_rows_ptr.assign(1, 0);
  lc->mg_rows_nnz.clear();
  for (int col = 0; col < lc->n_ecol; col++) {
    std::vector<int> list;
    for (int k = lc->ecol_ptr[col]; k < lc->ecol_ptr[col + 1]; k++) {
      const int e = lc->ecol_elems[k];
      for (int i = 0; i < nve; i++) {
        const int n = lc->conn[(size_t)e * nve + i];
        const int cnt = (i == cl) ? lc->first_unknown[n + 1] - lc->first_unknown[n] : lc->nv;
        for (int a = 0; a < cnt; a++) list.push_back(lc->first_unknown[n] + a);
      }
    }
    std::sort(list.begin(), list.end());
    list.erase(std::unique(list.begin(), list.end()), list.end());
    long long nnz = 0;
    for (int r : list) nnz += 1 + lc->bindx[r + 1] - lc->bindx[r];
    rows.insert(rows.end(), list.begin(), list.end());
    lc->mg_rows_ptr.push_back((int)rows.size());
    lc->mg_rows_nnz.push_back(nnz);
  }
  // multiplicity of every unknown among the element blocks
  std::vector<double> w(lc->N, 0.0);
  for (int e = 0; e < lc->num_elems; e++)
    for (int i = 0; i < nve; i++) {
      const int n = lc->conn[(size_t)e * nve + i];
      const int cnt = (i == cl) ? lc->first_unknown[n + 1] - lc->first_unknown[n] : lc->nv;
      for (int a = 0; a < cnt; a++) w[lc->first_unknown[n] + a] += 1.0;
    }
  for (double &v : w) v = v > 0.0 ? 1.0 / v : 0.0;
  int rc = upv(&lc->d_mg_wmul, w.data(), w.size());
  if (rc) return rc;
  // gather lists of the additive smoother: local dof li of element e (nodal
  // variables node-blocked by variable, then the element's P1 dofs) is unknown
  // g; its correction sits at e * nl + li
  const int nl = lc->nv * nve + lc->np;
  std::vector<int> gptr(lc->N + 1, 0), gent((size_t)lc->num_elems * nl);
  auto dof = [&](int e, int li) {
    if (li < lc->nv * nve) {
      const int a = li / nve, i = li - a * nve;
      return lc->first_unknown[lc->conn[(size_t)e * nve + i]] + a;
    }
    return lc->first_unknown[lc->elem_centroid[e]] + lc->dim + lc->energy + (li - lc->nv * nve);
  };
  for (int e = 0; e < lc->num_elems; e++)
    for (int li = 0; li < nl; li++) gptr[dof(e, li) + 1]++;
  for (int u = 0; u < lc->N; u++) gptr[u + 1] += gptr[u];
  std::vector<int> fill(gptr.begin(), gptr.end() - 1);
  for (int e = 0; e < lc->num_elems; e++)
    for (int li = 0; li < nl; li++) gent[fill[dof(e, li)]++] = e * nl + li;
  if ((rc = upv(&lc->d_mg_gptr, gptr.data(), gptr.size())) ||
      (rc = upv(&lc->d_mg_gent, gent.data(), gent.size())))
    return rc;
  GM_HIP(hipMalloc(&lc->d_mg_corr, gent.size() * sizeof(double)));
  return upv(&lc->d_mg_rows, rows.data(), rows.size());
}

// Dense inverse of the coarsest level's Jacobian (exact coarse solve),
// enqueued on stream s; coarse_inverse_finish waits for it and reports a
// singular matrix.
static int coarse_inverse(gm_ctx *c, MgLevel &L, int prec, hipStream_t s) {
  gm_ctx *lc = L.ctx;
  const int n = lc->N;
  if (n > GM_MG_DIRECT_MAX) {
    gm_set_error("coarsest multigrid level has %d unknowns (exact solve up to %d): coarsen "
                 "further or set mg_coarse_sweeps > 0", n, GM_MG_DIRECT_MAX);
    return GM_ERR_ARG;
  }
  if (L.dense_n != n) {
    for (void *p : {(void *)L.d_dense, (void *)L.d_pcol, (void *)L.d_perm})
      if (p) (void)hipFree(p);
    L.d_dense = L.d_pcol = nullptr;
    L.d_perm = nullptr;
    L.dense_n = 0;
    GM_HIP(hipMalloc(&L.d_dense, (size_t)n * n * sizeof(double)));
    GM_HIP(hipMalloc(&L.d_pcol, (size_t)GJ_NB * n * sizeof(double)));  // pivot rows B_K
    GM_HIP(hipMalloc(&L.d_perm, (size_t)n * sizeof(int)));
    L.dense_n = n;
  }
  const long long nn = (long long)n * n;
  k_set<double><<<1024, 256, 0, s>>>(L.d_dense, nn, 0.0);
  if (prec == GM_PREC_MIXED)
    k_msr_to_dense<float><<<n, 64, 0, s>>>(n, (const float *)L.d_val, lc->d_bindx, L.d_dense);
  else
    k_msr_to_dense<double><<<n, 64, 0, s>>>(n, (const double *)L.d_val, lc->d_bindx, L.d_dense);
  GM_HIP(hipMemsetAsync(c->d_flag + 3, 0, sizeof(int), s));
  const int nt = (n + 63) / 64;
  // panel width x rows per thread (512 threads; wider panels left the step
  // loop partly rolled, the panel in scratch)
  const int pw = n <= 1536 ? 16 : 8;  // 16-wide panels of more rows would spill
  for (int kb0 = 0; kb0 < n; kb0 += pw) {
    const int nb = std::min(pw, n - kb0);
    // panel: read and written once; swap: the panel rows of the other columns
    // read, written and copied; update: the matrix read and written
    const double pb = 16.0 * n * nb, sb = 24.0 * n * nb, ub = 16.0 * (double)n * n;
    auto panel = n <= 512 ? k_gjb_panel<16, 1> : n <= 1024 ? k_gjb_panel<16, 2>
               : n <= 1536 ? k_gjb_panel<16, 3> : k_gjb_panel<8, 8>;
    gm_launch(c, GM_K_COARSE_INV, pb, panel, dim3(1), dim3(512), 0, s, n, kb0, nb, L.d_dense,
              L.d_perm, c->d_flag);
    gm_launch(c, GM_K_COARSE_INV, sb, k_gjb_swap, dim3((n + 255) / 256), dim3(256), 0, s, n, kb0,
              nb, L.d_dense, (const int *)L.d_perm, L.d_pcol);
    gm_launch(c, GM_K_COARSE_INV, ub, k_gjb_update, dim3(nt, nt), dim3(256), 0, s, n, kb0, nb,
              L.d_dense, (const double *)L.d_pcol);
  }
  k_gj_unpermute<<<(n + 255) / 256, 256, 0, s>>>(n, L.d_perm, L.d_dense);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

static int coarse_inverse_finish(gm_ctx *c, hipStream_t s) {
  int flag = 0;
  GM_HIP(hipMemcpyAsync(&flag, c->d_flag + 3, sizeof(int), hipMemcpyDeviceToHost, s));
  GM_HIP(hipStreamSynchronize(s));
  if (flag) {
    gm_set_error("singular Jacobian on the coarsest multigrid level");
    return GM_ERR_PIVOT;
  }
  return GM_OK;
}

// Stream of the overlapped element-block inversions, kept off 1/8 of the CUs.
// The inversions fill every CU they may use (three workgroups of 147-VGPR
// waves each), so a workgroup of another stream is placed only where they
// leave room; the coarsest dense inverse is a chain of small launches, each
// of which waited on a full GPU for the inversions to drain.  The spared CUs
// keep that chain moving beside them.  Without the mask the main stream is
// used.
#ifndef GM_INV_SPARE_DIV
#define GM_INV_SPARE_DIV 8  // CUs spared: 1 in this many (1/16 and 1/12 measured slower)
#endif
static int inv_stream(gm_ctx *c, hipStream_t *out) {
  if (!c->s_inv) {
    int dev = 0, ncu = 0;
    GM_HIP(hipGetDevice(&dev));
    GM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    hipStream_t s = nullptr;
    if (ncu >= 64) {
      std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
      const int keep = ncu - ncu / GM_INV_SPARE_DIV;
      for (int i = 0; i < keep; i++) mask[i / 32] |= 1u << (i % 32);
      if (hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
        (void)hipGetLastError();
        s = nullptr;
      }
    }
    if (s && (hipEventCreateWithFlags(&c->ev_inv[0], hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&c->ev_inv[1], hipEventDisableTiming) != hipSuccess)) {
      (void)hipGetLastError();
      (void)hipStreamDestroy(s);
      s = nullptr;
    }
    c->s_inv = s ? s : c->stream;
  }
  *out = c->s_inv;
  return GM_OK;
}

// Per Newton step: coarse Jacobians at the injected state, local inverses on
// every level, and the dense inverse of the coarsest level when it is solved
// exactly (coarse_direct).  val / val32 is the level-0 (scaled) matrix, x the
// state.
int gm_mg_setup(gm_ctx *c, const double *x, const double *val, const float *val32, int prec,
                const double *scale, int coarse_direct, int galerkin) {
  c->mg_scale = scale;
  if (c->mg.empty()) {
    gm_set_error("multigrid preconditioner requested but no coarse level attached");
    return GM_ERR_ARG;
  }
  int rc;
  if ((rc = alloc_level_vectors(c->mg0, c, prec, false))) return rc;
  if ((rc = build_colour_rows(c))) return rc;
  for (MgLevel &L : c->mg)
    if ((rc = build_colour_rows(L.ctx))) return rc;
  const double *xs = x;
  gm_ctx *above = c;
  if (galerkin && !c->d_kel) {
    gm_set_error("Galerkin coarse operators need the element matrices of the last assembly");
    return GM_ERR_ARG;
  }
  // The level-0 block inverses need only the level-0 matrix: enqueued first,
  // they run beside the coarse levels' Galerkin products and scatters, which
  // use each coarse level's own stream and synchronise only that stream, and
  // beside the coarsest dense inverse.  Beside a long chain of dense-inverse
  // panels (over 32) they use the stream kept off some CUs, ordered after the
  // main stream's assembly; otherwise the main stream.  With kernel timing on,
  // the set-up runs in order on the main stream instead, so that each
  // launch's measured duration is its own.
  const bool overlap = !c->prof;
  hipStream_t si = c->stream;
  if (overlap && coarse_direct && c->mg.back().ctx->N > 512 && (rc = inv_stream(c, &si)))
    return rc;
  if (si != c->stream) {
    GM_HIP(hipEventRecord(c->ev_inv[0], c->stream));
    GM_HIP(hipStreamWaitEvent(si, c->ev_inv[0], 0));
  }
  // on every return from here, what the main stream queues next (the V-cycles,
  // or the next assembly over the matrix being inverted) waits for the
  // inversions
  struct Join {
    gm_ctx *c;
    hipStream_t si;
    ~Join() {
      if (si == c->stream) return;
      if (hipEventRecord(c->ev_inv[1], si) != hipSuccess ||
          hipStreamWaitEvent(c->stream, c->ev_inv[1], 0) != hipSuccess)
        (void)hipStreamSynchronize(si);
    }
  } join{c, si};
  auto invert0 = [&]() {
    if (prec == GM_PREC_MIXED)
      invert<float>(c, c, si, val32, (float *)c->mg0.d_ainv);
    else
      invert<double>(c, c, si, val, (double *)c->mg0.d_ainv);
  };
  if (overlap) invert0();
  GM_HIP(hipGetLastError());
  for (MgLevel &L : c->mg) {
    if (galerkin) {
      gm_ctx *lc = L.ctx;
      if (!L.gal_ok) {
        gm_set_error("Galerkin coarse operators unavailable: %s", L.gal_err.c_str());
        return GM_ERR_ARG;
      }
      if ((rc = alloc_level_vectors(L, lc, prec, true)) || (rc = gm_mg_kel_alloc(lc))) return rc;
      if (prec == GM_PREC_MIXED) {
        k_set<float><<<2048, 256, 0, lc->stream>>>((float *)L.d_val, lc->nval, 0.0f);
        galerkin_level<float>(c, above, L, (float *)L.d_val);
      } else {
        k_set<double><<<2048, 256, 0, lc->stream>>>((double *)L.d_val, lc->nval, 0.0);
        galerkin_level<double>(c, above, L, (double *)L.d_val);
      }
      GM_HIP(hipGetLastError());
      GM_HIP(hipStreamSynchronize(lc->stream));
      above = lc;
      continue;
    }
    gm_ctx *lc = L.ctx;
    if ((rc = alloc_level_vectors(L, lc, prec, true))) return rc;
    k_inject<<<1024, 256, 0, c->stream>>>(lc->num_nodes, L.d_inject, above->d_fu, lc->d_fu, lc->nv,
                                          xs, L.d_state);
    GM_HIP(hipGetLastError());
    GM_HIP(hipStreamSynchronize(c->stream));
    k_set<double><<<256, 256, 0, lc->stream>>>(L.d_rdummy, lc->N, 0.0);
    if (prec == GM_PREC_MIXED) {
      k_set<float><<<2048, 256, 0, lc->stream>>>((float *)L.d_val, lc->nval, 0.0f);
      rc = gm_fill_launch(lc, L.d_state, L.d_rdummy, nullptr, (float *)L.d_val);
    } else {
      k_set<double><<<2048, 256, 0, lc->stream>>>((double *)L.d_val, lc->nval, 0.0);
      rc = gm_fill_launch(lc, L.d_state, L.d_rdummy, (double *)L.d_val, nullptr);
    }
    if (rc) return rc;
    xs = L.d_state;
    above = lc;
  }
  auto invert_coarse = [&]() {
    if (prec == GM_PREC_MIXED) {
      for (MgLevel &L : c->mg)
        invert<float>(c, L.ctx, si, (const float *)L.d_val, (float *)L.d_ainv);
    } else {
      for (MgLevel &L : c->mg)
        invert<double>(c, L.ctx, si, (const double *)L.d_val, (double *)L.d_ainv);
    }
  };
  // The coarse levels' block inverses are queued behind the level-0 ones
  // before the dense inverse's chain of launches: queueing that chain holds the
  // host until most of it has run.  The coarsest level's dense inverse (one
  // workgroup per panel step, latency bound) runs on that level's own stream,
  // beside the element-block inversions; its matrix is complete here (the
  // Galerkin products were synchronised; a rediscretised matrix was filled and
  // checked on the host).
  if (overlap) invert_coarse();
  hipStream_t cs = overlap ? c->mg.back().ctx->stream : c->stream;
  if (coarse_direct && (rc = coarse_inverse(c, c->mg.back(), prec, cs))) return rc;
  if (!overlap) {
    invert0();
    invert_coarse();
  }
  GM_HIP(hipGetLastError());
  // the V-cycles on the main stream read the dense inverse after this host wait
  // and the block inverses after join's event
  if (coarse_direct && (rc = coarse_inverse_finish(c, cs))) return rc;
  return GM_OK;
}

static gm_ctx *g_prof_ctx = nullptr;  // context whose counters the V-cycle feeds

// lanes per row: 2D rows hold ~50-250 entries (8 lanes: 33 vs 36 us per 256^2 residual
// with 16), 3D rows ~500-2100
template <typename VT, typename BT, int G>
static void resid_g(gm_ctx *lc, hipStream_t s, const VT *a, const BT *x, const BT *b, BT *r) {
  const int blocks = (int)(((long long)lc->N * G + 255) / 256);
  const double bytes = (double)(lc->nval - 1) * (sizeof(VT) + 4) + 4.0 * (lc->N + 1) +
                       3.0 * lc->N * sizeof(BT);
  gm_launch(g_prof_ctx, GM_K_MG_RESID, bytes, k_mg_resid<VT, BT, G>, dim3(blocks), dim3(256), 0, s,
            lc->N, a, lc->d_bindx, x, b, r);
}

template <typename VT, typename BT>
static void resid(gm_ctx *lc, hipStream_t s, const VT *a, const BT *x, const BT *b, BT *r) {
  if (lc->dim == 2 || lc->nn_cap > GM_NODE_CAP_MAX) {  // G lanes per row
    if (lc->dim == 3) resid_g<VT, BT, 64>(lc, s, a, x, b, r);
    else resid_g<VT, BT, 8>(lc, s, a, x, b, r);
    return;
  }
  // HEX27: one wave per node (k_msr_nodes)
  using AT = typename std::conditional<sizeof(BT) == 8, double, float>::type;
  const double bytes = (double)(lc->nval - 1) * sizeof(VT) +
                       4.0 * (2.0 * (lc->num_nodes + 1) + lc->N + lc->nn_list.size()) +
                       3.0 * lc->N * sizeof(BT);
  gm_launch(g_prof_ctx, GM_K_MG_RESID, bytes, k_msr_nodes<AT, VT, BT, BT, true>,
            dim3((lc->num_nodes + 3) / 4), dim3(256), (size_t)4 * 2 * lc->nn_cap * sizeof(int), s,
            lc->num_nodes, lc->dim, lc->nv, lc->nn_cap, lc->d_fu, lc->d_nn_ptr, lc->d_nn_list,
            lc->d_bindx, a, x, b, r);
}

template <typename VT, typename BT, int G>
static void resid_rows_g(gm_ctx *lc, hipStream_t s, int col, const VT *a, const BT *x, const BT *b,
                         BT *r) {
  const int nr = lc->mg_rows_ptr[col + 1] - lc->mg_rows_ptr[col];
  if (!nr) return;
  const int blocks = (int)(((long long)nr * G + 255) / 256);
  const double nnz_rows = (double)lc->mg_rows_nnz[col];
  const double bytes = nnz_rows * (sizeof(VT) + 4) + nr * (4.0 * 3 + 3.0 * sizeof(BT));
  gm_launch(g_prof_ctx, GM_K_MG_RESID_ROWS, bytes, k_mg_resid_rows<VT, BT, G>, dim3(blocks), dim3(256),
            0, s, nr, (const int *)(lc->d_mg_rows + lc->mg_rows_ptr[col]), a, (const int *)lc->d_bindx,
            x, b, r);
}

template <typename VT, typename BT>
static void resid_rows(gm_ctx *lc, hipStream_t s, int col, const VT *a, const BT *x, const BT *b,
                       BT *r) {
  if (lc->dim == 3) resid_rows_g<VT, BT, 64>(lc, s, col, a, x, b, r);
  else resid_rows_g<VT, BT, 16>(lc, s, col, a, x, b, r);
}

template <int DIM, int EN, typename AT, typename BT>
static void vanka_colour(gm_ctx *lc, hipStream_t s, int col, const AT *ainv, const BT *r, BT *x,
                         double omega, const double *wmul) {
  constexpr int NT = DIM == 2 ? 64 : 128;
  constexpr int n = Loc<DIM, EN>::N;
  const int cnt = lc->ecol_ptr[col + 1] - lc->ecol_ptr[col];
  if (!cnt) return;
  // inverses, element lists/connectivity, residual gathered, x read and written
  const double bytes = (double)cnt * (n * n * sizeof(AT) + 4.0 * (Loc<DIM, EN>::NVE + 2) +
                                      n * 3.0 * sizeof(BT));
  constexpr int EPB = (2 * n <= NT) ? 2 : 1;
  gm_launch(g_prof_ctx, GM_K_VANKA, bytes, k_vanka_update<DIM, EN, AT, BT, NT>, dim3((cnt + EPB - 1) / EPB), dim3(NT),
            0, s, level_dev(lc), cnt, (const int *)(lc->d_ecol_elems + lc->ecol_ptr[col]), ainv, r, x,
            omega, wmul);
}

template <int DIM, int EN, typename AT, typename BT>
static void vanka_additive(gm_ctx *lc, hipStream_t s, const AT *ainv, const BT *r, BT *x, double omega,
                           bool assign) {
  constexpr int NT = DIM == 2 ? 64 : 128;
  constexpr int n = Loc<DIM, EN>::N;
  constexpr int EPB = (2 * n <= NT) ? 2 : 1;
  const int ne = lc->num_elems;
  // inverses, connectivity, residual gathered, corrections written
  const double bytes = (double)ne * (n * n * sizeof(AT) + 4.0 * (Loc<DIM, EN>::NVE + 1) +
                                     n * (sizeof(BT) + 8.0));
  gm_launch(g_prof_ctx, GM_K_VANKA, bytes, k_vanka_corr<DIM, EN, AT, BT, NT>, dim3((ne + EPB - 1) / EPB),
            dim3(NT), 0, s, level_dev(lc), ne, ainv, r, lc->d_mg_corr);
  // lists, corrections, weights, x read and written
  const double gbytes = (double)ne * n * 12.0 + 4.0 * (lc->N + 1) + lc->N * (8.0 + 2.0 * sizeof(BT));
  gm_launch(g_prof_ctx, GM_K_VANKA_GATHER, gbytes, k_vanka_gather<BT>, dim3((lc->N + 255) / 256),
            dim3(256), 0, s, lc->N, (const int *)lc->d_mg_gptr, (const int *)lc->d_mg_gent,
            (const double *)lc->d_mg_corr, (const double *)lc->d_mg_wmul, omega, x, assign ? 1 : 0);
}

// Multiplicative: coloured block Gauss-Seidel, the residual recomputed on each
// colour's rows.  Additive: one residual per sweep, every element's correction
// from it (one launch), then each unknown sums its corrections weighted by
// 1 / elements sharing it (a second launch; no colours needed).  x_zero:
// smoothing from x = 0 (every level's pre-smoothing in the V-cycle): the first
// additive sweep takes b as its residual and assigns its corrections to x,
// whose entries it never reads (the multiplicative smoother gets x cleared).
template <typename VT, typename BT>
static void smooth(gm_ctx *lc, hipStream_t s, const VT *a, const VT *ainv, const BT *b, BT *x, BT *r,
                   int sweeps, double omega, int additive, bool x_zero) {
  for (int it = 0; it < sweeps; it++) {
    if (additive) {
      const BT *rr = r;
      const bool first = it == 0 && x_zero;  // x = 0: residual b, and x is assigned
      if (first) rr = b;
      else resid<VT, BT>(lc, s, a, x, b, r);
      if (lc->dim == 2) {
        if (lc->energy) vanka_additive<2, 1, VT, BT>(lc, s, ainv, rr, x, omega, first);
        else vanka_additive<2, 0, VT, BT>(lc, s, ainv, rr, x, omega, first);
      } else {
        if (lc->energy) vanka_additive<3, 1, VT, BT>(lc, s, ainv, rr, x, omega, first);
        else vanka_additive<3, 0, VT, BT>(lc, s, ainv, rr, x, omega, first);
      }
      continue;
    }
    const double *w = nullptr;
    for (int col = 0; col < lc->n_ecol; col++) {
      resid_rows<VT, BT>(lc, s, col, a, x, b, r);
      if (lc->dim == 2) {
        if (lc->energy) vanka_colour<2, 1, VT, BT>(lc, s, col, ainv, r, x, omega, w);
        else vanka_colour<2, 0, VT, BT>(lc, s, col, ainv, r, x, omega, w);
      } else {
        if (lc->energy) vanka_colour<3, 1, VT, BT>(lc, s, col, ainv, r, x, omega, w);
        else vanka_colour<3, 0, VT, BT>(lc, s, col, ainv, r, x, omega, w);
      }
    }
  }
}

// z = M^-1 v: one V-cycle.  VT: matrix / inverse type, BT: vector type.
template <typename BT>
static void transfer(gm_ctx *c, hipStream_t s, int nrows, const int *rp, const int *ci,
                     const double *cv, int nnz, int ncols, const BT *v, const double *sc,
                     const signed char *mask, BT *y, int acc) {
  const double bytes = (double)nnz * 12.0 + 4.0 * (nrows + 1) + (double)ncols * sizeof(BT) +
                       (double)nrows * (1 + (acc ? 2 : 1) * sizeof(BT));
  const double avg = nrows ? (double)nnz / nrows : 0.0;  // entries per row
  auto launch = [&](auto kern, int G) {
    gm_launch(c, GM_K_TRANSFER, bytes, kern, dim3((int)(((long long)nrows * G + 255) / 256)), dim3(256),
              0, s, nrows, rp, ci, cv, v, sc, mask, y, acc);
  };
  if (avg <= 4.0) launch(k_csr_apply<BT, 4>, 4);
  else if (avg <= 12.0) launch(k_csr_apply<BT, 8>, 8);
  else launch(k_csr_apply<BT, 16>, 16);
}

template <typename VT, typename BT>
static void vcycle(gm_ctx *c, const VT *a0, const BT *v, BT *z, const gm_solver_opts *o) {
  hipStream_t s = c->stream;
  g_prof_ctx = c;
  const int nl = (int)c->mg.size();
  const int nu = o->mg_nu > 0 ? o->mg_nu : 1;
  const double om = o->mg_omega > 0 ? o->mg_omega : 1.0;
  const int cs = o->mg_coarse_sweeps;  // 0: exact solve with the dense inverse
  auto lv_ctx = [&](int l) { return l == 0 ? c : c->mg[l - 1].ctx; };
  auto lv = [&](int l) -> MgLevel & { return l == 0 ? c->mg0 : c->mg[l - 1]; };
  auto amat = [&](int l) -> const VT * { return l == 0 ? a0 : (const VT *)c->mg[l - 1].d_val; };
  // level 0 works on the caller's vectors
  BT *x0 = z;
  const BT *b0 = v;
  for (int l = 0; l <= nl; l++) {
    gm_ctx *lc = lv_ctx(l);
    MgLevel &L = lv(l);
    BT *x = l == 0 ? x0 : (BT *)L.d_x;
    const BT *b = l == 0 ? b0 : (const BT *)L.d_b;
    if (l == nl && cs <= 0 && L.dense_n == lc->N) {
      const int n = lc->N;
      gm_launch(c, GM_K_MG_DENSE, (double)n * n * 8.0 + 2.0 * n * sizeof(BT), k_dense_gemv<BT>,
                dim3((unsigned)(((long long)n * 64 + 255) / 256)), dim3(256), 0, s, n,
                (const double *)L.d_dense, b, x);
      break;
    }
    // pre-smoothing from x = 0 (the additive smoother's first sweep assigns x)
    if (o->mg_smoother != 1) k_set<BT><<<1024, 256, 0, s>>>(x, lc->N, (BT)0);
    smooth<VT, BT>(lc, s, amat(l), (const VT *)L.d_ainv, b, x, (BT *)L.d_r,
                   l == nl ? (cs > 0 ? cs : 20) : nu, om, o->mg_smoother == 1, true);
    if (l == nl) break;
    resid<VT, BT>(lc, s, amat(l), x, b, (BT *)L.d_r);
    MgLevel &C = c->mg[l];
    // restrict the (unscaled) residual: the level-0 matrix is row-sum scaled
    transfer<BT>(c, s, C.ctx->N, C.d_Rrow, C.d_Rcol, C.d_Rval, C.nnz, lc->N, (const BT *)L.d_r,
                 l == 0 ? c->mg_scale : nullptr, C.ctx->d_dbc, (BT *)C.d_b, 0);
  }
  for (int l = nl - 1; l >= 0; l--) {
    gm_ctx *lc = lv_ctx(l);
    MgLevel &L = lv(l);
    MgLevel &C = c->mg[l];
    BT *x = l == 0 ? x0 : (BT *)L.d_x;
    const BT *b = l == 0 ? b0 : (const BT *)L.d_b;
    transfer<BT>(c, s, lc->N, C.d_Prow, C.d_Pcol, C.d_Pval, C.nnz, C.ctx->N, (const BT *)C.d_x,
                 nullptr, lc->d_dbc, x, 1);
    smooth<VT, BT>(lc, s, amat(l), (const VT *)L.d_ainv, b, x, (BT *)L.d_r, nu, om,
                   o->mg_smoother == 1, false);
  }
}

int gm_mg_apply_d(gm_ctx *c, const double *a0, const double *v, double *z, const gm_solver_opts *o) {
  vcycle<double, double>(c, a0, v, z, o);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

int gm_mg_apply_f(gm_ctx *c, const float *a0, const double *v, double *z, const gm_solver_opts *o) {
  vcycle<float, double>(c, a0, v, z, o);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

void gm_mg_free(gm_ctx *c) {
  auto fr = [](MgLevel &L) {
    void *p[] = {L.d_Prow, L.d_Pcol, L.d_Pval, L.d_Rrow, L.d_Rcol, L.d_Rval, L.d_inject,
                 L.d_state, L.d_rdummy, L.d_x, L.d_b, L.d_r, L.d_val_p[0], L.d_val_p[1],
                 L.d_ainv_p[0], L.d_ainv_p[1],
                 L.d_dense, L.d_pcol, L.d_perm, L.d_child, L.d_Wp, L.d_Wi, L.d_Wv, L.d_Q};
    for (void *q : p)
      if (q) (void)hipFree(q);
  };
  for (MgLevel &L : c->mg) fr(L);
  fr(c->mg0);
  c->mg.clear();
  c->mg0 = MgLevel();
}

extern "C" int gm_mg_setup_dev(gm_ctx *c, const double *x, const double *val, const double *scale) {
  if (!c || !x || !val) return GM_ERR_ARG;
  int rc = gm_mg_setup(c, x, val, nullptr, GM_PREC_FP64, scale, 1, 0);
  if (rc) return rc;
  GM_HIP(hipStreamSynchronize(c->stream));
  return GM_OK;
}
