// Linear algebra on the MSR Jacobian: row-sum scaling, norms, SpMV,
// ILU(0) in patch order, restarted right-preconditioned GMRES.
//
// All reductions are two-stage (per-block partials, then a fixed-order sum on
// the host), so results are bitwise reproducible run to run.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <vector>

#include "gm_launch.h"

static constexpr int RED_BLOCKS = 1024;
static constexpr int RED_THREADS = 256;

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
__device__ inline T block_sum(T v, T *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += sh[i];
  return s;  // valid on thread 0
}

// ---------------------------------------------------------------- fills
template <typename T>
__global__ void k_fill_value(T *p, long long n, T v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

int gm_zero(gm_ctx *c, double *p, long long n) {
  k_fill_value<double><<<2048, 256, 0, c->stream>>>(p, n, 0.0);
  GM_HIP(hipGetLastError());
  return GM_OK;
}
int gm_zero32(gm_ctx *c, float *p, long long n) {
  k_fill_value<float><<<2048, 256, 0, c->stream>>>(p, n, 0.0f);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

// ---------------------------------------------------------------- row-sum scaling
// row_sum_scale_MSR (src/sl_matrix_util.c:506-600): s_i = sgn(a_ii) sum_j |a_ij|
// (sign only when |a_ii| > 1e-200); a_ij /= s_i; b_i /= s_i.  G lanes per row;
// the first G*KEEP entries of a row stay in registers between the sum and the
// scaling (one read of the matrix instead of two), longer rows re-read the
// rest.  Row sums accumulate in FP64 for either value type.
template <typename VT, int G, int KEEP>
__global__ __launch_bounds__(256) void k_row_sum_scale(int N, VT *__restrict__ a,
                                                       const int *__restrict__ bindx,
                                                       double *__restrict__ b,
                                                       double *__restrict__ scale) {
  const int gid = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  if (gid >= N) return;
  const int r = gid;
  const int k0 = bindx[r], k1 = bindx[r + 1];
  double s = 0.0, v[KEEP];
#pragma unroll
  for (int t = 0; t < KEEP; t++) {
    const int k = k0 + lane + G * t;
    v[t] = k < k1 ? (double)a[k] : 0.0;
    s += fabs(v[t]);
  }
  for (int k = k0 + lane + G * KEEP; k < k1; k += G) s += fabs((double)a[k]);
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  const double d = (double)a[r];
  double rs = fabs(d) + s;
  if (fabs(d) > 1.0e-200) rs = d > 0 ? rs : -rs;
#pragma unroll
  for (int t = 0; t < KEEP; t++) {
    const int k = k0 + lane + G * t;
    if (k < k1) a[k] = (VT)(v[t] / rs);
  }
  for (int k = k0 + lane + G * KEEP; k < k1; k += G) a[k] = (VT)((double)a[k] / rs);
  if (lane == 0) {
    scale[r] = rs;
    b[r] = b[r] / rs;
    a[r] = (VT)(d / rs);
  }
}

template <int G>
static void scale_g(gm_ctx *c, double *val, float *val32, double *b, double *scale) {
  const int blocks = (int)(((long long)c->N * G + 255) / 256);
  const double nnz = (double)c->nval - 1.0;
  // values read and written, the row pointers, b read and written, the scale
  const double bytes = nnz * 2.0 * (val32 ? 4.0 : 8.0) + 4.0 * (c->N + 1) + c->N * 24.0;
  constexpr int KEEP = G == 64 ? 9 : 4;  // HEX27 rows up to 576 entries, QUAD9 up to 64
  if (val32)
    gm_launch(c, GM_K_ROW_SCALE, bytes, k_row_sum_scale<float, G, KEEP>, dim3(blocks), dim3(256),
              0, c->stream, c->N, val32, c->d_bindx, b, scale);
  else
    gm_launch(c, GM_K_ROW_SCALE, bytes, k_row_sum_scale<double, G, KEEP>, dim3(blocks), dim3(256),
              0, c->stream, c->N, val, c->d_bindx, b, scale);
}

int gm_scale_launch(gm_ctx *c, double *val, float *val32, double *b, double *scale) {
  // a wave per HEX27 row (~250 entries, at most ~530), 16 lanes per QUAD9 row (~60)
  if (c->dim == 3) scale_g<64>(c, val, val32, b, scale);
  else scale_g<16>(c, val, val32, b, scale);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

// ---------------------------------------------------------------- norms
// L2_norm / Loo_norm (src/mm_sol_nonlinear.c:3177, 3320): partial sums of
// squares and partial maxima per block.
__global__ __launch_bounds__(RED_THREADS) void k_norms(const double *__restrict__ v, long long n,
                                                       double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  double s = 0.0, m = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double x = v[i];
    s += x * x;
    m = fmax(m, fabs(x));
  }
  double ss = block_sum(s, sh);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  __shared__ double shm[RED_THREADS / 64];
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int i = 0; i < RED_THREADS / 64; i++) mm = fmax(mm, shm[i]);
    part[blockIdx.x] = ss;
    part[gridDim.x + blockIdx.x] = mm;
  }
}

// L2_norm_r / Loo_norm_r (src/mm_sol_nonlinear.c:3249, 3376)
__global__ __launch_bounds__(RED_THREADS) void k_corr_norms(const double *__restrict__ d,
                                                            const double *__restrict__ x,
                                                            long long n, double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  double s = 0.0, m = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double di = d[i], xi = x[i];
    s += di * di / (1 + xi * xi);
    m = fmax(m, fabs(di) / (1.0 + fabs(xi)));
  }
  double ss = block_sum(s, sh);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  __shared__ double shm[RED_THREADS / 64];
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int i = 0; i < RED_THREADS / 64; i++) mm = fmax(mm, shm[i]);
    part[blockIdx.x] = ss;
    part[gridDim.x + blockIdx.x] = mm;
  }
}

static int finish_norms(gm_ctx *c, double *l2, double *loo) {
  GM_HIP(hipMemcpyAsync(c->h_red, c->d_red, 2 * RED_BLOCKS * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  GM_HIP(hipStreamSynchronize(c->stream));
  double s = 0.0, m = 0.0;
  for (int i = 0; i < RED_BLOCKS; i++) s += c->h_red[i];
  for (int i = 0; i < RED_BLOCKS; i++) m = std::max(m, c->h_red[RED_BLOCKS + i]);
  // NaN propagates through the sum; fmax drops it, so test the sum too
  *l2 = std::sqrt(s);
  *loo = std::isfinite(s) ? m : s;
  return GM_OK;
}

int gm_norms(gm_ctx *c, const double *v, long long n, double *l2, double *loo) {
  k_norms<<<RED_BLOCKS, RED_THREADS, 0, c->stream>>>(v, n, c->d_red);
  GM_HIP(hipGetLastError());
  return finish_norms(c, l2, loo);
}

int gm_corr_norms(gm_ctx *c, const double *dx, const double *x, double *l2r, double *loor) {
  k_corr_norms<<<RED_BLOCKS, RED_THREADS, 0, c->stream>>>(dx, x, c->N, c->d_red);
  GM_HIP(hipGetLastError());
  return finish_norms(c, l2r, loor);
}

// global_velocity_norm (src/mm_fill_aux.c:612-672): partial sums of x^2 over
// the velocity unknowns, the first dim of every node
__global__ __launch_bounds__(RED_THREADS) void k_velocity_sq(int num_nodes, int dim,
                                                             const int *__restrict__ fu,
                                                             const double *__restrict__ x,
                                                             double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  double s = 0.0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < (long long)num_nodes * dim;
       t += (long long)gridDim.x * blockDim.x) {
    const int n = (int)(t / dim), a = (int)(t - (long long)n * dim);
    const double v = x[fu[n] + a];
    s += v * v;
  }
  const double ss = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = ss;
}

int gm_velocity_norm(gm_ctx *c, const double *x, double *U) {
  k_velocity_sq<<<RED_BLOCKS, RED_THREADS, 0, c->stream>>>(c->num_nodes, c->dim, c->d_fu, x, c->d_red);
  GM_HIP(hipGetLastError());
  GM_HIP(hipMemcpyAsync(c->h_red, c->d_red, RED_BLOCKS * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  GM_HIP(hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (int i = 0; i < RED_BLOCKS; i++) s += c->h_red[i];
  *U = s / ((double)c->num_nodes * c->dim);
  return GM_OK;
}

// x -= damp * dx (NLS_FULL_STEP update, src/mm_sol_nonlinear.c:2553-2557)
__global__ void k_newton_update(double *__restrict__ x, const double *__restrict__ dx, int n,
                                double damp) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    x[i] -= damp * dx[i];
}

int gm_axpy_update(gm_ctx *c, double *x, const double *dx, double damp) {
  k_newton_update<<<2048, 256, 0, c->stream>>>(x, dx, c->N, damp);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

// ---------------------------------------------------------------- MSR SpMV
// y = A x (MV_MSR, src/sl_auxutil.c:373).  G lanes per row.
template <typename VT, typename XT, typename YT, int G>
__global__ __launch_bounds__(256) void k_spmv_msr(int N, const VT *__restrict__ a,
                                                  const int *__restrict__ bindx,
                                                  const XT *__restrict__ x, YT *__restrict__ y) {
  const int r = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x & (G - 1);
  if (r >= N) return;
  const int k0 = bindx[r], k1 = bindx[r + 1];
  // FP64 accumulation whenever the vector is FP64: an FP32 matrix times an
  // FP64 Krylov vector must not round the vector, or the Arnoldi relation
  // only holds to ~1e-7 |A||x| and GMRES stalls on ill-conditioned systems.
  using AT = typename std::conditional<sizeof(VT) == 8 || sizeof(XT) == 8, double, float>::type;
  AT s = row_dot<AT, VT, XT, G>(a, bindx, x, k0, k1, lane);
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (lane == 0) y[r] = (YT)(s + (AT)a[r] * (AT)x[r]);
}

template <typename VT, typename XT, typename YT>
static void spmv(gm_ctx *c, const VT *a, const XT *x, YT *y) {
  using AT = typename std::conditional<sizeof(VT) == 8 || sizeof(XT) == 8, double, float>::type;
  const double vec = (double)c->N * (sizeof(XT) + sizeof(YT));
  if (c->dim == 3 && c->nn_cap <= GM_NODE_CAP_MAX) {
    // HEX27: one wave per node (k_msr_nodes), values, node graph and row
    // pointers (QUAD9 rows, ~60 entries, are faster per row: 0.156 vs 0.345 ms)
    const double bytes = (double)(c->nval - 1) * sizeof(VT) +
                         4.0 * (2.0 * (c->num_nodes + 1) + c->N + c->nn_list.size()) + vec;
    gm_launch(c, GM_K_SPMV, bytes, k_msr_nodes<AT, VT, XT, YT, false>, dim3((c->num_nodes + 3) / 4),
              dim3(256), (size_t)4 * 2 * c->nn_cap * sizeof(int), c->stream, c->num_nodes, c->dim,
              c->nv, c->nn_cap, c->d_fu, c->d_nn_ptr, c->d_nn_list, c->d_bindx, a, x,
              (const XT *)nullptr, y);
    return;
  }
  // G lanes per row
  const double bytes = (double)(c->nval - 1) * (sizeof(VT) + 4) + 4.0 * (c->N + 1) + vec;
  if (c->dim == 2) {
    constexpr int G = 8;  // 2D rows hold ~50-250 entries: measured 156 vs 172 us at 256^2 (G = 16)
    const int blocks = (int)(((long long)c->N * G + 255) / 256);
    gm_launch(c, GM_K_SPMV, bytes, k_spmv_msr<VT, XT, YT, G>, dim3(blocks), dim3(256), 0, c->stream,
              c->N, a, c->d_bindx, x, y);
  } else {
    constexpr int G = 64;
    const int blocks = (int)(((long long)c->N * G + 255) / 256);
    gm_launch(c, GM_K_SPMV, bytes, k_spmv_msr<VT, XT, YT, G>, dim3(blocks), dim3(256), 0, c->stream,
              c->N, a, c->d_bindx, x, y);
  }
}

int gm_spmv_launch(gm_ctx *c, const double *val, const double *x, double *y) {
  spmv<double, double, double>(c, val, x, y);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

int gm_spmv32_launch(gm_ctx *c, const float *val, const double *x, double *y) {
  spmv<float, double, double>(c, val, x, y);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

__global__ void k_lincomb(int n, double *__restrict__ out, const double *__restrict__ a, double s,
                          const double *__restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = a[i] - s * b[i];
}

__global__ void k_div_scale(int n, double *__restrict__ v, const double *__restrict__ sc) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v[i] = v[i] / sc[i];
}

__global__ __launch_bounds__(RED_THREADS) void k_dot(int n, const double *__restrict__ a,
                                                     const double *__restrict__ b,
                                                     double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    s += a[i] * b[i];
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

int gm_lincomb(gm_ctx *c, double *out, const double *a, double s, const double *b) {
  k_lincomb<<<2048, 256, 0, c->stream>>>(c->N, out, a, s, b);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

// out = w0 v0 + w1 v1 + w2 v2 (null vectors and their weights skipped)
__global__ void k_lin3(int n, double *__restrict__ out, double w0, const double *__restrict__ v0,
                       double w1, const double *__restrict__ v1, double w2,
                       const double *__restrict__ v2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = w0 * v0[i];
    if (v1) s += w1 * v1[i];
    if (v2) s += w2 * v2[i];
    out[i] = s;
  }
}

int gm_lin3(gm_ctx *c, double *out, double w0, const double *v0, double w1, const double *v1,
            double w2, const double *v2) {
  k_lin3<<<2048, 256, 0, c->stream>>>(c->N, out, w0, v0, w1, v1, w2, v2);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

int gm_div_scale(gm_ctx *c, double *v, const double *scale) {
  k_div_scale<<<2048, 256, 0, c->stream>>>(c->N, v, scale);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

int gm_dot(gm_ctx *c, const double *a, const double *b, double *out) {
  k_dot<<<RED_BLOCKS, RED_THREADS, 0, c->stream>>>(c->N, a, b, c->d_red);
  GM_HIP(hipGetLastError());
  GM_HIP(hipMemcpyAsync(c->h_red, c->d_red, RED_BLOCKS * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  GM_HIP(hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (int i = 0; i < RED_BLOCKS; i++) s += c->h_red[i];
  *out = s;
  return GM_OK;
}

// ---------------------------------------------------------------- ILU(0)
// ILU(0) of P A P^T where P is the patch ordering.  Patches of one colour
// share no couplings, so each colour is one launch with one workgroup per
// patch.  Inside a patch the rows keep their sequential (IKJ) order of
// dependence; the host groups the patch's nodes into levels (a node depends
// on every earlier node of the patch its rows reference in L, or later node in
// U), and the workgroup walks the levels: all rows of a level first take the
// contributions from outside their own node in parallel, then one thread per
// node resolves the node's small diagonal block.  The arithmetic per entry is
// the sequential algorithm's; only the summation order inside a row differs.

static constexpr int ILU_MAXPATCH = 8192;  // rows of one patch held in LDS by the sweeps
static constexpr int ILU_NT = 256;         // threads per patch workgroup (sweeps)
static constexpr int ILU_FT = 512;         // threads per patch workgroup (factorisation)

struct IluLv {
  const int *rowptr, *col, *diag, *node_first, *node_last, *patch_row;
  const int *lv_ptr, *lv_nstart, *lv_node, *lv_rstart, *lv_row;
};

template <int G, typename T>
__device__ inline T group_sum(T v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

// forward sweep  y = L^-1 (P b),  L unit lower; y in permuted numbering
template <typename VT, typename BT, typename WT, int G>
__global__ __launch_bounds__(ILU_NT) void k_ilu_lower(IluLv S, const VT *__restrict__ lu,
                                                      int patch0, const int *__restrict__ perm,
                                                      const BT *__restrict__ b, WT *y) {
  __shared__ WT s_y[ILU_MAXPATCH];
  constexpr int NG = ILU_NT / G;
  const int P = patch0 + blockIdx.x;
  const int r0 = S.patch_row[P], r1 = S.patch_row[P + 1];
  const int tid = threadIdx.x, lane = tid % G, grp = tid / G;
  for (int l = S.lv_ptr[P]; l < S.lv_ptr[P + 1]; l++) {
    const int ra = S.lv_rstart[l], rb = S.lv_rstart[l + 1];
    for (int k = ra + grp; k < rb; k += NG) {
      const int r = S.lv_row[k];
      const int nf = S.node_first[r];
      WT s = 0;
      const int e = S.diag[r];
      for (int t = S.rowptr[r] + lane; t < e; t += G) {
        const int cc = S.col[t];
        if (cc < nf) s += (WT)lu[t] * (cc >= r0 ? s_y[cc - r0] : y[cc]);
      }
      s = group_sum<G>(s);
      if (lane == 0) s_y[r - r0] = (WT)b[perm[r]] - s;
    }
    __syncthreads();
    for (int k = S.lv_nstart[l] + tid; k < S.lv_nstart[l + 1]; k += ILU_NT) {
      const int nf = S.lv_node[k];
      const int nl = S.node_last[nf];
      for (int r = nf + 1; r <= nl; r++) {
        WT s = s_y[r - r0];
        for (int t = S.diag[r] - 1; t >= S.rowptr[r]; t--) {
          const int cc = S.col[t];
          if (cc < nf) break;
          s -= (WT)lu[t] * s_y[cc - r0];
        }
        s_y[r - r0] = s;
      }
    }
    __syncthreads();
  }
  for (int r = r0 + tid; r < r1; r += ILU_NT) y[r] = s_y[r - r0];
}

// backward sweep x = U^-1 y in place on y (permuted), plus the unpermuted copy
template <typename VT, typename WT, typename OT, int G>
__global__ __launch_bounds__(ILU_NT) void k_ilu_upper(IluLv S, const VT *__restrict__ lu,
                                                      int patch0, const int *__restrict__ perm,
                                                      WT *y, OT *__restrict__ out) {
  __shared__ WT s_x[ILU_MAXPATCH];
  constexpr int NG = ILU_NT / G;
  const int P = patch0 + blockIdx.x;
  const int r0 = S.patch_row[P], r1 = S.patch_row[P + 1];
  const int tid = threadIdx.x, lane = tid % G, grp = tid / G;
  for (int l = S.lv_ptr[P]; l < S.lv_ptr[P + 1]; l++) {
    const int ra = S.lv_rstart[l], rb = S.lv_rstart[l + 1];
    for (int k = ra + grp; k < rb; k += NG) {
      const int r = S.lv_row[k];
      const int nl = S.node_last[S.node_first[r]];
      WT s = 0;
      const int e = S.rowptr[r + 1];
      for (int t = S.diag[r] + 1 + lane; t < e; t += G) {
        const int cc = S.col[t];
        if (cc > nl) s += (WT)lu[t] * (cc < r1 ? s_x[cc - r0] : y[cc]);
      }
      s = group_sum<G>(s);
      if (lane == 0) s_x[r - r0] = y[r] - s;
    }
    __syncthreads();
    for (int k = S.lv_nstart[l] + tid; k < S.lv_nstart[l + 1]; k += ILU_NT) {
      const int nf = S.lv_node[k];
      const int nl = S.node_last[nf];
      for (int r = nl; r >= nf; r--) {
        WT s = s_x[r - r0];
        const int d = S.diag[r];
        for (int t = d + 1; t < S.rowptr[r + 1]; t++) {
          const int cc = S.col[t];
          if (cc > nl) break;
          s -= (WT)lu[t] * s_x[cc - r0];
        }
        s_x[r - r0] = s / (WT)lu[d];
      }
    }
    __syncthreads();
  }
  for (int r = r0 + tid; r < r1; r += ILU_NT) {
    const WT v = s_x[r - r0];
    y[r] = v;
    out[perm[r]] = (OT)v;
  }
}

// IKJ elimination of row r by the pivot rows whose column lies in [kb, ke),
// one wavefront per row, the row staged in LDS.
template <typename VT>
__device__ inline void ilu_row_eliminate(const IluLv &S, VT *lu, int r, int kb, int ke, int *s_col,
                                         VT *s_val, int *flag) {
  const int lane = threadIdx.x & 63;
  const int b = S.rowptr[r], len = S.rowptr[r + 1] - b, dpos = S.diag[r] - b;
  for (int t = lane; t < len; t += 64) {
    s_col[t] = S.col[b + t];
    s_val[t] = lu[b + t];
  }
  __builtin_amdgcn_wave_barrier();
  __threadfence_block();
  for (int kk = 0; kk < dpos; kk++) {
    const int k = s_col[kk];
    if (k < kb) continue;
    if (k >= ke) break;
    const int dk = S.diag[k];
    const VT l = s_val[kk] / lu[dk];
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) s_val[kk] = l;
    const int kend = S.rowptr[k + 1];
    for (int jj = dk + 1 + lane; jj < kend; jj += 64) {
      const int j = S.col[jj];
      int lo = kk + 1, hi = len - 1, pos = -1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = s_col[mid];
        if (v == j) {
          pos = mid;
          break;
        }
        if (v < j) lo = mid + 1;
        else hi = mid - 1;
      }
      if (pos >= 0) s_val[pos] -= l * lu[jj];
    }
    __builtin_amdgcn_wave_barrier();
    __threadfence_block();
  }
  for (int t = lane; t < len; t += 64) lu[b + t] = s_val[t];
  if (lane == 0 && ke > r && s_val[dpos] == (VT)0) atomicOr(flag + 2, 1);
  __threadfence_block();
  __builtin_amdgcn_wave_barrier();
}

template <typename VT>
__global__ __launch_bounds__(ILU_FT) void k_ilu_factor(IluLv S, VT *lu, int patch0, int *flag) {
  constexpr int NW = ILU_FT / 64;
  __shared__ int s_col[NW][GM_MAXROW];
  __shared__ VT s_val[NW][GM_MAXROW];
  const int P = patch0 + blockIdx.x;
  const int w = threadIdx.x >> 6;
  for (int l = S.lv_ptr[P]; l < S.lv_ptr[P + 1]; l++) {
    // rows of the level against pivot rows outside their own node
    for (int k = S.lv_rstart[l] + w; k < S.lv_rstart[l + 1]; k += NW) {
      const int r = S.lv_row[k];
      ilu_row_eliminate<VT>(S, lu, r, 0, S.node_first[r], s_col[w], s_val[w], flag);
    }
    __syncthreads();
    // each node's rows against the node's earlier rows, in order
    for (int k = S.lv_nstart[l] + w; k < S.lv_nstart[l + 1]; k += NW) {
      const int nf = S.lv_node[k];
      const int nl = S.node_last[nf];
      for (int r = nf; r <= nl; r++) ilu_row_eliminate<VT>(S, lu, r, nf, r + 1, s_col[w], s_val[w], flag);
    }
    __syncthreads();
  }
}

// gather MSR values into the permuted CSR
template <typename VT, typename DT>
__global__ void k_gather(long long n, const int *__restrict__ src, const VT *__restrict__ a,
                         DT *__restrict__ lu) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    lu[i] = (DT)a[src[i]];
}

static IluLv lv_args(const IluPlan &P, bool upper) {
  IluLv S;
  S.rowptr = P.d_rowptr;
  S.col = P.d_col;
  S.diag = P.d_diag;
  S.node_first = P.d_node_first;
  S.node_last = P.d_node_last;
  S.patch_row = P.d_patch_row;
  const int u = upper ? 1 : 0;
  S.lv_ptr = P.d_lv_ptr[u];
  S.lv_nstart = P.d_lv_nstart[u];
  S.lv_node = P.d_lv_node[u];
  S.lv_rstart = P.d_lv_rstart[u];
  S.lv_row = P.d_lv_row[u];
  return S;
}

static void free_plan(IluPlan &P) {
  void *ptrs[] = {P.d_perm, P.d_rowptr, P.d_col, P.d_diag, P.d_src, P.d_patch_row, P.d_lu,
                  P.d_lu32, P.d_y, P.d_y32, P.d_node_first, P.d_node_last};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (int u = 0; u < 2; u++) {
    void *q[] = {P.d_lv_ptr[u], P.d_lv_nstart[u], P.d_lv_node[u], P.d_lv_rstart[u], P.d_lv_row[u]};
    for (void *p : q)
      if (p) (void)hipFree(p);
  }
  P = IluPlan();
}

template <typename T>
static int up(T **d, const std::vector<T> &h) {
  GM_HIP(hipMalloc(d, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) GM_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return GM_OK;
}

int gm_ilu_plan(gm_ctx *c, int bins) {
  IluPlan &P = c->ilu;
  if (P.bins == bins) return GM_OK;
  if (P.d_perm) free_plan(P);
  std::vector<int> cptr;
  int ncol = gm_patch_order(c, bins, P.perm, cptr, P.patch_row);
  if (ncol < 0) return ncol;
  P.ncol = ncol;
  P.colour_ptr = cptr;
  P.npatch = (int)P.patch_row.size() - 1;
  for (int p = 0; p < P.npatch; p++)
    if (P.patch_row[p + 1] - P.patch_row[p] > ILU_MAXPATCH) {
      gm_set_error("ILU patch with %d rows exceeds %d; use more patch bins",
                   P.patch_row[p + 1] - P.patch_row[p], ILU_MAXPATCH);
      return GM_ERR_ARG;
    }
  const int N = c->N;
  std::vector<int> iperm(N);
  for (int i = 0; i < N; i++) iperm[P.perm[i]] = i;
  std::vector<int> rowptr(N + 1, 0);
  for (int r = 0; r < N; r++) {
    const int i = P.perm[r];
    rowptr[r + 1] = 1 + c->bindx[i + 1] - c->bindx[i];
  }
  for (int r = 0; r < N; r++) rowptr[r + 1] += rowptr[r];
  const long long nnz = rowptr[N];
  P.nnz = nnz;
  std::vector<int> col(nnz), src(nnz), diag(N);
#pragma omp parallel
  {
    std::vector<std::pair<int, int>> tmp;
#pragma omp for schedule(dynamic, 4096)
    for (int r = 0; r < N; r++) {
      const int i = P.perm[r];
      tmp.clear();
      tmp.emplace_back(r, i);
      for (int k = c->bindx[i]; k < c->bindx[i + 1]; k++) tmp.emplace_back(iperm[c->bindx[k]], k);
      std::sort(tmp.begin(), tmp.end());
      const int b = rowptr[r];
      for (size_t t = 0; t < tmp.size(); t++) {
        col[b + t] = tmp[t].first;
        src[b + t] = tmp[t].second;
        if (tmp[t].first == r) diag[r] = b + (int)t;
      }
    }
  }
  for (int r = 0; r < N; r++)
    if (rowptr[r + 1] - rowptr[r] > GM_MAXROW) {
      gm_set_error("Jacobian row longer than %d", GM_MAXROW);
      return GM_ERR_ARG;
    }
  // nodes in permuted numbering: the unknowns of a mesh node stay contiguous
  std::vector<int> node_first(N), node_last(N);
  {
    std::vector<int> node_of(N);
    for (int n = 0; n < c->num_nodes; n++)
      for (int u = c->first_unknown[n]; u < c->first_unknown[n + 1]; u++) node_of[u] = n;
    int r = 0;
    while (r < N) {
      const int n = node_of[P.perm[r]];
      const int cnt = c->first_unknown[n + 1] - c->first_unknown[n];
      for (int k = 0; k < cnt; k++) {
        node_first[r + k] = r;
        node_last[r + k] = r + cnt - 1;
      }
      r += cnt;
    }
  }
  // level schedules per patch (0: forward / factorisation, 1: backward)
  std::vector<int> lv_ptr[2], lv_nstart[2], lv_node[2], lv_rstart[2], lv_row[2];
  {
    std::vector<int> level(N, 0);
    for (int u = 0; u < 2; u++) {
      lv_ptr[u].assign(1, 0);
      lv_nstart[u].assign(1, 0);
      lv_rstart[u].assign(1, 0);
      for (int p = 0; p < P.npatch; p++) {
        const int r0 = P.patch_row[p], r1 = P.patch_row[p + 1];
        std::vector<int> nodes;
        for (int r = r0; r < r1; r = node_last[r] + 1) nodes.push_back(r);
        int maxlev = -1;
        if (u == 0) {
          for (int nf : nodes) {
            int lv = 0;
            for (int r = nf; r <= node_last[nf]; r++)
              for (int t = rowptr[r]; t < diag[r]; t++) {
                const int cc = col[t];
                if (cc >= r0 && cc < nf) lv = std::max(lv, level[node_first[cc]] + 1);
              }
            level[nf] = lv;
            maxlev = std::max(maxlev, lv);
          }
        } else {
          for (auto it = nodes.rbegin(); it != nodes.rend(); ++it) {
            const int nf = *it, nl = node_last[nf];
            int lv = 0;
            for (int r = nf; r <= nl; r++)
              for (int t = diag[r] + 1; t < rowptr[r + 1]; t++) {
                const int cc = col[t];
                if (cc > nl && cc < r1) lv = std::max(lv, level[node_first[cc]] + 1);
              }
            level[nf] = lv;
            maxlev = std::max(maxlev, lv);
          }
        }
        std::vector<std::vector<int>> byl(maxlev + 1);
        for (int nf : nodes) byl[level[nf]].push_back(nf);
        for (int lv = 0; lv <= maxlev; lv++) {
          for (int nf : byl[lv]) {
            lv_node[u].push_back(nf);
            for (int r = nf; r <= node_last[nf]; r++) lv_row[u].push_back(r);
          }
          lv_nstart[u].push_back((int)lv_node[u].size());
          lv_rstart[u].push_back((int)lv_row[u].size());
        }
        lv_ptr[u].push_back((int)lv_nstart[u].size() - 1);
      }
    }
  }
  P.levels = (int)lv_ptr[0].back();
  // per colour: rows and strictly-lower / upper (incl. diagonal) entries, for
  // the algorithmic-byte counts of the sweeps
  P.col_rows.assign(ncol, 0);
  P.col_nnzL.assign(ncol, 0);
  P.col_nnzU.assign(ncol, 0);
  for (int col = 0; col < ncol; col++) {
    const int ra = P.patch_row[P.colour_ptr[col]], rb = P.patch_row[P.colour_ptr[col + 1]];
    P.col_rows[col] = rb - ra;
    for (int r = ra; r < rb; r++) {
      P.col_nnzL[col] += diag[r] - rowptr[r];
      P.col_nnzU[col] += rowptr[r + 1] - diag[r];
    }
  }
  int rc;
  if ((rc = up(&P.d_perm, P.perm)) || (rc = up(&P.d_rowptr, rowptr)) || (rc = up(&P.d_col, col)) ||
      (rc = up(&P.d_diag, diag)) || (rc = up(&P.d_src, src)) || (rc = up(&P.d_patch_row, P.patch_row)) ||
      (rc = up(&P.d_node_first, node_first)) || (rc = up(&P.d_node_last, node_last)))
    return rc;
  for (int u = 0; u < 2; u++)
    if ((rc = up(&P.d_lv_ptr[u], lv_ptr[u])) || (rc = up(&P.d_lv_nstart[u], lv_nstart[u])) ||
        (rc = up(&P.d_lv_node[u], lv_node[u])) || (rc = up(&P.d_lv_rstart[u], lv_rstart[u])) ||
        (rc = up(&P.d_lv_row[u], lv_row[u])))
      return rc;
  GM_HIP(hipMalloc(&P.d_lu, nnz * sizeof(double)));
  GM_HIP(hipMalloc(&P.d_y, N * sizeof(double)));
  P.bins = bins;
  return GM_OK;
}

// factorise the (scaled) Jacobian; prec 0: FP64 values, 1: FP32 values
int gm_ilu_factor(gm_ctx *c, const double *val, const float *val32, int prec) {
  IluPlan &P = c->ilu;
  if (prec == GM_PREC_MIXED && !P.d_lu32) {
    GM_HIP(hipMalloc(&P.d_lu32, P.nnz * sizeof(float)));
    GM_HIP(hipMalloc(&P.d_y32, c->N * sizeof(float)));
  }
  GM_HIP(hipMemsetAsync(c->d_flag + 2, 0, sizeof(int), c->stream));
  if (prec == GM_PREC_MIXED) {
    if (val32) k_gather<float, float><<<4096, 256, 0, c->stream>>>(P.nnz, P.d_src, val32, P.d_lu32);
    else k_gather<double, float><<<4096, 256, 0, c->stream>>>(P.nnz, P.d_src, val, P.d_lu32);
  } else {
    k_gather<double, double><<<4096, 256, 0, c->stream>>>(P.nnz, P.d_src, val, P.d_lu);
  }
  const IluLv S = lv_args(P, false);
  for (int col = 0; col < P.ncol; col++) {
    const int p0 = P.colour_ptr[col], cnt = P.colour_ptr[col + 1] - p0;
    if (!cnt) continue;
    if (prec == GM_PREC_MIXED)
      gm_launch(c, GM_K_ILU_FACT, 0.0, k_ilu_factor<float>, dim3(cnt), dim3(ILU_FT), 0, c->stream,
                S, P.d_lu32, p0, c->d_flag);
    else
      gm_launch(c, GM_K_ILU_FACT, 0.0, k_ilu_factor<double>, dim3(cnt), dim3(ILU_FT), 0, c->stream,
                S, P.d_lu, p0, c->d_flag);
  }
  GM_HIP(hipGetLastError());
  int flag = 0;
  GM_HIP(hipMemcpyAsync(&flag, c->d_flag + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  GM_HIP(hipStreamSynchronize(c->stream));
  if (flag) {
    gm_set_error("zero pivot in ILU(0)");
    return GM_ERR_PIVOT;
  }
  return GM_OK;
}

// z = M^-1 r  (r, z unpermuted)
template <typename VT, typename RT, typename ZT>
static void ilu_apply(gm_ctx *c, const VT *lu, VT *ywork, const RT *r, ZT *z) {
  IluPlan &P = c->ilu;
  const IluLv SL = lv_args(P, false), SU = lv_args(P, true);
  for (int col = 0; col < P.ncol; col++) {
    const int p0 = P.colour_ptr[col], cnt = P.colour_ptr[col + 1] - p0;
    if (!cnt) continue;
    // L values + column ids, row pointers, diagonals, node ranges, level lists,
    // permutation, input (gathered) and output vectors
    const double bytes = (double)P.col_nnzL[col] * (sizeof(VT) + 4) +
                         (double)P.col_rows[col] * (6 * 4 + sizeof(RT) + sizeof(VT));
    if (c->dim == 2)
      gm_launch(c, GM_K_ILU_LOWER, bytes, k_ilu_lower<VT, RT, VT, 16>, dim3(cnt), dim3(ILU_NT), 0,
                c->stream, SL, lu, p0, P.d_perm, r, ywork);
    else
      gm_launch(c, GM_K_ILU_LOWER, bytes, k_ilu_lower<VT, RT, VT, 64>, dim3(cnt), dim3(ILU_NT), 0,
                c->stream, SL, lu, p0, P.d_perm, r, ywork);
  }
  for (int col = P.ncol - 1; col >= 0; col--) {
    const int p0 = P.colour_ptr[col], cnt = P.colour_ptr[col + 1] - p0;
    if (!cnt) continue;
    const double bytes = (double)P.col_nnzU[col] * (sizeof(VT) + 4) +
                         (double)P.col_rows[col] * (6 * 4 + 2 * sizeof(VT) + sizeof(ZT));
    if (c->dim == 2)
      gm_launch(c, GM_K_ILU_UPPER, bytes, k_ilu_upper<VT, VT, ZT, 16>, dim3(cnt), dim3(ILU_NT), 0,
                c->stream, SU, lu, p0, P.d_perm, ywork, z);
    else
      gm_launch(c, GM_K_ILU_UPPER, bytes, k_ilu_upper<VT, VT, ZT, 64>, dim3(cnt), dim3(ILU_NT), 0,
                c->stream, SU, lu, p0, P.d_perm, ywork, z);
  }
}

// ---------------------------------------------------------------- Krylov kernels
// partial dots: part[k*nb + blk] = sum_i V[k][i] w[i]  for k < nk; the last row
// (k = nk) holds w.w.
template <typename BT>
__global__ __launch_bounds__(RED_THREADS) void k_mdot(int N, int nk, const BT *__restrict__ V,
                                                      const BT *__restrict__ w,
                                                      double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  const int nb = gridDim.x;
  for (int k = 0; k <= nk; k++) {
    const BT *vk = k < nk ? V + (size_t)k * N : w;
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
      s += (double)vk[i] * (double)w[i];
    double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[(size_t)k * nb + blockIdx.x] = t;
  }
}

// w -= sum_k h[k] V[k]; part[blk] = partial of w.w after the update
template <typename BT>
__global__ __launch_bounds__(RED_THREADS) void k_orth_update(int N, int nk,
                                                             const BT *__restrict__ V,
                                                             const double *__restrict__ h,
                                                             BT *__restrict__ w,
                                                             double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  __shared__ double s_h[128];
  for (int k = threadIdx.x; k < nk; k += blockDim.x) s_h[k] = h[k];
  __syncthreads();
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double acc = (double)w[i];
    for (int k = 0; k < nk; k++) acc -= s_h[k] * (double)V[(size_t)k * N + i];
    const BT wn = (BT)acc;
    w[i] = wn;
    s += (double)wn * (double)wn;
  }
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <typename BT>
__global__ void k_scale_into(int N, const BT *__restrict__ w, double inv, BT *__restrict__ v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
    v[i] = (BT)((double)w[i] * inv);
}

// u = sum_k y[k] V[k]
template <typename BT>
__global__ void k_combine(int N, int nk, const BT *__restrict__ V, const double *__restrict__ y,
                          BT *__restrict__ u) {
  __shared__ double s_y[128];
  for (int k = threadIdx.x; k < nk; k += blockDim.x) s_y[k] = y[k];
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = 0; k < nk; k++) acc += s_y[k] * (double)V[(size_t)k * N + i];
    u[i] = (BT)acc;
  }
}

// x += z (x FP64, z any), r = b - r  (r holds A x on entry), both with partial r.r
template <typename ZT>
__global__ void k_add(int N, double *__restrict__ x, const ZT *__restrict__ z) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
    x[i] += (double)z[i];
}

template <typename BT>
__global__ __launch_bounds__(RED_THREADS) void k_resid(int N, const double *__restrict__ b,
                                                       const double *__restrict__ ax,
                                                       BT *__restrict__ r,
                                                       double *__restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const double v = b[i] - ax[i];
    r[i] = (BT)v;
    s += v * v;
  }
  double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

static int sum_parts(gm_ctx *c, const double *d_part, int rows, int nb, double *out) {
  GM_HIP(hipMemcpyAsync(c->h_red, d_part, (size_t)rows * nb * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  GM_HIP(hipStreamSynchronize(c->stream));
  for (int k = 0; k < rows; k++) {
    double s = 0.0;
    for (int i = 0; i < nb; i++) s += c->h_red[(size_t)k * nb + i];
    out[k] = s;
  }
  return GM_OK;
}

static constexpr int KB = 256;  // blocks of the Krylov reductions

static int ensure_krylov(gm_ctx *c, int m, int prec) {
  KrylovWork &K = c->kry;
  if (K.m >= m && K.prec == prec) return GM_OK;
  if (K.d_V) hipFree(K.d_V);
  if (K.d_w) hipFree(K.d_w);
  if (K.d_z) hipFree(K.d_z);
  if (K.d_part) hipFree(K.d_part);
  if (K.d_h) hipFree(K.d_h);
  if (K.d_V32) hipFree(K.d_V32);
  if (K.d_w32) hipFree(K.d_w32);
  if (K.d_z32) hipFree(K.d_z32);
  K = KrylovWork();
  const size_t N = c->N;
  if (prec == GM_PREC_MIXED) {
    GM_HIP(hipMalloc(&K.d_V32, (size_t)(m + 1) * N * sizeof(float)));
    GM_HIP(hipMalloc(&K.d_w32, N * sizeof(float)));
    GM_HIP(hipMalloc(&K.d_z32, 2 * N * sizeof(float)));
  } else {
    GM_HIP(hipMalloc(&K.d_V, (size_t)(m + 1) * N * sizeof(double)));
  }
  GM_HIP(hipMalloc(&K.d_w, N * sizeof(double)));
  GM_HIP(hipMalloc(&K.d_z, 2 * N * sizeof(double)));
  GM_HIP(hipMalloc(&K.d_part, (size_t)(m + 2) * KB * sizeof(double)));
  GM_HIP(hipMalloc(&K.d_h, (m + 2) * sizeof(double)));
  K.m = m;
  K.prec = prec;
  return GM_OK;
}

// Restarted GMRES(m), right preconditioned, classical Gram-Schmidt with a
// second pass when the first one cancels more than 1/sqrt(2) of the norm,
// convergence ||b - A x|| / ||b - A x0|| < tol (Goma's AztecOO defaults
// AZ_gmres / AZ_classic / AZ_r0, src/sl_util.c:392-960).
template <typename BT, typename VT>
static int gmres_t(gm_ctx *c, const VT *a, const VT *lu, VT *ywork, const double *b, double *x,
                   const gm_solver_opts *o, int *its_out, double *relres) {
  const int N = c->N, m = o->kspace;
  KrylovWork &K = c->kry;
  BT *V = sizeof(BT) == 8 ? (BT *)K.d_V : (BT *)K.d_V32;
  BT *w = sizeof(BT) == 8 ? (BT *)K.d_w : (BT *)K.d_w32;
  BT *z = sizeof(BT) == 8 ? (BT *)K.d_z : (BT *)K.d_z32;
  BT *u = z + N;
  double *ax = K.d_w;  // FP64 scratch for A x (FP64 mode reuses w after use)
  double *ax2 = K.d_z + N;
  const bool pc = o->precond != GM_PC_NONE;
  std::vector<double> H((size_t)(m + 1) * m, 0.0), g(m + 1), cs(m), sn(m), y(m), h(m + 2), h2(m + 2);
  const int grid = 2048;

  auto precond = [&](const BT *in, BT *out) {
    if (o->precond == GM_PC_MG) {
      if constexpr (sizeof(VT) == 8) gm_mg_apply_d(c, (const double *)a, (const double *)in, (double *)out, o);
      else gm_mg_apply_f(c, (const float *)a, (const double *)in, (double *)out, o);
    } else if (o->precond == GM_PC_BAND) {
      gm_band_apply(c, (const double *)in, (double *)out);
    } else if (pc) {
      ilu_apply<VT, BT, BT>(c, lu, ywork, in, out);
    } else {
      k_scale_into<BT><<<grid, 256, 0, c->stream>>>(N, in, 1.0, out);
    }
  };

  // r0 = b - A x
  double *axp = sizeof(BT) == 8 ? ax2 : ax;
  spmv<VT, double, double>(c, a, x, axp);
  k_resid<BT><<<KB, RED_THREADS, 0, c->stream>>>(N, b, axp, w, K.d_part);
  GM_HIP(hipGetLastError());
  double s;
  int rc = sum_parts(c, K.d_part, 1, KB, &s);
  if (rc) return rc;
  const double r0 = std::sqrt(s);
  double rn = r0;
  int its = 0;
  if (r0 == 0.0) {
    *its_out = 0;
    *relres = 0.0;
    return GM_OK;
  }
  const double eta2 = 0.5;  // (1/sqrt 2)^2
  while (its < o->max_its) {
    const double beta = rn;
    k_scale_into<BT><<<grid, 256, 0, c->stream>>>(N, w, 1.0 / beta, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j;
    for (j = 0; j < m && its < o->max_its; j++) {
      BT *vj = V + (size_t)j * N, *vn = V + (size_t)(j + 1) * N;
      precond(vj, z);
      spmv<VT, BT, BT>(c, a, z, w);
      gm_launch(c, GM_K_MDOT, (double)(j + 2) * N * sizeof(BT), k_mdot<BT>, dim3(KB),
                dim3(RED_THREADS), 0, c->stream, N, j + 1, (const BT *)V, (const BT *)w, K.d_part);
      GM_HIP(hipGetLastError());
      if ((rc = sum_parts(c, K.d_part, j + 2, KB, h.data()))) return rc;
      const double ww0 = h[j + 1];
      GM_HIP(hipMemcpyAsync(K.d_h, h.data(), (j + 1) * sizeof(double), hipMemcpyHostToDevice, c->stream));
      gm_launch(c, GM_K_ORTH, (double)(j + 3) * N * sizeof(BT), k_orth_update<BT>, dim3(KB),
                dim3(RED_THREADS), 0, c->stream, N, j + 1, (const BT *)V, (const double *)K.d_h, w,
                K.d_part);
      double ww;
      if ((rc = sum_parts(c, K.d_part, 1, KB, &ww))) return rc;
      if (ww < eta2 * ww0) {
        k_mdot<BT><<<KB, RED_THREADS, 0, c->stream>>>(N, j + 1, V, w, K.d_part);
        if ((rc = sum_parts(c, K.d_part, j + 2, KB, h2.data()))) return rc;
        GM_HIP(hipMemcpyAsync(K.d_h, h2.data(), (j + 1) * sizeof(double), hipMemcpyHostToDevice, c->stream));
        k_orth_update<BT><<<KB, RED_THREADS, 0, c->stream>>>(N, j + 1, V, K.d_h, w, K.d_part);
        if ((rc = sum_parts(c, K.d_part, 1, KB, &ww))) return rc;
        for (int k = 0; k <= j; k++) h[k] += h2[k];
      }
      const double hn = std::sqrt(ww);
      double *hj = H.data() + (size_t)j * (m + 1);
      for (int k = 0; k <= j; k++) hj[k] = h[k];
      hj[j + 1] = hn;
      if (hn != 0.0) k_scale_into<BT><<<grid, 256, 0, c->stream>>>(N, w, 1.0 / hn, vn);
      for (int k = 0; k < j; k++) {
        const double t0 = cs[k] * hj[k] + sn[k] * hj[k + 1];
        hj[k + 1] = -sn[k] * hj[k] + cs[k] * hj[k + 1];
        hj[k] = t0;
      }
      const double den = std::sqrt(hj[j] * hj[j] + hj[j + 1] * hj[j + 1]);
      cs[j] = den == 0.0 ? 1.0 : hj[j] / den;
      sn[j] = den == 0.0 ? 0.0 : hj[j + 1] / den;
      hj[j] = den;
      hj[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      its++;
      if (std::fabs(g[j + 1]) / r0 < o->lin_tol || hn == 0.0) {
        j++;
        break;
      }
    }
    for (int k = j - 1; k >= 0; k--) {
      double t = g[k];
      for (int l = k + 1; l < j; l++) t -= H[(size_t)l * (m + 1) + k] * y[l];
      y[k] = t / H[(size_t)k * (m + 1) + k];
    }
    GM_HIP(hipMemcpyAsync(K.d_h, y.data(), j * sizeof(double), hipMemcpyHostToDevice, c->stream));
    k_combine<BT><<<grid, 256, 0, c->stream>>>(N, j, V, K.d_h, u);
    precond(u, z);
    k_add<BT><<<grid, 256, 0, c->stream>>>(N, x, z);
    spmv<VT, double, double>(c, a, x, axp);
    k_resid<BT><<<KB, RED_THREADS, 0, c->stream>>>(N, b, axp, w, K.d_part);
    GM_HIP(hipGetLastError());
    if ((rc = sum_parts(c, K.d_part, 1, KB, &s))) return rc;
    rn = std::sqrt(s);
    if (rn / r0 < o->lin_tol) break;
  }
  *its_out = its;
  *relres = rn / r0;
  return GM_OK;
}

int gm_gmres_run(gm_ctx *c, const double *val, const float *val32, const double *b, double *x,
                 const gm_solver_opts *o, int *its, double *relres) {
  if (o->kspace < 1 || o->kspace > 120) {
    gm_set_error("kspace must be in [1, 120]");
    return GM_ERR_ARG;
  }
  const int prec = o->precision;
  int rc = ensure_krylov(c, o->kspace, prec);
  if (rc) return rc;
  if (o->precond == GM_PC_ILU0) {
    int bins = o->patch_bins > 0 ? o->patch_bins : gm_default_bins(c);
    if ((rc = gm_ilu_plan(c, bins))) return rc;
    if ((rc = gm_ilu_factor(c, val, val32, prec))) return rc;
  } else if (o->precond == GM_PC_BAND) {
    if ((rc = gm_band_factor(c, val, val32, prec))) return rc;
  } else if (o->precond == GM_PC_MG &&
             (!c->mg0.d_ainv || c->mg0.d_ainv != c->mg0.d_ainv_p[prec == GM_PREC_MIXED ? 1 : 0])) {
    gm_set_error("multigrid preconditioner not set up (gm_mg_setup_dev)");
    return GM_ERR_ARG;
  }
  return gm_gmres_run_nofactor(c, val, val32, b, x, o, its, relres);
}

int gm_gmres_run_nofactor(gm_ctx *c, const double *val, const float *val32, const double *b,
                          double *x, const gm_solver_opts *o, int *its, double *relres) {
  const int prec = o->precision;
  int rc = ensure_krylov(c, o->kspace, GM_PREC_FP64);
  if (rc) return rc;
  if (prec == GM_PREC_MIXED) {
    if (!val32) {
      gm_set_error("mixed precision needs FP32 Jacobian values");
      return GM_ERR_ARG;
    }
    // FP32 Jacobian values and preconditioner; Krylov basis, dot products
    // and the update in FP64
    return gmres_t<double, float>(c, val32, c->ilu.d_lu32, c->ilu.d_y32, b, x, o, its, relres);
  }
  return gmres_t<double, double>(c, val, c->ilu.d_lu, c->ilu.d_y, b, x, o, its, relres);
}

void gm_free_linalg(gm_ctx *c) {
  free_plan(c->ilu);
  KrylovWork &K = c->kry;
  void *ptrs[] = {K.d_V, K.d_w, K.d_z, K.d_part, K.d_h, K.d_V32, K.d_w32, K.d_z32};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  c->kry = KrylovWork();
}

int gm_ilu_factor_export(gm_ctx *c, const double *val, int bins, int *rowptr, int *col, double *lu) {
  int rc = gm_ilu_plan(c, bins);
  if (rc) return rc;
  if ((rc = gm_ilu_factor(c, val, nullptr, GM_PREC_FP64))) return rc;
  IluPlan &P = c->ilu;
  GM_HIP(hipMemcpy(rowptr, P.d_rowptr, (c->N + 1) * sizeof(int), hipMemcpyDeviceToHost));
  GM_HIP(hipMemcpy(col, P.d_col, P.nnz * sizeof(int), hipMemcpyDeviceToHost));
  GM_HIP(hipMemcpy(lu, P.d_lu, P.nnz * sizeof(double), hipMemcpyDeviceToHost));
  return (int)P.nnz;
}
