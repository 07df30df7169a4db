// Banded LU factorisation and solve of the row-sum-scaled Jacobian, without
// pivoting, in Goma's unknown order: a direct solver on the GPU for the
// Newton systems where iterative preconditioners stall -- the role of Goma's
// direct solvers (Solution Algorithm = lu, Goma's default, src/mm_input.c:5904;
// umf / lu dispatch in src/mm_sol_nonlinear.c:1624-1681) -- behind the same
// GMRES interface as the other preconditioners.
//
// Storage.  With kl / ku the lower / upper bandwidths of the MSR graph, the
// band lives column-major in ld = kl + ku + 1 + 2 MG entries per column,
// A(i, j) at ab[off + i + j * lda] with off = MG + ku and lda = ld - 1, for
// j - ku - MG <= i <= j + kl + MG.  The MG-entry margins above and below the
// band let every panel step treat its window as a plain dense matrix: window
// entries outside the band start at zero and stay exactly zero (each of their
// updates has a zero factor).  Without pivoting no fill leaves the band.
//
// Factorisation, right-looking in panels of NB = 64 columns:
//   k_band_a11     A11 = L11 U11 (one workgroup, in LDS), for the first
//                  panel; later ones are factored at the end of the previous
//                  update, by the workgroup whose tile holds them;
//   k_band_panel   L21 = A21 U11^-1 (a thread per row) and
//                  U12 = L11^-1 A12 (a thread per column);
//   k_band_update  A22 -= L21 U12 (64 x 64 tiles, 4 x 4 per thread).
// A launch never writes what another of its workgroups reads.
// Solve: forward and backward substitution, one launch per NB block: a wave
// solves the block's triangle, the workgroup updates the rows the block
// couples to.  Zero pivots are reported (GM_ERR_PIVOT), never divided by.
#include <algorithm>
#include <cmath>

#include "gm_launch.h"

namespace {

constexpr int NB = 64;  // panel width
constexpr int MG = NB;  // storage margin above and below the band

template <typename T>
__global__ __launch_bounds__(256) void k_band_zero(size_t n, T *__restrict__ p) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = T(0);
}

template <typename VT, typename T>
__global__ __launch_bounds__(256) void k_band_load(int N, const VT *__restrict__ val,
                                                   const int *__restrict__ bindx, T *__restrict__ ab,
                                                   long long off, long long lda) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  ab[off + i + (size_t)i * lda] = (T)val[i];
  for (int k = bindx[i]; k < bindx[i + 1]; k++) ab[off + i + (size_t)bindx[k] * lda] = (T)val[k];
}

// In-place LU (no pivoting) of the kb x kb block in LDS by all 256 threads of
// the workgroup (every thread must call it): on return s holds L (strictly
// lower, unit diagonal implied) and U.
template <typename T>
__device__ inline void block_lu(T (*s)[NB + 1], int kb, int *flag) {
  // thread (i = tid / 4, q = tid % 4) owns row i, columns q, q + 4, ..., q + 60
  const int tid = threadIdx.x, i = tid >> 2, q = tid & 3;
  for (int k = 0; k < kb; k++) {
    __syncthreads();
    T piv = s[k][k];
    if (piv == T(0)) {
      if (tid == 0) atomicOr(flag, 1);
      piv = T(1);
    }
    const bool act = i > k && i < kb;
    T l = T(0);
    if (act) l = s[i][k] / piv;
    __syncthreads();  // every thread has read s[i][k] and row k before they change
    if (act) {
      if (q == 0) s[i][k] = l;
#pragma unroll
      for (int c = 0; c < NB / 4; c++) {
        const int j = q + 4 * c;
        if (j > k && j < kb) s[i][j] -= l * s[k][j];
      }
    }
  }
  __syncthreads();
}

// A11 = L11 U11 in place (the first panel, or a panel whose predecessor
// launched no update).
template <typename T>
__global__ __launch_bounds__(256) void k_band_a11(int j0, int kb, T *__restrict__ ab, long long off,
                                                  long long lda, int *flag) {
  __shared__ T sA[NB][NB + 1];
  T *__restrict__ base = ab + off;
  const int tid = threadIdx.x;
  for (int t = tid; t < NB * NB; t += 256) {
    const int j = t >> 6, i = t & 63;
    sA[i][j] = (i < kb && j < kb) ? base[(j0 + i) + (size_t)(j0 + j) * lda] : T(0);
  }
  block_lu<T>(sA, kb, flag);
  for (int t = tid; t < NB * NB; t += 256) {
    const int j = t >> 6, i = t & 63;
    if (i < kb && j < kb) base[(j0 + i) + (size_t)(j0 + j) * lda] = sA[i][j];
  }
}

// With A11 factored: workgroups [0, nrb) compute 256 rows of L21 each,
// [nrb, gridDim.x) 256 columns of U12.  A11 is only read.
template <typename T>
__global__ __launch_bounds__(256) void k_band_panel(int j0, int kb, int R, int C, int nrb,
                                                    T *__restrict__ ab, long long off, long long lda) {
  __shared__ T sA[NB][NB + 1];
  __shared__ T sR[NB];  // 1 / U11 diagonal
  const int tid = threadIdx.x;
  T *__restrict__ base = ab + off;
  for (int t = tid; t < NB * NB; t += 256) {
    const int j = t >> 6, i = t & 63;
    sA[i][j] = (i < kb && j < kb) ? base[(j0 + i) + (size_t)(j0 + j) * lda] : T(0);
  }
  if (tid < NB) {
    const T d = tid < kb ? base[(j0 + tid) + (size_t)(j0 + tid) * lda] : T(1);
    sR[tid] = T(1) / (d == T(0) ? T(1) : d);
  }
  __syncthreads();
  const int b = blockIdx.x;
  if (b < nrb) {
    // L21 row: l U11 = a
    const int rr = b * 256 + tid;
    if (rr >= R) return;
    const size_t i = (size_t)j0 + kb + rr;
    T a[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) a[k] = k < kb ? base[i + (size_t)(j0 + k) * lda] : T(0);
#pragma unroll
    for (int k = 0; k < NB; k++) {
      if (k < kb) {
        a[k] *= sR[k];
#pragma unroll
        for (int j = k + 1; j < NB; j++) a[j] -= a[k] * sA[k][j];
      }
    }
#pragma unroll
    for (int k = 0; k < NB; k++)
      if (k < kb) base[i + (size_t)(j0 + k) * lda] = a[k];
  } else {
    // U12 column: L11 u = a (unit lower)
    const int cc = (b - nrb) * 256 + tid;
    if (cc >= C) return;
    const size_t j = (size_t)j0 + kb + cc;
    T a[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) a[k] = k < kb ? base[(j0 + k) + j * lda] : T(0);
#pragma unroll
    for (int k = 0; k < NB; k++) {
      if (k < kb) {
#pragma unroll
        for (int i = k + 1; i < NB; i++) a[i] -= sA[i][k] * a[k];
      }
    }
#pragma unroll
    for (int k = 0; k < NB; k++)
      if (k < kb) base[(j0 + k) + j * lda] = a[k];
  }
}

// A[rs..re) x [cs..ce) -= L(rows, k0..k0+K) U(k0..k0+K, cols), K taken in
// chunks of NB; 64 x 64 tiles, 4 x 4 per thread (measured faster than 128 x 128
// tiles with 8 x 8 per thread, whose LDS allows one workgroup per CU).  With
// fuse > 0, workgroup (0, 0) -- whose tile then holds the next panel's
// diagonal block, now final -- also factors that block's fuse x fuse corner in
// place (from its own tile in LDS).
template <typename T>
__global__ __launch_bounds__(256) void k_band_update(int k0, int K, int rs, int re, int cs, int ce,
                                                     int fuse, T *__restrict__ ab, long long off,
                                                     long long lda, int *flag) {
  __shared__ T sL[NB][NB + 1];  // [k][row]; then the updated tile of workgroup (0, 0), [row][col]
  __shared__ T sU[NB][NB + 1];  // [k][column]
  T *__restrict__ base = ab + off;
  const int tid = threadIdx.x;
  const int r0 = rs + blockIdx.x * NB, c0 = cs + blockIdx.y * NB;
  const int tr = tid & 15, tc = tid >> 4;
  T acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) acc[u][v] = T(0);
  for (int kc = 0; kc < K; kc += NB) {
    const int kn = min(NB, K - kc);
    if (kc) __syncthreads();
    for (int t = tid; t < NB * NB; t += 256) {
      const int k = t >> 6, i = t & 63;
      sL[k][i] = (k < kn && r0 + i < re) ? base[(size_t)(r0 + i) + (size_t)(k0 + kc + k) * lda] : T(0);
    }
    for (int t = tid; t < NB * NB; t += 256) {
      const int j = t >> 6, k = t & 63;
      sU[k][j] = (k < kn && c0 + j < ce) ? base[(size_t)(k0 + kc + k) + (size_t)(c0 + j) * lda] : T(0);
    }
    __syncthreads();
    for (int k = 0; k < kn; k++) {
      T a[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        a[u] = sL[k][tr + 16 * u];
        bv[u] = sU[k][tc + 16 * u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[u][v] += a[u] * bv[v];
    }
  }
  const bool fz = fuse > 0 && blockIdx.x == 0 && blockIdx.y == 0;
  if (fz) __syncthreads();  // uniform per workgroup: sL is reused below
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int cj = c0 + tc + 16 * v;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int ri = r0 + tr + 16 * u;
      if (ri < re && cj < ce) {
        T *p = base + (size_t)ri + (size_t)cj * lda;
        const T nv = *p - acc[u][v];
        *p = nv;
        if (fz) sL[tr + 16 * u][tc + 16 * v] = nv;
      }
    }
  }
  if (!fz) return;
  block_lu<T>(sL, fuse, flag);
  for (int t = tid; t < NB * NB; t += 256) {
    const int j = t >> 6, i = t & 63;
    if (i < fuse && j < fuse) base[(size_t)(r0 + i) + (size_t)(c0 + j) * lda] = sL[i][j];
  }
}

// Forward substitution of one block: y[blk] = L11^-1 w[blk] (w holds b minus
// the earlier blocks' contributions), then w[j0+kb .. j0+kb+R) -= L21 y[blk].
// w[blk] is only read in this launch, so every workgroup solves the block
// itself (from L11 staged in LDS); workgroup 0 writes it to y.
template <typename T>
__global__ __launch_bounds__(256) void k_band_fwd(int j0, int kb, int R, const T *__restrict__ ab,
                                                  long long off, long long lda, double *__restrict__ w,
                                                  double *__restrict__ y) {
  __shared__ T sT[NB][NB + 1];
  __shared__ double sy[NB];
  const T *__restrict__ base = ab + off;
  const int tid = threadIdx.x, lane = tid & 63;
  for (int t = tid; t < NB * NB; t += 256) {
    const int j = t >> 6, i = t & 63;
    sT[i][j] = (i < kb && j < kb && i > j) ? base[(j0 + i) + (size_t)(j0 + j) * lda] : T(0);
  }
  if (tid < NB) sy[tid] = tid < kb ? w[j0 + tid] : 0.0;
  __syncthreads();
  if (tid < 64) {
    double v = sy[lane];
    for (int k = 0; k < kb; k++) {
      const double vk = __shfl(v, k);
      v -= (double)sT[lane][k] * vk;  // zero for lane <= k
    }
    sy[lane] = v;
    if (blockIdx.x == 0 && lane < kb) y[j0 + lane] = v;
  }
  __syncthreads();
  const int rr = blockIdx.x * 256 + tid;
  if (rr >= R) return;
  const size_t i = (size_t)j0 + kb + rr;
  const T *__restrict__ col = base + i + (size_t)j0 * lda;
  double s0 = 0.0, s1 = 0.0;
  int k = 0;
#pragma unroll 8
  for (; k + 1 < kb; k += 2) {
    s0 += (double)col[(size_t)k * lda] * sy[k];
    s1 += (double)col[(size_t)(k + 1) * lda] * sy[k + 1];
  }
  if (k < kb) s0 += (double)col[(size_t)k * lda] * sy[k];
  w[i] -= s0 + s1;
}

// Backward substitution of one block: x[blk] = U11^-1 w[blk], then
// w[j0-R .. j0) -= U(.., blk) x[blk].
template <typename T>
__global__ __launch_bounds__(256) void k_band_bwd(int j0, int kb, int R, const T *__restrict__ ab,
                                                  long long off, long long lda, double *__restrict__ w,
                                                  double *__restrict__ x) {
  __shared__ T sT[NB][NB + 1];
  __shared__ double sx[NB], sr[NB];
  const T *__restrict__ base = ab + off;
  const int tid = threadIdx.x, lane = tid & 63;
  for (int t = tid; t < NB * NB; t += 256) {
    const int j = t >> 6, i = t & 63;
    sT[i][j] = (i < kb && j < kb && i <= j) ? base[(j0 + i) + (size_t)(j0 + j) * lda] : T(0);
  }
  if (tid < NB) {
    sx[tid] = tid < kb ? w[j0 + tid] : 0.0;
    const double d = tid < kb ? (double)base[(j0 + tid) + (size_t)(j0 + tid) * lda] : 1.0;
    sr[tid] = 1.0 / (d == 0.0 ? 1.0 : d);
  }
  __syncthreads();
  if (tid < 64) {
    double v = sx[lane];
    for (int k = kb - 1; k >= 0; k--) {
      if (lane == k) v *= sr[k];
      const double vk = __shfl(v, k);
      if (lane < k) v -= (double)sT[lane][k] * vk;
    }
    sx[lane] = v;
    if (blockIdx.x == 0 && lane < kb) x[j0 + lane] = v;
  }
  __syncthreads();
  const int rr = blockIdx.x * 256 + tid;
  if (rr >= R) return;
  const size_t i = (size_t)j0 - R + rr;
  const T *__restrict__ col = base + i + (size_t)j0 * lda;
  double s0 = 0.0, s1 = 0.0;
  int k = 0;
#pragma unroll 8
  for (; k + 1 < kb; k += 2) {
    s0 += (double)col[(size_t)k * lda] * sx[k];
    s1 += (double)col[(size_t)(k + 1) * lda] * sx[k + 1];
  }
  if (k < kb) s0 += (double)col[(size_t)k * lda] * sx[k];
  w[i] -= s0 + s1;
}

__global__ __launch_bounds__(256) void k_band_copy(int n, const double *__restrict__ a, double *__restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = a[i];
}

}  // namespace

static int band_plan(gm_ctx *c) {
  BandLU &B = c->band;
  if (B.kl >= 0) return GM_OK;
  int kl = 0, ku = 0;
  for (int i = 0; i < c->N; i++)
    for (int k = c->bindx[i]; k < c->bindx[i + 1]; k++) {
      const int j = c->bindx[k];
      kl = std::max(kl, i - j);
      ku = std::max(ku, j - i);
    }
  B.kl = kl;
  B.ku = ku;
  B.ld = (long long)kl + ku + 1 + 2 * MG;
  GM_HIP(hipMalloc(&B.d_w1, (size_t)c->N * sizeof(double)));
  GM_HIP(hipMalloc(&B.d_w2, (size_t)c->N * sizeof(double)));
  return GM_OK;
}

template <typename VT, typename T>
static int band_factor_t(gm_ctx *c, const VT *val, T *ab) {
  BandLU &B = c->band;
  const int N = c->N, kl = B.kl, ku = B.ku;
  const long long off = MG + ku, lda = B.ld - 1;
  const size_t total = (size_t)N * B.ld;
  hipStream_t s = c->stream;
  k_band_zero<T><<<4096, 256, 0, s>>>(total, ab);
  k_band_load<VT, T><<<(N + 255) / 256, 256, 0, s>>>(N, val, c->d_bindx, ab, off, lda);
  GM_HIP(hipMemsetAsync(c->d_flag + 4, 0, sizeof(int), s));
  int *d_pivflag = c->d_flag + 4;
  bool a11_done = false;  // the next panel's A11 was factored by the last update
  auto panel = [&](int j0, int kb) {
    const int R = std::min(kl, N - j0 - kb), C = std::min(ku, N - j0 - kb);
    const int nrb = (R + 255) / 256, ncb = (C + 255) / 256;
    if (!a11_done)
      gm_launch(c, GM_K_BAND_PANEL, 2.0 * kb * kb * sizeof(T), k_band_a11<T>, dim3(1), dim3(256), 0, s,
                j0, kb, ab, off, lda, d_pivflag);
    a11_done = false;
    if (nrb + ncb > 0) {
      const double pbytes = ((double)kb * kb + (double)(R + C) * kb * 2.0) * sizeof(T);
      gm_launch(c, GM_K_BAND_PANEL, pbytes, k_band_panel<T>, dim3(nrb + ncb), dim3(256), 0, s, j0, kb,
                R, C, nrb, ab, off, lda);
    }
  };
  auto update = [&](int k0, int K, int rs, int re, int cs, int ce, int fuse) {
    if (re <= rs || ce <= cs) return;
    const double ubytes = ((double)(re - rs + ce - cs) * K + 2.0 * (re - rs) * (ce - cs)) * sizeof(T);
    gm_launch(c, GM_K_BAND_UPDATE, ubytes, k_band_update<T>,
              dim3((re - rs + NB - 1) / NB, (ce - cs + NB - 1) / NB), dim3(256), 0, s, k0, K, rs, re, cs, ce,
              fuse, ab, off, lda, d_pivflag);
    if (fuse) a11_done = true;
  };
  for (int j0 = 0; j0 < N;) {
    const int kb = std::min(NB, N - j0);
    const int R = std::min(kl, N - j0 - kb), C = std::min(ku, N - j0 - kb);
    const int next_kb = std::min(NB, std::max(0, N - j0 - kb));
    panel(j0, kb);
    if (R > 0 && C > 0)
      update(j0, kb, j0 + kb, j0 + kb + R, j0 + kb, j0 + kb + C,
             (next_kb > 0 && R >= next_kb && C >= next_kb) ? next_kb : 0);
    j0 += kb;
  }
  GM_HIP(hipGetLastError());
  int flag = 0;
  GM_HIP(hipMemcpyAsync(&flag, c->d_flag + 4, sizeof(int), hipMemcpyDeviceToHost, s));
  GM_HIP(hipStreamSynchronize(s));
  if (flag) {
    gm_set_error("zero pivot in the banded LU (no pivoting)");
    return GM_ERR_PIVOT;
  }
  return GM_OK;
}

int gm_band_factor(gm_ctx *c, const double *val, const float *val32, int prec) {
  int rc = band_plan(c);
  if (rc) return rc;
  BandLU &B = c->band;
  const int pi = prec == GM_PREC_MIXED ? 1 : 0;
  const size_t es = pi ? sizeof(float) : sizeof(double);
  if (!B.d_ab[pi]) {
    const hipError_t e = hipMalloc(&B.d_ab[pi], (size_t)c->N * B.ld * es);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      B.d_ab[pi] = nullptr;
      gm_set_error("banded LU needs %.1f GB (N %d, bandwidths %d / %d)",
                   (double)c->N * B.ld * es / 1e9, c->N, B.kl, B.ku);
      return GM_ERR_ARG;
    }
  }
  B.prec = pi;
  if (pi) {
    if (!val32) {
      gm_set_error("mixed precision needs FP32 Jacobian values");
      return GM_ERR_ARG;
    }
    return band_factor_t<float, float>(c, val32, (float *)B.d_ab[1]);
  }
  return band_factor_t<double, double>(c, val, (double *)B.d_ab[0]);
}

template <typename T>
static void band_apply_t(gm_ctx *c, const T *ab, const double *in, double *out) {
  BandLU &B = c->band;
  const int N = c->N, kl = B.kl, ku = B.ku;
  const long long off = MG + ku, lda = B.ld - 1;
  hipStream_t s = c->stream;
  k_band_copy<<<1024, 256, 0, s>>>(N, in, B.d_w1);
  for (int j0 = 0; j0 < N; j0 += NB) {
    const int kb = std::min(NB, N - j0);
    const int R = std::min(kl, N - j0 - kb);
    const double bytes = ((double)kb * (kb + R)) * sizeof(T) + 16.0 * (kb + R);
    gm_launch(c, GM_K_BAND_SOLVE, bytes, k_band_fwd<T>, dim3(std::max(1, (R + 255) / 256)), dim3(256), 0,
              s, j0, kb, R, ab, off, lda, B.d_w1, B.d_w2);
  }
  const int nblk = (N + NB - 1) / NB;
  for (int b = nblk - 1; b >= 0; b--) {
    const int j0 = b * NB, kb = std::min(NB, N - j0);
    const int R = std::min(ku, j0);
    const double bytes = ((double)kb * (kb + R)) * sizeof(T) + 16.0 * (kb + R);
    gm_launch(c, GM_K_BAND_SOLVE, bytes, k_band_bwd<T>, dim3(std::max(1, (R + 255) / 256)), dim3(256), 0,
              s, j0, kb, R, ab, off, lda, B.d_w2, out);
  }
}

int gm_band_apply(gm_ctx *c, const double *in, double *out) {
  BandLU &B = c->band;
  if (B.prec < 0 || !B.d_ab[B.prec]) {
    gm_set_error("banded LU not factored");
    return GM_ERR_ARG;
  }
  if (B.prec == 1) band_apply_t<float>(c, (const float *)B.d_ab[1], in, out);
  else band_apply_t<double>(c, (const double *)B.d_ab[0], in, out);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

void gm_band_free(gm_ctx *c) {
  BandLU &B = c->band;
  void *p[] = {B.d_ab[0], B.d_ab[1], B.d_w1, B.d_w2};
  for (void *q : p)
    if (q) (void)hipFree(q);
  c->band = BandLU();
}
