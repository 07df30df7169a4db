// Learned initial guess: a coordinate-plus-parameter MLP trained and
// evaluated on the matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Layout.  Every layer is padded to multiples of 64 (inputs to 64, hidden
// width as given, outputs to 64) so one GEMM kernel serves all layers; the
// padded weights are zero and stay zero (their gradients vanish: padded
// inputs are zero, padded outputs carry a zero mask).  Activations of a batch
// are row-major [points][features] in BF16; master weights, Adam state and
// gradients are FP32; each update rewrites BF16 copies W [N][K] (forward) and
// W^T [K][N] (backward data GEMM), so every GEMM reads both operands
// K-contiguous.  The weight-gradient GEMM contracts over points: it stages
// 32-point tiles through LDS transposed, splits the batch into 512-point
// chunks and sums the chunk partials in a fixed order, so training is
// deterministic.
//
// MFMA 16x16x32 BF16 lane layout (wave64): lane l holds row l%16 of A and
// column l%16 of B, reduction indices 8*(l/16) .. 8*(l/16)+7 of both; the
// FP32 result holds column l%16, rows 4*(l/16) .. 4*(l/16)+3.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "gm_launch.h"

typedef uint16_t bf16_t;  // BF16 storage
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

static constexpr int MT = 64;           // feature tile: widths and padded layer sizes
static constexpr int PT = 128;          // point granularity of a padded batch
static constexpr int SPLIT = 512;       // points per weight-gradient partial
static constexpr int LOSS_BLOCKS = 256;
static constexpr int IN_PAD = 64, OUT_PAD = 64;
static constexpr int QBLK = 512;        // blocks of the subspace dot products
static constexpr int MAX_BASIS = 128;

// round to nearest even (NaN kept quiet), as the oracle's omlp_bf16
__device__ inline bf16_t f2bf(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) {
    if (u & 0x007fffffu) u |= 0x00400000u;
  } else {
    u += 0x7fffu + ((u >> 16) & 1u);
  }
  return (bf16_t)(u >> 16);
}
__device__ inline float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

__device__ inline bf8 load8(const bf16_t *p) { return __builtin_bit_cast(bf8, *(const u4v *)p); }

struct InNorm {
  float sh[GM_MLP_MAX_IN], sc[GM_MLP_MAX_IN];
};

// ------------------------------------------------------------------ kernels
// batch input: a0 [Mpad][64] = bf16((in - shift) * scale), zero padded;
// targets and masks of the batch (zero mask on padded points).
__global__ __launch_bounds__(256) void k_mlp_gather(int M, int Mpad, const int *__restrict__ idx,
                                                    const float *__restrict__ in,
                                                    const float *__restrict__ t,
                                                    const float *__restrict__ m, int n_in,
                                                    int n_out, InNorm nrm, bf16_t *__restrict__ a0,
                                                    float *__restrict__ bt, float *__restrict__ bm) {
  const long long total = (long long)Mpad * IN_PAD;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int p = (int)(i / IN_PAD), col = (int)(i % IN_PAD);
    const long long src = p < M ? (idx ? idx[p] : p) : -1;
    float v = 0.f;
    if (src >= 0 && col < n_in) v = (in[src * n_in + col] - nrm.sh[col]) * nrm.sc[col];
    a0[i] = f2bf(v);
    if (bt && col < n_out) {
      bt[(long long)p * n_out + col] = src >= 0 ? t[src * n_out + col] : 0.f;
      bm[(long long)p * n_out + col] = src >= 0 ? m[src * n_out + col] : 0.f;
    }
  }
}

enum { EPI_TANH = 0, EPI_LIN = 1, EPI_DTANH = 2 };

// C [M][N] = A [M][K] * B[N][K]^T, BF16 operands, FP32 accumulation.
// Four waves as 2 x 2; each wave computes (16 TM) x (16 TN) with TM x TN MFMA
// tiles, so one k-step issues TM x TN MFMAs for TM + TN operand loads; the
// block covers (32 TM) x (32 TN).  M a multiple of 32 TM, N of 32 TN, K of 32.
// Epilogues: tanh(C + bias) -> BF16; C + bias -> FP32; C * (1 - aux^2) -> BF16
// (back-propagation through the tanh that made aux).
template <int EPI, int TM, int TN>
__global__ __launch_bounds__(256) void k_mlp_gemm(int M, int N, int K, const bf16_t *__restrict__ A,
                                                  const bf16_t *__restrict__ B,
                                                  const float *__restrict__ bias,
                                                  const bf16_t *__restrict__ aux,
                                                  bf16_t *__restrict__ o16, float *__restrict__ o32) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = blockIdx.x * (32 * TM) + (wave >> 1) * (16 * TM);
  const int n0 = blockIdx.y * (32 * TN) + (wave & 1) * (16 * TN);
  const int r = lane & 15, kg = (lane >> 4) * 8;
  f4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const bf16_t *pa = A + (size_t)(m0 + r) * K + kg;
  const bf16_t *pb = B + (size_t)(n0 + r) * K + kg;
  for (int k = 0; k < K; k += 32) {
    bf8 a[TM], b[TN];
#pragma unroll
    for (int i = 0; i < TM; i++) a[i] = load8(pa + (size_t)16 * i * K + k);
#pragma unroll
    for (int j = 0; j < TN; j++) b[j] = load8(pb + (size_t)16 * j * K + k);
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int m = m0 + 16 * i + 4 * (lane >> 4) + q, n = n0 + 16 * j + r;
        const size_t o = (size_t)m * N + n;
        const float v = acc[i][j][q];
        if (EPI == EPI_TANH) {
          o16[o] = f2bf(tanhf(v + bias[n]));
        } else if (EPI == EPI_LIN) {
          o32[o] = v + bias[n];
        } else {
          const float a = bf2f(aux[o]);
          o16[o] = f2bf(v * (1.f - a * a));
        }
      }
}

// GEMM launch: 128 x 128 blocks when N allows, 128 x 64 otherwise
template <int EPI>
static void gemm(hipStream_t s, int M, int N, int K, const bf16_t *A, const bf16_t *B,
                 const float *bias, const bf16_t *aux, bf16_t *o16, float *o32) {
  if (N % 128 == 0)
    k_mlp_gemm<EPI, 4, 4><<<dim3(M / 128, N / 128), 256, 0, s>>>(M, N, K, A, B, bias, aux, o16, o32);
  else
    k_mlp_gemm<EPI, 4, 2><<<dim3(M / 128, N / 64), 256, 0, s>>>(M, N, K, A, B, bias, aux, o16, o32);
}

// Weight and bias gradients of one layer over one chunk of points:
//   part[s][n][k] = sum_{m in chunk s} D[m][n] X[m][k],  partb[s][n] = sum D[m][n]
// D [M][N] (deltas), X [M][K] (layer input).  Grid (N/64, K/64, chunks);
// 32-point tiles are transposed through LDS so the MFMA operands are
// contiguous along the point index.
__global__ __launch_bounds__(256) void k_mlp_wgrad(int M, int N, int K, int chunk,
                                                   const bf16_t *__restrict__ D,
                                                   const bf16_t *__restrict__ X,
                                                   float *__restrict__ part,
                                                   float *__restrict__ partb) {
  __shared__ __align__(16) bf16_t sD[MT][40];
  __shared__ __align__(16) bf16_t sX[MT][40];
  const int n0 = blockIdx.x * MT, k0 = blockIdx.y * MT, s = blockIdx.z;
  const int mbeg = s * chunk, mend = min(M, mbeg + chunk);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wn = (wave >> 1) * 32, wk = (wave & 1) * 32;
  const int r = lane & 15, kg = (lane >> 4) * 8;
  const int mm = tid >> 3, c = (tid & 7) * 8;  // staging: point mm, 8 features from c
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (int mb = mbeg; mb < mend; mb += 32) {
    const u4v vd = *(const u4v *)(D + (size_t)(mb + mm) * N + n0 + c);
    const u4v vx = *(const u4v *)(X + (size_t)(mb + mm) * K + k0 + c);
    __syncthreads();  // the previous tile has been consumed
#pragma unroll
    for (int e = 0; e < 4; e++) {
      sD[c + 2 * e][mm] = (bf16_t)(vd[e] & 0xffffu);
      sD[c + 2 * e + 1][mm] = (bf16_t)(vd[e] >> 16);
      sX[c + 2 * e][mm] = (bf16_t)(vx[e] & 0xffffu);
      sX[c + 2 * e + 1][mm] = (bf16_t)(vx[e] >> 16);
    }
    __syncthreads();
    bf8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++) a[i] = load8(&sD[wn + 16 * i + r][kg]);
#pragma unroll
    for (int j = 0; j < 2; j++) b[j] = load8(&sX[wk + 16 * j + r][kg]);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (blockIdx.y == 0 && tid < MT)
      for (int q = 0; q < 32; q++) bsum += bf2f(sD[tid][q]);
  }
  float *out = part + (size_t)s * N * K;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int n = n0 + wn + 16 * i + 4 * (lane >> 4) + q, k = k0 + wk + 16 * j + r;
        out[(size_t)n * K + k] = acc[i][j][q];
      }
  if (blockIdx.y == 0 && tid < MT) partb[(size_t)s * N + n0 + tid] = bsum;
}

// fixed-order sum of the chunk partials
__global__ __launch_bounds__(256) void k_mlp_reduce(int nsplit, int NK, int N,
                                                    const float *__restrict__ part,
                                                    const float *__restrict__ partb,
                                                    float *__restrict__ g, float *__restrict__ gb) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < NK) {
    float s = 0.f;
    for (int k = 0; k < nsplit; k++) s += part[(size_t)k * NK + i];
    g[i] = s;
  }
  if (i < N) {
    float s = 0.f;
    for (int k = 0; k < nsplit; k++) s += partb[(size_t)k * N + i];
    gb[i] = s;
  }
}

// masked MSE: loss partials and the output-layer deltas dz [Mpad][64]
__global__ __launch_bounds__(256) void k_mlp_loss(int M, int Mpad, int n_out, float norm,
                                                  const float *__restrict__ y,
                                                  const float *__restrict__ t,
                                                  const float *__restrict__ m,
                                                  bf16_t *__restrict__ dz, double *__restrict__ lpart) {
  __shared__ double sh[4];
  double s = 0.0;
  const long long total = (long long)Mpad * OUT_PAD;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int p = (int)(i / OUT_PAD), o = (int)(i % OUT_PAD);
    float d = 0.f;
    if (p < M && o < n_out) {
      const float e = y[i] - t[(long long)p * n_out + o], mk = m[(long long)p * n_out + o];
      s += (double)mk * (double)e * (double)e;
      d = 2.f * mk * e * norm;
    }
    dz[i] = f2bf(d);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) lpart[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void k_mlp_loss_sum(const double *__restrict__ lpart, int n, double norm,
                               double *__restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += lpart[i];
    *out = s * norm;
  }
}

struct AdamArgs {
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
};

// Adam on one layer's padded weights (real block Nr x Kr), then the BF16
// copies W [N][K] and W^T [K][N]; padded entries stay zero.
__global__ __launch_bounds__(256) void k_mlp_adam_w(int N, int K, int Nr, int Kr, float *__restrict__ W,
                                                    float *__restrict__ mW, float *__restrict__ vW,
                                                    const float *__restrict__ g, bf16_t *__restrict__ Wb,
                                                    bf16_t *__restrict__ WbT, AdamArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N * K) return;
  const int n = i / K, k = i - n * K;
  float w = W[i];
  if (n < Nr && k < Kr) {
    const float gi = g[i];
    const float mi = a.b1 * mW[i] + (1.f - a.b1) * gi;
    const float vi = a.b2 * vW[i] + (1.f - a.b2) * gi * gi;
    mW[i] = mi;
    vW[i] = vi;
    w -= a.lr * (mi / a.c1) / (sqrtf(vi / a.c2) + a.eps);
    W[i] = w;
  }
  const bf16_t h = f2bf(w);
  Wb[i] = h;
  WbT[(size_t)k * N + n] = h;
}

__global__ void k_mlp_adam_b(int N, int Nr, float *__restrict__ b, float *__restrict__ mb,
                             float *__restrict__ vb, const float *__restrict__ g, AdamArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Nr || i >= N) return;
  const float gi = g[i];
  const float mi = a.b1 * mb[i] + (1.f - a.b1) * gi;
  const float vi = a.b2 * vb[i] + (1.f - a.b2) * gi * gi;
  mb[i] = mi;
  vb[i] = vi;
  b[i] -= a.lr * (mi / a.c1) / (sqrtf(vi / a.c2) + a.eps);
}

__global__ void k_mlp_bf16_copies(int N, int K, const float *__restrict__ W, bf16_t *__restrict__ Wb,
                                  bf16_t *__restrict__ WbT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * K) return;
  const int n = i / K, k = i - n * K;
  const bf16_t h = f2bf(W[i]);
  Wb[i] = h;
  WbT[(size_t)k * N + n] = h;
}

__global__ void k_mlp_copy_out(int M, int n_out, const float *__restrict__ y, float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * n_out) return;
  const int p = i / n_out, o = i - p * n_out;
  out[i] = y[(size_t)p * OUT_PAD + o];
}

// Network outputs at the nodes and element sample points -> unknowns.
struct GuessMap {
  int num_nodes, num_elems, dim, energy, n_out, spe;  // spe: sample points per element
  double scale[GM_MLP_MAX_OUT], shift[GM_MLP_MAX_OUT];
};
// network input of a guess: the sample point's coordinates, then the query's
// parameters
struct GuessParams {
  float p[GM_MLP_MAX_IN];
};
__global__ void k_mlp_guess_input(long long np, int dim, int n_in, const float *__restrict__ pts,
                                  GuessParams q, float *__restrict__ in) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= np * n_in) return;
  const long long p = t / n_in;
  const int i = (int)(t - p * n_in);
  in[t] = i < dim ? pts[p * dim + i] : q.p[i - dim];
}

__global__ void k_mlp_guess_nodes(GuessMap g, const int *__restrict__ fu, const float *__restrict__ y,
                                  double *__restrict__ x) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.num_nodes) return;
  const float *yn = y + (size_t)n * g.n_out;
  const int u = fu[n];
  for (int d = 0; d < g.dim; d++) x[u + d] = g.shift[d] + g.scale[d] * (double)yn[d];
  if (g.energy) x[u + g.dim] = g.shift[g.dim + 1] + g.scale[g.dim + 1] * (double)yn[g.dim + 1];
}
__global__ void k_mlp_guess_elems(GuessMap g, const int *__restrict__ fu, const int *__restrict__ centroid,
                                  const float *__restrict__ y, double *__restrict__ x) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.num_elems) return;
  const int P = g.dim;  // pressure output
  auto p_at = [&](int j) {
    const size_t row = (size_t)g.num_nodes + (size_t)e * g.spe + j;
    return g.shift[P] + g.scale[P] * (double)y[row * g.n_out + P];
  };
  const int u = fu[centroid[e]] + g.dim + g.energy;
  x[u] = p_at(0);
  // points 1 + 2d and 2 + 2d sit at +1/2 and -1/2 along reference axis d
  for (int d = 0; d < g.dim; d++) x[u + 1 + d] = p_at(1 + 2 * d) - p_at(2 + 2 * d);
}
__global__ void k_mlp_delta(int N, const double *__restrict__ xprev, const double *__restrict__ gp,
                            double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) x[i] = xprev[i] + (x[i] - gp[i]);
}
__global__ void k_mlp_dirichlet(int N, const signed char *__restrict__ dbc,
                                const double *__restrict__ val, double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N && dbc[i]) x[i] = val[i];
}

// ---- snapshot subspace (weighted Gram-Schmidt, projection)
__global__ void k_mlp_wdiff(int N, const double *__restrict__ a, const double *__restrict__ b,
                            const double *__restrict__ w, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) out[i] = w[i] * (a[i] - b[i]);
}
// part[k][blk] = sum over the block's share of Q[k] . v, k < nb (fixed order)
__global__ __launch_bounds__(256) void k_mlp_qdots(int N, int nb, const double *__restrict__ Q,
                                                   const double *__restrict__ v,
                                                   double *__restrict__ part) {
  __shared__ double sh[4];
  for (int k = 0; k < nb; k++) {
    const double *q = Q + (size_t)k * N;
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) s += q[i] * v[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[(size_t)k * gridDim.x + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
  }
}
// v -= sum_k c_k Q[k]
__global__ void k_mlp_qsub(int N, int nb, const double *__restrict__ Q, const double *__restrict__ c,
                           double *__restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (int k = 0; k < nb; k++) s += c[k] * Q[(size_t)k * N + i];
  v[i] -= s;
}
__global__ void k_mlp_scale(int N, double *v, double s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) v[i] *= s;
}
// x = x_prev + W^-1 sum_k c_k Q[k]  (zero change where the weight is zero)
__global__ void k_mlp_qapply(int N, int nb, const double *__restrict__ Q, const double *__restrict__ c,
                             const double *__restrict__ w, const double *__restrict__ xprev,
                             double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (int k = 0; k < nb; k++) s += c[k] * Q[(size_t)k * N + i];
  x[i] = xprev[i] + (w[i] > 0.0 ? s / w[i] : 0.0);
}

// ---------------------------------------------- allocation-free inference
// Scratch of gm_mlp_predict_noalloc_dev, one copy per GPU for the life of
// the process: two BF16 activation buffers of NA_PTS points x NA_W features
// (ping-pong through the layers) and the FP32 outputs, NA_PTS x OUT_PAD.
// Points go through it in chunks of NA_PTS.  Every call on a GPU shares that
// GPU's copy: calls that overlap on one GPU race.  Each chunk overwrites every
// entry it reads (inputs and their zero padding, each layer's full output),
// so nothing needs clearing between calls.
static constexpr int NA_PTS = 2048, NA_W = 512;
static constexpr size_t NA_ACT = (size_t)NA_PTS * NA_W * sizeof(bf16_t);
static constexpr size_t NA_BYTES = 2 * NA_ACT + (size_t)NA_PTS * OUT_PAD * sizeof(float);
static_assert(NA_PTS % PT == 0 && NA_W % MT == 0, "scratch shape");
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

static constexpr int NA_MAX_DEV = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NA_MAX_DEV] = {};  // device address of `scratch` on each GPU

// Device address of this GPU's scratch, taken once per device with that
// device current (host code must not pass the symbol itself to a kernel).
static int scratch_init(int dev) {
  if (dev < 0 || dev >= NA_MAX_DEV) {
    gm_set_error("device %d outside the %d devices the inference scratch tracks", dev, NA_MAX_DEV);
    return GM_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return GM_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    gm_set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    return GM_ERR_HIP;
  }
  g_scratch_ptr[dev] = p;
  return GM_OK;
}

// ------------------------------------------------------------------ host side
struct gm_mlp {
  gm_mlp_shape s{};
  int L = 0;
  int dev = -1;  // the GPU the network lives on (current device at creation)
  std::vector<int> K, N, Kr, Nr;
  hipStream_t stream = nullptr;
  std::vector<float *> W, b, mW, vW, mb, vb, gW, gb;
  std::vector<bf16_t *> Wb, WbT;
  long long cap = 0;
  std::vector<bf16_t *> act;
  float *y = nullptr;
  bf16_t *dz[2] = {nullptr, nullptr};
  float *part = nullptr, *partb = nullptr;
  float *bt = nullptr, *bm = nullptr;
  double *lpart = nullptr, *lossbuf = nullptr;
  long long lossbuf_n = 0;
  long long npts = 0;
  float *din = nullptr, *dt = nullptr, *dm = nullptr;
  int *dperm = nullptr;
  long long adam_t = 0;
  // inference / guess scratch
  float *gin = nullptr, *gout = nullptr;
  long long gcap = 0;
  double *gx = nullptr;
  long long gx_n = 0;
  const gm_ctx *gctx = nullptr;       // context whose sample points are cached
  std::vector<float> gpts;            // [npts][dim] coordinates of nodes + element points
  float *d_gpts = nullptr;            // the same on the device
  // snapshot subspace of the delta form
  const gm_ctx *bctx = nullptr;
  int nbasis = 0;
  double *Q = nullptr;                // [nbasis][N], orthonormal in the weighted norm
  double *wgt = nullptr;              // [N]
  double *qpart = nullptr;            // [GM_MLP_MAX_BASIS][QBLK] partial dot products
  double *qc = nullptr;               // [GM_MLP_MAX_BASIS] coefficients
  double *qv = nullptr;               // [N] work vector
};

static uint64_t splitmix(uint64_t &st) {
  uint64_t z = (st += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static double uniform01(uint64_t &st) { return (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0); }

template <typename T>
static int dalloc(T **p, size_t n) {
  GM_HIP(hipMalloc(p, std::max<size_t>(n, 1) * sizeof(T)));
  return GM_OK;
}
static void dfree(void *p) {
  if (p) (void)hipFree(p);
}

static int maxdim(const gm_mlp *m) {
  int d = std::max(IN_PAD, OUT_PAD);
  for (int l = 0; l < m->L; l++) d = std::max(d, std::max(m->K[l], m->N[l]));
  return d;
}

static int ensure_cap(gm_mlp *m, long long M) {
  if (M <= m->cap) return GM_OK;
  const long long cap = std::max<long long>((M + PT - 1) / PT * PT, 4096);
  for (bf16_t *p : m->act) dfree(p);
  dfree(m->y);
  dfree(m->dz[0]);
  dfree(m->dz[1]);
  dfree(m->part);
  dfree(m->partb);
  dfree(m->bt);
  dfree(m->bm);
  m->act.assign(m->L, nullptr);
  int rc;
  for (int l = 0; l < m->L; l++)
    if ((rc = dalloc(&m->act[l], (size_t)cap * m->K[l]))) return rc;
  const int md = maxdim(m);
  const long long nsplit = (cap + SPLIT - 1) / SPLIT;
  if ((rc = dalloc(&m->y, (size_t)cap * OUT_PAD))) return rc;
  if ((rc = dalloc(&m->dz[0], (size_t)cap * md))) return rc;
  if ((rc = dalloc(&m->dz[1], (size_t)cap * md))) return rc;
  if ((rc = dalloc(&m->part, (size_t)nsplit * md * md))) return rc;
  if ((rc = dalloc(&m->partb, (size_t)nsplit * md))) return rc;
  if ((rc = dalloc(&m->bt, (size_t)cap * m->s.n_out))) return rc;
  if ((rc = dalloc(&m->bm, (size_t)cap * m->s.n_out))) return rc;
  m->cap = cap;
  return GM_OK;
}

static InNorm in_norm(const gm_mlp *m) {
  InNorm n{};
  for (int i = 0; i < m->s.n_in; i++) {
    n.sh[i] = (float)m->s.in_shift[i];
    n.sc[i] = (float)m->s.in_scale[i];
  }
  return n;
}

static void gather(gm_mlp *m, int M, int Mpad, const int *idx, const float *in, const float *t,
                   const float *mk) {
  const long long total = (long long)Mpad * IN_PAD;
  const int blocks = (int)std::min<long long>((total + 255) / 256, 4096);
  k_mlp_gather<<<blocks, 256, 0, m->stream>>>(M, Mpad, idx, in, t, mk, m->s.n_in, m->s.n_out,
                                              in_norm(m), m->act[0], t ? m->bt : nullptr,
                                              t ? m->bm : nullptr);
}

static void forward(gm_mlp *m, int Mpad) {
  for (int l = 0; l < m->L; l++) {
    if (l < m->L - 1)
      gemm<EPI_TANH>(m->stream, Mpad, m->N[l], m->K[l], m->act[l], m->Wb[l], m->b[l], nullptr,
                     m->act[l + 1], nullptr);
    else
      gemm<EPI_LIN>(m->stream, Mpad, m->N[l], m->K[l], m->act[l], m->Wb[l], m->b[l], nullptr,
                    nullptr, m->y);
  }
}

// loss of the batch into lossbuf[slot]; output deltas into dz[0]
static void loss(gm_mlp *m, int M, int Mpad, long long slot) {
  const double norm = 1.0 / ((double)M * m->s.n_out);
  k_mlp_loss<<<LOSS_BLOCKS, 256, 0, m->stream>>>(M, Mpad, m->s.n_out, (float)norm, m->y, m->bt,
                                                 m->bm, m->dz[0], m->lpart);
  k_mlp_loss_sum<<<1, 64, 0, m->stream>>>(m->lpart, LOSS_BLOCKS, norm, m->lossbuf + slot);
}

static void backward(gm_mlp *m, int Mpad) {
  int cur = 0;
  for (int l = m->L - 1; l >= 0; l--) {
    const int nsplit = (Mpad + SPLIT - 1) / SPLIT;
    k_mlp_wgrad<<<dim3(m->N[l] / MT, m->K[l] / MT, nsplit), 256, 0, m->stream>>>(
        Mpad, m->N[l], m->K[l], SPLIT, m->dz[cur], m->act[l], m->part, m->partb);
    const int NK = m->N[l] * m->K[l];
    k_mlp_reduce<<<(NK + 255) / 256, 256, 0, m->stream>>>(nsplit, NK, m->N[l], m->part, m->partb,
                                                          m->gW[l], m->gb[l]);
    if (l == 0) break;
    // dz_prev [Mpad][K_l] = (dz [Mpad][N_l] * W_l) * (1 - act_l^2)
    gemm<EPI_DTANH>(m->stream, Mpad, m->K[l], m->N[l], m->dz[cur], m->WbT[l], nullptr, m->act[l],
                    m->dz[1 - cur], nullptr);
    cur = 1 - cur;
  }
}

static void adam(gm_mlp *m, double lr, const gm_mlp_train_opts *o) {
  m->adam_t++;
  AdamArgs a;
  a.lr = (float)lr;
  a.b1 = (float)o->beta1;
  a.b2 = (float)o->beta2;
  a.eps = (float)o->eps;
  a.c1 = (float)(1.0 - std::pow(o->beta1, (double)m->adam_t));
  a.c2 = (float)(1.0 - std::pow(o->beta2, (double)m->adam_t));
  for (int l = 0; l < m->L; l++) {
    const int NK = m->N[l] * m->K[l];
    k_mlp_adam_w<<<(NK + 255) / 256, 256, 0, m->stream>>>(m->N[l], m->K[l], m->Nr[l], m->Kr[l],
                                                          m->W[l], m->mW[l], m->vW[l], m->gW[l],
                                                          m->Wb[l], m->WbT[l], a);
    k_mlp_adam_b<<<(m->N[l] + 255) / 256, 256, 0, m->stream>>>(m->N[l], m->Nr[l], m->b[l], m->mb[l],
                                                               m->vb[l], m->gb[l], a);
  }
}

static int refresh_bf16(gm_mlp *m) {
  for (int l = 0; l < m->L; l++) {
    const int NK = m->N[l] * m->K[l];
    k_mlp_bf16_copies<<<(NK + 255) / 256, 256, 0, m->stream>>>(m->N[l], m->K[l], m->W[l], m->Wb[l],
                                                                m->WbT[l]);
  }
  GM_HIP(hipGetLastError());
  return GM_OK;
}

static int reset_adam(gm_mlp *m) {
  for (int l = 0; l < m->L; l++) {
    const size_t NK = (size_t)m->N[l] * m->K[l];
    GM_HIP(hipMemsetAsync(m->mW[l], 0, NK * sizeof(float), m->stream));
    GM_HIP(hipMemsetAsync(m->vW[l], 0, NK * sizeof(float), m->stream));
    GM_HIP(hipMemsetAsync(m->mb[l], 0, m->N[l] * sizeof(float), m->stream));
    GM_HIP(hipMemsetAsync(m->vb[l], 0, m->N[l] * sizeof(float), m->stream));
  }
  m->adam_t = 0;
  return GM_OK;
}

static void free_mlp(gm_mlp *m) {
  if (!m) return;
  std::vector<void *> ps;
  for (auto *v : {&m->W, &m->b, &m->mW, &m->vW, &m->mb, &m->vb, &m->gW, &m->gb})
    for (float *p : *v) ps.push_back(p);
  for (bf16_t *p : m->Wb) ps.push_back(p);
  for (bf16_t *p : m->WbT) ps.push_back(p);
  for (bf16_t *p : m->act) ps.push_back(p);
  void *more[] = {m->y,   m->dz[0], m->dz[1], m->part, m->partb, m->bt,  m->bm,  m->lpart,
                  m->lossbuf, m->din, m->dt,  m->dm,   m->dperm, m->gin, m->gout, m->gx, m->d_gpts,
                  m->Q,   m->wgt,  m->qpart, m->qc,  m->qv};
  for (void *p : more) ps.push_back(p);
  for (void *p : ps) dfree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

static int create_impl(const gm_mlp_shape *s, gm_mlp *m) {
  if (s->n_in < 1 || s->n_in > GM_MLP_MAX_IN || s->n_out < 1 || s->n_out > GM_MLP_MAX_OUT ||
      s->hidden < 1 || s->width < MT || s->width % MT) {
    gm_set_error("invalid network shape (n_in %d, n_out %d, hidden %d, width %d)", s->n_in,
                 s->n_out, s->hidden, s->width);
    return GM_ERR_ARG;
  }
  m->s = *s;
  m->L = s->hidden + 1;
  GM_HIP(hipGetDevice(&m->dev));
  int rc0 = scratch_init(m->dev);
  if (rc0) return rc0;
  for (int l = 0; l < m->L; l++) {
    m->Kr.push_back(l == 0 ? s->n_in : s->width);
    m->K.push_back(l == 0 ? IN_PAD : s->width);
    m->Nr.push_back(l == m->L - 1 ? s->n_out : s->width);
    m->N.push_back(l == m->L - 1 ? OUT_PAD : s->width);
  }
  // a blocking stream: ordered after work the caller queued on the default
  // stream (e.g. zeroing the arrays it passes in) and before what it queues later
  GM_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamDefault));
  int rc;
  uint64_t st = s->seed;
  for (int l = 0; l < m->L; l++) {
    const size_t NK = (size_t)m->N[l] * m->K[l];
    // registered before allocation so that free_mlp releases a partial set
    for (auto *v : {&m->W, &m->mW, &m->vW, &m->gW, &m->b, &m->mb, &m->vb, &m->gb}) v->push_back(nullptr);
    if ((rc = dalloc(&m->W[l], NK)) || (rc = dalloc(&m->mW[l], NK)) || (rc = dalloc(&m->vW[l], NK)) ||
        (rc = dalloc(&m->gW[l], NK)) || (rc = dalloc(&m->b[l], m->N[l])) ||
        (rc = dalloc(&m->mb[l], m->N[l])) || (rc = dalloc(&m->vb[l], m->N[l])) ||
        (rc = dalloc(&m->gb[l], m->N[l])))
      return rc;
    float *W = m->W[l], *b = m->b[l];
    m->Wb.push_back(nullptr);
    m->WbT.push_back(nullptr);
    if ((rc = dalloc(&m->Wb[l], NK)) || (rc = dalloc(&m->WbT[l], NK))) return rc;
    // Glorot uniform on the real block, zero padding and biases
    std::vector<float> hw(NK, 0.f);
    const double a = std::sqrt(6.0 / (m->Kr[l] + m->Nr[l]));
    for (int n = 0; n < m->Nr[l]; n++)
      for (int k = 0; k < m->Kr[l]; k++)
        hw[(size_t)n * m->K[l] + k] = (float)((2.0 * uniform01(st) - 1.0) * a);
    GM_HIP(hipMemcpy(W, hw.data(), NK * sizeof(float), hipMemcpyHostToDevice));
    GM_HIP(hipMemset(b, 0, m->N[l] * sizeof(float)));
  }
  if ((rc = dalloc(&m->lpart, LOSS_BLOCKS))) return rc;
  if ((rc = reset_adam(m))) return rc;
  if ((rc = refresh_bf16(m))) return rc;
  GM_HIP(hipStreamSynchronize(m->stream));
  return GM_OK;
}

extern "C" int gm_mlp_create(const gm_mlp_shape *shape, gm_mlp **out) {
  if (!shape || !out) return GM_ERR_ARG;
  *out = nullptr;
  gm_mlp *m = new gm_mlp();
  const int rc = create_impl(shape, m);
  if (rc) {
    free_mlp(m);
    return rc;
  }
  *out = m;
  return GM_OK;
}

extern "C" void gm_mlp_destroy(gm_mlp *m) { free_mlp(m); }

extern "C" long long gm_mlp_num_params(const gm_mlp *m) {
  if (!m) return GM_ERR_ARG;
  long long n = 0;
  for (int l = 0; l < m->L; l++) n += (long long)m->Nr[l] * m->Kr[l] + m->Nr[l];
  return n;
}

extern "C" int gm_mlp_get_params(gm_mlp *m, float *params) {
  if (!m || !params) return GM_ERR_ARG;
  GM_HIP(hipStreamSynchronize(m->stream));
  size_t off = 0;
  for (int l = 0; l < m->L; l++) {
    std::vector<float> hw((size_t)m->N[l] * m->K[l]), hb(m->N[l]);
    GM_HIP(hipMemcpy(hw.data(), m->W[l], hw.size() * sizeof(float), hipMemcpyDeviceToHost));
    GM_HIP(hipMemcpy(hb.data(), m->b[l], hb.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int n = 0; n < m->Nr[l]; n++)
      for (int k = 0; k < m->Kr[l]; k++) params[off++] = hw[(size_t)n * m->K[l] + k];
    for (int n = 0; n < m->Nr[l]; n++) params[off++] = hb[n];
  }
  return GM_OK;
}

extern "C" int gm_mlp_set_params(gm_mlp *m, const float *params) {
  if (!m || !params) return GM_ERR_ARG;
  GM_HIP(hipStreamSynchronize(m->stream));
  size_t off = 0;
  for (int l = 0; l < m->L; l++) {
    std::vector<float> hw((size_t)m->N[l] * m->K[l], 0.f), hb(m->N[l], 0.f);
    for (int n = 0; n < m->Nr[l]; n++)
      for (int k = 0; k < m->Kr[l]; k++) hw[(size_t)n * m->K[l] + k] = params[off++];
    for (int n = 0; n < m->Nr[l]; n++) hb[n] = params[off++];
    GM_HIP(hipMemcpy(m->W[l], hw.data(), hw.size() * sizeof(float), hipMemcpyHostToDevice));
    GM_HIP(hipMemcpy(m->b[l], hb.data(), hb.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  int rc;
  if ((rc = reset_adam(m)) || (rc = refresh_bf16(m))) return rc;
  GM_HIP(hipStreamSynchronize(m->stream));
  return GM_OK;
}

extern "C" void gm_mlp_default_train_opts(gm_mlp_train_opts *o) {
  o->epochs = 100;
  o->batch = 8192;
  o->lr = 2e-3;
  o->lr_final = 2e-5;
  o->beta1 = 0.9;
  o->beta2 = 0.999;
  o->eps = 1e-8;
  o->seed = 1;
  o->verbose = 0;
}

extern "C" int gm_mlp_set_data(gm_mlp *m, long long npts, const float *in, const float *targ,
                               const float *mask) {
  if (!m || npts <= 0 || !in || !targ || !mask || npts >= (1LL << 31)) return GM_ERR_ARG;
  dfree(m->din);
  dfree(m->dt);
  dfree(m->dm);
  dfree(m->dperm);
  m->din = m->dt = m->dm = nullptr;
  m->dperm = nullptr;
  int rc;
  if ((rc = dalloc(&m->din, (size_t)npts * m->s.n_in)) ||
      (rc = dalloc(&m->dt, (size_t)npts * m->s.n_out)) ||
      (rc = dalloc(&m->dm, (size_t)npts * m->s.n_out)) || (rc = dalloc(&m->dperm, (size_t)npts)))
    return rc;
  GM_HIP(hipMemcpy(m->din, in, (size_t)npts * m->s.n_in * sizeof(float), hipMemcpyHostToDevice));
  GM_HIP(hipMemcpy(m->dt, targ, (size_t)npts * m->s.n_out * sizeof(float), hipMemcpyHostToDevice));
  GM_HIP(hipMemcpy(m->dm, mask, (size_t)npts * m->s.n_out * sizeof(float), hipMemcpyHostToDevice));
  m->npts = npts;
  return GM_OK;
}

static int ensure_lossbuf(gm_mlp *m, long long n) {
  if (n <= m->lossbuf_n) return GM_OK;
  dfree(m->lossbuf);
  m->lossbuf = nullptr;
  int rc;
  if ((rc = dalloc(&m->lossbuf, (size_t)n))) return rc;
  m->lossbuf_n = n;
  return GM_OK;
}

extern "C" int gm_mlp_train(gm_mlp *m, const gm_mlp_train_opts *o, double *loss_per_epoch) {
  if (!m || !o || o->epochs < 0 || o->batch < PT || o->batch % PT) return GM_ERR_ARG;
  if (!m->npts) {
    gm_set_error("gm_mlp_train: no training data");
    return GM_ERR_ARG;
  }
  int rc;
  if ((rc = ensure_cap(m, o->batch))) return rc;
  const long long steps = (m->npts + o->batch - 1) / o->batch;
  if ((rc = ensure_lossbuf(m, steps))) return rc;
  const long long total = steps * o->epochs;
  std::vector<int> perm(m->npts);
  std::vector<double> hl(steps);
  uint64_t st = o->seed;
  long long step = 0;
  for (int ep = 0; ep < o->epochs; ep++) {
    for (long long i = 0; i < m->npts; i++) perm[i] = (int)i;
    for (long long i = m->npts - 1; i > 0; i--) {
      const long long j = (long long)(splitmix(st) % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    GM_HIP(hipMemcpyAsync(m->dperm, perm.data(), m->npts * sizeof(int), hipMemcpyHostToDevice,
                          m->stream));
    for (long long b = 0; b < steps; b++, step++) {
      const int M = (int)std::min<long long>(o->batch, m->npts - b * o->batch);
      const int Mpad = (M + PT - 1) / PT * PT;
      const double frac = total > 1 ? (double)step / (double)(total - 1) : 1.0;
      const double lr = o->lr_final + 0.5 * (o->lr - o->lr_final) * (1.0 + std::cos(M_PI * frac));
      gather(m, M, Mpad, m->dperm + b * o->batch, m->din, m->dt, m->dm);
      forward(m, Mpad);
      loss(m, M, Mpad, b);
      backward(m, Mpad);
      adam(m, lr, o);
    }
    GM_HIP(hipGetLastError());
    GM_HIP(hipMemcpyAsync(hl.data(), m->lossbuf, steps * sizeof(double), hipMemcpyDeviceToHost,
                          m->stream));
    GM_HIP(hipStreamSynchronize(m->stream));
    double mean = 0.0;
    for (double v : hl) mean += v;
    mean /= (double)steps;
    if (loss_per_epoch) loss_per_epoch[ep] = mean;
    if (!std::isfinite(mean)) {
      gm_set_error("gm_mlp_train: non-finite loss at epoch %d", ep);
      return GM_ERR_NAN;
    }
    if (o->verbose && (ep % 10 == 0 || ep == o->epochs - 1))
      fprintf(stderr, "[mlp] epoch %4d loss %.4e\n", ep, mean);
  }
  return GM_OK;
}

extern "C" int gm_mlp_loss_grad(gm_mlp *m, int npts, const float *in, const float *targ,
                                const float *mask, double *loss_out, float *grad) {
  if (!m || npts <= 0 || npts > 65536 || !in || !targ || !mask || !loss_out || !grad)
    return GM_ERR_ARG;
  const int Mpad = (npts + PT - 1) / PT * PT;
  int rc;
  if ((rc = ensure_cap(m, Mpad)) || (rc = ensure_lossbuf(m, 1))) return rc;
  float *din, *dt, *dm;
  if ((rc = dalloc(&din, (size_t)npts * m->s.n_in))) return rc;
  if ((rc = dalloc(&dt, (size_t)npts * m->s.n_out))) {
    dfree(din);
    return rc;
  }
  if ((rc = dalloc(&dm, (size_t)npts * m->s.n_out))) {
    dfree(din);
    dfree(dt);
    return rc;
  }
  auto done = [&](int r) {
    dfree(din);
    dfree(dt);
    dfree(dm);
    return r;
  };
  if (hipMemcpy(din, in, (size_t)npts * m->s.n_in * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dt, targ, (size_t)npts * m->s.n_out * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dm, mask, (size_t)npts * m->s.n_out * 4, hipMemcpyHostToDevice) != hipSuccess) {
    gm_set_error("gm_mlp_loss_grad: upload failed");
    return done(GM_ERR_HIP);
  }
  gather(m, npts, Mpad, nullptr, din, dt, dm);
  forward(m, Mpad);
  loss(m, npts, Mpad, 0);
  backward(m, Mpad);
  if (hipStreamSynchronize(m->stream) != hipSuccess) {
    gm_set_error("gm_mlp_loss_grad: %s", hipGetErrorString(hipGetLastError()));
    return done(GM_ERR_HIP);
  }
  (void)hipMemcpy(loss_out, m->lossbuf, sizeof(double), hipMemcpyDeviceToHost);
  size_t off = 0;
  for (int l = 0; l < m->L; l++) {
    std::vector<float> hw((size_t)m->N[l] * m->K[l]), hb(m->N[l]);
    (void)hipMemcpy(hw.data(), m->gW[l], hw.size() * 4, hipMemcpyDeviceToHost);
    (void)hipMemcpy(hb.data(), m->gb[l], hb.size() * 4, hipMemcpyDeviceToHost);
    for (int n = 0; n < m->Nr[l]; n++)
      for (int k = 0; k < m->Kr[l]; k++) grad[off++] = hw[(size_t)n * m->K[l] + k];
    for (int n = 0; n < m->Nr[l]; n++) grad[off++] = hb[n];
  }
  return done(GM_OK);
}

extern "C" int gm_mlp_predict_dev(gm_mlp *m, long long npts, const float *in_dev, float *out_dev) {
  if (!m || npts < 0 || (npts && (!in_dev || !out_dev))) return GM_ERR_ARG;
  int rc;
  if ((rc = ensure_cap(m, std::min<long long>(std::max<long long>(npts, 1), 65536)))) return rc;
  for (long long off = 0; off < npts; off += m->cap) {
    const int M = (int)std::min<long long>(m->cap, npts - off);
    const int Mpad = (M + PT - 1) / PT * PT;
    gather(m, M, Mpad, nullptr, in_dev + off * m->s.n_in, nullptr, nullptr);
    forward(m, Mpad);
    k_mlp_copy_out<<<(M * m->s.n_out + 255) / 256, 256, 0, m->stream>>>(M, m->s.n_out, m->y,
                                                                        out_dev + off * m->s.n_out);
  }
  GM_HIP(hipGetLastError());
  GM_HIP(hipStreamSynchronize(m->stream));
  return GM_OK;
}

extern "C" int gm_mlp_predict_noalloc_dev(gm_mlp *m, long long npts, const float *in_dev,
                                          float *out_dev) {
  if (!m || npts < 0 || (npts && (!in_dev || !out_dev))) return GM_ERR_ARG;
  if (maxdim(m) > NA_W) {
    gm_set_error("gm_mlp_predict_noalloc_dev: layer width %d exceeds the scratch's %d features",
                 maxdim(m), NA_W);
    return GM_ERR_ARG;
  }
  void *base = (m->dev >= 0 && m->dev < NA_MAX_DEV) ? g_scratch_ptr[m->dev] : nullptr;
  if (!base) {
    gm_set_error("gm_mlp_predict_noalloc_dev: no scratch address for device %d", m->dev);
    return GM_ERR_HIP;
  }
  bf16_t *buf[2] = {(bf16_t *)base, (bf16_t *)((char *)base + NA_ACT)};
  float *y = (float *)((char *)base + 2 * NA_ACT);
  for (long long off = 0; off < npts; off += NA_PTS) {
    const int M = (int)std::min<long long>(NA_PTS, npts - off);
    const int Mpad = (M + PT - 1) / PT * PT;
    const long long total = (long long)Mpad * IN_PAD;
    const int blocks = (int)std::min<long long>((total + 255) / 256, 4096);
    k_mlp_gather<<<blocks, 256, 0, m->stream>>>(M, Mpad, nullptr, in_dev + off * m->s.n_in, nullptr,
                                                nullptr, m->s.n_in, m->s.n_out, in_norm(m), buf[0],
                                                nullptr, nullptr);
    for (int l = 0; l < m->L; l++) {
      if (l < m->L - 1)
        gemm<EPI_TANH>(m->stream, Mpad, m->N[l], m->K[l], buf[l & 1], m->Wb[l], m->b[l], nullptr,
                       buf[(l + 1) & 1], nullptr);
      else
        gemm<EPI_LIN>(m->stream, Mpad, m->N[l], m->K[l], buf[l & 1], m->Wb[l], m->b[l], nullptr,
                      nullptr, y);
    }
    k_mlp_copy_out<<<(M * m->s.n_out + 255) / 256, 256, 0, m->stream>>>(M, m->s.n_out, y,
                                                                        out_dev + off * m->s.n_out);
  }
  GM_HIP(hipGetLastError());
  GM_HIP(hipStreamSynchronize(m->stream));
  return GM_OK;
}

extern "C" int gm_mlp_predict(gm_mlp *m, long long npts, const float *in, float *out) {
  if (!m || npts < 0 || (npts && (!in || !out))) return GM_ERR_ARG;
  if (!npts) return GM_OK;
  float *di, *dout;
  int rc;
  if ((rc = dalloc(&di, (size_t)npts * m->s.n_in))) return rc;
  if ((rc = dalloc(&dout, (size_t)npts * m->s.n_out))) {
    dfree(di);
    return rc;
  }
  rc = GM_OK;
  if (hipMemcpy(di, in, (size_t)npts * m->s.n_in * 4, hipMemcpyHostToDevice) != hipSuccess) {
    gm_set_error("gm_mlp_predict: upload failed");
    rc = GM_ERR_HIP;
  }
  if (!rc) rc = gm_mlp_predict_dev(m, npts, di, dout);
  if (!rc && hipMemcpy(out, dout, (size_t)npts * m->s.n_out * 4, hipMemcpyDeviceToHost) != hipSuccess) {
    gm_set_error("gm_mlp_predict: download failed");
    rc = GM_ERR_HIP;
  }
  dfree(di);
  dfree(dout);
  return rc;
}

// Sample points of a context: the mesh nodes, then per element the centroid
// and the points at +-1/2 along each reference axis (Q2 map of the element).
static void sample_points(const gm_ctx *c, std::vector<float> &pts) {
  const int dim = c->dim, spe = 1 + 2 * dim, nve = c->nve;
  pts.assign(((size_t)c->num_nodes + (size_t)c->num_elems * spe) * dim, 0.f);
  for (int n = 0; n < c->num_nodes; n++)
    for (int d = 0; d < dim; d++) pts[(size_t)n * dim + d] = (float)c->coor[(size_t)d * c->num_nodes + n];
  std::vector<std::vector<double>> phi(spe, std::vector<double>(nve));
  for (int j = 0; j < spe; j++) {
    double xi[3] = {0.0, 0.0, 0.0};
    if (j > 0) xi[(j - 1) / 2] = (j % 2) ? 0.5 : -0.5;
    gm_q2_basis(dim, xi, phi[j].data());
  }
  for (int e = 0; e < c->num_elems; e++) {
    const int *cn = c->conn.data() + (size_t)e * nve;
    for (int j = 0; j < spe; j++) {
      const size_t row = (size_t)c->num_nodes + (size_t)e * spe + j;
      for (int d = 0; d < dim; d++) {
        double v = 0.0;
        for (int k = 0; k < nve; k++) v += phi[j][k] * c->coor[(size_t)d * c->num_nodes + cn[k]];
        pts[row * dim + d] = (float)v;
      }
    }
  }
}

// c[k] = Q[k] . v for k < nb, reduced in a fixed order on the host
static int qdots(gm_mlp *m, int N, int nb, const double *v, std::vector<double> &c) {
  c.assign(nb, 0.0);
  if (!nb) return GM_OK;
  k_mlp_qdots<<<QBLK, 256, 0, m->stream>>>(N, nb, m->Q, v, m->qpart);
  GM_HIP(hipGetLastError());
  std::vector<double> part((size_t)nb * QBLK);
  GM_HIP(hipMemcpyAsync(part.data(), m->qpart, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                        m->stream));
  GM_HIP(hipStreamSynchronize(m->stream));
  for (int k = 0; k < nb; k++)
    for (int b = 0; b < QBLK; b++) c[k] += part[(size_t)k * QBLK + b];
  return GM_OK;
}

extern "C" void gm_mlp_clear_basis(gm_mlp *m) {
  if (!m) return;
  m->bctx = nullptr;
  m->nbasis = 0;
}

extern "C" int gm_mlp_set_basis(gm_mlp *m, gm_ctx *c, int nsnap, const double *snaps,
                                const double *weight) {
  if (!m || !c || nsnap < 2 || !snaps || !weight) return GM_ERR_ARG;
  if (nsnap - 1 > MAX_BASIS) {
    gm_set_error("at most %d snapshot differences in the subspace", MAX_BASIS);
    return GM_ERR_ARG;
  }
  const int N = c->N;
  gm_mlp_clear_basis(m);
  for (void *p : {(void *)m->Q, (void *)m->wgt, (void *)m->qpart, (void *)m->qc, (void *)m->qv}) dfree(p);
  m->Q = m->wgt = m->qpart = m->qc = m->qv = nullptr;
  int rc;
  if ((rc = dalloc(&m->Q, (size_t)(nsnap - 1) * N)) || (rc = dalloc(&m->wgt, (size_t)N)) ||
      (rc = dalloc(&m->qpart, (size_t)MAX_BASIS * QBLK)) || (rc = dalloc(&m->qc, (size_t)MAX_BASIS)) ||
      (rc = dalloc(&m->qv, (size_t)N)))
    return rc;
  GM_HIP(hipStreamSynchronize(c->stream));
  GM_HIP(hipMemcpyAsync(m->wgt, weight, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice,
                        m->stream));
  const int g = (N + 255) / 256;
  int nb = 0;
  std::vector<double> cf;
  for (int k = 1; k < nsnap; k++) {
    double *v = m->Q + (size_t)nb * N;  // candidate written in place
    k_mlp_wdiff<<<g, 256, 0, m->stream>>>(N, snaps + (size_t)k * N, snaps, m->wgt, v);
    double n0 = 0.0;
    // two classical Gram-Schmidt passes against the accepted vectors
    for (int pass = 0; pass < 2; pass++) {
      if ((rc = qdots(m, N, nb, v, cf))) return rc;
      if (nb) {
        GM_HIP(hipMemcpyAsync(m->qc, cf.data(), nb * sizeof(double), hipMemcpyHostToDevice, m->stream));
        k_mlp_qsub<<<g, 256, 0, m->stream>>>(N, nb, m->Q, m->qc, v);
      }
      if (pass == 0)  // |v|^2 before = |v|^2 after the first pass + sum c^2
        for (double x : cf) n0 += x * x;
    }
    // norm of the orthogonalised vector: one dot with itself
    {
      k_mlp_qdots<<<QBLK, 256, 0, m->stream>>>(N, 1, v, v, m->qpart);
      std::vector<double> part(QBLK);
      GM_HIP(hipMemcpyAsync(part.data(), m->qpart, QBLK * sizeof(double), hipMemcpyDeviceToHost,
                            m->stream));
      GM_HIP(hipStreamSynchronize(m->stream));
      double s = 0.0;
      for (double x : part) s += x;
      const double nrm = std::sqrt(s);
      if (nrm > 1e-10 * std::sqrt(n0 + s) && nrm > 0.0) {
        k_mlp_scale<<<g, 256, 0, m->stream>>>(N, v, 1.0 / nrm);
        nb++;
      }
    }
  }
  GM_HIP(hipGetLastError());
  GM_HIP(hipStreamSynchronize(m->stream));
  m->nbasis = nb;
  m->bctx = nb ? c : nullptr;
  return nb;
}

static int guess_one(gm_mlp *m, gm_ctx *c, const gm_mlp_query *q, double *x) {
  const int dim = c->dim, spe = 1 + 2 * dim, n_in = m->s.n_in;
  const long long np = (long long)c->num_nodes + (long long)c->num_elems * spe;
  GuessParams gp{};
  for (int i = dim; i < n_in; i++) gp.p[i - dim] = (float)q->param[i - dim];
  const long long nt = np * n_in;
  k_mlp_guess_input<<<(unsigned)((nt + 255) / 256), 256, 0, m->stream>>>(np, dim, n_in, m->d_gpts, gp,
                                                                         m->gin);
  GM_HIP(hipGetLastError());
  int rc = gm_mlp_predict_dev(m, np, m->gin, m->gout);
  if (rc) return rc;
  GuessMap g{};
  g.num_nodes = c->num_nodes;
  g.num_elems = c->num_elems;
  g.dim = dim;
  g.energy = c->energy;
  g.n_out = m->s.n_out;
  g.spe = spe;
  for (int o = 0; o < m->s.n_out; o++) {
    g.scale[o] = q->out_scale[o];
    g.shift[o] = q->out_shift[o];
  }
  k_mlp_guess_nodes<<<(c->num_nodes + 255) / 256, 256, 0, m->stream>>>(g, c->d_fu, m->gout, x);
  k_mlp_guess_elems<<<(c->num_elems + 255) / 256, 256, 0, m->stream>>>(g, c->d_fu, c->d_centroid,
                                                                       m->gout, x);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

extern "C" int gm_mlp_guess_dev(gm_mlp *m, gm_ctx *c, const gm_mlp_query *q, const gm_mlp_query *q_prev,
                                const double *x_prev_dev, double *x_dev) {
  if (!m || !c || !q || !x_dev || ((q_prev == nullptr) != (x_prev_dev == nullptr))) return GM_ERR_ARG;
  if (m->s.n_out != c->dim + 1 + c->energy || m->s.n_in <= c->dim) {
    gm_set_error("network outputs (%d) / inputs (%d) do not match a %dD problem with%s energy",
                 m->s.n_out, m->s.n_in, c->dim, c->energy ? "" : "out");
    return GM_ERR_ARG;
  }
  const int spe = 1 + 2 * c->dim;
  const long long np = (long long)c->num_nodes + (long long)c->num_elems * spe;
  int rc;
  if (m->gctx != c) {
    sample_points(c, m->gpts);
    dfree(m->d_gpts);
    m->d_gpts = nullptr;
    if ((rc = dalloc(&m->d_gpts, m->gpts.size()))) return rc;
    GM_HIP(hipMemcpy(m->d_gpts, m->gpts.data(), m->gpts.size() * sizeof(float), hipMemcpyHostToDevice));
    m->gctx = c;
  }
  if (np > m->gcap) {
    dfree(m->gin);
    dfree(m->gout);
    m->gin = m->gout = nullptr;
    if ((rc = dalloc(&m->gin, (size_t)np * m->s.n_in)) || (rc = dalloc(&m->gout, (size_t)np * m->s.n_out)))
      return rc;
    m->gcap = np;
  }
  // the context's stream may hold work on x_prev: order the two streams
  GM_HIP(hipStreamSynchronize(c->stream));
  if ((rc = guess_one(m, c, q, x_dev))) return rc;
  const int N = c->N;
  if (q_prev) {
    if (N > m->gx_n) {
      dfree(m->gx);
      m->gx = nullptr;
      if ((rc = dalloc(&m->gx, (size_t)N))) return rc;
      m->gx_n = N;
    }
    if ((rc = guess_one(m, c, q_prev, m->gx))) return rc;
    if (m->bctx == c && m->nbasis > 0) {
      // change = G(q) - G(q_prev), projected onto the snapshot subspace
      k_mlp_wdiff<<<(N + 255) / 256, 256, 0, m->stream>>>(N, x_dev, m->gx, m->wgt, m->qv);
      std::vector<double> cf;
      if ((rc = qdots(m, N, m->nbasis, m->qv, cf))) return rc;
      GM_HIP(hipMemcpyAsync(m->qc, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice,
                            m->stream));
      k_mlp_qapply<<<(N + 255) / 256, 256, 0, m->stream>>>(N, m->nbasis, m->Q, m->qc, m->wgt,
                                                           x_prev_dev, x_dev);
    } else {
      k_mlp_delta<<<(N + 255) / 256, 256, 0, m->stream>>>(N, x_prev_dev, m->gx, x_dev);
    }
  }
  k_mlp_dirichlet<<<(N + 255) / 256, 256, 0, m->stream>>>(N, c->d_dbc, c->d_dbc_val, x_dev);
  GM_HIP(hipGetLastError());
  GM_HIP(hipStreamSynchronize(m->stream));
  return GM_OK;
}
