/* CPU restatement of the learned-guess MLP (test infrastructure; see
 * gm_mlp_oracle.h).  Straightforward loops in double; one point at a time. */
#include "gm_mlp_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

long long omlp_num_params(const omlp_shape *s) {
  long long n = 0;
  int k = s->n_in;
  for (int l = 0; l <= s->hidden; l++) {
    const int nn = l == s->hidden ? s->n_out : s->width;
    n += (long long)nn * k + nn;
    k = nn;
  }
  return n;
}

float omlp_bf16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) { /* inf / nan: truncate, keep nan quiet */
    if (u & 0x007fffffu) u |= 0x00400000u;
    u &= 0xffff0000u;
  } else {
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
  }
  float r;
  memcpy(&r, &u, 4);
  return r;
}

static double q(double v, int quant) { return quant ? (double)omlp_bf16((float)v) : v; }

static int check(const omlp_shape *s) {
  return s && s->n_in > 0 && s->n_out > 0 && s->hidden >= 1 && s->width > 0;
}

/* layer dims and parameter offsets */
static void layer(const omlp_shape *s, int l, int *K, int *N, long long *woff, long long *boff) {
  long long off = 0;
  int k = s->n_in;
  for (int i = 0; i <= l; i++) {
    const int n = i == s->hidden ? s->n_out : s->width;
    if (i == l) {
      *K = k;
      *N = n;
      *woff = off;
      *boff = off + (long long)n * k;
      return;
    }
    off += (long long)n * k + n;
    k = n;
  }
}

/* forward for one point; acts[l] holds the (rounded) input of layer l */
static void fwd_point(const omlp_shape *s, const float *p, const double *sh, const double *sc,
                      int quant, const float *in, double **acts, double *y) {
  for (int i = 0; i < s->n_in; i++) {
    double u;
    if (quant) u = (double)((in[i] - (float)sh[i]) * (float)sc[i]);
    else u = ((double)in[i] - sh[i]) * sc[i];
    acts[0][i] = q(u, quant);
  }
  for (int l = 0; l <= s->hidden; l++) {
    int K = 0, N = 0;
    long long wo = 0, bo = 0;
    layer(s, l, &K, &N, &wo, &bo);
    for (int n = 0; n < N; n++) {
      double z = 0.0;
      for (int k = 0; k < K; k++) z += q(p[wo + (long long)n * K + k], quant) * acts[l][k];
      z += p[bo + n];
      if (l < s->hidden) acts[l + 1][n] = q(tanh(z), quant);
      else y[n] = z;
    }
  }
}

static double **alloc_acts(const omlp_shape *s) {
  double **a = (double **)calloc((size_t)s->hidden + 1, sizeof(double *));
  if (!a) return NULL;
  for (int l = 0; l <= s->hidden; l++) {
    a[l] = (double *)calloc((size_t)(l == 0 ? s->n_in : s->width), sizeof(double));
    if (!a[l]) return NULL;
  }
  return a;
}

static void free_acts(const omlp_shape *s, double **a) {
  if (!a) return;
  for (int l = 0; l <= s->hidden; l++) free(a[l]);
  free(a);
}

int omlp_forward(const omlp_shape *s, const float *params, const double *in_shift,
                 const double *in_scale, int quant, long long npts, const float *in, double *y) {
  if (!check(s)) return -2;
  double **acts = alloc_acts(s);
  if (!acts) return -3;
  for (long long pt = 0; pt < npts; pt++)
    fwd_point(s, params, in_shift, in_scale, quant, in + pt * s->n_in, acts, y + pt * s->n_out);
  free_acts(s, acts);
  return 0;
}

int omlp_loss_grad(const omlp_shape *s, const float *params, const double *in_shift,
                   const double *in_scale, int quant, long long npts, const float *in,
                   const float *targ, const float *mask, double *loss, double *grad) {
  if (!check(s) || npts <= 0) return -2;
  const long long np = omlp_num_params(s);
  memset(grad, 0, (size_t)np * sizeof(double));
  double **acts = alloc_acts(s);
  double *y = (double *)calloc((size_t)s->n_out, sizeof(double));
  const int wmax = s->width > s->n_out ? s->width : s->n_out;
  double *dz = (double *)calloc((size_t)(wmax > s->n_in ? wmax : s->n_in), sizeof(double));
  double *da = (double *)calloc((size_t)(wmax > s->n_in ? wmax : s->n_in), sizeof(double));
  if (!acts || !y || !dz || !da) return -3;
  const double norm = 1.0 / ((double)npts * s->n_out);
  double L = 0.0;
  for (long long pt = 0; pt < npts; pt++) {
    fwd_point(s, params, in_shift, in_scale, quant, in + pt * s->n_in, acts, y);
    const float *t = targ + pt * s->n_out, *m = mask + pt * s->n_out;
    for (int o = 0; o < s->n_out; o++) {
      const double e = y[o] - (double)t[o];
      L += (double)m[o] * e * e;
      /* the GPU forms the delta in float before rounding it */
      dz[o] = quant ? q((double)(float)(2.0 * m[o] * e * norm), 1) : 2.0 * m[o] * e * norm;
    }
    for (int l = s->hidden; l >= 0; l--) {
      int K = 0, N = 0;
      long long wo = 0, bo = 0;
      layer(s, l, &K, &N, &wo, &bo);
      for (int n = 0; n < N; n++) {
        grad[bo + n] += dz[n];
        for (int k = 0; k < K; k++) grad[wo + (long long)n * K + k] += dz[n] * acts[l][k];
      }
      if (l == 0) break;
      for (int k = 0; k < K; k++) {
        double sum = 0.0;
        for (int n = 0; n < N; n++) sum += dz[n] * q(params[wo + (long long)n * K + k], quant);
        const double a = acts[l][k];
        da[k] = sum * (1.0 - a * a);
      }
      for (int k = 0; k < K; k++) dz[k] = q(da[k], quant);
    }
  }
  *loss = L * norm;
  free_acts(s, acts);
  free(y);
  free(dz);
  free(da);
  return 0;
}
