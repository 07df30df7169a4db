/*
 * gm_mlp_oracle.h -- CPU restatement of the learned-initial-guess network.
 *
 * TEST INFRASTRUCTURE ONLY.  Like gm_oracle.h, this is the checker: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product (goma_amd/, libgm.so) never does.
 *
 * The network has no counterpart in the reference (Goma's continuation,
 * src/ac_conti.c:760-787, only uses the previous converged solution as the
 * next guess).  It is a plain MLP:
 *     u0 = (raw - in_shift) * in_scale                      (per input)
 *     a_{l+1} = tanh(W_l a_l + b_l),  l = 0 .. hidden-1
 *     y = W_hidden a_hidden + b_hidden                      (linear output)
 *     L = sum_{p,o} mask_po (y_po - t_po)^2 / (npts * n_out)
 * Parameters are stored unpadded, layer by layer: W_l [N_l][K_l] row-major,
 * then b_l [N_l].
 *
 * quant = 1 reproduces the rounding points of libgm's BF16 kernels: the
 * normalised inputs, the weights, every hidden activation and every
 * back-propagated delta are rounded to bfloat16 (round to nearest even), the
 * input normalisation is evaluated in float; sums run in double here and in
 * float on the GPU.  quant = 0 is pure double (used to pin the gradients by
 * finite differences).
 */
#ifndef GM_MLP_ORACLE_H
#define GM_MLP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct omlp_shape {
  int n_in, n_out, hidden, width;
} omlp_shape;

long long omlp_num_params(const omlp_shape *s);

/* bfloat16 rounding (nearest even) of a float, returned as float */
float omlp_bf16(float x);

/* y [npts][n_out] */
int omlp_forward(const omlp_shape *s, const float *params, const double *in_shift,
                 const double *in_scale, int quant, long long npts, const float *in, double *y);

/* loss and dL/dparams (same layout as params) */
int omlp_loss_grad(const omlp_shape *s, const float *params, const double *in_shift,
                   const double *in_scale, int quant, long long npts, const float *in,
                   const float *targ, const float *mask, double *loss, double *grad);

#ifdef __cplusplus
}
#endif
#endif
