// Internal definitions shared by the libgm translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gm_capi.h"

#define GM_QS_MAX 5 /* QSIDE / QCONV side entries per element */
#define GM_MAXROW 1024 /* longest Jacobian row (HEX27 vertex row: 125 nodes x 4 + 8 x 4 P = 532) */

void gm_set_error(const char *fmt, ...);

#define GM_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      gm_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return GM_ERR_HIP;                                                             \
    }                                                                                \
  } while (0)

// Element-to-nonzero map entry layout for one element (see gm_fill.hip):
//   [nv*nv blocks][nve*nve node pairs]   nodal var A x nodal var B
//   [dim][nve*np]                        momentum a  x pressure dof  (J_m_P)
//   [dim][np*nve]                        continuity x velocity b     (J_c_v)
//   [np*np]                              continuity x pressure       (J_c_P)
struct ElemLayout {
  int dim, nve, np, nv;  // nv = dim + energy nodal variables
  int nq;
  long long off_mp, off_cv, off_cp, total;
};

struct IluPlan {
  int bins = -1;
  int ncol = 0, npatch = 0;
  std::vector<int> perm;         // perm[new] = old
  std::vector<int> colour_ptr;   // patches per colour: patch index ranges
  std::vector<int> patch_row;    // [npatch+1] first new row of each patch
  long long nnz = 0;
  int *d_perm = nullptr;      // [N]
  int *d_rowptr = nullptr;    // [N+1]
  int *d_col = nullptr;       // [nnz] new column ids, ascending per row
  int *d_diag = nullptr;      // [N] position of the diagonal
  int *d_src = nullptr;       // [nnz] MSR index of each entry
  int *d_patch_row = nullptr; // [npatch+1]
  double *d_lu = nullptr;     // [nnz]
  float *d_lu32 = nullptr;    // [nnz] (mixed precision)
  double *d_y = nullptr;      // [N] work (permuted numbering)
  float *d_y32 = nullptr;
  int *d_node_first = nullptr, *d_node_last = nullptr;  // [N] node row range of each row
  int levels = 0;             // total forward levels over all patches
  std::vector<long long> col_rows, col_nnzL, col_nnzU;  // per patch colour
  // per-patch level schedules, [0] forward / factorisation, [1] backward
  int *d_lv_ptr[2] = {nullptr, nullptr}, *d_lv_nstart[2] = {nullptr, nullptr};
  int *d_lv_node[2] = {nullptr, nullptr}, *d_lv_rstart[2] = {nullptr, nullptr};
  int *d_lv_row[2] = {nullptr, nullptr};
};

struct KrylovWork {
  int m = 0;
  int prec = -1;
  double *d_V = nullptr;   // [(m+1) * N]
  double *d_w = nullptr;   // [N]
  double *d_z = nullptr;   // [N]
  double *d_part = nullptr;  // partial dot products [(m+2) * nblk]
  double *d_h = nullptr;   // [m+2]
  float *d_V32 = nullptr;
  float *d_w32 = nullptr, *d_z32 = nullptr;
};

// banded LU (gm_band.hip)
struct BandLU {
  int kl = -1, ku = -1;       // lower / upper bandwidth of the MSR graph
  long long ld = 0;           // entries per stored column
  int prec = -1;              // precision of the last factorisation (0 FP64, 1 FP32)
  void *d_ab[2] = {nullptr, nullptr};  // [N][ld] per precision
  double *d_w1 = nullptr, *d_w2 = nullptr;  // solve work vectors
};

struct gm_ctx;
struct MgLevel {
  gm_ctx *ctx = nullptr;  // the level's context (not owned)
  // transfer between this level and the next finer one: P (fine x coarse), R = P^T
  int *d_Prow = nullptr, *d_Pcol = nullptr, *d_Rrow = nullptr, *d_Rcol = nullptr;
  double *d_Pval = nullptr, *d_Rval = nullptr;
  int *d_inject = nullptr;    // coarse node -> finer node at the same position
  // element-local form of P, for Galerkin coarse operators: the finer elements
  // inside each coarse element, and per finer element the nodal weights
  // W[i][m] (finer local node i <- coarse local node m) and P1 weights Q[k][l]
  int *d_child = nullptr;     // [coarse elems][2^dim]
  // nodal weight matrix W of each finer element by columns (coarse local
  // node m: the finer local nodes interpolating from it, ascending, and
  // their weights): pointers [finer elems][nve + 1], rows / weights
  // [finer elems][W_stride]
  int *d_Wp = nullptr, *d_Wi = nullptr;
  double *d_Wv = nullptr;
  int W_stride = 0;
  double *d_Q = nullptr;      // [finer elems][np * np]
  bool gal_ok = false;
  std::string gal_err;
  double *d_state = nullptr;  // injected state for re-assembly
  double *d_rdummy = nullptr; // residual written by the coarse assembly (unused)
  void *d_x = nullptr, *d_b = nullptr, *d_r = nullptr;  // level vectors
  void *d_val = nullptr;      // coarse Jacobian (MSR), in the active precision
  void *d_ainv = nullptr;     // element-block inverses, in the active precision
  void *d_val_p[2] = {nullptr, nullptr};   // storage per precision (FP64, mixed): kept, so
  void *d_ainv_p[2] = {nullptr, nullptr};  // switching precision never reallocates
  int vec_prec = -1;
  int nnz = 0;                // entries of P
  // coarsest level only: dense FP64 inverse for the exact coarse solve
  double *d_dense = nullptr;  // [n][n]
  double *d_pcol = nullptr;  // pivot rows of one panel of the coarsest inverse [GJ_NB][n]
  int *d_perm = nullptr;      // row interchange of each step
  int dense_n = 0;
  // host copy of P, kept by gm_mg_build (gm_mg_copy_prolongation)
  std::vector<int> P_host_rowptr, P_host_col;
  std::vector<double> P_host_val;
};

#define GM_MG_DIRECT_MAX 4096 /* largest coarsest level solved exactly */
#define GM_NODE_CAP_MAX 2048 /* longest node pattern the per-node row kernels take (LDS) */

struct gm_ctx {
  int dim = 0, nve = 0, np = 0, nv = 0, num_nodes = 0, num_elems = 0, energy = 0;
  gm_physics phys{};
  double h_elem_avg = 0.0;  // PSPG element size (global_h_elem_siz), from the mesh
  double U_norm = 0.0;      // PSPG velocity norm (global_velocity_norm) the fills use
  int N = 0;
  long long nval = 0;  // bindx[N]
  std::vector<double> coor;
  std::vector<int> conn;
  std::vector<int> first_unknown, unk_var, elem_centroid, is_centroid;
  std::vector<signed char> dbc;
  std::vector<double> dbc_val;
  std::vector<int> nn_ptr, nn_list;
  std::vector<int> bindx;
  int n_ecol = 0;
  std::vector<int> ecol_ptr, ecol_elems, elem_colour;
  ElemLayout lay{};

  hipStream_t stream = nullptr;
  // multigrid set-up: the element-block inversions' stream (kept off a few
  // CUs; equal to stream when that mask is unavailable) and its two events
  hipStream_t s_inv = nullptr;
  hipEvent_t ev_inv[2] = {nullptr, nullptr};
  double *d_coor = nullptr;
  int *d_conn = nullptr;
  int *d_fu = nullptr;
  int *d_centroid = nullptr;
  signed char *d_dbc = nullptr;
  double *d_dbc_val = nullptr;
  int *d_bindx = nullptr;
  int *d_nn_ptr = nullptr, *d_nn_list = nullptr;  // node-node graph (the row kernels' runs)
  int nn_cap = 0;  // longest column pattern of a node (sum of its neighbours' unknowns)
  int *d_ecol_elems = nullptr;
  int *d_e2nz = nullptr;        // [num_elems * lay.total]
  int *d_flag = nullptr;        // device error flags
  double *d_tables = nullptr;   // reference basis tables
  double *d_stab = nullptr;     // side basis tables (surface integrals)
  unsigned short *d_rank = nullptr;  // HEX27: per element and node pair, the rank of the
                                     // column node in the row node's rows (+ first-touch bit)
  int fill_overwrite = 0;       // the next assembly stores first touches instead of adding

  // theta-method state of the next assemblies (tr_x_old == nullptr: steady):
  // xdot = tr_cmass (x - x_old) - tr_cold xdot_old
  const double *tr_x_old = nullptr, *tr_xdot_old = nullptr;
  double tr_cmass = 0.0, tr_cold = 0.0;
  // QSIDE weak heat-flux entries (element, 1-based side, flux), host copy in
  // card order; on the device grouped per element: d_qs_ptr [num_elems+1]
  std::vector<int> qs_elem, qs_side;
  std::vector<double> qs_q, qs_h, qs_tc;  // flux; QCONV coefficient and bath temperature
  int *d_qs_ptr = nullptr, *d_qs_side = nullptr;
  double *d_qs_q = nullptr;  // (q, h, T_c) per entry

  // Newton work vectors
  double *d_val = nullptr, *d_resid = nullptr, *d_dx = nullptr, *d_scale = nullptr;
  double *d_kel = nullptr;  // element matrices before Dirichlet rows [num_elems][lay.total]
  int kel_on = 0;           // the next assembly also writes d_kel
  float *d_val32 = nullptr;
  double *d_red = nullptr;      // reduction scratch
  double *h_red = nullptr;      // pinned host mirror
  double *d_xwork = nullptr;
  double *d_xsave = nullptr, *d_rtrial = nullptr, *d_w = nullptr;  // line search work

  IluPlan ilu;
  KrylovWork kry;
  BandLU band;
  std::vector<MgLevel> mg;  // coarse levels, finest first
  std::vector<gm_ctx *> owned_levels;  // level contexts made by gm_mg_build (freed with this one)
  MgLevel mg0;              // level-0 vectors and element inverses
  const double *mg_scale = nullptr;  // row-sum scale of the level-0 matrix
  int *d_mg_rows = nullptr;          // unknowns of each element colour (smoother residuals)
  double *d_mg_wmul = nullptr;       // 1 / elements sharing each unknown (additive smoother)
  int *d_mg_gptr = nullptr;          // [N+1] additive smoother: corrections of each unknown,
  int *d_mg_gent = nullptr;          //   as indices e * n_loc + local dof into d_mg_corr
  double *d_mg_corr = nullptr;       // [num_elems][n_loc] element corrections
  std::vector<int> mg_rows_ptr;
  std::vector<long long> mg_rows_nnz;

  bool prof = false;
  long long prof_count[GM_K_COUNT] = {0};
  double prof_ms[GM_K_COUNT] = {0};
  double prof_bytes[GM_K_COUNT] = {0};
  double prof_sampled_bytes[GM_K_COUNT] = {0};
  long long prof_seen[GM_K_COUNT] = {0};
  int prof_every = 1;
  std::vector<hipEvent_t> ev_b[GM_K_COUNT], ev_e[GM_K_COUNT];
  int ev_used[GM_K_COUNT] = {0};
};

// set-up (gm_setup.cpp)
int gm_build_host(gm_ctx *c);
int gm_patch_order(const gm_ctx *c, int bins, std::vector<int> &perm, std::vector<int> &colour_ptr,
                   std::vector<int> &patch_row);
int gm_default_bins(const gm_ctx *c);

// device set-up and kernels (gm_fill.hip)
int gm_fill_setup_device(gm_ctx *c);
int gm_fill_launch(gm_ctx *c, const double *x, double *resid, double *val, float *val32);
void gm_q2_basis(int dim, const double *xi, double *phi);  // host: Q2 basis at one reference point
int gm_fill_launch_ex(gm_ctx *c, const double *x, double *resid, double *val, float *val32,
                      int resid_only);

// linear algebra (gm_linalg.hip)
int gm_zero(gm_ctx *c, double *p, long long n);
int gm_zero32(gm_ctx *c, float *p, long long n);
int gm_scale_launch(gm_ctx *c, double *val, float *val32, double *b, double *scale);
int gm_norms(gm_ctx *c, const double *v, long long n, double *l2, double *loo);
// PSPG: mean square of the velocity unknowns of x (device), the mesh's mean
// element size (host) and the global tau at c->U_norm
int gm_velocity_norm(gm_ctx *c, const double *x, double *U);
double gm_h_elem_size(const gm_ctx *c);
double gm_pspg_tau(const gm_ctx *c);
int gm_corr_norms(gm_ctx *c, const double *dx, const double *x, double *l2r, double *loor);
int gm_axpy_update(gm_ctx *c, double *x, const double *dx, double damp);
int gm_spmv_launch(gm_ctx *c, const double *val, const double *x, double *y);
int gm_spmv32_launch(gm_ctx *c, const float *val, const double *x, double *y);
int gm_lincomb(gm_ctx *c, double *out, const double *a, double s, const double *b);  // out = a - s b
int gm_div_scale(gm_ctx *c, double *v, const double *scale);                        // v /= scale
int gm_lin3(gm_ctx *c, double *out, double w0, const double *v0, double w1, const double *v1,
            double w2, const double *v2);  // out = w0 v0 + w1 v1 + w2 v2 (v1, v2 may be null)
int gm_dot(gm_ctx *c, const double *a, const double *b, double *out);
int gm_ilu_plan(gm_ctx *c, int bins);
int gm_ilu_factor(gm_ctx *c, const double *val, const float *val32, int prec);
int gm_gmres_run(gm_ctx *c, const double *val, const float *val32, const double *b, double *x,
                 const gm_solver_opts *o, int *its, double *relres);
int gm_gmres_run_nofactor(gm_ctx *c, const double *val, const float *val32, const double *b,
                          double *x, const gm_solver_opts *o, int *its, double *relres);
void gm_free_linalg(gm_ctx *c);

// multigrid (gm_mg.hip)
int gm_mg_setup(gm_ctx *c, const double *x, const double *val, const float *val32, int prec,
                const double *scale, int coarse_direct, int galerkin);
int gm_mg_kel_alloc(gm_ctx *c);  // element-matrix buffer of the finest level
int gm_mg_apply_d(gm_ctx *c, const double *a0, const double *v, double *z, const gm_solver_opts *o);
int gm_mg_apply_f(gm_ctx *c, const float *a0, const double *v, double *z, const gm_solver_opts *o);
void gm_mg_free(gm_ctx *c);

// banded LU (gm_band.hip)
int gm_band_factor(gm_ctx *c, const double *val, const float *val32, int prec);
int gm_band_apply(gm_ctx *c, const double *in, double *out);
void gm_band_free(gm_ctx *c);

// timing helper (used through gm_launch in gm_launch.h): adds `bytes` to the
// family's total; returns true, with the event pair that the launch must
// stamp, when this launch is one of the sampled ones.
bool gm_prof_take(gm_ctx *c, int id, double bytes, hipEvent_t *b, hipEvent_t *e);
int gm_ilu_factor_export(gm_ctx *c, const double *val, int bins, int *rowptr, int *col, double *lu);
