// Element assembly (matrix_fill_full) for QUAD9 / HEX27 Q2-P1 Navier-Stokes
// with optional energy, as HIP kernels for gfx950.
//
// One workgroup assembles one element.  Elements are launched one colour at a
// time (no two elements of a colour share a node), so every element owns the
// nonzeros it scatters into during its launch and the scatter is a plain
// read-modify-write: deterministic, no atomics.
//
// Per element:
//   1. nodal coordinates / unknowns -> LDS
//   2. per quadrature point: map Jacobian, inverse, dA = detJ*w  (beer_belly)
//   3. physical basis gradients grad_phi[q][k][p]                (load_bf_grad)
//   4. per quadrature point: v, grad v, T, grad T, P, mu(gammadot), body force
//   5. residual rows, then the Jacobian: thread t owns node pair (i,j) and
//      accumulates every (row variable, column variable) block of that pair;
//      pressure couplings are separate task lists.
//   6. Dirichlet rows: J row -> unit diagonal, R -> x - V (per element, as
//      put_dirichlet_in_matrix does, src/bc_dirich.c:44-153).
//   7. scatter through the element-to-nonzero map into the MSR arrays.
//
// Weak forms (Goma's, restated):
//   momentum  (src/mm_fill_momentum.c:474-672, J_m_v 1564-1735, J_m_T 746-915,
//             J_m_P 2052-2117; fluid_stress 2834-3737):
//     R_a,i = int[ -rho phi_i (v.grad v_a) - grad phi_i . Pi[a][.] + phi_i f_a ] dA
//     Pi = -P I + mu (grad v + grad v^T),  f_a = -rho g_a beta (T - Tref)
//   continuity (src/mm_fill_continuity.c:435-700):  R_k = int psi_k div v dA
//     with PSPG (global tau, src/mm_fill_continuity.c:600-612; calc_pspg
//     src/mm_fill_stabilization.c:1037-1049, 1100-1465, no polymer stress):
//     R_k += int grad psi_k . tau (rho v_dot + rho v.grad v + grad P - f) dA,
//     grad psi_0 = 0 and grad psi_k = B[.][k-1] for the P1 basis {1, xi_k}
//   energy   (src/mm_fill_energy.c:280-700, heat_flux 1629):
//     R_i = int[ -rho Cp w_i (v.grad T) - k grad phi_i . grad T + phi_i h ] dA
//     w_i = phi_i, or with SUPG (src/mm_fill_energy.c:237-259, 332-339)
//     w_i = phi_i + s/h (v.grad phi_i), h = sqrt(sum_p vc_p^2 / hsq_p) / 2
//     from the squared face-centroid widths hsq and the velocity vc at the
//     element's last local node (element_velocity, src/mm_fill_aux.c:798)
#include <cmath>
#include <vector>

#include "gm_launch.h"

template <int DIM>
struct El {
  static constexpr int NVE = DIM == 2 ? 9 : 27;
  static constexpr int NQ = DIM == 2 ? 9 : 27;
  static constexpr int NP = DIM + 1;
};

struct FillArgs {
  const double *coor;       // [dim][num_nodes]
  const int *conn;          // [num_elems][nve]
  const int *fu;            // first unknown per node
  const int *centroid;      // centroid node per element
  const signed char *dbc;   // [N]
  const double *dbc_val;    // [N]
  const double *tables;     // phi, dphidxi, psi, wt
  const int *e2nz;          // [num_elems][E]
  const double *x;
  double *resid;
  void *val;
  double *kel;  // optional [num_elems][E]: element matrices before the Dirichlet rows
                // (for Galerkin multigrid coarse operators), in the e2nz layout
  int *flag;
  // theta-method time step (nullptr x_old: steady): at each node
  // xdot = cmass (x - x_old) - cold xdot_old, cmass = (1 + 2 theta) / dt, cold = 2 theta
  const double *x_old, *xdot_old;
  double cmass, cold;
  // QSIDE / QCONV side entries: those of element e are [qs_ptr[e], qs_ptr[e+1]),
  // each a side number (1-based, Exodus order) and (q, h, T_c) at qs_q[3 k]:
  // inward heat flux q + h (T_c - T); nullptr qs_ptr: none
  const int *qs_ptr, *qs_side;
  const double *qs_q;
  const double *stab;  // side tables: [side][point] phi[NVE], dphidxi[NVE][DIM], weight
  const int *bindx;                 // MSR graph
  const unsigned short *rank;       // rank map [num_elems][2][nve * nve] (k_build_ranks)
  int overwrite;                    // 1: first touch of a node pair stores (no zeroing needed)
  int num_nodes;
  double tau_pspg;  // global PSPG tau, 0: no pressure stabilisation
  double supg_e;    // energy SUPG coefficient, 0: Galerkin weights
  int resid_only;  // 1: residual only (Assemble_Jacobian = FALSE)
  long long E;
  long long off_mp, off_cv, off_cp;
  double rho, mu0, muinf, lam, nexp, aexp, Cp, kcond, hsrc, g[3], beta, Tref;
};

// Reference tables: phi[NQ][NVE], dphidxi[NQ][NVE][DIM], psi[NQ][NP], wt[NQ].
template <int DIM>
struct Tab {
  static constexpr int NVE = El<DIM>::NVE, NQ = El<DIM>::NQ, NP = El<DIM>::NP;
  static constexpr int PHI = 0;
  static constexpr int DPHI = PHI + NQ * NVE;
  static constexpr int PSI = DPHI + NQ * NVE * DIM;
  static constexpr int WT = PSI + NQ * NP;
  static constexpr int SIZE = WT + NQ;
};

// Side tables for surface integrals (QSIDE): for each of the 2*DIM sides in
// Exodus order and each surface Gauss point (3-point rule per free axis),
// phi[NVE], dphidxi[NVE][DIM] and the weight.
template <int DIM>
struct STab {
  static constexpr int NVE = El<DIM>::NVE;
  static constexpr int NSIDE = 2 * DIM;
  static constexpr int NSP = DIM == 2 ? 3 : 9;
  static constexpr int REC = NVE + NVE * DIM + 1;
  static constexpr int SIZE = NSIDE * NSP * REC;
};

// Energy SUPG element parameters (src/mm_fill_energy.c:250-259, 644-651):
// out[0] = s / h, out[1 + b] = s d(1/h)/d vc_b (the derivative in the
// velocity component b at the last local node), all zero when h = 0.  x are
// the element's nodal coordinates (local order), vc the velocity at the last
// local node.  Face-centroid widths as h_elem_siz (src/mm_fill_aux.c:844).
template <int DIM, typename XA>
__device__ inline void supg_params(double s, const XA &x, const double *vc, double *out) {
  double hsq[3];
  if constexpr (DIM == 2) {
    double w0 = 0.0, w1 = 0.0;
#pragma unroll
    for (int a = 0; a < 2; a++) {
      const double d0 = 0.5 * (x[a][1] + x[a][2]) - 0.5 * (x[a][0] + x[a][3]);
      const double d1 = 0.5 * (x[a][0] + x[a][1]) - 0.5 * (x[a][2] + x[a][3]);
      w0 += d0 * d0;
      w1 += d1 * d1;
    }
    hsq[0] = w0;
    hsq[1] = w1;
  } else {
    double w[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double f1 = 0.25 * (x[a][0] + x[a][1] + x[a][2] + x[a][3]);
      const double f2 = 0.25 * (x[a][1] + x[a][2] + x[a][5] + x[a][6]);
      const double f3 = 0.25 * (x[a][2] + x[a][3] + x[a][6] + x[a][7]);
      const double f4 = 0.25 * (x[a][0] + x[a][1] + x[a][4] + x[a][5]);
      const double f5 = 0.25 * (x[a][0] + x[a][3] + x[a][4] + x[a][7]);
      const double f6 = 0.25 * (x[a][4] + x[a][5] + x[a][6] + x[a][7]);
      w[0] += (f2 - f5) * (f2 - f5);
      w[1] += (f3 - f4) * (f3 - f4);
      w[2] += (f1 - f6) * (f1 - f6);
    }
    hsq[0] = w[0];
    hsq[1] = w[1];
    hsq[2] = w[2];
  }
  double he = 0.0;
#pragma unroll
  for (int p = 0; p < DIM; p++) he += vc[p] * vc[p] / hsq[p];
  he = sqrt(he) / 2.0;
  const double hinv = he == 0.0 ? 0.0 : 1.0 / he;
  out[0] = s * hinv;
#pragma unroll
  for (int b = 0; b < DIM; b++) {
    const double dh = hsq[b] != 0.0 ? vc[b] * hinv / 4.0 / hsq[b] : 0.0;
    out[1 + b] = he != 0.0 ? s * (-dh / he / he) : 0.0;
  }
}

template <int DIM, int EN, int VISC, typename JT, int NT>
__global__ __launch_bounds__(NT) void k_fill(FillArgs A, const int *__restrict__ elems) {
  constexpr int NVE = El<DIM>::NVE, NQ = El<DIM>::NQ, NP = El<DIM>::NP;
  constexpr int NV = DIM + EN;
  constexpr int NVE2 = NVE * NVE;
  using T = Tab<DIM>;
  const int e = elems[xcd_block()];
  const int tid = threadIdx.x;
  const double *__restrict__ tb = A.tables;

  __shared__ double s_x[DIM][NVE];
  __shared__ double s_u[NV][NVE];
  __shared__ double s_p[NP];
  __shared__ int s_gun[NV][NVE];
  __shared__ int s_gp[NP];
  __shared__ double s_B[NQ][DIM][DIM];
  __shared__ double s_dA[NQ];
  __shared__ double s_g[NQ][NVE][DIM];
  __shared__ double s_v[NQ][DIM];
  __shared__ double s_gv[NQ][DIM][DIM];  // gv[p][a] = d v_a / d x_p
  __shared__ double s_gT[NQ][DIM];
  __shared__ double s_P[NQ];
  __shared__ double s_mu[NQ];
  __shared__ double s_dmu[NQ];           // (dmu/dgammadot) / gammadot, 0 if gammadot == 0
  __shared__ double s_f[NQ][DIM];
  __shared__ double s_div[NQ];
  __shared__ double s_ud[NV][NVE];       // nodal time derivatives (transient)
  __shared__ double s_vd[NQ][DIM];       // v_dot at the quadrature points
  __shared__ double s_Td[NQ];            // T_dot
  __shared__ double s_ps[NQ][DIM];       // PSPG: tau dA (rho v_dot + rho v.grad v + grad P - f)
  __shared__ double s_su[4];             // energy SUPG: s/h, s d(1/h)/d vc_b
  constexpr int NSP = STab<DIM>::NSP, SREC = STab<DIM>::REC;
  __shared__ double s_qw[GM_QS_MAX][NSP];  // QSIDE / QCONV: w |surface det| (q + h (T_c - T))
  __shared__ double s_qh[GM_QS_MAX][NSP];  // QCONV: w |surface det| h (the Jacobian's weight)
  __shared__ int s_qsd[GM_QS_MAX];         // side index (0-based) of each entry
  const int qs0 = A.qs_ptr ? A.qs_ptr[e] : 0;
  const int nqs = A.qs_ptr ? A.qs_ptr[e + 1] - qs0 : 0;

  // 1. gather
  if (tid < NVE) {
    const int n = A.conn[(size_t)e * NVE + tid];
    const int f = A.fu[n];
#pragma unroll
    for (int k = 0; k < DIM; k++) s_x[k][tid] = A.coor[(size_t)k * A.num_nodes + n];
#pragma unroll
    for (int a = 0; a < NV; a++) {
      s_gun[a][tid] = f + a;
      s_u[a][tid] = A.x[f + a];
      // xdot as solve_nonlinear_problem forms it after each update (src/rf_solve.c:2848)
      s_ud[a][tid] = A.x_old ? A.cmass * (A.x[f + a] - A.x_old[f + a]) - A.cold * A.xdot_old[f + a]
                             : 0.0;
    }
  } else if (tid >= 64 && tid < 64 + NP) {
    const int k = tid - 64;
    const int g = A.fu[A.centroid[e]] + DIM + EN + k;
    s_gp[k] = g;
    s_p[k] = A.x[g];
  }
  __syncthreads();

  // 2. map per quadrature point
  if (tid < NQ) {
    const int q = tid;
    double J[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
      for (int j = 0; j < DIM; j++) {
        double s = 0.0;
        for (int k = 0; k < NVE; k++) s += s_x[j][k] * tb[T::DPHI + (q * NVE + k) * DIM + i];
        J[i][j] = s;
      }
    double det;
    if constexpr (DIM == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      s_B[q][0][0] = J[1][1] / det;
      s_B[q][0][1] = -J[0][1] / det;
      s_B[q][1][0] = -J[1][0] / det;
      s_B[q][1][1] = J[0][0] / det;
    } else {
      det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
            J[0][1] * (J[1][0] * J[2][2] - J[2][0] * J[1][2]) +
            J[0][2] * (J[1][0] * J[2][1] - J[2][0] * J[1][1]);
      s_B[q][0][0] = (J[1][1] * J[2][2] - J[2][1] * J[1][2]) / det;
      s_B[q][0][1] = -(J[0][1] * J[2][2] - J[2][1] * J[0][2]) / det;
      s_B[q][0][2] = (J[0][1] * J[1][2] - J[1][1] * J[0][2]) / det;
      s_B[q][1][0] = -(J[1][0] * J[2][2] - J[2][0] * J[1][2]) / det;
      s_B[q][1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) / det;
      s_B[q][1][2] = -(J[0][0] * J[1][2] - J[1][0] * J[0][2]) / det;
      s_B[q][2][0] = (J[1][0] * J[2][1] - J[2][0] * J[1][1]) / det;
      s_B[q][2][1] = -(J[0][0] * J[2][1] - J[2][0] * J[0][1]) / det;
      s_B[q][2][2] = (J[0][0] * J[1][1] - J[1][0] * J[0][1]) / det;
    }
    if (!(det > 0.0)) atomicOr(&A.flag[0], 1);
    s_dA[q] = det * tb[T::WT + q];
  } else if (tid == NQ) {
    double vc[DIM], su[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int p = 0; p < DIM; p++) vc[p] = s_u[p][NVE - 1];
    if (EN && A.supg_e != 0.0) supg_params<DIM>(A.supg_e, s_x, vc, su);
#pragma unroll
    for (int k = 0; k < 4; k++) s_su[k] = su[k];
  }
  __syncthreads();

  // 3. physical gradients
  for (int idx = tid; idx < NQ * NVE; idx += NT) {
    const int q = idx / NVE, k = idx - q * NVE;
    double dxi[DIM];
#pragma unroll
    for (int r = 0; r < DIM; r++) dxi[r] = tb[T::DPHI + (q * NVE + k) * DIM + r];
#pragma unroll
    for (int p = 0; p < DIM; p++) {
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < DIM; r++) s += s_B[q][p][r] * dxi[r];
      s_g[q][k][p] = s;
    }
  }
  __syncthreads();

  // 4. fields at quadrature points
  if (tid < NQ) {
    const int q = tid;
    double v[DIM], gv[DIM][DIM], Tq = 0.0, gT[DIM];
#pragma unroll
    for (int a = 0; a < DIM; a++) {
      v[a] = 0.0;
      gT[a] = 0.0;
#pragma unroll
      for (int p = 0; p < DIM; p++) gv[p][a] = 0.0;
    }
    for (int k = 0; k < NVE; k++) {
      const double ph = tb[T::PHI + q * NVE + k];
#pragma unroll
      for (int a = 0; a < DIM; a++) {
        const double ua = s_u[a][k];
        v[a] += ua * ph;
#pragma unroll
        for (int p = 0; p < DIM; p++) gv[p][a] += ua * s_g[q][k][p];
      }
      if constexpr (EN) {
        const double tk = s_u[DIM][k];
        Tq += tk * ph;
#pragma unroll
        for (int p = 0; p < DIM; p++) gT[p] += tk * s_g[q][k][p];
      }
    }
    double vd[DIM], Td = 0.0;
#pragma unroll
    for (int a = 0; a < DIM; a++) vd[a] = 0.0;
    if (A.x_old) {
      for (int k = 0; k < NVE; k++) {
        const double ph = tb[T::PHI + q * NVE + k];
#pragma unroll
        for (int a = 0; a < DIM; a++) vd[a] += s_ud[a][k] * ph;
        if constexpr (EN) Td += s_ud[DIM][k] * ph;
      }
    }
#pragma unroll
    for (int a = 0; a < DIM; a++) s_vd[q][a] = vd[a];
    s_Td[q] = Td;
    double P = 0.0;
#pragma unroll
    for (int k = 0; k < NP; k++) P += s_p[k] * tb[T::PSI + q * NP + k];
    double mu = A.mu0, dmu = 0.0;
    if constexpr (VISC == 1) {
      // calc_shearrate + carreau_viscosity (src/mm_fill_util.c:4564,
      // src/mm_viscosity.c:1065-1116)
      double gd = 0.0;
#pragma unroll
      for (int a = 0; a < DIM; a++)
#pragma unroll
        for (int b = 0; b < DIM; b++) {
          const double gab = gv[a][b] + gv[b][a];
          gd += gab * gab;
        }
      gd = sqrt(0.5 * fabs(gd));
      const bool nz = fabs(gd) > 1.0e-298;
      const double val2 = nz ? pow(A.lam * gd, A.aexp) : 0.0;
      mu = A.muinf + (A.mu0 - A.muinf) * pow(1.0 + val2, (A.nexp - 1.0) / A.aexp);
      if (nz) {
        const double val = pow(A.lam * gd, A.aexp - 1.0);
        const double val1 = pow(1.0 + val2, (A.nexp - 1.0 - A.aexp) / A.aexp);
        const double dmu_gd = (A.mu0 - A.muinf) * (A.nexp - 1.0) * A.lam * val * val1;
        dmu = dmu_gd / gd;
      }
    }
    double div = 0.0;
#pragma unroll
    for (int a = 0; a < DIM; a++) {
      s_v[q][a] = v[a];
      s_gT[q][a] = gT[a];
      div += gv[a][a];
#pragma unroll
      for (int p = 0; p < DIM; p++) s_gv[q][p][a] = gv[p][a];
      s_f[q][a] = EN ? A.rho * A.g[a] * (-A.beta * (Tq - A.Tref)) : 0.0;
    }
    s_P[q] = P;
    s_mu[q] = mu;
    s_dmu[q] = dmu;
    s_div[q] = div;
    const double w = A.tau_pspg * s_dA[q];
#pragma unroll
    for (int a = 0; a < DIM; a++) {
      double adv = 0.0, gP = 0.0;
#pragma unroll
      for (int p = 0; p < DIM; p++) adv += v[p] * gv[p][a];
#pragma unroll
      for (int k = 1; k < NP; k++) gP += s_p[k] * s_B[q][a][k - 1];
      s_ps[q][a] = w * (A.rho * (adv + vd[a]) + gP - s_f[q][a]);
    }
  } else if (EN && tid >= NT / 2 && tid < NT / 2 + nqs * NSP) {
    // QSIDE (src/bc_integ.c:1518-1519, 2156-2157): w * |surface Jacobian| * q
    // at one surface Gauss point of one side entry; the side's free reference
    // axes give the tangents, as get_side_info orders them
    // (src/mm_fill_fill.c:3049-3072, 3182-3250)
    const int k = (tid - NT / 2) / NSP, g = tid - NT / 2 - k * NSP;
    const int sd = A.qs_side[qs0 + k] - 1;
    const double *st = A.stab + (size_t)(sd * NSP + g) * SREC;
    const int ax = (sd == 0 || sd == 2) ? 1 : (sd == 1 || sd == 3) ? 0 : 2;
    int fr[2], nf = 0;
#pragma unroll
    for (int a = 0; a < DIM; a++)
      if (a != ax) fr[nf++] = a;
    double t1[DIM], t2[DIM];
#pragma unroll
    for (int j = 0; j < DIM; j++) {
      t1[j] = 0.0;
      t2[j] = 0.0;
      for (int i = 0; i < NVE; i++) {
        t1[j] += s_x[j][i] * st[NVE + i * DIM + fr[0]];
        if constexpr (DIM == 3) t2[j] += s_x[j][i] * st[NVE + i * DIM + fr[1]];
      }
    }
    double sdet;
    if constexpr (DIM == 2) {
      sdet = sqrt(t1[0] * t1[0] + t1[1] * t1[1]);
    } else {
      const double c0 = t1[1] * t2[2] - t1[2] * t2[1], c1 = t1[2] * t2[0] - t1[0] * t2[2],
                   c2 = t1[0] * t2[1] - t1[1] * t2[0];
      sdet = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    }
    const double *qk = A.qs_q + 3 * (qs0 + k), wd = st[NVE + NVE * DIM] * sdet;
    double Tg = 0.0;
    if constexpr (EN)
      for (int j = 0; j < NVE; j++) Tg += st[j] * s_u[DIM][j];
    s_qw[k][g] = wd * (qk[0] + qk[1] * (qk[2] - Tg));
    s_qh[k][g] = wd * qk[1];
    if (g == 0) s_qsd[k] = sd;
  }
  __syncthreads();

  const double rho = A.rho, rCp = A.rho * A.Cp, kc = A.kcond;
  const double cm = A.x_old ? A.cmass : 0.0;  // d xdot / d x: the mass-matrix coefficient
  JT *__restrict__ out = (JT *)A.val;
  const int *__restrict__ map = A.e2nz + (size_t)e * A.E;

  // 5a. residual rows
  for (int t = tid; t < NV * NVE + NP; t += NT) {
    int row;
    double R = 0.0;
    if (t < NV * NVE) {
      const int a = t / NVE, i = t - a * NVE;
      row = s_gun[a][i];
      if (a < DIM) {
        for (int q = 0; q < NQ; q++) {
          const double ph = tb[T::PHI + q * NVE + i];
          double adv = 0.0, dif = 0.0;
#pragma unroll
          for (int p = 0; p < DIM; p++) {
            adv += s_v[q][p] * s_gv[q][p][a];
            const double Pi_ap = (a == p ? -s_P[q] : 0.0) + s_mu[q] * (s_gv[q][a][p] + s_gv[q][p][a]);
            dif += s_g[q][i][p] * Pi_ap;
          }
          R += s_dA[q] * (-rho * ph * adv - dif + ph * s_f[q][a] - rho * ph * s_vd[q][a]);
        }
      } else {
        const double sw = s_su[0];
        for (int q = 0; q < NQ; q++) {
          const double ph = tb[T::PHI + q * NVE + i];
          double adv = 0.0, dif = 0.0, vgi = 0.0;
#pragma unroll
          for (int p = 0; p < DIM; p++) {
            adv += s_v[q][p] * s_gT[q][p];
            dif += s_g[q][i][p] * s_gT[q][p];
            vgi += s_v[q][p] * s_g[q][i][p];
          }
          R += s_dA[q] * (-rCp * (ph + sw * vgi) * adv - kc * dif + ph * A.hsrc - rCp * ph * s_Td[q]);
        }
        for (int k = 0; k < nqs; k++)
          for (int g = 0; g < NSP; g++) R += s_qw[k][g] * A.stab[(size_t)(s_qsd[k] * NSP + g) * SREC + i];
      }
    } else {
      const int k = t - NV * NVE;
      row = s_gp[k];
      for (int q = 0; q < NQ; q++) R += s_dA[q] * tb[T::PSI + q * NP + k] * s_div[q];
      if (A.tau_pspg != 0.0 && k > 0)
        for (int q = 0; q < NQ; q++)
#pragma unroll
          for (int a = 0; a < DIM; a++) R += s_B[q][a][k - 1] * s_ps[q][a];
    }
    if (A.dbc[row]) R = A.x[row] - A.dbc_val[row];
    A.resid[row] += R;
  }

  if (A.resid_only) return;

  // 5b. nodal-variable blocks, one node pair (i,j) per task
  for (int t = tid; t < NVE2; t += NT) {
    const int i = t / NVE, j = t - i * NVE;
    double acc[NV][NV];
#pragma unroll
    for (int a = 0; a < NV; a++)
#pragma unroll
      for (int b = 0; b < NV; b++) acc[a][b] = 0.0;
    for (int q = 0; q < NQ; q++) {
      const double dA = s_dA[q];
      const double phi_i = tb[T::PHI + q * NVE + i];
      const double phi_j = tb[T::PHI + q * NVE + j];
      double Gi[DIM], gj[DIM];
#pragma unroll
      for (int p = 0; p < DIM; p++) {
        Gi[p] = s_g[q][i][p];
        gj[p] = s_g[q][j][p];
      }
      double advj = 0.0, GG = 0.0;
#pragma unroll
      for (int p = 0; p < DIM; p++) {
        advj += s_v[q][p] * gj[p];
        GG += Gi[p] * gj[p];
      }
      const double mu = s_mu[q];
      const double pij = phi_i * phi_j;
      double Gg[DIM], dmb[DIM];  // sum_p Gi[p] gamma[a][p], c * sum_p gj[p] gamma[b][p]
      if constexpr (VISC == 1) {
        const double c = s_dmu[q];
#pragma unroll
        for (int a = 0; a < DIM; a++) {
          double s1 = 0.0, s2 = 0.0;
#pragma unroll
          for (int p = 0; p < DIM; p++) {
            const double gam = s_gv[q][a][p] + s_gv[q][p][a];
            s1 += Gi[p] * gam;
            s2 += gj[p] * gam;
          }
          Gg[a] = s1;
          dmb[a] = c * s2;
        }
      }
      const double diag = rho * phi_i * advj + mu * GG + rho * cm * pij;
#pragma unroll
      for (int a = 0; a < DIM; a++) {
#pragma unroll
        for (int b = 0; b < DIM; b++) {
          double v = -rho * pij * s_gv[q][b][a] - mu * gj[a] * Gi[b];
          if (a == b) v -= diag;
          if constexpr (VISC == 1) v -= dmb[b] * Gg[a];
          acc[a][b] += dA * v;
        }
      }
      if constexpr (EN) {
        // SUPG weight w_i = phi_i + sw (v.G_i) on the advection terms, and
        // the weight's own derivative in v (d_wt_func)
        const double sw = s_su[0];
        double vgi = 0.0, vgT = 0.0;
#pragma unroll
        for (int p = 0; p < DIM; p++) {
          vgi += s_v[q][p] * Gi[p];
          vgT += s_v[q][p] * s_gT[q][p];
        }
        const double wi = phi_i + sw * vgi;
#pragma unroll
        for (int a = 0; a < DIM; a++) {
          acc[a][DIM] += dA * (-rho * A.g[a] * A.beta * pij);
          double dw = sw * phi_j * Gi[a];
          if (j == NVE - 1) dw += s_su[1 + a] * s_v[q][a] * Gi[a];
          acc[DIM][a] += dA * (-rCp * (wi * phi_j * s_gT[q][a] + dw * vgT));
        }
        acc[DIM][DIM] += dA * (-rCp * wi * advj - kc * GG - rCp * cm * pij);
      }
    }
    if (A.kel) {
      double *__restrict__ ke = A.kel + (size_t)e * A.E;
#pragma unroll
      for (int a = 0; a < NV; a++)
#pragma unroll
        for (int b = 0; b < NV; b++) ke[(a * NV + b) * NVE2 + t] = acc[a][b];
    }
#pragma unroll
    for (int a = 0; a < NV; a++) {
      const int row = s_gun[a][i];
      const bool dir = A.dbc[row] != 0;
#pragma unroll
      for (int b = 0; b < NV; b++) {
        double v = acc[a][b];
        if (dir) {
          if (a == b && i == j) v = 1.0;
          else continue;
        }
        const int idx = map[(a * NV + b) * NVE2 + t];
        out[idx] += (JT)v;
      }
    }
  }

  // QCONV: -h w |surface Jacobian| phi_i phi_j over the element's sides in the
  // T-T block, on top of the entries scattered above
  if constexpr (EN) {
    bool conv = false;
    for (int k = 0; k < nqs; k++) conv |= A.qs_q[3 * (qs0 + k) + 1] != 0.0;
    if (conv) {
      __syncthreads();
      for (int t = tid; t < NVE2; t += NT) {
        const int i = t / NVE, j = t - i * NVE;
        double v = 0.0;
#pragma unroll 1
        for (int k = 0; k < nqs; k++) {
          if (A.qs_q[3 * (qs0 + k) + 1] == 0.0) continue;
#pragma unroll 1
          for (int g = 0; g < NSP; g++) {
            const double *st = A.stab + (size_t)(s_qsd[k] * NSP + g) * SREC;
            v -= s_qh[k][g] * st[i] * st[j];
          }
        }
        if (v == 0.0) continue;  // a node pair off the sides
        if (A.kel) A.kel[(size_t)e * A.E + (size_t)(DIM * NV + DIM) * NVE2 + t] += v;
        if (A.dbc[s_gun[DIM][i]]) continue;  // a Dirichlet row keeps its unit diagonal only
        out[map[(DIM * NV + DIM) * NVE2 + t]] += (JT)v;
      }
    }
  }

  // 5c. momentum x pressure (J_m_P): int grad_a phi_i psi_k dA
  for (int t = tid; t < NVE * NP; t += NT) {
    const int i = t / NP, k = t - i * NP;
    double acc[DIM];
#pragma unroll
    for (int a = 0; a < DIM; a++) acc[a] = 0.0;
    for (int q = 0; q < NQ; q++) {
      const double w = s_dA[q] * tb[T::PSI + q * NP + k];
#pragma unroll
      for (int a = 0; a < DIM; a++) acc[a] += w * s_g[q][i][a];
    }
#pragma unroll
    for (int a = 0; a < DIM; a++) {
      if (A.kel) A.kel[(size_t)e * A.E + A.off_mp + a * NVE * NP + t] = acc[a];
      if (A.dbc[s_gun[a][i]]) continue;
      const int idx = map[A.off_mp + a * NVE * NP + t];
      out[idx] += (JT)acc[a];
    }
  }

  // 5d. continuity x velocity (J_c_v): int psi_k grad_b phi_j dA; PSPG adds
  // tau rho [phi_j grad psi_k . d_b v + grad_b psi_k (v.grad phi_j + c phi_j)]
  // and continuity x temperature tau rho beta (grad psi_k . g) phi_j
  const bool ps = A.tau_pspg != 0.0;
  for (int t = tid; t < NP * NVE; t += NT) {
    const int k = t / NVE, j = t - k * NVE;
    const bool dir = A.dbc[s_gp[k]] != 0;
    if (dir && !A.kel) continue;
    double acc[DIM], accT = 0.0;
#pragma unroll
    for (int b = 0; b < DIM; b++) acc[b] = 0.0;
    for (int q = 0; q < NQ; q++) {
      const double w = s_dA[q] * tb[T::PSI + q * NP + k];
#pragma unroll
      for (int b = 0; b < DIM; b++) acc[b] += w * s_g[q][j][b];
      if (ps && k > 0) {
        const double wp = A.tau_pspg * s_dA[q] * rho, phj = tb[T::PHI + q * NVE + j];
        double vg = 0.0, gk[DIM], gkg = 0.0;
#pragma unroll
        for (int p = 0; p < DIM; p++) {
          vg += s_v[q][p] * s_g[q][j][p];
          gk[p] = s_B[q][p][k - 1];
          gkg += gk[p] * A.g[p];
        }
#pragma unroll
        for (int b = 0; b < DIM; b++) {
          double gv = 0.0;
#pragma unroll
          for (int a = 0; a < DIM; a++) gv += gk[a] * s_gv[q][b][a];
          acc[b] += wp * (phj * gv + gk[b] * (vg + cm * phj));
        }
        accT += wp * A.beta * gkg * phj;
      }
    }
#pragma unroll
    for (int b = 0; b < DIM; b++) {
      if (A.kel) A.kel[(size_t)e * A.E + A.off_cv + b * NP * NVE + t] = acc[b];
      if (dir) continue;
      const int idx = map[A.off_cv + b * NP * NVE + t];
      out[idx] += (JT)acc[b];
    }
    if (EN && ps && k > 0 && !dir) {
      // the element layout holds no continuity x temperature entries: find
      // the column in the row (sorted off-diagonal list)
      const int row = s_gp[k], col = s_gun[DIM][j];
      int lo = A.bindx[row], hi = A.bindx[row + 1] - 1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1, v = A.bindx[mid];
        if (v == col) {
          out[mid] += (JT)accT;
          break;
        }
        if (v < col) lo = mid + 1;
        else hi = mid - 1;
      }
    }
  }

  // 5e. continuity x pressure: zero for Galerkin Q2-P1; PSPG adds
  // tau grad psi_k . grad psi_l; a unit diagonal on a Dirichlet (datum) row
  if (tid < NP * NP) {
    const int k = tid / NP, k2 = tid - k * NP;
    double acc = 0.0;
    if (ps && k > 0 && k2 > 0)
      for (int q = 0; q < NQ; q++) {
        double g2 = 0.0;
#pragma unroll
        for (int a = 0; a < DIM; a++) g2 += s_B[q][a][k - 1] * s_B[q][a][k2 - 1];
        acc += A.tau_pspg * s_dA[q] * g2;
      }
    if (A.kel) A.kel[(size_t)e * A.E + A.off_cp + tid] = acc;
    if (A.dbc[s_gp[k]]) {
      if (k == k2) out[map[A.off_cp + tid]] += (JT)1.0;
    } else if (acc != 0.0) {
      out[map[A.off_cp + tid]] += (JT)acc;
    }
  }
}

// ---------------------------------------------------------------------------
// HEX27 assembly on the matrix cores.
//
// One workgroup (four waves) per element.  The basis tables, the physical
// gradients and the quadrature-point fields are staged in LDS; every
// node-pair block of the element Jacobian is a sum of products
//   E(X, w, Y)[i][j] = sum_q X[q][i] w[q] Y[q][j]
// of 27 x 27 operand tables (X, Y in {phi, d phi / d x_p}, weights from the
// fields), computed with v_mfma_f64_16x16x4f64 on the 27 x 27 node-pair
// matrix padded to 32 x 32 (two 16-wide tiles per side, one tile per wave)
// and 27 quadrature points padded to 28 (seven k-steps).  Per block (a, b),
// with c = d xdot / d x (transient), mu constant or Carreau mu(q):
//   J_ab = E(phi, -rho dA d_b v_a, phi) - E(G_b, mu dA, G_a)
//          - delta_ab [rho E(phi, dA, v.G) + sum_p E(G_p, mu dA, G_p) + rho c E(phi, dA, phi)]
//          - E(H_a, c_mu dA, H_b)           (Carreau: H_a = sum_p G_p gamma_ap,
//                                            c_mu = (d mu / d gammadot) / gammadot)
//   J_aT = -rho g_a beta E(phi, dA, phi),   J_Ta = E(phi, -rho Cp dA d_a T, phi)
//   J_TT = -rho Cp E(phi, dA, v.G) - k sum_p E(G_p, dA, G_p) - rho Cp c E(phi, dA, phi)
// (the terms of src/mm_fill_momentum.c:1564-1735, 746-915 and
// src/mm_fill_energy.c:425-487, 620-690 regrouped as products over q).
// MFMA lane layout (checked on the device): lane l supplies A[l%16][l/16]
// and B[l/16][l%16] and holds D[l/16 + 4r][l%16], r = 0..3 (rows interleaved).
//
// Scatter: each entry's MSR position comes from the element's rank map
// (rank of node j's first column in the rows of node i, see
// k_build_ranks), so a node pair's NV x NV block lands at
// bindx[row] + rank + b - (col > row).  With `overwrite` the first element
// (in colour order) to touch a node pair stores its values, later ones add:
// no zeroing pass and no read of the old value on first touch.
typedef double d4v __attribute__((ext_vector_type(4)));

template <int EN, int VISC, typename JT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VISC ? 3 : 4, VISC ? 3 : 4))) void k_fill3m(
    FillArgs A, const int *__restrict__ elems) {
  constexpr int DIM = 3, NVE = 27, NQ = 27, NP = 4, NV = DIM + EN, NVE2 = NVE * NVE;
  constexpr int KP = 28, LD = 32, NT = 256, CL = 20;
  using T = Tab<3>;
  constexpr int NSP = STab<3>::NSP, SREC = STab<3>::REC;
  const int e = elems[xcd_block()];
  const int tid = threadIdx.x;
  const double *__restrict__ tb = A.tables;

  __shared__ double s_x[DIM][NVE];
  __shared__ double s_u[NV][NVE];
  __shared__ double s_ud[NV][NVE];
  __shared__ double s_p[NP];
  __shared__ int s_gun[NV][NVE];   // global row of each nodal unknown
  __shared__ int s_brow[NV][NVE];  // bindx[row]: start of its off-diagonal list
  __shared__ signed char s_dir[NV][NVE];
  __shared__ int s_gp[NP], s_bp[NP];
  __shared__ signed char s_dirp[NP];
  __shared__ double s_B[NQ][DIM][DIM];
  __shared__ double s_dA[KP];
  __shared__ double s_Phi[KP][LD];
  __shared__ double s_G[DIM][KP][LD];
  __shared__ double s_v[KP][DIM];
  __shared__ double s_gv[KP][DIM][DIM];  // gv[p][a] = d v_a / d x_p
  __shared__ double s_gT[KP][DIM];
  __shared__ double s_P[KP], s_mu[KP], s_dmu[KP], s_div[KP], s_Td[KP];
  __shared__ double s_f[KP][DIM], s_vd[KP][DIM];
  __shared__ double s_su[4];  // energy SUPG: s/h, s d(1/h)/d vc_b
  __shared__ double s_qw[GM_QS_MAX][NSP], s_qh[GM_QS_MAX][NSP];
  __shared__ int s_qsd[GM_QS_MAX];
  const int qs0 = A.qs_ptr ? A.qs_ptr[e] : 0;
  const int nqs = A.qs_ptr ? A.qs_ptr[e + 1] - qs0 : 0;

  // 1. gather; basis values into the padded operand table
  if (tid < NVE) {
    const int n = A.conn[(size_t)e * NVE + tid];
    const int f = A.fu[n];
#pragma unroll
    for (int k = 0; k < DIM; k++) s_x[k][tid] = A.coor[(size_t)k * A.num_nodes + n];
#pragma unroll
    for (int a = 0; a < NV; a++) {
      const int g = f + a;
      s_gun[a][tid] = g;
      s_brow[a][tid] = A.bindx[g];
      s_dir[a][tid] = A.dbc[g];
      s_u[a][tid] = A.x[g];
      s_ud[a][tid] = A.x_old ? A.cmass * (A.x[g] - A.x_old[g]) - A.cold * A.xdot_old[g] : 0.0;
    }
  } else if (tid >= 64 && tid < 64 + NP) {
    const int k = tid - 64;
    const int g = A.fu[A.centroid[e]] + DIM + EN + k;
    s_gp[k] = g;
    s_bp[k] = A.bindx[g];
    s_dirp[k] = A.dbc[g];
    s_p[k] = A.x[g];
  }
  for (int idx = tid; idx < KP * LD; idx += NT) {
    const int q = idx / LD, i = idx - q * LD;
    s_Phi[q][i] = (q < NQ && i < NVE) ? tb[T::PHI + q * NVE + i] : 0.0;
  }
  __syncthreads();

  // 2. map per quadrature point (beer_belly)
  if (tid < NQ) {
    const int q = tid;
    double J[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
      for (int j = 0; j < DIM; j++) {
        double s = 0.0;
#pragma unroll 3
        for (int k = 0; k < NVE; k++) s += s_x[j][k] * tb[T::DPHI + (q * NVE + k) * DIM + i];
        J[i][j] = s;
      }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[2][0] * J[1][2]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[2][0] * J[1][1]);
    s_B[q][0][0] = (J[1][1] * J[2][2] - J[2][1] * J[1][2]) / det;
    s_B[q][0][1] = -(J[0][1] * J[2][2] - J[2][1] * J[0][2]) / det;
    s_B[q][0][2] = (J[0][1] * J[1][2] - J[1][1] * J[0][2]) / det;
    s_B[q][1][0] = -(J[1][0] * J[2][2] - J[2][0] * J[1][2]) / det;
    s_B[q][1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) / det;
    s_B[q][1][2] = -(J[0][0] * J[1][2] - J[1][0] * J[0][2]) / det;
    s_B[q][2][0] = (J[1][0] * J[2][1] - J[2][0] * J[1][1]) / det;
    s_B[q][2][1] = -(J[0][0] * J[2][1] - J[2][0] * J[0][1]) / det;
    s_B[q][2][2] = (J[0][0] * J[1][1] - J[1][0] * J[0][1]) / det;
    if (!(det > 0.0)) atomicOr(&A.flag[0], 1);
    s_dA[q] = det * tb[T::WT + q];
  } else if (tid == NQ) {
    s_dA[NQ] = 0.0;  // the padding point carries no weight
  } else if (tid == NQ + 1) {
    double vc[DIM], su[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int p = 0; p < DIM; p++) vc[p] = s_u[p][NVE - 1];
    if (EN && A.supg_e != 0.0) supg_params<DIM>(A.supg_e, s_x, vc, su);
#pragma unroll
    for (int k = 0; k < 4; k++) s_su[k] = su[k];
  } else if (EN && tid >= 64 && tid < 64 + nqs * NSP) {
    // QSIDE (src/bc_integ.c:1518-1519, 2156-2157): w |surface Jacobian| q at
    // one surface Gauss point of one side entry
    const int t2 = tid - 64;
    const int k = t2 / NSP, g = t2 - k * NSP;
    const int sd = A.qs_side[qs0 + k] - 1;
    const double *st = A.stab + (size_t)(sd * NSP + g) * SREC;
    const int ax = (sd == 0 || sd == 2) ? 1 : (sd == 1 || sd == 3) ? 0 : 2;
    int fr[2], nf = 0;
#pragma unroll
    for (int a = 0; a < DIM; a++)
      if (a != ax) fr[nf++] = a;
    double t1[DIM], tt2[DIM];
#pragma unroll
    for (int j = 0; j < DIM; j++) {
      t1[j] = 0.0;
      tt2[j] = 0.0;
      for (int i = 0; i < NVE; i++) {
        t1[j] += s_x[j][i] * st[NVE + i * DIM + fr[0]];
        tt2[j] += s_x[j][i] * st[NVE + i * DIM + fr[1]];
      }
    }
    const double c0 = t1[1] * tt2[2] - t1[2] * tt2[1], c1 = t1[2] * tt2[0] - t1[0] * tt2[2],
                 c2 = t1[0] * tt2[1] - t1[1] * tt2[0];
    const double *qk = A.qs_q + 3 * (qs0 + k);
    const double wd = st[NVE + NVE * DIM] * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    double Tg = 0.0;
    if constexpr (EN)
      for (int j = 0; j < NVE; j++) Tg += st[j] * s_u[DIM][j];
    s_qw[k][g] = wd * (qk[0] + qk[1] * (qk[2] - Tg));
    s_qh[k][g] = wd * qk[1];
    if (g == 0) s_qsd[k] = sd;
  }
  __syncthreads();

  // 3. physical gradients into the padded operand tables (load_bf_grad)
  for (int idx = tid; idx < DIM * KP * LD; idx += NT) {
    const int p = idx / (KP * LD), r0 = idx - p * KP * LD, q = r0 / LD, i = r0 - q * LD;
    double s = 0.0;
    if (q < NQ && i < NVE) {
#pragma unroll
      for (int r = 0; r < DIM; r++) s += s_B[q][p][r] * tb[T::DPHI + (q * NVE + i) * DIM + r];
    }
    s_G[p][q][i] = s;
  }
  __syncthreads();

  // 4. fields at the quadrature points (load_fv, load_fv_grads): eight
  // threads per point, each summing a quarter of the nodes, then a
  // butterfly over the eight
  if (tid < NQ * 8) {
    const int q = tid >> 3, sub = tid & 7;
    double v[DIM] = {0.0, 0.0, 0.0}, gv[DIM][DIM], Tq = 0.0, gT[DIM] = {0.0, 0.0, 0.0};
    double vd[DIM] = {0.0, 0.0, 0.0}, Td = 0.0;
#pragma unroll
    for (int a = 0; a < DIM; a++)
#pragma unroll
      for (int p = 0; p < DIM; p++) gv[p][a] = 0.0;
#pragma unroll 1
    for (int k = sub; k < NVE; k += 8) {
      const double ph = s_Phi[q][k];
      double g[DIM];
#pragma unroll
      for (int p = 0; p < DIM; p++) g[p] = s_G[p][q][k];
#pragma unroll
      for (int a = 0; a < DIM; a++) {
        const double ua = s_u[a][k];
        v[a] += ua * ph;
        vd[a] += s_ud[a][k] * ph;
#pragma unroll
        for (int p = 0; p < DIM; p++) gv[p][a] += ua * g[p];
      }
      if constexpr (EN) {
        const double tk = s_u[DIM][k];
        Tq += tk * ph;
        Td += s_ud[DIM][k] * ph;
#pragma unroll
        for (int p = 0; p < DIM; p++) gT[p] += tk * g[p];
      }
    }
#pragma unroll
    for (int o = 4; o > 0; o >>= 1) {
#pragma unroll
      for (int a = 0; a < DIM; a++) {
        v[a] += __shfl_xor(v[a], o, 8);
        vd[a] += __shfl_xor(vd[a], o, 8);
        gT[a] += __shfl_xor(gT[a], o, 8);
#pragma unroll
        for (int p = 0; p < DIM; p++) gv[p][a] += __shfl_xor(gv[p][a], o, 8);
      }
      Tq += __shfl_xor(Tq, o, 8);
      Td += __shfl_xor(Td, o, 8);
    }
    if (sub == 0) {
      double P = 0.0;
#pragma unroll
      for (int k = 0; k < NP; k++) P += s_p[k] * tb[T::PSI + q * NP + k];
      double mu = A.mu0, dmu = 0.0;
      if constexpr (VISC == 1) {
        // calc_shearrate + carreau_viscosity (src/mm_fill_util.c:4564,
        // src/mm_viscosity.c:1065-1116)
        double gd = 0.0;
#pragma unroll
        for (int a = 0; a < DIM; a++)
#pragma unroll
          for (int b = 0; b < DIM; b++) {
            const double gab = gv[a][b] + gv[b][a];
            gd += gab * gab;
          }
        gd = sqrt(0.5 * fabs(gd));
        const bool nz = fabs(gd) > 1.0e-298;
        const double val2 = nz ? pow(A.lam * gd, A.aexp) : 0.0;
        mu = A.muinf + (A.mu0 - A.muinf) * pow(1.0 + val2, (A.nexp - 1.0) / A.aexp);
        if (nz) {
          const double val = pow(A.lam * gd, A.aexp - 1.0);
          const double val1 = pow(1.0 + val2, (A.nexp - 1.0 - A.aexp) / A.aexp);
          dmu = (A.mu0 - A.muinf) * (A.nexp - 1.0) * A.lam * val * val1 / gd;
        }
      }
      double div = 0.0;
#pragma unroll
      for (int a = 0; a < DIM; a++) {
        s_v[q][a] = v[a];
        s_vd[q][a] = vd[a];
        s_gT[q][a] = gT[a];
        div += gv[a][a];
#pragma unroll
        for (int p = 0; p < DIM; p++) s_gv[q][p][a] = gv[p][a];
        s_f[q][a] = EN ? A.rho * A.g[a] * (-A.beta * (Tq - A.Tref)) : 0.0;
      }
      s_P[q] = P;
      s_mu[q] = mu;
      s_dmu[q] = dmu;
      s_div[q] = div;
      s_Td[q] = Td;
    }
  } else if (tid == NT - 1) {
    // the padding point: finite zeros for the weights read by the products
#pragma unroll
    for (int a = 0; a < DIM; a++) {
      s_v[NQ][a] = 0.0;
      s_vd[NQ][a] = 0.0;
      s_gT[NQ][a] = 0.0;
      s_f[NQ][a] = 0.0;
#pragma unroll
      for (int p = 0; p < DIM; p++) s_gv[NQ][p][a] = 0.0;
    }
    s_P[NQ] = s_mu[NQ] = s_dmu[NQ] = s_div[NQ] = s_Td[NQ] = 0.0;
  }
  __syncthreads();

  const double rho = A.rho, rCp = A.rho * A.Cp, kc = A.kcond;
  const double cm = A.x_old ? A.cmass : 0.0;
  JT *__restrict__ out = (JT *)A.val;
  const bool ow = A.overwrite != 0;
  const unsigned short *__restrict__ rk = A.rank + (size_t)e * 2 * NVE2;

  // 5a. residual rows
  for (int t = tid; t < NV * NVE + NP; t += NT) {
    int row;
    double R = 0.0;
    if (t < NV * NVE) {
      const int a = t / NVE, i = t - a * NVE;
      row = s_gun[a][i];
      if (a < DIM) {
#pragma unroll 1
        for (int q = 0; q < NQ; q++) {
          const double ph = s_Phi[q][i];
          double adv = 0.0, dif = 0.0;
#pragma unroll
          for (int p = 0; p < DIM; p++) {
            adv += s_v[q][p] * s_gv[q][p][a];
            const double Pi_ap = (a == p ? -s_P[q] : 0.0) + s_mu[q] * (s_gv[q][a][p] + s_gv[q][p][a]);
            dif += s_G[p][q][i] * Pi_ap;
          }
          R += s_dA[q] * (-rho * ph * adv - dif + ph * s_f[q][a] - rho * ph * s_vd[q][a]);
        }
      } else {
        const double sw = s_su[0];
#pragma unroll 1
        for (int q = 0; q < NQ; q++) {
          const double ph = s_Phi[q][i];
          double adv = 0.0, dif = 0.0;
#pragma unroll
          for (int p = 0; p < DIM; p++) {
            adv += s_v[q][p] * s_gT[q][p];
            dif += s_G[p][q][i] * s_gT[q][p];
          }
          double wi = ph;
          if (sw != 0.0) {
            double vgi = 0.0;
#pragma unroll
            for (int p = 0; p < DIM; p++) vgi += s_v[q][p] * s_G[p][q][i];
            wi += sw * vgi;
          }
          R += s_dA[q] * (-rCp * wi * adv - kc * dif + ph * A.hsrc - rCp * ph * s_Td[q]);
        }
        for (int k = 0; k < nqs; k++)
          for (int g = 0; g < NSP; g++) R += s_qw[k][g] * A.stab[(size_t)(s_qsd[k] * NSP + g) * SREC + i];
      }
      if (s_dir[a][i]) R = s_u[a][i] - A.dbc_val[row];
    } else {
      const int k = t - NV * NVE;
      row = s_gp[k];
#pragma unroll 1
      for (int q = 0; q < NQ; q++) R += s_dA[q] * tb[T::PSI + q * NP + k] * s_div[q];
      if (A.tau_pspg != 0.0 && k > 0) {
        // PSPG: grad psi_k . tau (rho v_dot + rho v.grad v + grad P - f) dA,
        // formed per point from the staged fields (four threads do this)
#pragma unroll 1
        for (int q = 0; q < NQ; q++) {
          double r = 0.0;
#pragma unroll
          for (int a = 0; a < DIM; a++) {
            double adv = 0.0, gP = 0.0;
#pragma unroll
            for (int p = 0; p < DIM; p++) adv += s_v[q][p] * s_gv[q][p][a];
#pragma unroll
            for (int l = 1; l < NP; l++) gP += s_p[l] * s_B[q][a][l - 1];
            r += s_B[q][a][k - 1] * (rho * (adv + s_vd[q][a]) + gP - s_f[q][a]);
          }
          R += A.tau_pspg * s_dA[q] * r;
        }
      }
      if (s_dirp[k]) R = s_p[k] - A.dbc_val[row];
    }
    A.resid[row] += R;
  }

  if (A.resid_only) return;

  // 5b. nodal blocks on the matrix cores: wave w computes rows 16 (w >> 1) ..
  // +15 and columns 16 (w & 1) .. +15 of every block
  {
    const int lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, kk = lane >> 4;
    const int ii = 16 * (wave >> 1) + r16;  // row operand index (A)
    const int jj = 16 * (wave & 1) + r16;   // column operand index (B)
#define GM_MFMA64(a, b, c) __builtin_amdgcn_mfma_f64_16x16x4f64((a), (b), (c), 0, 0, 0)
    const d4v z = {0.0, 0.0, 0.0, 0.0};
    // shared products, four independent chains per k-step: E(phi, dA, phi),
    // E(phi, dA, v.G), sum_p E(G_p, dA, G_p) and (Carreau) its mu-weighted
    // form, folded into the diagonal-block terms
    d4v A1 = z, Dadv = z, L1 = z, Lmu = z;
    const bool supg = EN && s_su[0] != 0.0;
#pragma unroll 1
    for (int q0 = 0; q0 < KP; q0 += 4) {
      const int q = q0 + kk;
      const double dA = s_dA[q];
      const double phA = s_Phi[q][ii] * dA, phB = s_Phi[q][jj];
      double vB = 0.0;
#pragma unroll
      for (int p = 0; p < DIM; p++) vB += s_v[q][p] * s_G[p][q][jj];
      A1 = GM_MFMA64(phA, phB, A1);
      Dadv = GM_MFMA64(phA, vB, Dadv);
#pragma unroll
      for (int p = 0; p < DIM; p++) {
        const double gA = s_G[p][q][ii] * dA, gB = s_G[p][q][jj];
        L1 = GM_MFMA64(gA, gB, L1);
        if constexpr (VISC == 1) Lmu = GM_MFMA64(gA * s_mu[q], gB, Lmu);
      }
    }
    if constexpr (VISC == 0) Lmu = A.mu0 * L1;
    const d4v Dm = -(rho * Dadv + Lmu + (rho * cm) * A1);
    const d4v Tt = -(rCp * Dadv + kc * L1 + (rCp * cm) * A1);

    // this lane's rows i = 16 (wave >> 1) + kk + 4 r, column j = jj: rank
    // words of the four node pairs
    const int j = jj;
    unsigned rmw[4], rew[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = 16 * (wave >> 1) + kk + 4 * r;
      const bool in = i < NVE && j < NVE;
      rmw[r] = in ? rk[i * NVE + j] : 0u;
      rew[r] = (in && EN) ? rk[NVE2 + i * NVE + j] : 0u;
    }
    // scatter of one block row a (put_dirichlet_in_matrix, load_lec)
    auto scatter_row = [&](int a, const d4v *Jr) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * (wave >> 1) + kk + 4 * r;
        if (i >= NVE || j >= NVE) continue;
        const int t = i * NVE + j;
        const bool first = ow && (rmw[r] & 0x8000u);
        const int row = s_gun[a][i];
        const bool dir = s_dir[a][i] != 0;
        const int base = s_brow[a][i] + (a < DIM ? (int)(rmw[r] & 0x7fffu) : (int)rew[r]);
#pragma unroll
        for (int b = 0; b < NV; b++) {
          double v = Jr[b][r];
          if (A.kel) A.kel[(size_t)e * A.E + (size_t)(a * NV + b) * NVE2 + t] = v;
          if (dir) {
            if (a == b && i == j) v = 1.0;
            else if (first) v = 0.0;
            else continue;
          }
          const int col = s_gun[b][j];
          const long long pos = row == col ? row : base + b - (col > row ? 1 : 0);
          if (first) out[pos] = (JT)v;
          else out[pos] += (JT)v;
        }
      }
    };
    // block row a of the momentum equation: three independent chains (b)
#pragma unroll 1
    for (int a = 0; a < DIM; a++) {
      d4v Jr[NV];
#pragma unroll
      for (int b = 0; b < DIM; b++) Jr[b] = a == b ? Dm : z;
#pragma unroll 1
      for (int q0 = 0; q0 < KP; q0 += 4) {
        const int q = q0 + kk;
        const double dA = s_dA[q];
        const double phA = s_Phi[q][ii] * (-rho * dA), phB = s_Phi[q][jj];
        const double gaB = s_G[a][q][jj];
        const double wmu = -dA * (VISC ? s_mu[q] : A.mu0);
        double hA = 0.0;
        if constexpr (VISC == 1) {
          // Carreau: (d mu / d gammadot) / gammadot (G_i . gamma_a) (g_j . gamma_b)
#pragma unroll
          for (int p = 0; p < DIM; p++) hA += s_G[p][q][ii] * (s_gv[q][a][p] + s_gv[q][p][a]);
          hA *= -dA * s_dmu[q];
        }
#pragma unroll
        for (int b = 0; b < DIM; b++) {
          // advection through d_b v_a, then mu g_j,a G_i,b
          Jr[b] = GM_MFMA64(phA * s_gv[q][b][a], phB, Jr[b]);
          Jr[b] = GM_MFMA64(s_G[b][q][ii] * wmu, gaB, Jr[b]);
          if constexpr (VISC == 1) {
            double hB = 0.0;
#pragma unroll
            for (int p = 0; p < DIM; p++) hB += s_G[p][q][jj] * (s_gv[q][b][p] + s_gv[q][p][b]);
            Jr[b] = GM_MFMA64(hA, hB, Jr[b]);
          }
        }
      }
      if constexpr (EN) Jr[DIM] = (-rho * A.g[a] * A.beta) * A1;
      scatter_row(a, Jr);
    }
    if constexpr (EN) {
      d4v Jr[NV];
#pragma unroll
      for (int b = 0; b < DIM; b++) Jr[b] = z;
      const double sw = s_su[0];
      Jr[DIM] = Tt;
      if (supg) {
#pragma unroll 1
        for (int q0 = 0; q0 < KP; q0 += 4) {
          const int q = q0 + kk;
          const double w = -rCp * s_dA[q], phB = s_Phi[q][jj];
          // (phi_i + sw v.G_i) d_b T + sw (v.grad T) G_b,i: the SUPG weight on
          // the advection and its derivative in v at the column node; and
          // the weight's part of J_TT, -rho Cp sw E(v.G_i dA, v.G_j)
          double vA = 0.0, vgT = 0.0, vB = 0.0;
#pragma unroll
          for (int p = 0; p < DIM; p++) {
            vA += s_v[q][p] * s_G[p][q][ii];
            vgT += s_v[q][p] * s_gT[q][p];
            vB += s_v[q][p] * s_G[p][q][jj];
          }
          const double wi = s_Phi[q][ii] + sw * vA;
#pragma unroll
          for (int b = 0; b < DIM; b++)
            Jr[b] = GM_MFMA64(w * (wi * s_gT[q][b] + sw * vgT * s_G[b][q][ii]), phB, Jr[b]);
          Jr[DIM] = GM_MFMA64(w * sw * vA, vB, Jr[DIM]);
        }
      } else {
#pragma unroll 1
        for (int q0 = 0; q0 < KP; q0 += 4) {
          const int q = q0 + kk;
          const double phA = s_Phi[q][ii] * (-rCp * s_dA[q]), phB = s_Phi[q][jj];
#pragma unroll
          for (int b = 0; b < DIM; b++) Jr[b] = GM_MFMA64(phA * s_gT[q][b], phB, Jr[b]);
        }
      }
      scatter_row(DIM, Jr);
    }
#undef GM_MFMA64
  }

  // energy SUPG: d(1/h)/dv at the last local node adds
  // -rho Cp s_d,b sum_q dA (v.grad T) v_b G_b,i to the energy rows' columns of
  // that node's velocities, on top of the entries scattered above
  if (EN && s_su[0] != 0.0) {
    __syncthreads();
    for (int t = tid; t < DIM * NVE; t += NT) {
      const int b = t / NVE, i = t - b * NVE;
      double c = 0.0;
#pragma unroll 1
      for (int q = 0; q < NQ; q++) {
        double vgT = 0.0;
#pragma unroll
        for (int p = 0; p < DIM; p++) vgT += s_v[q][p] * s_gT[q][p];
        c += s_dA[q] * vgT * s_v[q][b] * s_G[b][q][i];
      }
      const double v = -rCp * s_su[1 + b] * c;
      const int ij = i * NVE + (NVE - 1);
      if (A.kel) A.kel[(size_t)e * A.E + (size_t)(DIM * NV + b) * NVE2 + ij] += v;
      if (s_dir[DIM][i]) continue;  // a Dirichlet row keeps its unit diagonal only
      const int row = s_gun[DIM][i], col = s_gun[b][NVE - 1];
      const long long pos = s_brow[DIM][i] + (int)rk[NVE2 + ij] + b - (col > row ? 1 : 0);
      out[pos] += (JT)v;
    }
  }

  // QCONV: -h w |surface Jacobian| phi_i phi_j over the element's sides in the
  // T-T block, on top of the entries scattered above
  if (EN && nqs > 0) {
    bool conv = false;
    for (int k = 0; k < nqs; k++) conv |= A.qs_q[3 * (qs0 + k) + 1] != 0.0;
    if (conv) {
      __syncthreads();
      for (int t = tid; t < NVE2; t += NT) {
        const int i = t / NVE, j = t - i * NVE;
        double v = 0.0;
#pragma unroll 1
        for (int k = 0; k < nqs; k++) {
          if (A.qs_q[3 * (qs0 + k) + 1] == 0.0) continue;
#pragma unroll 1
          for (int g = 0; g < NSP; g++) {
            const double *st = A.stab + (size_t)(s_qsd[k] * NSP + g) * SREC;
            v -= s_qh[k][g] * st[i] * st[j];
          }
        }
        if (v == 0.0) continue;  // a node pair off the sides
        if (A.kel) A.kel[(size_t)e * A.E + (size_t)(DIM * NV + DIM) * NVE2 + t] += v;
        if (s_dir[DIM][i]) continue;  // a Dirichlet row keeps its unit diagonal only
        const int row = s_gun[DIM][i], col = s_gun[DIM][j];
        const long long pos =
            row == col ? row : s_brow[DIM][i] + (int)rk[NVE2 + t] + DIM - (col > row ? 1 : 0);
        out[pos] += (JT)v;
      }
    }
  }

  // 5c. momentum x pressure (J_m_P): int grad_a phi_i psi_k dA
  for (int t = tid; t < NVE * NP; t += NT) {
    const int i = t / NP, k = t - i * NP;
    double acc[DIM] = {0.0, 0.0, 0.0};
#pragma unroll 1
    for (int q = 0; q < NQ; q++) {
      const double w = s_dA[q] * tb[T::PSI + q * NP + k];
#pragma unroll
      for (int a = 0; a < DIM; a++) acc[a] += w * s_G[a][q][i];
    }
    const unsigned rm = rk[i * NVE + CL];
    const bool first = ow && (rm & 0x8000u);
    const int col = s_gp[k];
#pragma unroll
    for (int a = 0; a < DIM; a++) {
      if (A.kel) A.kel[(size_t)e * A.E + A.off_mp + a * NVE * NP + t] = acc[a];
      double v = acc[a];
      if (s_dir[a][i]) {
        if (!first) continue;
        v = 0.0;
      }
      const int row = s_gun[a][i];
      const long long pos = s_brow[a][i] + (int)(rm & 0x7fffu) + NV + k - (col > row ? 1 : 0);
      if (first) out[pos] = (JT)v;
      else out[pos] += (JT)v;
    }
  }

  // 5d. continuity x velocity (J_c_v): int psi_k grad_b phi_j dA; PSPG adds
  // tau rho [phi_j grad psi_k . d_b v + grad_b psi_k (v.grad phi_j + c phi_j)]
  // and continuity x temperature tau rho beta (grad psi_k . g) phi_j
  const bool ps = A.tau_pspg != 0.0;
  for (int t = tid; t < NP * NVE; t += NT) {
    const int k = t / NVE, j = t - k * NVE;
    const bool dir = s_dirp[k] != 0;
    double acc[DIM] = {0.0, 0.0, 0.0}, accT = 0.0;
#pragma unroll 1
    for (int q = 0; q < NQ; q++) {
      const double w = s_dA[q] * tb[T::PSI + q * NP + k];
#pragma unroll
      for (int b = 0; b < DIM; b++) acc[b] += w * s_G[b][q][j];
    }
    if (ps && k > 0) {
#pragma unroll 1
      for (int q = 0; q < NQ; q++) {
        const double wp = A.tau_pspg * s_dA[q] * rho, phj = s_Phi[q][j];
        double vg = 0.0, gk[DIM], gkg = 0.0;
#pragma unroll
        for (int p = 0; p < DIM; p++) {
          vg += s_v[q][p] * s_G[p][q][j];
          gk[p] = s_B[q][p][k - 1];
          gkg += gk[p] * A.g[p];
        }
#pragma unroll
        for (int b = 0; b < DIM; b++) {
          double gv = 0.0;
#pragma unroll
          for (int a = 0; a < DIM; a++) gv += gk[a] * s_gv[q][b][a];
          acc[b] += wp * (phj * gv + gk[b] * (vg + cm * phj));
        }
        accT += wp * A.beta * gkg * phj;
      }
    }
    const unsigned rm = rk[CL * NVE + j];
    const bool first = ow && (rm & 0x8000u);
    const int row = s_gp[k];
#pragma unroll
    for (int b = 0; b < DIM; b++) {
      if (A.kel) A.kel[(size_t)e * A.E + A.off_cv + b * NP * NVE + t] = acc[b];
      double v = acc[b];
      if (dir) {
        if (!first) continue;
        v = 0.0;
      }
      const int col = s_gun[b][j];
      const long long pos = s_bp[k] + (int)(rm & 0x7fffu) + b - (col > row ? 1 : 0);
      if (first) out[pos] = (JT)v;
      else out[pos] += (JT)v;
    }
    // continuity x temperature: in the graph (the continuity rows couple to
    // every unknown of the element's nodes); zero in the Galerkin weak form,
    // PSPG's Boussinesq term otherwise
    if (EN && (first || (ps && !dir))) {
      const int col = s_gun[DIM][j];
      const long long pos = s_bp[k] + (int)(rm & 0x7fffu) + DIM - (col > row ? 1 : 0);
      const double v = dir ? 0.0 : accT;
      if (first) out[pos] = (JT)v;
      else out[pos] += (JT)v;
    }
  }

  // 5e. continuity x pressure: zero for Galerkin Q2-P1; PSPG adds
  // tau grad psi_k . grad psi_l; a unit diagonal on a Dirichlet (datum) row
  if (tid < NP * NP) {
    const int k = tid / NP, k2 = tid - k * NP;
    double acc = 0.0;
    if (ps && k > 0 && k2 > 0)
#pragma unroll 1
      for (int q = 0; q < NQ; q++) {
        double g2 = 0.0;
#pragma unroll
        for (int a = 0; a < DIM; a++) g2 += s_B[q][a][k - 1] * s_B[q][a][k2 - 1];
        acc += A.tau_pspg * s_dA[q] * g2;
      }
    if (A.kel) A.kel[(size_t)e * A.E + A.off_cp + tid] = acc;
    const unsigned rm = rk[CL * NVE + CL];
    const bool first = ow && (rm & 0x8000u);
    const double v = s_dirp[k] ? (k == k2 ? 1.0 : 0.0) : acc;
    const long long pos = k == k2 ? (long long)s_gp[k] : s_bp[k] + (int)(rm & 0x7fffu) + NV + k2 - (k2 > k ? 1 : 0);
    if (first) out[pos] = (JT)v;
    else if (v != 0.0) out[pos] += (JT)v;
  }
}

// Rank map of every element: for node pair (i, j) (local order, i row node),
//   rank[e][0][i*nve + j] = number of columns below fu[n_j] in the full column
//     set (off-diagonals and the diagonal) of the momentum rows of node n_i,
//     with bit 15 set when e is the first element, in colour order, holding
//     both nodes (filled in by k_first_touch);
//   rank[e][1][i*nve + j] = the same for the energy row (no pressure columns).
// Every momentum row of a node has the same full column set, so the off-
// diagonal position of column fu[n_j] + b in row fu[n_i] + a is
// bindx[row] + rank + b - (col > row).  Continuity rows live on the centroid
// node, whose neighbours are the element's own nodes, so they share the
// column set of the centroid's momentum rows.
template <int DIM>
__global__ __launch_bounds__(256) void k_build_ranks(const int *__restrict__ conn,
                                                     const int *__restrict__ fu,
                                                     const int *__restrict__ bindx, int en,
                                                     unsigned short *__restrict__ rank, int *flag) {
  constexpr int NVE = El<DIM>::NVE, NVE2 = NVE * NVE;
  const int e = blockIdx.x;
  __shared__ int s_f[NVE];
  if (threadIdx.x < NVE) s_f[threadIdx.x] = fu[conn[(size_t)e * NVE + threadIdx.x]];
  __syncthreads();
  for (int t = threadIdx.x; t < NVE2; t += blockDim.x) {
    const int i = t / NVE, j = t - i * NVE;
    const int col = s_f[j];
    for (int w = 0; w < 1 + en; w++) {
      const int row = s_f[i] + (w ? DIM : 0);
      // lower bound of col in the sorted off-diagonal list of row
      int lo = bindx[row], hi = bindx[row + 1];
      const int start = lo;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (bindx[mid] < col) lo = mid + 1;
        else hi = mid;
      }
      const int r = lo - start + (row < col ? 1 : 0);
      if (r > 0x7fff) atomicOr(&flag[1], 1);
      rank[(size_t)e * 2 * NVE2 + w * NVE2 + t] = (unsigned short)r;
    }
    if (!en) rank[(size_t)e * 2 * NVE2 + NVE2 + t] = 0;
  }
}

// First touch, one element colour per launch (colours in order): the element
// of this colour holding both nodes of a pair marks it when no earlier colour
// has; `touched` is indexed by the MSR position of the pair's (U1, U1) entry.
template <int DIM>
__global__ __launch_bounds__(256) void k_first_touch(const int *__restrict__ elems,
                                                     const int *__restrict__ conn,
                                                     const int *__restrict__ fu,
                                                     const int *__restrict__ bindx,
                                                     unsigned short *__restrict__ rank,
                                                     unsigned char *__restrict__ touched,
                                                     int *flag) {
  constexpr int NVE = El<DIM>::NVE, NVE2 = NVE * NVE;
  const int e = elems[blockIdx.x];
  __shared__ int s_f[NVE];
  if (threadIdx.x < NVE) s_f[threadIdx.x] = fu[conn[(size_t)e * NVE + threadIdx.x]];
  __syncthreads();
  for (int t = threadIdx.x; t < NVE2; t += blockDim.x) {
    const int i = t / NVE, j = t - i * NVE;
    const int row = s_f[i], col = s_f[j];
    unsigned short *rp = rank + (size_t)e * 2 * NVE2 + t;
    const int r = *rp & 0x7fff;
    long long pos = row;
    if (row != col) {
      pos = (long long)bindx[row] + r - (col > row ? 1 : 0);
      if (pos < bindx[row] || pos >= bindx[row + 1] || bindx[pos] != col) {
        atomicOr(&flag[1], 1);
        continue;
      }
    }
    if (!touched[pos]) {
      touched[pos] = 1;
      *rp = (unsigned short)(r | 0x8000);
    }
  }
}

// Element-to-nonzero map: for every local entry of every element, the MSR
// index of the global (row, column) pair; the diagonal lives at index row.
template <int DIM>
__global__ __launch_bounds__(256) void k_build_e2nz(const int *__restrict__ conn,
                                                    const int *__restrict__ fu,
                                                    const int *__restrict__ centroid,
                                                    const int *__restrict__ bindx, int en,
                                                    long long E, long long off_mp,
                                                    long long off_cv, long long off_cp,
                                                    int *__restrict__ e2nz, int *flag) {
  constexpr int NVE = El<DIM>::NVE, NP = El<DIM>::NP;
  const int NV = DIM + en;
  const int e = blockIdx.x;
  __shared__ int s_f[NVE];
  __shared__ int s_pc;
  if (threadIdx.x < NVE) s_f[threadIdx.x] = fu[conn[(size_t)e * NVE + threadIdx.x]];
  if (threadIdx.x == 0) s_pc = fu[centroid[e]] + DIM + en;
  __syncthreads();
  for (long long idx = threadIdx.x; idx < E; idx += blockDim.x) {
    int row, col;
    if (idx < off_mp) {
      const int blk = (int)(idx / (NVE * NVE)), t = (int)(idx - (long long)blk * NVE * NVE);
      const int a = blk / NV, b = blk - a * NV;
      row = s_f[t / NVE] + a;
      col = s_f[t % NVE] + b;
    } else if (idx < off_cv) {
      const int r = (int)(idx - off_mp);
      const int a = r / (NVE * NP), t = r - a * NVE * NP;
      row = s_f[t / NP] + a;
      col = s_pc + t % NP;
    } else if (idx < off_cp) {
      const int r = (int)(idx - off_cv);
      const int b = r / (NP * NVE), t = r - b * NP * NVE;
      row = s_pc + t / NVE;
      col = s_f[t % NVE] + b;
    } else {
      const int r = (int)(idx - off_cp);
      row = s_pc + r / NP;
      col = s_pc + r % NP;
    }
    int pos;
    if (row == col) {
      pos = row;
    } else {
      int lo = bindx[row], hi = bindx[row + 1] - 1;
      pos = -1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = bindx[mid];
        if (v == col) {
          pos = mid;
          break;
        }
        if (v < col) lo = mid + 1;
        else hi = mid - 1;
      }
      if (pos < 0) atomicOr(&flag[1], 1);
    }
    e2nz[(size_t)e * E + idx] = pos;
  }
}

// Reference basis tables on the host.  QUAD9 closed forms of
// src/rf_shape.c:361-470; HEX27 as products of 1D quadratic Lagrange factors
// (identical to the closed forms of src/rf_shape.c:1105-1460); Gauss points
// ordered as find_stu (src/el_elm_info.c:1748-1823).
static const signed char kHex27[27][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
    {-1, 1, 1},   {0, -1, -1}, {1, 0, -1}, {0, 1, -1},  {-1, 0, -1}, {-1, -1, 0}, {1, -1, 0},
    {1, 1, 0},    {-1, 1, 0},  {0, -1, 1}, {1, 0, 1},   {0, 1, 1},   {-1, 0, 1},  {0, 0, 0},
    {0, 0, -1},   {0, 0, 1},   {-1, 0, 0}, {1, 0, 0},   {0, -1, 0},  {0, 1, 0}};
static const signed char kQuad9[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1},
                                        {1, 0},   {0, 1},  {-1, 0}, {0, 0}};

static void lagrange1(int node, double s, double &v, double &dv) {
  if (node < 0) {
    v = -0.5 * s * (1.0 - s);
    dv = -0.5 + s;
  } else if (node > 0) {
    v = 0.5 * s * (1.0 + s);
    dv = 0.5 + s;
  } else {
    v = 1.0 - s * s;
    dv = -2.0 * s;
  }
}

// Q2 basis (QUAD9 / HEX27, Goma local node order) at one reference point.
void gm_q2_basis(int dim, const double *xi, double *phi) {
  const int nve = dim == 2 ? 9 : 27;
  for (int k = 0; k < nve; k++) {
    const signed char *pos = dim == 2 ? kQuad9[k] : kHex27[k];
    double v = 1.0, a, da;
    for (int d = 0; d < dim; d++) {
      lagrange1(pos[d], xi[d], a, da);
      v *= a;
    }
    phi[k] = v;
  }
}

static std::vector<double> make_tables(int dim) {
  const int nve = dim == 2 ? 9 : 27, nq = nve, np = dim + 1;
  std::vector<double> t((size_t)nq * nve + (size_t)nq * nve * dim + (size_t)nq * np + nq);
  double *phi = t.data(), *dphi = phi + nq * nve, *psi = dphi + nq * nve * dim, *wt = psi + nq * np;
  const double F2 = 0.77459666924148340428, w1 = 0.55555555555555555556, w2 = 0.88888888888888888888;
  const double pt[3] = {F2, 0.0, -F2}, ww[3] = {w1, w2, w1};
  for (int q = 0; q < nq; q++) {
    const int is = q % 3, it = (q % 9) / 3, iu = q / 9;
    const double s = pt[is], tt = pt[it], u = dim == 3 ? pt[iu] : 0.0;
    wt[q] = ww[is] * ww[it] * (dim == 3 ? ww[iu] : 1.0);
    for (int k = 0; k < nve; k++) {
      double a, da, b, db, cc = 1.0, dc = 0.0;
      const signed char *pos = dim == 2 ? kQuad9[k] : kHex27[k];
      lagrange1(pos[0], s, a, da);
      lagrange1(pos[1], tt, b, db);
      if (dim == 3) lagrange1(pos[2], u, cc, dc);
      phi[q * nve + k] = a * b * cc;
      dphi[(q * nve + k) * dim + 0] = da * b * cc;
      dphi[(q * nve + k) * dim + 1] = a * db * cc;
      if (dim == 3) dphi[(q * nve + k) * dim + 2] = a * b * dc;
    }
    // P1 discontinuous basis {1, s, t[, u]} (src/mm_fill_util.c:3802-3858, 4205-4276)
    psi[q * np + 0] = 1.0;
    psi[q * np + 1] = s;
    psi[q * np + 2] = tt;
    if (dim == 3) psi[q * np + 3] = u;
  }
  return t;
}

// Side tables (STab layout): the Q2 basis and its reference derivatives at
// the 3 (2D) or 3 x 3 (3D) Gauss points of each side, sides in Exodus order
// (get_side_info, src/mm_fill_fill.c:3049-3072, 3182-3250): the fixed
// reference axis and value of sides 1..2*dim are (t,-1) (s,+1) (t,+1) (s,-1)
// (u,-1) (u,+1); the free axes run in ascending order, the first slowest.
static std::vector<double> make_side_tables(int dim) {
  const int nve = dim == 2 ? 9 : 27, nside = 2 * dim, nsp = dim == 2 ? 3 : 9;
  const int rec = nve + nve * dim + 1;
  static const int axis[6] = {1, 0, 1, 0, 2, 2};
  static const double fixv[6] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0};
  const double F2 = 0.77459666924148340428, w1 = 0.55555555555555555556, w2 = 0.88888888888888888888;
  const double gp[3] = {-F2, 0.0, F2}, gw[3] = {w1, w2, w1};
  std::vector<double> t((size_t)nside * nsp * rec);
  for (int sd = 0; sd < nside; sd++) {
    int fr[2], nf = 0;
    for (int a = 0; a < dim; a++)
      if (a != axis[sd]) fr[nf++] = a;
    for (int g = 0; g < nsp; g++) {
      const int g1 = dim == 2 ? g : g / 3, g2 = dim == 2 ? 0 : g % 3;
      double ref[3] = {0.0, 0.0, 0.0};
      ref[axis[sd]] = fixv[sd];
      ref[fr[0]] = gp[g1];
      double w = gw[g1];
      if (dim == 3) {
        ref[fr[1]] = gp[g2];
        w *= gw[g2];
      }
      double *r = t.data() + (size_t)(sd * nsp + g) * rec;
      for (int k = 0; k < nve; k++) {
        const signed char *pos = dim == 2 ? kQuad9[k] : kHex27[k];
        double v[3], dv[3];
        for (int d = 0; d < dim; d++) lagrange1(pos[d], ref[d], v[d], dv[d]);
        double ph = 1.0;
        for (int d = 0; d < dim; d++) ph *= v[d];
        r[k] = ph;
        for (int d = 0; d < dim; d++) {
          double dd = dv[d];
          for (int e = 0; e < dim; e++)
            if (e != d) dd *= v[e];
          r[nve + k * dim + d] = dd;
        }
      }
      r[nve + nve * dim] = w;
    }
  }
  return t;
}

int gm_fill_setup_device(gm_ctx *c) {
  ElemLayout &L = c->lay;
  L.dim = c->dim;
  L.nve = c->nve;
  L.np = c->np;
  L.nv = c->nv;
  L.nq = c->nve;
  const long long nve2 = (long long)L.nve * L.nve;
  L.off_mp = (long long)L.nv * L.nv * nve2;
  L.off_cv = L.off_mp + (long long)L.dim * L.nve * L.np;
  L.off_cp = L.off_cv + (long long)L.dim * L.np * L.nve;
  L.total = L.off_cp + (long long)L.np * L.np;

  std::vector<double> tab = make_tables(c->dim);
  GM_HIP(hipMalloc(&c->d_tables, tab.size() * sizeof(double)));
  GM_HIP(hipMemcpy(c->d_tables, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> stab = make_side_tables(c->dim);
  GM_HIP(hipMalloc(&c->d_stab, stab.size() * sizeof(double)));
  GM_HIP(hipMemcpy(c->d_stab, stab.data(), stab.size() * sizeof(double), hipMemcpyHostToDevice));
  const size_t ne = (size_t)c->num_elems;
  GM_HIP(hipMalloc(&c->d_e2nz, ne * L.total * sizeof(int)));
  GM_HIP(hipMemsetAsync(c->d_flag, 0, 4 * sizeof(int), c->stream));
  if (c->dim == 2)
    k_build_e2nz<2><<<c->num_elems, 256, 0, c->stream>>>(c->d_conn, c->d_fu, c->d_centroid, c->d_bindx,
                                                        c->energy, L.total, L.off_mp, L.off_cv,
                                                        L.off_cp, c->d_e2nz, c->d_flag);
  else
    k_build_e2nz<3><<<c->num_elems, 256, 0, c->stream>>>(c->d_conn, c->d_fu, c->d_centroid, c->d_bindx,
                                                        c->energy, L.total, L.off_mp, L.off_cv,
                                                        L.off_cp, c->d_e2nz, c->d_flag);
  GM_HIP(hipGetLastError());
  int flag[4];
  GM_HIP(hipMemcpyAsync(flag, c->d_flag, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  GM_HIP(hipStreamSynchronize(c->stream));
  if (flag[1]) {
    gm_set_error("element-to-nonzero map: entry missing from the MSR graph");
    return GM_ERR_ARG;
  }
  if (c->dim == 3) {
    // rank map and first touches (the HEX27 assembly kernel's scatter)
    const size_t nr = ne * 2 * (size_t)L.nve * L.nve;
    GM_HIP(hipMalloc(&c->d_rank, nr * sizeof(unsigned short)));
    k_build_ranks<3><<<c->num_elems, 256, 0, c->stream>>>(c->d_conn, c->d_fu, c->d_bindx, c->energy,
                                                         c->d_rank, c->d_flag);
    GM_HIP(hipGetLastError());
    unsigned char *touched = nullptr;
    GM_HIP(hipMalloc(&touched, (size_t)c->nval));
    GM_HIP(hipMemsetAsync(touched, 0, (size_t)c->nval, c->stream));
    for (int col = 0; col < c->n_ecol; col++) {
      const int cnt = c->ecol_ptr[col + 1] - c->ecol_ptr[col];
      if (cnt)
        k_first_touch<3><<<cnt, 256, 0, c->stream>>>(c->d_ecol_elems + c->ecol_ptr[col], c->d_conn,
                                                    c->d_fu, c->d_bindx, c->d_rank, touched,
                                                    c->d_flag);
    }
    const hipError_t le = hipGetLastError();
    GM_HIP(hipMemcpyAsync(flag, c->d_flag, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GM_HIP(hipStreamSynchronize(c->stream));
    (void)hipFree(touched);
    GM_HIP(le);
    if (flag[1]) {
      gm_set_error("rank map: a node pair's position does not match the MSR graph");
      return GM_ERR_ARG;
    }
  }
  return GM_OK;
}

template <int DIM, int EN, int VISC, typename JT>
static void launch_fill(gm_ctx *c, FillArgs &a) {
  constexpr int NT = DIM == 2 ? 128 : 256;
  void (*kern)(FillArgs, const int *);
  if constexpr (DIM == 3) kern = k_fill3m<EN, VISC, JT>;
  else kern = k_fill<DIM, EN, VISC, JT, NT>;
  for (int col = 0; col < c->n_ecol; col++) {
    const int cnt = c->ecol_ptr[col + 1] - c->ecol_ptr[col];
    if (cnt == 0) continue;
    // this colour's share of B_asm = nnz*s + 16 N + 8 dim n_nodes + 4 nve n_el
    const double share = (double)cnt / c->num_elems;
    const double basm = (double)(c->nval - 1) * sizeof(JT) + 16.0 * c->N +
                        8.0 * DIM * c->num_nodes + 4.0 * El<DIM>::NVE * c->num_elems;
    const int id = a.resid_only ? GM_K_FILL_RESID : GM_K_FILL;
    gm_launch(c, id, share * (a.resid_only ? basm - (double)(c->nval - 1) * sizeof(JT) : basm),
              kern, dim3(cnt), dim3(NT), 0, c->stream, a, c->d_ecol_elems + c->ecol_ptr[col]);
  }
}

template <int DIM, typename JT>
static void dispatch_fill(gm_ctx *c, FillArgs &a) {
  const int en = c->energy, vm = c->phys.visc_model;
  if (en && vm) launch_fill<DIM, 1, 1, JT>(c, a);
  else if (en) launch_fill<DIM, 1, 0, JT>(c, a);
  else if (vm) launch_fill<DIM, 0, 1, JT>(c, a);
  else launch_fill<DIM, 0, 0, JT>(c, a);
}

int gm_fill_launch(gm_ctx *c, const double *x, double *resid, double *val, float *val32) {
  return gm_fill_launch_ex(c, x, resid, val, val32, 0);
}

int gm_fill_launch_ex(gm_ctx *c, const double *x, double *resid, double *val, float *val32,
                      int resid_only) {
  FillArgs a;
  a.resid_only = resid_only;
  a.coor = c->d_coor;
  a.conn = c->d_conn;
  a.fu = c->d_fu;
  a.centroid = c->d_centroid;
  a.dbc = c->d_dbc;
  a.dbc_val = c->d_dbc_val;
  a.tables = c->d_tables;
  a.e2nz = c->d_e2nz;
  a.x = x;
  a.resid = resid;
  a.val = val32 ? (void *)val32 : (void *)val;
  a.kel = c->kel_on ? c->d_kel : nullptr;
  a.flag = c->d_flag;
  a.x_old = c->tr_x_old;
  a.xdot_old = c->tr_xdot_old;
  a.cmass = c->tr_cmass;
  a.cold = c->tr_cold;
  a.qs_ptr = c->qs_side.empty() ? nullptr : c->d_qs_ptr;
  a.qs_side = c->d_qs_side;
  a.qs_q = c->d_qs_q;
  a.stab = c->d_stab;
  a.bindx = c->d_bindx;
  a.rank = c->d_rank;
  a.overwrite = c->fill_overwrite && c->d_rank ? 1 : 0;
  a.num_nodes = c->num_nodes;
  a.tau_pspg = c->phys.pspg ? gm_pspg_tau(c) : 0.0;
  a.supg_e = c->energy ? c->phys.energy_supg : 0.0;
  a.E = c->lay.total;
  a.off_mp = c->lay.off_mp;
  a.off_cv = c->lay.off_cv;
  a.off_cp = c->lay.off_cp;
  const gm_physics &p = c->phys;
  a.rho = p.rho;
  a.mu0 = p.mu0;
  a.muinf = p.muinf;
  a.lam = p.lam;
  a.nexp = p.nexp;
  a.aexp = p.aexp;
  a.Cp = p.Cp;
  a.kcond = p.kcond;
  a.hsrc = p.hsrc;
  for (int k = 0; k < 3; k++) a.g[k] = p.g[k];
  a.beta = p.beta;
  a.Tref = p.Tref;
  GM_HIP(hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
  if (c->dim == 2) {
    if (val32) dispatch_fill<2, float>(c, a);
    else dispatch_fill<2, double>(c, a);
  } else {
    if (val32) dispatch_fill<3, float>(c, a);
    else dispatch_fill<3, double>(c, a);
  }
  GM_HIP(hipGetLastError());
  int flag = 0;
  GM_HIP(hipMemcpyAsync(&flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  GM_HIP(hipStreamSynchronize(c->stream));
  if (flag) {
    gm_set_error("zero or negative element Jacobian determinant");
    return GM_ERR_ELEM;
  }
  return GM_OK;
}
