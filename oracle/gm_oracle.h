/*
 * gm_oracle.h -- FP64 CPU restatement of Goma's Newton inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product (goma_amd/, libgm.so) never links or calls it.
 *
 * It restates, in plain C and double precision, the reference path:
 *   solve_nonlinear_problem   src/mm_sol_nonlinear.c:596-2913
 *   matrix_fill_full          src/mm_fill.c:158-312  (element loop)
 *   matrix_fill               src/mm_fill.c:317-3760 (one element)
 *   assemble_momentum         src/mm_fill_momentum.c:100-2474
 *   fluid_stress              src/mm_fill_momentum.c:2834-3737
 *   assemble_continuity       src/mm_fill_continuity.c:119-1401
 *   assemble_energy           src/mm_fill_energy.c:109-1152, heat_flux 1629
 *   carreau_viscosity         src/mm_viscosity.c:981-1119
 *   calc_shearrate            src/mm_fill_util.c:4564-4660
 *   bouss_momentum_source     src/mm_std_models.c:125-284
 *   beer_belly                src/mm_fill_util.c:140-634
 *   shape (QUAD9 / HEX27)     src/rf_shape.c:361-470, 1105-1460
 *   P1 basis                  src/mm_fill_util.c:3802-3858, 4205-4276
 *   find_stu / Gq_weight      src/el_elm_info.c:1615-1830, 3469-3690
 *   put_dirichlet_in_matrix   src/bc_dirich.c:44-153
 *   load_lec (MSR)            src/mm_fill.c:5241-5483, in_list include/el_elm_info.h:110
 *   find_MSR_problem_graph    src/mm_fill_util.c:3229-3445
 *   build_node_node           src/exo_conn.c:204-377
 *   row_sum_scale_MSR         src/sl_matrix_util.c:506-600
 *   QSIDE weak heat flux      src/bc_integ.c:1518-1519, 2156-2157, 2207
 *   transient mass terms      src/mm_fill_momentum.c:534-539, 1586-1590;
 *                             src/mm_fill_energy.c:322-328, 435-441
 *   L2/L1/Loo norms           src/mm_sol_nonlinear.c:3177-3420
 * GMRES(m)/ILU(0) restate the AztecOO behaviour Goma selects
 * (src/sl_util.c:392-960); AztecOO itself is not in the reference tree.  They
 * are pinned by what defines them: each GMRES iterate is the minimal-residual
 * point of its restart cycle's (right-preconditioned) Krylov space, and the
 * ILU(0) factors reproduce the matrix on its pattern (test_oracle_gmres.py).
 *
 * The pins of this restatement are the tests in tests/test_oracle_*.py.
 */
#ifndef GM_ORACLE_H
#define GM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Goma variable types (include/rf_fem_const.h:174-200). */
#define OM_U1 0
#define OM_U2 1
#define OM_U3 2
#define OM_T 3
#define OM_P 9

typedef struct om_mesh {
  int dim;            /* 2: QUAD9, 3: HEX27 */
  int num_nodes;
  int num_elems;
  const double *coor; /* [dim][num_nodes], like Goma's Coor */
  const int *connect; /* [num_elems][9 or 27], Goma local node order */
} om_mesh;

typedef struct om_physics {
  int energy;     /* 1: energy equation and temperature unknown active */
  int visc_model; /* 0: Newtonian (mu0), 1: Carreau(-Yasuda) */
  double rho;
  double mu0, muinf, lam, nexp, aexp;
  double Cp, kcond, hsrc;
  double g[3];    /* momentum_source (gravity) */
  double beta;    /* Volume_Expansion */
  double Tref;
  int pspg;           /* Pressure Stabilization: 0 no, 1 global (PSPG = 1) */
  double ps_scaling;  /* Pressure Stabilization Scaling */
  double energy_supg; /* energy weight function SUPG coefficient (Ewt_func), 0: Galerkin */
} om_physics;

typedef struct om_bc {
  int num_dbc;        /* Dirichlet conditions (node, variable type, value) */
  const int *node;
  const int *var;
  const double *value;
  int datum_elem;     /* pressure datum element, -1 for none */
  double datum_value;
} om_bc;

typedef struct om_solver_opts {
  int max_newton;     /* Max_Newton_Steps */
  double eps_resid;   /* Epsilon[0], Normalized Residual Tolerance */
  double eps_corr;    /* Epsilon[2], Normalized Correction Tolerance (1e10 default) */
  double damp;        /* damp_factor (1: full Newton step) */
  int kspace;         /* GMRES restart length */
  int max_its;        /* GMRES max iterations per Newton step */
  double lin_tol;     /* ||r||/||r0|| tolerance (AZ_r0) */
  int precond;        /* 0: none, 1: ILU(0) natural order, 2: ILU(0) node-colour order,
                         b+2 (b >= 1): ILU(0) in patch order with b patch bins per axis */
  int line_search;    /* 0: NLS_FULL_STEP, 1: NLS_BACKTRACK */
  double min_damp;    /* Line_Search_Minimum_Damping */
} om_solver_opts;

typedef struct om_report {
  int newton_its;     /* value Goma returns: number of passes through the loop */
  int converged;
  int lin_its_total;
  int lin_its[64];
  double resid_L2[64];   /* row-sum-scaled ||R||_2 at each pass */
  double resid_Loo[64];
  double unscaled_L2[64];
  double t_assemble, t_scale, t_precond, t_solve, t_total; /* seconds */
  double damp[64];
} om_report;

typedef struct om_problem om_problem;

/* Q2 shape functions phi[nve] and dphi[nve][3] at (s,t,u); Gauss point iq
 * of the 3-point tensor rule: stuw = {s, t, u, weight}. */
void om_shape(int dim, double s, double t, double u, double *phi, double *dphi);
void om_gauss(int dim, int iq, double *stuw);

om_problem *om_create(const om_mesh *mesh, const om_physics *phys, const om_bc *bc);
void om_destroy(om_problem *p);

int om_num_unknowns(const om_problem *p);
int om_ija_size(const om_problem *p); /* ija[N] == length of ija and val */
void om_copy_ija(const om_problem *p, int *ija_out);
void om_copy_first_unknown(const om_problem *p, int *fu_out); /* [num_nodes+1] */
void om_copy_unknown_var(const om_problem *p, int *var_out);  /* [N] variable type */
void om_copy_dbc(const om_problem *p, char *mask_out, double *val_out);

/* matrix_fill_full: resid and val are ACCUMULATED (caller zeroes them).
 * Returns 0, or -1 on a zero or negative element Jacobian determinant. */
int om_fill(const om_problem *p, const double *x, double *resid, double *val);
/* Same, restricted to elements [e0, e1) (used for bounded CPU timing samples). */
int om_fill_range(const om_problem *p, const double *x, double *resid, double *val, int e0, int e1);
/* matrix_fill_full of a transient step (theta method): as om_fill, plus the
 * mass terms of momentum and energy (src/mm_fill_momentum.c:534-539,
 * 1586-1590; src/mm_fill_energy.c:322-328, 435-441).  xdot is the time
 * derivative at x, cmass = (1 + 2 theta) / dt its derivative in x; the caller
 * forms xdot = cmass (x - x_old) - 2 theta xdot_old (src/rf_solve.c:2848). */
int om_fill_transient(const om_problem *p, const double *x, const double *xdot, double cmass,
                      double *resid, double *val);
/* Weak heat-flux condition QSIDE (src/bc_integ.c:1518-1519, 2207) on element
 * sides (Exodus numbering, 1-based): the energy residual gains the integral
 * of phi_i q over each side, q being the heat flux into the domain.  Calls
 * append.  Returns 0, or -1 without the energy equation or for an element or
 * side out of range. */
int om_add_qside(om_problem *p, int n, const int *elems, const int *sides, double q);
/* QCONV: heat transfer coefficient h and bath temperature T_c on the given sides */
int om_add_qconv(om_problem *p, int n, const int *elems, const int *sides, double h, double tc);
/* One element's dense local residual and Jacobian in the oracle's local order:
 * local dofs = [var slot][dof], var slots U1,U2,(U3),(T),P.  R[nloc], J[nloc*nloc]. */
int om_element(const om_problem *p, int elem, const double *x, double *R, double *J, int *nloc);

void om_row_sum_scale(int N, double *val, const int *ija, double *b, double *scale);
double om_norm_L2(const double *v, int n);
double om_norm_L1(const double *v, int n);
double om_norm_Loo(const double *v, int n);
double om_norm_L2_r(const double *d, const double *x, int n);
double om_norm_Loo_r(const double *d, const double *x, int n);
void om_msr_matvec(int N, const double *val, const int *ija, const double *x, double *y);

/* Node colouring used by the GPU preconditioner (greedy first-fit over the
 * node-node graph in node order, element-centroid nodes forced into one last
 * colour) and the resulting unknown permutation perm[new] = old. */
int om_node_colours(const om_problem *p, int *colour_out);
void om_colour_permutation(const om_problem *p, int *perm_out, int *colour_ptr_out, int *ncolours);

int om_patch_permutation(const om_problem *p, int nbins, int *perm, int *colour_ptr, int *npatch,
                         int *patch_ptr);

/* ILU(0) + GMRES on an MSR matrix.  precond 0/1/2 as in om_solver_opts.
 * x is the initial guess (Goma passes zero); returns iterations, or -1 on a
 * zero pivot.  *rel_resid = final ||b-Ax|| / ||b - A x0||. */
int om_gmres_msr(int N, const double *val, const int *ija, const double *b, double *x,
                 const om_problem *p_for_colour, int precond, int kspace, int max_its,
                 double tol, double *rel_resid);

int om_gmres_perm(int N, const double *val, const int *ija, const double *b, double *x,
                  const int *perm, const int *part, int kspace, int max_its, double tol,
                  double *rel_resid);

/* ILU(0) factors of the (optionally permuted) matrix in CSR form (for tests):
 * returns nnz; rowptr[N+1], col[nnz], lu[nnz] (L strictly lower, unit diag;
 * U upper incl. diag), all in permuted numbering when perm != NULL. */
int om_ilu0_csr(int N, const double *val, const int *ija, const int *perm, int *rowptr,
                int *col, double *lu);

/* Newton loop (solve_nonlinear_problem, steady, full step).  x is in/out.
 * Returns Goma's return value (passes through the loop) or -1. */
int om_newton(om_problem *p, double *x, const om_solver_opts *opt, om_report *rep);
/* PSPG inputs.  global_h_elem_siz (src/mm_fill_aux.c:1128-1200): the mean
 * over elements of sqrt(sum_p hsquared[p] / dim), hsquared from h_elem_siz's
 * face-centroid widths (src/mm_fill_aux.c:844-1125).  global_velocity_norm
 * (src/mm_fill_aux.c:612-672): the mean of x^2 over the velocity unknowns
 * (a mean square, not its root).  om_newton sets the velocity norm from x at
 * the top of each pass (src/mm_sol_nonlinear.c:1183-1192); a standalone
 * om_fill uses the last value set. */
double om_h_elem_avg(const om_problem *p);
double om_velocity_norm(const om_problem *p, const double *x);
void om_set_velocity_norm(om_problem *p, double U_norm);
/* tau_pspg of the global scheme at the current velocity norm
 * (src/mm_fill_stabilization.c:1037-1049). */
double om_pspg_tau(const om_problem *p);

#ifdef __cplusplus
}
#endif
#endif
