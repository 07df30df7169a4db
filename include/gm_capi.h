/*
 * gm_capi.h -- C ABI of libgm.so, the MI355X-native replacement for Goma's
 * Newton inner loop (steady incompressible Navier-Stokes + energy on
 * QUAD9 / HEX27 Q2-P1 elements).
 *
 * The library keeps a context (mesh, unknown map, MSR graph, element
 * colouring, element-to-nonzero map, preconditioner ordering) resident in
 * HBM and runs the element assembly, the scatter, row-sum scaling, norms,
 * ILU(0)-preconditioned GMRES and the Newton update as HIP kernels on gfx950.
 *
 * Conventions (Goma's, src/mm_sol_nonlinear.c:1108-2913, src/mm_fill.c:158-312):
 *   - Unknowns are node-major, at each node U1,U2,(U3),(T),P (P1 pressure dofs
 *     on the element centroid node); variable type ids are Goma's
 *     (include/rf_fem_const.h:174-200): U1=0 U2=1 U3=2 T=3 P=9.
 *   - The Jacobian is stored in Goma's MSR layout: val[0..N) diagonal,
 *     bindx[0..N] row pointers into bindx/val for the off-diagonals,
 *     bindx[N+1..] column indices; arrays have length bindx[N].
 *   - matrix_fill accumulates (+=) into resid and val; callers zero them.
 *   - Newton solves J dx = R (R row-sum scaled) and updates x -= damp*dx.
 *   - No entry point aborts: every call returns GM_OK (0) or a negative
 *     status; gm_last_error() gives a message.
 *   - A context is not re-entrant: one host thread per context.  Every call
 *     is synchronous: device work is complete when it returns.  The library
 *     works on its own blocking stream, so its kernels start after the work
 *     a caller queued on the default (NULL) stream, such as zeroing the
 *     arrays it passes in; work on other non-blocking streams must be
 *     complete before the call.  The multigrid set-up also runs kernels on
 *     streams of its own (per coarse level, and the block inversions on one
 *     kept off 1/8 of the CUs), ordered after that stream's work and joined
 *     back into it.
 *
 * Pointer arguments ending in _dev are device (HBM) pointers; the others are
 * host pointers.
 */
#ifndef GM_CAPI_H
#define GM_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

#define GM_OK 0
#define GM_ERR_ARG -2      /* invalid argument */
#define GM_ERR_HIP -3      /* HIP runtime error */
#define GM_ERR_ELEM -1     /* zero or negative element Jacobian (matrix_fill_full's -1) */
#define GM_ERR_PIVOT -4    /* zero pivot in ILU(0) */
#define GM_ERR_NAN -5      /* non-finite residual norm (solve_nonlinear_problem's -1) */
#define GM_ERR_CONV -6     /* a time step's Newton solve did not converge */

#define GM_VAR_U1 0
#define GM_VAR_U2 1
#define GM_VAR_U3 2
#define GM_VAR_T 3
#define GM_VAR_P 9

#define GM_PREC_FP64 0   /* Jacobian, ILU and Krylov in FP64 */
#define GM_PREC_MIXED 1  /* Jacobian values, ILU and inner Krylov FP32; R, dots, update FP64 */

#define GM_PC_NONE 0
#define GM_PC_ILU0 1     /* ILU(0) in patch order (see gm_solver_opts.patch_bins) */
#define GM_PC_MG 2       /* geometric multigrid V-cycle, element-block (Vanka) smoothing */
#define GM_PC_BAND 3     /* banded LU of the scaled matrix in Goma's unknown order, without
                            pivoting: a direct solve (the role of Goma's direct solvers,
                            Solution Algorithm = lu / umf, src/mm_sol_nonlinear.c:1624-1681);
                            memory N (kl + ku + 129) x 8 bytes (x 4 in mixed precision),
                            kl / ku the graph's bandwidths */

typedef struct gm_mesh {
  int dim;             /* 2: QUAD9, 3: HEX27 */
  int num_nodes;
  int num_elems;
  const double *coor;  /* [dim][num_nodes] (Goma's Coor) */
  const int *connect;  /* [num_elems][9|27], Goma local node order (rf_shape.c) */
} gm_mesh;

typedef struct gm_physics {
  int energy;          /* 1: energy equation and T unknowns active */
  int visc_model;      /* 0: Newtonian (mu0); 1: Carreau(-Yasuda) */
  double rho;
  double mu0, muinf, lam, nexp, aexp; /* Carreau: mu = muinf + (mu0-muinf)(1+(lam gd)^a)^((n-1)/a) */
  double Cp, kcond, hsrc;
  double g[3];         /* momentum_source (gravity vector) */
  double beta;         /* Volume_Expansion (Boussinesq) */
  double Tref;         /* reference[TEMPERATURE] */
  int pspg;            /* Pressure Stabilization: 0 no, 1 global (Goma's PSPG = 1) */
  double ps_scaling;   /* Pressure Stabilization Scaling */
  double energy_supg;  /* energy weight function SUPG coefficient (Ewt_funcModel = SUPG,
                          Ewt_func; src/mm_fill_energy.c:237-259), 0: Galerkin */
} gm_physics;

typedef struct gm_bc {
  int num_dbc;         /* Dirichlet conditions: (node, variable type, value) */
  const int *node;     /* later entries override earlier ones on one unknown */
  const int *var;
  const double *value;
  int datum_elem;      /* pressure datum: element whose first P dof is fixed, -1: none */
  double datum_value;
} gm_bc;

typedef struct gm_solver_opts {
  int max_newton;      /* Max_Newton_Steps */
  double eps_resid;    /* Normalized Residual Tolerance (Epsilon[0]) */
  double eps_corr;     /* Normalized Correction Tolerance (Epsilon[2]); >1 disables */
  double damp;         /* full-step damping factor */
  int kspace;          /* GMRES restart length */
  int max_its;         /* GMRES iteration limit per Newton step */
  double lin_tol;      /* ||r||/||r0|| tolerance */
  int precond;         /* GM_PC_NONE | GM_PC_ILU0 | GM_PC_MG | GM_PC_BAND */
  int patch_bins;      /* ILU(0) patches per axis (0: library default) */
  int precision;       /* GM_PREC_FP64 | GM_PREC_MIXED */
  int verbose;         /* 1: one line per Newton step on stderr */
  int mg_nu;           /* multigrid: smoothing sweeps before and after the coarse correction */
  double mg_omega;     /* multigrid: damping of the element-block updates */
  int mg_coarse_sweeps;/* multigrid: 0 = exact solve on the coarsest level (dense
                          inverse, at most 4096 unknowns); k > 0 = k smoothing sweeps */
  int line_search;     /* 0: full step (NLS_FULL_STEP); 1: backtracking (NLS_BACKTRACK) */
  double min_damp;     /* Line_Search_Minimum_Damping */
  int pc_reuse;        /* Matrix_Factorization_Reuse (src/mm_sol_nonlinear.c:1687-1708):
                          0 "recalc": rebuild the preconditioner at every Newton pass;
                          1 "reuse": build it at the first pass and keep it;
                          2 adaptive: keep it until GMRES misses lin_tol or needs
                            more than 1.5x the iterations of its first use */
  int mg_smoother;     /* multigrid element-block smoother: 0 multiplicative (coloured block
                          Gauss-Seidel, a residual on each colour's rows); 1 additive (one
                          residual per sweep, corrections weighted by 1 / elements sharing
                          the unknown) */
  int mg_galerkin;     /* multigrid coarse operators: 0 re-assembled on each coarse mesh at the
                          injected state; 1 Galerkin R A P of the finer level's element
                          matrices (P element-local: each finer element inside one coarser
                          one), with each coarse level's own Dirichlet rows */
  int lin_tol_adapt;   /* 0: every linear solve to lin_tol (Goma's fixed Residual Ratio
                          Tolerance); 1: Eisenstat-Walker forcing terms (choice 2, gamma 0.9,
                          alpha (1+sqrt 5)/2, at most 0.1, at least lin_tol) from the ratio of successive
                          scaled residual norms -- loose solves far from the solution */
} gm_solver_opts;

#define GM_MAX_REPORT 64
typedef struct gm_newton_report {
  int newton_its;      /* Goma's return value: passes through the Newton loop */
  int converged;
  int lin_its_total;
  int lin_its[GM_MAX_REPORT];
  double resid_L2[GM_MAX_REPORT];    /* row-sum-scaled ||R||_2 per pass */
  double resid_Loo[GM_MAX_REPORT];
  double unscaled_L2[GM_MAX_REPORT];
  double lin_relres[GM_MAX_REPORT];
  double damp[GM_MAX_REPORT];        /* step length taken at each pass */
  double t_assemble, t_scale, t_precond, t_solve, t_update, t_total; /* seconds */
  int fell_back;       /* gm_solve_nonlinear_guess_dev: 1 if the guess was abandoned */
  int discarded_its;   /* Newton passes spent on the abandoned guess */
  int pc_builds;       /* preconditioner set-ups (see pc_reuse) */
  int resid_fills;     /* residual-only assemblies (line-search trials) */
} gm_newton_report;

typedef struct gm_ctx gm_ctx;

/* Builds the unknown map, the MSR graph (find_MSR_problem_graph,
 * src/mm_fill_util.c:3229), the element colouring and element-to-nonzero
 * map, and the ILU patch ordering; uploads everything to the current HIP
 * device.  Replaces the set-up Goma does in setup_problem / solve_problem
 * (src/rf_setup_problem.c:194, src/rf_solve.c:729-769). */
int gm_ctx_create(const gm_mesh *mesh, const gm_physics *phys, const gm_bc *bc, gm_ctx **out);
void gm_ctx_destroy(gm_ctx *ctx);
const char *gm_last_error(void);

int gm_num_unknowns(const gm_ctx *ctx);
int gm_num_nodes(const gm_ctx *ctx);
int gm_num_elems(const gm_ctx *ctx);
int gm_msr_size(const gm_ctx *ctx);                /* bindx[N] = length of bindx and val */
int gm_copy_msr_graph(const gm_ctx *ctx, int *bindx);        /* host copy, length bindx[N] */
int gm_copy_first_unknown(const gm_ctx *ctx, int *fu);       /* [num_nodes+1] */
int gm_copy_dirichlet(const gm_ctx *ctx, signed char *mask, double *value); /* [N] each */
int gm_num_elem_colours(const gm_ctx *ctx);
int gm_copy_elem_colours(const gm_ctx *ctx, int *colour);    /* [num_elems] */
/* ILU patch ordering: perm[new] = old; colour_ptr[ncol+1]; returns ncol. */
int gm_copy_patch_permutation(gm_ctx *ctx, int patch_bins, int *perm, int *colour_ptr);

/* Replaces the material constants (continuation parameter update,
 * update_parameterC, src/ac_update_parameter.c:48), on this context and on
 * the multigrid levels attached to it. */
int gm_set_physics(gm_ctx *ctx, const gm_physics *phys);

/* Pressure stabilisation (PSPG, global tau; calc_pspg,
 * src/mm_fill_stabilization.c:1037-1049): with phys.pspg = 1 the continuity
 * rows gain int grad psi_k . tau (rho v_dot + rho v.grad v + grad P - f) dA
 * (src/mm_fill_continuity.c:600-612) and its Jacobian, with
 * Re = rho U h / (2 mu_avg), tau = s h^2 / (12 mu_avg) (Re <= 3) or
 * s h / (2 rho U), mu_avg the viscosity at a shear-rate tensor of 1e-5
 * entries (element_viscosity, src/mm_fill_aux.c:687).
 * gm_h_elem_avg replaces global_h_elem_siz (src/mm_fill_aux.c:1128): the
 * mean over elements of sqrt(sum_p hsquared[p] / dim), from the mesh.
 * gm_velocity_norm_dev replaces global_velocity_norm (src/mm_fill_aux.c:612):
 * the mean of x^2 over the velocity unknowns of x_dev.
 * gm_set_velocity_norm is matrix_fill_full's U_norm argument
 * (src/mm_fill.c:174) for the assembly entry points; the Newton drivers set
 * it themselves from x at the top of every pass (src/mm_sol_nonlinear.c:
 * 1183-1192), and their line-search residuals keep it. */
int gm_h_elem_avg(gm_ctx *ctx, double *h);
int gm_velocity_norm_dev(gm_ctx *ctx, const double *x_dev, double *U_norm);
int gm_set_velocity_norm(gm_ctx *ctx, double U_norm);

/* matrix_fill_full (src/mm_fill.c:158-312): accumulates the residual and the
 * analytic Jacobian of every element, with Dirichlet rows replaced per element
 * (put_dirichlet_in_matrix, src/bc_dirich.c:44).  val is MSR (gm_msr_size
 * doubles).  Returns GM_OK or GM_ERR_ELEM. */
int gm_matrix_fill_full_dev(gm_ctx *ctx, const double *x_dev, double *resid_dev, double *val_dev);
int gm_matrix_fill_full(gm_ctx *ctx, const double *x, double *resid, double *val);
/* The same assembly, with the Jacobian stored rather than accumulated: val
 * needs no zeroing beforehand (the HEX27 kernel writes each entry's first
 * element contribution and adds the others; QUAD9 zeroes val first).  resid
 * is still accumulated.  Saves the caller's pass over val (8 GB at 48^3). */
int gm_matrix_fill_full_assign_dev(gm_ctx *ctx, const double *x_dev, double *resid_dev,
                                   double *val_dev);
/* The same assembly with the Jacobian values in FP32 (the mixed-precision
 * path, GM_PREC_MIXED): val32 holds gm_msr_size floats and is accumulated
 * (+=); the residual stays FP64.  Replaces matrix_fill_full
 * (src/mm_fill.c:158-312) where the Jacobian is stored in single precision. */
int gm_matrix_fill_full32_dev(gm_ctx *ctx, const double *x_dev, double *resid_dev, float *val32_dev);

/* matrix_fill_full of one theta-method time step: the steady terms plus the
 * transient mass terms of momentum and energy (src/mm_fill_momentum.c:534-539,
 * J 1586-1590; src/mm_fill_energy.c:322-328, J 435-441), with the time
 * derivative that solve_nonlinear_problem keeps in step with x,
 *   xdot = (1 + 2 theta) / delta_t (x - x_old) - 2 theta xdot_old
 * (src/rf_solve.c:2848), and its derivative (1 + 2 theta) / delta_t in J.
 * theta = 0 is backward Euler, 0.5 Crank-Nicolson.  delta_t > 0. */
int gm_matrix_fill_full_transient_dev(gm_ctx *ctx, const double *x_dev, const double *x_old_dev,
                                      const double *xdot_old_dev, double delta_t, double theta,
                                      double *resid_dev, double *val_dev);

/* QSIDE weak heat flux on element sides (BC = QSIDE SS <id> <q>,
 * src/bc_integ.c:1518-1519, 2156-2157): every later assembly adds
 * q * phi_i * |surface Jacobian| integrated over side sides[k] (1-based,
 * Exodus numbering as get_side_info, src/mm_fill_fill.c:3049-3072, 3182-3250)
 * of element elems[k] to the energy rows of that element's nodes (no Jacobian
 * term; Dirichlet rows keep x - V).  Entries accumulate over calls, at most
 * 8 per element.  Needs the energy equation. */
int gm_add_qside(gm_ctx *ctx, int n, const int *elems, const int *sides, double q);
/* QCONV on element sides (same numbering and limits as gm_add_qside): adds
 * h (t_bath - T) to the inward heat flux, i.e. the side integral of
 * phi_i h (t_bath - T) to the energy residual and -h times the side integral
 * of phi_i phi_j to its temperature Jacobian.  Replaces the QCONV_BC case of
 * apply_integrated_bc (src/bc_integ.c:1492-1494: qrad_surf with zero
 * emissivity, src/mm_ns_bc.c:15145-15255).  Cleared by gm_clear_qside. */
int gm_add_qconv(gm_ctx *ctx, int n, const int *elems, const int *sides, double h, double t_bath);
/* Drops every QSIDE and QCONV entry. */
int gm_clear_qside(gm_ctx *ctx);

/* row_sum_scale_MSR (src/sl_matrix_util.c:506): scales val and b in place. */
int gm_row_sum_scale_dev(gm_ctx *ctx, double *val_dev, double *b_dev, double *scale_dev);
/* MV_MSR (src/sl_auxutil.c:373): y = A x. */
int gm_msr_matvec_dev(gm_ctx *ctx, const double *val_dev, const double *x_dev, double *y_dev);
/* Preconditioned GMRES on an MSR matrix (the AZ_solve call,
 * src/mm_sol_nonlinear.c:1736): x_dev holds the initial guess. */
int gm_gmres_dev(gm_ctx *ctx, const double *val_dev, const double *b_dev, double *x_dev,
                 const gm_solver_opts *opt, int *its, double *relres);
/* ILU(0) factors of the patch-ordered matrix, as CSR in permuted numbering
 * (for tests): rowptr[N+1], col[nnz], lu[nnz]; returns nnz or <0. */
int gm_ilu0_factor_dev(gm_ctx *ctx, const double *val_dev, int patch_bins, int *rowptr, int *col,
                       double *lu);

/* solve_nonlinear_problem (src/mm_sol_nonlinear.c:596): steady Newton to
 * ||D^-1 R||_2 < eps_resid and ||D^-1 R||_oo < eps_resid.  x is the initial
 * guess on entry and the solution on exit.  Returns the number of passes
 * through the Newton loop, or a negative status. */
int gm_solve_nonlinear_dev(gm_ctx *ctx, double *x_dev, const gm_solver_opts *opt,
                           gm_newton_report *rep);
int gm_solve_nonlinear(gm_ctx *ctx, double *x, const gm_solver_opts *opt, gm_newton_report *rep);

/* Newton from a predicted guess, with the previous-step guess as fallback (no
 * counterpart in the reference: Goma's zeroth-order continuation,
 * src/ac_conti.c:760-787, starts each step from the previous solution, which
 * is what x_prev_dev is here).  x_dev holds the predicted guess on entry and
 * the solution on exit.  If the scaled residual ||D^-1 R||_2 has not dropped
 * below its value at the guess divided by `drop` within `watch_its` Newton
 * passes, the guess is abandoned and Newton restarts from x_prev_dev;
 * rep->fell_back and rep->discarded_its record that, and the timings cover
 * both attempts. */
int gm_solve_nonlinear_guess_dev(gm_ctx *ctx, double *x_dev, const double *x_prev_dev,
                                 const gm_solver_opts *opt, int watch_its, double drop,
                                 gm_newton_report *rep);

/* solve_nonlinear_problem for one theta-method time step
 * (src/mm_sol_nonlinear.c:596, called from src/rf_solve.c:1905): Newton on
 * the transient residual of gm_matrix_fill_full_transient_dev.  x_dev holds
 * the predicted solution on entry and the solution at t + delta_t on exit.
 * With multigrid, Galerkin coarse operators (mg_galerkin = 1) carry the mass
 * terms to the coarse levels; re-assembled coarse operators omit them. */
int gm_solve_nonlinear_transient_dev(gm_ctx *ctx, double *x_dev, const double *x_old_dev,
                                     const double *xdot_old_dev, double delta_t, double theta,
                                     const gm_solver_opts *opt, gm_newton_report *rep);

typedef struct gm_transient_opts {
  double delta_t;      /* fixed time step */
  double theta;        /* Time step parameter: 0 backward Euler, 0.5 Crank-Nicolson */
  int num_steps;       /* time steps to take */
  int startup_steps;   /* first steps taken with theta = 0 (Goma: 3, src/rf_solve.c:1378-1386) */
} gm_transient_opts;

typedef struct gm_transient_report {
  int steps;           /* converged time steps taken */
  double time;         /* time reached */
  int newton_its_total, lin_its_total;
  int newton_its[GM_MAX_REPORT];  /* per step (first GM_MAX_REPORT steps) */
  double t_total;      /* seconds */
} gm_transient_report;

/* The fixed-step time loop of solve_problem (src/rf_solve.c:1307-2372):
 * per step the predictor of predict_solution (src/rf_solve.c:2767-2830;
 * second-order extrapolation for theta = 0.5, first-order otherwise), a
 * transient Newton solve, then xdot = (1 + 2 theta)/dt (x - x_old) - 2 theta
 * xdot_old and the history shift x_oldest <- x_older <- x_old <- x.
 * x_dev is the initial condition on entry and the solution at
 * t0 + steps * delta_t on exit; xdot_dev the initial time derivative
 * (Goma starts from zero) on entry and the last one on exit.  Returns the
 * number of steps taken, or a negative status; a step whose Newton solve does
 * not converge ends the loop with GM_ERR_CONV (Goma would cut the step). */
int gm_solve_transient_dev(gm_ctx *ctx, double *x_dev, double *xdot_dev,
                           const gm_transient_opts *topt, const gm_solver_opts *opt,
                           gm_transient_report *rep);

void gm_default_solver_opts(gm_solver_opts *opt);

/* Multigrid hierarchy for GM_PC_MG (the reference has no counterpart: Goma's
 * AztecOO path offers one-level preconditioners only).  Attaches `coarse` as
 * the next level below the current coarsest one.  `coarse` is a context on the
 * coarsened mesh with the same physics and the Dirichlet masks of its nodes;
 * P (CSR: P_rowptr[N_fine+1], P_col, P_val) interpolates coarse unknowns to
 * the finer level; inject[coarse node] is the finer node at the same place.
 * The fine context keeps a pointer to `coarse`, which must outlive it. */
int gm_mg_add_level(gm_ctx *fine, gm_ctx *coarse, const int *P_rowptr, const int *P_col,
                    const double *P_val, const int *inject);
int gm_mg_num_levels(const gm_ctx *ctx);
/* Builds and attaches the whole hierarchy of a structured lattice mesh (nodes
 * on a (2n+1)^dim lattice and elements numbered x fastest, as the generated
 * workloads are): coarse meshes keeping every other element vertex, down to
 * min_elems_per_side elements per side while n is even; each level a context
 * owned by ctx (freed with it) with the physics of ctx and the Dirichlet masks
 * of the finer level at the same nodes; Q2 / P1 prolongations exact for
 * Q2 x P1 fields on affine meshes; injection at shared nodes.  Returns the
 * number of levels including ctx, or GM_ERR_ARG for a mesh with another
 * numbering (ILU(0) or the banded LU serve those) or when a hierarchy is
 * already attached.  No counterpart in the reference (Goma's AztecOO path has
 * one-level preconditioners only); it is what selects GM_PC_MG at Goma's
 * solve_nonlinear_problem call sites. */
int gm_mg_build(gm_ctx *ctx, int min_elems_per_side);
/* The context of multigrid level `level` (0 = ctx itself), borrowed: valid
 * while ctx lives; NULL if there is no such level. */
gm_ctx *gm_mg_level(gm_ctx *ctx, int level);
/* Host copy of the prolongation of level `level` >= 1 (finer unknowns x this
 * level's unknowns) built by gm_mg_build, as CSR: rowptr [N_finer + 1],
 * col, val [nnz] (any may be NULL); returns nnz. */
int gm_mg_copy_prolongation(const gm_ctx *ctx, int level, int *rowptr, int *col, double *val);
/* Coarse Jacobians at the injected state x_dev and element-block inverses for
 * the level-0 matrix val_dev (row-sum scaled by scale_dev, or NULL). */
int gm_mg_setup_dev(gm_ctx *ctx, const double *x_dev, const double *val_dev, const double *scale_dev);
/* Host copy (as FP64) of the element-block inverse of element `elem` on
 * multigrid level `level` (0 = the context's own mesh) from the last set-up:
 * n x n with n = the element's unknowns (local order: each nodal variable
 * over the element's nodes, then its P1 dofs), stored transposed
 * (out[j * n + i] = (A_e^-1)[i][j]).  Returns n. */
int gm_mg_copy_block_inverse(const gm_ctx *ctx, int level, int elem, double *out);
/* Host copy (as FP64) of the Jacobian of multigrid level `level` (1 = first
 * coarse level) from the last set-up, in that level's MSR layout. */
int gm_mg_copy_level_matrix(const gm_ctx *ctx, int level, double *val);
/* Host copy of the dense inverse of the coarsest multigrid level's Jacobian
 * from the last set-up with an exact coarse solve: row-major n x n, n = that
 * level's unknowns (at most 4096).  Returns n.  No counterpart in the
 * reference (Goma has no multigrid); it exposes the exact coarse solve for
 * inspection, as gm_mg_copy_block_inverse does the smoother's blocks. */
int gm_mg_copy_coarse_inverse(const gm_ctx *ctx, double *out);

/* Kernel timing: when enabled, launches of the kernels below carry HIP
 * events that their own dispatch stamps (hipExtLaunchKernel), so a duration
 * is the kernel's execution alone; gm_prof_get returns the launch count, the
 * summed duration and the summed algorithmic HBM bytes (compulsory traffic,
 * each array counted once per launch) of those launches. */
#define GM_K_FILL 0       /* k_fill: assembly + scatter of one element colour */
#define GM_K_SPMV 1       /* k_spmv_msr: MSR matvec */
#define GM_K_ILU_LOWER 2  /* k_ilu_lower: forward sweep of one patch colour */
#define GM_K_ILU_UPPER 3  /* k_ilu_upper: backward sweep of one patch colour */
#define GM_K_ILU_FACT 4   /* k_ilu_factor: ILU(0) of one patch colour */
#define GM_K_MDOT 5       /* k_mdot: Gram-Schmidt projections */
#define GM_K_ORTH 6       /* k_orth_update: Gram-Schmidt update */
#define GM_K_MG_RESID 7   /* k_mg_resid: multigrid residual r = b - A x, all rows of one level */
#define GM_K_VANKA 8      /* k_vanka_update: element-block updates of one colour (multiplicative);
                             k_vanka_corr: every element's correction (additive) */
#define GM_K_TRANSFER 9   /* k_csr_apply: restriction or prolongation */
#define GM_K_VANKA_INV 10 /* k_vanka_invert: element-block inverses of one level */
#define GM_K_FILL_RESID 11 /* k_fill, residual only (line-search trial), one colour */
#define GM_K_MG_RESID_ROWS 12 /* k_mg_resid_rows: smoother residual on one element colour's rows */
#define GM_K_MG_DENSE 13  /* k_dense_gemv: exact solve on the coarsest level (dense inverse) */
#define GM_K_GALERKIN 14  /* k_galerkin: coarse element matrices P^T K P of one level */
#define GM_K_ELEM_SCATTER 15 /* k_elem_scatter: element matrices of one colour into MSR */
#define GM_K_BAND_PANEL 16 /* k_band_a11 / k_band_panel: one 64-column panel of the banded LU */
#define GM_K_BAND_UPDATE 17 /* k_band_update: trailing update of one panel */
#define GM_K_BAND_SOLVE 18 /* k_band_fwd / k_band_bwd: one 64-row block of the triangular solves */
#define GM_K_VANKA_GATHER 19 /* k_vanka_gather: additive smoother, element corrections summed per unknown */
#define GM_K_ROW_SCALE 20 /* k_row_sum_scale: row sums and scaling of the Jacobian and residual */
#define GM_K_COARSE_INV 21 /* k_gjb_panel / k_gjb_swap / k_gjb_update: one panel of the coarsest level's dense inverse */
#define GM_K_COUNT 22
/* on = 0: off; 1: time every launch; k > 1: time a pseudo-random one launch
 * in k of each kernel (lower overhead), durations scaled to all launches by
 * the launch count. */
int gm_prof_enable(gm_ctx *ctx, int on);
int gm_prof_get(gm_ctx *ctx, int kernel_id, int *count, double *total_ms, double *bytes);

/* ------------------------------------------------------------------------
 * Learned initial guess: a coordinate-plus-parameter MLP trained on the
 * solutions of a continuation sweep (no counterpart in the reference; it
 * replaces the previous-step predictor of src/ac_conti.c:760-787 when it does
 * better, and gm_solve_nonlinear_guess_dev falls back to that predictor).
 *
 *   u = (raw - in_shift) * in_scale,  a_{l+1} = tanh(W_l a_l + b_l),
 *   y = W_H a_H + b_H,  loss = sum mask (y - t)^2 / (npts * n_out).
 * GEMMs run on the matrix cores in BF16 with FP32 accumulation (inputs,
 * weights, activations and back-propagated deltas rounded to BF16; biases,
 * outputs, gradients, Adam state and master weights FP32).
 * Parameters are exchanged unpadded, layer by layer: W_l [N_l][K_l]
 * row-major, then b_l [N_l].
 * ------------------------------------------------------------------------ */
typedef struct gm_mlp gm_mlp;

#define GM_MLP_MAX_IN 8
#define GM_MLP_MAX_OUT 8
typedef struct gm_mlp_shape {
  int n_in;             /* inputs: coordinates then parameters, <= GM_MLP_MAX_IN */
  int n_out;            /* outputs, <= GM_MLP_MAX_OUT */
  int hidden;           /* hidden layers (>= 1) */
  int width;            /* hidden width, a multiple of 64 */
  unsigned long long seed; /* Glorot-uniform initialisation */
  double in_shift[GM_MLP_MAX_IN], in_scale[GM_MLP_MAX_IN];
} gm_mlp_shape;

typedef struct gm_mlp_train_opts {
  int epochs;
  int batch;            /* points per step, multiple of 128 */
  double lr, lr_final;  /* Adam step size, cosine-decayed from lr to lr_final */
  double beta1, beta2, eps;
  unsigned long long seed; /* point shuffling */
  int verbose;
} gm_mlp_train_opts;

int gm_mlp_create(const gm_mlp_shape *shape, gm_mlp **out);
void gm_mlp_destroy(gm_mlp *m);
long long gm_mlp_num_params(const gm_mlp *m);
int gm_mlp_get_params(gm_mlp *m, float *params);
int gm_mlp_set_params(gm_mlp *m, const float *params); /* also resets the Adam state */
void gm_mlp_default_train_opts(gm_mlp_train_opts *o);
/* Training set (host arrays, copied to HBM): in [npts][n_in],
 * targ and mask [npts][n_out]. */
int gm_mlp_set_data(gm_mlp *m, long long npts, const float *in, const float *targ,
                    const float *mask);
/* Adam on the training set; loss_per_epoch[epochs] (may be NULL) receives
 * the mean batch loss of each epoch. */
int gm_mlp_train(gm_mlp *m, const gm_mlp_train_opts *o, double *loss_per_epoch);
/* Loss and gradient (unpadded parameter layout) of one batch, no update:
 * host arrays, npts <= 65536. */
int gm_mlp_loss_grad(gm_mlp *m, int npts, const float *in, const float *targ, const float *mask,
                     double *loss, float *grad);
/* y [npts][n_out] = network(in [npts][n_in]); device or host arrays. */
int gm_mlp_predict_dev(gm_mlp *m, long long npts, const float *in_dev, float *out_dev);
int gm_mlp_predict(gm_mlp *m, long long npts, const float *in, float *out);
/* Allocation-free gm_mlp_predict_dev: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Activations and
 * outputs pass through a scratch array the library keeps on each GPU for the
 * life of the process (two 2048-point x 512-feature BF16 buffers and
 * 2048 x 64 FP32 outputs, 4.5 MiB), 2048 points at a time, so any npts is
 * accepted; a network with a layer wider than 512 features is refused with
 * GM_ERR_ARG.  Calls on one GPU share its scratch: calls that overlap on one
 * GPU race.  Makes no hipMalloc or hipFree call.  Same outputs as
 * gm_mlp_predict_dev, bit for bit. */
int gm_mlp_predict_noalloc_dev(gm_mlp *m, long long npts, const float *in_dev, float *out_dev);

/* One query of the network for a context: the parameter inputs that follow
 * the coordinates, and the affine map from network outputs to fields. */
typedef struct gm_mlp_query {
  double param[GM_MLP_MAX_IN];     /* inputs n_in - dim .. n_in - 1 */
  double out_scale[GM_MLP_MAX_OUT];
  double out_shift[GM_MLP_MAX_OUT];
} gm_mlp_query;
/* Initial guess in the context's unknown layout.  Outputs are
 * (U1, U2[, U3], P[, T]): nodal values at the mesh nodes; each element's P1
 * coefficients {P0, P_s, P_t[, P_u]} from the pressure output at the element
 * centroid and at reference points +-1/2 along each axis (central
 * differences, exact for pressure fields linear along those lines).
 * Dirichlet unknowns take their prescribed values.  With q_prev and
 * x_prev_dev (delta form): x = x_prev + G(q) - G(q_prev). */
int gm_mlp_guess_dev(gm_mlp *m, gm_ctx *ctx, const gm_mlp_query *q, const gm_mlp_query *q_prev,
                     const double *x_prev_dev, double *x_dev);
/* Snapshot subspace for the delta form: an orthonormal basis (in the norm
 * weighted by weight_dev[N], zero on excluded unknowns) of the differences
 * snaps[k] - snaps[0] of the sweep solutions snaps_dev [nsnap][N].  While it is
 * set, gm_mlp_guess_dev on this context projects the network's predicted
 * change G(q) - G(q_prev) onto the subspace before adding it to x_prev: the
 * network supplies the change, the subspace removes its mesh-scale noise.
 * Returns the basis size; gm_mlp_clear_basis drops it. */
int gm_mlp_set_basis(gm_mlp *m, gm_ctx *ctx, int nsnap, const double *snaps_dev,
                     const double *weight_dev);
void gm_mlp_clear_basis(gm_mlp *m);

#ifdef __cplusplus
}
#endif
#endif
