// C ABI of libgm (include/gm_capi.h): context life cycle, the
// matrix_fill_full and solve_nonlinear_problem replacements, kernel timing.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "gm_internal.h"

static thread_local char g_err[512] = "";

void gm_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char *gm_last_error(void) { return g_err; }

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static constexpr int kProfPool = 2048;

// fold recorded event pairs of one kernel id into the totals
static void prof_fold(gm_ctx *c, int id) {
  if (!c->ev_used[id]) return;
  hipEventSynchronize(c->ev_e[id][c->ev_used[id] - 1]);
  for (int i = 0; i < c->ev_used[id]; i++) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev_b[id][i], c->ev_e[id][i]);
    c->prof_ms[id] += ms;
  }
  c->prof_count[id] += c->ev_used[id];
  c->ev_used[id] = 0;
}

// Every launch adds its bytes to the family's total.  With prof_every = k > 1
// a pseudo-random one launch in k (and the family's first) is timed, by
// events that its own dispatch stamps; a fixed stride could alias with the
// repeating launch pattern of a V-cycle and bias the average.
static bool sampled_launch(long long seen, int id, int every) {
  if (every <= 1 || seen == 1) return true;
  uint64_t z = (uint64_t)seen * 0x9e3779b97f4a7c15ull + (uint64_t)id * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return z % (uint64_t)every == 0;
}

bool gm_prof_take(gm_ctx *c, int id, double bytes, hipEvent_t *b, hipEvent_t *e) {
  if (!c->prof) return false;
  c->prof_bytes[id] += bytes;
  c->prof_seen[id]++;
  if (!sampled_launch(c->prof_seen[id], id, c->prof_every)) return false;
  if (c->ev_used[id] == kProfPool) prof_fold(c, id);
  const int k = c->ev_used[id];
  if ((int)c->ev_b[id].size() <= k) {
    hipEvent_t eb, ee;
    if (hipEventCreate(&eb) != hipSuccess) return false;
    if (hipEventCreate(&ee) != hipSuccess) {
      (void)hipEventDestroy(eb);
      return false;
    }
    c->ev_b[id].push_back(eb);
    c->ev_e[id].push_back(ee);
  }
  c->prof_sampled_bytes[id] += bytes;
  c->ev_used[id]++;
  *b = c->ev_b[id][k];
  *e = c->ev_e[id][k];
  return true;
}

extern "C" void gm_default_solver_opts(gm_solver_opts *o) {
  o->max_newton = 25;
  o->eps_resid = 1e-10;
  o->eps_corr = 1e10;
  o->damp = 1.0;
  o->kspace = 50;
  o->max_its = 500;
  o->lin_tol = 1e-6;
  o->precond = GM_PC_ILU0;
  o->patch_bins = 0;
  o->precision = GM_PREC_FP64;
  o->verbose = 0;
  o->mg_nu = 1;
  o->mg_omega = 0.7;
  o->mg_coarse_sweeps = 0;
  o->line_search = 0;
  o->min_damp = 0.05;
  o->pc_reuse = 0;
  o->mg_smoother = 0;
  o->mg_galerkin = 0;
  o->lin_tol_adapt = 0;
}

static void free_ctx(gm_ctx *c) {
  if (!c) return;
  gm_free_linalg(c);
  gm_mg_free(c);
  gm_band_free(c);
  void *ptrs[] = {c->d_coor, c->d_conn,   c->d_fu,    c->d_centroid, c->d_dbc,   c->d_dbc_val,
                  c->d_bindx, c->d_ecol_elems, c->d_e2nz, c->d_flag, c->d_tables, c->d_val,
                  c->d_resid, c->d_dx,    c->d_scale, c->d_val32,    c->d_red,   c->d_xwork,
                  c->d_xsave, c->d_rtrial, c->d_w, c->d_mg_rows, c->d_mg_wmul, c->d_kel,
                  c->d_mg_gptr, c->d_mg_gent, c->d_mg_corr, c->d_stab, c->d_qs_ptr,
                  c->d_qs_side, c->d_qs_q, c->d_rank, c->d_nn_ptr, c->d_nn_list};
  for (void *p : ptrs)
    if (p) hipFree(p);
  if (c->h_red) hipHostFree(c->h_red);
  for (int i = 0; i < GM_K_COUNT; i++) {
    for (hipEvent_t e : c->ev_b[i]) hipEventDestroy(e);
    for (hipEvent_t e : c->ev_e[i]) hipEventDestroy(e);
  }
  for (hipEvent_t e : c->ev_inv)
    if (e) hipEventDestroy(e);
  if (c->s_inv && c->s_inv != c->stream) hipStreamDestroy(c->s_inv);
  if (c->stream) hipStreamDestroy(c->stream);
  // the level contexts gm_mg_build made, after this context's levels were released
  for (gm_ctx *l : c->owned_levels) free_ctx(l);
  delete c;
}

template <typename T>
static int upload(T **d, const T *h, size_t n) {
  GM_HIP(hipMalloc(d, std::max<size_t>(n, 1) * sizeof(T)));
  if (n) GM_HIP(hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return GM_OK;
}

static int create_impl(const gm_mesh *m, const gm_physics *ph, const gm_bc *bc, gm_ctx *c) {
  if (!m || !ph || !bc || (m->dim != 2 && m->dim != 3) || m->num_nodes <= 0 || m->num_elems <= 0 ||
      !m->coor || !m->connect) {
    gm_set_error("invalid mesh description");
    return GM_ERR_ARG;
  }
  c->dim = m->dim;
  c->nve = m->dim == 2 ? 9 : 27;
  c->np = m->dim + 1;
  c->energy = ph->energy ? 1 : 0;
  c->nv = c->dim + c->energy;
  c->num_nodes = m->num_nodes;
  c->num_elems = m->num_elems;
  c->phys = *ph;
  c->coor.assign(m->coor, m->coor + (size_t)m->dim * m->num_nodes);
  c->conn.assign(m->connect, m->connect + (size_t)m->num_elems * c->nve);
  for (int v : c->conn)
    if (v < 0 || v >= m->num_nodes) {
      gm_set_error("connectivity references node %d outside [0, %d)", v, m->num_nodes);
      return GM_ERR_ARG;
    }
  int rc = gm_build_host(c);
  if (rc) return rc;
  c->h_elem_avg = gm_h_elem_size(c);
  // Dirichlet map per unknown; pressure datum on the datum element's first
  // P dof (find_and_set_Dirichlet, src/mm_bc.c:132, 386-457).
  c->dbc.assign(c->N, 0);
  c->dbc_val.assign(c->N, 0.0);
  for (int i = 0; i < bc->num_dbc; i++) {
    const int n = bc->node[i], v = bc->var[i];
    if (n < 0 || n >= c->num_nodes) {
      gm_set_error("Dirichlet condition %d on node %d outside the mesh", i, n);
      return GM_ERR_ARG;
    }
    int off = -1;
    if (v >= GM_VAR_U1 && v < GM_VAR_U1 + c->dim) off = v - GM_VAR_U1;
    else if (v == GM_VAR_T && c->energy) off = c->dim;
    if (off < 0) {
      gm_set_error("Dirichlet condition %d on variable %d that is not solved for", i, v);
      return GM_ERR_ARG;
    }
    const int u = c->first_unknown[n] + off;
    c->dbc[u] = 1;
    c->dbc_val[u] = bc->value[i];
  }
  if (bc->datum_elem >= 0) {
    if (bc->datum_elem >= c->num_elems) {
      gm_set_error("pressure datum element %d outside the mesh", bc->datum_elem);
      return GM_ERR_ARG;
    }
    const int u = c->first_unknown[c->elem_centroid[bc->datum_elem]] + c->dim + c->energy;
    c->dbc[u] = 1;
    c->dbc_val[u] = bc->datum_value;
  }

  // a blocking stream: ordered after work the caller queued on the default
  // stream (e.g. zeroing the arrays it passes in) and before what it queues later
  GM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamDefault));
  if ((rc = upload(&c->d_coor, c->coor.data(), c->coor.size()))) return rc;
  if ((rc = upload(&c->d_conn, c->conn.data(), c->conn.size()))) return rc;
  if ((rc = upload(&c->d_fu, c->first_unknown.data(), c->first_unknown.size()))) return rc;
  if ((rc = upload(&c->d_centroid, c->elem_centroid.data(), c->elem_centroid.size()))) return rc;
  if ((rc = upload(&c->d_dbc, c->dbc.data(), c->dbc.size()))) return rc;
  if ((rc = upload(&c->d_dbc_val, c->dbc_val.data(), c->dbc_val.size()))) return rc;
  if ((rc = upload(&c->d_bindx, c->bindx.data(), c->bindx.size()))) return rc;
  c->nn_cap = 0;
  for (int n = 0; n < c->num_nodes; n++) {
    int len = 0;
    for (int q = c->nn_ptr[n]; q < c->nn_ptr[n + 1]; q++)
      len += c->first_unknown[c->nn_list[q] + 1] - c->first_unknown[c->nn_list[q]];
    c->nn_cap = std::max(c->nn_cap, len);
  }
  if ((rc = upload(&c->d_nn_ptr, c->nn_ptr.data(), c->nn_ptr.size()))) return rc;
  if ((rc = upload(&c->d_nn_list, c->nn_list.data(), c->nn_list.size()))) return rc;
  if ((rc = upload(&c->d_ecol_elems, c->ecol_elems.data(), c->ecol_elems.size()))) return rc;
  GM_HIP(hipMalloc(&c->d_flag, 8 * sizeof(int)));
  GM_HIP(hipMemset(c->d_flag, 0, 8 * sizeof(int)));
  GM_HIP(hipMalloc(&c->d_red, 2 * 1024 * sizeof(double)));
  GM_HIP(hipHostMalloc(&c->h_red, 128 * 1024 * sizeof(double), hipHostMallocDefault));
  return gm_fill_setup_device(c);
}

extern "C" int gm_ctx_create(const gm_mesh *m, const gm_physics *ph, const gm_bc *bc, gm_ctx **out) {
  if (!out) {
    gm_set_error("null output pointer");
    return GM_ERR_ARG;
  }
  *out = nullptr;
  gm_ctx *c = new gm_ctx();
  int rc = create_impl(m, ph, bc, c);
  if (rc) {
    free_ctx(c);
    return rc;
  }
  *out = c;
  return GM_OK;
}

extern "C" void gm_ctx_destroy(gm_ctx *c) { free_ctx(c); }

extern "C" int gm_num_unknowns(const gm_ctx *c) { return c ? c->N : GM_ERR_ARG; }
extern "C" int gm_num_nodes(const gm_ctx *c) { return c ? c->num_nodes : GM_ERR_ARG; }
extern "C" int gm_num_elems(const gm_ctx *c) { return c ? c->num_elems : GM_ERR_ARG; }
extern "C" int gm_msr_size(const gm_ctx *c) { return c ? (int)c->nval : GM_ERR_ARG; }

extern "C" int gm_copy_msr_graph(const gm_ctx *c, int *bindx) {
  if (!c || !bindx) return GM_ERR_ARG;
  std::memcpy(bindx, c->bindx.data(), c->bindx.size() * sizeof(int));
  return GM_OK;
}

extern "C" int gm_copy_first_unknown(const gm_ctx *c, int *fu) {
  if (!c || !fu) return GM_ERR_ARG;
  std::memcpy(fu, c->first_unknown.data(), c->first_unknown.size() * sizeof(int));
  return GM_OK;
}

extern "C" int gm_copy_dirichlet(const gm_ctx *c, signed char *mask, double *value) {
  if (!c || !mask || !value) return GM_ERR_ARG;
  std::memcpy(mask, c->dbc.data(), c->N);
  std::memcpy(value, c->dbc_val.data(), c->N * sizeof(double));
  return GM_OK;
}

extern "C" int gm_num_elem_colours(const gm_ctx *c) { return c ? c->n_ecol : GM_ERR_ARG; }

extern "C" int gm_copy_elem_colours(const gm_ctx *c, int *colour) {
  if (!c || !colour) return GM_ERR_ARG;
  std::memcpy(colour, c->elem_colour.data(), c->num_elems * sizeof(int));
  return GM_OK;
}

extern "C" int gm_copy_patch_permutation(gm_ctx *c, int bins, int *perm, int *colour_ptr) {
  if (!c || !perm || !colour_ptr) return GM_ERR_ARG;
  std::vector<int> p, cp, pr;
  int ncol = gm_patch_order(c, bins > 0 ? bins : gm_default_bins(c), p, cp, pr);
  if (ncol < 0) return ncol;
  std::memcpy(perm, p.data(), c->N * sizeof(int));
  for (int k = 0; k <= ncol; k++) colour_ptr[k] = pr[cp[k]];  // unknown offsets
  return ncol;
}

extern "C" int gm_set_physics(gm_ctx *c, const gm_physics *ph) {
  if (!c || !ph) return GM_ERR_ARG;
  if ((ph->energy ? 1 : 0) != c->energy) {
    gm_set_error("the energy equation cannot be switched on an existing context");
    return GM_ERR_ARG;
  }
  c->phys = *ph;
  // the multigrid levels re-assemble the same physics on coarser meshes
  for (MgLevel &L : c->mg) L.ctx->phys = *ph;
  return GM_OK;
}

extern "C" int gm_h_elem_avg(gm_ctx *c, double *h) {
  if (!c || !h) return GM_ERR_ARG;
  *h = c->h_elem_avg;
  return GM_OK;
}

extern "C" int gm_velocity_norm_dev(gm_ctx *c, const double *x, double *U) {
  if (!c || !x || !U) return GM_ERR_ARG;
  return gm_velocity_norm(c, x, U);
}

extern "C" int gm_set_velocity_norm(gm_ctx *c, double U) {
  if (!c || !(U >= 0.0)) {
    gm_set_error("the velocity norm must be a non-negative number");
    return GM_ERR_ARG;
  }
  c->U_norm = U;
  for (MgLevel &L : c->mg) L.ctx->U_norm = U;
  return GM_OK;
}

extern "C" int gm_matrix_fill_full_dev(gm_ctx *c, const double *x, double *resid, double *val) {
  if (!c || !x || !resid || !val) return GM_ERR_ARG;
  return gm_fill_launch(c, x, resid, val, nullptr);
}

extern "C" int gm_matrix_fill_full32_dev(gm_ctx *c, const double *x, double *resid, float *val32) {
  if (!c || !x || !resid || !val32) return GM_ERR_ARG;
  return gm_fill_launch(c, x, resid, nullptr, val32);
}

extern "C" int gm_matrix_fill_full(gm_ctx *c, const double *x, double *resid, double *val) {
  if (!c || !x || !resid || !val) return GM_ERR_ARG;
  double *dx = nullptr, *dr = nullptr, *dv = nullptr;
  int rc = GM_OK;
  auto cleanup = [&]() {
    if (dx) hipFree(dx);
    if (dr) hipFree(dr);
    if (dv) hipFree(dv);
  };
  if (hipMalloc(&dx, c->N * sizeof(double)) != hipSuccess ||
      hipMalloc(&dr, c->N * sizeof(double)) != hipSuccess ||
      hipMalloc(&dv, c->nval * sizeof(double)) != hipSuccess) {
    cleanup();
    gm_set_error("device allocation failed");
    return GM_ERR_HIP;
  }
  hipMemcpy(dx, x, c->N * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(dr, resid, c->N * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(dv, val, c->nval * sizeof(double), hipMemcpyHostToDevice);
  rc = gm_fill_launch(c, dx, dr, dv, nullptr);
  if (rc == GM_OK) {
    hipMemcpy(resid, dr, c->N * sizeof(double), hipMemcpyDeviceToHost);
    hipMemcpy(val, dv, c->nval * sizeof(double), hipMemcpyDeviceToHost);
  }
  cleanup();
  return rc;
}

// Sets the theta-method state that the next assemblies read (x_old null: steady).
static int set_transient(gm_ctx *c, const double *x_old, const double *xdot_old, double dt,
                         double theta) {
  if (!x_old) {
    c->tr_x_old = c->tr_xdot_old = nullptr;
    c->tr_cmass = c->tr_cold = 0.0;
    return GM_OK;
  }
  if (!xdot_old || !(dt > 0.0) || !std::isfinite(dt) || !(theta >= 0.0 && theta <= 1.0)) {
    gm_set_error("transient step needs x_old, xdot_old, delta_t > 0 and theta in [0, 1]");
    return GM_ERR_ARG;
  }
  c->tr_x_old = x_old;
  c->tr_xdot_old = xdot_old;
  c->tr_cmass = (1.0 + 2.0 * theta) / dt;
  c->tr_cold = 2.0 * theta;
  return GM_OK;
}

extern "C" int gm_matrix_fill_full_assign_dev(gm_ctx *c, const double *x, double *resid,
                                              double *val) {
  if (!c || !x || !resid || !val) return GM_ERR_ARG;
  int rc;
  if (!c->d_rank && (rc = gm_zero(c, val, c->nval))) return rc;
  c->fill_overwrite = 1;
  rc = gm_fill_launch(c, x, resid, val, nullptr);
  c->fill_overwrite = 0;
  return rc;
}

extern "C" int gm_matrix_fill_full_transient_dev(gm_ctx *c, const double *x, const double *x_old,
                                                 const double *xdot_old, double dt, double theta,
                                                 double *resid, double *val) {
  if (!c || !x || !resid || !val) return GM_ERR_ARG;
  int rc = set_transient(c, x_old, xdot_old, dt, theta);
  if (rc == GM_OK && !x_old) rc = GM_ERR_ARG;
  if (rc) return rc;
  rc = gm_fill_launch(c, x, resid, val, nullptr);
  set_transient(c, nullptr, nullptr, 0.0, 0.0);
  return rc;
}

// Groups the QSIDE / QCONV entries per element (card order kept within an
// element) and uploads them, each with its (q, h, T_c).
static int upload_qside(gm_ctx *c) {
  if (c->d_qs_ptr) hipFree(c->d_qs_ptr);
  if (c->d_qs_side) hipFree(c->d_qs_side);
  if (c->d_qs_q) hipFree(c->d_qs_q);
  c->d_qs_ptr = c->d_qs_side = nullptr;
  c->d_qs_q = nullptr;
  const size_t n = c->qs_elem.size();
  if (n == 0) return GM_OK;
  std::vector<int> ptr(c->num_elems + 1, 0), side(n);
  std::vector<double> q(3 * n);
  for (int e : c->qs_elem) ptr[e + 1]++;
  for (int e = 0; e < c->num_elems; e++) ptr[e + 1] += ptr[e];
  std::vector<int> pos(ptr.begin(), ptr.end() - 1);
  for (size_t k = 0; k < n; k++) {
    const int at = pos[c->qs_elem[k]]++;
    side[at] = c->qs_side[k];
    q[3 * at] = c->qs_q[k];
    q[3 * at + 1] = c->qs_h[k];
    q[3 * at + 2] = c->qs_tc[k];
  }
  int rc;
  if ((rc = upload(&c->d_qs_ptr, ptr.data(), ptr.size()))) return rc;
  if ((rc = upload(&c->d_qs_side, side.data(), n))) return rc;
  return upload(&c->d_qs_q, q.data(), 3 * n);
}

// Side entries adding q + h (T_c - T) to the inward heat flux: QSIDE (h = 0)
// and QCONV (q = 0).
static int add_sides(gm_ctx *c, const char *card, int n, const int *elems, const int *sides,
                     double q, double h, double tc) {
  if (!c->energy) {
    gm_set_error("%s needs the energy equation", card);
    return GM_ERR_ARG;
  }
  std::vector<int> count(c->num_elems, 0);
  for (int e : c->qs_elem) count[e]++;
  for (int k = 0; k < n; k++) {
    if (elems[k] < 0 || elems[k] >= c->num_elems || sides[k] < 1 || sides[k] > 2 * c->dim) {
      gm_set_error("%s entry %d: element %d side %d outside the mesh", card, k, elems[k], sides[k]);
      return GM_ERR_ARG;
    }
    if (++count[elems[k]] > GM_QS_MAX) {
      gm_set_error("%s: more than %d side entries on element %d", card, GM_QS_MAX, elems[k]);
      return GM_ERR_ARG;
    }
  }
  for (int k = 0; k < n; k++) {
    c->qs_elem.push_back(elems[k]);
    c->qs_side.push_back(sides[k]);
    c->qs_q.push_back(q);
    c->qs_h.push_back(h);
    c->qs_tc.push_back(tc);
  }
  return upload_qside(c);
}

extern "C" int gm_add_qside(gm_ctx *c, int n, const int *elems, const int *sides, double q) {
  if (!c || n < 0 || (n > 0 && (!elems || !sides)) || !std::isfinite(q)) return GM_ERR_ARG;
  return add_sides(c, "QSIDE", n, elems, sides, q, 0.0, 0.0);
}

extern "C" int gm_add_qconv(gm_ctx *c, int n, const int *elems, const int *sides, double h,
                            double t_bath) {
  if (!c || n < 0 || (n > 0 && (!elems || !sides)) || !std::isfinite(h) || !std::isfinite(t_bath))
    return GM_ERR_ARG;
  return add_sides(c, "QCONV", n, elems, sides, 0.0, h, t_bath);
}

extern "C" int gm_clear_qside(gm_ctx *c) {
  if (!c) return GM_ERR_ARG;
  c->qs_elem.clear();
  c->qs_side.clear();
  c->qs_q.clear();
  c->qs_h.clear();
  c->qs_tc.clear();
  return upload_qside(c);
}

extern "C" int gm_row_sum_scale_dev(gm_ctx *c, double *val, double *b, double *scale) {
  if (!c || !val || !b || !scale) return GM_ERR_ARG;
  int rc = gm_scale_launch(c, val, nullptr, b, scale);
  if (rc) return rc;
  GM_HIP(hipStreamSynchronize(c->stream));
  return GM_OK;
}

extern "C" int gm_msr_matvec_dev(gm_ctx *c, const double *val, const double *x, double *y) {
  if (!c || !val || !x || !y) return GM_ERR_ARG;
  int rc = gm_spmv_launch(c, val, x, y);
  if (rc) return rc;
  GM_HIP(hipStreamSynchronize(c->stream));
  return GM_OK;
}

extern "C" int gm_gmres_dev(gm_ctx *c, const double *val, const double *b, double *x,
                            const gm_solver_opts *o, int *its, double *relres) {
  if (!c || !val || !b || !x || !o || !its || !relres) return GM_ERR_ARG;
  if (o->precision != GM_PREC_FP64) {
    gm_set_error("gm_gmres_dev takes an FP64 matrix");
    return GM_ERR_ARG;
  }
  int rc = gm_gmres_run(c, val, nullptr, b, x, o, its, relres);
  if (rc) return rc;
  GM_HIP(hipStreamSynchronize(c->stream));
  return GM_OK;
}

extern "C" int gm_ilu0_factor_dev(gm_ctx *c, const double *val, int bins, int *rowptr, int *col,
                                  double *lu) {
  if (!c || !val || !rowptr || !col || !lu) return GM_ERR_ARG;
  return gm_ilu_factor_export(c, val, bins > 0 ? bins : gm_default_bins(c), rowptr, col, lu);
}

static int ensure_newton_buffers(gm_ctx *c, int prec) {
  if (!c->d_resid) {
    GM_HIP(hipMalloc(&c->d_resid, c->N * sizeof(double)));
    GM_HIP(hipMalloc(&c->d_dx, c->N * sizeof(double)));
    GM_HIP(hipMalloc(&c->d_scale, c->N * sizeof(double)));
  }
  if (!c->d_xsave) {
    GM_HIP(hipMalloc(&c->d_xsave, c->N * sizeof(double)));
    GM_HIP(hipMalloc(&c->d_rtrial, c->N * sizeof(double)));
    GM_HIP(hipMalloc(&c->d_w, c->N * sizeof(double)));
  }
  if (prec == GM_PREC_MIXED) {
    if (!c->d_val32) GM_HIP(hipMalloc(&c->d_val32, c->nval * sizeof(float)));
  } else {
    if (!c->d_val) GM_HIP(hipMalloc(&c->d_val, c->nval * sizeof(double)));
  }
  return GM_OK;
}

// perform_backtrack_line_search (src/mm_sol_nonlinear.c:344-590): try the
// full step, then halve the step while the scaled residual fails
// 0.5|R|^2 <= 0.5|R0|^2 + 0.5 slope damp, slope = -|(J dx).R(x - dx)|, down to
// the minimum damping; keep the step with the smallest residual norm.
// Residual-only assemblies, scaled with the current row sums.
static int backtrack(gm_ctx *c, double *x, const gm_solver_opts *o, double L2, double *best_out,
                     int *trials) {
  int rc;
  const double eps0 = o->eps_resid;
  GM_HIP(hipMemcpyAsync(c->d_xsave, x, c->N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  auto trial = [&](double damp, double *g) -> int {
    int r;
    (*trials)++;
    if ((r = gm_lincomb(c, x, c->d_xsave, damp, c->d_dx))) return r;
    if ((r = gm_zero(c, c->d_rtrial, c->N))) return r;
    if ((r = gm_fill_launch_ex(c, x, c->d_rtrial, nullptr, nullptr, 1))) return r;
    if ((r = gm_div_scale(c, c->d_rtrial, c->d_scale))) return r;
    double loo;
    return gm_norms(c, c->d_rtrial, c->N, g, &loo);
  };
  double damp = 1.0, g;
  if ((rc = trial(damp, &g))) return rc;
  double best_damp = damp, best_norm = g;
  const double r_check = 0.5 * L2 * L2;
  if (o->precision == GM_PREC_MIXED) rc = gm_spmv32_launch(c, c->d_val32, c->d_dx, c->d_w);
  else rc = gm_spmv_launch(c, c->d_val, c->d_dx, c->d_w);
  if (rc) return rc;
  double slope;
  if ((rc = gm_dot(c, c->d_w, c->d_rtrial, &slope))) return rc;
  if (slope > 0) slope = -slope;
  else if (slope == 0) slope = -1;
  bool skip = best_norm < eps0;
  while (!skip) {
    damp *= 0.5;
    if (damp < o->min_damp) break;
    if ((rc = trial(damp, &g))) return rc;
    if (std::isnan(g)) break;
    if (g < best_norm) {
      best_damp = damp;
      best_norm = g;
    }
    if (best_norm < eps0) break;
    if (0.5 * g * g <= r_check + 0.5 * slope * damp) break;
  }
  if ((rc = gm_lincomb(c, x, c->d_xsave, best_damp, c->d_dx))) return rc;
  *best_out = best_damp;
  return GM_OK;
}

// solve_nonlinear_problem, steady, NLS_FULL_STEP or NLS_BACKTRACK
// (src/mm_sol_nonlinear.c:1108-2913)
// watch_its > 0: stop early (*stalled = true) when, at pass watch_its, no
// pass since the first has brought ||D^-1 R||_2 below its first value / drop.
static int solve_impl(gm_ctx *c, double *x, const gm_solver_opts *o, gm_newton_report *rep,
                      int watch_its, double drop, bool *stalled) {
  *stalled = false;
  std::memset(rep, 0, sizeof(*rep));
  const int prec = o->precision;
  int rc = ensure_newton_buffers(c, prec);
  if (rc) return rc;
  const double t0 = now_s();
  int converged = 0, inewton = 0;
  bool have_pc = false, last_ok = true;  // preconditioner reuse state
  int its_ref = 0, last_its = 0;
  double eta = 0.1, L2_prev = -1.0;  // Eisenstat-Walker forcing term and the last scaled norm
  const bool galerkin = o->precond == GM_PC_MG && o->mg_galerkin && !c->mg.empty();
  if (galerkin && (rc = gm_mg_kel_alloc(c))) return rc;
  while (!converged && inewton < o->max_newton) {
    double ta = now_s();
    // the PSPG velocity norm of this pass (src/mm_sol_nonlinear.c:1183-1192);
    // the line search's residuals keep it
    if (c->phys.pspg) {
      if ((rc = gm_velocity_norm(c, x, &c->U_norm))) return rc;
      for (MgLevel &L : c->mg) L.ctx->U_norm = c->U_norm;
    }
    if ((rc = gm_zero(c, c->d_resid, c->N))) return rc;
    if ((rc = gm_zero(c, c->d_dx, c->N))) return rc;
    // the HEX27 assembly stores each node pair's first contribution, so the
    // Jacobian needs no zeroing pass
    const bool assign = c->d_rank != nullptr;
    if (!assign) {
      if (prec == GM_PREC_MIXED) rc = gm_zero32(c, c->d_val32, c->nval);
      else rc = gm_zero(c, c->d_val, c->nval);
      if (rc) return rc;
    }
    c->kel_on = galerkin;
    c->fill_overwrite = assign;
    rc = gm_fill_launch(c, x, c->d_resid, c->d_val, prec == GM_PREC_MIXED ? c->d_val32 : nullptr);
    c->kel_on = 0;
    c->fill_overwrite = 0;
    if (rc) return rc;
    double un, unoo;
    if ((rc = gm_norms(c, c->d_resid, c->N, &un, &unoo))) return rc;
    double tb = now_s();
    rep->t_assemble += tb - ta;
    if ((rc = gm_scale_launch(c, c->d_val, prec == GM_PREC_MIXED ? c->d_val32 : nullptr, c->d_resid,
                              c->d_scale)))
      return rc;
    double L2, Loo;
    if ((rc = gm_norms(c, c->d_resid, c->N, &L2, &Loo))) return rc;
    double tc = now_s();
    rep->t_scale += tc - tb;
    if (inewton < GM_MAX_REPORT) {
      rep->resid_L2[inewton] = L2;
      rep->resid_Loo[inewton] = Loo;
      rep->unscaled_L2[inewton] = un;
    }
    if (inewton > 0 && L2 < o->eps_resid && Loo < o->eps_resid) converged = o->eps_corr > 1.0;
    if (watch_its > 0 && inewton == watch_its && !converged && watch_its < GM_MAX_REPORT) {
      double best = L2;
      for (int i = 1; i < inewton; i++) best = std::min(best, rep->resid_L2[i]);
      if (!(best * drop <= rep->resid_L2[0])) {
        *stalled = true;
        rep->newton_its = inewton;
        rep->t_total = now_s() - t0;
        return inewton;
      }
    }
    if (!std::isfinite(L2) || !std::isfinite(Loo)) {
      gm_set_error("non-finite residual norm");
      rep->newton_its = inewton;
      rep->t_total = now_s() - t0;
      return GM_ERR_NAN;
    }
    if (!converged) {
      int its = 0;
      double rr = 0.0;
      gm_solver_opts oo = *o;
      if (o->lin_tol_adapt) {
        // Eisenstat-Walker choice 2 with its safeguard (SIAM J. Sci. Comput. 17, 1996),
        // exponent (1 + sqrt 5) / 2 (measured: fewer GMRES iterations than 2 at 256^2 and 48^3)
        const double emax = 0.1, gam = 0.9, alpha = 1.6180339887498949;
        double e = emax;
        if (L2_prev > 0.0) {
          e = gam * std::pow(L2 / L2_prev, alpha);
          const double safe = gam * std::pow(eta, alpha);
          if (safe > 0.1) e = std::max(e, safe);
        }
        eta = std::min(emax, std::max(e, o->lin_tol));
        oo.lin_tol = eta;
      }
      L2_prev = L2;
      const bool rebuild = !have_pc || o->pc_reuse == 0 ||
                           (o->pc_reuse == 2 && (!last_ok || last_its > 1.5 * its_ref + 2));
      // preconditioner set-up timed separately, then GMRES
      if (rebuild && o->precond == GM_PC_ILU0) {
        const int bins = o->patch_bins > 0 ? o->patch_bins : gm_default_bins(c);
        if ((rc = gm_ilu_plan(c, bins))) return rc;
        if ((rc = gm_ilu_factor(c, c->d_val, c->d_val32, prec))) return rc;
      } else if (rebuild && o->precond == GM_PC_BAND) {
        if ((rc = gm_band_factor(c, c->d_val, c->d_val32, prec))) return rc;
      } else if (rebuild && o->precond == GM_PC_MG) {
        if ((rc = gm_mg_setup(c, x, c->d_val, c->d_val32, prec, c->d_scale,
                              o->mg_coarse_sweeps <= 0, galerkin)))
          return rc;
        GM_HIP(hipStreamSynchronize(c->stream));
      }
      double td = now_s();
      rep->t_precond += td - tc;
      oo.precond = o->precond;
      rc = gm_gmres_run_nofactor(c, c->d_val, c->d_val32, c->d_resid, c->d_dx, &oo, &its, &rr);
      if (rc) return rc;
      double te = now_s();
      rep->t_solve += te - td;
      if (inewton < GM_MAX_REPORT) {
        rep->lin_its[inewton] = its;
        rep->lin_relres[inewton] = rr;
      }
      rep->lin_its_total += its;
      if (rebuild) {
        have_pc = true;
        its_ref = its;
        rep->pc_builds++;
      }
      last_its = its;
      last_ok = rr <= o->lin_tol;
      double cL2r, cLoor;
      if ((rc = gm_corr_norms(c, c->d_dx, x, &cL2r, &cLoor))) return rc;
      double taken = o->damp;
      if (o->line_search) {
        if ((rc = backtrack(c, x, o, L2, &taken, &rep->resid_fills))) return rc;
      } else if ((rc = gm_axpy_update(c, x, c->d_dx, o->damp))) {
        return rc;
      }
      if (inewton < GM_MAX_REPORT) rep->damp[inewton] = taken;
      GM_HIP(hipStreamSynchronize(c->stream));
      rep->t_update += now_s() - te;
      converged = (L2 < o->eps_resid) && (Loo < o->eps_resid) && (cL2r < o->eps_corr) &&
                  (cLoor < o->eps_corr);
    }
    if (o->verbose)
      fprintf(stderr,
              "[gm] newton %2d |R|2 %.3e |R|oo %.3e unscaled %.3e lin_its %d relres %.2e damp %.4f\n",
              inewton, L2, Loo, un, inewton < GM_MAX_REPORT ? rep->lin_its[inewton] : -1,
              inewton < GM_MAX_REPORT ? rep->lin_relres[inewton] : 0.0,
              inewton < GM_MAX_REPORT ? rep->damp[inewton] : 0.0);
    inewton++;
  }
  rep->newton_its = inewton;
  rep->converged = converged;
  rep->t_total = now_s() - t0;
  return inewton;
}

extern "C" int gm_solve_nonlinear_dev(gm_ctx *c, double *x, const gm_solver_opts *o,
                                      gm_newton_report *rep) {
  if (!c || !x || !o || !rep) return GM_ERR_ARG;
  bool stalled = false;
  return solve_impl(c, x, o, rep, 0, 0.0, &stalled);
}

extern "C" int gm_solve_nonlinear_transient_dev(gm_ctx *c, double *x, const double *x_old,
                                                const double *xdot_old, double dt, double theta,
                                                const gm_solver_opts *o, gm_newton_report *rep) {
  if (!c || !x || !x_old || !o || !rep) return GM_ERR_ARG;
  int rc = set_transient(c, x_old, xdot_old, dt, theta);
  if (rc) return rc;
  bool stalled = false;
  rc = solve_impl(c, x, o, rep, 0, 0.0, &stalled);
  set_transient(c, nullptr, nullptr, 0.0, 0.0);
  return rc;
}

extern "C" int gm_solve_transient_dev(gm_ctx *c, double *x, double *xdot,
                                      const gm_transient_opts *t, const gm_solver_opts *o,
                                      gm_transient_report *rep) {
  if (!c || !x || !xdot || !t || !o || !rep) return GM_ERR_ARG;
  std::memset(rep, 0, sizeof(*rep));
  const double dt = t->delta_t;
  if (!(dt > 0.0) || !std::isfinite(dt) || !(t->theta >= 0.0 && t->theta <= 1.0) ||
      t->num_steps < 0 || t->startup_steps < 0) {
    gm_set_error("transient options: delta_t > 0, theta in [0, 1], step counts >= 0");
    return GM_ERR_ARG;
  }
  const double t0 = now_s();
  const size_t nb = (size_t)c->N * sizeof(double);
  // history: x_old, x_older, x_oldest, xdot_old, xdot_older
  double *h[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  auto release = [&]() {
    for (double *p : h)
      if (p) hipFree(p);
  };
  for (double *&p : h)
    if (hipMalloc(&p, nb) != hipSuccess) {
      release();
      gm_set_error("device allocation failed");
      return GM_ERR_HIP;
    }
  double *x_old = h[0], *x_older = h[1], *x_oldest = h[2], *xd_old = h[3], *xd_older = h[4];
  int rc = GM_OK;
  auto run = [&]() -> int {
    // the initial condition fills the history (src/rf_solve.c:1222-1229)
    GM_HIP(hipMemcpyAsync(x_old, x, nb, hipMemcpyDeviceToDevice, c->stream));
    GM_HIP(hipMemcpyAsync(x_older, x, nb, hipMemcpyDeviceToDevice, c->stream));
    GM_HIP(hipMemcpyAsync(x_oldest, x, nb, hipMemcpyDeviceToDevice, c->stream));
    GM_HIP(hipMemcpyAsync(xd_old, xdot, nb, hipMemcpyDeviceToDevice, c->stream));
    int r;
    if ((r = gm_zero(c, xd_older, c->N))) return r;
    for (int n = 0; n < t->num_steps; n++) {
      // backward Euler for the start-up steps, then the requested scheme
      // (src/rf_solve.c:1378-1386)
      const double th = n < t->startup_steps ? 0.0 : t->theta;
      // predict_solution (src/rf_solve.c:2767-2830) with equal steps
      if (th == 0.5) r = gm_lin3(c, x, 3.0, x_old, -3.0, x_older, 1.0, x_oldest);
      else r = gm_lin3(c, x, 1.0, x_old, dt * (1.0 + th), xd_old, -th * dt, xd_older);
      if (r) return r;
      gm_newton_report nr;
      if ((r = gm_solve_nonlinear_transient_dev(c, x, x_old, xd_old, dt, th, o, &nr)) < 0) return r;
      rep->newton_its_total += nr.newton_its;
      rep->lin_its_total += nr.lin_its_total;
      if (n < GM_MAX_REPORT) rep->newton_its[n] = nr.newton_its;
      if (!nr.converged) {
        gm_set_error("time step %d (t = %g): Newton did not converge", n, (n + 1) * dt);
        return GM_ERR_CONV;
      }
      // xdot consistent with the converged x, then the history shift
      // (src/rf_solve.c:2359-2365)
      const double cm = (1.0 + 2.0 * th) / dt;
      if ((r = gm_lin3(c, xdot, cm, x, -cm, x_old, -2.0 * th, xd_old))) return r;
      std::swap(xd_older, xd_old);
      GM_HIP(hipMemcpyAsync(xd_old, xdot, nb, hipMemcpyDeviceToDevice, c->stream));
      std::swap(x_oldest, x_older);
      std::swap(x_older, x_old);
      GM_HIP(hipMemcpyAsync(x_old, x, nb, hipMemcpyDeviceToDevice, c->stream));
      rep->steps = n + 1;
      rep->time = (n + 1) * dt;
    }
    GM_HIP(hipStreamSynchronize(c->stream));
    return GM_OK;
  };
  rc = run();
  hipStreamSynchronize(c->stream);
  release();
  rep->t_total = now_s() - t0;
  return rc < 0 ? rc : rep->steps;
}

extern "C" int gm_solve_nonlinear_guess_dev(gm_ctx *c, double *x, const double *x_prev,
                                            const gm_solver_opts *o, int watch_its, double drop,
                                            gm_newton_report *rep) {
  if (!c || !x || !x_prev || !o || !rep || watch_its < 1 || watch_its >= GM_MAX_REPORT ||
      !(drop > 1.0))
    return GM_ERR_ARG;
  bool stalled = false;
  int rc = solve_impl(c, x, o, rep, watch_its, drop, &stalled);
  if (rc < 0 || !stalled) return rc;
  const gm_newton_report first = *rep;
  GM_HIP(hipMemcpyAsync(x, x_prev, (size_t)c->N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  rc = solve_impl(c, x, o, rep, 0, 0.0, &stalled);
  rep->fell_back = 1;
  rep->discarded_its = first.newton_its;
  rep->lin_its_total += first.lin_its_total;
  rep->pc_builds += first.pc_builds;
  rep->resid_fills += first.resid_fills;
  rep->t_assemble += first.t_assemble;
  rep->t_scale += first.t_scale;
  rep->t_precond += first.t_precond;
  rep->t_solve += first.t_solve;
  rep->t_update += first.t_update;
  rep->t_total += first.t_total;
  return rc;
}

extern "C" int gm_solve_nonlinear(gm_ctx *c, double *x, const gm_solver_opts *o,
                                  gm_newton_report *rep) {
  if (!c || !x) return GM_ERR_ARG;
  if (!c->d_xwork) GM_HIP(hipMalloc(&c->d_xwork, c->N * sizeof(double)));
  GM_HIP(hipMemcpy(c->d_xwork, x, c->N * sizeof(double), hipMemcpyHostToDevice));
  int rc = gm_solve_nonlinear_dev(c, c->d_xwork, o, rep);
  GM_HIP(hipMemcpy(x, c->d_xwork, c->N * sizeof(double), hipMemcpyDeviceToHost));
  return rc;
}

extern "C" int gm_prof_enable(gm_ctx *c, int on) {
  if (!c) return GM_ERR_ARG;
  if (c->stream) hipStreamSynchronize(c->stream);
  c->prof = on != 0;
  c->prof_every = on > 1 ? on : 1;
  for (int i = 0; i < GM_K_COUNT; i++) {
    c->prof_count[i] = 0;
    c->prof_ms[i] = 0.0;
    c->prof_bytes[i] = 0.0;
    c->prof_sampled_bytes[i] = 0.0;
    c->prof_seen[i] = 0;
    c->ev_used[i] = 0;
  }
  return GM_OK;
}

extern "C" int gm_prof_get(gm_ctx *c, int id, int *count, double *ms, double *bytes) {
  if (!c || id < 0 || id >= GM_K_COUNT || !count || !ms) return GM_ERR_ARG;
  prof_fold(c, id);
  // time of all launches = sampled time x launches / sampled launches (the
  // sample is pseudo-random, so this is unbiased over launches of any size)
  const long long seen = c->prof_seen[id], sampled = c->prof_count[id];
  const double scale = sampled ? (double)seen / sampled : 0.0;
  *count = (int)seen;
  *ms = c->prof_ms[id] * scale;
  if (bytes) *bytes = c->prof_bytes[id];
  return GM_OK;
}
