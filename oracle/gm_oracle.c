/*
 * gm_oracle.c -- FP64 CPU restatement of Goma's Newton inner loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never part of the product).
 * Every function cites the reference file:line it restates.  The structure
 * deliberately follows Goma's: a dense per-element store (lec) filled term by
 * term per quadrature point, Dirichlet rows replaced per element, then a
 * scatter into the MSR matrix that finds each column with a linear scan.
 */
#include "gm_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define MDE 27  /* max dofs per variable per element (HEX27) */
#define NSLOT 5 /* local variable slots: U1 U2 U3 T P */
#define SLOT_T 3
#define SLOT_P 4

struct om_problem {
  int dim, nve, num_nodes, num_elems;
  double *coor;
  int *conn;
  om_physics ph;
  int N;
  int *first_unknown; /* [num_nodes+1] */
  int *unk_var;       /* [N] */
  int *is_centroid;   /* [num_nodes] element index or -1 */
  int *elem_centroid; /* [num_elems] centroid node */
  char *dbc;
  double *dbc_val;
  int *nn_ptr, *nn_list; /* node-node graph, sorted, self included */
  int *ija;              /* MSR */
  int n_qs;              /* QSIDE heat-flux sides: element, side (1-based), flux */
  int *qs_elem, *qs_side;
  double *qs_q, *qs_h, *qs_tc; /* flux; QCONV heat transfer coefficient and bath temperature */
  double h_elem_avg; /* PSPG: global mean element size */
  double U_norm;     /* PSPG: global velocity norm (mean square) */
};

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Basis functions and quadrature                                            */
/* ------------------------------------------------------------------------ */

/* QUAD9 shape functions, closed forms of src/rf_shape.c:361-470. */
static void shape_quad9(double s, double t, double *phi, double dphi[][3]) {
  phi[0] = .25 * (1.0 - s) * (1.0 - t) * (-s - t - 1.0) + .25 * (1.0 - s * s) * (1.0 - t * t);
  phi[1] = .25 * (1.0 + s) * (1.0 - t) * (s - t - 1.0) + .25 * (1.0 - s * s) * (1.0 - t * t);
  phi[2] = .25 * (1.0 + s) * (1.0 + t) * (s + t - 1.0) + .25 * (1.0 - s * s) * (1.0 - t * t);
  phi[3] = .25 * (1.0 - s) * (1.0 + t) * (-s + t - 1.0) + .25 * (1.0 - s * s) * (1.0 - t * t);
  phi[4] = .5 * (1.0 - s * s) * (1.0 - t) - .5 * (1.0 - s * s) * (1.0 - t * t);
  phi[5] = .5 * (1.0 + s) * (1.0 - t * t) - .5 * (1.0 - s * s) * (1.0 - t * t);
  phi[6] = .5 * (1.0 - s * s) * (1.0 + t) - .5 * (1.0 - s * s) * (1.0 - t * t);
  phi[7] = .5 * (1.0 - s) * (1.0 - t * t) - .5 * (1.0 - s * s) * (1.0 - t * t);
  phi[8] = (1.0 - s * s) * (1.0 - t * t);
  dphi[0][0] = -.25 * (1.0 - t) * (-2.0 * s - t) - 0.5 * s * (1 - t * t);
  dphi[1][0] = .25 * (1.0 - t) * (2.0 * s - t) - 0.5 * s * (1 - t * t);
  dphi[2][0] = .25 * (1.0 + t) * (2.0 * s + t) - 0.5 * s * (1 - t * t);
  dphi[3][0] = -.25 * (1.0 + t) * (-2.0 * s + t) - 0.5 * s * (1 - t * t);
  dphi[4][0] = -s * (1.0 - t) + s * (1 - t * t);
  dphi[5][0] = .5 * (1.0 - t * t) + s * (1 - t * t);
  dphi[6][0] = -s * (1.0 + t) + s * (1 - t * t);
  dphi[7][0] = -.5 * (1.0 - t * t) + s * (1 - t * t);
  dphi[8][0] = -2.0 * s * (1.0 - t * t);
  dphi[0][1] = -.25 * (1.0 - s) * (-s - 2.0 * t) - 0.5 * t * (1.0 - s * s);
  dphi[1][1] = -.25 * (1.0 + s) * (s - 2.0 * t) - 0.5 * t * (1.0 - s * s);
  dphi[2][1] = .25 * (1.0 + s) * (s + 2.0 * t) - 0.5 * t * (1.0 - s * s);
  dphi[3][1] = .25 * (1.0 - s) * (-s + 2.0 * t) - 0.5 * t * (1.0 - s * s);
  dphi[4][1] = -.5 * (1.0 - s * s) + t * (1.0 - s * s);
  dphi[5][1] = -(1.0 + s) * t + t * (1.0 - s * s);
  dphi[6][1] = .5 * (1.0 - s * s) + t * (1.0 - s * s);
  dphi[7][1] = -(1.0 - s) * t + t * (1.0 - s * s);
  dphi[8][1] = -2.0 * t * (1.0 - s * s);
  for (int i = 0; i < 9; i++) dphi[i][2] = 0.0;
}

/* HEX27 local node positions in reference coordinates, read off the closed
 * forms at src/rf_shape.c:1105-1196 (node 20 is the centroid). */
static const signed char hex27_pos[27][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
    {-1, 1, 1},   {0, -1, -1}, {1, 0, -1}, {0, 1, -1},  {-1, 0, -1}, {-1, -1, 0}, {1, -1, 0},
    {1, 1, 0},    {-1, 1, 0},  {0, -1, 1}, {1, 0, 1},   {0, 1, 1},   {-1, 0, 1},  {0, 0, 0},
    {0, 0, -1},   {0, 0, 1},   {-1, 0, 0}, {1, 0, 0},   {0, -1, 0},  {0, 1, 0}};

/* 1D quadratic Lagrange factors on {-1,0,1}; the HEX27 closed forms of
 * src/rf_shape.c:1105-1460 are products of these. */
static void lag1(int node, double s, double *v, double *dv) {
  if (node < 0) {
    *v = -0.5 * s * (1.0 - s);
    *dv = -0.5 + s;
  } else if (node > 0) {
    *v = 0.5 * s * (1.0 + s);
    *dv = 0.5 + s;
  } else {
    *v = 1.0 - s * s;
    *dv = -2.0 * s;
  }
}

static void shape_hex27(double s, double t, double u, double *phi, double dphi[][3]) {
  for (int i = 0; i < 27; i++) {
    double a, da, b, db, c, dc;
    lag1(hex27_pos[i][0], s, &a, &da);
    lag1(hex27_pos[i][1], t, &b, &db);
    lag1(hex27_pos[i][2], u, &c, &dc);
    phi[i] = a * b * c;
    dphi[i][0] = da * b * c;
    dphi[i][1] = a * db * c;
    dphi[i][2] = a * b * dc;
  }
}

/* find_stu / Gq_weight for BIQUAD_QUAD and TRIQUAD_HEX,
 * src/el_elm_info.c:1748-1763, 1802-1823, 3555-3571, 3658-3680. */
static void gauss_point(int dim, int iq, double *s, double *t, double *u, double *w) {
  static const double F2 = 0.77459666924148340428;
  static const double w1 = 0.55555555555555555556;
  static const double w2 = 0.88888888888888888888;
  int is = iq % 3, it = (iq % 9) / 3, iu = iq / 9;
  const double pt[3] = {F2, 0.0, -F2};
  const double wt[3] = {w1, w2, w1};
  *s = pt[is];
  *t = pt[it];
  *u = (dim == 3) ? pt[iu] : 0.0;
  *w = wt[is] * wt[it] * ((dim == 3) ? wt[iu] : 1.0);
}


/* Basis/quadrature accessors for tests. */
void om_shape(int dim, double s, double t, double u, double *phi, double *dphi) {
  double d[MDE][3];
  if (dim == 2) shape_quad9(s, t, phi, d);
  else shape_hex27(s, t, u, phi, d);
  int n = dim == 2 ? 9 : 27;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) dphi[i * 3 + k] = d[i][k];
}

void om_gauss(int dim, int iq, double *stuw) {
  gauss_point(dim, iq, &stuw[0], &stuw[1], &stuw[2], &stuw[3]);
}

/* ------------------------------------------------------------------------ */
/* Problem set-up: unknown map, node graph, MSR graph, Dirichlet map         */
/* ------------------------------------------------------------------------ */

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* Interaction mask restated from set_interaction_masks,
 * src/mm_unknown_map.c:1226-1260 (momentum), 1692-1720 (energy: no P),
 * 2276-2310 (continuity: U, T, P). */
static int inter_mask(int row_var, int col_var) {
  if (row_var == OM_T) return col_var != OM_P;
  return 1;
}

static int node_nunk(const om_problem *p, int node) {
  return p->first_unknown[node + 1] - p->first_unknown[node];
}

om_problem *om_create(const om_mesh *m, const om_physics *ph, const om_bc *bc) {
  om_problem *p = (om_problem *)calloc(1, sizeof(om_problem));
  p->dim = m->dim;
  p->nve = (m->dim == 2) ? 9 : 27;
  p->num_nodes = m->num_nodes;
  p->num_elems = m->num_elems;
  p->ph = *ph;
  size_t nc = (size_t)m->dim * m->num_nodes, ne = (size_t)m->num_elems * p->nve;
  p->coor = (double *)malloc(nc * sizeof(double));
  memcpy(p->coor, m->coor, nc * sizeof(double));
  p->conn = (int *)malloc(ne * sizeof(int));
  memcpy(p->conn, m->connect, ne * sizeof(int));

  /* P1 pressure dofs live on the element centroid node
   * (QUAD9 local node 8, HEX27 local node 20; src/el_elm_info.c:935-938, 1109). */
  int cl = (m->dim == 2) ? 8 : 20;
  p->is_centroid = (int *)malloc(m->num_nodes * sizeof(int));
  for (int n = 0; n < m->num_nodes; n++) p->is_centroid[n] = -1;
  p->elem_centroid = (int *)malloc(m->num_elems * sizeof(int));
  for (int e = 0; e < m->num_elems; e++) {
    int c = p->conn[(size_t)e * p->nve + cl];
    p->elem_centroid[e] = c;
    p->is_centroid[c] = e;
  }
  /* Node-major unknown numbering: at each node the variable types in
   * ascending order U1,U2,(U3),(T),P (fill_variable_vector,
   * src/mm_fill_util.c:3516; Index_Solution src/mm_unknown_map.c:5268). */
  int d = m->dim, en = ph->energy ? 1 : 0;
  p->first_unknown = (int *)malloc((m->num_nodes + 1) * sizeof(int));
  p->first_unknown[0] = 0;
  for (int n = 0; n < m->num_nodes; n++) {
    int k = d + en + ((p->is_centroid[n] >= 0) ? d + 1 : 0);
    p->first_unknown[n + 1] = p->first_unknown[n] + k;
  }
  p->N = p->first_unknown[m->num_nodes];
  p->unk_var = (int *)malloc(p->N * sizeof(int));
  for (int n = 0; n < m->num_nodes; n++) {
    int f = p->first_unknown[n], k = 0;
    for (int a = 0; a < d; a++) p->unk_var[f + k++] = OM_U1 + a;
    if (en) p->unk_var[f + k++] = OM_T;
    if (p->is_centroid[n] >= 0)
      for (int a = 0; a <= d; a++) p->unk_var[f + k++] = OM_P;
  }

  /* build_node_node (src/exo_conn.c:204-377): nodes sharing an element,
   * sorted, the node itself included. */
  int *ne_cnt = (int *)calloc(m->num_nodes + 1, sizeof(int));
  for (size_t i = 0; i < ne; i++) ne_cnt[p->conn[i] + 1]++;
  for (int n = 0; n < m->num_nodes; n++) ne_cnt[n + 1] += ne_cnt[n];
  int *ne_list = (int *)malloc(ne * sizeof(int));
  int *fillp = (int *)malloc(m->num_nodes * sizeof(int));
  memcpy(fillp, ne_cnt, m->num_nodes * sizeof(int));
  for (int e = 0; e < m->num_elems; e++)
    for (int i = 0; i < p->nve; i++) ne_list[fillp[p->conn[(size_t)e * p->nve + i]]++] = e;
  free(fillp);
  p->nn_ptr = (int *)malloc((m->num_nodes + 1) * sizeof(int));
  size_t cap = (size_t)m->num_nodes * 32, len = 0;
  p->nn_list = (int *)malloc(cap * sizeof(int));
  int *tmp = (int *)malloc(8 * 27 * sizeof(int) + 64);
  p->nn_ptr[0] = 0;
  for (int n = 0; n < m->num_nodes; n++) {
    int c = 0;
    for (int k = ne_cnt[n]; k < ne_cnt[n + 1]; k++) {
      int e = ne_list[k];
      for (int i = 0; i < p->nve; i++) tmp[c++] = p->conn[(size_t)e * p->nve + i];
    }
    qsort(tmp, c, sizeof(int), cmp_int);
    int u = 0;
    for (int i = 0; i < c; i++)
      if (u == 0 || tmp[i] != tmp[u - 1]) tmp[u++] = tmp[i];
    if (len + u > cap) {
      cap = cap * 2 + u;
      p->nn_list = (int *)realloc(p->nn_list, cap * sizeof(int));
    }
    memcpy(p->nn_list + len, tmp, u * sizeof(int));
    len += u;
    p->nn_ptr[n + 1] = (int)len;
  }
  free(tmp);
  free(ne_cnt);
  free(ne_list);

  /* find_MSR_problem_graph (src/mm_fill_util.c:3229-3445): for each row
   * (node-major), the columns are the unknowns of every neighbour node in
   * sorted node order, filtered by the interaction mask, diagonal excluded.
   * MSR: ija[0..N] are row pointers into ija/val, ija[N+1..] column ids. */
  size_t nnz_off = 0;
  for (int n = 0; n < m->num_nodes; n++) {
    for (int u = p->first_unknown[n]; u < p->first_unknown[n + 1]; u++) {
      int rv = p->unk_var[u];
      for (int k = p->nn_ptr[n]; k < p->nn_ptr[n + 1]; k++) {
        int J = p->nn_list[k];
        for (int w = p->first_unknown[J]; w < p->first_unknown[J + 1]; w++)
          if (inter_mask(rv, p->unk_var[w]) && w != u) nnz_off++;
      }
    }
  }
  size_t tot = (size_t)p->N + 1 + nnz_off;
  p->ija = (int *)malloc(tot * sizeof(int));
  size_t ptr = (size_t)p->N + 1;
  for (int n = 0; n < m->num_nodes; n++) {
    for (int u = p->first_unknown[n]; u < p->first_unknown[n + 1]; u++) {
      int rv = p->unk_var[u];
      p->ija[u] = (int)ptr;
      for (int k = p->nn_ptr[n]; k < p->nn_ptr[n + 1]; k++) {
        int J = p->nn_list[k];
        for (int w = p->first_unknown[J]; w < p->first_unknown[J + 1]; w++)
          if (inter_mask(rv, p->unk_var[w]) && w != u) p->ija[ptr++] = w;
      }
    }
  }
  p->ija[p->N] = (int)ptr;

  /* Dirichlet map per unknown.  Later conditions on the same unknown
   * override earlier ones.  The pressure datum is a Dirichlet condition on
   * the first pressure dof of the datum element (find_and_set_Dirichlet,
   * src/mm_bc.c:386-457). */
  p->dbc = (char *)calloc(p->N, 1);
  p->dbc_val = (double *)calloc(p->N, sizeof(double));
  for (int i = 0; i < bc->num_dbc; i++) {
    int n = bc->node[i], v = bc->var[i], off = -1;
    if (n < 0 || n >= m->num_nodes) continue;
    if (v >= OM_U1 && v < OM_U1 + d) off = v - OM_U1;
    else if (v == OM_T && en) off = d;
    if (off < 0) continue;
    int u = p->first_unknown[n] + off;
    p->dbc[u] = 1;
    p->dbc_val[u] = bc->value[i];
  }
  if (bc->datum_elem >= 0 && bc->datum_elem < m->num_elems) {
    int c = p->elem_centroid[bc->datum_elem];
    int u = p->first_unknown[c] + d + en;
    p->dbc[u] = 1;
    p->dbc_val[u] = bc->datum_value;
  }
  p->h_elem_avg = om_h_elem_avg(p);
  return p;
}

void om_destroy(om_problem *p) {
  if (!p) return;
  free(p->coor);
  free(p->conn);
  free(p->first_unknown);
  free(p->unk_var);
  free(p->is_centroid);
  free(p->elem_centroid);
  free(p->dbc);
  free(p->dbc_val);
  free(p->nn_ptr);
  free(p->nn_list);
  free(p->ija);
  free(p->qs_elem);
  free(p->qs_side);
  free(p->qs_q);
  free(p->qs_h);
  free(p->qs_tc);
  free(p);
}

/* Side entries of the weak energy conditions: each adds q + h (T_c - T) to
 * the inward heat flux over the side.  QSIDE: h = 0 (src/bc_integ.c:1518-1519);
 * QCONV: q = 0 (src/bc_integ.c:1492-1494, qrad_surf with zero emissivity,
 * src/mm_ns_bc.c:15145-15255: func += h (T_c - T), d func / d T_j = -h phi_j). */
static int add_sides(om_problem *p, int n, const int *elems, const int *sides, double q, double h,
                     double tc) {
  const int nsides = p->dim == 2 ? 4 : 6;
  if (!p->ph.energy || n < 0) return -1;
  for (int k = 0; k < n; k++)
    if (elems[k] < 0 || elems[k] >= p->num_elems || sides[k] < 1 || sides[k] > nsides) return -1;
  int m = p->n_qs + n;
  int *qe = (int *)realloc(p->qs_elem, (size_t)(m > 0 ? m : 1) * sizeof(int));
  if (qe) p->qs_elem = qe;
  int *qd = (int *)realloc(p->qs_side, (size_t)(m > 0 ? m : 1) * sizeof(int));
  if (qd) p->qs_side = qd;
  double *qv = (double *)realloc(p->qs_q, (size_t)(m > 0 ? m : 1) * sizeof(double));
  if (qv) p->qs_q = qv;
  double *qh = (double *)realloc(p->qs_h, (size_t)(m > 0 ? m : 1) * sizeof(double));
  if (qh) p->qs_h = qh;
  double *qt = (double *)realloc(p->qs_tc, (size_t)(m > 0 ? m : 1) * sizeof(double));
  if (qt) p->qs_tc = qt;
  if (!qe || !qd || !qv || !qh || !qt) return -1;
  for (int k = 0; k < n; k++) {
    p->qs_elem[p->n_qs + k] = elems[k];
    p->qs_side[p->n_qs + k] = sides[k];
    p->qs_q[p->n_qs + k] = q;
    p->qs_h[p->n_qs + k] = h;
    p->qs_tc[p->n_qs + k] = tc;
  }
  p->n_qs = m;
  return 0;
}

int om_add_qside(om_problem *p, int n, const int *elems, const int *sides, double q) {
  return add_sides(p, n, elems, sides, q, 0.0, 0.0);
}

int om_add_qconv(om_problem *p, int n, const int *elems, const int *sides, double h, double tc) {
  return add_sides(p, n, elems, sides, 0.0, h, tc);
}

int om_num_unknowns(const om_problem *p) { return p->N; }
int om_ija_size(const om_problem *p) { return p->ija[p->N]; }
void om_copy_ija(const om_problem *p, int *o) { memcpy(o, p->ija, (size_t)p->ija[p->N] * sizeof(int)); }
void om_copy_first_unknown(const om_problem *p, int *o) {
  memcpy(o, p->first_unknown, (p->num_nodes + 1) * sizeof(int));
}
void om_copy_unknown_var(const om_problem *p, int *o) { memcpy(o, p->unk_var, p->N * sizeof(int)); }
void om_copy_dbc(const om_problem *p, char *mask, double *val) {
  memcpy(mask, p->dbc, p->N);
  memcpy(val, p->dbc_val, p->N * sizeof(double));
}

/* ------------------------------------------------------------------------ */
/* One element: matrix_fill for the NS + continuity + energy terms           */
/* ------------------------------------------------------------------------ */

typedef struct {
  double R[NSLOT][MDE];
  double J[NSLOT][NSLOT][MDE][MDE];
  int ndof[NSLOT];
  int gun[NSLOT][MDE]; /* global unknown of each local dof (-1: inactive slot) */
} lec_t;

#define DELTA(a, b) ((a) == (b) ? 1.0 : 0.0)

/* QSIDE and QCONV (src/bc_integ.c:1492-1494, 1518-1519 func = q + h (T_c - T);
 * 2156-2157, 2207: R[T][i] += w phi_i sdet func over the side, J[T][T][i][j]
 * += w phi_i sdet d func / d T_j = -w phi_i sdet h phi_j).  Exodus side
 * numbering, as Goma's get_side_info (src/mm_fill_fill.c:3049-3072,
 * 3182-3250): the reference coordinate fixed on each side (axis 0 = s, 1 = t,
 * 2 = u) and its value. */
static void qside_elem(const om_problem *p, int e, double xc[3][MDE], const double *Tn, lec_t *L) {
  static const int side_axis[6] = {1, 0, 1, 0, 2, 2};
  static const double side_val[6] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0};
  const double gp[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
  const double gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  const int d = p->dim, nve = p->nve;
  for (int k = 0; k < p->n_qs; k++) {
    if (p->qs_elem[k] != e) continue;
    const int sd = p->qs_side[k] - 1, ax = side_axis[sd];
    int fr[2], nf = 0;
    for (int a = 0; a < d; a++)
      if (a != ax) fr[nf++] = a;
    const int n1 = 3, n2 = (d == 3) ? 3 : 1;
    for (int g1 = 0; g1 < n1; g1++)
      for (int g2 = 0; g2 < n2; g2++) {
        double ref[3] = {0, 0, 0}, phi[MDE], dphi[MDE][3];
        ref[ax] = side_val[sd];
        ref[fr[0]] = gp[g1];
        double w = gw[g1];
        if (d == 3) {
          ref[fr[1]] = gp[g2];
          w *= gw[g2];
        }
        if (d == 2) shape_quad9(ref[0], ref[1], phi, dphi);
        else shape_hex27(ref[0], ref[1], ref[2], phi, dphi);
        double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
        for (int j = 0; j < d; j++)
          for (int i = 0; i < nve; i++) {
            t1[j] += xc[j][i] * dphi[i][fr[0]];
            if (d == 3) t2[j] += xc[j][i] * dphi[i][fr[1]];
          }
        double sdet;
        if (d == 2) {
          sdet = sqrt(t1[0] * t1[0] + t1[1] * t1[1]);
        } else {
          double c0 = t1[1] * t2[2] - t1[2] * t2[1], c1 = t1[2] * t2[0] - t1[0] * t2[2],
                 c2 = t1[0] * t2[1] - t1[1] * t2[0];
          sdet = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        }
        const double h = p->qs_h[k];
        double Tg = 0.0;
        for (int j = 0; j < nve; j++) Tg += phi[j] * Tn[j];
        const double func = p->qs_q[k] + h * (p->qs_tc[k] - Tg);
        for (int i = 0; i < nve; i++) {
          L->R[SLOT_T][i] += w * phi[i] * sdet * func;
          if (h != 0.0)
            for (int j = 0; j < nve; j++) L->J[SLOT_T][SLOT_T][i][j] -= w * phi[i] * sdet * h * phi[j];
        }
      }
  }
}

/* h_elem_siz (src/mm_fill_aux.c:844-1125), QUAD and HEX branches: the
 * squared widths between the centroids of opposite faces, from the vertex
 * nodes (local 0-3 / 0-7). */
static void h_elem_siz(const om_problem *p, int e, double hsq[3]) {
  const int d = p->dim, nve = p->nve;
  double xn[3][8];
  const int nvert = d == 2 ? 4 : 8;
  for (int a = 0; a < d; a++)
    for (int i = 0; i < nvert; i++)
      xn[a][i] = p->coor[(size_t)a * p->num_nodes + p->conn[(size_t)e * nve + i]];
  if (d == 2) {
    double hh0[2], hh1[2];
    for (int a = 0; a < 2; a++) {
      const double p1 = 0.5 * (xn[a][0] + xn[a][3]), p2 = 0.5 * (xn[a][1] + xn[a][2]);
      const double p3 = 0.5 * (xn[a][0] + xn[a][1]), p4 = 0.5 * (xn[a][2] + xn[a][3]);
      hh0[a] = p2 - p1;
      hh1[a] = p3 - p4;
    }
    hsq[0] = hh0[0] * hh0[0] + hh0[1] * hh0[1];
    hsq[1] = hh1[0] * hh1[0] + hh1[1] * hh1[1];
    hsq[2] = 0.0;
  } else {
    double hh[3][3];
    for (int a = 0; a < 3; a++) {
      const double f1 = 0.25 * (xn[a][0] + xn[a][1] + xn[a][2] + xn[a][3]);
      const double f2 = 0.25 * (xn[a][1] + xn[a][2] + xn[a][5] + xn[a][6]);
      const double f3 = 0.25 * (xn[a][2] + xn[a][3] + xn[a][6] + xn[a][7]);
      const double f4 = 0.25 * (xn[a][0] + xn[a][1] + xn[a][4] + xn[a][5]);
      const double f5 = 0.25 * (xn[a][0] + xn[a][3] + xn[a][4] + xn[a][7]);
      const double f6 = 0.25 * (xn[a][4] + xn[a][5] + xn[a][6] + xn[a][7]);
      hh[0][a] = f2 - f5;
      hh[1][a] = f3 - f4;
      hh[2][a] = f1 - f6;
    }
    for (int r = 0; r < 3; r++) hsq[r] = hh[r][0] * hh[r][0] + hh[r][1] * hh[r][1] + hh[r][2] * hh[r][2];
  }
}

double om_h_elem_avg(const om_problem *p) {
  /* global_h_elem_siz (src/mm_fill_aux.c:1128-1200), one process */
  double h = 0.0;
  const double w = 1.0 / (double)p->num_elems;
  for (int e = 0; e < p->num_elems; e++) {
    double hsq[3];
    h_elem_siz(p, e, hsq);
    double s = 0.0;
    for (int r = 0; r < p->dim; r++) s += hsq[r];
    h += sqrt(s / p->dim) * w;
  }
  return h;
}

double om_velocity_norm(const om_problem *p, const double *x) {
  /* global_velocity_norm (src/mm_fill_aux.c:612-672): every node carries
   * the velocity components */
  double s = 0.0;
  long long cnt = 0;
  for (int a = 0; a < p->dim; a++)
    for (int n = 0; n < p->num_nodes; n++) {
      const double v = x[p->first_unknown[n] + a];
      s += v * v;
      cnt++;
    }
  return s / (double)cnt;
}

void om_set_velocity_norm(om_problem *p, double U_norm) { p->U_norm = U_norm; }

/* element_viscosity (src/mm_fill_aux.c:687-751): the viscosity model at a
 * shear-rate tensor with every one of its dim x dim entries 1e-5 (VIM =
 * Num_Dim in Cartesian coordinates, src/mm_prob_def.c:96). */
static double element_viscosity(const om_physics *ph, int dim) {
  if (ph->visc_model != 1) return ph->mu0;
  const double gd = sqrt(0.5 * fabs((double)(dim * dim) * 1e-5 * 1e-5));
  const double val2 = pow(ph->lam * gd, ph->aexp);
  return ph->muinf + (ph->mu0 - ph->muinf) * pow(1. + val2, (ph->nexp - 1.) / ph->aexp);
}

double om_pspg_tau(const om_problem *p) {
  /* calc_pspg, pspg_global (src/mm_fill_stabilization.c:1037-1049); the
   * global scheme's tau carries no Jacobian dependence */
  const om_physics *ph = &p->ph;
  const double mu_avg = element_viscosity(ph, p->dim), h = p->h_elem_avg, U = p->U_norm;
  const double Re = ph->rho * U * h / (2.0 * mu_avg);
  if (Re <= 3.0) return ph->ps_scaling * h * h / (12.0 * mu_avg);
  return ph->ps_scaling * h / (2.0 * ph->rho * U);
}

/* Assembles one element into lec.  Returns 0, or -1 on detJ <= 0.
 * Follows matrix_fill's quadrature loop (src/mm_fill.c:1253-2791).  With
 * xdot != NULL the transient mass terms are added: xdot is the time
 * derivative of x and cmass = d xdot / d x = (1 + 2 theta) / dt. */
static int elem_fill(const om_problem *p, int e, const double *x, const double *xdot,
                     double cmass, lec_t *L) {
  const int d = p->dim, nve = p->nve, np = d + 1, en = p->ph.energy ? 1 : 0;
  const om_physics *ph = &p->ph;
  const int nq = (d == 2) ? 9 : 27;
  const double tau_pspg = ph->pspg ? om_pspg_tau(p) : 0.0;
  double xc[3][MDE], vel[3][MDE], Tn[MDE], Pn[4], veld[3][MDE], Tdn[MDE];
  /* energy SUPG (src/mm_fill_energy.c:237-259): h_elem from the element's
   * face-centroid widths and the element velocity of element_velocity for Q2
   * (src/mm_fill_aux.c:798-806: the velocity at local dof dofs - 1, with
   * dv_dnode = 1 there); supg_w = Ewt_func / h_elem, supg_d[b] its derivative
   * in the velocity component b at that dof (src/mm_fill_energy.c:644-651) */
  double supg_w = 0.0, supg_d[3] = {0.0, 0.0, 0.0};
  memset(L->ndof, 0, sizeof(L->ndof));
  for (int a = 0; a < d; a++) L->ndof[a] = nve;
  if (en) L->ndof[SLOT_T] = nve;
  L->ndof[SLOT_P] = np;
  for (int i = 0; i < nve; i++) {
    int n = p->conn[(size_t)e * nve + i];
    int f = p->first_unknown[n];
    for (int j = 0; j < d; j++) xc[j][i] = p->coor[(size_t)j * p->num_nodes + n];
    for (int a = 0; a < d; a++) {
      vel[a][i] = x[f + a];
      L->gun[a][i] = f + a;
    }
    Tn[i] = en ? x[f + d] : 0.0;
    if (en) L->gun[SLOT_T][i] = f + d;
    for (int a = 0; a < d; a++) veld[a][i] = xdot ? xdot[f + a] : 0.0;
    Tdn[i] = (xdot && en) ? xdot[f + d] : 0.0;
  }
  {
    int c = p->elem_centroid[e];
    int f = p->first_unknown[c] + d + en;
    for (int k = 0; k < np; k++) {
      Pn[k] = x[f + k];
      L->gun[SLOT_P][k] = f + k;
    }
  }
  if (en && ph->energy_supg != 0.0) {
    double hsq[3], he = 0.0;
    h_elem_siz(p, e, hsq);
    for (int pp = 0; pp < d; pp++) he += vel[pp][nve - 1] * vel[pp][nve - 1] / hsq[pp];
    he = sqrt(he) / 2.0;
    const double hinv = he == 0.0 ? 0.0 : 1.0 / he;
    supg_w = ph->energy_supg * hinv;
    if (he != 0.0)
      for (int b = 0; b < d; b++) {
        const double dh = hsq[b] != 0.0 ? vel[b][nve - 1] * hinv / 4.0 / hsq[b] : 0.0;
        supg_d[b] = ph->energy_supg * (-dh / he / he);
      }
  }
  /* zero_lec (src/mm_fill.c:5641), restricted to the active blocks */
  for (int A = 0; A < NSLOT; A++) {
    if (!L->ndof[A]) continue;
    memset(L->R[A], 0, sizeof(L->R[A]));
    for (int B = 0; B < NSLOT; B++)
      if (L->ndof[B]) memset(L->J[A][B], 0, sizeof(L->J[A][B]));
  }

  for (int q = 0; q < nq; q++) {
    double s, t, u, wt;
    gauss_point(d, q, &s, &t, &u, &wt);
    double phi[MDE], dphidxi[MDE][3], psi[4];
    if (d == 2) shape_quad9(s, t, phi, dphidxi);
    else shape_hex27(s, t, u, phi, dphidxi);
    /* P1 discontinuous pressure basis {1, s, t[, u]} in reference
     * coordinates (src/mm_fill_util.c:3802-3858, 4205-4276). */
    psi[0] = 1.0;
    psi[1] = s;
    psi[2] = t;
    psi[3] = u;

    /* beer_belly (src/mm_fill_util.c:266-273, 384-392): J[i][j] = sum_k
     * x_j(k) dphi_k/dxi_i, B = J^-1, detJ. */
    double Jm[3][3] = {{0}}, B[3][3] = {{0}}, detJ;
    for (int i = 0; i < d; i++)
      for (int j = 0; j < d; j++)
        for (int k = 0; k < nve; k++) Jm[i][j] += xc[j][k] * dphidxi[k][i];
    if (d == 2) {
      detJ = Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0];
      if (detJ <= 0.0) return -1;
      B[0][0] = Jm[1][1] / detJ;
      B[0][1] = -Jm[0][1] / detJ;
      B[1][0] = -Jm[1][0] / detJ;
      B[1][1] = Jm[0][0] / detJ;
    } else {
      detJ = Jm[0][0] * (Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1]) -
             Jm[0][1] * (Jm[1][0] * Jm[2][2] - Jm[2][0] * Jm[1][2]) +
             Jm[0][2] * (Jm[1][0] * Jm[2][1] - Jm[2][0] * Jm[1][1]);
      if (detJ <= 0.0) return -1;
      B[0][0] = (Jm[1][1] * Jm[2][2] - Jm[2][1] * Jm[1][2]) / detJ;
      B[0][1] = -(Jm[0][1] * Jm[2][2] - Jm[2][1] * Jm[0][2]) / detJ;
      B[0][2] = (Jm[0][1] * Jm[1][2] - Jm[1][1] * Jm[0][2]) / detJ;
      B[1][0] = -(Jm[1][0] * Jm[2][2] - Jm[2][0] * Jm[1][2]) / detJ;
      B[1][1] = (Jm[0][0] * Jm[2][2] - Jm[2][0] * Jm[0][2]) / detJ;
      B[1][2] = -(Jm[0][0] * Jm[1][2] - Jm[1][0] * Jm[0][2]) / detJ;
      B[2][0] = (Jm[1][0] * Jm[2][1] - Jm[2][0] * Jm[1][1]) / detJ;
      B[2][1] = -(Jm[0][0] * Jm[2][1] - Jm[2][0] * Jm[0][1]) / detJ;
      B[2][2] = (Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1]) / detJ;
    }
    /* load_bf_grad (src/mm_fill_util.c:1634-1852): grad_phi[i][p] =
     * sum_q B[p][q] dphidxi[i][q]; grad_phi_e[i][a][p][q] = grad_phi[i][p]
     * delta(a,q) is used implicitly below. */
    double gphi[MDE][3];
    for (int i = 0; i < nve; i++)
      for (int pp = 0; pp < 3; pp++) {
        double acc = 0.0;
        if (pp < d)
          for (int qq = 0; qq < d; qq++) acc += B[pp][qq] * dphidxi[i][qq];
        gphi[i][pp] = acc;
      }

    /* load_fv / load_fv_grads (src/load_field_variables.c:128, 2049):
     * grad_v[p][q] = d v_q / d x_p, grad_T, P, div_v. */
    double v[3] = {0, 0, 0}, grad_v[3][3] = {{0}}, Tq = 0, grad_T[3] = {0, 0, 0}, Pq = 0;
    for (int i = 0; i < nve; i++) {
      for (int a = 0; a < d; a++) {
        v[a] += vel[a][i] * phi[i];
        for (int pp = 0; pp < d; pp++) grad_v[pp][a] += vel[a][i] * gphi[i][pp];
      }
      if (en) {
        Tq += Tn[i] * phi[i];
        for (int pp = 0; pp < d; pp++) grad_T[pp] += Tn[i] * gphi[i][pp];
      }
    }
    for (int k = 0; k < np; k++) Pq += Pn[k] * psi[k];
    double v_dot[3] = {0, 0, 0}, T_dot = 0.0;
    if (xdot)
      for (int i = 0; i < nve; i++) {
        for (int a = 0; a < d; a++) v_dot[a] += veld[a][i] * phi[i];
        T_dot += Tdn[i] * phi[i];
      }
    double div_v = 0.0;
    for (int pp = 0; pp < d; pp++) div_v += grad_v[pp][pp];
    const double d_area = detJ * wt; /* h3 = 1 (Cartesian), src/mm_fill_momentum.c:282 */

    /* fluid_stress (src/mm_fill_momentum.c:3003-3010, 3270, 3466, 3704) */
    double gamma[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) gamma[a][b] = grad_v[a][b] + grad_v[b][a];
    double mu, dmu_v[3][MDE];
    memset(dmu_v, 0, sizeof(dmu_v));
    if (ph->visc_model == 1) {
      /* calc_shearrate (src/mm_fill_util.c:4564-4660) */
      double gd = 0.0;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) gd += gamma[a][b] * gamma[b][a];
      gd = sqrt(0.5 * fabs(gd));
      double d_gd_dv[3][MDE];
      memset(d_gd_dv, 0, sizeof(d_gd_dv));
      if (gd != 0.0) {
        for (int a = 0; a < d; a++)
          for (int i = 0; i < nve; i++) {
            double acc = 0.0;
            for (int pp = 0; pp < 3; pp++)
              for (int qq = 0; qq < 3; qq++) {
                /* grad_phi_e[i][a][p][q] + grad_phi_e[i][a][q][p] */
                double ge = gphi[i][pp] * DELTA(a, qq) + gphi[i][qq] * DELTA(a, pp);
                acc += 0.5 * ge * gamma[qq][pp] / gd;
              }
            d_gd_dv[a][i] = acc;
          }
      }
      /* carreau_viscosity (src/mm_viscosity.c:1065-1116), Include_Visc_Sens */
      const double mu0 = ph->mu0, muinf = ph->muinf, lam = ph->lam, nexp = ph->nexp,
                   aexp = ph->aexp;
      const int nz = fabs(gd) > 1.0e-298;
      double val2 = nz ? pow(lam * gd, aexp) : 0.0;
      double val = pow(1. + val2, (nexp - 1.) / aexp);
      mu = muinf + (mu0 - muinf) * val;
      val = nz ? pow(lam * gd, aexp - 1.) : 0.0;
      double val1 = pow(1. + val2, (nexp - 1. - aexp) / aexp);
      double dmu_gd = (mu0 - muinf) * (nexp - 1.) * lam * val * val1;
      if (nz)
        for (int a = 0; a < d; a++)
          for (int i = 0; i < nve; i++) dmu_v[a][i] = dmu_gd * d_gd_dv[a][i];
    } else {
      mu = ph->mu0;
    }
    double Pi[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) Pi[a][b] = -Pq * DELTA(a, b) + mu * gamma[a][b];

    /* BOUSSINESQ momentum source (src/mm_std_models.c:256-284),
     * f_a = rho g_a (-beta (T - Tref)); df->T[a][j] = -rho g_a beta phi_j */
    double f[3] = {0, 0, 0};
    if (en)
      for (int a = 0; a < d; a++) f[a] = ph->rho * ph->g[a] * (-ph->beta * (Tq - ph->Tref));

    /* ---------------- assemble_momentum (src/mm_fill_momentum.c:474-2117) */
    for (int a = 0; a < d; a++) {
      for (int i = 0; i < nve; i++) {
        double phi_i = phi[i];
        double advection = 0.0;
        for (int pp = 0; pp < d; pp++) advection += v[pp] * grad_v[pp][a];
        advection *= ph->rho;
        advection *= -phi_i * d_area;
        double diffusion = 0.0;
        for (int pp = 0; pp < 3; pp++)
          for (int qq = 0; qq < 3; qq++)
            diffusion += gphi[i][pp] * DELTA(a, qq) * Pi[qq][pp];
        diffusion *= -d_area;
        double source = f[a] * phi_i * d_area;
        L->R[a][i] += advection + diffusion + source;
        /* transient mass (src/mm_fill_momentum.c:534-539, J 1586-1590):
         * R -= rho v_dot_a phi_i dA;  J_m_v[a][a] -= rho c phi_j phi_i dA */
        if (xdot) {
          L->R[a][i] += -ph->rho * v_dot[a] * phi_i * d_area;
          for (int j = 0; j < nve; j++) L->J[a][a][i][j] += -ph->rho * cmass * phi[j] * phi_i * d_area;
        }

        /* J_m_T (src/mm_fill_momentum.c:746-915): Boussinesq source */
        if (en)
          for (int j = 0; j < nve; j++) {
            double dfT = -ph->rho * ph->g[a] * ph->beta * phi[j];
            L->J[a][SLOT_T][i][j] += phi_i * dfT * d_area;
          }
        /* J_m_v (src/mm_fill_momentum.c:1564-1735) */
        for (int b = 0; b < d; b++) {
          for (int j = 0; j < nve; j++) {
            double phi_j = phi[j];
            double adv = phi_j * grad_v[b][a];
            for (int pp = 0; pp < d; pp++) adv += v[pp] * gphi[j][pp] * DELTA(b, a);
            adv *= ph->rho * -phi_i * d_area;
            double dif = 0.0;
            for (int pp = 0; pp < 3; pp++) {
              for (int qq = 0; qq < 3; qq++) {
                /* d_Pi->v[q][p][b][j] = mu (grad_phi_e[j][b][q][p] +
                 * grad_phi_e[j][b][p][q]) + d_mu->v[b][j] gamma[q][p] */
                double dPi = mu * (gphi[j][qq] * DELTA(b, pp) + gphi[j][pp] * DELTA(b, qq)) +
                             dmu_v[b][j] * gamma[qq][pp];
                dif += gphi[i][pp] * DELTA(a, qq) * dPi;
              }
            }
            dif *= -d_area;
            L->J[a][b][i][j] += adv + dif;
          }
        }
        /* J_m_P (src/mm_fill_momentum.c:2052-2117): d_Pi->P[q][p][k] =
         * -delta(p,q) psi_k */
        for (int k = 0; k < np; k++) {
          double dif = 0.0;
          for (int pp = 0; pp < 3; pp++)
            for (int qq = 0; qq < 3; qq++)
              dif -= gphi[i][pp] * DELTA(a, qq) * (-DELTA(pp, qq) * psi[k]);
          dif *= d_area;
          L->J[a][SLOT_P][i][k] += dif;
        }
      }
    }

    /* ---------------- assemble_continuity (src/mm_fill_continuity.c:435-700) */
    for (int k = 0; k < np; k++) {
      L->R[SLOT_P][k] += div_v * psi[k] * d_area;
      for (int b = 0; b < d; b++)
        for (int j = 0; j < nve; j++) {
          /* div(phi_j e_b) = grad_phi_e[j][b][p][p] summed */
          double div_phi_j_e_b = gphi[j][b];
          L->J[SLOT_P][b][k][j] += psi[k] * div_phi_j_e_b * d_area;
        }
    }
    /* PSPG, global tau (src/mm_fill_continuity.c:305-307, 600-612, 742-749,
     * 792-809, 890-905; calc_pspg src/mm_fill_stabilization.c:1100-1465 with
     * no polymer stress, so div_s = div_G = 0): R_c[k] += grad psi_k . pspg dA,
     * pspg_a = tau (rho v_dot_a + rho v . grad v_a + dP/dx_a - f_a).  The P1
     * gradients: grad psi_0 = 0, grad psi_k[a] = B[a][k-1]. */
    if (ph->pspg) {
      double gpsi[4][3], gP[3], mom[3];
      for (int a = 0; a < 3; a++) gpsi[0][a] = 0.0;
      for (int k = 1; k < np; k++)
        for (int a = 0; a < 3; a++) gpsi[k][a] = a < d ? B[a][k - 1] : 0.0;
      for (int a = 0; a < d; a++) {
        gP[a] = 0.0;
        for (int k = 0; k < np; k++) gP[a] += Pn[k] * gpsi[k][a];
        double adv = 0.0;
        for (int pp = 0; pp < d; pp++) adv += v[pp] * grad_v[pp][a];
        mom[a] = ph->rho * adv + gP[a] - f[a] + (xdot ? ph->rho * v_dot[a] : 0.0);
      }
      const double w = tau_pspg * d_area;
      for (int k = 0; k < np; k++) {
        double r = 0.0;
        for (int a = 0; a < d; a++) r += gpsi[k][a] * mom[a];
        L->R[SLOT_P][k] += w * r;
        /* J_c_v: d pspg_a / d v_b,j = tau rho (phi_j dv_a/dx_b + delta_ab
         * (v . grad phi_j + c phi_j)), c = (1 + 2 theta) / dt (transient) */
        for (int b = 0; b < d; b++)
          for (int j = 0; j < nve; j++) {
            double vg = 0.0;
            for (int pp = 0; pp < d; pp++) vg += v[pp] * gphi[j][pp];
            double acc = 0.0;
            for (int a = 0; a < d; a++) acc += gpsi[k][a] * phi[j] * grad_v[b][a];
            acc += gpsi[k][b] * (vg + (xdot ? cmass * phi[j] : 0.0));
            L->J[SLOT_P][b][k][j] += w * ph->rho * acc;
          }
        /* J_c_T: d pspg_a / d T_j = -tau df_a/dT_j = tau rho g_a beta phi_j */
        if (en)
          for (int j = 0; j < nve; j++) {
            double acc = 0.0;
            for (int a = 0; a < d; a++) acc += gpsi[k][a] * ph->rho * ph->g[a] * ph->beta;
            L->J[SLOT_P][SLOT_T][k][j] += w * acc * phi[j];
          }
        /* J_c_P: d pspg_a / d P_l = tau grad psi_l[a] */
        for (int l = 0; l < np; l++) {
          double acc = 0.0;
          for (int a = 0; a < d; a++) acc += gpsi[k][a] * gpsi[l][a];
          L->J[SLOT_P][SLOT_P][k][l] += w * acc;
        }
      }
    }

    /* ---------------- assemble_energy (src/mm_fill_energy.c:280-700) */
    if (en) {
      const double k = ph->kcond, rCp = ph->rho * ph->Cp;
      double qf[3]; /* heat_flux, Fourier (src/mm_fill_energy.c:1652) */
      for (int pp = 0; pp < 3; pp++) qf[pp] = -k * grad_T[pp];
      double vgT = 0.0; /* v . grad T */
      for (int pp = 0; pp < d; pp++) vgT += v[pp] * grad_T[pp];
      for (int i = 0; i < nve; i++) {
        double phi_i = phi[i];
        /* weight of the advection term: phi_i + supg / h_elem (v . grad phi_i)
         * (src/mm_fill_energy.c:332-339, 403-411) */
        double vgi = 0.0;
        for (int pp = 0; pp < d; pp++) vgi += v[pp] * gphi[i][pp];
        const double wt_func = phi_i + supg_w * vgi;
        double advection = vgT;
        advection *= -wt_func * rCp * detJ * wt;
        double diffusion = 0.0;
        for (int pp = 0; pp < d; pp++) diffusion += gphi[i][pp] * qf[pp];
        diffusion *= detJ * wt;
        double source = phi_i * ph->hsrc * detJ * wt;
        L->R[SLOT_T][i] += advection + diffusion + source;
        /* transient mass (src/mm_fill_energy.c:322-328, J 435-441), constant
         * rho and Cp: R -= rho Cp T_dot phi_i dA;  J_e_T -= rho Cp c phi_j phi_i dA */
        if (xdot) {
          L->R[SLOT_T][i] += -rCp * T_dot * phi_i * detJ * wt;
          for (int j = 0; j < nve; j++)
            L->J[SLOT_T][SLOT_T][i][j] += -rCp * cmass * phi[j] * phi_i * detJ * wt;
        }
        /* J_e_T (src/mm_fill_energy.c:425-487) */
        for (int j = 0; j < nve; j++) {
          double adv = 0.0, dif = 0.0;
          for (int pp = 0; pp < d; pp++) adv += rCp * v[pp] * gphi[j][pp];
          adv *= -wt_func * detJ * wt;
          for (int pp = 0; pp < d; pp++) dif += (-k * gphi[j][pp]) * gphi[i][pp];
          dif *= detJ * wt;
          L->J[SLOT_T][SLOT_T][i][j] += adv + dif;
        }
        /* J_e_v (src/mm_fill_energy.c:620-690): d_vconv->v[b][b][j] = phi_j;
         * with SUPG the weight's own derivative: d_wt_func = supg/h phi_j
         * d_b phi_i + supg d(1/h)/dv_b,j v_b d_b phi_i (the second only at
         * the element-velocity dof) times the advection rho Cp v . grad T */
        for (int b = 0; b < d; b++)
          for (int j = 0; j < nve; j++) {
            double adv = wt_func * rCp * phi[j] * grad_T[b];
            adv *= -detJ * wt;
            if (supg_w != 0.0 || supg_d[b] != 0.0) {
              double dwt = supg_w * phi[j] * gphi[i][b];
              if (j == nve - 1) dwt += supg_d[b] * v[b] * gphi[i][b];
              adv += -dwt * rCp * vgT * detJ * wt;
            }
            L->J[SLOT_T][b][i][j] += adv;
          }
      }
    }
  }

  if (p->n_qs && en) qside_elem(p, e, xc, Tn, L);

  /* put_dirichlet_in_matrix (src/bc_dirich.c:44-153): zero the local row,
   * J_ii = DIRICHLET_PENALTY (1, include/rf_bc_const.h:152), R = x - V. */
  for (int A = 0; A < NSLOT; A++) {
    for (int i = 0; i < L->ndof[A]; i++) {
      int g = L->gun[A][i];
      if (!p->dbc[g]) continue;
      for (int B = 0; B < NSLOT; B++)
        for (int j = 0; j < L->ndof[B]; j++) L->J[A][B][i][j] = 0.0;
      L->J[A][A][i][i] = 1.0;
      L->R[A][i] = 1.0 * (x[g] - p->dbc_val[g]);
    }
  }
  return 0;
}

static int slot_var(int A) { return A == SLOT_P ? OM_P : (A == SLOT_T ? OM_T : OM_U1 + A); }

/* in_list (include/el_elm_info.h:110-119): linear scan */
static int in_list(int value, int start, int end, const int *list) {
  for (int i = start; i < end; i++)
    if (list[i] == value) return i;
  return -1;
}

/* load_lec, MSR branch (src/mm_fill.c:5241-5483) */
static int load_lec(const om_problem *p, const lec_t *L, double *resid, double *a) {
  const int *ija = p->ija;
  for (int A = 0; A < NSLOT; A++) {
    for (int i = 0; i < L->ndof[A]; i++) {
      int ie = L->gun[A][i];
      resid[ie] += L->R[A][i];
      for (int B = 0; B < NSLOT; B++) {
        if (!L->ndof[B] || !inter_mask(slot_var(A), slot_var(B))) continue;
        for (int j = 0; j < L->ndof[B]; j++) {
          int je = L->gun[B][j];
          int ja = (ie == je) ? ie : in_list(je, ija[ie], ija[ie + 1], ija);
          if (ja < 0) return -1;
          a[ja] += L->J[A][B][i][j];
        }
      }
    }
  }
  return 0;
}

static int fill_range(const om_problem *p, const double *x, const double *xdot, double cmass,
                      double *resid, double *val, int e0, int e1) {
  lec_t *L = (lec_t *)malloc(sizeof(lec_t));
  int rc = 0;
  for (int e = e0; e < e1; e++) {
    if (elem_fill(p, e, x, xdot, cmass, L) < 0) {
      rc = -1; /* negative / zero element volume: matrix_fill_full returns -1 */
      break;
    }
    if (load_lec(p, L, resid, val) < 0) {
      rc = -1;
      break;
    }
  }
  free(L);
  return rc;
}

int om_fill_range(const om_problem *p, const double *x, double *resid, double *val, int e0,
                  int e1) {
  return fill_range(p, x, NULL, 0.0, resid, val, e0, e1);
}

int om_fill(const om_problem *p, const double *x, double *resid, double *val) {
  return om_fill_range(p, x, resid, val, 0, p->num_elems);
}

int om_fill_transient(const om_problem *p, const double *x, const double *xdot, double cmass,
                      double *resid, double *val) {
  return fill_range(p, x, xdot, cmass, resid, val, 0, p->num_elems);
}

int om_element(const om_problem *p, int e, const double *x, double *R, double *J, int *nloc) {
  lec_t *L = (lec_t *)malloc(sizeof(lec_t));
  int rc = elem_fill(p, e, x, NULL, 0.0, L);
  int off[NSLOT], n = 0;
  for (int A = 0; A < NSLOT; A++) {
    off[A] = n;
    n += L->ndof[A];
  }
  *nloc = n;
  if (rc == 0) {
    for (int A = 0; A < NSLOT; A++)
      for (int i = 0; i < L->ndof[A]; i++) {
        R[off[A] + i] = L->R[A][i];
        for (int B = 0; B < NSLOT; B++)
          for (int j = 0; j < L->ndof[B]; j++) J[(size_t)(off[A] + i) * n + off[B] + j] = L->J[A][B][i][j];
      }
  }
  free(L);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Scaling, norms, MSR matvec                                                */
/* ------------------------------------------------------------------------ */

/* row_sum_scale_MSR (src/sl_matrix_util.c:506-600) */
void om_row_sum_scale(int N, double *a, const int *ija, double *b, double *scale) {
  for (int irow = 0; irow < N; irow++) {
    int j_last = ija[irow + 1] - ija[irow], ija_row = ija[irow];
    double row_sum = fabs(a[irow]);
    for (int j = 0; j < j_last; j++) row_sum += fabs(a[ija_row + j]);
    if (fabs(a[irow]) > 1.0E-200) row_sum = row_sum * (a[irow] > 0 ? 1.0 : -1.0);
    scale[irow] = row_sum;
    b[irow] = b[irow] / row_sum;
    a[irow] = a[irow] / row_sum;
    for (int j = 0; j < j_last; j++) a[ija_row + j] /= row_sum;
  }
}

/* L2_norm / L1_norm / Loo_norm / L2_norm_r / Loo_norm_r
 * (src/mm_sol_nonlinear.c:3177-3420) */
double om_norm_L2(const double *v, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += v[i] * v[i];
  return sqrt(s);
}
double om_norm_L1(const double *v, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += fabs(v[i]);
  return s;
}
double om_norm_Loo(const double *v, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++)
    if (fabs(v[i]) > s) s = fabs(v[i]);
  return s;
}
double om_norm_L2_r(const double *d, const double *x, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += d[i] * d[i] / (1 + x[i] * x[i]);
  return sqrt(s);
}
double om_norm_Loo_r(const double *d, const double *x, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) {
    double q = fabs(d[i]) / (1 + fabs(x[i]));
    if (q > s) s = q;
  }
  return s;
}

/* MV_MSR (src/sl_auxutil.c:373) */
void om_msr_matvec(int N, const double *val, const int *ija, const double *x, double *y) {
  for (int i = 0; i < N; i++) {
    double s = val[i] * x[i];
    for (int k = ija[i]; k < ija[i + 1]; k++) s += val[k] * x[ija[k]];
    y[i] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* Node colouring for the parallel ILU(0) ordering                           */
/* ------------------------------------------------------------------------ */

int om_node_colours(const om_problem *p, int *col) {
  int nn = p->num_nodes, maxc = 0;
  int *mark = (int *)malloc(1024 * sizeof(int));
  for (int c = 0; c < 1024; c++) mark[c] = -1;
  for (int n = 0; n < nn; n++) col[n] = -1;
  for (int n = 0; n < nn; n++) {
    if (p->is_centroid[n] >= 0) continue;
    for (int k = p->nn_ptr[n]; k < p->nn_ptr[n + 1]; k++) {
      int J = p->nn_list[k];
      if (col[J] >= 0) mark[col[J]] = n;
    }
    int c = 0;
    while (mark[c] == n) c++;
    col[n] = c;
    if (c + 1 > maxc) maxc = c + 1;
  }
  for (int n = 0; n < nn; n++)
    if (p->is_centroid[n] >= 0) col[n] = maxc;
  free(mark);
  return maxc + 1;
}

void om_colour_permutation(const om_problem *p, int *perm, int *cptr, int *ncol) {
  int *col = (int *)malloc(p->num_nodes * sizeof(int));
  int nc = om_node_colours(p, col);
  int k = 0;
  for (int c = 0; c < nc; c++) {
    if (cptr) cptr[c] = k;
    for (int n = 0; n < p->num_nodes; n++)
      if (col[n] == c)
        for (int u = p->first_unknown[n]; u < p->first_unknown[n + 1]; u++) perm[k++] = u;
  }
  if (cptr) cptr[nc] = k;
  *ncol = nc;
  free(col);
}


/* Patch ordering for the GPU ILU(0): nodes are binned into patches by their
 * coordinates (nbins boxes per axis over the mesh bounding box), patches are
 * coloured greedily in patch order so that no two patches of one colour hold
 * coupled nodes, and unknowns are ordered colour-major, then patch, then
 * node, then unknown at the node.  Fills perm[new] = old, pcol_ptr (unknown
 * offset of each colour), returns the number of colours; *npatch_out and
 * patch_ptr (unknown offset of each patch in the new order) when non-NULL. */
int om_patch_permutation(const om_problem *p, int nbins, int *perm, int *cptr, int *npatch_out,
                         int *patch_ptr) {
  int nn = p->num_nodes, d = p->dim;
  double lo[3], hi[3];
  for (int k = 0; k < d; k++) {
    lo[k] = 1e300;
    hi[k] = -1e300;
    for (int n = 0; n < nn; n++) {
      double v = p->coor[(size_t)k * nn + n];
      if (v < lo[k]) lo[k] = v;
      if (v > hi[k]) hi[k] = v;
    }
  }
  int *patch = (int *)malloc(nn * sizeof(int));
  for (int n = 0; n < nn; n++) {
    int id = 0, mul = 1;
    for (int k = 0; k < d; k++) {
      double w = hi[k] - lo[k];
      int b = w > 0 ? (int)floor((p->coor[(size_t)k * nn + n] - lo[k]) * nbins / w) : 0;
      if (b < 0) b = 0;
      if (b > nbins - 1) b = nbins - 1;
      id += b * mul;
      mul *= nbins;
    }
    patch[n] = id;
  }
  int np = 1;
  for (int k = 0; k < d; k++) np *= nbins;
  /* patch adjacency via node-node graph, greedy first-fit colouring */
  int *pc = (int *)malloc(np * sizeof(int));
  for (int i = 0; i < np; i++) pc[i] = -1;
  int *pn_cnt = (int *)calloc(np + 1, sizeof(int));
  for (int n = 0; n < nn; n++) pn_cnt[patch[n] + 1]++;
  for (int i = 0; i < np; i++) pn_cnt[i + 1] += pn_cnt[i];
  int *pn = (int *)malloc(nn * sizeof(int));
  int *fp = (int *)malloc(np * sizeof(int));
  memcpy(fp, pn_cnt, np * sizeof(int));
  for (int n = 0; n < nn; n++) pn[fp[patch[n]]++] = n; /* nodes ascending within patch */
  free(fp);
  int *mark = (int *)malloc(4096 * sizeof(int));
  for (int c = 0; c < 4096; c++) mark[c] = -1;
  int ncol = 0;
  for (int P = 0; P < np; P++) {
    for (int k = pn_cnt[P]; k < pn_cnt[P + 1]; k++) {
      int n = pn[k];
      for (int q = p->nn_ptr[n]; q < p->nn_ptr[n + 1]; q++) {
        int Q = patch[p->nn_list[q]];
        if (Q != P && pc[Q] >= 0) mark[pc[Q]] = P;
      }
    }
    int c = 0;
    while (mark[c] == P) c++;
    pc[P] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  int k = 0, npatch = 0;
  for (int c = 0; c < ncol; c++) {
    cptr[c] = k;
    for (int P = 0; P < np; P++) {
      if (pc[P] != c || pn_cnt[P] == pn_cnt[P + 1]) continue;
      if (patch_ptr) patch_ptr[npatch] = k;
      npatch++;
      for (int q = pn_cnt[P]; q < pn_cnt[P + 1]; q++) {
        int n = pn[q];
        for (int u = p->first_unknown[n]; u < p->first_unknown[n + 1]; u++) perm[k++] = u;
      }
    }
  }
  cptr[ncol] = k;
  if (patch_ptr) patch_ptr[npatch] = k;
  if (npatch_out) *npatch_out = npatch;
  free(mark);
  free(pn);
  free(pn_cnt);
  free(pc);
  free(patch);
  return ncol;
}

/* ------------------------------------------------------------------------ */
/* ILU(0) and GMRES                                                          */
/* ------------------------------------------------------------------------ */

/* MSR -> CSR (diagonal merged, columns sorted) in permuted numbering. */
static int msr_to_csr_perm(int N, const double *val, const int *ija, const int *perm, int *rowptr,
                           int *col, double *a) {
  int *iperm = NULL;
  if (perm) {
    iperm = (int *)malloc(N * sizeof(int));
    for (int i = 0; i < N; i++) iperm[perm[i]] = i;
  }
  rowptr[0] = 0;
  int nz = 0;
  for (int inew = 0; inew < N; inew++) {
    int i = perm ? perm[inew] : inew;
    int start = nz;
    col[nz] = perm ? iperm[i] : i;
    a[nz++] = val[i];
    for (int k = ija[i]; k < ija[i + 1]; k++) {
      col[nz] = perm ? iperm[ija[k]] : ija[k];
      a[nz++] = val[k];
    }
    /* insertion sort of the row by column */
    for (int x = start + 1; x < nz; x++) {
      int c = col[x];
      double v = a[x];
      int y = x - 1;
      while (y >= start && col[y] > c) {
        col[y + 1] = col[y];
        a[y + 1] = a[y];
        y--;
      }
      col[y + 1] = c;
      a[y + 1] = v;
    }
    rowptr[inew + 1] = nz;
  }
  free(iperm);
  return nz;
}

/* ILU(0), IKJ variant, in place on a CSR with sorted columns. */
static int ilu0_inplace(int N, const int *rowptr, const int *col, double *a, int *diag) {
  int *pos = (int *)malloc(N * sizeof(int));
  for (int i = 0; i < N; i++) pos[i] = -1;
  for (int i = 0; i < N; i++) {
    diag[i] = -1;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
      if (col[k] == i) diag[i] = k;
    }
  }
  int rc = 0;
  for (int i = 0; i < N && rc == 0; i++) {
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) pos[col[k]] = k;
    for (int kk = rowptr[i]; kk < rowptr[i + 1] && col[kk] < i; kk++) {
      int k = col[kk];
      double piv = a[diag[k]];
      if (piv == 0.0) {
        rc = -1;
        break;
      }
      a[kk] /= piv;
      double lik = a[kk];
      for (int jj = diag[k] + 1; jj < rowptr[k + 1]; jj++) {
        int pj = pos[col[jj]];
        if (pj >= 0) a[pj] -= lik * a[jj];
      }
    }
    if (a[diag[i]] == 0.0) rc = -1;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) pos[col[k]] = -1;
  }
  free(pos);
  return rc;
}

int om_ilu0_csr(int N, const double *val, const int *ija, const int *perm, int *rowptr, int *col,
                double *lu) {
  int nz = msr_to_csr_perm(N, val, ija, perm, rowptr, col, lu);
  int *diag = (int *)malloc(N * sizeof(int));
  int rc = ilu0_inplace(N, rowptr, col, lu, diag);
  free(diag);
  return rc < 0 ? rc : nz;
}

typedef struct {
  int N, use;
  int *rowptr, *col, *diag, *perm;
  double *lu, *tmp;
} ilu_t;

static void ilu_apply(const ilu_t *M, const double *r, double *z) {
  int N = M->N;
  if (!M->use) {
    memcpy(z, r, N * sizeof(double));
    return;
  }
  double *y = M->tmp;
  for (int i = 0; i < N; i++) y[i] = M->perm ? r[M->perm[i]] : r[i];
  for (int i = 0; i < N; i++) {
    double s = y[i];
    for (int k = M->rowptr[i]; k < M->diag[i]; k++) s -= M->lu[k] * y[M->col[k]];
    y[i] = s;
  }
  for (int i = N - 1; i >= 0; i--) {
    double s = y[i];
    for (int k = M->diag[i] + 1; k < M->rowptr[i + 1]; k++) s -= M->lu[k] * y[M->col[k]];
    y[i] = s / M->lu[M->diag[i]];
  }
  for (int i = 0; i < N; i++) {
    if (M->perm) z[M->perm[i]] = y[i];
    else z[i] = y[i];
  }
}

/* Restarted GMRES(m), right preconditioned, classical Gram-Schmidt applied
 * twice, convergence ||r||/||r0|| < tol (Goma's defaults: AZ_gmres,
 * AZ_classic, AZ_r0; src/sl_util.c:670, 810-821; src/mm_input.c:5905-5920). */
static int gmres_core(int N, const double *val, const int *ija, const ilu_t *M, const double *b,
                      double *x, int m, int maxits, double tol, double *relres) {
  double *V = (double *)malloc((size_t)(m + 1) * N * sizeof(double));
  double *Z = (double *)malloc((size_t)N * sizeof(double));
  double *w = (double *)malloc((size_t)N * sizeof(double));
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *h2 = (double *)malloc((m + 1) * sizeof(double));
  double *cs = (double *)malloc(m * sizeof(double)), *sn = (double *)malloc(m * sizeof(double));
  double *g = (double *)malloc((m + 1) * sizeof(double)), *y = (double *)malloc(m * sizeof(double));
  int its = 0;
  om_msr_matvec(N, val, ija, x, w);
  for (int i = 0; i < N; i++) w[i] = b[i] - w[i];
  double r0 = om_norm_L2(w, N), rn = r0;
  if (r0 == 0.0) {
    *relres = 0.0;
    goto done;
  }
  while (its < maxits) {
    double beta = om_norm_L2(w, N);
    for (int i = 0; i < N; i++) V[i] = w[i] / beta;
    memset(g, 0, (m + 1) * sizeof(double));
    g[0] = beta;
    int j;
    for (j = 0; j < m && its < maxits; j++) {
      double *vj = V + (size_t)j * N, *vn = V + (size_t)(j + 1) * N;
      ilu_apply(M, vj, Z);
      om_msr_matvec(N, val, ija, Z, vn);
      double *hj = H + (size_t)j * (m + 1); /* column j */
      for (int k = 0; k <= j; k++) hj[k] = 0.0;
      for (int pass = 0; pass < 2; pass++) {
        for (int k = 0; k <= j; k++) {
          double s = 0.0;
          const double *vk = V + (size_t)k * N;
          for (int i = 0; i < N; i++) s += vk[i] * vn[i];
          h2[k] = s;
        }
        for (int k = 0; k <= j; k++) {
          const double *vk = V + (size_t)k * N;
          for (int i = 0; i < N; i++) vn[i] -= h2[k] * vk[i];
          hj[k] += h2[k];
        }
      }
      double hn = om_norm_L2(vn, N);
      hj[j + 1] = hn;
      if (hn != 0.0)
        for (int i = 0; i < N; i++) vn[i] /= hn;
      for (int k = 0; k < j; k++) {
        double t0 = cs[k] * hj[k] + sn[k] * hj[k + 1];
        hj[k + 1] = -sn[k] * hj[k] + cs[k] * hj[k + 1];
        hj[k] = t0;
      }
      double den = sqrt(hj[j] * hj[j] + hj[j + 1] * hj[j + 1]);
      cs[j] = den == 0.0 ? 1.0 : hj[j] / den;
      sn[j] = den == 0.0 ? 0.0 : hj[j + 1] / den;
      hj[j] = den;
      hj[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      its++;
      if (fabs(g[j + 1]) / r0 < tol || hn == 0.0) {
        j++;
        break;
      }
    }
    /* x += M^-1 V y */
    for (int k = j - 1; k >= 0; k--) {
      double s = g[k];
      for (int l = k + 1; l < j; l++) s -= H[(size_t)l * (m + 1) + k] * y[l];
      y[k] = s / H[(size_t)k * (m + 1) + k];
    }
    memset(w, 0, N * sizeof(double));
    for (int k = 0; k < j; k++) {
      const double *vk = V + (size_t)k * N;
      for (int i = 0; i < N; i++) w[i] += y[k] * vk[i];
    }
    ilu_apply(M, w, Z);
    for (int i = 0; i < N; i++) x[i] += Z[i];
    om_msr_matvec(N, val, ija, x, w);
    for (int i = 0; i < N; i++) w[i] = b[i] - w[i];
    rn = om_norm_L2(w, N);
    if (rn / r0 < tol) break;
  }
  *relres = rn / r0;
done:
  free(V);
  free(Z);
  free(w);
  free(H);
  free(h2);
  free(cs);
  free(sn);
  free(g);
  free(y);
  return its;
}

static int ilu_setup(ilu_t *M, int N, const double *val, const int *ija, const om_problem *p,
                     int precond) {
  memset(M, 0, sizeof(*M));
  M->N = N;
  M->use = precond != 0;
  if (!M->use) return 0;
  size_t nnz = (size_t)ija[N] - 1; /* diag + off-diag */
  M->rowptr = (int *)malloc((N + 1) * sizeof(int));
  M->col = (int *)malloc(nnz * sizeof(int));
  M->lu = (double *)malloc(nnz * sizeof(double));
  M->diag = (int *)malloc(N * sizeof(int));
  M->tmp = (double *)malloc(N * sizeof(double));
  if ((precond == 2 || precond >= 3) && p) {
    int nc;
    M->perm = (int *)malloc(N * sizeof(int));
    int *cptr = (int *)malloc((N + 2) * sizeof(int));
    if (precond == 2) om_colour_permutation(p, M->perm, cptr, &nc);
    else om_patch_permutation(p, precond - 2, M->perm, cptr, NULL, NULL);
    free(cptr);
  }
  msr_to_csr_perm(N, val, ija, M->perm, M->rowptr, M->col, M->lu);
  return ilu0_inplace(N, M->rowptr, M->col, M->lu, M->diag);
}

static void ilu_free(ilu_t *M) {
  free(M->rowptr);
  free(M->col);
  free(M->lu);
  free(M->diag);
  free(M->tmp);
  free(M->perm);
}

int om_gmres_msr(int N, const double *val, const int *ija, const double *b, double *x,
                 const om_problem *p, int precond, int kspace, int max_its, double tol,
                 double *rel_resid) {
  ilu_t M;
  if (ilu_setup(&M, N, val, ija, p, precond) < 0) {
    ilu_free(&M);
    return -1;
  }
  int its = gmres_core(N, val, ija, &M, b, x, kspace, max_its, tol, rel_resid);
  ilu_free(&M);
  return its;
}


/* GMRES with ILU(0) of P A P^T for an explicit permutation perm[new] = old;
 * with part != NULL, couplings between unknowns of different parts are
 * dropped from the preconditioner (block Jacobi over parts). */
int om_gmres_perm(int N, const double *val, const int *ija, const double *b, double *x,
                  const int *perm, const int *part, int kspace, int max_its, double tol,
                  double *rel_resid) {
  ilu_t M;
  memset(&M, 0, sizeof(M));
  M.N = N;
  M.use = 1;
  size_t nnz = (size_t)ija[N] - 1;
  M.rowptr = (int *)malloc((N + 1) * sizeof(int));
  M.col = (int *)malloc(nnz * sizeof(int));
  M.lu = (double *)malloc(nnz * sizeof(double));
  M.diag = (int *)malloc(N * sizeof(int));
  M.tmp = (double *)malloc(N * sizeof(double));
  if (perm) {
    M.perm = (int *)malloc(N * sizeof(int));
    memcpy(M.perm, perm, N * sizeof(int));
  }
  double *v2 = (double *)malloc((size_t)ija[N] * sizeof(double));
  memcpy(v2, val, (size_t)ija[N] * sizeof(double));
  if (part)
    for (int i = 0; i < N; i++)
      for (int k = ija[i]; k < ija[i + 1]; k++)
        if (part[i] != part[ija[k]]) v2[k] = 0.0;
  msr_to_csr_perm(N, v2, ija, M.perm, M.rowptr, M.col, M.lu);
  free(v2);
  if (ilu0_inplace(N, M.rowptr, M.col, M.lu, M.diag) < 0) {
    ilu_free(&M);
    return -1;
  }
  int its = gmres_core(N, val, ija, &M, b, x, kspace, max_its, tol, rel_resid);
  ilu_free(&M);
  return its;
}

/* ------------------------------------------------------------------------ */
/* Newton loop                                                               */
/* ------------------------------------------------------------------------ */

/* perform_backtrack_line_search (src/mm_sol_nonlinear.c:344-590). */
static double backtrack(const om_problem *p, double *x, const double *dx, const double *val,
                        const double *scale, double L2, double eps0, double min_damp) {
  const int N = p->N;
  const int nv = p->ija[N];
  double *xs = (double *)malloc(N * sizeof(double)), *R = (double *)malloc(N * sizeof(double));
  double *w = (double *)malloc(N * sizeof(double)), *vd = (double *)malloc((size_t)nv * sizeof(double));
  memcpy(xs, x, N * sizeof(double));
  double damp = 1.0;
  for (int j = 0; j < N; j++) x[j] = xs[j] - damp * dx[j];
  memset(R, 0, N * sizeof(double));
  om_fill(p, x, R, vd); /* vd receives a Jacobian this restatement ignores */
  for (int j = 0; j < N; j++) R[j] = R[j] / scale[j];
  double g = om_norm_L2(R, N);
  double best_damp = damp, best_norm = g;
  double r_check = 0.5 * (L2 * L2);
  om_msr_matvec(N, val, p->ija, dx, w);
  double slope = 0.0;
  for (int j = 0; j < N; j++) slope += w[j] * R[j];
  if (slope > 0) slope = -slope;
  else if (slope == 0) slope = -1;
  int skip = best_norm < eps0;
  while (!skip) {
    damp *= 0.5;
    if (damp < min_damp) break;
    for (int j = 0; j < N; j++) x[j] = xs[j] - damp * dx[j];
    memset(R, 0, N * sizeof(double));
    memset(vd, 0, (size_t)nv * sizeof(double));
    om_fill(p, x, R, vd);
    for (int j = 0; j < N; j++) R[j] = R[j] / scale[j];
    g = om_norm_L2(R, N);
    if (isnan(g)) break;
    if (g < best_norm) {
      best_damp = damp;
      best_norm = g;
    }
    if (best_norm < eps0) break;
    g = 0.5 * (g * g);
    if (g <= r_check + 0.5 * slope * damp) break;
  }
  for (int j = 0; j < N; j++) x[j] = xs[j] - best_damp * dx[j];
  free(xs);
  free(R);
  free(w);
  free(vd);
  return best_damp;
}

/* solve_nonlinear_problem, steady, NLS_FULL_STEP or NLS_BACKTRACK
 * (src/mm_sol_nonlinear.c:1108-2913). */
int om_newton(om_problem *p, double *x, const om_solver_opts *o, om_report *rep) {
  const int N = p->N;
  const int nv = p->ija[N];
  double *resid = (double *)malloc(N * sizeof(double));
  double *dx = (double *)malloc(N * sizeof(double));
  double *scale = (double *)malloc(N * sizeof(double));
  double *val = (double *)malloc((size_t)nv * sizeof(double));
  memset(rep, 0, sizeof(*rep));
  double t0 = now_s();
  int converged = 0, inewton = 0, ret = 0;
  double initial_norm = 0.0;
  while (!converged && inewton < o->max_newton) {
    memset(resid, 0, N * sizeof(double));
    memset(dx, 0, N * sizeof(double));
    memset(val, 0, (size_t)nv * sizeof(double));
    /* the PSPG velocity norm from this pass's x (src/mm_sol_nonlinear.c:
     * 1183-1192); the line search's residual fills keep it */
    if (p->ph.pspg) p->U_norm = om_velocity_norm(p, x);
    double ta = now_s();
    if (om_fill(p, x, resid, val) < 0) {
      ret = -1;
      break;
    }
    double tb = now_s();
    rep->t_assemble += tb - ta;
    double un = om_norm_L2(resid, N);
    om_row_sum_scale(N, val, p->ija, resid, scale);
    rep->t_scale += now_s() - tb;
    double Loo = om_norm_Loo(resid, N), L2 = om_norm_L2(resid, N);
    if (inewton < 64) {
      rep->resid_L2[inewton] = L2;
      rep->resid_Loo[inewton] = Loo;
      rep->unscaled_L2[inewton] = un;
    }
    if (inewton > 0 && L2 < o->eps_resid && Loo < o->eps_resid) converged = o->eps_corr > 1.0;
    if (!isfinite(Loo) || !isfinite(L2)) {
      ret = -1;
      break;
    }
    if (inewton == 0) initial_norm = L2;
    (void)initial_norm;
    if (!converged) {
      double tc = now_s(), rr = 0.0;
      ilu_t M;
      if (ilu_setup(&M, N, val, p->ija, p, o->precond) < 0) {
        ilu_free(&M);
        ret = -1;
        break;
      }
      double td = now_s();
      rep->t_precond += td - tc;
      int its = gmres_core(N, val, p->ija, &M, resid, dx, o->kspace, o->max_its, o->lin_tol, &rr);
      ilu_free(&M);
      rep->t_solve += now_s() - td;
      if (inewton < 64) rep->lin_its[inewton] = its;
      rep->lin_its_total += its;
      double cL2r = om_norm_L2_r(dx, x, N), cLoor = om_norm_Loo_r(dx, x, N);
      double taken = o->damp;
      if (o->line_search)
        taken = backtrack(p, x, dx, val, scale, L2, o->eps_resid, o->min_damp);
      else
        for (int i = 0; i < N; i++) x[i] -= o->damp * dx[i];
      if (inewton < 64) rep->damp[inewton] = taken;
      converged = (L2 < o->eps_resid) && (Loo < o->eps_resid) && (cL2r < o->eps_corr) &&
                  (cLoor < o->eps_corr);
    }
    inewton++;
  }
  rep->t_total = now_s() - t0;
  rep->newton_its = inewton;
  rep->converged = converged;
  free(resid);
  free(dx);
  free(scale);
  free(val);
  return ret < 0 ? -1 : inewton;
}
