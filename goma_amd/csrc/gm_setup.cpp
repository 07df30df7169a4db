// Host-side set-up of a gm_ctx: unknown map, node-node graph, MSR graph,
// Dirichlet map, element colouring and the ILU(0) patch ordering.
//
// These run once per mesh, on the host, before anything touches the GPU.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "gm_internal.h"

// Interaction mask, set_interaction_masks (src/mm_unknown_map.c:1226-1260
// momentum rows couple to U,T,P; 1692-1720 energy rows to U,T only;
// 2276-2310 continuity rows to U,T,P).
static inline bool inter_mask(int row_var, int col_var) {
  return !(row_var == GM_VAR_T && col_var == GM_VAR_P);
}

int gm_build_host(gm_ctx *c) {
  const int d = c->dim, nve = c->nve, nn = c->num_nodes, ne = c->num_elems;
  const int en = c->energy;
  const int cl = (d == 2) ? 8 : 20;  // centroid local node (el_elm_info.c:935-938, 1109)

  c->is_centroid.assign(nn, -1);
  c->elem_centroid.resize(ne);
  for (int e = 0; e < ne; e++) {
    int n = c->conn[(size_t)e * nve + cl];
    c->elem_centroid[e] = n;
    c->is_centroid[n] = e;
  }
  // Node-major unknowns, variable types ascending at each node
  // (fill_variable_vector, src/mm_fill_util.c:3516).
  c->first_unknown.resize(nn + 1);
  c->first_unknown[0] = 0;
  for (int n = 0; n < nn; n++)
    c->first_unknown[n + 1] = c->first_unknown[n] + d + en + (c->is_centroid[n] >= 0 ? d + 1 : 0);
  long long Nl = c->first_unknown[nn];
  if (Nl >= (1LL << 31)) {
    gm_set_error("too many unknowns (%lld)", Nl);
    return GM_ERR_ARG;
  }
  c->N = (int)Nl;
  c->unk_var.resize(c->N);
  for (int n = 0; n < nn; n++) {
    int f = c->first_unknown[n], k = 0;
    for (int a = 0; a < d; a++) c->unk_var[f + k++] = GM_VAR_U1 + a;
    if (en) c->unk_var[f + k++] = GM_VAR_T;
    if (c->is_centroid[n] >= 0)
      for (int a = 0; a <= d; a++) c->unk_var[f + k++] = GM_VAR_P;
  }

  // node -> elements
  std::vector<int> ne_ptr(nn + 1, 0);
  for (size_t i = 0; i < (size_t)ne * nve; i++) ne_ptr[c->conn[i] + 1]++;
  for (int n = 0; n < nn; n++) ne_ptr[n + 1] += ne_ptr[n];
  std::vector<int> ne_list(ne_ptr[nn]);
  {
    std::vector<int> fp(ne_ptr.begin(), ne_ptr.end() - 1);
    for (int e = 0; e < ne; e++)
      for (int i = 0; i < nve; i++) ne_list[fp[c->conn[(size_t)e * nve + i]]++] = e;
  }
  // Node-node graph: sorted neighbours incl. self (build_node_node,
  // src/exo_conn.c:204-377).
  std::vector<int> cnt(nn + 1, 0);
#pragma omp parallel
  {
    std::vector<int> tmp;
#pragma omp for schedule(dynamic, 1024)
    for (int n = 0; n < nn; n++) {
      tmp.clear();
      for (int k = ne_ptr[n]; k < ne_ptr[n + 1]; k++) {
        int e = ne_list[k];
        tmp.insert(tmp.end(), c->conn.begin() + (size_t)e * nve, c->conn.begin() + (size_t)(e + 1) * nve);
      }
      std::sort(tmp.begin(), tmp.end());
      cnt[n + 1] = (int)(std::unique(tmp.begin(), tmp.end()) - tmp.begin());
    }
  }
  c->nn_ptr.assign(nn + 1, 0);
  for (int n = 0; n < nn; n++) c->nn_ptr[n + 1] = c->nn_ptr[n] + cnt[n + 1];
  c->nn_list.resize(c->nn_ptr[nn]);
#pragma omp parallel
  {
    std::vector<int> tmp;
#pragma omp for schedule(dynamic, 1024)
    for (int n = 0; n < nn; n++) {
      tmp.clear();
      for (int k = ne_ptr[n]; k < ne_ptr[n + 1]; k++) {
        int e = ne_list[k];
        tmp.insert(tmp.end(), c->conn.begin() + (size_t)e * nve, c->conn.begin() + (size_t)(e + 1) * nve);
      }
      std::sort(tmp.begin(), tmp.end());
      int u = (int)(std::unique(tmp.begin(), tmp.end()) - tmp.begin());
      std::copy(tmp.begin(), tmp.begin() + u, c->nn_list.begin() + c->nn_ptr[n]);
    }
  }

  // MSR graph (find_MSR_problem_graph, src/mm_fill_util.c:3229-3445):
  // columns of a row = unknowns of every neighbour node, sorted node order,
  // masked by the interaction mask, diagonal excluded.
  std::vector<long long> rowlen(c->N + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int n = 0; n < nn; n++) {
    for (int u = c->first_unknown[n]; u < c->first_unknown[n + 1]; u++) {
      int rv = c->unk_var[u];
      long long k = 0;
      for (int q = c->nn_ptr[n]; q < c->nn_ptr[n + 1]; q++) {
        int J = c->nn_list[q];
        for (int w = c->first_unknown[J]; w < c->first_unknown[J + 1]; w++)
          if (inter_mask(rv, c->unk_var[w]) && w != u) k++;
      }
      rowlen[u + 1] = k;
    }
  }
  for (int i = 0; i < c->N; i++) rowlen[i + 1] += rowlen[i];
  long long tot = (long long)c->N + 1 + rowlen[c->N];
  if (tot >= (1LL << 31)) {
    gm_set_error("MSR graph too large for 32-bit bindx (%lld)", tot);
    return GM_ERR_ARG;
  }
  c->nval = tot;
  c->bindx.resize(tot);
  for (int i = 0; i <= c->N; i++) c->bindx[i] = (int)(c->N + 1 + rowlen[i]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int n = 0; n < nn; n++) {
    for (int u = c->first_unknown[n]; u < c->first_unknown[n + 1]; u++) {
      int rv = c->unk_var[u];
      long long p = c->bindx[u];
      for (int q = c->nn_ptr[n]; q < c->nn_ptr[n + 1]; q++) {
        int J = c->nn_list[q];
        for (int w = c->first_unknown[J]; w < c->first_unknown[J + 1]; w++)
          if (inter_mask(rv, c->unk_var[w]) && w != u) c->bindx[p++] = w;
      }
    }
  }

  // Element colouring: greedy first-fit in element order; elements of one
  // colour share no node, so their scatters never touch the same nonzero.
  c->elem_colour.assign(ne, -1);
  {
    std::vector<int> mark(256, -1);
    int ncol = 0;
    for (int e = 0; e < ne; e++) {
      for (int i = 0; i < nve; i++) {
        int n = c->conn[(size_t)e * nve + i];
        for (int k = ne_ptr[n]; k < ne_ptr[n + 1]; k++) {
          int f = ne_list[k];
          if (c->elem_colour[f] >= 0) mark[c->elem_colour[f]] = e;
        }
      }
      int col = 0;
      while (mark[col] == e) col++;
      c->elem_colour[e] = col;
      ncol = std::max(ncol, col + 1);
    }
    c->n_ecol = ncol;
    c->ecol_ptr.assign(ncol + 1, 0);
    for (int e = 0; e < ne; e++) c->ecol_ptr[c->elem_colour[e] + 1]++;
    for (int k = 0; k < ncol; k++) c->ecol_ptr[k + 1] += c->ecol_ptr[k];
    c->ecol_elems.resize(ne);
    std::vector<int> fp(c->ecol_ptr.begin(), c->ecol_ptr.end() - 1);
    for (int e = 0; e < ne; e++) c->ecol_elems[fp[c->elem_colour[e]]++] = e;
  }
  return GM_OK;
}

// global_h_elem_siz (src/mm_fill_aux.c:1128-1200) with the QUAD / HEX
// widths of h_elem_siz (src/mm_fill_aux.c:844-1125): along each reference
// direction the squared distance between the centroids of the two opposite
// faces, from the vertex nodes (local 0-3 or 0-7); an element contributes
// sqrt(mean of the squared widths), the result is the mean over elements.
double gm_h_elem_size(const gm_ctx *c) {
  // vertex sets of the faces at xi_r = -1 and xi_r = +1 (Goma's local order)
  static const int q_lo[2][2] = {{0, 3}, {0, 1}}, q_hi[2][2] = {{1, 2}, {2, 3}};
  static const int h_lo[3][4] = {{0, 3, 4, 7}, {0, 1, 4, 5}, {4, 5, 6, 7}};
  static const int h_hi[3][4] = {{1, 2, 5, 6}, {2, 3, 6, 7}, {0, 1, 2, 3}};
  const int d = c->dim, nf = d == 2 ? 2 : 4;
  double sum = 0.0;
  for (int e = 0; e < c->num_elems; e++) {
    const int *en = &c->conn[(size_t)e * c->nve];
    double hsq = 0.0;
    for (int r = 0; r < d; r++) {
      double w2 = 0.0;
      for (int a = 0; a < d; a++) {
        const double *xa = &c->coor[(size_t)a * c->num_nodes];
        double lo = 0.0, hi = 0.0;
        for (int k = 0; k < nf; k++) {
          lo += xa[en[d == 2 ? q_lo[r][k] : h_lo[r][k]]];
          hi += xa[en[d == 2 ? q_hi[r][k] : h_hi[r][k]]];
        }
        const double w = (hi - lo) / nf;
        w2 += w * w;
      }
      hsq += w2;
    }
    sum += std::sqrt(hsq / d);
  }
  return sum / c->num_elems;
}

// tau of the global PSPG scheme (calc_pspg, src/mm_fill_stabilization.c:
// 1037-1049) at c->U_norm; mu_avg is element_viscosity's (src/mm_fill_aux.c:
// 687-751): the viscosity model at a dim x dim shear-rate tensor of 1e-5
// entries, so gammadot = sqrt(dim^2 / 2) 1e-5.
double gm_pspg_tau(const gm_ctx *c) {
  const gm_physics &p = c->phys;
  double mu = p.mu0;
  if (p.visc_model == 1) {
    const double gd = std::sqrt(0.5 * c->dim * c->dim) * 1e-5;
    mu = p.muinf + (p.mu0 - p.muinf) * std::pow(1.0 + std::pow(p.lam * gd, p.aexp), (p.nexp - 1.0) / p.aexp);
  }
  const double h = c->h_elem_avg, U = c->U_norm;
  if (p.rho * U * h / (2.0 * mu) <= 3.0) return p.ps_scaling * h * h / (12.0 * mu);
  return p.ps_scaling * h / (2.0 * p.rho * U);
}

int gm_default_bins(const gm_ctx *c) {
  // aim at patches of about 8x8 elements in 2D and 4x4x4 in 3D
  double per_dir = std::pow((double)c->num_elems, 1.0 / c->dim);
  int target = (c->dim == 2) ? 8 : 4;
  int b = (int)std::lround(per_dir / target);
  return std::max(1, b);
}

// ILU(0) patch ordering.  Nodes are binned into `bins` boxes per axis over
// the mesh bounding box; patches are coloured greedily in patch order so that
// no two patches of one colour hold coupled nodes; unknowns are ordered
// colour-major, then patch, then node (ascending), then unknown at the node.
int gm_patch_order(const gm_ctx *c, int bins, std::vector<int> &perm, std::vector<int> &colour_ptr,
                   std::vector<int> &patch_row) {
  const int nn = c->num_nodes, d = c->dim;
  double lo[3], hi[3];
  for (int k = 0; k < d; k++) {
    lo[k] = 1e300;
    hi[k] = -1e300;
    for (int n = 0; n < nn; n++) {
      double v = c->coor[(size_t)k * nn + n];
      lo[k] = std::min(lo[k], v);
      hi[k] = std::max(hi[k], v);
    }
  }
  std::vector<int> patch(nn);
  for (int n = 0; n < nn; n++) {
    int id = 0, mul = 1;
    for (int k = 0; k < d; k++) {
      double w = hi[k] - lo[k];
      int b = w > 0 ? (int)std::floor((c->coor[(size_t)k * nn + n] - lo[k]) * bins / w) : 0;
      b = std::max(0, std::min(bins - 1, b));
      id += b * mul;
      mul *= bins;
    }
    patch[n] = id;
  }
  int np = 1;
  for (int k = 0; k < d; k++) np *= bins;
  std::vector<int> pn_ptr(np + 1, 0), pn(nn);
  for (int n = 0; n < nn; n++) pn_ptr[patch[n] + 1]++;
  for (int i = 0; i < np; i++) pn_ptr[i + 1] += pn_ptr[i];
  {
    std::vector<int> fp(pn_ptr.begin(), pn_ptr.end() - 1);
    for (int n = 0; n < nn; n++) pn[fp[patch[n]]++] = n;
  }
  std::vector<int> pc(np, -1), mark(4096, -1);
  int ncol = 0;
  for (int P = 0; P < np; P++) {
    for (int k = pn_ptr[P]; k < pn_ptr[P + 1]; k++) {
      int n = pn[k];
      for (int q = c->nn_ptr[n]; q < c->nn_ptr[n + 1]; q++) {
        int Q = patch[c->nn_list[q]];
        if (Q != P && pc[Q] >= 0) mark[pc[Q]] = P;
      }
    }
    int col = 0;
    while (mark[col] == P) col++;
    if (col >= 4095) {
      gm_set_error("patch colouring overflow");
      return GM_ERR_ARG;
    }
    pc[P] = col;
    ncol = std::max(ncol, col + 1);
  }
  perm.resize(c->N);
  colour_ptr.assign(ncol + 1, 0);
  patch_row.clear();
  int k = 0, npatch = 0;
  for (int col = 0; col < ncol; col++) {
    colour_ptr[col] = npatch;  // patch index where this colour starts
    for (int P = 0; P < np; P++) {
      if (pc[P] != col || pn_ptr[P] == pn_ptr[P + 1]) continue;
      patch_row.push_back(k);
      npatch++;
      for (int q = pn_ptr[P]; q < pn_ptr[P + 1]; q++) {
        int n = pn[q];
        for (int u = c->first_unknown[n]; u < c->first_unknown[n + 1]; u++) perm[k++] = u;
      }
    }
  }
  colour_ptr[ncol] = npatch;
  patch_row.push_back(k);
  return ncol;
}
