// Geometric multigrid hierarchy of a structured QUAD9 / HEX27 mesh, built
// behind the C ABI (gm_mg_build): coarsened meshes, their contexts with the
// Dirichlet masks of the finer level, the Q2 / P1 prolongations and the
// injection maps, attached level by level with gm_mg_add_level.
//
// A mesh has a hierarchy when it is the structured lattice mesh of n^dim
// elements: nodes on a (2n+1)^dim lattice numbered x fastest (then y, then
// z), elements numbered x fastest, each element's local nodes at its centre
// lattice point plus Goma's local node offsets.  Coarsening keeps every other
// element vertex, so jittered meshes coarsen consistently; coarse mid-edge,
// mid-face and centroid nodes are the means of their vertices, axis by axis.
//
// Prolongation (coarse unknowns -> finer unknowns), exact for Q2 x P1 fields
// on affine meshes:
//   * nodal unknowns: coarse Q2 interpolation in reference coordinates; a
//     fine lattice point at a half-integer coarse position takes the 1D
//     weights of the quadratic Lagrange basis at s = -1/2 or +1/2 per axis,
//     tensorised (zero weights dropped);
//   * P1 pressure {p0, p_s, p_t[, p_u]} of a finer element at offset
//     sigma = +-1 per axis inside its coarse parent: p0 = P0 + sum sigma_k P_k / 2,
//     p_k = P_k / 2.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gm_internal.h"

namespace {

const signed char kQuad9Pos[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1},
                                     {1, 0},   {0, 1},  {-1, 0}, {0, 0}};
const signed char kHex27Pos[27][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
    {-1, 1, 1},   {0, -1, -1}, {1, 0, -1}, {0, 1, -1},  {-1, 0, -1}, {-1, -1, 0}, {1, -1, 0},
    {1, 1, 0},    {-1, 1, 0},  {0, -1, 1}, {1, 0, 1},   {0, 1, 1},   {-1, 0, 1},  {0, 0, 0},
    {0, 0, -1},   {0, 0, 1},   {-1, 0, 0}, {1, 0, 0},   {0, -1, 0},  {0, 1, 0}};

int pos(int dim, int k, int a) { return dim == 2 ? kQuad9Pos[k][a] : kHex27Pos[k][a]; }

long long ipow(long long b, int e) {
  long long r = 1;
  for (int i = 0; i < e; i++) r *= b;
  return r;
}

// Elements per side if the mesh is the structured lattice mesh, else 0.
int lattice_n(const gm_ctx *c) {
  const int d = c->dim, nve = c->nve;
  int n = (int)std::llround(std::pow((double)c->num_elems, 1.0 / d));
  while (n > 1 && ipow(n, d) > c->num_elems) n--;
  while (ipow(n + 1, d) <= c->num_elems) n++;
  if (n < 1 || ipow(n, d) != c->num_elems || ipow(2 * n + 1, d) != c->num_nodes) return 0;
  const long long m = 2 * n + 1;
  for (int e = 0; e < c->num_elems; e++) {
    long long ex[3] = {e % n, (e / n) % n, d == 3 ? e / ((long long)n * n) : 0};
    for (int k = 0; k < nve; k++) {
      long long id = 0;
      for (int a = d - 1; a >= 0; a--) id = id * m + (2 * ex[a] + 1 + pos(d, k, a));
      if (c->conn[(size_t)e * nve + k] != id) return 0;
    }
  }
  return n;
}

struct Coarse {
  int n = 0;
  std::vector<double> coor;  // [dim][nodes]
  std::vector<int> conn;     // [elems][nve]
};

// Coarse mesh of a structured fine mesh with n (even) elements per side.
Coarse coarsen(const gm_ctx *f, int n) {
  const int d = f->dim, nve = f->nve, nc = n / 2;
  const long long mf = 2 * n + 1, mc = 2 * nc + 1;
  Coarse C;
  C.n = nc;
  // vertex grid (nc+1)^d, then each axis refined in turn: even index = the
  // coordinate already there, odd index = mean of its two neighbours
  std::vector<long long> ext(d, nc + 1);
  std::vector<double> X((size_t)ipow(nc + 1, d) * d);
  for (long long v = 0; v < ipow(nc + 1, d); v++) {
    long long fid = 0, r = v;
    long long idx[3] = {0, 0, 0};
    for (int a = 0; a < d; a++) {
      idx[a] = r % (nc + 1);
      r /= (nc + 1);
    }
    for (int a = d - 1; a >= 0; a--) fid = fid * mf + 4 * idx[a];
    for (int k = 0; k < d; k++) X[(size_t)v * d + k] = f->coor[(size_t)k * f->num_nodes + fid];
  }
  for (int ax = 0; ax < d; ax++) {
    std::vector<long long> next = ext;
    next[ax] = mc;
    long long tot = 1;
    for (int a = 0; a < d; a++) tot *= next[a];
    std::vector<double> Y((size_t)tot * d);
    for (long long p = 0; p < tot; p++) {
      long long idx[3] = {0, 0, 0}, r = p;
      for (int a = 0; a < d; a++) {
        idx[a] = r % next[a];
        r /= next[a];
      }
      auto flat = [&](long long i_ax) {
        long long q = 0;
        for (int a = d - 1; a >= 0; a--) q = q * ext[a] + (a == ax ? i_ax : idx[a]);
        return q;
      };
      const long long i = idx[ax];
      for (int k = 0; k < d; k++) {
        if (i % 2 == 0) {
          Y[(size_t)p * d + k] = X[(size_t)flat(i / 2) * d + k];
        } else {
          Y[(size_t)p * d + k] =
              0.5 * (X[(size_t)flat((i - 1) / 2) * d + k] + X[(size_t)flat((i + 1) / 2) * d + k]);
        }
      }
    }
    X.swap(Y);
    ext = next;
  }
  const long long nn = ipow(mc, d);
  C.coor.assign((size_t)d * nn, 0.0);
  for (long long p = 0; p < nn; p++)
    for (int k = 0; k < d; k++) C.coor[(size_t)k * nn + p] = X[(size_t)p * d + k];
  const long long ne = ipow(nc, d);
  C.conn.assign((size_t)ne * nve, 0);
  for (long long e = 0; e < ne; e++) {
    long long ex[3] = {e % nc, (e / nc) % nc, d == 3 ? e / ((long long)nc * nc) : 0};
    for (int k = 0; k < nve; k++) {
      long long id = 0;
      for (int a = d - 1; a >= 0; a--) id = id * mc + (2 * ex[a] + 1 + pos(d, k, a));
      C.conn[(size_t)e * nve + k] = (int)id;
    }
  }
  return C;
}

// Quadratic Lagrange weights at reference s of the nodes at -1, 0, +1.
void lag1(double s, double w[3]) {
  w[0] = -0.5 * s * (1.0 - s);
  w[1] = 1.0 - s * s;
  w[2] = 0.5 * s * (1.0 + s);
}

struct Transfer {
  std::vector<int> rowptr, col, inject;
  std::vector<double> val;
};

// Prolongation CSR (rows sorted by column) from the coarse context cc to the
// finer context fc, structured with n (fine) and n / 2 (coarse) elements per side.
Transfer prolongation(const gm_ctx *fc, const gm_ctx *cc, int n) {
  const int d = fc->dim, nv = fc->nv, nc = n / 2, cl = d == 2 ? 8 : 20;
  const long long mf = 2 * n + 1, mc = 2 * nc + 1;
  std::vector<long long> ci(mf * 3);
  std::vector<double> cw(mf * 3);
  for (long long g = 0; g < mf; g++) {
    if (g % 2 == 0) {
      for (int j = 0; j < 3; j++) {
        ci[g * 3 + j] = g / 2;
        cw[g * 3 + j] = j == 0 ? 1.0 : 0.0;
      }
    } else {
      const double h = g / 2.0;
      const long long c = std::min<long long>((long long)std::floor(h / 2.0) * 2 + 1, mc - 2);
      double w[3];
      lag1(h - (double)c, w);
      for (int j = 0; j < 3; j++) {
        ci[g * 3 + j] = c - 1 + j;
        cw[g * 3 + j] = w[j];
      }
    }
  }
  struct Ent {
    int r, c;
    double v;
  };
  std::vector<Ent> ents;
  const std::vector<int> &ff = fc->first_unknown, &cf = cc->first_unknown;
  for (long long node = 0; node < fc->num_nodes; node++) {
    long long g[3] = {node % mf, (node / mf) % mf, d == 3 ? node / (mf * mf) : 0};
    const int nt = d == 2 ? 9 : 27;
    for (int t = 0; t < nt; t++) {
      int j[3] = {t % 3, (t / 3) % 3, t / 9};
      double w = 1.0;
      long long id = 0;
      for (int a = 0; a < d; a++) w *= cw[g[a] * 3 + j[a]];
      for (int a = d - 1; a >= 0; a--) id = id * mc + ci[g[a] * 3 + j[a]];
      if (w == 0.0) continue;
      for (int a = 0; a < nv; a++) ents.push_back({ff[node] + a, cf[id] + a, w});
    }
  }
  for (long long e = 0; e < fc->num_elems; e++) {
    long long ex[3] = {e % n, (e / n) % n, d == 3 ? e / ((long long)n * n) : 0};
    long long ec = 0;
    for (int a = d - 1; a >= 0; a--) ec = ec * nc + ex[a] / 2;
    const int pf = ff[fc->conn[(size_t)e * fc->nve + cl]] + nv;
    const int pc = cf[cc->conn[(size_t)ec * cc->nve + cl]] + nv;
    ents.push_back({pf, pc, 1.0});
    for (int k = 0; k < d; k++) {
      const double sig = ex[k] % 2 == 1 ? 1.0 : -1.0;
      ents.push_back({pf, pc + 1 + k, 0.5 * sig});
      ents.push_back({pf + 1 + k, pc + 1 + k, 0.5});
    }
  }
  std::sort(ents.begin(), ents.end(),
            [](const Ent &a, const Ent &b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
  Transfer T;
  T.rowptr.assign(fc->N + 1, 0);
  T.col.resize(ents.size());
  T.val.resize(ents.size());
  for (size_t k = 0; k < ents.size(); k++) {
    T.rowptr[ents[k].r + 1]++;
    T.col[k] = ents[k].c;
    T.val[k] = ents[k].v;
  }
  for (int i = 0; i < fc->N; i++) T.rowptr[i + 1] += T.rowptr[i];
  T.inject.resize(cc->num_nodes);
  for (long long node = 0; node < cc->num_nodes; node++) {
    long long g[3] = {node % mc, (node / mc) % mc, d == 3 ? node / (mc * mc) : 0};
    long long id = 0;
    for (int a = d - 1; a >= 0; a--) id = id * mf + 2 * g[a];
    T.inject[node] = (int)id;
  }
  return T;
}

}  // namespace

extern "C" int gm_mg_build(gm_ctx *c, int min_elems_per_side) {
  if (!c || min_elems_per_side < 1) return GM_ERR_ARG;
  if (!c->mg.empty()) {
    gm_set_error("a multigrid hierarchy is already attached to this context");
    return GM_ERR_ARG;
  }
  int n = lattice_n(c);
  if (n == 0) {
    gm_set_error("geometric multigrid needs the structured lattice mesh (nodes and elements "
                 "numbered x fastest); this mesh has another numbering: use ILU(0) or the "
                 "banded LU");
    return GM_ERR_ARG;
  }
  gm_ctx *above = c;
  while (n % 2 == 0 && n / 2 >= min_elems_per_side) {
    Coarse cm = coarsen(above, n);
    // Dirichlet conditions of the coarse level: a coarse node is Dirichlet for
    // a nodal variable when the finer node at the same place is (value 0:
    // the levels carry corrections); pressure datum on coarse element 0
    const int d = above->dim, mc = 2 * cm.n + 1, mf = 2 * n + 1;
    const long long ncn = ipow(mc, d);
    std::vector<int> bnode, bvar;
    for (long long node = 0; node < ncn; node++) {
      long long g[3] = {node % mc, (node / mc) % mc, d == 3 ? node / ((long long)mc * mc) : 0};
      long long fid = 0;
      for (int a = d - 1; a >= 0; a--) fid = fid * mf + 2 * g[a];
      const int fu = above->first_unknown[fid];
      for (int a = 0; a < above->nv; a++)
        if (above->dbc[fu + a]) {
          bnode.push_back((int)node);
          bvar.push_back(a < d ? GM_VAR_U1 + a : GM_VAR_T);
        }
    }
    std::vector<double> bval(bnode.size(), 0.0);
    gm_mesh mesh{d, (int)ncn, (int)ipow(cm.n, d), cm.coor.data(), cm.conn.data()};
    gm_bc bc{(int)bnode.size(), bnode.data(), bvar.data(), bval.data(), 0, 0.0};
    gm_ctx *cc = nullptr;
    int rc = gm_ctx_create(&mesh, &c->phys, &bc, &cc);
    if (rc) return rc;
    c->owned_levels.push_back(cc);
    Transfer T = prolongation(above, cc, n);
    if ((rc = gm_mg_add_level(c, cc, T.rowptr.data(), T.col.data(), T.val.data(),
                              T.inject.data())))
      return rc;
    c->mg.back().P_host_rowptr = std::move(T.rowptr);
    c->mg.back().P_host_col = std::move(T.col);
    c->mg.back().P_host_val = std::move(T.val);
    above = cc;
    n = cm.n;
  }
  return (int)c->mg.size() + 1;
}

extern "C" gm_ctx *gm_mg_level(gm_ctx *c, int level) {
  if (!c || level < 0 || level > (int)c->mg.size()) return nullptr;
  return level == 0 ? c : c->mg[level - 1].ctx;
}

extern "C" int gm_mg_copy_prolongation(const gm_ctx *c, int level, int *rowptr, int *col,
                                       double *val) {
  if (!c || level < 1 || level > (int)c->mg.size()) return GM_ERR_ARG;
  const MgLevel &L = c->mg[level - 1];
  if (L.P_host_rowptr.empty()) {
    gm_set_error("level %d was attached without a host copy of its prolongation", level);
    return GM_ERR_ARG;
  }
  const int nnz = (int)L.P_host_col.size();
  if (rowptr) std::memcpy(rowptr, L.P_host_rowptr.data(), L.P_host_rowptr.size() * sizeof(int));
  if (col) std::memcpy(col, L.P_host_col.data(), nnz * sizeof(int));
  if (val) std::memcpy(val, L.P_host_val.data(), nnz * sizeof(double));
  return nnz;
}
