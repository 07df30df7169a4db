"""Structured mesh hierarchy and Q2/P1 transfer operators for the multigrid
preconditioner.

A coarse mesh keeps every other element vertex of the fine mesh (so jittered
meshes coarsen consistently); its mid-edge, mid-face and centroid nodes are
the means of its vertices, as in :func:`goma_amd.mesh.structured_mesh`.

The prolongation P maps coarse unknowns to fine unknowns:
  * nodal unknowns (U1..Ud, T): coarse Q2 interpolation in reference
    coordinates -- a fine lattice point at half-integer coarse position takes
    the 1D weights (3/8, 3/4, -1/8) per axis, tensorised;
  * P1 pressure {p0, ps, pt[, pu]} of a fine element inside coarse element E,
    at offset sigma = +-1 per axis: p(s_c) = P0 + sum Pk s_c,k with
    s_c = (s_f + sigma)/2, i.e. p0_f = P0 + sum sigma_k Pk / 2 and pk_f = Pk / 2.
Both are exact for Q2 x P1 fields on affine meshes.
"""
from __future__ import annotations

import numpy as np

from .mesh import CENTROID_LOCAL, HEX27_POS, QUAD9_POS, Mesh


def coarsen(fine: Mesh) -> Mesh:
    if fine.n % 2:
        raise ValueError("mesh must have an even number of elements per side")
    d, n = fine.dim, fine.n
    nc = n // 2
    mf, mc = 2 * n + 1, 2 * nc + 1
    latf = fine.lattice()
    # fine vertices sit at lattice multiples of 2; coarse vertices at multiples of 4
    shape = (nc + 1,) * d
    V = np.zeros(shape + (d,))
    idx = np.indices(shape).reshape(d, -1).T  # (ix, iy[, iz])
    fid = np.zeros(len(idx), dtype=np.int64)
    for k in range(d):
        fid += (4 * idx[:, k]) * (mf**k)
    V.reshape(-1, d)[:] = fine.coor[:, fid].T
    X = V
    for k in range(d):
        sh = list(X.shape)
        sh[k] = mc
        Y = np.empty(sh)
        ev = [slice(None)] * (d + 1)
        ev[k] = slice(0, mc, 2)
        Y[tuple(ev)] = X
        od = [slice(None)] * (d + 1)
        od[k] = slice(1, mc, 2)
        a = [slice(None)] * (d + 1)
        b = [slice(None)] * (d + 1)
        a[k] = slice(0, nc)
        b[k] = slice(1, nc + 1)
        Y[tuple(od)] = 0.5 * (X[tuple(a)] + X[tuple(b)])
        X = Y
    Xt = np.transpose(X, tuple(range(d - 1, -1, -1)) + (d,))
    coor = np.ascontiguousarray(Xt.reshape(-1, d).T)
    pos = QUAD9_POS if d == 2 else HEX27_POS
    e = np.arange(nc**d, dtype=np.int64)
    base = [2 * (e % nc) + 1, 2 * ((e // nc) % nc) + 1]
    if d == 3:
        base.append(2 * (e // (nc * nc)) + 1)
    conn = np.zeros((nc**d, pos.shape[0]), dtype=np.int64)
    for k in range(d):
        conn += (base[k][:, None] + pos[None, :, k]) * (mc**k)
    return Mesh(dim=d, n=nc, coor=coor, connect=conn.astype(np.int32))


def _lag1(s):
    return np.array([-0.5 * s * (1 - s), 1 - s * s, 0.5 * s * (1 + s)])


def prolongation(fine_fu: np.ndarray, coarse_fu: np.ndarray, fine: Mesh, coarse: Mesh,
                 energy: int):
    """CSR (rowptr, col, val) of P: coarse unknowns -> fine unknowns, and the
    injection map coarse node -> fine node at the same lattice point."""
    d, n = fine.dim, fine.n
    nc = coarse.n
    mf, mc = 2 * n + 1, 2 * nc + 1
    nv = d + energy
    # per-axis weights: fine lattice g -> (coarse lattice index, weight) x 3
    ci = np.zeros((mf, 3), np.int64)
    cw = np.zeros((mf, 3))
    for g in range(mf):
        if g % 2 == 0:
            ci[g] = [g // 2, g // 2, g // 2]
            cw[g] = [1.0, 0.0, 0.0]
        else:
            h = g / 2.0
            c = min(int(np.floor(h / 2.0)) * 2 + 1, mc - 2)
            ci[g] = [c - 1, c, c + 1]
            cw[g] = _lag1(h - c)
    latf = fine.lattice()
    ids = np.zeros((fine.num_nodes, 1), np.int64)
    ws = np.ones((fine.num_nodes, 1))
    for k in range(d):
        g = latf[:, k]
        ids = (ids[:, :, None] + ci[g][:, None, :] * (mc**k)).reshape(fine.num_nodes, -1)
        ws = (ws[:, :, None] * cw[g][:, None, :]).reshape(fine.num_nodes, -1)
    keep = ws != 0.0
    node_r = np.repeat(np.arange(fine.num_nodes), keep.sum(axis=1))
    node_c = ids[keep]
    node_w = ws[keep]
    rows = [fine_fu[node_r][None, :] + np.arange(nv)[:, None]]
    cols = [coarse_fu[node_c][None, :] + np.arange(nv)[:, None]]
    vals = [np.broadcast_to(node_w, (nv, len(node_w)))]
    rows = [r.ravel() for r in rows]
    cols = [c.ravel() for c in cols]
    vals = [v.ravel() for v in vals]
    # pressure: fine element -> parent coarse element
    cl = CENTROID_LOCAL[d]
    fe = np.arange(fine.num_elems)
    ex = [fe % n, (fe // n) % n] + ([fe // (n * n)] if d == 3 else [])
    ec = sum((ex[k] // 2) * (nc**k) for k in range(d))
    pf = fine_fu[fine.connect[:, cl]] + nv
    pc = coarse_fu[coarse.connect[ec, cl]] + nv
    rows.append(pf)
    cols.append(pc)
    vals.append(np.ones(len(fe)))
    for k in range(d):
        sig = np.where(ex[k] % 2 == 1, 1.0, -1.0)
        rows.append(pf)
        cols.append(pc + 1 + k)
        vals.append(0.5 * sig)
        rows.append(pf + 1 + k)
        cols.append(pc + 1 + k)
        vals.append(np.full(len(fe), 0.5))
    rows = np.concatenate(rows).astype(np.int64)
    cols = np.concatenate(cols).astype(np.int64)
    vals = np.concatenate(vals).astype(np.float64)
    Nf = int(fine_fu[-1])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rowptr = np.zeros(Nf + 1, np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr)
    latc = coarse.lattice()
    inj = np.zeros(coarse.num_nodes, np.int64)
    for k in range(d):
        inj += (2 * latc[:, k]) * (mf**k)
    return rowptr.astype(np.int32), cols.astype(np.int32), vals, inj.astype(np.int32)


def coarse_problem(prob, cmesh):
    """The problem's physics and Dirichlet masks on a coarsened mesh: a
    coarse node is Dirichlet for a variable when the fine node at the same
    position is."""
    from .problems import Problem

    fine = prob.mesh
    d = fine.dim
    mf = 2 * fine.n + 1
    latc = cmesh.lattice()
    fnode = np.zeros(cmesh.num_nodes, np.int64)
    for k in range(d):
        fnode += (2 * latc[:, k]) * (mf**k)
    inv = -np.ones(fine.num_nodes, np.int64)
    inv[fnode] = np.arange(cmesh.num_nodes)
    cn = inv[prob.dbc_node]
    keep = cn >= 0
    return Problem(name=prob.name + "_c", mesh=cmesh, phys=prob.phys,
                   dbc_node=cn[keep].astype(np.int32), dbc_var=prob.dbc_var[keep],
                   dbc_value=np.zeros(int(keep.sum())), datum_elem=0, datum_value=0.0,
                   meta=dict(prob.meta))


def hierarchy(prob, min_elems_per_side: int = 2):
    """Coarse problems down to ``min_elems_per_side``."""
    out = []
    cur = prob
    while cur.mesh.n % 2 == 0 and cur.mesh.n // 2 >= min_elems_per_side:
        cm = coarsen(cur.mesh)
        cur = coarse_problem(cur, cm)
        out.append(cur)
    return out
