"""Structured QUAD9 / HEX27 meshes on the unit square / cube.

Goma reads its meshes from Exodus II files (rd_mesh.c:152); the workloads
here are synthetic, so this module writes the same two arrays Goma keeps after
reading a mesh: nodal coordinates ``coor[dim][num_nodes]`` (Goma's ``Coor``)
and element connectivity ``connect[num_elems][nodes_per_elem]`` in Goma's
local node order (src/rf_shape.c:361-470 for QUAD9, 1105-1460 for HEX27).

Node numbering: nodes sit on a (2n+1)^d lattice numbered x fastest, then y,
then z.  Element numbering: x fastest.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# Reference-coordinate position of each local node (s, t[, u]) in {-1, 0, 1}.
QUAD9_POS = np.array(
    [[-1, -1], [1, -1], [1, 1], [-1, 1], [0, -1], [1, 0], [0, 1], [-1, 0], [0, 0]], dtype=np.int64
)
HEX27_POS = np.array(
    [
        [-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
        [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
        [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1],
        [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0],
        [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1],
        [0, 0, 0],
        [0, 0, -1], [0, 0, 1], [-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0],
    ],
    dtype=np.int64,
)
CENTROID_LOCAL = {2: 8, 3: 20}


@dataclass
class Mesh:
    dim: int
    n: int  # elements per direction
    coor: np.ndarray  # float64 [dim, num_nodes]
    connect: np.ndarray  # int32 [num_elems, 9 or 27]

    @property
    def num_nodes(self) -> int:
        return self.coor.shape[1]

    @property
    def num_elems(self) -> int:
        return self.connect.shape[0]

    @property
    def nodes_per_elem(self) -> int:
        return self.connect.shape[1]

    def lattice(self) -> np.ndarray:
        """Integer lattice index (gx, gy[, gz]) of every node, shape [num_nodes, dim]."""
        m = 2 * self.n + 1
        ids = np.arange(self.num_nodes, dtype=np.int64)
        out = [ids % m, (ids // m) % m]
        if self.dim == 3:
            out.append(ids // (m * m))
        return np.stack(out, axis=1)


def structured_mesh(dim: int, n: int, jitter: float = 0.0, seed: int = 20261202) -> Mesh:
    """Unit square (dim 2, QUAD9) or cube (dim 3, HEX27) with n elements per side.

    With ``jitter > 0`` every interior element vertex moves by an independent
    uniform offset in [-jitter*h, jitter*h] per coordinate (h = 1/n, numpy
    ``default_rng(seed)``); mid-edge, mid-face and centroid nodes then sit at
    the mean of the vertices around them, so element edges stay straight.
    """
    if dim not in (2, 3):
        raise ValueError("dim must be 2 or 3")
    if n < 1:
        raise ValueError("n must be >= 1")
    h = 1.0 / n
    m = 2 * n + 1
    # vertex lattice (n+1)^dim
    vx = np.linspace(0.0, 1.0, n + 1)
    grids = np.meshgrid(*([vx] * dim), indexing="ij")  # grids[k][i0, i1, (i2)] along axis k
    V = np.stack(grids, axis=-1)  # [n+1]*dim + [dim], index order (ix, iy, iz)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        off = rng.uniform(-jitter * h, jitter * h, size=V.shape)
        interior = np.ones(V.shape[:-1], dtype=bool)
        for k in range(dim):
            sl = [slice(None)] * dim
            sl[k] = 0
            interior[tuple(sl)] = False
            sl[k] = n
            interior[tuple(sl)] = False
        V = V + off * interior[..., None]
    # full Q2 lattice: even index -> vertex, odd index -> mean of the two neighbours
    X = V
    for k in range(dim):
        shape = list(X.shape)
        shape[k] = m
        Y = np.empty(shape)
        sl_even = [slice(None)] * (dim + 1)
        sl_even[k] = slice(0, m, 2)
        Y[tuple(sl_even)] = X
        sl_odd = [slice(None)] * (dim + 1)
        sl_odd[k] = slice(1, m, 2)
        a = [slice(None)] * (dim + 1)
        b = [slice(None)] * (dim + 1)
        a[k] = slice(0, n)
        b[k] = slice(1, n + 1)
        Y[tuple(sl_odd)] = 0.5 * (X[tuple(a)] + X[tuple(b)])
        X = Y
    # flatten with x fastest: X is indexed (ix, iy, iz) -> transpose to (iz, iy, ix)
    Xt = np.transpose(X, tuple(range(dim - 1, -1, -1)) + (dim,))
    coor = np.ascontiguousarray(Xt.reshape(-1, dim).T)

    pos = QUAD9_POS if dim == 2 else HEX27_POS
    e = np.arange(n**dim, dtype=np.int64)
    ex = e % n
    ey = (e // n) % n
    base = [2 * ex + 1, 2 * ey + 1]
    if dim == 3:
        base.append(2 * (e // (n * n)) + 1)
    conn = np.zeros((n**dim, pos.shape[0]), dtype=np.int64)
    for k in range(dim):
        g = base[k][:, None] + pos[None, :, k]
        conn += g * (m**k)
    return Mesh(dim=dim, n=n, coor=coor, connect=conn.astype(np.int32))


def renumber(mesh: Mesh, node_order: np.ndarray, elem_order: np.ndarray) -> Mesh:
    """The same mesh with new numbering: new node k is old node
    ``node_order[k]``, new element j is old element ``elem_order[j]``.  The
    result has no structured hierarchy (``n == 0``)."""
    node_order = np.asarray(node_order, np.int64)
    elem_order = np.asarray(elem_order, np.int64)
    if not np.array_equal(np.sort(node_order), np.arange(mesh.num_nodes)):
        raise ValueError("node_order is not a permutation of the nodes")
    if not np.array_equal(np.sort(elem_order), np.arange(mesh.num_elems)):
        raise ValueError("elem_order is not a permutation of the elements")
    new_id = np.empty_like(node_order)
    new_id[node_order] = np.arange(node_order.size)
    return Mesh(dim=mesh.dim, n=0, coor=np.ascontiguousarray(mesh.coor[:, node_order]),
                connect=new_id[mesh.connect[elem_order]].astype(np.int32))


def _node_graph(mesh: Mesh):
    """Sparse node adjacency: nodes sharing an element."""
    import scipy.sparse as sp

    ne, npe = mesh.connect.shape
    B = sp.csr_matrix((np.ones(ne * npe, np.int8), (mesh.connect.reshape(-1).astype(np.int64),
                                                   np.repeat(np.arange(ne), npe))),
                      shape=(mesh.num_nodes, ne))
    return (B @ B.T).tocsr()


def _dof_half_bandwidth(mesh: Mesh, new_id: np.ndarray) -> int:
    """Half bandwidth of the unknowns' matrix under node renumbering
    ``new_id`` (unknowns node by node: dim velocities + one temperature at
    every node, dim + 1 P1 coefficients more at each centroid).  Two nodes
    couple only inside an element, so the widest coupling of an element is
    from the first unknown of its lowest node to the last of its highest."""
    count = np.full(mesh.num_nodes, mesh.dim + 1, np.int64)
    count[mesh.connect[:, CENTROID_LOCAL[mesh.dim]]] += mesh.dim + 1
    fu = np.zeros(mesh.num_nodes + 1, np.int64)
    np.cumsum(count[np.argsort(new_id)], out=fu[1:])
    ids = new_id[mesh.connect]
    return int((fu[ids.max(axis=1) + 1] - 1 - fu[ids.min(axis=1)]).max())


def band_order(mesh: Mesh):
    """A node and element numbering that keeps the Jacobian's band narrow.

    Unknowns are numbered node by node, so the band (the banded LU's storage
    and work) follows the node numbering, and a mesh from a file can come in
    any numbering.  Two candidates are tried and the narrower band kept:
    a coordinate sweep (nodes sorted with the longest extent slowest, which
    on a box reproduces row-by-row numbering) and reverse Cuthill-McKee of
    the node graph (for other shapes).  Elements follow their lowest new
    node.  Returns (node_order, elem_order) for ``renumber``."""
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    ext = mesh.coor.max(axis=1) - mesh.coor.min(axis=1)
    axes = np.argsort(ext, kind="stable")  # np.lexsort: last key is the slowest
    cands = [np.lexsort(tuple(mesh.coor[a] for a in axes)).astype(np.int64),
             np.asarray(reverse_cuthill_mckee(_node_graph(mesh), symmetric_mode=True), np.int64)]
    best, best_bw = None, None
    for order in cands:
        new_id = np.empty_like(order)
        new_id[order] = np.arange(order.size)
        bw = _dof_half_bandwidth(mesh, new_id)
        if best_bw is None or bw < best_bw:
            best, best_bw, best_id = order, bw, new_id
    elem_order = np.argsort(best_id[mesh.connect].min(axis=1), kind="stable")
    return best, elem_order
