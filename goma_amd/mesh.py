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
