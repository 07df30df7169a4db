"""Learned initial guess for a Rayleigh-number continuation.

A coordinate-plus-parameter MLP (inputs x, y[, z], log10 Ra; outputs u, v[, w],
p, T) is trained by libgm's own BF16 matrix-core kernels on the converged
solutions of a continuation sweep (goma_amd.continuation).  At a new Ra it
predicts the Newton initial guess; gm_solve_nonlinear_guess_dev falls back to
the previous-step guess of zeroth-order continuation (src/ac_conti.c:760-787)
when the scaled residual has not dropped 10x within 3 Newton passes.

Training data.  For each sweep solution: the nodal (u, v[, w], T) at every
mesh node, and the discontinuous P1 pressure P0 + P_s s + P_t t[ + P_u u] at
the element centroid and at the reference points +-1/2 along each axis.  These
are the points at which gm_mlp_guess_dev evaluates the network, in the same
order (nodes, then per element: centroid, +s, -s, +t, -t[, +u, -u]).

Normalisation.  Targets are normalised per snapshot and per field (shift =
mean, scale = standard deviation over the field's points).  At a query Ra the
shift and log(scale) are interpolated piecewise-linearly in log10 Ra between
the bracketing snapshots, so the network learns field shapes while the
amplitudes, which grow with Ra by orders of magnitude, come from the sweep.

Guess forms.  ``direct``: x = G(Ra).  ``delta``: x = x_prev + G(Ra) - G(Ra_prev),
the previous solution moved by the network's predicted change, in which the
network's own fitting error largely cancels.
"""
from __future__ import annotations

import numpy as np

from . import capi
from .mesh import CENTROID_LOCAL, HEX27_POS, QUAD9_POS


def _lagrange1(pos, s):
    if pos < 0:
        return -0.5 * s * (1.0 - s)
    if pos > 0:
        return 0.5 * s * (1.0 + s)
    return 1.0 - s * s


def q2_basis(dim: int, xi) -> np.ndarray:
    """Q2 basis (QUAD9 / HEX27, Goma local order) at one reference point."""
    pos = QUAD9_POS if dim == 2 else HEX27_POS
    out = np.ones(len(pos))
    for k, p in enumerate(pos):
        for d in range(dim):
            out[k] *= _lagrange1(int(p[d]), float(xi[d]))
    return out


def sample_refs(dim: int) -> np.ndarray:
    """Reference points per element: centroid, then +-1/2 along each axis."""
    refs = [np.zeros(dim)]
    for d in range(dim):
        for sgn in (0.5, -0.5):
            r = np.zeros(dim)
            r[d] = sgn
            refs.append(r)
    return np.array(refs)


def sample_points(mesh) -> np.ndarray:
    """[num_nodes + num_elems * (1 + 2 dim), dim] coordinates: the nodes, then
    each element's pressure sample points through its Q2 map."""
    dim = mesh.dim
    refs = sample_refs(dim)
    phi = np.array([q2_basis(dim, r) for r in refs])  # [spe, nve]
    X = mesh.coor.T[mesh.connect]  # [ne, nve, dim]
    elem = np.einsum("jk,ekd->ejd", phi, X).reshape(-1, dim)
    return np.concatenate([mesh.coor.T, elem], axis=0)


def field_values(prob, first_unknown: np.ndarray, x: np.ndarray):
    """Targets and masks at the sample points for solution x.

    Output order (u, v[, w], p[, T]); nodes carry the nodal fields, element
    points the pressure."""
    mesh = prob.mesh
    dim, en = mesh.dim, int(prob.phys.energy)
    n_out = dim + 1 + en
    nn, ne = mesh.num_nodes, mesh.num_elems
    spe = 1 + 2 * dim
    t = np.zeros((nn + ne * spe, n_out), np.float64)
    m = np.zeros_like(t, dtype=np.float32)
    fu = first_unknown[:nn].astype(np.int64)
    for d in range(dim):
        t[:nn, d] = x[fu + d]
        m[:nn, d] = 1
    if en:
        t[:nn, dim + 1] = x[fu + dim]
        m[:nn, dim + 1] = 1
    cen = fu[mesh.connect[:, CENTROID_LOCAL[dim]]] + dim + en  # P0 unknown of each element
    P = np.stack([x[cen + k] for k in range(dim + 1)], axis=1)  # [ne, dim+1]
    refs = sample_refs(dim)  # [spe, dim]
    pv = P[:, :1] + P[:, 1:] @ refs.T  # [ne, spe]
    t[nn:, dim] = pv.reshape(-1)
    m[nn:, dim] = 1
    return t, m


class LearnedGuess:
    """Network + per-snapshot normalisation for one problem's Ra sweep."""

    def __init__(self, prob, ctx, ras, sols, hidden: int = 4, width: int = 256,
                 seed: int = 20261202):
        self.prob, self.ctx = prob, ctx
        self.dim = prob.mesh.dim
        self.n_out = self.dim + 1 + int(prob.phys.energy)
        self.lra = np.log10(np.asarray(ras, dtype=np.float64))
        fu = ctx.first_unknown()
        pts = sample_points(prob.mesh)
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        llo, lhi = self.lra.min(), self.lra.max()
        in_shift = list(0.5 * (lo + hi)) + [0.5 * (llo + lhi)]
        in_scale = list(2.0 / np.maximum(hi - lo, 1e-30)) + [2.0 / max(lhi - llo, 1e-30)]
        self.net = capi.Mlp(self.dim + 1, self.n_out, hidden=hidden, width=width, seed=seed,
                            in_shift=in_shift, in_scale=in_scale)
        shifts, scales, ins, tgs, mks = [], [], [], [], []
        for lr, x in zip(self.lra, sols):
            t, m = field_values(prob, fu, x)
            sh = np.array([t[m[:, k] > 0, k].mean() for k in range(self.n_out)])
            sc = np.array([t[m[:, k] > 0, k].std() for k in range(self.n_out)])
            sc = np.maximum(sc, 1e-12 * np.maximum(np.abs(sh), 1.0))
            shifts.append(sh)
            scales.append(sc)
            ins.append(np.concatenate([pts, np.full((len(pts), 1), lr)], axis=1))
            tgs.append((t - sh) / sc)
            mks.append(m)
        self.shift = np.array(shifts)
        self.logscale = np.log(np.array(scales))
        self.train_in = np.concatenate(ins).astype(np.float32)
        self.train_t = np.concatenate(tgs).astype(np.float32)
        self.train_m = np.concatenate(mks).astype(np.float32)

    @property
    def num_points(self) -> int:
        return self.train_in.shape[0]

    def train(self, **opts) -> np.ndarray:
        self.net.set_data(self.train_in, self.train_t, self.train_m)
        return self.net.train(**opts)

    def query(self, Ra: float) -> capi.GmMlpQuery:
        lr = float(np.log10(Ra))
        sh = np.array([np.interp(lr, self.lra, self.shift[:, k]) for k in range(self.n_out)])
        sc = np.exp([np.interp(lr, self.lra, self.logscale[:, k]) for k in range(self.n_out)])
        return capi.mlp_query([lr], sc, sh)

    def guess(self, x_ptr: int, Ra: float, Ra_prev: float | None = None,
              x_prev_ptr: int | None = None):
        """Writes the guess at Ra into x (direct form, or delta form from the
        solution x_prev at Ra_prev)."""
        if Ra_prev is None:
            self.net.guess_dev(self.ctx, self.query(Ra), x_ptr)
        else:
            self.net.guess_dev(self.ctx, self.query(Ra), x_ptr, self.query(Ra_prev), x_prev_ptr)

    def set_basis(self, sols):
        """Project delta-form guesses onto the span of the sweep's solution
        differences, weighted per field by 1 / RMS of those differences
        (Dirichlet unknowns excluded).  Returns the subspace dimension."""
        import torch

        from .mesh import CENTROID_LOCAL

        S = np.asarray(sols, dtype=np.float64)
        D = S[1:] - S[:1]
        m = self.prob.mesh
        d, en = m.dim, int(self.prob.phys.energy)
        fu = self.ctx.first_unknown()[: m.num_nodes].astype(np.int64)
        cen = fu[m.connect[:, CENTROID_LOCAL[d]]] + d + en
        groups = [fu + k for k in range(d)] + ([fu + d] if en else []) + \
                 [np.concatenate([cen + k for k in range(d + 1)])]
        w = np.zeros(self.ctx.N)
        for g in groups:
            rms = np.sqrt(np.mean(D[:, g] ** 2))
            w[g] = 1.0 / rms if rms > 0 else 0.0
        mask, _ = self.ctx.dirichlet()
        w[mask > 0] = 0.0
        self._snaps = torch.tensor(S, device="cuda")
        self._w = torch.tensor(w, device="cuda")
        nb = self.net.set_basis(self.ctx, self._snaps.data_ptr(), len(S), self._w.data_ptr())
        self._snaps = None
        return nb

    def close(self):
        self.net.close()
