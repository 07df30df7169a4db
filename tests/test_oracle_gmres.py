"""The oracle's GMRES(m) and ILU(0), pinned by what the mathematics fixes (CPU).

Restarted, right-preconditioned GMRES is defined by its iterates: after j
steps of a cycle started at x_s with r_s = b - A x_s, the iterate is the
minimiser of ||b - A x||_2 over x in x_s + M^-1 K_j(A M^-1, r_s).  Any
orthogonalisation (Aztec's classical Gram-Schmidt, the oracle's twice-applied
one) gives the same iterates in exact arithmetic, so this pins the iterate
sequence Goma's AZ_gmres produces, up to rounding.  The minimisers here come
from an independent route: SVD least squares on the explicitly formed,
column-normalised Krylov power basis.  ILU(0) is pinned by its defining
property: (L U)_ij = A_ij on the pattern of A, L unit lower, U upper.
"""
import numpy as np
import pytest

from goma_amd import problems


def _system(oracle_mod, seed=1):
    """Row-sum-scaled Jacobian and residual of a small Boussinesq problem."""
    prob = problems.natural_convection(n=2, Ra=1e4, jitter=0.1)
    orc = oracle_mod.Oracle(prob)
    x = np.random.default_rng(seed).standard_normal(orc.N) * 0.3
    rc, r, v = orc.fill(x)
    assert rc == 0
    s = np.zeros(orc.N)
    oracle_mod.lib().om_row_sum_scale(orc.N, v, orc.ija, r, s)
    A = oracle_mod.msr_to_scipy(orc.ija, v, orc.N).toarray()
    return orc, v, A, r


def _krylov_minimiser(B, r, j):
    """argmin_c ||r - B K c|| with K = [r, B r, .., B^(j-1) r] (columns scaled
    to unit norm), by SVD least squares.  Returns (K c, residual norm)."""
    cols = [r / np.linalg.norm(r)]
    for _ in range(j - 1):
        w = B @ cols[-1]
        cols.append(w / np.linalg.norm(w))
    K = np.stack(cols, axis=1)
    c, *_ = np.linalg.lstsq(B @ K, r, rcond=None)
    z = K @ c
    return z, np.linalg.norm(r - B @ z)


def _ilu_dense(orc, v):
    rc, rowptr, col, lu = orc.ilu0(v)
    assert rc == len(lu)  # the factor's entry count; negative on a zero pivot
    N = orc.N
    LU = np.zeros((N, N))
    for i in range(N):
        LU[i, col[rowptr[i]:rowptr[i + 1]]] = lu[rowptr[i]:rowptr[i + 1]]
    L = np.tril(LU, -1) + np.eye(N)
    U = np.triu(LU)
    return L, U


def test_ilu0_reproduces_the_matrix_on_its_pattern(oracle_mod):
    orc, v, A, _ = _system(oracle_mod)
    L, U = _ilu_dense(orc, v)
    P = A != 0
    np.fill_diagonal(P, True)
    LU = L @ U
    assert np.abs(LU[P] - A[P]).max() <= 1e-12 * np.abs(A).max()
    # and it is an incomplete factorisation: fill outside the pattern is dropped
    assert np.abs(LU[~P]).max() > 1e-8


def test_unrestarted_gmres_iterates_are_krylov_minimisers(oracle_mod):
    orc, v, A, b = _system(oracle_mod)
    for j in range(1, 9):
        its, x, _ = orc.gmres(v, b, precond=0, kspace=30, max_its=j, tol=1e-300)
        assert its == j
        z, rmin = _krylov_minimiser(A, b, j)
        rn = np.linalg.norm(b - A @ x)
        assert abs(rn - rmin) <= 1e-9 * np.linalg.norm(b), j
        assert np.linalg.norm(x - z) <= 1e-6 * max(np.linalg.norm(z), 1e-30), j


def test_restarted_gmres_minimises_cycle_by_cycle(oracle_mod):
    orc, v, A, b = _system(oracle_mod, seed=2)
    its, x, _ = orc.gmres(v, b, precond=0, kspace=3, max_its=7, tol=1e-300)
    assert its == 7
    xs = np.zeros(orc.N)
    for length in (3, 3, 1):  # GMRES(3): two full cycles, then one step
        z, _ = _krylov_minimiser(A, b - A @ xs, length)
        xs = xs + z
    assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)
    assert abs(np.linalg.norm(b - A @ x) - np.linalg.norm(b - A @ xs)) <= 1e-9 * np.linalg.norm(b)


def test_ilu_gmres_is_right_preconditioned(oracle_mod):
    """With ILU(0), the iterates minimise the true residual over
    M^-1 K_j(A M^-1, b); a left-preconditioned method (minimising
    ||M^-1 (b - A x)||) would give other iterates, and the test tells them apart."""
    orc, v, A, b = _system(oracle_mod, seed=3)
    L, U = _ilu_dense(orc, v)
    Minv = np.linalg.inv(L @ U)
    AM = A @ Minv
    left_gap = 0.0
    for j in range(1, 7):
        its, x, _ = orc.gmres(v, b, precond=1, kspace=30, max_its=j, tol=1e-300)
        assert its == j
        z, rmin = _krylov_minimiser(AM, b, j)
        xr = Minv @ z
        assert abs(np.linalg.norm(b - A @ x) - rmin) <= 1e-9 * np.linalg.norm(b), j
        assert np.linalg.norm(x - xr) <= 1e-6 * np.linalg.norm(xr), j
        zl, _ = _krylov_minimiser(Minv @ A, Minv @ b, j)  # left-preconditioned iterate
        left_gap = max(left_gap, np.linalg.norm(zl - xr) / np.linalg.norm(xr))
    assert left_gap > 1e-4
