"""Galerkin multigrid coarse operators (gm_solver_opts.mg_galerkin) on the GPU.

The coarse matrices are checked against R A P formed with SciPy from the
oracle's natural Jacobian (the same problem without Dirichlet rows, so that
A = sum of element matrices) and the prolongations the hierarchy attaches;
coarse Dirichlet rows must be Goma's per-element unit rows (k x the unit row,
k = coarse elements sharing the node).  Bars: 1e-12 of the largest entry in
FP64; Newton with Galerkin multigrid reaches the oracle's solution to 1e-8.
"""
from dataclasses import replace

import numpy as np
import pytest

from goma_amd import problems

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def capi():
    from goma_amd import capi as M

    M.lib()
    assert torch.cuda.is_available()
    return M


def _natural(prob):
    return replace(prob, dbc_node=prob.dbc_node[:0], dbc_var=prob.dbc_var[:0],
                   dbc_value=prob.dbc_value[:0], datum_elem=-1)


CASES = {
    "carreau2d_jitter": lambda: problems.natural_convection(n=8, Ra=1e4, carreau=True, jitter=0.15),
    "cavity3d_energy": lambda: problems.lid_cavity(n=4, Re=400.0, dim=3, energy=True, Ra=1e4),
}


@pytest.mark.parametrize("name", list(CASES))
def test_galerkin_coarse_matrices_equal_rap(name, capi, oracle_mod):
    import scipy.sparse as sp

    prob = CASES[name]()
    print(f"[case] galerkin coarse matrices {name}", flush=True)
    ctx = capi.Context(prob)
    nl = ctx.attach_multigrid(2)
    assert nl >= 2
    x0 = np.random.default_rng(11).standard_normal(ctx.N) * 0.3
    mask, val = ctx.dirichlet()
    x0[mask > 0] = val[mask > 0]
    xd = torch.from_numpy(x0.copy()).cuda()
    # one Newton pass: assembly at x0, then the multigrid set-up from it
    ctx.solve_dev(xd.data_ptr(), max_newton=1, precond=2, mg_galerkin=1, kspace=10, max_its=2)
    orc = oracle_mod.Oracle(_natural(prob))
    rc, _, v = orc.fill(x0)
    assert rc == 0
    A = oracle_mod.msr_to_scipy(orc.ija, v, orc.N)
    n_prev = ctx.N
    for lev in range(1, nl):
        rp, ci, cv = ctx.mg_prolongation(lev)
        cc = ctx._coarse[lev - 1]
        P = sp.csr_matrix((cv, ci, rp), shape=(n_prev, cc.N))
        A = (P.T @ A @ P).tocsr()  # natural Galerkin operator of this level
        G = oracle_mod.msr_to_scipy(cc.bindx(), ctx.mg_level_matrix(lev), cc.N).tocsr()
        cmask, _ = cc.dirichlet()
        free = np.flatnonzero(cmask == 0)
        fixed = np.flatnonzero(cmask > 0)
        scale = abs(A).max()
        err = abs(G[free] - A[free]).max() if len(free) else 0.0
        assert err <= 1e-12 * scale, f"level {lev}: {err} vs {scale}"
        D = G[fixed].toarray()
        diag = D[np.arange(len(fixed)), fixed]
        assert np.all(diag >= 1.0) and np.all(diag == np.round(diag))
        D[np.arange(len(fixed)), fixed] = 0.0
        assert not D.any()
        n_prev = cc.N
    ctx.close()


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("name", ["natconv2d", "carreau2d_jitter"])
def test_newton_galerkin_multigrid_matches_oracle(name, prec, capi, oracle_mod):
    prob = (problems.natural_convection(n=8, Ra=1e4) if name == "natconv2d" else
            problems.natural_convection(n=8, Ra=1e3, carreau=True, jitter=0.15))
    print(f"[case] newton galerkin multigrid {name} prec={prec}", flush=True)
    ctx = capi.Context(prob)
    ctx.attach_multigrid(2)
    orc = oracle_mod.Oracle(prob)
    rc, x_ref, rep_ref = orc.newton(precond=1, kspace=50, max_its=1000, line_search=1)
    assert rep_ref.converged
    xd = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc, rep = ctx.solve_dev(xd.data_ptr(), kspace=50, max_its=200, precond=2, precision=prec,
                            line_search=1, mg_smoother=1, mg_nu=2, mg_galerkin=1)
    assert rep.converged
    x = xd.cpu().numpy()
    assert np.abs(x - x_ref).max() <= 1e-8 * np.abs(x_ref).max()
    ctx.close()
