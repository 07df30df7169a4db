"""Newton's post-update correction test and the error returns on the GPU.

* Normalized Correction Tolerance eps2 <= 1 (src/mm_sol_nonlinear.c:1545-1550,
  ~2704-2710): the pre-solve test no longer ends the loop; a pass converges
  only when the residual norms are below eps0 and the relative correction
  norms L2_norm_r / Loo_norm_r (3249) are below eps2.  The GPU takes the same
  number of passes as the oracle and reaches its solution; with an
  unreachable eps2 it runs to the pass limit unconverged, like the oracle.
* matrix_fill_full's -1 on a zero or negative element Jacobian
  (src/mm_fill.c:285-311) is GM_ERR_ELEM from the fill and from Newton.
* A non-finite residual norm ends Newton with GM_ERR_NAN (the reference's
  -1 through goto free_and_clear).
"""
import ctypes as C

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


@pytest.mark.parametrize("eps_corr", [1e-6, 1e-30])
def test_correction_tolerance_matches_oracle(eps_corr, capi, oracle_mod):
    prob = problems.lid_cavity(n=8, Re=100.0)
    print(f"[case] newton cavity n=8 eps_corr={eps_corr}", flush=True)
    ctx = capi.Context(prob)
    orc = oracle_mod.Oracle(prob)
    rc, x_ref, rep_ref = orc.newton(precond=1, kspace=50, max_its=1000, eps_corr=eps_corr,
                                    max_newton=12)
    xd = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    rc, rep = ctx.solve_dev(xd.data_ptr(), kspace=50, max_its=1000, patch_bins=2,
                            eps_corr=eps_corr, max_newton=12)
    assert bool(rep.converged) == bool(rep_ref.converged)
    assert rep.newton_its == rep_ref.newton_its
    if eps_corr < 1e-20:
        assert not rep.converged and rep.newton_its == 12
    else:
        # one pass more than the residual test alone needs
        rc0, rep0 = ctx.solve_dev(torch.zeros_like(xd).data_ptr(), kspace=50, max_its=1000,
                                  patch_bins=2)
        assert rep.converged and rep.newton_its >= rep0.newton_its
        x = xd.cpu().numpy()
        assert np.abs(x - x_ref).max() <= 1e-9 * np.abs(x_ref).max()
    ctx.close()


def _inverted(prob):
    """Element 0 mirrored in s: local nodes 0<->1, 3<->2, 7<->5 (detJ < 0)."""
    conn = prob.mesh.connect.copy()
    e = conn[0].copy()
    for a, b in ((0, 1), (3, 2), (7, 5)):
        e[a], e[b] = e[b], e[a]
    conn[0] = e
    prob.mesh.connect = conn
    return prob


def test_inverted_element_returns_elem_error(capi, oracle_mod):
    prob = _inverted(problems.lid_cavity(n=3, Re=10.0))
    print("[case] inverted element", flush=True)
    orc = oracle_mod.Oracle(prob)
    rc, _, _ = orc.fill(np.zeros(orc.N))
    assert rc == -1
    ctx = capi.Context(prob)
    xd = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    rd = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    vd = torch.zeros(ctx.nval, dtype=torch.float64, device="cuda")
    assert ctx.fill_dev(xd.data_ptr(), rd.data_ptr(), vd.data_ptr()) == capi.GM_ERR_ELEM
    rep = capi.GmNewtonReport()
    o = capi.solver_opts(patch_bins=1)
    rc = capi.lib().gm_solve_nonlinear_dev(ctx.h, xd.data_ptr(), C.byref(o), C.byref(rep))
    assert rc == capi.GM_ERR_ELEM
    ctx.close()


def test_nan_state_returns_nan_error(capi):
    prob = problems.natural_convection(n=4, Ra=1e4)
    print("[case] non-finite state", flush=True)
    ctx = capi.Context(prob)
    mask, _ = ctx.dirichlet()
    x = np.zeros(ctx.N)
    x[np.nonzero(mask == 0)[0][5]] = np.nan
    xd = torch.from_numpy(x).cuda()
    rep = capi.GmNewtonReport()
    o = capi.solver_opts(patch_bins=1)
    rc = capi.lib().gm_solve_nonlinear_dev(ctx.h, xd.data_ptr(), C.byref(o), C.byref(rep))
    assert rc == capi.GM_ERR_NAN
    assert rep.newton_its == 0
    ctx.close()
