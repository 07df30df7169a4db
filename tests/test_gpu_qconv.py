"""QCONV convective heat transfer on the GPU against the oracle.

QCONV (src/bc_integ.c:1492-1494: qrad_surf with zero emissivity,
src/mm_ns_bc.c:15145-15255) adds h (T_c - T) to the inward heat flux over
element sides: the side integral of phi_i h (T_c - T) in the energy residual
and -h times the side integral of phi_i phi_j in the T-T Jacobian (added in
both fill kernels after the volume terms, and to the element matrices the
Galerkin coarse operators are built from).  Bars: residual and Jacobian
within 1e-12 of the largest entry (FP32 Jacobian 1e-6), also in the stored
(not accumulated) assembly, with QSIDE on a neighbouring side; conduction
with T = 0 at x = 0 and QCONV at x = 1 solves to T = h T_c x / (k + h) at the
nodes (within 1e-10); a 3D cavity with energy and a convective wall through
Galerkin multigrid matches the oracle's Newton solution within 1e-9 of max|x|.
"""
from dataclasses import replace

import numpy as np
import pytest

from goma_amd import exodus, problems
from goma_amd.problems import TEMP

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BAND = dict(precond=3, kspace=20, max_its=40, lin_tol=1e-10, max_newton=12)


@pytest.fixture(scope="module")
def capi():
    from goma_amd import capi as M

    M.lib()
    assert torch.cuda.is_available()
    return M


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("dim,n", [(2, 4), (3, 2)])
def test_qconv_fill_matches_oracle(dim, n, capi, oracle_mod):
    print(f"[case] qconv fill dim {dim}", flush=True)
    prob = (problems.natural_convection(n=n, Ra=1e4, jitter=0.15) if dim == 2 else
            problems.lid_cavity(n=n, Re=100.0, dim=3, energy=True, Ra=1e3, jitter=0.15))
    sides = exodus.wall_side_sets(prob.mesh)
    ctx = capi.Context(prob)
    orc = oracle_mod.Oracle(prob)
    # QCONV on a face without temperature Dirichlet rows (2D: y = 0; 3D: x = 1),
    # QSIDE on a face sharing an edge with it
    e, sd = sides[3 if dim == 2 else 2]
    ctx.add_qconv(e, sd, 1.3, 0.4)
    assert orc.add_qconv(e, sd, 1.3, 0.4) == 0
    e, sd = sides[2 if dim == 2 else 5]
    ctx.add_qside(e, sd, -0.3)
    assert orc.add_qside(e, sd, -0.3) == 0
    x = np.random.default_rng(5).standard_normal(orc.N) * 0.3
    _, R_ref, V_ref = orc.fill(x)
    _, R_plain, V_plain = oracle_mod.Oracle(prob).fill(x)
    assert _relerr(V_ref, V_plain) > 1e-6  # the condition is in the Jacobian
    r = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    v = torch.zeros(ctx.nval, dtype=torch.float64, device="cuda")
    xd = _dev(x)
    assert ctx.fill_dev(xd.data_ptr(), r.data_ptr(), v.data_ptr()) == 0
    assert _relerr(r.cpu().numpy(), R_ref) < 1e-12 and _relerr(v.cpu().numpy(), V_ref) < 1e-12
    # stored, not accumulated, over NaN (val[N] is never written)
    vn = torch.full((ctx.nval,), float("nan"), dtype=torch.float64, device="cuda")
    r.zero_()
    assert ctx.fill_assign_dev(xd.data_ptr(), r.data_ptr(), vn.data_ptr()) == 0
    va = vn.cpu().numpy()
    va[ctx.N] = V_ref[ctx.N]
    assert np.isfinite(va).all() and _relerr(va, V_ref) < 1e-12
    # FP32 Jacobian values
    v32 = torch.zeros(ctx.nval, dtype=torch.float32, device="cuda")
    r.zero_()
    assert ctx.fill32_dev(xd.data_ptr(), r.data_ptr(), v32.data_ptr()) == 0
    assert _relerr(r.cpu().numpy(), R_ref) < 1e-12
    assert _relerr(v32.cpu().numpy().astype(np.float64), V_ref) < 1e-6
    ctx.close()


def _conduction(dim, n, k):
    """No flow (g = 0), T = 0 on x = 0 only."""
    prob = (problems.natural_convection(n=n, Ra=0.0, jitter=0.1) if dim == 2 else
            problems.lid_cavity(n=n, dim=3, energy=True, Ra=0.0, Re=1.0, jitter=0.1))
    keep = prob.dbc_var != TEMP
    tn = np.flatnonzero(prob.mesh.lattice()[:, 0] == 0)
    return replace(prob, dbc_node=np.concatenate([prob.dbc_node[keep], tn]).astype(np.int32),
                   dbc_var=np.concatenate([prob.dbc_var[keep], np.full(tn.size, TEMP)]).astype(np.int32),
                   dbc_value=np.zeros(int(keep.sum()) + tn.size),
                   phys=replace(prob.phys, kcond=k, g=(0.0, 0.0, 0.0)))


@pytest.mark.parametrize("dim,n,k,h,tc", [(2, 4, 2.5, 3.0, 0.8), (3, 2, 0.4, 1.5, -0.6)])
def test_qconv_conduction_is_linear(dim, n, k, h, tc, capi):
    print(f"[case] qconv conduction dim {dim}", flush=True)
    prob = _conduction(dim, n, k)
    e, sd = exodus.wall_side_sets(prob.mesh)[2]
    ctx = capi.Context(prob)
    ctx.add_qconv(e, sd, h, tc)
    x = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    rc, rep = ctx.solve_dev(x.data_ptr(), **BAND)
    assert rep.converged
    tdof = ctx.first_unknown().astype(np.int64)[:-1] + dim
    T = x.cpu().numpy()[tdof]
    assert np.abs(T - h * tc * prob.mesh.coor[0] / (k + h)).max() <= 1e-10 * abs(tc)
    ctx.close()


@pytest.mark.parametrize("galerkin", [0, 1])
def test_qconv_newton_multigrid_3d(galerkin, capi, oracle_mod):
    """HEX27 cavity with energy and a convective wall through multigrid-
    preconditioned GMRES; with Galerkin coarse operators the term reaches the
    coarse levels through the element matrices."""
    print(f"[case] qconv newton multigrid 3d galerkin={galerkin}", flush=True)
    prob = problems.lid_cavity(n=4, Re=100.0, dim=3, energy=True, Ra=1e3)
    e, sd = exodus.wall_side_sets(prob.mesh)[5]  # z = 0: no temperature Dirichlet rows inside
    ctx, orc = capi.Context(prob), oracle_mod.Oracle(prob)
    ctx.add_qconv(e, sd, 2.0, 0.3)
    assert orc.add_qconv(e, sd, 2.0, 0.3) == 0
    rc, x_ref, rep_ref = orc.newton(precond=1, kspace=100, max_its=2000, max_newton=25)
    assert rep_ref.converged
    assert ctx.attach_multigrid(2) == 2
    xd = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    rc, rep = ctx.solve_dev(xd.data_ptr(), kspace=50, max_its=200, precond=2, mg_galerkin=galerkin,
                            max_newton=25, line_search=1)
    assert rep.converged
    assert np.abs(xd.cpu().numpy() - x_ref).max() <= 1e-9 * np.abs(x_ref).max()
    ctx.close()
