"""The multigrid hierarchy built behind the C ABI (gm_mg_build) against the
NumPy restatement in goma_amd/multigrid.py, whose transfers are pinned on the
CPU (test_multigrid_transfer.py: exact reproduction of Q2 x P1 fields).

For every level: the prolongation CSR is identical to multigrid.prolongation
(bit for bit); the level's context has the unknown map, MSR graph and
Dirichlet map of a context built on multigrid.coarse_problem; and both
assemble bit-identical residuals and Jacobians at the same state, which
fixes the coarse coordinates and connectivity.  Then a Newton solve through
the hierarchy, reached through gm_* symbols only, matches the oracle.
"""
import ctypes as C

import numpy as np
import pytest

from goma_amd import multigrid, problems

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def capi():
    from goma_amd import capi as M

    M.lib()
    assert torch.cuda.is_available()
    return M


CASES = {
    "natconv2d_n8_jitter": lambda: problems.natural_convection(n=8, Ra=1e4, jitter=0.15),
    "cavity3d_n4_energy": lambda: problems.lid_cavity(n=4, Re=100.0, dim=3, energy=True, Ra=1e3,
                                                      jitter=0.1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_hierarchy_matches_numpy_restatement(name, capi):
    print(f"[case] mg_build {name}", flush=True)
    prob = CASES[name]()
    ctx = capi.Context(prob)
    nl = ctx.attach_multigrid(2)
    levels = multigrid.hierarchy(prob, 2)
    assert nl == len(levels) + 1
    above_prob, above = prob, ctx
    rng = np.random.default_rng(3)
    for lev, cp in enumerate(levels, start=1):
        cc = ctx._coarse[lev - 1]
        ref = capi.Context(cp)
        assert cc.N == ref.N and cc.nval == ref.nval
        np.testing.assert_array_equal(cc.bindx(), ref.bindx())
        np.testing.assert_array_equal(cc.first_unknown(), ref.first_unknown())
        m1, v1 = cc.dirichlet()
        m2, v2 = ref.dirichlet()
        np.testing.assert_array_equal(m1, m2)
        np.testing.assert_array_equal(v1, v2)
        rp, ci, cv = ctx.mg_prolongation(lev)
        rp2, ci2, cv2, _ = multigrid.prolongation(above.first_unknown(), ref.first_unknown(),
                                                  above_prob.mesh, cp.mesh, cp.phys.energy)
        np.testing.assert_array_equal(rp, rp2)
        np.testing.assert_array_equal(ci, ci2)
        np.testing.assert_array_equal(cv, cv2)
        x = torch.from_numpy(rng.standard_normal(cc.N) * 0.3).cuda()
        out = []
        for c in (cc, ref):
            r = torch.zeros(c.N, dtype=torch.float64, device="cuda")
            v = torch.zeros(c.nval, dtype=torch.float64, device="cuda")
            assert c.fill_dev(x.data_ptr(), r.data_ptr(), v.data_ptr()) == 0
            out.append((r.cpu(), v.cpu()))
        assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
        ref.close()
        above_prob, above = cp, cc
    ctx.close()


def test_multigrid_through_the_c_abi_only(capi, oracle_mod):
    """Context creation, gm_mg_build and the Newton solve, all through gm_*
    symbols (ctypes), as a C caller would make them."""
    print("[case] mg_build C ABI only", flush=True)
    L = capi.lib()
    prob = problems.natural_convection(n=8, Ra=1e4)
    m = prob.mesh
    coor = np.ascontiguousarray(m.coor, np.float64)
    conn = np.ascontiguousarray(m.connect, np.int32)
    bn = np.ascontiguousarray(prob.dbc_node, np.int32)
    bv = np.ascontiguousarray(prob.dbc_var, np.int32)
    bx = np.ascontiguousarray(prob.dbc_value, np.float64)
    gm = capi.GmMesh(m.dim, m.num_nodes, m.num_elems, coor.ctypes.data, conn.ctypes.data)
    ph = capi.physics_struct(prob.phys)
    bc = capi.GmBC(len(bn), bn.ctypes.data, bv.ctypes.data, bx.ctypes.data, prob.datum_elem,
                   prob.datum_value)
    h = C.c_void_p()
    assert L.gm_ctx_create(C.byref(gm), C.byref(ph), C.byref(bc), C.byref(h)) == 0
    assert L.gm_mg_build(h, 2) == 3  # 8 -> 4 -> 2 elements per side
    assert L.gm_mg_num_levels(h) == 3
    assert L.gm_mg_build(h, 2) == capi.GM_ERR_ARG  # a second hierarchy is refused
    assert L.gm_mg_level(h, 3) is None
    N = L.gm_num_unknowns(h)
    x = torch.zeros(N, dtype=torch.float64, device="cuda")
    o = capi.solver_opts(kspace=50, max_its=200, precond=2, line_search=1, mg_smoother=1,
                         mg_galerkin=1, lin_tol_adapt=1, max_newton=30)
    rep = capi.GmNewtonReport()
    rc = L.gm_solve_nonlinear_dev(h, x.data_ptr(), C.byref(o), C.byref(rep))
    assert rc > 0 and rep.converged
    orc = oracle_mod.Oracle(prob)
    rc, x_ref, rep_ref = orc.newton(precond=1, kspace=50, max_its=1000, line_search=1)
    assert rep_ref.converged
    assert np.abs(x.cpu().numpy() - x_ref).max() <= 1e-8 * np.abs(x_ref).max()
    L.gm_ctx_destroy(h)


def test_renumbered_mesh_has_no_hierarchy(capi):
    """A mesh whose nodes are not numbered on the lattice is refused by the
    library itself (the Python wrapper's own check bypassed)."""
    prob = problems.natural_convection(n=4, Ra=1e4)
    perm = np.random.default_rng(1).permutation(prob.mesh.num_nodes)
    inv = np.argsort(perm)
    from dataclasses import replace

    mesh = replace(prob.mesh, coor=prob.mesh.coor[:, perm], connect=inv[prob.mesh.connect].astype(np.int32))
    rp = replace(prob, mesh=mesh, dbc_node=inv[prob.dbc_node].astype(np.int32))
    ctx = capi.Context(rp)
    with pytest.raises(capi.GomaError, match="structured lattice"):
        capi._check(capi.lib().gm_mg_build(ctx.h, 2), "gm_mg_build")
    ctx.close()
