"""Energy SUPG (Goma's Ewt_funcModel = SUPG) on the GPU against the oracle.

The advection term of the energy rows is weighted by phi_i + s/h (v.grad
phi_i), with h from the element's face-centroid widths and the velocity at
its last local node (src/mm_fill_energy.c:237-259, 332-339, 640-660).  Bars:
residual and Jacobian within 1e-12 of the largest entry in FP64 (1e-6 for
FP32 values), with and without PSPG, over a NaN-filled array for the stored
form, the transient fill, and Newton solutions within 1e-9 of max|x|.
"""
from dataclasses import replace

import numpy as np
import pytest

from goma_amd import problems

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _supg(prob, s=0.7, pspg=0):
    return replace(prob, phys=replace(prob.phys, energy_supg=s, pspg=pspg, ps_scaling=0.5))


CASES = {
    "natconv2d_jitter": lambda: problems.natural_convection(n=4, Ra=1e4, jitter=0.15),
    "carreau2d_jitter_pspg": lambda: problems.natural_convection(n=4, Ra=1e4, carreau=True, jitter=0.15),
    "cavity3d_energy_jitter": lambda: problems.lid_cavity(n=2, Re=400.0, dim=3, energy=True, Ra=1e4,
                                                          jitter=0.1),
    "natconv3d_carreau_pspg": lambda: problems.natural_convection(n=2, Ra=1e4, dim=3, carreau=True,
                                                                  jitter=0.1),
}


@pytest.fixture(scope="module")
def capi():
    from goma_amd import capi as M

    M.lib()
    assert torch.cuda.is_available()
    return M


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _z(n):
    return torch.zeros(n, dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("name", list(CASES))
def test_supg_fill_matches_oracle(name, capi, oracle_mod):
    prob = _supg(CASES[name](), pspg=1 if name.endswith("pspg") else 0)
    print(f"[case] supg fill {name}", flush=True)
    ctx, orc = capi.Context(prob), oracle_mod.Oracle(prob)
    x = np.random.default_rng(9).standard_normal(orc.N) * 0.5
    u = orc.velocity_norm(x)
    orc.set_velocity_norm(u)
    ctx.set_velocity_norm(u)
    rc, r_ref, v_ref = orc.fill(x)
    assert rc == 0
    xd, rd, vd = _dev(x), _z(ctx.N), _z(ctx.nval)
    torch.cuda.synchronize()
    assert ctx.fill_dev(xd.data_ptr(), rd.data_ptr(), vd.data_ptr()) == 0
    assert np.abs(rd.cpu().numpy() - r_ref).max() <= 1e-12 * np.abs(r_ref).max()
    assert np.abs(vd.cpu().numpy() - v_ref).max() <= 1e-12 * np.abs(v_ref).max()
    # the weight is in the result
    _, r0, _ = oracle_mod.Oracle(replace(prob, phys=replace(prob.phys, energy_supg=0.0))).fill(x)
    assert np.abs(r_ref - r0).max() > 1e-6 * np.abs(r_ref).max()
    vn = torch.full((ctx.nval,), float("nan"), dtype=torch.float64, device="cuda")
    assert ctx.fill_assign_dev(xd.data_ptr(), _z(ctx.N).data_ptr(), vn.data_ptr()) == 0
    va = vn.cpu().numpy()
    va[ctx.N] = v_ref[ctx.N]
    assert np.isfinite(va).all()
    assert np.abs(va - v_ref).max() <= 1e-12 * np.abs(v_ref).max()
    v32 = torch.zeros(ctx.nval, dtype=torch.float32, device="cuda")
    assert ctx.fill32_dev(xd.data_ptr(), _z(ctx.N).data_ptr(), v32.data_ptr()) == 0
    assert np.abs(v32.cpu().numpy().astype(np.float64) - v_ref).max() <= 1e-6 * np.abs(v_ref).max()
    ctx.close()


@pytest.mark.parametrize("name", ["natconv2d_jitter", "cavity3d_energy_jitter"])
def test_supg_transient_fill_matches_oracle(name, capi, oracle_mod):
    prob = _supg(CASES[name]())
    print(f"[case] supg transient fill {name}", flush=True)
    ctx, orc = capi.Context(prob), oracle_mod.Oracle(prob)
    rng = np.random.default_rng(12)
    x, x_old, xdot_old = (rng.standard_normal(orc.N) * 0.5 for _ in range(3))
    dt, theta = 0.05, 0.5
    c = (1.0 + 2.0 * theta) / dt
    rc, r_ref, v_ref = orc.fill_transient(x, c * (x - x_old) - 2.0 * theta * xdot_old, c)
    assert rc == 0
    xd, xo, xdo, rd, vd = _dev(x), _dev(x_old), _dev(xdot_old), _z(ctx.N), _z(ctx.nval)
    torch.cuda.synchronize()
    assert ctx.fill_transient_dev(xd.data_ptr(), xo.data_ptr(), xdo.data_ptr(), dt, theta,
                                  rd.data_ptr(), vd.data_ptr()) == 0
    assert np.abs(rd.cpu().numpy() - r_ref).max() <= 1e-12 * np.abs(r_ref).max()
    assert np.abs(vd.cpu().numpy() - v_ref).max() <= 1e-12 * np.abs(v_ref).max()
    ctx.close()


def test_supg_newton_2d_matches_oracle(capi, oracle_mod):
    prob = _supg(problems.natural_convection(n=8, Ra=1e5), s=1.0)
    print("[case] supg newton natconv n=8", flush=True)
    ctx, orc = capi.Context(prob), oracle_mod.Oracle(prob)
    rc, x_ref, rep_ref = orc.newton(precond=1, kspace=60, max_its=1000, max_newton=30)
    assert rep_ref.converged
    xd = _z(ctx.N)
    torch.cuda.synchronize()
    rc, rep = ctx.solve_dev(xd.data_ptr(), kspace=60, max_its=1000, patch_bins=2, max_newton=30)
    assert rep.converged
    assert np.abs(xd.cpu().numpy() - x_ref).max() <= 1e-9 * np.abs(x_ref).max()
    ctx.close()


def test_supg_pspg_newton_3d_multigrid(capi, oracle_mod):
    prob = _supg(problems.lid_cavity(n=4, Re=400.0, dim=3, energy=True, Ra=1e4), s=1.0, pspg=1)
    print("[case] supg+pspg newton multigrid 3d", flush=True)
    ctx, orc = capi.Context(prob), oracle_mod.Oracle(prob)
    rc, x_ref, rep_ref = orc.newton(precond=1, kspace=100, max_its=2000, max_newton=30)
    assert rep_ref.converged
    assert ctx.attach_multigrid(2) == 2
    xd = _z(ctx.N)
    torch.cuda.synchronize()
    rc, rep = ctx.solve_dev(xd.data_ptr(), kspace=50, max_its=200, precond=2, max_newton=30,
                            line_search=1)
    assert rep.converged
    assert np.abs(xd.cpu().numpy() - x_ref).max() <= 1e-9 * np.abs(x_ref).max()
    ctx.close()
