"""Known answers from the literature: they pin the discretisation (assembly,
boundary conditions, Newton) against independent solutions, not bits.

* Lid-driven cavity, Re = 100 (Ghia, Ghia & Shin 1982, J. Comput. Phys. 48,
  Table I; u along the vertical centreline).  Ghia's 129^2 vorticity-stream
  function values carry ~1e-2 relative error near the minimum; spectral and
  high-resolution solutions give u_min = -0.2140 (Botella & Peyret 1998).
* Differentially heated cavity, Ra = 1e5, Pr = 0.71 (de Vahl Davis 1983,
  Int. J. Numer. Meth. Fluids 3): mean hot-wall Nusselt number 4.519,
  u_max = 34.73 on x = 1/2, v_max = 68.59 on y = 1/2 (alpha/L units).
"""
import numpy as np
import pytest

from goma_amd import problems
from oracle import Oracle, shape

GHIA_Y = [0.0547, 0.0625, 0.0703, 0.1016, 0.1719, 0.2813, 0.4531, 0.5, 0.6172, 0.7344, 0.8516,
          0.9531, 0.9609, 0.9688, 0.9766]
GHIA_U = [-0.03717, -0.04192, -0.04775, -0.06434, -0.10150, -0.15662, -0.21090, -0.20581,
          -0.13641, 0.00332, 0.23151, 0.68717, 0.73722, 0.78871, 0.84123]


def _centreline(prob, fu, x, axis, comp):
    """Nodal values of velocity component comp on the line where coordinate
    `axis` is 1/2, sorted along the other coordinate."""
    lat = prob.mesh.lattice()
    sel = np.nonzero(lat[:, axis] == prob.mesh.n)[0]
    s = prob.mesh.coor[1 - axis, sel]
    o = np.argsort(s)
    return s[o], x[fu[sel] + comp][o]


def test_oracle_cavity_re100_matches_ghia():
    prob = problems.lid_cavity(n=32, Re=100.0)
    o = Oracle(prob)
    rc, x, rep = o.newton(precond=3, kspace=50, max_its=500, lin_tol=1e-8, max_newton=20)
    assert rep.converged
    y, u = _centreline(prob, o.first_unknown, x, 0, 0)
    ui = np.interp(GHIA_Y, y, u)
    assert np.max(np.abs(ui - GHIA_U)) < 0.01  # Ghia's own accuracy
    assert abs(u.min() - (-0.2140)) < 0.3e-3
    assert abs(y[np.argmin(u)] - 0.4531) < 0.02


def nusselt_hot_wall(prob, x, fu):
    """Mean of -dT/dx over x = 0 through the Q2 field of the first element column."""
    m = prob.mesh
    n = m.n
    g = np.sqrt(0.6)
    pts, wts = [-g, 0.0, g], [5 / 9, 8 / 9, 5 / 9]
    total = 0.0
    for iy in range(n):
        e = iy * n
        conn = m.connect[e]
        X = m.coor[:, conn].T  # [9, 2]
        T = x[fu[conn] + 2]
        for t, w in zip(pts, wts):
            _, dphi = shape(2, -1.0, t)
            J = X.T @ dphi  # J[i][j] = dx_i / dxi_j
            gT = np.linalg.solve(J.T, dphi.T @ T)
            total += -gT[0] * w * J[1, 1]
    return total


@pytest.mark.gpu
def test_natural_convection_ra1e5_matches_de_vahl_davis():
    torch = pytest.importorskip("torch")
    from goma_amd import capi
    print("[case] de Vahl Davis Ra=1e5 n=64", flush=True)
    prob = problems.natural_convection(n=64, Ra=1e5, Pr=0.71)
    ctx = capi.Context(prob)
    ctx.attach_multigrid(2)
    x = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    _, rep = ctx.solve_dev(x.data_ptr(), precond=capi.GM_PC_MG, line_search=1, max_newton=30)
    assert rep.converged
    xh = x.cpu().numpy()
    fu = ctx.first_unknown()
    _, u = _centreline(prob, fu, xh, 0, 0)
    _, v = _centreline(prob, fu, xh, 1, 1)
    nu = nusselt_hot_wall(prob, xh, fu)
    print(f"  Nu {nu:.4f}  u_max {u.max():.3f}  v_max {v.max():.3f}", flush=True)
    assert abs(nu - 4.519) < 0.01 * 4.519
    assert abs(u.max() - 34.73) < 0.01 * 34.73
    assert abs(v.max() - 68.59) < 0.01 * 68.59
    ctx.close()


def test_nusselt_of_conduction_profile_is_one():
    """T = 1/2 - x (pure conduction) has -dT/dx = 1 on the hot wall: Nu = 1,
    also on a jittered mesh (the Q2 map reproduces linear fields)."""
    prob = problems.natural_convection(n=6, Ra=1e3, jitter=0.15)
    o = Oracle(prob)
    fu = o.first_unknown
    x = np.zeros(o.N)
    nn = prob.mesh.num_nodes
    x[fu[:nn] + 2] = 0.5 - prob.mesh.coor[0]
    assert abs(nusselt_hot_wall(prob, x, fu) - 1.0) < 1e-12
