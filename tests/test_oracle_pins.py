"""Pins of three oracle functions that the GPU-vs-oracle parity cannot pin by
itself (a mistake made in both would pass it): the Carreau-Yasuda viscosity
value, row-sum scaling, and the backtracking line search.  CPU only.

* Carreau-Yasuda mu(gammadot) (src/mm_viscosity.c:1065-1072, shear rate
  src/mm_fill_util.c:4564): a homogeneous simple shear u = G y, v = 0 lies in
  the Q2 space, has no advection and a uniform stress Pi_xy = mu(G) G, so the
  element residual is mu(G) times the Newtonian (mu = 1) one.  The measured
  ratio is checked at points where the law has closed values that do not need
  the formula typed out: (lam G)^a = 3 with a = 2, n = 1/2 gives
  (1 + 3)^(-1/4) = 1/sqrt(2); a = 1, n = 0, lam G = 1 gives 1/2; n = 1 is
  Newtonian for every G; at rest the Jacobian is mu0 times the Newtonian one.
* Row-sum scaling (src/sl_matrix_util.c:506-560): every scaled row has
  sum |a_ij| = 1, a nonzero diagonal ends positive, a (near-)zero diagonal
  row is divided by the plain positive sum, and b is divided by the same
  factor as its row.
* Backtracking (src/mm_sol_nonlinear.c:344-590): one Newton pass of the
  oracle with the line search takes the first of 1, 1/2, 1/4, ... (>= the
  minimum damping) whose scaled residual satisfies
  0.5 |R|^2 <= 0.5 |R0|^2 + 0.5 slope damp (slope = -|(J dx) . R(x - dx)|),
  or else the trial with the smallest norm; the trials are evaluated here
  with independent oracle assemblies and a direct solve.

Each pin is shown to fail on a planted mistake.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from goma_amd import problems
from goma_amd.mesh import structured_mesh
from goma_amd.problems import Physics, Problem


def _shear_problem(visc_model, mu0=1.0, muinf=0.0, lam=0.0, nexp=1.0, aexp=2.0, jitter=0.0):
    """One 2x2 patch of QUAD9 elements, no boundary conditions, no datum."""
    mesh = structured_mesh(2, 2, jitter=jitter)
    phys = Physics(energy=0, visc_model=visc_model, rho=1.0, mu0=mu0, muinf=muinf, lam=lam,
                   nexp=nexp, aexp=aexp)
    z = np.zeros(0, np.int32)
    return Problem(name="shear", mesh=mesh, phys=phys, dbc_node=z, dbc_var=z,
                   dbc_value=np.zeros(0), datum_elem=-1, datum_value=0.0)


def _shear_state(o, prob, G):
    """x with u = G y, v = 0, P = 0 at every node."""
    x = np.zeros(o.N)
    fu = o.first_unknown[: prob.mesh.num_nodes]
    x[fu + 0] = G * prob.mesh.coor[1]
    return x


def _measured_mu(oracle_mod, G, **law):
    """Effective viscosity from the ratio of the Carreau residual to the
    Newtonian (mu = 1) residual of the same simple shear."""
    pc = _shear_problem(1, **law)
    pn = _shear_problem(0, mu0=1.0)
    oc, on = oracle_mod.Oracle(pc), oracle_mod.Oracle(pn)
    x = _shear_state(oc, pc, G)
    rc, Rc, _ = oc.fill(x)
    rn, Rn, _ = on.fill(x)
    assert rc == 0 and rn == 0
    big = np.abs(Rn) > 1e-3 * np.abs(Rn).max()
    ratio = Rc[big] / Rn[big]
    # homogeneous shear: one viscosity everywhere
    assert np.ptp(ratio) <= 1e-12 * np.abs(ratio).max()
    return float(ratio.mean())


LAW = dict(mu0=1.0, muinf=0.01, lam=1.0, nexp=0.5, aexp=2.0)  # the Carreau-Yasuda fluid of config 4


def test_carreau_mu_closed_values(oracle_mod):
    # (lam G)^2 = 3: (1 + 3)^((0.5 - 1)/2) = 4^(-1/4) = 1/sqrt(2)
    mu = _measured_mu(oracle_mod, np.sqrt(3.0), **LAW)
    assert abs(mu - (0.01 + 0.99 / np.sqrt(2.0))) <= 1e-13
    # lam = 2, G = 1/4: (lam G)^2 = 1/4, 1.25^(-1/4)
    mu = _measured_mu(oracle_mod, 0.25, **dict(LAW, lam=2.0))
    assert abs(mu - (0.01 + 0.99 * 0.8 ** 0.25)) <= 1e-13
    # Yasuda a = 1, n = 0, lam G = 1: (1 + 1)^(-1) = 1/2
    mu = _measured_mu(oracle_mod, 0.5, **dict(LAW, lam=2.0, nexp=0.0, aexp=1.0))
    assert abs(mu - (0.01 + 0.99 * 0.5)) <= 1e-13
    # n = 1: Newtonian at every shear rate
    for G in (0.1, 3.0, 40.0):
        assert abs(_measured_mu(oracle_mod, G, **dict(LAW, nexp=1.0)) - 1.0) <= 1e-13


def test_carreau_at_rest_is_mu0(oracle_mod):
    """At gammadot = 0 the law gives mu0 and no viscosity sensitivity: the
    Jacobian equals the Newtonian one with mu = mu0."""
    pc = _shear_problem(1, **dict(LAW, mu0=2.5))
    pn = _shear_problem(0, mu0=2.5)
    oc, on = oracle_mod.Oracle(pc), oracle_mod.Oracle(pn)
    x = np.zeros(oc.N)
    _, _, vc = oc.fill(x)
    _, _, vn = on.fill(x)
    assert np.abs(vc - vn).max() <= 1e-14 * np.abs(vn).max()


def test_carreau_pin_catches_planted_exponent(oracle_mod):
    """A wrong power-law index in the oracle's law (n = 0.25 where 0.5 is
    meant) moves the measured value off the closed one."""
    mu = _measured_mu(oracle_mod, np.sqrt(3.0), **dict(LAW, nexp=0.25))
    assert abs(mu - (0.01 + 0.99 / np.sqrt(2.0))) > 1e-2


def _random_msr(rng, N, per_row):
    """A random MSR matrix with mixed-sign diagonal, one zero diagonal and
    one tiny (< 1e-200) diagonal."""
    cols, ija = [], [N + 1]
    for i in range(N):
        c = np.sort(rng.choice([j for j in range(N) if j != i], per_row, replace=False))
        cols.append(c)
        ija.append(ija[-1] + per_row)
    ija = np.array(ija + [int(c) for cc in cols for c in cc], np.int32)
    val = np.zeros(len(ija))
    val[:N] = rng.standard_normal(N) * 3
    val[3] = 0.0
    val[5] = 1e-250
    val[N + 1:] = rng.standard_normal(len(ija) - N - 1)
    return ija, val


def _check_scaled(ija, a0, b0, a, b, s, N):
    for i in range(N):
        row = np.concatenate([[a[i]], a[ija[i]: ija[i + 1]]])
        row0 = np.concatenate([[a0[i]], a0[ija[i]: ija[i + 1]]])
        if abs(np.abs(row).sum() - 1.0) > 1e-14:
            return f"row {i}: sum |a| = {np.abs(row).sum()}"
        if abs(a0[i]) > 1e-200 and not a[i] > 0:
            return f"row {i}: diagonal {a[i]} not positive"
        if abs(a0[i]) <= 1e-200 and not s[i] > 0:
            return f"row {i}: zero-diagonal row scaled by {s[i]}"
        if np.abs(row * s[i] - row0).max() > 1e-14 * np.abs(row0).max():
            return f"row {i}: a not divided by its scale"
        if abs(b[i] * s[i] - b0[i]) > 1e-14 * abs(b0[i]):
            return f"row {i}: b not divided by its row's scale"
    return None


def test_row_sum_scaling_invariants(oracle_mod):
    rng = np.random.default_rng(21)
    N = 40
    ija, a0 = _random_msr(rng, N, 6)
    b0 = rng.standard_normal(N)
    a, b, s = a0.copy(), b0.copy(), np.zeros(N)
    oracle_mod.lib().om_row_sum_scale(N, a, ija, b, s)
    assert _check_scaled(ija, a0, b0, a, b, s, N) is None
    # on an assembled Jacobian too (Dirichlet rows, pressure rows)
    prob = problems.natural_convection(n=3, Ra=1e4, carreau=True, jitter=0.15)
    o = oracle_mod.Oracle(prob)
    x = np.random.default_rng(2).standard_normal(o.N) * 0.4
    _, r0, v0 = o.fill(x)
    r, v, s = r0.copy(), v0.copy(), np.zeros(o.N)
    oracle_mod.lib().om_row_sum_scale(o.N, v, o.ija, r, s)
    assert _check_scaled(o.ija, v0, r0, v, r, s, o.N) is None


def test_row_sum_scaling_catches_planted_mistakes(oracle_mod):
    rng = np.random.default_rng(21)
    N = 40
    ija, a0 = _random_msr(rng, N, 6)
    b0 = rng.standard_normal(N)
    a, b, s = a0.copy(), b0.copy(), np.zeros(N)
    oracle_mod.lib().om_row_sum_scale(N, a, ija, b, s)
    neg = np.nonzero(a0[:N] < -1e-3)[0][0]
    # sign rule dropped: the scale of a negative-diagonal row loses its sign
    a1, b1, s1 = a.copy(), b.copy(), s.copy()
    a1[neg], b1[neg], s1[neg] = -a1[neg], -b1[neg], -s1[neg]
    a1[ija[neg]: ija[neg + 1]] *= -1
    assert "not positive" in _check_scaled(ija, a0, b0, a1, b1, s1, N)
    # b left unscaled
    b2 = b.copy()
    b2[7] = b0[7]
    assert "b not divided" in _check_scaled(ija, a0, b0, a, b2, s, N)


def _scaled_system(o, x):
    rc, r, v = o.fill(x)
    assert rc == 0
    s = np.zeros(o.N)
    import oracle

    oracle.lib().om_row_sum_scale(o.N, v, o.ija, r, s)
    return r, v, s


def _expected_backtrack(o, x0, min_damp=0.05, eps0=1e-10, factor=0.5):
    """The backtracking rule evaluated with independent assemblies: returns
    (damping, trial norms, Newton correction).  ``factor`` is the step
    reduction (0.5 in Goma; another value plants a mistake)."""
    import oracle

    R0, v, s = _scaled_system(o, x0)
    A = oracle.msr_to_scipy(o.ija, v, o.N)
    dx = spla.spsolve(A.tocsc(), R0)
    L2 = np.linalg.norm(R0)

    def g_of(damp):
        rc, R, _ = o.fill(x0 - damp * dx)
        return R / s

    R1 = g_of(1.0)
    slope = float((A @ dx) @ R1)
    slope = -slope if slope > 0 else (-1.0 if slope == 0 else slope)
    norms = {1.0: np.linalg.norm(R1)}
    best, best_norm = 1.0, norms[1.0]
    damp = 1.0
    if best_norm >= eps0:
        while True:
            damp *= factor
            if damp < min_damp:
                break
            g = np.linalg.norm(g_of(damp))
            norms[damp] = g
            if g < best_norm:
                best, best_norm = damp, g
            if best_norm < eps0:
                break
            if 0.5 * g * g <= 0.5 * L2 * L2 + 0.5 * slope * damp:
                break
    return best, norms, dx


@pytest.mark.parametrize("Ra,x_scale", [(1e5, 0.0), (1e4, 1.0)])
def test_backtracking_follows_armijo_rule(oracle_mod, Ra, x_scale):
    """From rest at Ra = 1e5 the full step overshoots and is cut back; from a
    far-off random state at Ra = 1e4 too.  The oracle's one-pass Newton with
    the line search lands where the independently evaluated rule says."""
    prob = problems.natural_convection(n=4, Ra=Ra)
    o = oracle_mod.Oracle(prob)
    m, vals = o.dbc_mask, o.dbc_val
    x0 = np.random.default_rng(9).standard_normal(o.N) * x_scale
    x0[m != 0] = vals[m != 0]
    best, norms, dx = _expected_backtrack(o, x0)
    assert len(norms) >= 2, "the case must exercise at least one step reduction"
    rc, x1, rep = o.newton(x0=x0, max_newton=1, lin_tol=1e-14, kspace=200, max_its=4000,
                           line_search=1)
    assert rep.damp[0] == best
    assert np.abs(x1 - (x0 - best * dx)).max() <= 1e-8 * np.abs(x1).max()


def test_backtracking_pin_catches_planted_reduction(oracle_mod):
    """With the step cut by 1/4 instead of 1/2 (a mistake planted in the rule)
    the expected damping differs from what the oracle takes."""
    prob = problems.natural_convection(n=4, Ra=1e4)
    o = oracle_mod.Oracle(prob)
    x0 = np.random.default_rng(9).standard_normal(o.N)
    x0[o.dbc_mask != 0] = o.dbc_val[o.dbc_mask != 0]
    rc, x1, rep = o.newton(x0=x0, max_newton=1, lin_tol=1e-14, kspace=200, max_its=4000,
                           line_search=1)
    best, _, _ = _expected_backtrack(o, x0)
    planted, _, _ = _expected_backtrack(o, x0, factor=0.25)
    assert rep.damp[0] == best
    assert rep.damp[0] != planted
