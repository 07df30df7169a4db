"""The oracle's restatement of soln_sens (src/mm_sol_nonlinear.c:3638-3930),
pinned on the CPU: dx/dRa = -J^-1 dR/dRa at a converged state equals the
central difference of the converged solutions at Ra -+ 1 (the implicit
function theorem), and the two differ by far more when the sign or the
parameter step is wrong.
"""
from dataclasses import replace

import numpy as np

from goma_amd import problems


def test_sensitivity_is_the_derivative_of_the_solution(oracle_mod):
    prob = problems.natural_convection(n=4, Ra=1e4, jitter=0.1)
    at = lambda ra: replace(prob, phys=problems.physics_at(prob, ra))
    Ra = 1e4
    rc, x, rep = oracle_mod.Oracle(at(Ra)).newton(precond=1, kspace=50, max_its=1000,
                                                  line_search=1, eps_resid=1e-12)
    assert rep.converged
    xs = []
    for ra in (Ra - 1.0, Ra + 1.0):
        rc, xr, rp = oracle_mod.Oracle(at(ra)).newton(x0=x, precond=1, kspace=50, max_its=1000,
                                                     line_search=1, eps_resid=1e-12)
        assert rp.converged
        xs.append(xr)
    fd = (xs[1] - xs[0]) / 2.0
    t = oracle_mod.sensitivity(at, x, Ra, 1e-3 * Ra)
    scale = np.abs(fd).max()
    assert np.abs(t - fd).max() <= 1e-7 * scale
    assert np.abs(-t - fd).max() > 0.5 * scale
    # a parameter step that is not the one differenced scales the answer
    t2 = oracle_mod.sensitivity(lambda ra: at(Ra + 2.0 * (ra - Ra)), x, Ra, 1e-3 * Ra)
    assert np.abs(t2 - fd).max() > 0.5 * scale
