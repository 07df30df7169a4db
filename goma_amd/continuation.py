"""Zeroth-order natural-parameter continuation over the Rayleigh number.

Restates the steady branch of Goma's continue_problem (src/ac_conti.c:106-1100)
for one parameter:

* predictor of a zeroth-order step: the previous converged solution,
  x_{n+1}^0 = x_n (ac_conti.c:760-787);
* corrector: Newton at the new parameter value (solve_nonlinear_problem,
  ac_conti.c:856), here gm_solve_nonlinear_dev after gm_set_physics
  (update_parameterC, src/ac_update_parameter.c:48);
* a step whose Newton fails is retried from the last converged solution with
  half the parameter increment, up to five times (ac_conti.c:968-1045).

Steps are uniform in log10 Ra, the natural spacing of the sweep.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import problems


@dataclass
class Step:
    Ra: float
    newton_its: int
    lin_its: int
    seconds: float
    converged: bool
    retries: int


def _solve(ctx, prob, x, Ra, opts):
    ctx.set_physics(problems.physics_at(prob, Ra))
    t0 = time.perf_counter()
    _, rep = ctx.solve_dev(x.data_ptr(), **opts)
    return rep, time.perf_counter() - t0


def sweep(ctx, prob, ras, opts: dict, x=None, max_retries: int = 5, keep: bool = True,
          log=None):
    """Continuation through the values ``ras`` (increasing or decreasing).

    Returns (steps, solutions): one Step per requested value and, with
    ``keep``, the converged solutions as host float64 arrays.  ``x`` (a
    float64 CUDA tensor of ctx.N entries, zero if omitted) is the predictor
    of the first step and holds the last solution on return.
    """
    import torch

    if x is None:
        x = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    steps, sols = [], []
    last_ok = x.clone()
    lam_prev = None
    for Ra in ras:
        target = float(Ra)
        retries = 0
        its = lin = 0
        secs = 0.0
        lam = target
        while True:
            x.copy_(last_ok)
            rep, dt = _solve(ctx, prob, x, lam, opts)
            its += rep.newton_its
            lin += rep.lin_its_total
            secs += dt
            if rep.converged:
                last_ok.copy_(x)
                lam_prev = lam
                if lam == target:
                    break
                lam = target  # reduced step succeeded: continue to the requested value
                continue
            if lam_prev is None or retries >= max_retries:
                break
            retries += 1
            # half the increment, in log10 Ra
            lam = float(10 ** (0.5 * (np.log10(lam) + np.log10(lam_prev))))
            if log:
                log(f"[conti] Ra={target:.4g}: Newton failed, retrying at {lam:.4g}")
        ok = bool(rep.converged) and lam == target
        steps.append(Step(target, its, lin, secs, ok, retries))
        if log:
            log(f"[conti] Ra={target:.4g} newton {its} gmres {lin} {secs:.3f}s"
                + ("" if ok else " FAILED"))
        if keep:
            sols.append(x.cpu().numpy().copy())
        if not ok:
            break
    return steps, sols
