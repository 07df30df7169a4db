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

First-order continuation (``order=1``, Goma's ALC_FIRST, ac_conti.c:766-773,
1035-1041) predicts x_{n+1}^0 = x_n + (Ra_{n+1} - Ra_n) dx/dRa, with the
sensitivity from J dx/dRa = -dR/dRa at the converged x_n, dR/dRa by central
differences of the residual (soln_sens, src/mm_sol_nonlinear.c:3638-3930).

Steps are uniform in log10 Ra, the natural spacing of the sweep.

``continue_problem`` is the adaptive form of the same loop: the path
parameter s = log10 Ra advances by a step that path_step_control
(src/rf_util.c:1312-1365) rescales after every step that converged at the
first try, from the number of Newton passes it took, between Delta_s_min and
Delta_s_max (ac_conti.c:1091-1100); a failed step is halved and retried
(ac_conti.c:968-1045).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import problems


def _dlagrange1(pos, s):
    if pos < 0:
        return -0.5 + s
    if pos > 0:
        return 0.5 + s
    return -2.0 * s


def cell_reynolds(prob, first_unknown, x):
    """Largest element Reynolds number rho |u| h / mu over the mesh, with u and
    the shear rate at each element centroid (Q2 fields, 2D) and mu from the
    problem's viscosity model (Carreau-Yasuda, src/mm_viscosity.c:1065-1072).
    Returns (max cell Re, min mu, max |u|)."""
    from .mesh import QUAD9_POS

    m, ph = prob.mesh, prob.phys
    if m.dim != 2:
        raise ValueError("2D meshes only")
    dphi = np.array([[_dlagrange1(p[0], 0.0) * (1.0 if p[1] == 0 else 0.0),
                      (1.0 if p[0] == 0 else 0.0) * _dlagrange1(p[1], 0.0)] for p in QUAD9_POS])
    fu = first_unknown[: m.num_nodes].astype(np.int64)
    conn = m.connect
    X = m.coor.T[conn]  # [ne, 9, 2]
    U = np.stack([x[fu[conn] + 0], x[fu[conn] + 1]], axis=-1)  # [ne, 9, 2]
    J = np.einsum("enk,nd->ekd", X, dphi)  # dx_k / dxi_d
    Jinv = np.linalg.inv(J)
    G = np.einsum("enk,nd,edj->ekj", U, dphi, Jinv)  # du_k / dx_j
    gam = G + np.transpose(G, (0, 2, 1))
    gdot = np.sqrt(0.5 * np.einsum("eij,eji->e", gam, gam))
    if ph.visc_model == 1:
        mu = ph.muinf + (ph.mu0 - ph.muinf) * (1 + (ph.lam * gdot) ** ph.aexp) ** ((ph.nexp - 1) / ph.aexp)
    else:
        mu = np.full(len(gdot), ph.mu0)
    uc = np.linalg.norm(U[:, 8], axis=1)
    h = np.sqrt(np.abs(np.linalg.det(J)) * 4.0)
    re = ph.rho * uc * h / mu
    return float(re.max()), float(mu.min()), float(uc.max())


@dataclass
class Step:
    Ra: float
    newton_its: int
    lin_its: int
    seconds: float
    converged: bool
    retries: int


def sensitivity(ctx, prob, x, Ra, opts):
    """dx/dRa at the converged state x (float64 CUDA tensor): J t = -dR/dRa on
    the row-sum-scaled system, solved with the solver options ``opts``."""
    import torch

    z = lambda n: torch.zeros(n, dtype=torch.float64, device="cuda")
    N, nv = ctx.N, ctx.nval
    r, rp, rm, sc, t = z(N), z(N), z(N), z(N), z(N)
    v = z(nv)
    dRa = 1e-3 * Ra
    for R, ra in ((rp, Ra + dRa), (rm, Ra - dRa)):
        ctx.set_physics(problems.physics_at(prob, ra))
        v.zero_()
        if ctx.fill_dev(x.data_ptr(), R.data_ptr(), v.data_ptr()) != 0:
            raise RuntimeError("assembly failed in the sensitivity")
    ctx.set_physics(problems.physics_at(prob, Ra))
    v.zero_()
    if ctx.fill_dev(x.data_ptr(), r.data_ptr(), v.data_ptr()) != 0:
        raise RuntimeError("assembly failed in the sensitivity")
    ctx.row_sum_scale_dev(v.data_ptr(), r.data_ptr(), sc.data_ptr())
    b = (rp - rm) / (2.0 * dRa) / sc
    lin = {k: opts[k] for k in ("precond", "precision", "kspace", "patch_bins", "mg_nu", "mg_omega",
                                "mg_smoother", "mg_galerkin") if k in opts}
    if lin.get("precision", 0) == 1:
        lin["precision"] = 0  # the public GMRES entry point takes FP64 values
    ctx.gmres_dev(v.data_ptr(), b.data_ptr(), t.data_ptr(), max_its=opts.get("max_its", 200),
                  lin_tol=1e-8, **lin)
    return -t


def _solve(ctx, prob, x, Ra, opts):
    ctx.set_physics(problems.physics_at(prob, Ra))
    t0 = time.perf_counter()
    _, rep = ctx.solve_dev(x.data_ptr(), **opts)
    return rep, time.perf_counter() - t0


def sweep(ctx, prob, ras, opts: dict, x=None, max_retries: int = 5, keep: bool = True,
          log=None, order: int = 0):
    """Continuation through the values ``ras`` (increasing or decreasing).

    Returns (steps, solutions): one Step per requested value and, with
    ``keep``, the converged solutions as host float64 arrays.  ``x`` (a
    float64 CUDA tensor of ctx.N entries, zero if omitted) is the predictor
    of the first step and holds the last solution on return.  ``order`` 1
    uses the first-order predictor from the second step on.
    """
    import torch

    if x is None:
        x = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    steps, sols = [], []
    last_ok = x.clone()
    lam_prev = None
    tan = None  # dx/dRa at last_ok
    for Ra in ras:
        target = float(Ra)
        retries = 0
        its = lin = 0
        secs = 0.0
        lam = target
        while True:
            x.copy_(last_ok)
            if tan is not None:
                x.add_(tan, alpha=lam - lam_prev)
            rep, dt = _solve(ctx, prob, x, lam, opts)
            its += rep.newton_its
            lin += rep.lin_its_total
            secs += dt
            if rep.converged:
                last_ok.copy_(x)
                lam_prev = lam
                if order == 1:
                    t0 = time.perf_counter()
                    tan = sensitivity(ctx, prob, last_ok, lam, opts)
                    secs += time.perf_counter() - t0
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


def path_step_control(delta_s_old: float, inewton: int, max_newton: int) -> float:
    """Goma's path_step_control (src/rf_util.c:1312-1365): the next step from
    the last one and the Newton passes it needed, aiming at half of
    Max_Newton_Steps.  beta = (2^(d - inewton) - 1) / d with d = max_newton // 2;
    beta > 1 gives the largest growth, 3; otherwise the factor is 10^beta,
    clipped to [0.5, 3]."""
    d = max_newton // 2
    if d < 1:
        return delta_s_old
    beta = (2.0 ** float(d - inewton) - 1.0) / d
    beta = 3.0 if beta > 1.0 else 10.0 ** beta
    return min(max(beta, 0.5), 3.0) * delta_s_old


@dataclass
class PathStep:
    Ra: float
    delta_s: float      # log10 Ra increment of the accepted step
    newton_its: int
    tries: int          # attempts at this step (1: converged at once)
    seconds: float


def continue_problem(ctx, prob, ra_start: float, ra_end: float, delta_s: float, opts: dict,
                     delta_s_min: float = 1e-3, delta_s_max: float = 0.5, order: int = 1,
                     x=None, max_steps: int = 200, log=None):
    """Adaptive natural-parameter continuation in s = log10 Ra from ra_start
    to ra_end (continue_problem, src/ac_conti.c:106-1100, steady branch).

    The first solve is at ra_start from ``x`` (zero if omitted).  Each step
    predicts the next state (zeroth order: the last solution; first order:
    plus the sensitivity times the increment), solves, and on failure halves
    the step and retries from the last solution, giving up after five tries
    or below delta_s_min.  After a step that converged at its first try, the
    next increment is path_step_control's, capped at delta_s_max.  Returns
    (accepted steps, converged: bool); ``x`` holds the last solution.
    """
    import torch

    if x is None:
        x = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    s, s_end = float(np.log10(ra_start)), float(np.log10(ra_end))
    sign = 1.0 if s_end >= s else -1.0
    max_newton = int(opts.get("max_newton", 25))
    rep, dt = _solve(ctx, prob, x, ra_start, opts)
    steps = [PathStep(ra_start, 0.0, rep.newton_its, 1, dt)]
    if not rep.converged:
        return steps, False
    last = x.clone()
    tan = sensitivity(ctx, prob, last, ra_start, opts) if order == 1 else None
    ds = abs(delta_s)
    n = 0
    while sign * (s_end - s) > 1e-12 and n < max_steps:
        ds = min(ds, abs(s_end - s))
        tries, secs = 0, 0.0
        while True:
            tries += 1
            s1 = s + sign * ds
            if abs(s1 - s_end) < 1e-12:
                s1 = s_end
            ra0, ra1 = 10.0 ** s, 10.0 ** s1
            x.copy_(last)
            if tan is not None:
                x.add_(tan, alpha=ra1 - ra0)
            rep, dt = _solve(ctx, prob, x, ra1, opts)
            secs += dt
            if rep.converged:
                break
            if tries > 5:
                return steps, False
            ds *= 0.5
            if log:
                log(f"[conti] Ra={ra1:.4g}: Newton failed, step {ds:.4g} in log10 Ra")
            if ds < delta_s_min:
                return steps, False
        steps.append(PathStep(ra1, ds, rep.newton_its, tries, secs))
        if log:
            log(f"[conti] Ra={ra1:.4g} step {ds:.4g} newton {rep.newton_its} tries {tries}")
        last.copy_(x)
        if order == 1:
            t0 = time.perf_counter()
            tan = sensitivity(ctx, prob, last, ra1, opts)
            steps[-1].seconds += time.perf_counter() - t0
        # path_step_control after a step that converged at once (ac_conti.c:1091-1100)
        if tries == 1 and n != 0:
            ds = min(path_step_control(ds, rep.newton_its, max_newton), delta_s_max)
        s = s1
        n += 1
    return steps, abs(s - s_end) < 1e-12
