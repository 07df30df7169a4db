#!/usr/bin/env python3
"""Benchmark: seconds per converged steady Newton solve on one MI355X.

A step is one converged Newton solve (||D^-1 R||_2 < 1e-10 and
||D^-1 R||_oo < 1e-10, Goma's test) from the zero initial guess, with the
mesh, matrix graph and all work arrays already resident in HBM.  The default
workload is 2D Boussinesq natural convection on a 256x256 QUAD9 mesh
(Q2/P1 velocity-pressure plus Q2 temperature, Pr = 0.71, Ra = 1e5), FP64.

Prints one JSON line.  Besides the headline value (wall time of the timed
solves, no instrumentation) it reports, for the kernel that takes the most
time, its algorithmic bytes per launch over its average launch duration as a
fraction of the MI355X HBM peak.  Durations come from a repeat of the timed
solve in which every launch carries HIP events stamped by its own dispatch on
the library's stream.  It also reports the mixed-precision solve of the same
workload with its nodal error against the FP64 solution, and the oracle's
FP64 CPU path timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6   # MI355X FP64 (vector and matrix)
FP32_PEAK_TFS = 157.3  # MI355X FP32 vector
METRIC = ("s per converged Newton solve (‖R‖<1e-10) vs plain FP64 GPU; "
          "assembly HBM GB/s")


def make_problem(name: str, n: int | None):
    from goma_amd import problems

    if name == "natconv2d":
        return problems.natural_convection(n=n or 256, Ra=1e5, Pr=0.71), dict(Ra=1e5, Pr=0.71)
    if name == "cavity2d":
        return problems.lid_cavity(n=n or 32, Re=100.0), dict(Re=100.0)
    if name == "cavity3d":
        return (problems.lid_cavity(n=n or 48, Re=400.0, dim=3, energy=True, Ra=1e4, Pr=0.71),
                dict(Re=400.0, Ra=1e4, Pr=0.71))
    if name == "carreau2d":
        return (problems.natural_convection(n=n or 128, Ra=1e3, carreau=True, jitter=0.15),
                dict(Ra=1e3, jitter=0.15))
    raise SystemExit(f"unknown workload {name}")


def algorithmic_bytes(ctx, prob, prec: int):
    """Compulsory HBM bytes per launch family (counted once)."""
    m = prob.mesh
    N, nval = ctx.N, ctx.nval
    nnz = nval - 1  # diagonal + off-diagonals
    sv = 4 if prec == 1 else 8
    d = m.dim
    nve = m.nodes_per_elem
    # assembly: write J once, write R and read x in FP64, coordinates, connectivity
    b_asm = nnz * sv + 16 * N + 8 * d * m.num_nodes + 4 * nve * m.num_elems
    # SpMV y = A x: values + column ids, row pointers, x and y
    b_spmv = nnz * (sv + 4) + 4 * (N + 1) + N * (sv + sv)
    # one ILU sweep (L or U half): its values and column ids, row pointers,
    # diagonal pointers, permutation, input and output vectors
    half = (nnz - N) / 2
    b_sweep_L = half * (sv + 4) + 4 * (N + 1) + 4 * N + 4 * N + N * (sv + sv)
    b_sweep_U = (half + N) * (sv + 4) + 4 * (N + 1) + 4 * N + 4 * N + N * (sv + sv)
    return dict(asm=b_asm, spmv=b_spmv, ilu_lower=b_sweep_L, ilu_upper=b_sweep_U)


def cpu_baseline(prob, gpu_counts: dict, budget_s: float):
    """The oracle -- a single-threaded FP64 restatement of Goma's CPU path
    (element loop with in_list scatter, GMRES(50) + natural-order ILU(0)) --
    timed on a bounded sample of the same workload.

    Times one assembly over a prefix of the elements (scaled to the mesh), one
    ILU(0) factorisation and ten GMRES iterations on the full system at the
    zero guess, then scales each by the counts of the GPU solve.  The GPU's
    multigrid needs far fewer GMRES iterations than ILU(0) does, so the
    figure is a lower bound on the CPU path's time.
    """
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C

    import oracle as O

    O.build()
    o = O.Oracle(prob)
    ne = prob.mesh.num_elems
    x = np.zeros(o.N)
    e1 = min(ne, 4096)
    t0 = time.perf_counter()
    rc, r, v = o.fill(x, 0, e1)
    t_asm_elem = (time.perf_counter() - t0) / e1
    t0 = time.perf_counter()
    rc, r, v = o.fill(x)
    t_full = time.perf_counter() - t0
    s = np.zeros(o.N)
    O.lib().om_row_sum_scale(o.N, v, o.ija, r, s)
    rr = C.c_double()
    t0 = time.perf_counter()
    its1 = O.lib().om_gmres_msr(o.N, v, o.ija, r, np.zeros(o.N), o.h, 1, 50, 1, 1e-30, C.byref(rr))
    t_one = time.perf_counter() - t0
    t0 = time.perf_counter()
    its = O.lib().om_gmres_msr(o.N, v, o.ija, r, np.zeros(o.N), o.h, 1, 50, 11, 1e-30, C.byref(rr))
    t_gm = time.perf_counter() - t0
    t_it = max((t_gm - t_one) / max(its - its1, 1), 1e-9)
    t_fact = max(t_one - t_it, 0.0)
    nasm = gpu_counts["assemblies"]
    nres = gpu_counts["residual_fills"]
    est = ((nasm + nres) * t_asm_elem * ne + gpu_counts["factorisations"] * t_fact
           + gpu_counts["gmres_its"] * t_it)
    return dict(value=round(est, 2), unit="s per converged solve (extrapolated lower bound)",
                cores=1, kind="port",
                sample=(f"oracle FP64 C, 1 thread: assembly of {e1} of {ne} elements "
                        f"({t_asm_elem * 1e6:.1f} us/element; full assembly {t_full:.2f} s), one "
                        f"natural-order ILU(0) factorisation ({t_fact:.2f} s) and {its - its1} "
                        f"GMRES(50) iterations ({t_it * 1e3:.1f} ms each) on the full "
                        f"{o.N}-unknown system at the zero guess; scaled by the GPU solve's "
                        f"{nasm} Jacobian and {nres} residual assemblies, "
                        f"{gpu_counts['factorisations']} factorisations and "
                        f"{gpu_counts['gmres_its']} GMRES iterations (ILU(0) needs more)"),
                assembly_elems_per_s=round(1.0 / t_asm_elem, 1), gmres_it_s=t_it,
                ilu_factor_s=t_fact)


def recorded_traffic(workload: str, kernel: str):
    """HBM bytes per launch of a kernel on a workload, measured with
    rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (calibrated on a device copy of
    known size) and recorded in profiles/traffic.txt; None if not recorded."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and not line.startswith("#") and f[0] == workload and f[1] == kernel:
            return float(f[2])
    return None


def field_groups(ctx, prob):
    """Unknown indices of the velocity, temperature and pressure fields."""
    from goma_amd.mesh import CENTROID_LOCAL

    m = prob.mesh
    d, en = m.dim, int(prob.phys.energy)
    fu = ctx.first_unknown()[: m.num_nodes].astype(np.int64)
    groups = {"velocity": np.concatenate([fu + k for k in range(d)])}
    if en:
        groups["temperature"] = fu + d
    cen = fu[m.connect[:, CENTROID_LOCAL[d]]] + d + en
    groups["pressure"] = np.concatenate([cen + k for k in range(d + 1)])
    return groups


def mixed_variant(ctx, prob, x_fp64, opts, torch):
    """The mixed-precision solve of the same workload at the same tolerance,
    and its nodal error per field against the FP64 solution."""
    ref = x_fp64.cpu().numpy()
    xm = torch.zeros_like(x_fp64)
    o = dict(opts)
    o["precision"] = 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc, rep = ctx.solve_dev(xm.data_ptr(), **o)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    xh = xm.cpu().numpy()
    err = {k: float(np.max(np.abs(xh[i] - ref[i])) / max(np.max(np.abs(ref[i])), 1e-300))
           for k, i in field_groups(ctx, prob).items()}
    print(f"[bench] mixed precision: newton {rep.newton_its} gmres {rep.lin_its_total} {dt:.2f}s "
          f"nodal error {err}", file=sys.stderr, flush=True)
    return {"fp64_s": None, "mixed_s": round(dt, 4), "mixed_converged": bool(rep.converged),
            "mixed_newton_its": rep.newton_its, "mixed_gmres_its": rep.lin_its_total,
            "mixed_nodal_error_vs_fp64": err}


def mlp_flops_per_point(n_in, n_out, hidden, width):
    """Algorithmic (unpadded) forward flops of the network per point."""
    dims = [n_in] + [width] * hidden + [n_out]
    return 2 * sum(a * b for a, b in zip(dims[:-1], dims[1:]))


def sweep_bench(args, torch, capi):
    """Carreau-Yasuda Ra continuation and the learned initial guess.

    Sweep: zeroth-order continuation over 64 log-spaced Ra in [1e3, 1e6] on
    the 128^2 jittered mesh (previous-step guess, FP64 Newton to 1e-10, each
    Newton system solved by the banded LU, or by multigrid-GMRES).
    Network: trained on the GPU on the converged sweep solutions.
    Evaluation: at up to 32 held-out geometric midpoints of sweep intervals
    spread over the converged range, Newton from the previous-step guess (the
    solution at the interval's lower end) against Newton from the learned
    guess (delta form from the same solution, with fallback), same tolerance.
    """
    from goma_amd import continuation, learned, problems

    n = args.n or 128
    gal = args.mg_galerkin
    band = args.sweep_precond == "band"
    if not band and not gal:
        # the shear-thinning flow defeats deep re-discretised hierarchies: stop at 16^2 (solved exactly)
        args.mg_coarse = max(args.mg_coarse, 16)
    prob = problems.natural_convection(n=n, Ra=1e3, carreau=True, jitter=0.15)
    ctx = capi.Context(prob)
    nlev = 0 if band else ctx.attach_multigrid(args.mg_coarse)
    prec = 1 if args.precision == "mixed" else 0
    if band:
        # element Reynolds numbers reach the hundreds: each Newton system is solved directly
        opts = dict(kspace=20, max_its=40, precision=prec, max_newton=15, precond=3,
                    line_search=args.line_search, pc_reuse=args.pc_reuse)
    else:
        # shear-thinning flow weakens the V(1,1) cycle as Ra grows: V(2,2), larger Krylov space
        opts = dict(kspace=max(args.kspace, 100), max_its=max(args.max_its, 400), precision=prec,
                    max_newton=25, precond=2, mg_nu=max(args.mg_nu, 2), mg_omega=args.mg_omega,
                    line_search=args.line_search, pc_reuse=args.pc_reuse, mg_smoother=0,
                    mg_galerkin=gal)
    log = lambda s: print(s, file=sys.stderr, flush=True)
    ras = problems.sweep_ra(args.sweep_count, 1e3, 1e6)
    log(f"[sweep] {prob.name}: N={ctx.N} " + ("banded LU" if band else f"multigrid, {nlev} levels")
        + f", {len(ras)} Ra values")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps, sols = continuation.sweep(ctx, prob, ras, opts, log=log if args.verbose else None,
                                     order=args.sweep_order)
    torch.cuda.synchronize()
    t_sweep = time.perf_counter() - t0
    ok = 0
    while ok < len(steps) and steps[ok].converged:
        ok += 1
    fu_h = ctx.first_unknown()
    diag = []
    for r, xs in zip(ras[:ok], sols[:ok]):
        re, mu_min, umax = continuation.cell_reynolds(prob, fu_h, xs)
        diag.append(dict(Ra=round(float(r), 1), cell_Re=round(re, 2), mu_min=round(mu_min, 4),
                         u_max=round(umax, 2)))
    if diag:
        log(f"[sweep] last converged step: {diag[-1]}")
    log(f"[sweep] {ok}/{len(ras)} steps converged, {t_sweep:.1f}s, "
        f"{sum(s.newton_its for s in steps)} Newton passes")
    ras_ok, sols_ok = ras[:ok], sols[:ok]
    lg = learned.LearnedGuess(prob, ctx, ras_ok, sols_ok, hidden=4, width=256, seed=20261202)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = lg.train(epochs=args.epochs, batch=args.batch, lr=2e-3, lr_final=2e-5, seed=1,
                    verbose=1 if args.verbose else 0)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    fpp = mlp_flops_per_point(3, 4, 4, 256)
    train_flops = 3.0 * fpp * lg.num_points * args.epochs  # forward + backward (2x forward)
    log(f"[sweep] trained on {lg.num_points} points x {args.epochs} epochs in {t_train:.1f}s "
        f"({train_flops / t_train / 1e12:.1f} TFLOP/s), loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    # the delta-form guess is projected onto the span of nearby sweep solutions: a wide sweep's
    # full snapshot span filters little of the network's mesh-scale noise
    half = args.basis_half
    log(f"[sweep] snapshot subspace from the {2 * half} sweep solutions nearest each held-out value")
    nb = 0
    rows = []
    # held-out values: geometric midpoints of sweep intervals, spread over the converged range
    picks = np.unique(np.linspace(0, ok - 2, min(args.held_out, ok - 1)).round().astype(int)) if ok > 1 else []
    for k in picks:
        Ra = float(np.sqrt(ras_ok[k] * ras_ok[k + 1]))
        ctx.set_physics(problems.physics_at(prob, Ra))
        xprev = torch.tensor(sols_ok[k], device="cuda")
        x = xprev.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, rb = ctx.solve_dev(x.data_ptr(), **opts)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        nb = lg.set_basis(sols_ok[max(0, k - half + 1): min(ok, k + half + 1)])
        x = torch.empty_like(xprev)
        t0 = time.perf_counter()
        lg.guess(x.data_ptr(), Ra, float(ras_ok[k]), xprev.data_ptr())
        _, rl = ctx.solve_guess_dev(x.data_ptr(), xprev.data_ptr(), 3, 10.0, **opts)
        torch.cuda.synchronize()
        tl = time.perf_counter() - t0
        rows.append(dict(Ra=Ra, prev_its=rb.newton_its, prev_s=tb, prev_ok=bool(rb.converged),
                         prev_r0=rb.resid_L2[0], learned_its=rl.newton_its + rl.discarded_its,
                         learned_s=tl, learned_ok=bool(rl.converged), fell_back=bool(rl.fell_back),
                         learned_r0=rl.resid_L2[0]))
        if args.verbose:
            log(f"[eval] Ra={Ra:.4g} previous-step {rb.newton_its} its {tb:.3f}s r0 {rb.resid_L2[0]:.2e}"
                f" | learned {rl.newton_its}+{rl.discarded_its} its {tl:.3f}s "
                f"r0 {rl.resid_L2[0]:.2e}{' FELL BACK' if rl.fell_back else ''}")
    prev_s = sum(r["prev_s"] for r in rows)
    learn_s = sum(r["learned_s"] for r in rows)
    out = {
        "metric": "s per held-out converged Newton solve (‖R‖<1e-10): learned vs previous-step guess",
        "value": round(learn_s / max(len(rows), 1), 4), "unit": "s", "n_gpus": 1,
        "higher_is_better": False, "dtype": "f64" if prec == 0 else "f32/f64",
        "data": "synthetic",
        "config": {"workload": f"{prob.name} Carreau-Yasuda natural convection, jitter 0.15h, "
                               f"Ra sweep [1e3, 1e6] x {len(ras)}", "unknowns": ctx.N,
                   "held_out": len(rows), "network": "3-256-256-256-256-4 tanh, BF16 MFMA",
                   "linear_solver": "banded LU (direct)" if band else "GMRES + multigrid",
                   "predictor": ("zeroth", "first")[args.sweep_order] + "-order continuation"},
        "sweep": {"converged_steps": ok, "seconds": round(t_sweep, 2), "cell_reynolds": diag[::4] + diag[-1:],
                  "newton_its": sum(s.newton_its for s in steps),
                  "retries": sum(s.retries for s in steps)},
        "subspace_dim": nb, "subspace_snapshots": 2 * half,
        "training": {"points": lg.num_points, "epochs": args.epochs, "seconds": round(t_train, 2),
                     "algorithmic_TFLOPs": round(train_flops / t_train / 1e12, 2),
                     "bf16_peak_TFLOPs": 2500.0, "loss_first": hist[0], "loss_last": hist[-1]},
        "previous_step": {"newton_its": sum(r["prev_its"] for r in rows),
                          "seconds": round(prev_s, 3),
                          "all_converged": all(r["prev_ok"] for r in rows)},
        "learned": {"newton_its": sum(r["learned_its"] for r in rows), "seconds": round(learn_s, 3),
                    "fallbacks": sum(r["fell_back"] for r in rows),
                    "all_converged": all(r["learned_ok"] for r in rows)},
        "speedup": round(prev_s / learn_s, 3) if learn_s > 0 else None,
        "per_ra": rows if args.verbose else None,
    }
    print(json.dumps(out), flush=True)
    lg.close()
    ctx.close()
    return 0


def learned_bench(args, torch, capi):
    """Mixed precision plus the learned initial guess against the plain FP64
    path, on one target Rayleigh number.

    Plain path: FP64 Newton from the zero guess at the target Ra.  Learned
    path: a continuation sweep over Ra values a quarter decade apart, offset
    so that the target is the geometric midpoint of two of them (the target is
    held out), trains the network on the GPU; at the target, Newton in mixed
    precision starts from the learned guess (delta form from the sweep
    solution just below, projected onto the sweep's subspace) with the
    previous-step guess as fallback.  The sweep and the training are reported
    separately: they are done once per parameter study, not per solve.
    """
    from goma_amd import continuation, learned, problems

    log = lambda s: print(s, file=sys.stderr, flush=True)
    prob, params = make_problem(args.learned_base, args.n)
    target = float(params["Ra"])
    ctx = capi.Context(prob)
    nlev = ctx.attach_multigrid(args.mg_coarse)
    base = dict(kspace=args.kspace, max_its=args.max_its, max_newton=40, precond=2,
                mg_nu=args.mg_nu, mg_omega=args.mg_omega, line_search=args.line_search,
                pc_reuse=args.pc_reuse, mg_smoother=args.mg_smoother, mg_galerkin=args.mg_galerkin,
                lin_tol_adapt=args.lin_adapt)
    fp64, mixed = dict(base, precision=0), dict(base, precision=1)
    groups = field_groups(ctx, prob)
    log(f"[learned] {prob.name}: N={ctx.N} levels {nlev}, target Ra={target:g}")

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    def nodal_error(x, ref):
        xh, rh = x.cpu().numpy(), ref.cpu().numpy()
        return {k: float(np.max(np.abs(xh[i] - rh[i])) / max(np.max(np.abs(rh[i])), 1e-300))
                for k, i in groups.items()}

    # plain FP64 path at the target, from zero (warm-up solve first)
    ctx.set_physics(problems.physics_at(prob, target))
    xw = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    ctx.solve_dev(xw.data_ptr(), **fp64)
    x_ref = torch.zeros_like(xw)
    (_, rp), t_plain = timed(lambda: ctx.solve_dev(x_ref.data_ptr(), **fp64))
    log(f"[learned] plain FP64 from zero: {rp.newton_its} passes, {rp.lin_its_total} GMRES, "
        f"{t_plain:.2f}s")
    # sweep with the target held out
    lt = np.log10(target)
    ks = np.arange(args.sweep_count) - args.sweep_count // 2
    ras = 10 ** (lt + (ks + 0.5) * args.sweep_step)
    (steps, sols), t_sweep = timed(lambda: continuation.sweep(ctx, prob, ras, fp64, log=log if args.verbose else None))
    if not all(s.converged for s in steps):
        raise SystemExit("[learned] the sweep did not converge")
    log(f"[learned] sweep of {len(ras)} values [{ras[0]:.4g}, {ras[-1]:.4g}] in {t_sweep:.1f}s")
    lg = learned.LearnedGuess(prob, ctx, ras, sols, hidden=4, width=256, seed=20261202)
    hist, t_train = timed(lambda: lg.train(epochs=args.epochs, batch=args.batch, lr=2e-3, lr_final=2e-5,
                                           seed=1))
    fpp = mlp_flops_per_point(prob.mesh.dim + 1, lg.n_out, 4, 256)
    tflops = 3.0 * fpp * lg.num_points * args.epochs / t_train / 1e12
    nb = lg.set_basis(sols)
    log(f"[learned] trained on {lg.num_points} points x {args.epochs} epochs in {t_train:.1f}s "
        f"({tflops:.1f} TFLOP/s), loss {hist[0]:.2e} -> {hist[-1]:.2e}, subspace {nb}")
    kp = int(np.searchsorted(ras, target)) - 1
    ra_prev = float(ras[kp])
    xprev = torch.tensor(sols[kp], device="cuda")
    ctx.set_physics(problems.physics_at(prob, target))
    # warm-up of the mixed-precision buffers and of the guess path (allocated on
    # first use, then kept for every later solve), untimed
    xw = xprev.clone()
    lg.guess(xw.data_ptr(), target, ra_prev, xprev.data_ptr())
    ctx.solve_dev(xw.data_ptr(), **mixed)
    out_v = {}
    for name, opts, use_net in (("mixed+learned", mixed, True), ("fp64+learned", fp64, True),
                                ("mixed+previous", mixed, False), ("fp64+previous", fp64, False)):
        x = torch.empty_like(xprev)

        def run():
            if use_net:
                lg.guess(x.data_ptr(), target, ra_prev, xprev.data_ptr())
                return ctx.solve_guess_dev(x.data_ptr(), xprev.data_ptr(), 3, 10.0, **opts)
            x.copy_(xprev)
            return ctx.solve_dev(x.data_ptr(), **opts)

        (_, rep), dt = timed(run)
        out_v[name] = dict(seconds=round(dt, 4), newton_its=rep.newton_its + rep.discarded_its,
                           gmres_its=rep.lin_its_total, converged=bool(rep.converged),
                           fell_back=bool(rep.fell_back), r0=rep.resid_L2[0],
                           phases_s={k: round(getattr(rep, k), 4) for k in
                                     ("t_assemble", "t_scale", "t_precond", "t_solve", "t_update")},
                           nodal_error_vs_fp64=nodal_error(x, x_ref))
        log(f"[learned] {name}: {dt:.2f}s, {out_v[name]['newton_its']} passes, r0 "
            f"{rep.resid_L2[0]:.2e}, fallback {bool(rep.fell_back)}, phases {out_v[name]['phases_s']}")
    t_ml = out_v["mixed+learned"]["seconds"]
    out = {
        "metric": "s per converged Newton solve (‖R‖<1e-10): mixed precision + learned guess vs "
                  "plain FP64 GPU",
        "value": t_ml, "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f32/f64",
        "data": "synthetic",
        "config": {"workload": f"{prob.name} at Ra={target:g} " +
                   ", ".join(f"{k}={v}" for k, v in params.items()),
                   "unknowns": ctx.N, "nnz": ctx.nval - 1, "elements": prob.mesh.num_elems,
                   "sweep": [round(float(r), 1) for r in ras], "network":
                   f"{prob.mesh.dim + 1}-256-256-256-256-{lg.n_out} tanh, BF16 MFMA"},
        "plain_fp64": {"seconds": round(t_plain, 4), "newton_its": rp.newton_its,
                       "gmres_its": rp.lin_its_total, "initial_guess": "zero"},
        "speedup_vs_plain_fp64": round(t_plain / t_ml, 3),
        "variants": out_v,
        "sweep_seconds": round(t_sweep, 2),
        "training": {"points": lg.num_points, "epochs": args.epochs, "seconds": round(t_train, 2),
                     "algorithmic_TFLOPs": round(tflops, 2), "loss_last": hist[-1]},
        "subspace_dim": nb,
    }
    print(json.dumps(out), flush=True)
    lg.close()
    ctx.close()
    return 0


def copy_calibration(torch):
    """A plain device copy of known size (1 GiB read, 1 GiB written, three
    times): the reference for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE
    counters to bytes."""
    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 3
    print(json.dumps({"workload": "copy calibration", "bytes_read": n * 8, "bytes_written": n * 8,
                      "launches": 3, "GBps": round(2 * n * 8 / dt / 1e9, 1)}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="natconv2d")
    ap.add_argument("--n", type=int, default=None, help="elements per side (default: workload's)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--kspace", type=int, default=50)
    ap.add_argument("--max-its", type=int, default=200)
    ap.add_argument("--patch-bins", type=int, default=0)
    ap.add_argument("--precond", choices=["mg", "ilu", "band"], default="mg")
    ap.add_argument("--mg-nu", type=int, default=None,
                    help="smoothing sweeps per level (default 1)")
    ap.add_argument("--mg-omega", type=float, default=None,
                    help="element-block update damping (default 0.65 in 2D, 0.7 in 3D)")
    ap.add_argument("--line-search", type=int, default=1, help="1: Goma's backtracking line search")
    ap.add_argument("--mg-coarse", type=int, default=2, help="coarsest elements per side")
    ap.add_argument("--mg-smoother", type=int, default=1,
                    help="element-block smoother: 0 multiplicative, 1 additive (one residual per sweep)")
    ap.add_argument("--mg-galerkin", type=int, default=1,
                    help="1: Galerkin R A P coarse operators; 0: re-assembled on each coarse mesh")
    ap.add_argument("--lin-adapt", type=int, default=1,
                    help="1: Eisenstat-Walker linear tolerances; 0: every solve to 1e-6 (Goma's fixed)")
    ap.add_argument("--sweep-precond", choices=["band", "mg"], default="band",
                    help="sweep workload: banded LU direct solve or multigrid")
    ap.add_argument("--held-out", type=int, default=32, help="sweep workload: held-out Ra values")
    ap.add_argument("--sweep-order", type=int, default=1,
                    help="sweep workload: continuation predictor order (0: previous solution, "
                         "1: first-order, with the sensitivity dx/dRa)")
    ap.add_argument("--basis-half", type=int, default=3,
                    help="sweep workload: sweep solutions on each side of a held-out value in its subspace")
    ap.add_argument("--pc-reuse", type=int, default=0,
                    help="preconditioner reuse across Newton passes: 0 recalc, 1 reuse, 2 adaptive")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--prof-every", type=int, default=1,
                    help="time one launch in this many (pseudo-random), per kernel")
    ap.add_argument("--variants", type=int, default=1,
                    help="1: also time the mixed-precision solve and report its nodal error")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep-count", type=int, default=64, help="sweep workload: Ra values")
    ap.add_argument("--epochs", type=int, default=60, help="sweep / learned: training epochs")
    ap.add_argument("--batch", type=int, default=8192, help="sweep / learned: points per training step")
    ap.add_argument("--learned-base", default="cavity3d", help="learned workload: the problem")
    ap.add_argument("--sweep-step", type=float, default=0.25,
                    help="learned workload: log10 Ra spacing of the sweep")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args(argv)
    if args.mg_nu is None:  # measured with Eisenstat-Walker tolerances: V(1,1) additive everywhere
        args.mg_nu = 1
    if args.mg_omega is None:  # measured: 0.65 best of 0.5-0.9 at 256^2; 0.7 at 48^3
        args.mg_omega = 0.7 if args.workload in ("cavity3d", "learned") else 0.65
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU decomposition is a design, not a build, in this project.",
              file=sys.stderr)
        return 2

    import torch

    from goma_amd import capi

    if args.workload == "copycal":
        return copy_calibration(torch)
    if args.workload == "sweep":
        return sweep_bench(args, torch, capi)
    if args.workload == "learned":
        if args.sweep_count == 64:
            args.sweep_count = 8
        return learned_bench(args, torch, capi)
    prec = 1 if args.precision == "mixed" else 0
    prob, params = make_problem(args.workload, args.n)
    t0 = time.perf_counter()
    ctx = capi.Context(prob)
    t_setup = time.perf_counter() - t0
    bins = args.patch_bins or 0
    pc = {"ilu": 1, "mg": 2, "band": 3}[args.precond]
    nlev = ctx.attach_multigrid(args.mg_coarse) if pc == 2 else 0
    t_setup = time.perf_counter() - t0
    print(f"[bench] {prob.name}: N={ctx.N} nnz={ctx.nval - 1} levels {nlev} setup {t_setup:.2f}s",
          file=sys.stderr, flush=True)
    x = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    opts = dict(kspace=args.kspace, max_its=args.max_its, patch_bins=bins, precision=prec,
                max_newton=40, verbose=1 if args.verbose else 0, precond=pc, mg_nu=args.mg_nu,
                mg_omega=args.mg_omega, line_search=args.line_search, pc_reuse=args.pc_reuse, mg_smoother=args.mg_smoother,
                mg_galerkin=args.mg_galerkin, lin_tol_adapt=args.lin_adapt)
    for w in range(args.warmup):
        x.zero_()
        torch.cuda.synchronize()
        rc, rep = ctx.solve_dev(x.data_ptr(), **opts)
        print(f"[bench] warmup {w}: newton {rep.newton_its} gmres {rep.lin_its_total} "
              f"{rep.t_total:.2f}s", file=sys.stderr, flush=True)
    reps = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        x.zero_()
        rc, rep = ctx.solve_dev(x.data_ptr(), **opts)
        reps.append(rep.summary())
        print(f"[bench] step {k}: newton {rep.newton_its} gmres {rep.lin_its_total} "
              f"converged {bool(rep.converged)} {rep.t_total:.2f}s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if not all(r["converged"] for r in reps):
        print("[bench] WARNING: a step did not converge", file=sys.stderr, flush=True)
    # Kernel timing on a repeat of the timed solve (same inputs, same launches):
    # every launch carries events stamped by its own dispatch.  Kept out of the
    # timed steps because per-launch events slow this launch-bound solve.
    ctx.prof_enable(True, every=args.prof_every)
    x.zero_()
    torch.cuda.synchronize()
    tp = time.perf_counter()
    rc, prep = ctx.solve_dev(x.data_ptr(), **opts)
    torch.cuda.synchronize()
    t_prof = time.perf_counter() - tp
    kern = ctx.prof_all()
    ctx.prof_enable(False)
    print(f"[bench] profiled repeat: newton {prep.newton_its} gmres {prep.lin_its_total} "
          f"{t_prof:.2f}s", file=sys.stderr, flush=True)
    variants = None
    if args.variants and prec == 0:
        variants = mixed_variant(ctx, prob, x, opts, torch)

    ms_step = elapsed / args.steps * 1e3
    n_ecol = capi.lib().gm_num_elem_colours(ctx.h)
    assemblies = kern["k_fill"]["launches"] // max(n_ecol, 1) if "k_fill" in kern else 0
    # dominant kernel (largest total time) and its roofline
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    roof = None
    if dom == "k_vanka_invert":
        # Gauss-Jordan on n x n element blocks: 2 n^3 flops per element; the
        # library's byte count is elements x n^2 x (2 s + 4), so flops follow
        # from it.  Bound: FP64 (or FP32) vector arithmetic through LDS.
        nloc = (9 * (prob.mesh.dim + int(prob.phys.energy)) + 3 if prob.mesh.dim == 2 else
                27 * (prob.mesh.dim + int(prob.phys.energy)) + 4)
        s = 8 if prec == 0 else 4
        flops = kern[dom]["bytes_per_launch"] * 2.0 * nloc / (2 * s + 4)
        tf = flops / (kern[dom]["avg_ms"] * 1e-3) / 1e12
        peak = FP64_PEAK_TFS if prec == 0 else FP32_PEAK_TFS
        roof = dict(kernel=dom, bound="fp64" if prec == 0 else "fp32", achieved=round(tf, 2),
                    peak=peak, unit="TFLOP/s", frac=round(tf / peak, 4), traffic=None,
                    flops_per_launch=round(flops), avg_launch_ms=round(kern[dom]["avg_ms"], 5))
    elif kern[dom]["GBps"]:
        a = kern[dom]["GBps"]
        tr = recorded_traffic(prob.name, dom) if prec == 0 else None
        roof = dict(kernel=dom, bound="hbm", achieved=round(a, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(a / HBM_PEAK_GBS, 4), traffic=round(tr) if tr else None,
                    bytes_per_launch=round(kern[dom]["bytes_per_launch"]),
                    avg_launch_ms=round(kern[dom]["avg_ms"], 5))
    gm_its = sum(r["lin_its_total"] for r in reps) // len(reps)
    newton = reps[0]["newton_its"]
    asm_s = kern["k_fill"]["total_ms"] * 1e-3 / max(assemblies, 1)
    byt = algorithmic_bytes(ctx, prob, prec)
    out = {
        "metric": METRIC,
        "value": round(elapsed / args.steps, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else "f32/f64",
        "data": "synthetic",
        "config": {"workload": f"{prob.name} ({args.workload}) " + ", ".join(f"{k}={v}" for k, v in params.items()),
                   "unknowns": ctx.N, "nnz": ctx.nval - 1, "elements": prob.mesh.num_elems,
                   "precision": args.precision, "solver": ("Eisenstat-Walker forcing terms (>= 1e-6), " if args.lin_adapt else "") + f"GMRES({args.kspace}) + " + ("multigrid V(%d,%d) %s element-block smoothing, %s coarse operators, %d levels, exact coarsest solve" % (args.mg_nu, args.mg_nu, ("multiplicative", "additive")[args.mg_smoother], ("re-assembled", "Galerkin")[args.mg_galerkin], nlev) if pc == 2 else "ILU(0) patch order" if pc == 1 else "banded LU") + ", preconditioner reuse " + ("recalc", "reuse", "adaptive")[args.pc_reuse],
                   "initial_guess": "zero"},
        "newton_its": newton,
        "gmres_its": gm_its,
        "final_resid_L2": reps[0]["resid_L2"][-1] if reps[0]["resid_L2"] else None,
        "assembly_s": round(asm_s, 6),
        "assembly_GBps": round(byt["asm"] / asm_s / 1e9, 1) if asm_s > 0 else None,
        "element_assemblies_per_s": round(prob.mesh.num_elems / asm_s, 1) if asm_s > 0 else None,
        "spmv_GBps": round(kern["k_spmv_msr"]["GBps"], 1) if "k_spmv_msr" in kern else None,
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kern.items()},
        "roofline": roof,
        "kernel_timing": ("HIP events stamped by each launch's own dispatch, on the library "
                          f"stream, over a repeat of the timed solve ({t_prof:.2f} s with events)"),
    }
    if variants:
        variants["fp64_s"] = out["value"]
        out["variants"] = variants
    if not args.no_cpu_baseline:
        nres = kern.get("k_fill(residual only)", {}).get("launches", 0) // max(n_ecol, 1)
        counts = dict(assemblies=assemblies // max(args.steps, 1),
                      residual_fills=nres // max(args.steps, 1), gmres_its=gm_its,
                      factorisations=max(newton - 1, 1))
        out["cpu_baseline"] = cpu_baseline(prob, counts, args.cpu_budget)
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
