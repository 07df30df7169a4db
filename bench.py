#!/usr/bin/env python3
"""Benchmark: seconds per converged steady Newton solve on one MI355X.

A step is one converged Newton solve (||D^-1 R||_2 < 1e-10 and
||D^-1 R||_oo < 1e-10, Goma's test) from the zero initial guess, in plain
FP64, with the mesh, matrix graph and all work arrays already resident in
HBM.  The default workload is the 3D lid-driven cavity with energy on a
48x48x48 HEX27 mesh (Q2/P1 velocity-pressure plus Q2 temperature, Re = 400,
Ra = 1e4, Pr = 0.71; 4,093,060 unknowns).

Prints one JSON line.  Besides the headline value (wall time of the timed
solves, no instrumentation) it reports:
  * for the kernel that takes the most time, its algorithmic bytes (or flops)
    per launch over its average launch duration as a fraction of the MI355X
    peak, and the same for the element assembly; durations come from a
    repeat of the timed solve in which every launch carries HIP events
    stamped by its own dispatch on the library's stream;
  * variants on the same GPU at the same tolerance, each with its nodal error
    against the FP64 solution: mixed precision from the zero guess; and, at
    Ra = 1e4 held out of a continuation sweep, FP64 and mixed precision from
    the previous-step guess and from the learned guess (the network's guess
    evaluation inside the timed region; the sweep and the training reported
    separately with the number of solves after which training pays off);
  * the oracle's FP64 CPU path timed on the GPU box's host on a bounded sample
    of the same workload, single-threaded and on all host cores, scaled by the
    counts of one GPU solve.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6   # MI355X FP64 (vector and matrix)
FP32_PEAK_TFS = 157.3  # MI355X FP32 vector / matrix
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
    return dict(asm=b_asm, spmv=b_spmv)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    """Host cores this process may use (the box's CPU share caps it at 16)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _cpu_sample(prob, out: dict, key, gmres_its: int = 10):
    """The oracle on one sample problem: one full assembly, one natural-order
    ILU(0) factorisation and gmres_its GMRES(50) iterations (FP64, 1 thread)."""
    import ctypes as C

    import oracle as O

    o = O.Oracle(prob)
    x = np.random.default_rng(1).standard_normal(o.N) * 0.1
    t0 = time.perf_counter()
    rc, r, v = o.fill(x)
    t_fill = time.perf_counter() - t0
    s = np.zeros(o.N)
    O.lib().om_row_sum_scale(o.N, v, o.ija, r, s)
    rr = C.c_double()
    t0 = time.perf_counter()
    its1 = O.lib().om_gmres_msr(o.N, v, o.ija, r, np.zeros(o.N), o.h, 1, 50, 1, 1e-30, C.byref(rr))
    t_one = time.perf_counter() - t0
    t0 = time.perf_counter()
    its = O.lib().om_gmres_msr(o.N, v, o.ija, r, np.zeros(o.N), o.h, 1, 50, 1 + gmres_its, 1e-30,
                               C.byref(rr))
    t_gm = time.perf_counter() - t0
    t_it = max((t_gm - t_one) / max(its - its1, 1), 1e-9)
    out[key] = dict(fill=t_fill, fact=max(t_one - t_it, 0.0), it=t_it, its=its - its1,
                    elems=prob.mesh.num_elems, nnz=o.nval - 1)


def cpu_baseline(workload: str, ctx, prob, counts: dict, threads: int):
    """The oracle -- a single-threaded FP64 restatement of Goma's CPU path
    (element loop with in_list scatter, GMRES(50) with natural-order ILU(0)) --
    timed on a bounded sample of the same workload, then scaled by the counts
    of one GPU solve.

    The sample is the same physics and element type on a small mesh (8^3 in
    3D, 32^2 in 2D): one assembly, one ILU(0) factorisation and ten GMRES(50)
    iterations.  Assembly scales with the element count, the factorisation
    and the iterations with the nonzeros.  It runs once on one thread, then
    as ``threads`` concurrent copies (one per core, the multi-rank pattern
    without its communication), whose aggregate rates give the all-core
    figure.  The GPU's multigrid needs far fewer GMRES iterations than
    ILU(0) would, so both figures are lower bounds on the CPU path's time.
    """
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    small, _ = make_problem(workload, 8 if prob.mesh.dim == 3 else 32)
    one = {}
    _cpu_sample(small, one, 0)
    s1 = one[0]
    par = {}
    ths = [threading.Thread(target=_cpu_sample, args=(small, par, k)) for k in range(threads)]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    wall = time.perf_counter() - t0
    # aggregate rates of the concurrent copies: each phase's wall time inside
    # a thread includes the others' contention, so T copies do T samples of
    # work in the mean phase time
    mean = {k: float(np.mean([p[k] for p in par.values()])) for k in ("fill", "fact", "it")}
    ne, nnz = prob.mesh.num_elems, ctx.nval - 1
    nasm = counts["assemblies"] + counts["residual_fills"]

    def estimate(fill_elem, fact_nnz, it_nnz):
        return (nasm * fill_elem * ne + counts["factorisations"] * fact_nnz * nnz
                + counts["gmres_its"] * it_nnz * nnz)

    e1 = estimate(s1["fill"] / s1["elems"], s1["fact"] / s1["nnz"], s1["it"] / s1["nnz"])
    eT = estimate(mean["fill"] / s1["elems"] / threads, mean["fact"] / s1["nnz"] / threads,
                  mean["it"] / s1["nnz"] / threads)
    return dict(
        value=round(eT, 1), unit="s per converged solve (extrapolated lower bound)",
        cores=threads, kind="port", cpu=cpu_model(),
        sample=(f"oracle FP64 C on {small.name} ({s1['elems']} elements, {s1['nnz']} nonzeros): "
                f"one assembly, one natural-order ILU(0) factorisation and {s1['its']} GMRES(50) "
                f"iterations, run by {threads} concurrent threads (one copy each, "
                f"{wall:.1f} s wall); scaled to {ne} elements and {nnz} nonzeros by the GPU "
                f"solve's {counts['assemblies']} Jacobian and {counts['residual_fills']} "
                f"residual assemblies, {counts['factorisations']} preconditioner set-ups and "
                f"{counts['gmres_its']} GMRES iterations (ILU(0) needs more)"),
        single_thread=dict(value=round(e1, 1), cores=1,
                           assembly_elems_per_s=round(s1["elems"] / s1["fill"], 1),
                           gmres_it_s_per_1e9_nnz=round(s1["it"] / s1["nnz"] * 1e9, 3)),
        all_cores_assembly_elems_per_s=round(s1["elems"] * threads / mean["fill"], 1),
        counts=counts)


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


def nodal_error(groups, x, ref):
    xh, rh = x.cpu().numpy(), ref.cpu().numpy()
    return {k: float(np.max(np.abs(xh[i] - rh[i])) / max(np.max(np.abs(rh[i])), 1e-300))
            for k, i in groups.items()}


def mlp_flops_per_point(n_in, n_out, hidden, width):
    """Algorithmic (unpadded) forward flops of the network per point."""
    dims = [n_in] + [width] * hidden + [n_out]
    return 2 * sum(a * b for a, b in zip(dims[:-1], dims[1:]))


def kernel_roofline(name, k, prob, prec, workload):
    """Roofline entry of one kernel family from its live timing."""
    if name == "k_vanka_invert":
        # factorisation of n x n element blocks: 2 n^3 flops per element (LU
        # plus the n right-hand sides of the explicit inverse); the library's
        # byte count is elements x n^2 x (2 s + 4), so flops follow from it
        d, en = prob.mesh.dim, int(prob.phys.energy)
        nloc = 9 * (d + en) + 3 if d == 2 else 27 * (d + en) + 4
        s = 8 if prec == 0 else 4
        flops = k["bytes_per_launch"] * 2.0 * nloc / (2 * s + 4)
        tf = flops / (k["avg_ms"] * 1e-3) / 1e12
        peak = FP64_PEAK_TFS if prec == 0 else FP32_PEAK_TFS
        tr = recorded_traffic(workload, name) if prec == 0 else None
        return dict(kernel=name, bound="mfma", achieved=round(tf, 2), peak=peak, unit="TFLOP/s",
                    frac=round(tf / peak, 4), traffic=round(tr) if tr else None,
                    flops_per_launch=round(flops),
                    avg_launch_ms=round(k["avg_ms"], 5),
                    precision="fp64" if prec == 0 else "fp32")
    a = k["GBps"] or 0.0
    tr = recorded_traffic(workload, name) if prec == 0 else None
    return dict(kernel=name, bound="hbm", achieved=round(a, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(a / HBM_PEAK_GBS, 4), traffic=round(tr) if tr else None,
                bytes_per_launch=round(k["bytes_per_launch"] or 0),
                avg_launch_ms=round(k["avg_ms"], 5))


def learned_variants(args, torch, ctx, prob, target, x_ref, t_plain, fp64, mixed, log):
    """Previous-step and learned guesses at a Rayleigh number held out of a
    continuation sweep, in FP64 and mixed precision.

    Sweep: Ra values a quarter decade apart, offset so that the target is the
    geometric midpoint of two of them (FP64, zeroth-order continuation).  The
    network is trained on the GPU on the sweep's solutions.  At the target:
    previous-step guess = the sweep solution just below; learned guess = that
    solution moved by the network's predicted change (delta form, projected
    onto the sweep's subspace), with the previous-step guess as fallback."""
    from goma_amd import continuation, learned, problems

    groups = field_groups(ctx, prob)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    lt = np.log10(target)
    ks = np.arange(args.sweep_count) - args.sweep_count // 2
    ras = 10 ** (lt + (ks + 0.5) * args.sweep_step)
    (steps, sols), t_sweep = timed(lambda: continuation.sweep(ctx, prob, ras, fp64))
    if not all(s.converged for s in steps):
        raise SystemExit("[bench] the continuation sweep did not converge")
    log(f"[bench] sweep of {len(ras)} values [{ras[0]:.4g}, {ras[-1]:.4g}] in {t_sweep:.1f}s")
    lg = learned.LearnedGuess(prob, ctx, ras, sols, hidden=4, width=256, seed=20261202)
    hist, t_train = timed(lambda: lg.train(epochs=args.epochs, batch=args.batch, lr=2e-3,
                                           lr_final=2e-5, seed=1))
    fpp = mlp_flops_per_point(prob.mesh.dim + 1, lg.n_out, 4, 256)
    tflops = 3.0 * fpp * lg.num_points * args.epochs / t_train / 1e12
    nb = lg.set_basis(sols)
    log(f"[bench] trained on {lg.num_points} points x {args.epochs} epochs in {t_train:.1f}s "
        f"({tflops:.1f} TFLOP/s), loss {hist[0]:.2e} -> {hist[-1]:.2e}, subspace {nb}")
    kp = int(np.searchsorted(ras, target)) - 1
    ra_prev = float(ras[kp])
    xprev = torch.tensor(sols[kp], device="cuda")
    ctx.set_physics(problems.physics_at(prob, target))
    # warm-up of the guess path (buffers allocated on first use), untimed
    xw = xprev.clone()
    lg.guess(xw.data_ptr(), target, ra_prev, xprev.data_ptr())
    # the guess alone (median of three): two forward passes, at the target and
    # the previous Ra, over every sample point, with the inputs built on the
    # device and the result projected onto the sweep's subspace
    tg = float(np.median([timed(lambda: lg.guess(xw.data_ptr(), target, ra_prev,
                                                 xprev.data_ptr()))[1] for _ in range(3)]))
    pts = lg.num_points // len(ras)
    guess = dict(ms=round(1e3 * tg, 2), passes=2, points_per_pass=pts,
                 algorithmic_TFLOPs=round(2 * fpp * pts / tg / 1e12, 1), bf16_peak_TFLOPs=2500.0)
    log(f"[bench] guess: {1e3 * tg:.1f} ms, {guess['algorithmic_TFLOPs']} TFLOP/s")
    out = {}
    for name, opts, use_net in (("fp64+previous", fp64, False), ("mixed+previous", mixed, False),
                                ("fp64+learned", fp64, True), ("mixed+learned", mixed, True)):
        x = torch.empty_like(xprev)

        def run():
            if use_net:
                lg.guess(x.data_ptr(), target, ra_prev, xprev.data_ptr())
                return ctx.solve_guess_dev(x.data_ptr(), xprev.data_ptr(), 3, 10.0, **opts)
            x.copy_(xprev)
            return ctx.solve_dev(x.data_ptr(), **opts)

        # median of a few runs from the same guess (each run starts over)
        runs = [timed(run) for _ in range(max(1, args.learned_reps))]
        times = [t for _, t in runs]
        dt = float(np.median(times))
        rep = runs[-1][0][1]
        out[name] = dict(seconds=round(dt, 4), runs=[round(t, 4) for t in times],
                         newton_its=rep.newton_its + rep.discarded_its,
                         gmres_its=rep.lin_its_total, converged=bool(rep.converged),
                         fell_back=bool(rep.fell_back), r0=rep.resid_L2[0],
                         nodal_error_vs_fp64=nodal_error(groups, x, x_ref))
        log(f"[bench] {name}: {dt:.3f}s, {out[name]['newton_its']} passes, "
            f"{rep.lin_its_total} GMRES, r0 {rep.resid_L2[0]:.2e}, fallback {bool(rep.fell_back)}")
    t_prev = out["fp64+previous"]["seconds"]
    t_ml = out["mixed+learned"]["seconds"]
    gain = out["mixed+previous"]["seconds"] - t_ml
    res = dict(
        target_Ra=target, sweep_Ra=[round(float(r), 1) for r in ras], held_out_between=[
            round(ra_prev, 1), round(float(ras[kp + 1]), 1)],
        variants=out,
        speedup_mixed_learned_vs_fp64_previous=round(t_prev / t_ml, 3),
        speedup_mixed_learned_vs_fp64_zero=round(t_plain / t_ml, 3),
        sweep_seconds=round(t_sweep, 2),
        training=dict(points=lg.num_points, epochs=args.epochs, seconds=round(t_train, 2),
                      algorithmic_TFLOPs=round(tflops, 2), bf16_peak_TFLOPs=2500.0,
                      loss_last=hist[-1],
                      break_even_solves=(round(t_train / gain, 1) if gain > 0 else None)),
        guess=guess,
        network=f"{prob.mesh.dim + 1}-256-256-256-256-{lg.n_out} tanh, BF16 MFMA",
        subspace_dim=nb)
    lg.close()
    return res


def headline(args, torch, capi):
    """The timed FP64 solves, the profiled repeat, the variants and the CPU
    baseline of one workload."""
    from goma_amd import problems

    log = lambda s: print(s, file=sys.stderr, flush=True)
    prob, params = make_problem(args.workload, args.n)
    t0 = time.perf_counter()
    ctx = capi.Context(prob)
    pc = {"ilu": 1, "mg": 2, "band": 3}[args.precond]
    nlev = ctx.attach_multigrid(args.mg_coarse) if pc == 2 else 0
    t_setup = time.perf_counter() - t0
    log(f"[bench] {prob.name}: N={ctx.N} nnz={ctx.nval - 1} levels {nlev} setup {t_setup:.2f}s")
    base = dict(kspace=args.kspace, max_its=args.max_its, patch_bins=args.patch_bins,
                max_newton=40, verbose=1 if args.verbose else 0, precond=pc, mg_nu=args.mg_nu,
                mg_omega=args.mg_omega, line_search=args.line_search, pc_reuse=args.pc_reuse,
                mg_smoother=args.mg_smoother, mg_galerkin=args.mg_galerkin,
                lin_tol_adapt=args.lin_adapt)
    fp64, mixed = dict(base, precision=0), dict(base, precision=1)
    x = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    for w in range(args.warmup):
        x.zero_()
        torch.cuda.synchronize()
        rc, rep = ctx.solve_dev(x.data_ptr(), **fp64)
        log(f"[bench] warmup {w}: newton {rep.newton_its} gmres {rep.lin_its_total} "
            f"{rep.t_total:.2f}s")
    reps = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        x.zero_()
        rc, rep = ctx.solve_dev(x.data_ptr(), **fp64)
        reps.append(rep.summary())
        log(f"[bench] step {k}: newton {rep.newton_its} gmres {rep.lin_its_total} "
            f"converged {bool(rep.converged)} {rep.t_total:.2f}s")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if not all(r["converged"] for r in reps):
        log("[bench] WARNING: a step did not converge")
    t_plain = elapsed / args.steps
    # Kernel timing on a repeat of the timed solve (same inputs, same launches):
    # every launch carries events stamped by its own dispatch.  Kept out of the
    # timed steps because per-launch events slow the launch-bound parts.
    ctx.prof_enable(True, every=args.prof_every)
    x.zero_()
    torch.cuda.synchronize()
    tp = time.perf_counter()
    rc, prep = ctx.solve_dev(x.data_ptr(), **fp64)
    torch.cuda.synchronize()
    t_prof = time.perf_counter() - tp
    kern = ctx.prof_all()
    ctx.prof_enable(False)
    log(f"[bench] profiled repeat: newton {prep.newton_its} gmres {prep.lin_its_total} "
        f"{t_prof:.2f}s")
    x_ref = x.clone()
    groups = field_groups(ctx, prob)

    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    roof = kernel_roofline(dom, kern[dom], prob, 0, prob.name)
    # assembly: a Jacobian + residual assembly is one launch of k_fill per
    # Newton pass (per element colour if the kernel is coloured)
    nasm = prep.newton_its
    k_fill = kern.get("k_fill")
    asm_s = k_fill["total_ms"] * 1e-3 / max(nasm, 1) if k_fill else None
    byt = algorithmic_bytes(ctx, prob, 0)
    asm_roof = kernel_roofline("k_fill", k_fill, prob, 0, prob.name) if k_fill else None
    if asm_roof:
        asm_roof["per_assembly_s"] = round(asm_s, 6)
        if prob.mesh.dim == 3:
            # HEX27 node-pair blocks on the FP64 matrix cores (k_fill3m): per
            # element and wave, 7 k-steps of 5 shared products (3 more with
            # Carreau), 6 per momentum block row and 3 for the energy row;
            # four waves; 2 * 16 * 16 * 4 flop per v_mfma_f64_16x16x4f64
            en, vm = int(prob.phys.energy), int(prob.phys.visc_model)
            per_wave = 7 * ((5 + 3 * vm) + 3 * (6 + 3 * vm) + 3 * en)
            ncol = int(ctx.elem_colours().max()) + 1
            flops = 4 * per_wave * 2048.0 * prob.mesh.num_elems / ncol
            tf = flops / (k_fill["avg_ms"] * 1e-3) / 1e12
            asm_roof["mfma"] = dict(bound="mfma", achieved=round(tf, 2), peak=FP64_PEAK_TFS,
                                    unit="TFLOP/s", frac=round(tf / FP64_PEAK_TFS, 4),
                                    flops_per_launch=round(flops),
                                    note="padded 32 x 32 x 28 products; the map, field, "
                                         "residual and scatter phases run beside them")
    out = {
        "metric": METRIC,
        "value": round(t_plain, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_plain * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{prob.name}: " + ", ".join(f"{k}={v}" for k, v in params.items())
                   + ", zero initial guess, plain FP64",
                   "unknowns": ctx.N, "nnz": ctx.nval - 1, "elements": prob.mesh.num_elems,
                   "solver": ("Eisenstat-Walker forcing terms (>= 1e-6), " if args.lin_adapt else "")
                   + f"GMRES({args.kspace}) + " + (
                       "multigrid V(%d,%d) %s element-block smoothing, %s coarse operators, %d levels,"
                       " exact coarsest solve" % (args.mg_nu, args.mg_nu,
                                                  ("multiplicative", "additive")[args.mg_smoother],
                                                  ("re-assembled", "Galerkin")[args.mg_galerkin], nlev)
                       if pc == 2 else "ILU(0) patch order" if pc == 1 else "banded LU")
                   + ", backtracking line search" * bool(args.line_search)},
        "newton_its": reps[0]["newton_its"],
        "gmres_its": reps[0]["lin_its_total"],
        "final_resid_L2": reps[0]["resid_L2"][-1] if reps[0]["resid_L2"] else None,
        "roofline": roof,
        "assembly": dict(asm_roof or {}, element_assemblies_per_s=(
            round(prob.mesh.num_elems / asm_s, 1) if asm_s else None),
            GBps=round(byt["asm"] / asm_s / 1e9, 1) if asm_s else None),
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kern.items()},
        "kernel_timing": ("HIP events stamped by each launch's own dispatch, on the library "
                          f"stream, over a repeat of the timed solve ({t_prof:.2f} s with events); "
                          "with timing on, the preconditioner set-up runs its kernels in order "
                          "(the timed solves run the block inversions, on a stream kept off 1/8 "
                          "of the CUs, beside the coarse Galerkin products and the coarsest dense "
                          "inverse), so each duration is the kernel's own"),
    }
    if args.variants:
        xm = torch.zeros_like(x)
        torch.cuda.synchronize()
        tm = time.perf_counter()
        rc, rm = ctx.solve_dev(xm.data_ptr(), **mixed)
        torch.cuda.synchronize()
        t_mixed = time.perf_counter() - tm
        # the first mixed solve allocates its FP32 buffers: time a second one
        xm.zero_()
        torch.cuda.synchronize()
        tm = time.perf_counter()
        rc, rm = ctx.solve_dev(xm.data_ptr(), **mixed)
        torch.cuda.synchronize()
        t_mixed = time.perf_counter() - tm
        var = {"fp64+zero": dict(seconds=round(t_plain, 4), newton_its=reps[0]["newton_its"],
                                 gmres_its=reps[0]["lin_its_total"]),
               "mixed+zero": dict(seconds=round(t_mixed, 4), newton_its=rm.newton_its,
                                  gmres_its=rm.lin_its_total, converged=bool(rm.converged),
                                  nodal_error_vs_fp64=nodal_error(groups, xm, x_ref),
                                  speedup_vs_fp64_zero=round(t_plain / t_mixed, 3))}
        log(f"[bench] mixed precision: newton {rm.newton_its} gmres {rm.lin_its_total} "
            f"{t_mixed:.2f}s nodal error {var['mixed+zero']['nodal_error_vs_fp64']}")
        out["variants"] = var
        if args.learned and "Ra" in params:
            out["learned"] = learned_variants(args, torch, ctx, prob, float(params["Ra"]), x_ref,
                                              t_plain, fp64, mixed, log)
            ctx.set_physics(prob.phys)
    if not args.no_cpu_baseline:
        counts = dict(assemblies=prep.newton_its, residual_fills=prep.resid_fills,
                      factorisations=prep.pc_builds, gmres_its=prep.lin_its_total)
        out["cpu_baseline"] = cpu_baseline(args.workload, ctx, prob, counts, args.cpu_threads
                                           or host_threads())
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


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


def ilu_config3(args, torch, capi):
    """Config 3's solver as specified: FP32 Jacobian values and ILU(0), inner
    GMRES(50), FP64 residual and Newton update (precision = mixed,
    preconditioner ILU(0) in patch order), from the zero guess on the 48^3
    cavity, GMRES capped at max-its iterations per pass and the fixed linear
    tolerance 1e-6 (Goma's Residual Ratio Tolerance default).  Prints one JSON
    line with the per-pass GMRES iterations and scaled residual norms; not a
    timed benchmark (the multigrid solve is the headline)."""
    prob, _ = make_problem("cavity3d", args.n)
    ctx = capi.Context(prob)
    x = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc, rep = ctx.solve_dev(x.data_ptr(), precond=1, precision=1, kspace=50, max_its=args.max_its,
                            max_newton=15, line_search=1, lin_tol=1e-6, lin_tol_adapt=0,
                            patch_bins=args.patch_bins, verbose=1 if args.verbose else 0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    n = min(rep.newton_its, 64)
    print(json.dumps({"workload": f"{prob.name}: FP32 ILU(0) + GMRES(50), FP64 residual",
                      "converged": bool(rep.converged), "newton_passes": rep.newton_its,
                      "gmres_its_total": rep.lin_its_total, "seconds": round(dt, 2),
                      "gmres_its": [int(rep.lin_its[i]) for i in range(n)],
                      "resid_L2": [float(rep.resid_L2[i]) for i in range(n)],
                      "max_its_per_pass": args.max_its}))
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
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cavity3d",
                    help="cavity3d (48^3 HEX27 lid cavity + energy, the default), natconv2d "
                         "(256^2 Boussinesq), carreau2d, cavity2d, sweep (config-4 Ra sweep and "
                         "learned guess), copycal (PMC counter calibration copy), cavity3d-ilu "
                         "(config 3's FP32 ILU(0) + GMRES(50) solver at 48^3, one solve)")
    ap.add_argument("--n", type=int, default=None, help="elements per side (default: workload's)")
    ap.add_argument("--kspace", type=int, default=50)
    ap.add_argument("--max-its", type=int, default=200)
    ap.add_argument("--patch-bins", type=int, default=0)
    ap.add_argument("--precond", choices=["mg", "ilu", "band"], default="mg")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="sweep workload: Newton precision")
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
    ap.add_argument("--pc-reuse", type=int, default=0,
                    help="preconditioner reuse across Newton passes: 0 recalc, 1 reuse, 2 adaptive")
    ap.add_argument("--prof-every", type=int, default=1,
                    help="time one launch in this many (pseudo-random), per kernel")
    ap.add_argument("--variants", type=int, default=1,
                    help="1: also time mixed precision and the previous-step / learned guesses")
    ap.add_argument("--learned", type=int, default=1,
                    help="1: continuation sweep, network training and the guess variants")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the all-core CPU sample (default: the host cores, at most 16)")
    ap.add_argument("--sweep-count", type=int, default=8,
                    help="Ra values of the continuation sweep (64 for the sweep workload)")
    ap.add_argument("--sweep-step", type=float, default=0.25,
                    help="log10 Ra spacing of the continuation sweep around the target")
    ap.add_argument("--epochs", type=int, default=60, help="training epochs")
    ap.add_argument("--learned-reps", type=int, default=3,
                    help="runs of each learned-guess variant (the median is reported)")
    ap.add_argument("--batch", type=int, default=8192, help="points per training step")
    ap.add_argument("--sweep-precond", choices=["band", "mg"], default="band",
                    help="sweep workload: banded LU direct solve or multigrid")
    ap.add_argument("--held-out", type=int, default=32, help="sweep workload: held-out Ra values")
    ap.add_argument("--sweep-order", type=int, default=1,
                    help="sweep workload: continuation predictor order (0: previous solution, "
                         "1: first-order, with the sensitivity dx/dRa)")
    ap.add_argument("--basis-half", type=int, default=3,
                    help="sweep workload: sweep solutions on each side of a held-out value in its subspace")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args(argv)
    if args.mg_nu is None:  # measured with Eisenstat-Walker tolerances: V(1,1) additive everywhere
        args.mg_nu = 1
    if args.mg_omega is None:  # measured: 0.65 best of 0.5-0.9 at 256^2; 0.7 at 48^3
        args.mg_omega = 0.7 if args.workload == "cavity3d" else 0.65
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
    if args.workload == "cavity3d-ilu":
        return ilu_config3(args, torch, capi)
    if args.workload == "sweep":
        if args.sweep_count == 8:
            args.sweep_count = 64
        return sweep_bench(args, torch, capi)
    return headline(args, torch, capi)


if __name__ == "__main__":
    sys.exit(main())
