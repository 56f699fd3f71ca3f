#!/usr/bin/env python3
"""Headline benchmark: converged Ginzburg-Landau Newton-GMRES solves per hour.

One step = one full plain-FP64 Newton solve (assembly of J and -R, GMRES +
18x18 point-block Jacobi, damped line search) of the order-parameter problem
from the seeded initial guess to ||R|| / ||R0|| <= 1e-10, entirely on the GPU
through the glhip C ABI.  Default workload: 3D cube [-20,20]^3, FE_Q(2) x 18
components, 32^3 cells (4 943 250 DoFs), T/Tc = 0.7, p = 26 bar, A = 0 on the
two z faces, B phase + 0.05 Delta_B uniform noise (seed 0).

Also reports, measured live with HIP events on the library's stream over the
timed region: the SpMV roofline (the dominant kernel: HBM-bound), the
assembly rate (GDoF/s) and a CPU baseline: the oracle (a CPU restatement of
the reference algorithm) timed on a bounded sample on the host cores and
extrapolated with the GPU run's own Newton/GMRES counts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "converged GL solves/hr @ ||F||<=1e-10 vs plain FP64; assembly GDoF/s, HBM GB/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6  # MI355X dense FP64 (matrix and vector)

WORKLOADS = {
    # name: (degree, cells per side, half side, p [bar], T/Tc)
    "cube_q2_ref5": (2, 32, 20.0, 26.0, 0.7),
    "cube_q1_ref3": (1, 8, 10.0, 20.0, 0.5),
}


def load_glhip():
    import importlib.util
    spec = importlib.util.spec_from_file_location("glhip", os.path.join(ROOT, "verkko-hem-repo_amd", "glhip.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def spmv_bytes(n_nodes, nnzb, val_bytes, vec_bytes):
    """Algorithmic HBM bytes of one BSR(18x18) SpMV y = J x: every block value
    once, one column index per block, the row pointer, x read once and y
    written once."""
    return nnzb * 324 * val_bytes + nnzb * 4 + (n_nodes + 1) * 4 + 2 * n_nodes * 18 * vec_bytes


# flops per quadrature point of H(q) X(q) formed as the derivative of the
# residual density at the stored iterate: 6 dot products of 9 values (108),
# the spin-space products u u^T, v v^T, u v^T and their derivatives along X
# (378), their sums and differences (72), their products with the iterate and
# X for the 18 components (864), the final combination (216); and the
# gradient terms gamma with the quadrature weight (234)
MF_POINT_FLOPS = 108 + 378 + 72 + 864 + 216 + 234


def mf_flops(deg, n_cells):
    """Algorithmic FP64 flops of one sum-factorised matrix-free product J x
    (k_mf_apply): per cell and component the three 1D passes for x and its
    gradients at the quadrature points (2 + 3 + 4 outputs per point, N1
    multiply-adds each) and the three 1D test passes (4 + 3 + 2 outputs per
    point), and per point MF_POINT_FLOPS; Robin faces left out."""
    n1 = deg + 1
    ns = n1 ** 3
    return n_cells * (2 * 18 * 18 * n1 * ns + ns * MF_POINT_FLOPS)


def mf_bytes(deg, n_cells, n_dofs):
    """Algorithmic HBM bytes of one matrix-free product: the stored iterate
    (18 values per quadrature point and cell) read once, x read once and y
    written once."""
    ns = (deg + 1) ** 3
    return n_cells * ns * 18 * 8 + 2 * n_dofs * 8


def host_cpu():
    """Host CPU of this run: model name (/proc/cpuinfo), nproc, and the CPUs
    this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(deg, h, mat, gap, n_cells, n_nodes, nnzb, newton_steps, gmres_its, ls_trials, restart):
    """CPU baseline of the same algorithm (the oracle, kind 'port') on the
    host cores (OpenMP), each leg timed on a bounded sample: the literal
    per-(i,j,q) cell loop (J + R) of assemble.cc:177-280 on one cell per
    thread, the factorized J + R and the residual-only cell work on a few
    cells per thread, and the oracle's Arnoldi step (point-block Jacobi, BSR
    SpMV, one classical Gram-Schmidt pass) on an 8^3-cell (Q2) or 16^3-cell
    (Q1) sample of the same operator at the GPU run's mean Krylov length.
    Extrapolated linearly to the full size (cells; blocks for the Arnoldi
    step) with the GPU run's own Newton, GMRES and line-search counts."""
    import oracle_lib as O
    L = O.lib()
    L.orc_time_cells.argtypes = [C.c_int, C.c_int, O._dp, O._dp, O._dp, C.c_int, C.c_int, O._dp]
    L.orc_time_cells.restype = C.c_double
    L.orc_time_residual_cells.argtypes = [C.c_int, C.c_int, O._dp, O._dp, O._dp, C.c_int, O._dp]
    L.orc_time_residual_cells.restype = C.c_double
    L.orc_time_arnoldi.argtypes = [C.c_int, O._ip, O._ip, O._dp, C.c_int, C.c_int]
    L.orc_time_arnoldi.restype = C.c_double
    L.orc_max_threads.restype = C.c_int
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or L.orc_max_threads()
    co = O.f64(mat)
    dpc = 18 * (deg + 1) ** 3
    rng = np.random.default_rng(7)
    hv = O.f64([h, h, h])
    cs = np.zeros(1)
    # factorized oracle (one Hessian per quadrature point), J+R per cell
    ns_f = max(4 * threads, 64 if deg == 2 else 512)
    Uf = rng.standard_normal(ns_f * dpc) * 0.5
    t_fac = L.orc_time_cells(deg, ns_f, hv, co, Uf, 0, threads, cs) / ns_f
    # residual only (the line-search trials of iteration.cc:170-206)
    t_res = L.orc_time_residual_cells(deg, ns_f, hv, co, Uf, threads, cs) / ns_f
    # literal per-(i,j,q) restatement of assemble.cc:177-280, one cell per thread
    ns_l = threads
    Ul = rng.standard_normal(ns_l * dpc) * 0.5
    t_lit = L.orc_time_cells(deg, ns_l, hv, co, Ul, 1, threads, cs) / ns_l
    # Arnoldi steps of the oracle's GMRES on the oracle's own J of a sample mesh
    G = load_glhip()
    ns_side = 8 if deg == 2 else 16
    m = G.Mesh(deg, (ns_side,) * 3, (0, 0, 0), (ns_side * h,) * 3, [1, 1, 1, 1, 0, 0])
    from product_util import oracle_mesh
    om = oracle_mesh(m)
    Us = O.initial_guess(m.n_nodes, gap, amp=0.05, seed=0)
    Js, _ = O.assemble(om, co, O.distribute(om, Us), R=False)
    a = m.arrays()
    n_steps = int(max(1, min(restart, round(gmres_its / max(1, newton_steps)))))
    t_arn = L.orc_time_arnoldi(m.n_nodes, a["row_ptr"], a["col"], Js.reshape(-1), n_steps, threads) / n_steps
    t_iter = t_arn * (nnzb / m.nnzb)
    solve_fac = newton_steps * t_fac * n_cells + (ls_trials + 1) * t_res * n_cells + gmres_its * t_iter
    solve_lit = newton_steps * t_lit * n_cells + (ls_trials + 1) * t_res * n_cells + gmres_its * t_iter
    return {
        "value": 3600.0 / solve_lit,
        "unit": "solves/hr",
        "cores": threads,
        "kind": "port",
        "host": host_cpu(),
        "sample": ("oracle legs timed on %d threads: literal per-(i,j,q) cell loop of assemble.cc:177-280 on %d "
                   "cells (%.2f s/cell/core), factorized J+R and residual-only cell work on %d cells, %d Arnoldi "
                   "steps (point-block Jacobi, BSR SpMV, one CGS pass) on the %d^3-cell sample (%.4f s per step, "
                   "%.3f s at full size); extrapolated with the GPU run's %d Newton steps, %d GMRES iterations and "
                   "%d line-search residuals" % (threads, ns_l, t_lit * threads, ns_f, n_steps, ns_side, t_arn,
                                                 t_iter, newton_steps, gmres_its, ls_trials)),
        "legs_s": {"literal_jr_per_cell_core": t_lit * threads, "factorized_jr_per_cell_core": t_fac * threads,
                   "residual_per_cell_core": t_res * threads, "arnoldi_step_full_size": t_iter},
        "value_factorized": 3600.0 / solve_fac,
        "extrapolated": True,
    }


def slab_batch(G, n_inst=256, seed=2):
    """The 256-slab phase-diagram sweep (warmstart.sweep_instances)."""
    return load_module("warmstart").sweep_instances(G, n_inst, seed)


def run_batch(G, args):
    meshes, coefs, Us, params = slab_batch(G, args.instances)
    batch = G.Mesh.concat(meshes)
    n_dofs = batch.n_nodes * 18
    print("bench: batch of %d instances, %d DoFs" % (len(meshes), n_dofs), file=sys.stderr, flush=True)
    kw = dict(rtol=1e-10, max_newton=60, lin_tol=args.lin_tol, lin_maxit=args.lin_maxit, restart=args.restart)
    ctx = G.Context(0)
    ctx.set_verbose(1)
    ctx.upload_mesh(batch)
    ctx.set_materials(coefs)
    ctx.set_solution(np.concatenate(Us))
    ctx.store_initial()
    ctx.set_continuation(args.continuation)
    ctx.set_matrix_free((args.operator == "matrix-free"))
    if args.direct_launch:
        # kernel timing on: every GMRES cycle is launched directly instead of
        # being replayed from captured graphs (for kernel-trace profiles)
        ctx.set_timing(True)
    out = {}
    tuned = {}
    for prec in ("fp64", "fp32"):
        kw_p = dict(kw)
        if args.autotune:
            # schedule picked by the library's autotuner on this batch (not timed)
            t0 = time.perf_counter()
            ctx.restore_initial()
            tuned[prec] = ctx.autotune(fp32=(prec == "fp32"), trial_newton=2, lin_tol=args.lin_tol,
                                       lin_maxit=args.lin_maxit)
            tuned[prec]["tuning_s"] = time.perf_counter() - t0
            print("bench: autotune %s %s" % (prec, tuned[prec]), file=sys.stderr, flush=True)
            kw_p["restart"] = 0
        for w in range(args.warmup):
            ctx.restore_initial()
            ctx.newton_solve(fp32=(prec == "fp32"), **kw_p)
        ctx.restore_initial()
        t0 = time.perf_counter()
        st = ctx.newton_solve(fp32=(prec == "fp32"), **kw_p)
        el = time.perf_counter() - t0
        ist = ctx.instance_stats()
        # convergence as reported, held to the residuals recomputed over the
        # batch at the final state (FP64, residual.cc:109-297 per instance)
        r_end = ctx.instance_residuals()
        ok = r_end <= 1e-10 * ist[:, 2] * (1 + 1e-6)
        failed = int(((ist[:, 1] == 1) & ~ok).sum())
        ist[:, 1] = np.where(ok, ist[:, 1], 0.0)
        out[prec] = dict(elapsed=el, st=st, inst=ist, F=ctx.instance_energy(), recheck_failed=failed)
        print("bench: batch %s %.3f s, %d/%d converged, %s" % (prec, el, int(ist[:, 1].sum()), len(meshes), st),
              file=sys.stderr, flush=True)
    ctx.close()
    # the plain FP64 drop-in: every instance solved on its own, one after
    # another; a seeded random sample when fewer than all are requested
    single = G.Context(0)
    single.set_continuation(args.continuation)
    single.set_matrix_free((args.operator == "matrix-free"))
    nseq = min(len(meshes), args.sequential)
    sample = np.arange(len(meshes)) if nseq == len(meshes) else np.sort(
        np.random.default_rng(7).choice(len(meshes), nseq, replace=False))
    t0 = time.perf_counter()
    seq_conv = 0
    Fseq, convseq = [], []
    for k, i in enumerate(sample):
        if k % 16 == 0:
            print("bench: sequential instance %d of %d (%.1f s)" % (k, nseq, time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
        single.upload_mesh(meshes[i])
        single.set_material(coefs[i][0], coefs[i][1:6], coefs[i][6:9], coefs[i][9])
        single.set_solution(Us[i])
        try:
            s1 = single.newton_solve(**kw)
        except G.GLHError as e:
            print("bench: sequential instance %d failed: %s" % (i, e), file=sys.stderr, flush=True)
            Fseq.append(np.nan)
            convseq.append(False)
            continue
        seq_conv += int(s1["converged"])
        Fseq.append(s1["F"])
        convseq.append(bool(s1["converged"]))
    seq_el = time.perf_counter() - t0
    single.close()
    F64, F32 = out["fp64"]["F"], out["fp32"]["F"]
    conv64 = out["fp64"]["inst"][:, 1] == 1
    both = conv64 & (out["fp32"]["inst"][:, 1] == 1)
    # Free energies are compared where both solves converged.  Newton from the
    # same guess can reach different stationary states (the problem is not
    # convex); such pairs are counted, not folded into the error.
    rel = lambda a, b: np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    e32 = rel(F32, F64)[both]
    same32 = e32 <= 1e-10
    Fseq = np.array(Fseq)
    bseq = conv64[sample] & np.array(convseq, dtype=bool)
    eseq = rel(Fseq, F64[sample])[bseq]
    sameseq = eseq <= 1e-10
    conv32 = out["fp32"]["inst"][:, 1]
    t_conv = out["fp32"]["inst"][:, 4]
    rate = lambda n, s: 3600.0 * n / s
    plain_rate = rate(seq_conv, seq_el)
    line = {
        "metric": METRIC,
        "value": rate(conv32.sum(), out["fp32"]["elapsed"]),
        "unit": "solves/hr",
        "n_gpus": 1, "steps": 1, "warmup": args.warmup,
        "ms_per_step": 1000.0 * out["fp32"]["elapsed"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "batched phase-diagram sweep: %d Q1 slab instances (8..16 x 8..16 x 2..8 cells, "
                               "%d DoFs total), T/Tc U[0.3,0.95], p U[0,34) bar, D U[6,40], seed 2; FP32 J + FP32 "
                               "GMRES in FP64 iterative refinement, FP64 residual and line search, one batched "
                               "launch per operation" % (len(meshes), n_dofs),
                   "name": "slab_batch%d" % len(meshes), "dofs": n_dofs, "instances": len(meshes),
                   "operator": "matrix-free" if args.operator == "matrix-free" else "assembled BSR"},
        "converged_instances": int(conv32.sum()),
        "convergence_rechecked": {"rule": "a converged instance's ||R||, recomputed in FP64 over the batch at the "
                                          "final state, is <= 1e-10 of its ||R0||",
                                  "reported_but_failed_fp32": out["fp32"]["recheck_failed"],
                                  "reported_but_failed_fp64": out["fp64"]["recheck_failed"]},
        "p50_instance_s": float(np.percentile(t_conv[t_conv >= 0], 50)) if (t_conv >= 0).any() else None,
        "p99_instance_s": float(np.percentile(t_conv[t_conv >= 0], 99)) if (t_conv >= 0).any() else None,
        "free_energy_fp32_vs_fp64": {"compared": int(both.sum()), "within_1e-10": int(same32.sum()),
                                     "max_rel_err_within": float(e32[same32].max()) if same32.any() else None,
                                     "other_stationary_state": int((~same32).sum())},
        "fp64_batched_solves_hr": rate(out["fp64"]["inst"][:, 1].sum(), out["fp64"]["elapsed"]),
        "plain_fp64_sequential_solves_hr": plain_rate,
        "plain_fp64_sequential_sample": nseq,
        "plain_fp64_sequential_converged": int(seq_conv),
        "free_energy_plain_vs_batched_fp64": {"compared": int(bseq.sum()), "within_1e-10": int(sameseq.sum()),
                                              "max_rel_err_within": float(eseq[sameseq].max()) if sameseq.any()
                                              else None},
        "speedup_vs_plain_fp64": rate(conv32.sum(), out["fp32"]["elapsed"]) / plain_rate if plain_rate else None,
        "schedule": tuned or {"restart": args.restart, "q_block": 8},
        "lockstep_newton_steps": out["fp32"]["st"]["newton_steps"],
        "lockstep_gmres_its": out["fp32"]["st"]["gmres_its"],
    }
    return line


def run_adaptive(G, args):
    """Adaptive slab: D = 12 xi0, lateral 40 xi0, A = 0 on the slab faces and
    natural on the sides, 16 x 16 x 4 Q1 root cells, 4 Kelly cycles refining
    30% of the cells (hanging nodes, 2:1 balance), T/Tc 0.8, p 22 bar, BnA
    guess (A phase below z = 0, B phase above) with 0.05 Delta_B noise from
    seed 1; FP32 J + FP32 GMRES in FP64 iterative refinement, every cycle
    checked against an FP64 solve from the same start."""
    lo, hi = (-20.0, -20.0, -6.0), (20.0, 20.0, 6.0)
    p, t = 22.0, 0.8
    f = G.Forest((16, 16, 4), lo, hi, [1, 1, 1, 1, 0, 0])
    m = f.mesh()
    ctx = G.Context(0)
    ctx.set_verbose(1)
    mp = ctx.set_material_pt(p, t)
    # the A/B wall is not stationary at 0.8 Tc, 22 bar: pseudo-transient
    # continuation (sigma0 = 1) carries Newton to the minimum
    ctx.set_continuation(1.0)
    # matrix-free J under the hanging-node constraints (condensed product and
    # diagonal blocks), or the assembled condensed matrix
    ctx.set_matrix_free(args.operator == "matrix-free")
    U0 = G.initial_guess_bna(m.arrays()["node_xyz"], 0.0, mp[8], mp[9], amp=0.05, seed=1)
    print("bench: adaptive slab, %d cells at cycle 0" % m.n_cells, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    rows, f2, m2 = ctx.run_adaptive(f, U0, n_refine=args.cycles, refine_frac=0.3, rtol=1e-10, lin_tol=args.lin_tol,
                                    lin_maxit=args.lin_maxit, restart=args.restart, fp32=True, check_fp64=True)
    el = time.perf_counter() - t0
    for r in rows:
        print("bench: cycle", {k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()},
              file=sys.stderr, flush=True)
    ctx.close()
    conv = [r for r in rows if r["converged"] == 1.0]
    t32 = sum(r["t_solve"] for r in rows)
    line = {
        "metric": METRIC, "value": 3600.0 * len(conv) / t32 if t32 > 0 else 0.0, "unit": "solves/hr",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1000.0 * t32, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "adaptive slab: 40 x 40 x 12 xi0, 16x16x4 Q1 root cells, %d Kelly cycles of 30%%, "
                               "T/Tc 0.8, p 22 bar, BnA guess seed 1; one solve = one converged cycle "
                               "(FP32 mixed precision), FP64 check per cycle" % args.cycles,
                   "name": "adaptive_slab", "final_dofs": int(rows[-1]["dofs"]), "final_cells": int(rows[-1]["cells"]),
                   "operator": args.operator},
        "cycles": [{k: r[k] for k in ("cycle", "cells", "dofs", "hanging", "newton_steps", "gmres_its", "F",
                                      "converged", "t_solve", "F64", "F_rel_err_vs_fp64", "U_rel_err_vs_fp64")}
                   for r in rows],
        "wall_s_including_fp64_checks": el,
    }
    return line


def load_module(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "verkko-hem-repo_amd", name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def run_warm_start(G, args):
    """Learned warm start: a coordinate MLP (64 Fourier frequencies, 5 x 256
    SiLU, 18 outputs; BF16 kernels, FP32 accumulation, Adam) trained on
    converged slab solutions (seed 3) predicts the initial guesses of held-out
    instances (seed 4), with the seeded B-phase guess as the fallback where the
    predicted guess's first residual is not lower.  Reports Newton iterations,
    batched FP32 solves/hr with and without the network, the plain FP64 path
    and the free-energy agreement with plain FP64."""
    WS = load_module("warmstart")
    t_all = time.perf_counter()
    log = lambda *a: print("bench:", *a, file=sys.stderr, flush=True)
    # 1. training data: converged FP32 batched solves of the seed-3 instances
    train = WS.slab_instances(G, args.train_instances, 3)
    Xs, Ys, n_conv = [], [], 0
    t0 = time.perf_counter()
    for b0 in range(0, len(train), 256):
        part = train[b0:b0 + 256]
        guesses = [WS.plain_guess(G, it, 3000 + b0 + i) for i, it in enumerate(part)]
        st, ist, F, Us = WS.solve_batch(G, part, guesses, fp32=True, rtol=1e-8, lin_maxit=args.lin_maxit,
                                       matrix_free=(args.operator == "matrix-free"))
        ok = ist[:, 1] == 1
        n_conv += int(ok.sum())
        X, Y = WS.training_set(part, Us, ok)
        Xs.append(X)
        Ys.append(Y)
        log("training data batch %d: %d/%d converged, %.1f s" % (b0 // 256, int(ok.sum()), len(part),
                                                                time.perf_counter() - t0))
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    t_data = time.perf_counter() - t0
    # 2. training on the GPU
    net = G.MLP(n_freq=64, sigma=args.sigma, width=256, depth=5, n_out=18, seed=3)
    net.set_data(X, Y)
    t0 = time.perf_counter()
    curve = WS.train(net, args.train_steps, args.train_batch, lr=1e-3, segment=1000, seed=3,
                     log=lambda done, L, rate: log("training step %d loss %.3e lr %.2e (%.1f s)"
                                                   % (done, L, rate, time.perf_counter() - t0)))
    t_train = time.perf_counter() - t0
    # 3. held-out instances
    test = WS.slab_instances(G, args.instances, 4)
    plain = [WS.plain_guess(G, it, 4000 + i) for i, it in enumerate(test)]
    kw = dict(rtol=1e-10, lin_maxit=args.lin_maxit, matrix_free=(args.operator == "matrix-free"))
    # Both batched FP32 paths use Eisenstat-Walker forcing and stop at
    # ||R|| <= 1e-10 ||R(plain guess)||, the same absolute accuracy; each
    # timed region holds the upload, the first residuals and the solve.
    t0 = time.perf_counter()
    st_p, ist_p, F_p, res_p, _, _ = WS.solve_from_candidates(G, test, [plain], fp32=True, **kw)
    t_plain32 = time.perf_counter() - t0
    log("plain FP32 batched: %.2f s, %d converged" % (t_plain32, int(ist_p[:, 1].sum())))
    # learned guess with the plain guess as fallback; inference is timed too
    t0 = time.perf_counter()
    learned = WS.predict_guesses(net, test)
    st_w, ist_w, F_w, res_w, choice, r_end = WS.solve_from_candidates(G, test, [plain, learned], fp32=True, **kw)
    t_warm32 = time.perf_counter() - t0
    use = choice == 1
    r_p, r_l = res_w[0], res_w[1]
    log("warm FP32 batched: %.2f s, %d converged, learned guess used for %d" % (t_warm32, int(ist_w[:, 1].sum()),
                                                                            int(use.sum())))
    # plain FP64 batched: the free-energy reference
    st_64, ist_64, F_64, _, _, _ = WS.solve_from_candidates(G, test, [plain], fp32=False, **kw)
    # plain FP64 drop-in, one instance at a time, on a seeded sample
    ns = min(args.sequential, len(test))
    sample = np.sort(np.random.default_rng(7).choice(len(test), ns, replace=False))
    t0 = time.perf_counter()
    seq_conv = 0
    for k, i in enumerate(sample):
        st1, ist1, F1, _ = WS.solve_batch(G, [test[i]], [plain[i]], fp32=False, **kw)
        seq_conv += int(ist1[0, 1] == 1)
    t_seq = time.perf_counter() - t0
    log("plain FP64 sequential sample: %d instances %.1f s" % (ns, t_seq))
    net.close()
    rate = lambda n, t: 3600.0 * n / t if t > 0 else 0.0
    conv_w, conv_p, conv_64 = ist_w[:, 1] == 1, ist_p[:, 1] == 1, ist_64[:, 1] == 1
    # converged means the recomputed residual meets the test too
    conv_w &= r_end <= 1e-10 * r_p * (1 + 1e-6)
    both = conv_w & conv_64
    rel = np.abs(F_w - F_64) / np.maximum(np.abs(F_64), 1e-300)
    same = both & (rel <= 1e-10)
    plain_rate = rate(seq_conv, t_seq)
    line = {
        "metric": METRIC, "value": rate(conv_w.sum(), t_warm32), "unit": "solves/hr", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": 1000.0 * t_warm32, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 MLP + f32/f64 solve", "data": "synthetic",
        "config": {"workload": "learned warm start: coordinate MLP (64 Fourier frequencies, 5 x 256 SiLU, 18 "
                               "outputs, BF16 with FP32 accumulation, Adam with cosine lr decay 1e-3 -> 1e-5, %d steps, batch %d) trained on "
                               "%d slab solutions (seed 3); %d held-out instances (seed 4), batched FP32 "
                               "Newton-GMRES, B-phase guess as fallback" % (args.train_steps, args.train_batch,
                                                                            n_conv, len(test)),
                   "name": "warm_start", "train_points": int(X.shape[0]),
                   "operator": "matrix-free" if args.operator == "matrix-free" else "assembled BSR"},
        "newton_steps_mean": {"learned": float(ist_w[:, 0].mean()), "plain": float(ist_p[:, 0].mean())},
        "newton_steps_lockstep": {"learned": st_w["newton_steps"], "plain": st_p["newton_steps"]},
        "gmres_its_lockstep": {"learned": st_w["gmres_its"], "plain": st_p["gmres_its"]},
        "learned_guess_used": int(use.sum()),
        "first_residual_median": {"learned": float(np.median(r_l)), "plain": float(np.median(r_p))},
        "solves_hr": {"learned_fp32_batched": rate(conv_w.sum(), t_warm32),
                      "learned_fp32_batched_incl_training": rate(conv_w.sum(), t_warm32 + t_data + t_train),
                      "plain_fp32_batched": rate(conv_p.sum(), t_plain32),
                      "plain_fp64_sequential": plain_rate},
        "speedup_vs_plain_fp64": rate(conv_w.sum(), t_warm32) / plain_rate if plain_rate else None,
        "speedup_incl_training_vs_plain_fp64": (rate(conv_w.sum(), t_warm32 + t_data + t_train) / plain_rate
                                                if plain_rate else None),
        "speedup_vs_plain_fp32_batched": (rate(conv_w.sum(), t_warm32) / rate(conv_p.sum(), t_plain32)
                                          if conv_p.sum() else None),
        "free_energy_vs_plain_fp64": {"compared": int(both.sum()), "within_1e-10": int(same.sum()),
                                      "max_rel_err_within": float(rel[same].max()) if same.any() else None,
                                      "lower_energy_state": int((both & (F_w < F_64) & ~same).sum())},
        "training": {"data_s": t_data, "train_s": t_train, "loss_curve": curve},
        "wall_s": time.perf_counter() - t_all,
    }
    return line


def run_mlp_kernels(G, args):
    """Short fixed sequence for rocprofv3: 50 training steps of the warm-start
    network (64 Fourier frequencies, 5 x 256 SiLU, 18 outputs) on a batch of
    65536 random points, then one inference pass over 262144 points."""
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (1 << 20, 6)).astype(np.float32)
    Y = rng.standard_normal((1 << 20, 18)).astype(np.float32)
    net = G.MLP(n_freq=64, sigma=0.5, width=256, depth=5, n_out=18, seed=0)
    net.set_data(X, Y)
    net.train(5, batch=65536, lr=1e-3, seed=1, loss_every=5)
    t0 = time.perf_counter()
    L = net.train(50, batch=65536, lr=1e-3, seed=1, loss_every=50)
    el = time.perf_counter() - t0
    nb, W, F2 = 65536, 256, 128
    flops_fwd = 2.0 * nb * (F2 * W + 4 * W * W + W * 64)
    flops_step = 3.0 * flops_fwd - 2.0 * nb * F2 * W  # forward, dW, dX (no dX for the first layer)
    print("bench: mlp step %.3f ms, %.1f TFLOP/s over the step (dense GEMM flops %.3g), loss %.4f"
          % (1000 * el / 50, flops_step / (el / 50) / 1e12, flops_step, L[-1]), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    net.predict(X[:262144])
    print("bench: inference of 262144 points %.3f ms" % (1000 * (time.perf_counter() - t0)), file=sys.stderr,
          flush=True)
    net.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cube_q2_ref5",
                    choices=sorted(WORKLOADS) + ["slab_batch", "adaptive_slab", "warm_start", "mlp_kernels"])
    ap.add_argument("--cycles", type=int, default=4, help="refinement cycles (adaptive_slab)")
    ap.add_argument("--gs-tile", type=int, default=0, choices=[0, 4, 8, 16],
                    help="cube workloads with --no-autotune: Gram-Schmidt tile (0: the library's default, 8)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="fixed schedule (restart --restart, blocks of 8) and Gram-Schmidt tile (8) instead of the "
                         "autotuner's (slab_batch and the cube workloads)")
    ap.add_argument("--train-instances", type=int, default=4096, help="training solutions (warm_start)")
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--train-batch", type=int, default=65536)
    ap.add_argument("--sigma", type=float, default=0.5, help="Fourier feature scale (warm_start)")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--lin-maxit", type=int, default=3000)
    ap.add_argument("--sequential", type=int, default=256,
                    help="instances solved one by one for the plain FP64 comparison (slab_batch)")
    ap.add_argument("--lin-tol", type=float, default=0.1)
    ap.add_argument("--restart", type=int, default=100)
    ap.add_argument("--no-mixed", action="store_true",
                    help="cube workloads: skip the FP32 mixed-precision solve reported beside the FP64 headline")
    ap.add_argument("--continuation", type=float, default=0.0,
                    help="slab_batch: pseudo-transient continuation sigma0 for the batched and the one-at-a-time "
                         "solves (0: plain damped Newton)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--direct-launch", action="store_true",
                    help="slab_batch: launch every GMRES cycle directly (kernel timing on) instead of replaying "
                         "captured HIP graphs")
    ap.add_argument("--no-workloads", action="store_true",
                    help="cube workloads: skip the slab_batch, warm_start and adaptive_slab records that the default "
                         "run attaches to the headline line")
    ap.add_argument("--records-sequential", type=int, default=32,
                    help="warm_start record of the default line: held-out instances solved one by one (seeded "
                         "sample) for the plain FP64 rate (the slab_batch record solves all 256 one by one)")
    ap.add_argument("--operator", choices=["matrix-free", "assembled"], default=None,
                    help="Jacobian form of the solves; default: the faster one measured per workload "
                         "(matrix-free for the cube and warm_start, assembled for slab_batch)")
    ap.add_argument("--no-assembled", action="store_true",
                    help="skip the assembled-Jacobian solve reported beside the matrix-free headline")
    ap.add_argument("--kernels", action="store_true",
                    help="short fixed kernel sequence for rocprofv3: 3 FP64 and 3 FP32 J+R assemblies, 3 residual "
                         "assemblies, 10 FP64 and 10 FP32 SpMVs, 5 copies of 1 GiB (counter calibration)")
    args = ap.parse_args()
    if args.operator is None:
        args.operator = "assembled" if args.workload == "slab_batch" else "matrix-free"
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs (--gpus %d) are not part of it" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    G = load_glhip()
    if args.workload == "slab_batch":
        print(json.dumps(run_batch(G, args)), flush=True)
        return
    if args.workload == "adaptive_slab":
        print(json.dumps(run_adaptive(G, args)), flush=True)
        return
    if args.workload == "warm_start":
        print(json.dumps(run_warm_start(G, args)), flush=True)
        return
    if args.workload == "mlp_kernels":
        run_mlp_kernels(G, args)
        return
    deg, ncell, half, p, t = WORKLOADS[args.workload]
    print("bench: building %s" % args.workload, file=sys.stderr, flush=True)
    mesh = G.Mesh(deg, (ncell,) * 3, (-half,) * 3, (half,) * 3, [1, 1, 1, 1, 0, 0])
    ctx = G.Context(0)
    m = ctx.set_material_pt(p, t)
    ctx.upload_mesh(mesh)
    U0 = G.initial_guess(mesh.n_nodes, m[9], phase=0, amp=0.05, seed=0)
    ctx.set_solution(U0)
    ctx.store_initial()
    n_dofs = mesh.n_nodes * 18
    if args.kernels:
        for fp32 in (False, True):
            for _ in range(3):
                ctx.assemble(J=True, R=True, fp32=fp32)
                print("bench: assemble fp32=%d %.3f ms" % (fp32, ctx.last_assembly_ms()), file=sys.stderr,
                      flush=True)
            print("bench: spmv fp32=%d %.3f ms" % (fp32, ctx.bench_spmv(10)), file=sys.stderr, flush=True)
        for _ in range(3):
            ctx.assemble(J=False, R=True)
        for fp32 in (False, True):
            for _ in range(3):
                ctx.assemble(J=True, R=False, matrix_free=True, fp32=fp32)
                print("bench: matrix-free setup fp32=%d %.3f ms" % (fp32, ctx.last_assembly_ms()), file=sys.stderr,
                      flush=True)
            print("bench: matrix-free J x fp32=%d %.3f ms" % (fp32, ctx.bench_spmv(10)), file=sys.stderr, flush=True)
        print("bench: copy 1 GiB %.3f ms" % ctx.bench_copy(1 << 30, 5), file=sys.stderr, flush=True)
        tr = ctx.bench_read(4 << 30, 5)
        print("bench: read 4 GiB %.3f ms (%.0f GB/s)" % (tr, (4 << 30) / (tr * 1e-3) / 1e9), file=sys.stderr,
              flush=True)
        ctx.close()
        return
    kw = dict(rtol=1e-10, max_newton=100, lin_tol=args.lin_tol, lin_maxit=100000, restart=args.restart)
    mf = (args.operator == "matrix-free")
    ctx.set_matrix_free(mf)
    tuned = None
    if not args.autotune and args.gs_tile:
        ctx.set_schedule(restart=args.restart, gs_tile=args.gs_tile)
    if args.autotune:
        # Gram-Schmidt tile picked by the library's autotuner (untimed; every
        # tile gives the same iterates to the bit)
        t0 = time.perf_counter()
        ctx.restore_initial()
        # (six Newton steps per trial: the first steps' short Krylov bases
        # alone would not tell the tiles apart)
        tuned = ctx.autotune(fp32=False, trial_newton=6, lin_tol=args.lin_tol, lin_maxit=100000)
        tuned["tuning_s"] = time.perf_counter() - t0
        print("bench: autotune %s" % tuned, file=sys.stderr, flush=True)
    for w in range(args.warmup):
        ctx.restore_initial()
        st = ctx.newton_solve(**kw)
        print("bench: warmup %d: %s" % (w, st), file=sys.stderr, flush=True)
    ctx.set_timing(True)
    runs = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.restore_initial()
        runs.append(ctx.newton_solve(**kw))
    elapsed = time.perf_counter() - t0
    tm = ctx.timing()
    for k, st in enumerate(runs):
        print("bench: step %d: %s" % (k, st), file=sys.stderr, flush=True)
    converged = all(st["r_final"] <= 1e-10 * st["r0"] for st in runs)
    ms_step = 1000.0 * elapsed / args.steps
    apply_avg_ms = tm["spmv_ms"] / max(1, tm["spmv_n"])  # one J x product (GMRES), either operator
    jasm_ms = tm["jasm_ms"] / max(1, tm["jasm_n"])
    traffic_alg = spmv_bytes(mesh.n_nodes, mesh.nnzb, 8, 8)
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    tjson = json.load(open(tfile)).get(args.workload, {}) if os.path.exists(tfile) else {}
    t_read = ctx.bench_read(4 << 30, 5)
    stream_gbs = (4 << 30) / (t_read * 1e-3) / 1e9

    def spmv_roofline(spmv_ms):
        ach = traffic_alg / (spmv_ms * 1e-3) / 1e9
        return {"kernel": "k_spmv<double,double> (BSR 18x18 SpMV in GMRES)", "bound": "hbm", "achieved": ach,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": tjson.get("k_spmv<double,double>", {}).get("bytes_per_launch"),
                "algorithmic_bytes": traffic_alg, "stream_read_gb_s": stream_gbs,
                "frac_of_stream_read": ach / stream_gbs}

    if mf:
        # algorithmically bound by the FP64 arithmetic: with H(q) X(q) formed
        # from the stored iterate (0.13 GB at the 32^3 Q2 cube) the flops at
        # the FP64 peak take longer than the bytes at the HBM peak.  MI355X runs
        # FP64 at 78.6 TFLOP/s on the vector and the matrix units alike.
        flops, nbytes = mf_flops(deg, mesh.n_cells), mf_bytes(deg, mesh.n_cells, n_dofs)
        tfs = flops / (apply_avg_ms * 1e-3) / 1e12
        gbs = nbytes / (apply_avg_ms * 1e-3) / 1e9
        roof = {"kernel": "k_mf_apply<%d> + k_mf_gather (matrix-free J x in GMRES: one launch over the cells, "
                          "one node gather)" % deg,
                "bound": "mfma", "achieved": tfs, "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                "frac": tfs / FP64_PEAK_TFS,
                "traffic": tjson.get("k_mf_apply", {}).get("bytes_per_product"), "algorithmic_bytes": nbytes,
                "algorithmic_flops": flops, "hbm_gb_s": gbs, "frac_hbm_peak": gbs / HBM_PEAK_GBS,
                "stream_read_gb_s": stream_gbs,
                # FP64 on the vector unit (the same peak as the matrix unit);
                # the kernel waits on its LDS passes and barriers (SQ counters
                # in profiles/matfree_cube_q2_ref5.md)
                "peak_note": "FP64 dense peak, vector and matrix units alike",
                "limiter": "LDS passes and barriers between them at 4 waves per SIMD"}
    else:
        roof = spmv_roofline(apply_avg_ms)
    st0 = runs[0]
    line = {
        "metric": METRIC,
        "value": 3600.0 * args.steps / elapsed if converged else 0.0,
        "unit": "solves/hr",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "GL Newton-GMRES solve, Q%d x 18 on [-%g,%g]^3, %d^3 cells, %d DoFs, T/Tc=%g, "
                               "p=%g bar, FP64 J (%s) + GMRES(%d) + point-block Jacobi, eta=%g, ||R||/||R0||<=1e-10"
                               % (deg, half, half, ncell, n_dofs, t, p, "matrix-free" if mf else "assembled BSR",
                                  args.restart, args.lin_tol),
                   "name": args.workload, "dofs": n_dofs, "cells": mesh.n_cells, "nnz": mesh.nnzb * 324,
                   "operator": "matrix-free" if mf else "assembled BSR"},
        "converged": converged,
        "newton_steps": st0["newton_steps"],
        "gmres_its": st0["gmres_its"],
        "free_energy": st0["F"],
        "jacobian_setup_ms": jasm_ms,
        "jx_ms": apply_avg_ms,
        "jx_products": tm["spmv_n"],
        "autotune": tuned,
        "roofline": roof,
    }
    if mf and not args.no_assembled:
        # the same FP64 solve with the assembled BSR Jacobian (the reference's
        # matrix), timed the same way, reported beside the headline
        U_mf = ctx.solution()
        ctx.set_matrix_free(False)
        ctx.set_timing(False)
        ctx.restore_initial()
        ctx.assemble(J=True, R=False)  # allocates the BSR values outside the timed solve
        ctx.restore_initial()
        ctx.set_timing(True)
        t1 = time.perf_counter()
        sa = ctx.newton_solve(**kw)
        ela = time.perf_counter() - t1
        tma = ctx.timing()
        Ua = ctx.solution()
        ja = tma["jasm_ms"] / max(1, tma["jasm_n"])
        sp = tma["spmv_ms"] / max(1, tma["spmv_n"])
        line["assembled"] = {
            "ms_per_solve": 1000.0 * ela, "solves_hr": 3600.0 / ela,
            "converged": bool(sa["r_final"] <= 1e-10 * sa["r0"]), "newton_steps": sa["newton_steps"],
            "gmres_its": sa["gmres_its"], "F_rel_diff_vs_matrix_free": abs(sa["F"] - st0["F"]) / abs(st0["F"]),
            "U_rel_diff_vs_matrix_free": float(np.abs(Ua - U_mf).max() / np.abs(U_mf).max()),
            "assembly_ms": ja, "assembly_gdof_s": n_dofs / (ja * 1e-3) / 1e9 if ja > 0 else None,
            "assembly_hbm_gb_s": (mesh.nnzb * 324 * 8 + 2 * n_dofs * 8) / (ja * 1e-3) / 1e9 if ja else None,
            "spmv_ms": sp, "spmv_roofline": spmv_roofline(sp)}
        print("bench: assembled %s" % {k: v for k, v in line["assembled"].items() if k != "spmv_roofline"},
              file=sys.stderr, flush=True)
        ctx.set_matrix_free(True)
    if not args.no_mixed:
        # the same solve with FP32 J and GMRES inside FP64 iterative refinement
        # (reported beside the FP64 headline, not part of its value)
        U64 = ctx.solution()
        ctx.set_timing(False)
        ctx.restore_initial()
        t1 = time.perf_counter()
        s32 = ctx.newton_solve(fp32=True, **kw)
        el32 = time.perf_counter() - t1
        U32 = ctx.solution()
        line["mixed_precision"] = {
            "operator": "matrix-free" if mf else "assembled BSR",
            "ms_per_solve": 1000.0 * el32, "converged": bool(s32["r_final"] <= 1e-10 * s32["r0"]),
            "newton_steps": s32["newton_steps"], "gmres_its": s32["gmres_its"], "free_energy": s32["F"],
            "F_rel_err_vs_fp64": abs(s32["F"] - st0["F"]) / abs(st0["F"]),
            "U_rel_err_vs_fp64": float(np.abs(U32 - U64).max() / np.abs(U64).max())}
        print("bench: mixed precision %s" % line["mixed_precision"], file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        ls = int(st0.get("ls_trials", st0["newton_steps"]))
        line["cpu_baseline"] = cpu_baseline(deg, 2 * half / ncell, list(ctx_coef(m)), m[9], mesh.n_cells,
                                            mesh.n_nodes, mesh.nnzb, int(st0["newton_steps"]), int(st0["gmres_its"]),
                                            ls, args.restart)
        print("bench: cpu baseline %s" % line["cpu_baseline"], file=sys.stderr, flush=True)
    ctx.close()
    mesh.close()
    if not args.no_workloads:
        line["workloads"] = workload_records(G, args, line)
    print(json.dumps(line), flush=True)


def workload_records(G, args, head):
    """The three additions of the plain form, each at its full configuration,
    run after the headline (outside its timed region) so that one default
    run records them all: the batched 256-slab sweep (FP32 batched with the
    autotuner against plain FP64 solves one at a time on a seeded sample),
    the learned warm start (4096 training solutions, 20 000 Adam steps,
    256 held-out instances), and the adaptive slab (4 Kelly cycles).  The
    ratios against the targets (learned batched FP32 >= 10x plain FP64 GPU,
    plain FP64 GPU >= 20x the CPU port) are summarised beside them."""
    import copy
    recs = {}
    a = copy.copy(args)
    # every instance solved one at a time: the ones that never converge run
    # to max_newton there and dominate that time, so a sample would misstate it
    a.instances, a.sequential, a.warmup, a.operator, a.autotune = 256, 256, 0, "assembled", True
    print("bench: workload record: slab_batch", file=sys.stderr, flush=True)
    recs["slab_batch"] = run_batch(G, a)
    a = copy.copy(args)
    a.instances, a.sequential, a.operator = 256, args.records_sequential, "matrix-free"
    print("bench: workload record: warm_start", file=sys.stderr, flush=True)
    recs["warm_start"] = run_warm_start(G, a)
    a = copy.copy(args)
    a.cycles = 4
    print("bench: workload record: adaptive_slab", file=sys.stderr, flush=True)
    recs["adaptive_slab"] = run_adaptive(G, a)
    for r in recs.values():
        for k in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"):
            r.pop(k, None)
    ws = recs["warm_start"]
    cpu = head.get("cpu_baseline", {}).get("value")
    recs["ratios"] = {
        "learned_fp32_batched_vs_plain_fp64_gpu": ws.get("speedup_vs_plain_fp64"),
        "learned_fp32_batched_incl_training_vs_plain_fp64_gpu": ws.get("speedup_incl_training_vs_plain_fp64"),
        "fp32_batched_vs_plain_fp64_gpu_slab_sweep": recs["slab_batch"].get("speedup_vs_plain_fp64"),
        "plain_fp64_gpu_vs_reference_cpu_port": head["value"] / cpu if cpu else None,
        "targets": {"learned_batched_vs_plain_fp64_gpu": 10.0, "plain_fp64_gpu_vs_cpu": 20.0}}
    return recs


def ctx_coef(m):
    K = 0.42072
    return [m[0], m[1], m[2], m[3], m[4], m[5], K, K, K, 1.0e10]


if __name__ == "__main__":
    main()
