#!/usr/bin/env python3
"""SIMPLE iterations/s of the device path on the precision-form lid-driven
cavity: 128^3 hexahedra on the unit cube, interior nodes perturbed by
U(-0.15, 0.15) of the spacing (seed 2026), Re = 1000 (rho = 1, mu = 1e-3,
lid u = 1), FP64 assembly, BiCGSTAB for p' to 1e-8 (L1, relative) with an
aggregation-AMG V-cycle preconditioner whose hierarchy runs in FP32
(``--precision fp64`` keeps it in FP64; one damped-Jacobi pre-sweep and two
post-sweeps per level, damping 0.85, over-correction 1.9), momentum by multicolour
symmetric Gauss-Seidel (rrF 0.1, at most 10 sweeps), urf 0.7 / 0.3.

One step = one SIMPLE outer iteration (three momentum assemble+solve passes,
Rhie-Chow, p' assembly and solve, corrections), everything resident in HBM.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def _reference_rate(pm, bu, bv, bw, types, iters, **kw):
    """SIMPLE iterations per second of the reference's own code
    (oracle/_ref/ref_simple.x, one thread): the wall time of a run of 2 + iters
    outer iterations less that of a run of 2 (start-up, geometry and the first
    iterations), so the rate is of iterations alone."""
    import pyoracle as po
    mesh = po.OrMesh(pm)
    t = [po.run_reference_simple(pm, mesh, bu, bv, bw, patch_types=types, niter=n, seconds=True,
                                 **kw)["seconds"] for n in (2, 2 + iters)]
    return iters / (t[1] - t[0]), t[1] - t[0]


def cpu_baseline(args):
    """The reference's own SIMPLE (SolveMomentumX/Y/Z and SolveContinuity
    compiled from its sources, oracle/_ref/ref_simple.x, one thread) on a
    bounded sample: the same cavity and settings on a --cpu-n^3 mesh, p' by
    the reference's algebraic multigrid with SOR smoothing (its fastest p'
    path) to rrF 1e-8.  Rate reported per 128^3-equivalent iteration by
    scaling with the cell count (a lower bound of the CPU time: the
    multigrid's cycle count also grows with the mesh).  Without the
    reference build, the oracle's restatement stands in (kind "port")."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from flowsolversourcecode_amd import meshgen
    n = args.cpu_n
    pm = meshgen.box_hex(n, n, n, perturb=0.15, seed=2026)
    bu, bv, bw = meshgen.lid_driven_cavity_walls(pm)
    scale = (n ** 3) / (args.n ** 3)
    if po.ref_simple_available():
        rate, el = _reference_rate(pm, bu, bv, bw, None, args.cpu_iters, rho=1.0, mu=1e-3, urf_u=0.7, urf_p=0.3,
                                   msolver=po.SOR, mrrf=0.1, mitmax=10, psolver=po.SOR, prrf=1e-8, pitmax=200,
                                   plmg=1)
        return {"value": rate * scale, "unit": "SIMPLE iter/s (128^3-equivalent)", "cores": 1, "kind": "reference",
                "sample": f"the reference's SIMPLE built from its sources on a {n}^3 perturbed cavity, same physics, "
                          f"p' by its algebraic multigrid (SOR smoother) to 1e-8: {args.cpu_iters} iterations in "
                          f"{el:.1f} s ({rate:.3f} it/s), scaled by cell count {n}^3/{args.n}^3"}
    oc = po.OrCase(pm, bu, bv, bw, rho=1.0, mu=1e-3, urf_u=0.7, urf_p=0.3, msolver=po.SOR, mrrf=0.1, mitmax=10,
                   psolver=po.SOR, prrf=1e-8, pitmax=20000)
    oc.iterate(2)
    k = 0
    t0 = time.perf_counter()
    while True:
        oc.iterate(1)
        k += 1
        el = time.perf_counter() - t0
        if el > args.cpu_seconds or k >= 50:
            break
    rate = k / el
    return {"value": rate * scale, "unit": "SIMPLE iter/s (128^3-equivalent)", "cores": 1, "kind": "port",
            "sample": f"oracle SIMPLE on a {n}^3 perturbed cavity, same physics, p' SOR to 1e-8, "
                      f"{k} iterations in {el:.1f} s ({rate:.3f} it/s), scaled by cell count {n}^3/{args.n}^3"}


def cpu_baseline_cavity2d(args):
    """The reference's own SIMPLE (oracle/_ref/ref_simple.x, one thread) on
    the two-dimensional cavity of the cavity2d variant with its own p' solver
    (PBCG with its D-ILU, rrF 1e-8, itmax as the device run); without the
    reference build, the oracle's restatement (kind "port")."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from flowsolversourcecode_amd import meshgen
    m, (bu, bv, bw), types = meshgen.cavity_2d(64)
    if po.ref_simple_available():
        rate, el = _reference_rate(m, bu, bv, bw, types, args.cavity2d_cpu_iters, rho=1.0, mu=0.01, urf_u=0.9,
                                   urf_p=0.1, msolver=po.SOR, mrrf=0.1, mitmax=10, psolver=po.PBCG, prrf=1e-8,
                                   pitmax=args.cavity2d_pitmax)
        return {"value": rate, "unit": "SIMPLE iter/s", "cores": 1, "kind": "reference",
                "sample": f"the reference's SIMPLE built from its sources, 64x64x1 cavity, p' PBCG "
                          f"(itmax {args.cavity2d_pitmax}), {args.cavity2d_cpu_iters} iterations in {el:.1f} s"}
    oc = po.OrCase(m, bu, bv, bw, rho=1.0, mu=0.01, urf_u=0.9, urf_p=0.1, msolver=po.SOR, mrrf=0.1, mitmax=10,
                   psolver=po.PBCG, prrf=1e-8, pitmax=args.cavity2d_pitmax)
    oc.set_patch_types(m.patch_sizes, types)
    oc.iterate(1)
    k, t0 = 0, time.perf_counter()
    while True:
        oc.iterate(1)
        k += 1
        el = time.perf_counter() - t0
        if el > args.cavity2d_cpu or k >= 500:
            break
    return {"value": k / el, "unit": "SIMPLE iter/s", "cores": 1, "kind": "port",
            "sample": f"oracle, 64x64x1 cavity, p' PBCG (itmax {args.cavity2d_pitmax}), {k} iterations in {el:.1f} s"}


COPY_BYTES = 1 << 30
HBM_PEAK_GBPS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_spmv_dot_cavity128_hbm_traffic.md")


def measured_traffic(cells):
    """HBM bytes per k_spmv_dot launch from the committed FETCH_SIZE /
    WRITE_SIZE counter summary of this workload (a line
    'traffic_bytes_per_launch cells=<N>: <bytes>'), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    head, val = line.split(":", 1)
                    if int(head.split("cells=")[1]) == cells:
                        return float(val)
    except (OSError, ValueError, IndexError):
        pass
    return None


def index_bytes(case):
    """Bytes of the ELL neighbour indices the row loops read: 16-bit offsets
    packed two slots to a word when every neighbour lies within +-32767
    cells of its row, 32-bit indices otherwise."""
    enbr = case.topology("enbr").astype(np.int64)  # (W, N), 1-based, 0 = none
    rows = np.arange(case.N)[None, :]
    off = np.where(enbr > 0, enbr - 1 - rows, 0)
    if np.abs(off).max() <= 32767:
        return 4 * ((case.W + 1) // 2) * case.N
    return 4 * case.W * case.N


def traffic_probe(case, args):
    """Launch sequence for the FETCH_SIZE / WRITE_SIZE counter runs: one SIMPLE
    iteration (to assemble a real p' system), then the ELL product alone, then
    a plain device-to-device copy of COPY_BYTES (k_copy16, 4 launches), whose
    counters give the scale from counter units to bytes."""
    import flowsolversourcecode_amd as fs
    case.iterate(1)
    ms = case.time_spmv(args.spmv_reps)
    copy_ms = fs.time_device_copy(COPY_BYTES, 3)
    print(json.dumps({"probe": "traffic", "cells": case.N, "W": case.W, "spmv_ms": ms, "spmv_reps": args.spmv_reps,
                      "spmv_bytes": 8 * case.N + index_bytes(case) + 8 * 2 * case.Fi + 24 * case.N,
                      "copy_bytes": COPY_BYTES,
                      "copy_ms": copy_ms, "copy_GBps": 2 * COPY_BYTES / (copy_ms * 1e-3) / 1e9}), flush=True)


def cavity2d(args, fs, meshgen):
    """The two-dimensional cavity (64 x 64 x 1 hexahedra, frontBack symmetry,
    Re = 100, urf 0.9 / 0.1, momentum SOR rrF 0.1 / 10 sweeps), --cavity2d-iters
    outer iterations from rest with three p' solvers to rrF 1e-8: the
    reference's PBCG (D-ILU), BiCGSTAB + FP64 AMG and BiCGSTAB + FP32 AMG.
    Reports iter/s, p' ms and iterations per solve, solves stopped at itmax,
    and the centre-line deviation from the Re = 100 benchmark."""
    m, (bu, bv, bw), types = meshgen.cavity_2d(64)
    out = {"workload": "lid-driven cavity 64x64x1 hex, frontBack symmetry, Re=100, "
                       f"{args.cavity2d_iters} SIMPLE iterations from rest"}
    for name, ps, lmg, fp32, pit in (("pbcg_reference", "pbcg", 0, 0, args.cavity2d_pitmax),
                                     ("bicgs_amg_fp64", "bicgs", fs.MG_PRECOND, 0, 300),
                                     ("bicgs_amg_fp32", "bicgs", fs.MG_PRECOND, 1, 300)):
        c = fs.Case(m, bu, bv, bw, rho=1.0, mu=0.01)
        c.set_patch_types(m.patch_sizes, types)
        c.configure(urf_u=0.9, urf_p=0.1, msolver="sor", mrrf=0.1, mitmax=10, psolver=ps, prrf=1e-8,
                    pitmax=pit, plmg=lmg)
        c.set_option("mg_fp32", fp32)
        st = c.iterate(args.cavity2d_iters)
        eu, ev = meshgen.centreline_errors(c.get("u"), c.get("v"), 64)
        k = args.cavity2d_iters
        out[name] = {"value": k / st["seconds"], "unit": "SIMPLE iter/s", "pc_solve_ms": 1e3 * st["pc_seconds"] / k,
                     "pc_iterations_per_solve": st["pc_iterations"] / k, "pc_stopped_at_itmax": st["pc_unconverged"],
                     "pc_itmax": pit, "mass_imbalance_l1": st["mass_l1"],
                     "centreline_max_dev_u_v": [eu, ev]}
        print(f"# cavity2d {name}: {k / st['seconds']:.1f} it/s, p' {1e3 * st['pc_seconds'] / k:.3f} ms "
              f"({st['pc_iterations'] / k:.1f} its, {st['pc_unconverged']} at itmax), Ghia dev {eu:.4f} {ev:.4f}",
              file=sys.stderr, flush=True)
        del c
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_cavity2d(args)
    return out


def tet_cavity(args, fs, meshgen, pm=None, second_order=False):
    """The lid-driven cavity on tetrahedra: --tet-n^3 hexahedra of the unit
    cube (interior nodes perturbed +-15 %, seed 2026) each split into 6
    tetrahedra (Kuhn), Re = 1000, the headline form's settings and cycle;
    second_order: TVD second-order upwind with the min/max limiter on the
    velocity gradients."""
    n = args.tet_n
    if pm is None:
        pm = meshgen.box_tet(n, n, n, perturb=0.15, seed=2026)
    bu, bv, bw = meshgen.lid_driven_cavity_walls(pm)
    c = fs.Case(pm, bu, bv, bw, rho=1.0, mu=1e-3)
    if second_order:
        c.set_convection("tvd", "sou")
        c.set_gradient_limiters(momentum=1)
    c.configure(urf_u=0.7, urf_p=0.3, msolver="sor", mrrf=0.1, mitmax=10, psolver="bicgs", prrf=1e-8, pitmax=200,
                plmg=fs.MG_PRECOND)
    for k, v in (("mg_fp32", 1), ("mg_alpha", args.mg_alpha), ("mg_presweep", args.mg_pre),
                 ("mg_postsweep", args.mg_post), ("mg_omega", args.mg_omega)):
        c.set_option(k, v)
    c.iterate(args.warmup)
    st = c.iterate(args.steps)
    scheme = ", second-order upwind (TVD sou), min/max-limited velocity gradients" if second_order else ""
    out = {"workload": f"lid-driven cavity {n}^3 hex x 6 tetrahedra (Kuhn), +/-15% node perturbation, Re=1000"
                       + scheme,
           "cells": c.N, "faces": c.F, "value": args.steps / st["seconds"], "unit": "SIMPLE iter/s",
           "ms_per_step": 1e3 * st["seconds"] / args.steps, "pc_solve_ms": 1e3 * st["pc_seconds"] / args.steps,
           "pc_iterations_per_solve": st["pc_iterations"] / args.steps,
           "momentum_ms": 1e3 * st["momentum_seconds"] / args.steps, "mass_imbalance_l1": st["mass_l1"]}
    print(f"# tet cavity{' second-order' if second_order else ''} {c.N} cells: {out['value']:.1f} it/s, "
          f"p' {out['pc_solve_ms']:.2f} ms ({out['pc_iterations_per_solve']:.1f} its), "
          f"momentum {out['momentum_ms']:.2f} ms", file=sys.stderr, flush=True)
    return out, pm


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="cells per side")
    ap.add_argument("--precision", choices=["fp32amg", "fp64"], default="fp32amg")
    ap.add_argument("--cpu-n", type=int, default=32, help="cells per side of the CPU baseline's sample")
    ap.add_argument("--cpu-iters", type=int, default=20, help="timed iterations of the reference's CPU sample")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="seconds of the port's CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=20, help="ELL-product launches of --traffic-probe")
    ap.add_argument("--no-variants", dest="variants", action="store_false",
                    help="skip the FP64-plain and learned-warm-start runs")
    ap.add_argument("--learn-samples", type=int, default=8)
    ap.add_argument("--learn-steps", type=int, default=160)
    ap.add_argument("--mg-alpha", type=float, default=1.9, help="over-correction of the headline V-cycle")
    ap.add_argument("--mg-pre", type=int, default=1, help="pre-sweeps per level of the headline V-cycle")
    ap.add_argument("--mg-post", type=int, default=2, help="post-sweeps per level of the headline V-cycle")
    ap.add_argument("--mg-omega", type=float, default=0.85, help="Jacobi damping of the headline V-cycle")
    ap.add_argument("--tet-n", type=int, default=64, help="hexahedra per side of the tetrahedral cavity variant")
    ap.add_argument("--clock-stride", type=int, default=5,
                    help="time every n-th launch of the roofline kernels with HIP events")
    ap.add_argument("--cavity2d-iters", type=int, default=500, help="outer iterations of the 2D-cavity variant")
    ap.add_argument("--cavity2d-pitmax", type=int, default=500, help="p' itmax of the 2D cavity's PBCG runs")
    ap.add_argument("--cavity2d-cpu", type=float, default=6.0, help="seconds of the 2D cavity's port CPU sample")
    ap.add_argument("--cavity2d-cpu-iters", type=int, default=60,
                    help="timed iterations of the 2D cavity's reference CPU sample")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="set up the workload, run one iteration, then only the ELL product (--spmv-reps launches) "
                         "and a 1 GiB device copy; for hardware-counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding of the mesh is a design "
                                   "note only (no collectives are built)", "n_gpus": args.gpus}))
        sys.exit(2)

    import flowsolversourcecode_amd as fs
    from flowsolversourcecode_amd import meshgen

    def make_case(fp32, alpha, pre=2, post=2, omega=0.8):
        pm = meshgen.box_hex(args.n, args.n, args.n, perturb=0.15, seed=2026)
        bu, bv, bw = meshgen.lid_driven_cavity_walls(pm)
        case = fs.Case(pm, bu, bv, bw, rho=1.0, mu=1e-3)
        case.configure(urf_u=0.7, urf_p=0.3, msolver="sor", mrrf=0.1, mitmax=10, psolver="bicgs", prrf=1e-8,
                       pitmax=200, plmg=fs.MG_PRECOND)
        case.set_option("mg_fp32", fp32)
        case.set_option("mg_alpha", alpha)
        case.set_option("mg_presweep", pre)
        case.set_option("mg_postsweep", post)
        case.set_option("mg_omega", omega)
        return case

    # the tuned V-cycle of the headline form: one pre-sweep and two post-sweeps
    # on every level, over-correction 1.9, Jacobi damping 0.85
    tuned = dict(alpha=args.mg_alpha, pre=args.mg_pre, post=args.mg_post, omega=args.mg_omega)
    t0 = time.perf_counter()
    case = make_case(1 if args.precision == "fp32amg" else 0, **tuned)
    setup_s = time.perf_counter() - t0
    n_cells, n_faces = case.N, case.F
    print(f"# setup {setup_s:.1f} s: N={n_cells} faces={n_faces}", file=sys.stderr, flush=True)

    if args.traffic_probe:
        traffic_probe(case, args)
        return
    if args.warmup:
        st = case.iterate(args.warmup)
        print(f"# warmup {args.warmup} its: {st['seconds']:.2f} s pc_its={st['pc_iterations']}",
              file=sys.stderr, flush=True)
    # per-launch HIP events (recorded by the dispatches themselves on the
    # case's stream) for the ELL product of the Krylov loop and the sweeps of
    # the finest AMG level, over the timed region, on every --clock-stride-th
    # launch of each (the events cost a few microseconds per timed launch)
    case.kernel_clock_start(["spmv", "fine_sweep"], capacity=16384, stride=args.clock_stride)
    st = case.iterate(args.steps)  # synchronises before and after
    el = st["seconds"]
    value = args.steps / el
    print(f"# timed {args.steps} its: {el:.3f} s, p' its {st['pc_iterations']}, p' s {st['pc_seconds']:.3f}, "
          f"mom s {st['momentum_seconds']:.3f}, mass {st['mass_l1']:.3e}", file=sys.stderr, flush=True)
    clk = {k: case.kernel_clock_read(k) for k in ("spmv", "fine_sweep")}
    case.kernel_clock_start("off")

    def roof(c, kernel, per_cell):
        if not c["launches"]:
            return None
        achieved = c["bytes"] / (c["avg_ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBPS, "traffic": None, "kernel": kernel,
                "bytes_per_launch": c["bytes"], "bytes_per_cell": per_cell, "avg_ms": c["avg_ms"],
                "launches": c["launches"], "idle_launches": c["untimed"], "sampled_every": args.clock_stride,
                "share_of_step": c["launches"] * c["avg_ms"] * 1e-3 / el}

    roof_spmv = roof(clk["spmv"], "k_spmv_dot (FP64 ELL y=A*x with (r0,y) partials, BiCGSTAB loop)",
                     clk["spmv"]["bytes"] / case.N if clk["spmv"]["launches"] else None)
    traffic = measured_traffic(case.N)
    if roof_spmv is not None and traffic is not None:
        roof_spmv["traffic"] = traffic
    roof_sweep = roof(clk["fine_sweep"], "k_sweep<1, float, *> (FP32 damped-Jacobi sweeps of the finest AMG level)",
                      clk["fine_sweep"]["bytes"] / case.N if clk["fine_sweep"]["launches"] else None)

    variants = {}
    if args.variants and args.precision == "fp32amg":
        head_fields = {k: case.get(k) for k in ("u", "v", "w", "p")}
        head_pc = st["pc_iterations"] / args.steps
        del case
        # the plain form: FP64 hierarchy and the reference's additive correction (alpha 1)
        plain = make_case(0, 1.0)
        plain.iterate(args.warmup)
        sp = plain.iterate(args.steps)
        err = {k: float(np.abs(plain.get(k) - head_fields[k]).max() / np.abs(plain.get(k)).max())
               for k in head_fields}
        variants["fp64_plain"] = {"value": args.steps / sp["seconds"], "ms_per_step": 1e3 * sp["seconds"] / args.steps,
                                  "pc_solve_ms": 1e3 * sp["pc_seconds"] / args.steps,
                                  "pc_iterations_per_solve": sp["pc_iterations"] / args.steps,
                                  "note": "FP64 AMG hierarchy, additive correction alpha = 1, same BiCGSTAB"}
        variants["field_error_fp32amg_vs_fp64_plain"] = err
        print(f"# fp64 plain: {args.steps / sp['seconds']:.2f} it/s, field error {err}", file=sys.stderr, flush=True)
        del plain
        # FP32 AMG + learned warm start, trained on the run's own first iterations
        from flowsolversourcecode_amd import learn
        lc = make_case(1, **tuned)
        net = learn.Mpnn.for_case(lc, layers=3, seed=7, capacity=args.learn_samples, use_prev=True)
        learn.attach(lc, sink=net)
        t1 = time.perf_counter()
        lc.iterate(args.learn_samples)
        learn.attach(lc)
        hist = net.train(epochs=max(1, args.learn_steps // args.learn_samples), batch=1, lr=3e-3, seed=0)
        t_train = time.perf_counter() - t1
        learn.attach(lc, warm=net)
        lc.iterate(1)
        sl = lc.iterate(args.steps)
        variants["fp32amg_learned"] = {
            "value": args.steps / sl["seconds"], "ms_per_step": 1e3 * sl["seconds"] / args.steps,
            "pc_solve_ms": 1e3 * sl["pc_seconds"] / args.steps,
            "pc_iterations_per_solve": sl["pc_iterations"] / args.steps,
            "pc_iterations_per_solve_zero_start": head_pc,
            "train": {"samples": args.learn_samples, "steps": len(hist) * args.learn_samples,
                      "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1], "seconds": t_train},
            "note": "3-layer MPNN (hidden 64, BF16 MFMA) with the previous p' among its inputs, trained on the "
                    "run's own first iterations; timed iterations start after those"}
        print(f"# learned: {args.steps / sl['seconds']:.2f} it/s, p' its {sl['pc_iterations'] / args.steps:.1f} "
              f"(zero start {head_pc:.1f}), train {t_train:.1f} s", file=sys.stderr, flush=True)
        del lc, net
        variants["cavity2d_64x64x1_re100"] = cavity2d(args, fs, meshgen)
        variants["tet_cavity"], tet_mesh = tet_cavity(args, fs, meshgen)
        variants["tet_cavity_second_order"], _ = tet_cavity(args, fs, meshgen, tet_mesh, second_order=True)
        del tet_mesh

    cpu = None if args.no_cpu else cpu_baseline(args)
    out = {
        "metric": "SIMPLE iter/s, p' solve ms to 1e-8 resid: FP64 plain vs FP32-AMG+learned",
        "value": value,
        "unit": "SIMPLE iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 (p' AMG hierarchy f32)",
        "data": "synthetic seeded mesh",
        "config": {"workload": f"lid-driven cavity {args.n}^3 hex, +/-15% node perturbation, Re=1000, "
                               f"p' BiCGSTAB+AMG({args.precision}) to 1e-8",
                   "cells": n_cells, "faces": n_faces, "parallelism": "single GPU",
                   "vcycle": {"presweeps": args.mg_pre, "postsweeps": args.mg_post, "alpha": args.mg_alpha,
                              "omega": args.mg_omega}},
        "pc_solve_ms": 1e3 * st["pc_seconds"] / args.steps,
        "pc_iterations_per_solve": st["pc_iterations"] / args.steps,
        "momentum_ms": 1e3 * st["momentum_seconds"] / args.steps,
        "pc_levels": st["pc_levels"],
        "mass_imbalance_l1": st["mass_l1"],
        "roofline": roof_spmv,
        "roofline_fine_sweep": roof_sweep,
        "variants": variants or None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
