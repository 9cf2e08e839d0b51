#!/usr/bin/env python3
"""SIMPLE iterations/s of the device path on the RANS channel: 64x128x160
hexahedra graded 1.05 toward both walls, each split into 6 tetrahedra (Kuhn;
7,864,320 cells), k-omega SST with low-Reynolds-number walls, Re_tau = 550 by
the prescribed mean pressure gradient, periodic streamwise, symmetry
spanwise, second-order upwind momentum with the min/max gradient limiter,
FP64 assembly, p' by CG to 1e-8 (L1, relative) preconditioned by an
aggregation-AMG F-cycle whose hierarchy runs in FP32; variants on the same
line: the plain FP64 form (BiCGSTAB + FP64 V-cycle), the headline's cycle in
FP64, and the learned warm start.  ``--workload cavity`` runs the 128^3
perturbed lid-driven cavity instead (Re = 1000, BiCGSTAB + FP32 AMG).

One step = one SIMPLE outer iteration (three momentum assemble+solve passes,
Rhie-Chow, p' assembly and solve, corrections, the k and omega equations and
mu_t), everything resident in HBM.  Prints one JSON line.
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
CHANNEL_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "channel_64x128x160x6_hbm_traffic.md")


def measured_traffic(cells, site=None, path=TRAFFIC_PROFILE):
    """HBM bytes per launch from a committed FETCH_SIZE / WRITE_SIZE counter
    summary (a line 'traffic_bytes_per_launch cells=<N>: <bytes>', or
    'traffic_bytes_per_launch site=<site> cells=<N>: <bytes>'), or None."""
    try:
        with open(path) as f:
            for line in f:
                if not line.startswith("traffic_bytes_per_launch"):
                    continue
                head, val = line.split(":", 1)
                if int(head.split("cells=")[1]) != cells:
                    continue
                if site is not None and f"site={site} " not in head + " ":
                    continue
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


# The RANS channel of the headline: half-height delta = 1 between walls at
# y = 0 and y = 2, streamwise 2 pi (periodic), spanwise pi (symmetry planes),
# rho = 1, u_tau = 1 through the prescribed pressure gradient beta =
# tau_w / delta = 1 (the fixed-flux equilibrium), nu = 1 / Re_tau.
RE_TAU = 550.0
# velocity and pressure relaxation of the channel: at 0.7 / 0.3 the reference's
# own SIMPLE + SST diverges within 20 iterations on this mesh family once the
# streamwise cells are 5x the spanwise ones (16x32x40 and finer)
URF_U, URF_P = 0.5, 0.2
# k and omega: first-order upwind (with second-order upwind on them as well
# the same iteration diverges after 60-80 iterations at 16x32x40, in the
# oracle and on the device alike), at most 10 sweeps to rrF 0.1, relaxed by 0.5
CHANNEL_SST = dict(kitmax=10, witmax=10, kframework="none", kscheme="upwind", wframework="none", wscheme="upwind",
                   urf_k=0.5, urf_w=0.5, urf_mut=0.5, krrf=0.1, wrrf=0.1)


def channel_mesh(dims, meshgen):
    return meshgen.channel_tet(*dims, lx=2.0 * np.pi, ly=2.0, lz=np.pi, grading=1.05)


def channel_initial(pm, types, yc, wd):
    """Start of the channel runs, in wall units (u_tau = 1, nu = 1/Re_tau):
    u from Reichardt's law of the wall at y+ = wd Re_tau; k rising as
    (y+/10)^2 from the wall to u_tau^2/sqrt(C_mu) and falling as (1 - y)^2
    toward the centre plane; omega the larger of the log-layer value
    sqrt(k)/(C_mu^(1/4) kappa y) and the viscous-sublayer value
    6 nu/(beta_1 y^2); mu_t = k/omega.  On the boundary faces k = mu_t = 0 at
    the walls and the owner cell's values elsewhere.  (A uniform start, k = 1
    and omega = 10 with a power-law u, sends the reference's SST iteration
    into divergence within a few hundred iterations on this mesh family.)"""
    nb = pm.n_faces - pm.n_internal
    kinds = np.concatenate([[t] * n for t, n in zip(types, pm.patch_sizes)])
    wall = kinds == "wall"
    kappa, cmu, beta1 = 0.41, 0.09, 0.075
    nu = 1.0 / RE_TAU
    yp = wd / nu
    u = np.log1p(kappa * yp) / kappa + 7.8 * (1.0 - np.exp(-yp / 11.0) - yp / 11.0 * np.exp(-yp / 3.0))
    k = np.maximum(np.minimum(1.0, (yp / 10.0) ** 2) * (1.0 - wd) ** 2 / np.sqrt(cmu), 1e-8)
    om = np.maximum(np.sqrt(k) / (cmu ** 0.25 * kappa * wd), 6.0 * nu / (beta1 * wd ** 2))
    mut = k / om
    ob = np.asarray(pm.owner)[pm.n_internal:]
    sst = dict(k=k, omega=om, mut=mut, wd=wd, bk=np.where(wall, 0.0, k[ob]), bomega=om[ob],
               bmut=np.where(wall, 0.0, mut[ob]))
    return u, sst


def cell_heights(pm):
    """y of every cell's centre, taken as the mean of its faces' node means
    (the initial profile only)."""
    fp = np.asarray(pm.face_ptr, np.int64)
    nodes = np.asarray(pm.face_nodes, np.int64)
    y = np.asarray(pm.points)[:, 1]
    cnt = np.diff(fp)
    fy = np.add.reduceat(y[nodes], fp[:-1]) / cnt
    n = int(max(np.max(pm.owner), np.max(pm.neighbour[:pm.n_internal]))) + 1
    s, c = np.zeros(n), np.zeros(n)
    np.add.at(s, pm.owner, fy)
    np.add.at(c, pm.owner, 1.0)
    nb = np.asarray(pm.neighbour)[:pm.n_internal]
    np.add.at(s, nb, fy[:pm.n_internal])
    np.add.at(c, nb, 1.0)
    return s / c


def channel_case(args, fs, pm, types, pairs, tr, fp32, alpha, pre=2, post=2, omega=0.8, wd=None, agg=0,
                 psolver="bicgs", fuse=1, maxlevels=10, cycle=0, csweeps=None):
    """The headline's case: SIMPLE with second-order upwind (TVD sou) and the
    min/max limiter on the velocity gradients, momentum and k, omega by
    multicolour Gauss-Seidel (rrF 0.1, <= 10 sweeps), p' BiCGSTAB + AMG to
    1e-8, k-omega SST with low-Reynolds-number walls.  wd None takes the
    plane channel's wall distance min(y, 2 - y), the solution the Poisson
    method (CalculateNormalDistanceToWall) approximates there, or with
    --poisson-wall-distance the device's Poisson method to a relative outer
    residual of 1e-6 (its deferred non-orthogonal correction needs hundreds of
    passes on the graded tetrahedra).  Returns the case, wd and the passes."""
    z = np.zeros(pm.n_faces - pm.n_internal)
    c = fs.Case(pm, z, z, z, rho=1.0, mu=1.0 / RE_TAU)
    c.configure(urf_u=URF_U, urf_p=URF_P, msolver="sor", mrrf=0.1, mitmax=10, psolver=psolver, prrf=1e-8,
                pitmax=200, plmg=fs.MG_PRECOND)
    c.set_patch_types(pm.patch_sizes, types)
    c.set_periodic(pm.patch_sizes, pairs, tr, beta=1.0, calc_beta=False)
    c.set_convection("tvd", "sou")
    c.set_gradient_limiters(momentum=1)
    for key, v in (("mg_fp32", fp32), ("mg_alpha", alpha), ("mg_presweep", pre), ("mg_postsweep", post),
                   ("mg_omega", omega), ("mg_aggregation", agg), ("mg_fuse_prolong", fuse),
                   ("mg_maxlevels", maxlevels), ("mg_cycle", cycle)):
        c.set_option(key, v)
    if csweeps is not None:
        c.set_option("mg_coarse_presweep", csweeps[0])
        c.set_option("mg_coarse_postsweep", csweeps[1])
    passes = 0
    if wd is None and args.poisson_wall_distance:
        t0 = time.perf_counter()
        wd, passes = c.wall_distance(tol=1e-6, maxouter=5000)
        print(f"# wall distance: {passes} passes, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    elif wd is None:
        yc = cell_heights(pm)
        wd = np.minimum(yc, 2.0 - yc)
    u, sst = channel_initial(pm, types, cell_heights(pm), wd)
    c.set("u", u)
    c.set_sst(**sst, ksolver="sor", wsolver="sor", **CHANNEL_SST)
    return c, wd, passes


def cpu_baseline_channel(args, full_cells):
    """The reference's own SIMPLE with CalculateEffectsOfTurbulence (k-omega
    SST, compiled from its sources, oracle/_ref/ref_simple.x, one thread) on a
    bounded sample: the same channel, start and settings on a coarser mesh
    (--cpu-channel hexahedra x 6 tetrahedra), its SOR for momentum, k and
    omega and its ILU for p' (rrF 1e-8, itmax 200): its algebraic multigrid
    diverges on this singular periodic p' system in the first iteration.
    Rate reported per full-mesh-equivalent iteration by scaling with the cell
    count (a lower bound of the CPU time)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from flowsolversourcecode_amd import meshgen
    dims = tuple(args.cpu_channel)
    pm, types, pairs, tr = channel_mesh(dims, meshgen)
    mesh = po.OrMesh(pm)
    yc = mesh.cell["yc"]
    wd = np.minimum(yc, 2.0 - yc)
    u0, sst = channel_initial(pm, types, yc, wd)
    sst.update(ksolver=po.SOR, wsolver=po.SOR, klim=0, wlim=0, **CHANNEL_SST)
    z = np.zeros(pm.n_faces - pm.n_internal)
    per = dict(pairs=pairs, translations=tr, beta=1.0, mdot=0.0, calc_beta=0)
    kw = dict(rho=1.0, mu=1.0 / RE_TAU, urf_u=URF_U, urf_p=URF_P, msolver=po.SOR, mrrf=0.1, mitmax=10, psolver=po.ILU,
              prrf=1e-8, pitmax=200, plmg=0, framework="tvd", scheme="sou", lim_mom=1, periodic=per, sst=sst,
              u0=(u0, np.zeros_like(u0), np.zeros_like(u0)))
    scale = mesh.N / full_cells
    if po.ref_simple_available():
        rate, el = _reference_rate(pm, z, z, z, types, args.cpu_iters, **kw)
        return {"value": rate * scale, "unit": "SIMPLE iter/s (full-mesh-equivalent)", "cores": 1,
                "host_cpus": os.cpu_count(), "kind": "reference",
                "sample": f"the reference's SIMPLE + k-omega SST built from its sources on the {dims} x 6 channel "
                          f"({mesh.N} tetrahedra), same physics and start, p' by its ILU to 1e-8 (itmax 200; its "
                          f"algebraic multigrid diverges on this singular periodic system): {args.cpu_iters} "
                          f"iterations in {el:.1f} s ({rate:.3f} it/s), scaled by cell count "
                          f"{mesh.N}/{full_cells}"}
    return None


def learn_probe(args):
    """The network's whole cycle for the MFMA counter runs: a 64^3 perturbed
    cavity (Re 1000) writes 8 snapshots, the network trains 4 epochs on them
    and then gives the initial guess of 4 more p' solves."""
    import flowsolversourcecode_amd as fs
    from flowsolversourcecode_amd import learn
    c = learn.cavity_case(64, 1000.0, 0.15, 2026)
    net = learn.Mpnn.for_case(c, layers=3, seed=7, capacity=8, use_prev=True)
    learn.attach(c, sink=net)
    c.iterate(8)
    learn.attach(c)
    hist = net.train(epochs=4, batch=1, lr=3e-3, seed=0)
    learn.attach(c, warm=net)
    st = c.iterate(4)
    print(json.dumps({"probe": "learn", "cells": c.N, "loss_last_epoch": hist[-1],
                      "pc_iterations_per_solve": st["pc_iterations"] / 4}))


def channel_main(args):
    import flowsolversourcecode_amd as fs
    from flowsolversourcecode_amd import meshgen
    dims = tuple(args.channel)
    t0 = time.perf_counter()
    pm, types, pairs, tr = channel_mesh(dims, meshgen)
    print(f"# mesh {dims} x 6: {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    tuned = dict(alpha=args.ch_alpha, pre=args.ch_pre, post=args.ch_post, omega=args.ch_omega,
                 agg=args.mg_aggregation, psolver=args.ch_psolver, fuse=args.ch_fuse_prolong,
                 maxlevels=args.ch_maxlevels, cycle=args.ch_cycle, csweeps=args.ch_coarse_sweeps)
    t0 = time.perf_counter()
    case, wd, passes = channel_case(args, fs, pm, types, pairs, tr, 1 if args.precision == "fp32amg" else 0,
                                    **tuned)
    case.set_option("warm_previous", args.ch_warm_previous)
    setup_s = time.perf_counter() - t0
    n_cells, n_faces = case.N, case.F
    print(f"# setup {setup_s:.1f} s: N={n_cells} faces={n_faces}", file=sys.stderr, flush=True)
    if args.warmup:
        st = case.iterate(args.warmup)
        print(f"# warmup {args.warmup} its: {st['seconds']:.2f} s pc_its={st['pc_iterations']}",
              file=sys.stderr, flush=True)
    case.kernel_clock_start(["spmv", "fine_sweep"], capacity=16384, stride=args.clock_stride)
    st = case.iterate(args.steps)
    el = st["seconds"]
    value = args.steps / el
    print(f"# timed {args.steps} its: {el:.3f} s, p' its {st['pc_iterations']}, p' s {st['pc_seconds']:.3f}, "
          f"mom s {st['momentum_seconds']:.3f}, k/omega its {st['turbulence_iterations']}, mass {st['mass_l1']:.3e}",
          file=sys.stderr, flush=True)
    clk = {k: case.kernel_clock_read(k) for k in ("spmv", "fine_sweep")}
    case.kernel_clock_start("off")
    roofs = {}
    for site, name in (("spmv", "k_spmv_dot (FP64 ELL y=A*x with a dot-product partial, Krylov loop)"),
                       ("fine_sweep", "k_sweep<1, float, *> (FP32 damped-Jacobi sweeps of the finest AMG level)")):
        c = clk[site]
        if c["launches"]:
            achieved = c["bytes"] / (c["avg_ms"] * 1e-3) / 1e9
            roofs[site] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBPS,
                           "traffic": measured_traffic(n_cells, site, CHANNEL_TRAFFIC_PROFILE), "kernel": name,
                           "bytes_per_launch": c["bytes"], "bytes_per_cell": c["bytes"] / n_cells,
                           "avg_ms": c["avg_ms"], "timed_launches": c["launches"],
                           "sampled_every": args.clock_stride,
                           "share_of_step": c["launches"] * args.clock_stride * c["avg_ms"] * 1e-3 / el}
    # the line's roofline: the kernel with the larger share of the step
    head = max(roofs.values(), key=lambda r: r["share_of_step"]) if roofs else None
    fields = {k: case.get(k) for k in ("u", "v", "w", "p", "k", "omega")}
    head_pc = st["pc_iterations"] / args.steps
    variants = {}
    if args.variants and args.precision == "fp32amg":
        del case
        fields32 = fields
        if args.ch_warm_previous:
            z32, _, _ = channel_case(args, fs, pm, types, pairs, tr, 1, wd=wd, **tuned)
            z32.iterate(args.warmup)
            sz = z32.iterate(args.steps)
            variants["fp32amg_zero_start"] = {
                "value": args.steps / sz["seconds"], "ms_per_step": 1e3 * sz["seconds"] / args.steps,
                "pc_solve_ms": 1e3 * sz["pc_seconds"] / args.steps,
                "pc_iterations_per_solve": sz["pc_iterations"] / args.steps,
                "note": "the headline without the previous-p' guess: every p' solve from zero"}
            fields32 = {k: z32.get(k) for k in fields}
            print(f"# fp32 zero start: {args.steps / sz['seconds']:.2f} it/s, p' its "
                  f"{sz['pc_iterations'] / args.steps:.1f}", file=sys.stderr, flush=True)
            del z32
        plain, _, _ = channel_case(args, fs, pm, types, pairs, tr, 0, 1.0, wd=wd)
        plain.iterate(args.warmup)
        sp = plain.iterate(args.steps)
        err = {k: float(np.abs(plain.get(k) - fields[k]).max() / max(np.abs(plain.get(k)).max(), 1e-300))
               for k in fields}
        variants["fp64_plain"] = {"value": args.steps / sp["seconds"], "ms_per_step": 1e3 * sp["seconds"] / args.steps,
                                  "pc_solve_ms": 1e3 * sp["pc_seconds"] / args.steps,
                                  "pc_iterations_per_solve": sp["pc_iterations"] / args.steps,
                                  "note": "FP64 AMG hierarchy by the reference's greedy agglomeration, 2/2 sweeps, "
                                          "additive correction alpha = 1, same BiCGSTAB"}
        variants["field_error_fp32amg_vs_fp64_plain"] = err
        print(f"# fp64 plain: {args.steps / sp['seconds']:.2f} it/s, field error {err}", file=sys.stderr, flush=True)
        del plain
        tuned64, _, _ = channel_case(args, fs, pm, types, pairs, tr, 0, wd=wd, **tuned)
        tuned64.iterate(args.warmup)
        s6 = tuned64.iterate(args.steps)
        variants["fp64_tuned"] = {"value": args.steps / s6["seconds"], "ms_per_step": 1e3 * s6["seconds"] / args.steps,
                                  "pc_solve_ms": 1e3 * s6["pc_seconds"] / args.steps,
                                  "pc_iterations_per_solve": s6["pc_iterations"] / args.steps,
                                  "note": "the headline's cycle (aggregation, sweeps, alpha, omega) with the hierarchy "
                                          "in FP64, every p' solve from zero: against fp32amg_zero_start it "
                                          "isolates the FP32 gain"}
        err64 = {k: float(np.abs(tuned64.get(k) - fields32[k]).max() / max(np.abs(tuned64.get(k)).max(), 1e-300))
                 for k in fields32}
        variants["field_error_fp32amg_vs_fp64_same_cycle"] = err64
        print(f"# fp64 tuned: {args.steps / s6['seconds']:.2f} it/s, field error {err64}", file=sys.stderr, flush=True)
        del tuned64
        from flowsolversourcecode_amd import learn
        lc, _, _ = channel_case(args, fs, pm, types, pairs, tr, 1, wd=wd, **tuned)
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
            "pc_iterations_per_solve_zero_start": variants.get("fp32amg_zero_start", {}).get(
                "pc_iterations_per_solve", head_pc),
            "train": {"samples": args.learn_samples, "steps": len(hist) * args.learn_samples,
                      "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1], "seconds": t_train},
            "note": "3-layer MPNN (hidden 64, BF16 MFMA) with the previous p' among its inputs, trained on the "
                    "run's own first iterations; timed iterations start after those"}
        print(f"# learned: {args.steps / sl['seconds']:.2f} it/s, p' its {sl['pc_iterations'] / args.steps:.1f} "
              f"(zero start {head_pc:.1f})", file=sys.stderr, flush=True)
        del lc, net
    cpu = None if args.no_cpu else cpu_baseline_channel(args, n_cells)
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
        "config": {"workload": f"RANS channel {dims[0]}x{dims[1]}x{dims[2]} hex graded 1.05 toward both walls x 6 "
                               f"tetrahedra (Kuhn), k-omega SST (low-Re walls), Re_tau={RE_TAU:.0f} by the "
                               f"prescribed pressure gradient tau_w/delta, periodic streamwise, symmetry spanwise, "
                               f"second-order upwind momentum with min/max-limited velocity gradients, upwind k "
                               f"and omega, urf {URF_U}/{URF_P}, start from the law of the wall, "
                               f"p' {args.ch_psolver.upper()}+AMG({args.precision}) to 1e-8",
                   "cells": n_cells, "faces": n_faces, "parallelism": "single GPU",
                   "wall_distance": (f"Poisson method on the device, {passes} passes to 1e-6"
                                     if args.poisson_wall_distance else "min(y, 2 - y) (plane channel)"),
                   "p_solver": args.ch_psolver,
                   "p_initial_guess": {0: "zero", 1: "previous p', scaled"}.get(
                       args.ch_warm_previous, f"projection on the last {args.ch_warm_previous} p' fields, scaled"),
                   "vcycle": {"cycle": "VFW"[args.ch_cycle], "coarse_sweeps": args.ch_coarse_sweeps,
                              "presweeps": args.ch_pre, "postsweeps": args.ch_post, "alpha": args.ch_alpha,
                              "omega": args.ch_omega, "fused_prolongation": bool(args.ch_fuse_prolong),
                              "max_levels": args.ch_maxlevels,
                              "aggregation": ["reference greedy", "pairwise, 1 pass",
                                              "pairwise, 2 passes"][args.mg_aggregation]}},
        "pc_solve_ms": 1e3 * st["pc_seconds"] / args.steps,
        "pc_iterations_per_solve": head_pc,
        "momentum_ms": 1e3 * st["momentum_seconds"] / args.steps,
        "pc_levels": st["pc_levels"],
        "mass_imbalance_l1": st["mass_l1"],
        "roofline": head,
        "roofline_kernels": roofs or None,
        "variants": variants or None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="cells per side")
    ap.add_argument("--precision", choices=["fp32amg", "fp64"], default="fp32amg")
    ap.add_argument("--cpu-n", type=int, default=32, help="cells per side of the CPU baseline's sample")
    ap.add_argument("--cpu-iters", type=int, default=12, help="timed iterations of the reference's CPU sample")
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
    ap.add_argument("--mg-aggregation", type=int, default=0, choices=[0, 1, 2],
                    help="aggregation of the channel's headline hierarchy: 0 the reference's greedy agglomeration, "
                         "1 or 2 pairwise passes per level")
    # the channel's headline p' solve: CG (the p' matrix is symmetric) with an
    # F-cycle, 1/1 damped-Jacobi sweeps (0.9) on every level, coarse
    # corrections over-relaxed by 1.3, the prolongation as a pass of its own.
    # On the 7.86M-tet channel (10 iterations after 3, ms per p' solve): the
    # V-cycle with 2/2 sweeps and alpha 1 76, F with 1/1 sweeps 64, F with
    # alpha 1.2 48 and 1.4 43; from alpha 1.6 the cycle stops being a usable
    # CG preconditioner (solves hit itmax), so 1.3 keeps a margin.  Among V
    # cycles CG beats BiCGSTAB (81 against 91 ms) and every tuned or pairwise
    # variant measured; damping 0.9 and the unfused prolongation each take a
    # further 5-6 %
    ap.add_argument("--ch-psolver", default="cg", choices=["cg", "bicgs", "gmres", "pbcg"])
    ap.add_argument("--ch-alpha", type=float, default=1.3)
    ap.add_argument("--ch-pre", type=int, default=1)
    ap.add_argument("--ch-post", type=int, default=1)
    ap.add_argument("--ch-omega", type=float, default=0.9)
    ap.add_argument("--ch-fuse-prolong", type=int, default=0, choices=[0, 1])
    # each p' solve of the headline starts from the previous iteration's p'
    # (scaled to its best multiple, dropped when worse than zero): 10.05
    # against 9.13 it/s from zero on the 7.86M-tet channel
    ap.add_argument("--ch-coarse-sweeps", type=int, nargs=2, default=[1, 1], metavar=("PRE", "POST"),
                    help="channel p' AMG sweeps on the levels below the finest")
    ap.add_argument("--ch-cycle", type=int, default=1, choices=[0, 1, 2],
                    help="channel p' AMG cycle: 0 V, 1 F, 2 W")
    ap.add_argument("--ch-warm-previous", type=int, default=2, choices=[0, 1, 2, 3, 4])
    # 12 levels bring the coarsest level of the 7.86M-tet hierarchy under the
    # 64 cells of the direct FP64 solve (10 leave about 130 cells to 16 sweeps):
    # 57 instead of 61 CG iterations per solve
    ap.add_argument("--ch-maxlevels", type=int, default=12)
    ap.add_argument("--tet-n", type=int, default=64, help="hexahedra per side of the tetrahedral cavity variant")
    ap.add_argument("--clock-stride", type=int, default=5,
                    help="time every n-th launch of the roofline kernels with HIP events")
    ap.add_argument("--cavity2d-iters", type=int, default=500, help="outer iterations of the 2D-cavity variant")
    ap.add_argument("--cavity2d-pitmax", type=int, default=500, help="p' itmax of the 2D cavity's PBCG runs")
    ap.add_argument("--cavity2d-cpu", type=float, default=6.0, help="seconds of the 2D cavity's port CPU sample")
    ap.add_argument("--cavity2d-cpu-iters", type=int, default=60,
                    help="timed iterations of the 2D cavity's reference CPU sample")
    ap.add_argument("--workload", choices=["channel", "cavity"], default="channel",
                    help="the RANS channel (the headline) or the 128^3 precision-form cavity")
    ap.add_argument("--channel", type=int, nargs=3, default=[64, 128, 160],
                    help="hexahedra of the channel (x streamwise, y wall-normal, z spanwise), each split into 6")
    ap.add_argument("--cpu-channel", type=int, nargs=3, default=[16, 32, 40],
                    help="hexahedra of the CPU baseline's channel sample")
    ap.add_argument("--poisson-wall-distance", action="store_true",
                    help="the channel's wall distance by the device's Poisson method instead of min(y, 2 - y)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, 4 launches of the plain copy kernel over 1 GiB (counter calibration)")
    ap.add_argument("--learn-probe", action="store_true",
                    help="the network's training and inference on a 64^3 cavity (MFMA counter runs) instead")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="set up the workload, run one iteration, then only the ELL product (--spmv-reps launches) "
                         "and a 1 GiB device copy; for hardware-counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding of the mesh is a design "
                                   "note only (no collectives are built)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.learn_probe:
        learn_probe(args)
        return
    if args.workload == "channel" and not args.traffic_probe:
        channel_main(args)
        if args.copy_calibration:
            import flowsolversourcecode_amd as fs
            print(json.dumps({"probe": "copy", "copy_bytes": COPY_BYTES, "copy_ms": fs.time_device_copy(COPY_BYTES, 3)}))
        return

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
                               f"p' {args.ch_psolver.upper()}+AMG({args.precision}) to 1e-8",
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
