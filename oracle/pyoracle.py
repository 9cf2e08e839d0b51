"""ctypes view of the C restatement (oracle/build/libfs_oracle.so) and a runner
for the reference build (oracle/_ref/ref_solve.x).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.

Array conventions are the reference's: ``nbr``/``anb`` are (W, N) C-ordered numpy
arrays, i.e. the Fortran column-major ``ElementNeighbor(N,W)`` / ``anb(N,W)``
memory image; neighbour values are 1-based with 0 for a boundary slot.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libfs_oracle.so")
REF_SOLVE = os.path.join(HERE, "_ref", "ref_solve.x")

SOR, ILU, PBCG = 0, 1, 2

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
        L = _lib
        L.or_calculate_residuals.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp]
        L.or_check_convergence.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
        L.or_check_convergence.restype = C.c_double
        L.or_ilu_factor.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.or_solve_equation.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, C.c_double,
                                        C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.or_solve_equation.restype = C.c_int
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def residuals(nbr, nnbr, ac, anb, bc, dphi, fit, nf=4):
    w, n = nbr.shape
    out = np.zeros(4)
    lib().or_calculate_residuals(n, w, _c(nbr, np.int32), _c(nnbr, np.int32), _c(ac, np.float64),
                                 _c(anb, np.float64), _c(bc, np.float64), _c(dphi, np.float64),
                                 _c(fit, np.float64), nf, out)
    return out


def ilu_factor(nbr, nnbr, ac, anb):
    w, n = nbr.shape
    dc = np.zeros(n)
    lib().or_ilu_factor(n, w, _c(nbr, np.int32), _c(nnbr, np.int32), _c(ac, np.float64), _c(anb, np.float64), dc)
    return dc


def solve_equation(nbr, nnbr, ac, anb, bc, fit, rrf, itmax, solver, nf=4):
    """SolveEquation without multigrid. Returns (fit_new, dphi, res[4], history)."""
    w, n = nbr.shape
    fit = _c(fit, np.float64).copy()
    dphi = np.zeros(n)
    res = np.zeros(4)
    hist = np.zeros(max(itmax, 1))
    it = lib().or_solve_equation(n, w, _c(nbr, np.int32), _c(nnbr, np.int32), _c(ac, np.float64),
                                 _c(anb, np.float64), _c(bc, np.float64), fit, rrf, itmax, solver, nf,
                                 dphi, res, hist)
    return fit, dphi, res, hist[:it].copy()


# ----------------------------------------------------------------------------
# Reference build runner (this container only: /root/reference is needed to
# build _ref/ref_solve.x; the GPU box never runs it).


def ref_available():
    return os.path.exists(REF_SOLVE)


def run_reference(nbr, nnbr, ac, anb, bc, fit, rrf, itmax, solver, mode=0, lmg=0,
                  cycle=0, npre=3, npost=5, maxcycles=30, maxlevels=10, minparents=10, ncalls=1):
    """Run the reference's SolveEquation (mode 0) or its per-iteration loop
    (mode 1) on one ELL system.  Returns a list (one per call) of dicts."""
    w, n = nbr.shape
    with tempfile.TemporaryDirectory(dir=os.path.join(HERE, "_ref")) as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([n, w, mode, solver, itmax, lmg, cycle, npre, npost, maxcycles, maxlevels,
                              minparents, ncalls], np.int32).tobytes())
            f.write(np.array([rrf], np.float64).tobytes())
            for a, dt in ((nbr, np.int32), (nnbr, np.int32), (ac, np.float64), (anb, np.float64),
                          (bc, np.float64), (fit, np.float64)):
                f.write(_c(a, dt).tobytes())
        r = subprocess.run([REF_SOLVE, fin, fout], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed: {r.stderr[-2000:]}")
        d = open(fout, "rb").read()
    out, o = [], 0

    def take(fmt_dtype, count):
        nonlocal o
        a = np.frombuffer(d, dtype=fmt_dtype, count=count, offset=o).copy()
        o += a.nbytes
        return a

    for _ in range(ncalls):
        n_, nh = take(np.int32, 2)
        rec = {"dphi": take(np.float64, n), "fit": take(np.float64, n), "res": take(np.float64, 4)}
        rec["hist"] = take(np.float64, max(int(nh), 1))[: int(nh)]
        L = int(take(np.int32, 1)[0])
        if L:
            hdr = take(np.int32, 4 * L).reshape(4, L)
            ne, ije, ijn, nfaces = hdr
            ntot = int(ne.sum())
            parents, nchild, eparent, nnbr_mg = take(np.int32, 4 * ntot).reshape(4, ntot)
            acmg = take(np.float64, ntot)
            sizes = take(np.int32, 3)
            enbr = take(np.int32, int(sizes[0]))
            children = take(np.int32, int(sizes[1]))
            anbmg = take(np.float64, int(sizes[2]))
            rec["mg"] = _unflatten_mg(L, ne, ije, ijn, parents, nchild, eparent, nnbr_mg, acmg, enbr,
                                      children, anbmg)
        out.append(rec)
    return out


REF_CONV = os.path.join(HERE, "_ref", "ref_conv.x")


def ref_conv_available():
    return os.path.exists(REF_CONV)


def _patch_matrix(vals, patch_sizes, nbfmax, dtype):
    """Boundary-face values (patch by patch) as the reference's (NumberOfBCSets,
    NBFacesMax) array, in Fortran memory order."""
    out = np.zeros((len(patch_sizes), nbfmax), dtype)
    o = 0
    for q, n in enumerate(patch_sizes):
        out[q, :n] = vals[o:o + n]
        o += n
    return np.asfortranarray(out)


def _patch_unmatrix(a, patch_sizes):
    return np.concatenate([a[q, :n] for q, n in enumerate(patch_sizes)]) if len(patch_sizes) else a[:0, 0]


def run_reference_conv(pm, mesh, mode, phi, grad, bphi, mdot=None, ftf=None, method=1, scheme="upwind",
                       framework="tvd", bleed=0.0):
    """Run the reference's LimitGradientVariable (mode 1, ``method`` 1 or 2) or
    HRSchemesCorrections for 'velx' (mode 2) on a mesh slice of ``mesh`` (an
    OrMesh of the polyMesh ``pm``).  grad = [gx | gy | gz] (3N); the boundary
    gradients are the owners' (the Green-Gauss boundary copies).  Returns the
    limited (grad, bgrad) for mode 1 and FluxTf (Fi) for mode 2."""
    N, W, Fi = mesh.N, mesh.W, mesh.nif
    sizes = [int(x) for x in pm.patch_sizes]
    nbc, nbfmax = len(sizes), max(sizes)
    own = np.asarray(pm.owner, np.int64)
    bown = own[Fi:]
    grad = np.asarray(grad, np.float64)
    fc = {k: mesh.face[k] for k in ("fcx", "fcy", "fcz")}
    if mdot is None:
        mdot = np.zeros(Fi)
    if ftf is None:
        ftf = np.zeros(Fi)
    with tempfile.TemporaryDirectory(dir=os.path.join(HERE, "_ref")) as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([N, W, Fi, nbc, nbfmax, mode, method], np.int32).tobytes())
            f.write(scheme.ljust(20).encode() + framework.ljust(4).encode())
            f.write(np.array([bleed], np.float64).tobytes())
            f.write(np.array(sizes, np.int32).tobytes())
            f.write(_patch_matrix(bown + 1, sizes, nbfmax, np.int32).tobytes(order="F"))
            f.write(_c(mesh.nef, np.int32).tobytes())
            f.write(_c(mesh.enbr, np.int32).tobytes())
            f.write(_c(mesh.gef, np.int32).tobytes())
            f.write(_c(own[:Fi] + 1, np.int32).tobytes())
            f.write(_c(np.asarray(pm.neighbour, np.int64)[:Fi] + 1, np.int32).tobytes())
            for k in ("xc", "yc", "zc"):
                f.write(_c(mesh.cell[k], np.float64).tobytes())
            for k in ("fcx", "fcy", "fcz"):
                f.write(_c(fc[k][:Fi], np.float64).tobytes())
            for k in ("fcx", "fcy", "fcz"):
                f.write(_patch_matrix(fc[k][Fi:], sizes, nbfmax, np.float64).tobytes(order="F"))
            for k in ("g", "dx", "dy", "dz"):
                f.write(_c(mesh.face[k][:Fi], np.float64).tobytes())
            f.write(_c(mdot, np.float64).tobytes())
            f.write(_c(ftf, np.float64).tobytes())
            f.write(_c(phi, np.float64).tobytes())
            for d in range(3):
                f.write(_c(grad[d * N:(d + 1) * N], np.float64).tobytes())
            f.write(_patch_matrix(bphi, sizes, nbfmax, np.float64).tobytes(order="F"))
            for d in range(3):
                f.write(_patch_matrix(grad[d * N:(d + 1) * N][bown], sizes, nbfmax, np.float64).tobytes(order="F"))
        r = subprocess.run([REF_CONV, fin, fout], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed: {r.stderr[-2000:]}")
        d = np.fromfile(fout, dtype=np.float64)
    if mode == 2:
        return d[:Fi].copy()
    g = d[:3 * N].copy()
    bm = d[3 * N:].reshape(3, nbfmax, nbc)  # three Fortran (nbc, nbfmax) blocks
    bg = np.concatenate([_patch_unmatrix(bm[k].T, sizes) for k in range(3)])
    return g, bg


REF_SIMPLE = os.path.join(HERE, "_ref", "ref_simple.x")

REF_GEOM_FACE = ("Sx", "Sy", "Sz", "A", "fcx", "fcy", "fcz", "nx", "ny", "nz", "dcf", "dx", "dy", "dz", "ex",
                 "ey", "ez", "Ex", "Ey", "Ez", "E", "gdiff", "Tx", "Ty", "Tz", "T")


def ref_simple_available():
    return os.path.exists(REF_SIMPLE)


def _simple_input(f, pm, mesh, mode, extra_ints=(), extra_doubles=(), tail=(), decomp=3, after_doubles=b""):
    """The polyMesh and the Readit topology of ``mesh`` (an OrMesh of ``pm``)
    in ref_simple_driver.f90's input layout."""
    s = mesh.ptr.contents
    N, W, nf, nif = mesh.N, mesh.W, mesh.nf, mesh.nif
    lenw = s.len_w
    fptr = np.asarray(pm.face_ptr, np.int64)
    fnn = np.diff(fptr).astype(np.int32)
    maxfn = int(fnn.max())
    fn = np.zeros((nf, maxfn), np.int32)
    nodes = np.asarray(pm.face_nodes, np.int64)
    for k in range(nf):
        fn[k, :fnn[k]] = nodes[fptr[k]:fptr[k + 1]] + 1
    pts = np.ascontiguousarray(pm.points, np.float64)
    nen = np.ctypeslib.as_array(s.nen, shape=(N,)).astype(np.int32)
    lst = np.ctypeslib.as_array(s.len, shape=(N * lenw,)).astype(np.int32).reshape(lenw, N).T
    lst = np.where(np.arange(lenw)[None, :] < nen[:, None], lst + 1, 0).astype(np.int32)
    f.write(np.array([len(pts), nf, nif, N, len(pm.patch_sizes), W, lenw, maxfn, mode, decomp] + list(extra_ints),
                     np.int32).tobytes())
    if extra_doubles:
        f.write(np.array(extra_doubles, np.float64).tobytes())
    f.write(after_doubles)
    for d in range(3):
        f.write(_c(pts[:, d], np.float64).tobytes())
    f.write(fnn.tobytes())
    f.write(np.asfortranarray(fn).tobytes(order="F"))
    f.write(_c(np.asarray(pm.owner) + 1, np.int32).tobytes())
    f.write(_c(np.asarray(pm.neighbour)[:nif] + 1, np.int32).tobytes())
    f.write(_c(pm.patch_sizes, np.int32).tobytes())
    for a in tail:
        f.write(a)
    f.write(_c(mesh.nef, np.int32).tobytes())
    f.write(_c(mesh.gef, np.int32).tobytes())
    f.write(_c(mesh.enbr, np.int32).tobytes())
    f.write(_c(mesh.iown, np.int32).tobytes())
    f.write(_c(mesh.inbr, np.int32).tobytes())
    f.write(nen.tobytes())
    f.write(np.asfortranarray(lst).tobytes(order="F"))


def run_reference_geometry(pm, mesh, decomp=3):
    """The reference's ProcessGeometry on the polyMesh ``pm`` with the Readit
    topology of ``mesh``: a dict of face fields over all faces (interior,
    then boundary) and cell fields, named as OrMesh names them, plus "g"
    (interior faces)."""
    nf, nif, N = mesh.nf, mesh.nif, mesh.N
    with tempfile.TemporaryDirectory(dir=os.path.join(HERE, "_ref")) as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            _simple_input(f, pm, mesh, 1, tail=(_c(np.zeros(len(pm.patch_sizes)), np.int32).tobytes(),),
                          decomp=decomp)
        r = subprocess.run([REF_SIMPLE, fin, fout], capture_output=True, text=True, cwd=td)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed: {r.stdout[-1000:]} {r.stderr[-2000:]}")
        d = np.fromfile(fout, dtype=np.float64)
    out, o = {}, 0
    for k in REF_GEOM_FACE:
        out[k] = d[o:o + nf].copy()
        o += nf
    out["g"] = d[o:o + nif].copy()
    o += nif
    for k in ("vol", "xc", "yc", "zc"):
        out[k] = d[o:o + N].copy()
        o += N
    return out


# The k-omega SST case (ref_simple_driver.f90 set_turbulence_case): solvers,
# iteration limits, convection (framework, scheme), MethodCalcGradient and
# nIterGradient, gradient limiter (0 off) of k and omega, the wall distance
# (0 the input's wd, 1 or 2 the reference's MethodCalculateNormalDistance);
# urfTKE, urfTOmega, urfTViscosity, rrFTKE, rrFTOmega, MaximumViscosityRatio.
SST_DEFAULTS = dict(ksolver=SOR, kitmax=20, wsolver=SOR, witmax=20, kframework="none", kscheme="upwind",
                    wframework="none", wscheme="upwind", kgrad=(2, 2), wgrad=(2, 2), klim=0, wlim=0, wdmethod=0,
                    urf_k=0.5, urf_w=0.5, urf_mut=0.5, krrf=0.1, wrrf=0.1, max_visc_ratio=1.0e6)


def sst_settings_blob(t):
    ints = (t["ksolver"], t["kitmax"], t["wsolver"], t["witmax"], HR_FRAMEWORKS[t["kframework"]],
            CONV_SCHEMES[t["kscheme"]], HR_FRAMEWORKS[t["wframework"]], CONV_SCHEMES[t["wscheme"]],
            *t["kgrad"], *t["wgrad"], t["klim"], t["wlim"], t["wdmethod"])
    dbl = (t["urf_k"], t["urf_w"], t["urf_mut"], t["krrf"], t["wrrf"], t["max_visc_ratio"])
    return np.array(ints, np.int32).tobytes() + np.array(dbl, np.float64).tobytes()


def run_reference_simple(pm, mesh, wall_u, wall_v, wall_w, patch_types=None, niter=1, rho=1.0, mu=0.01,
                         urf_u=0.7, urf_p=0.3, msolver=SOR, mrrf=0.1, mitmax=20, psolver=SOR, prrf=0.01,
                         pitmax=200, framework="none", scheme="upwind", bleed=0.0, start_iter=0, lim_mom=0,
                         lim_cont=0, plmg=0, algorithm="simple", grad_mom=(2, 2), grad_cont=(2, 2),
                         decomp=3, gi_mom="average", gi_cont="averagecorrected", restart=False, seconds=False,
                         periodic=None, sst=None, u0=None):
    """niter outer iterations of the reference's own SIMPLE (SolveMomentumX/Y/Z,
    SolveContinuity) from rest on the polyMesh ``pm`` with the Readit topology of
    ``mesh``; patch_types per patch "wall"/"symmetry" (default all walls);
    plmg = 1 solves p' with the reference's algebraic multigrid
    (LMultigridContinuity).  Returns a dict u, v, w, p (N) and mdot (Fi), and
    with seconds=True also the run's wall time ("seconds").  periodic: a
    dict pairs (partner patch per patch, -1 none), translations (3 per
    patch), beta (initial periodicBeta), mdot (periodicMdot), calc_beta,
    relax, iters, every (LcalculateBeta, relaxBeta, BetaIterations,
    nIterCorrectBeta); the run then also returns "beta" and "bmdot" (the
    boundary mass fluxes in boundary-face order).  sst: a dict of the k-omega
    SST case (see SST_DEFAULTS) with the initial cell fields k, omega, mut,
    wd and boundary-face fields bk, bomega, bmut; the run then also returns
    k, omega, mut (N), bk, bomega, bmut (boundary faces), F1 and the wall
    distance the run used, wd (N).  u0: an initial velocity, (u, v, w) cell
    arrays (default rest)."""
    nf, nif, N = mesh.nf, mesh.nif, mesh.N
    codes = np.array([PATCH_CODES[t] for t in (patch_types or ["wall"] * len(pm.patch_sizes))], np.int32)
    ints = (msolver, mitmax, psolver, pitmax, niter, HR_FRAMEWORKS[framework], CONV_SCHEMES[scheme], lim_mom,
            lim_cont, start_iter, plmg, {"simple": 0, "simplec": 1}[algorithm], *grad_mom, *grad_cont,
            GRADIENT_INTERPOLATION[gi_mom], GRADIENT_INTERPOLATION[gi_cont], 1 if periodic else 0,
            1 if sst else 0)
    dbl = (rho, mu, urf_u, urf_p, mrrf, prrf, bleed)
    blob = b""
    if periodic:
        blob = (np.asarray(periodic["pairs"], np.int32) + 1).tobytes() + \
            np.asarray(periodic["translations"], np.float64).ravel().tobytes() + \
            np.array([periodic.get("beta", 0.0), periodic["mdot"], periodic.get("relax", 1.0)], np.float64).tobytes() + \
            np.array([int(periodic.get("calc_beta", 1)), periodic.get("iters", 1), periodic.get("every", 1)],
                     np.int32).tobytes()
    if sst:
        t = dict(SST_DEFAULTS, **sst)
        blob += sst_settings_blob(t)
    with tempfile.TemporaryDirectory(dir=os.path.join(HERE, "_ref")) as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            _simple_input(f, pm, mesh, 2, extra_ints=ints, extra_doubles=dbl, tail=(codes.tobytes(),),
                          decomp=decomp, after_doubles=blob)
            for a in (wall_u, wall_v, wall_w):
                f.write(_c(a, np.float64).tobytes())
            if sst:
                for k in ("k", "omega", "mut", "wd", "bk", "bomega", "bmut"):
                    f.write(_c(t[k], np.float64).tobytes())
            f.write(np.array([0 if u0 is None else 1], np.int32).tobytes())
            if u0 is not None:
                for a in u0:
                    f.write(_c(a, np.float64).tobytes())
        t0 = time.perf_counter()
        r = subprocess.run([REF_SIMPLE, fin, fout], capture_output=True, text=True, cwd=td,
                           env=dict(os.environ, OMP_NUM_THREADS="1"))
        el = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed ({r.returncode}): {r.stdout[-1500:]} {r.stderr[-1500:]}")
        d = np.fromfile(fout, dtype=np.float64)
        o = len(REF_GEOM_FACE) * nf + nif + 4 * N
        out = {}
        for k, n in (("u", N), ("v", N), ("w", N), ("p", N), ("mdot", nif)):
            out[k] = d[o:o + n].copy()
            o += n
        if sst:
            for k, n in (("k", N), ("omega", N), ("mut", N), ("bk", nf - nif), ("bomega", nf - nif),
                         ("bmut", nf - nif), ("F1", N), ("wd", N)):
                out[k] = d[o:o + n].copy()
                o += n
        if periodic:
            out["beta"] = float(d[o])
            out["bmdot"] = d[o + 1:o + 1 + nf - nif].copy()
        if restart:
            out["restart"] = open(fout + ".restart", "rb").read()
    if seconds:
        out["seconds"] = el
    return out


def fortran_records(blob):
    """The records of a Fortran unformatted sequential file (4-byte length
    markers before and after each record)."""
    recs, o = [], 0
    while o < len(blob):
        n = int(np.frombuffer(blob, np.int32, 1, o)[0])
        recs.append(blob[o + 4:o + 4 + n])
        assert int(np.frombuffer(blob, np.int32, 1, o + 4 + n)[0]) == n
        o += n + 8
    return recs


def _unflatten_mg(L, ne, ije, ijn, parents, nchild, eparent, nnbr_mg, acmg, enbr, children, anbmg):
    """Turn the reference's flat MultiGrid1 arrays into per-level maps:
    parent[l] (level l cells -> 1-based level l+1 cell), seed/children of level
    l+1 cells (1-based level-l numbers), coarse neighbour lists and ac/anb."""
    levels = []
    for l in range(L):
        nl = int(ne[l])
        e0, n0 = int(ije[l]), int(ijn[l])
        lev = {"n": nl, "nnbr": nnbr_mg[e0:e0 + nl].copy(), "ac": acmg[e0:e0 + nl].copy()}
        nb, an = [], []
        for i in range(nl):
            k = int(lev["nnbr"][i])
            idx = [n0 + i + j * nl for j in range(k)]
            nb.append([int(enbr[t]) for t in idx])
            an.append([float(anbmg[t]) for t in idx])
        lev["nbr"], lev["anb"] = nb, an
        if l < L - 1:
            lev["parent"] = parents[e0:e0 + nl].copy()
        levels.append(lev)
    # seed / children of level l+1 aggregates are stored in level-l slots
    for l in range(L - 1):
        nl = int(ne[l])
        e0, n0 = int(ije[l]), int(ijn[l])
        nc = int(ne[l + 1])
        seeds, kids = [], []
        for p in range(nc):
            s = int(eparent[e0 + p])
            seeds.append(s - e0 if s > 0 else 0)
            k = int(nchild[e0 + p])
            kids.append([int(children[n0 + p + j * nl]) for j in range(k)])
        levels[l + 1]["seed"], levels[l + 1]["children"] = seeds, kids
    return levels


class Amg:
    """Reference-semantics AMG hierarchy (or_amg.c) built on the fine ELL system."""

    def __init__(self, nbr, nnbr, ac, anb, maxlevels=10, minparents=10):
        L = lib()
        L.or_amg_new.restype = C.c_void_p
        L.or_amg_new.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, C.c_int, C.c_int]
        L.or_amg_delete.argtypes = [C.c_void_p]
        L.or_amg_nlevels.argtypes = [C.c_void_p]
        L.or_amg_isolated.argtypes = [C.c_void_p]
        L.or_amg_level_dims.argtypes = [C.c_void_p, C.c_int, _ip]
        L.or_amg_level_copy.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp, _dp, _ip, _ip, _ip, _ip, _ip]
        L.or_amg_solve.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.or_amg_solve.restype = C.c_int
        L.or_amg_solve_cycle.argtypes = [C.c_void_p, C.c_int, _dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                         C.c_int, _dp, _dp]
        L.or_amg_solve_cycle.restype = C.c_int
        L.or_amg_update.argtypes = [C.c_void_p, _dp, _dp]
        w, n = nbr.shape
        self.n = n
        self.h = L.or_amg_new(n, w, _c(nbr, np.int32), _c(nnbr, np.int32), _c(ac, np.float64),
                              _c(anb, np.float64), maxlevels, minparents)
        self.nlevels = L.or_amg_nlevels(self.h)
        self.isolated = L.or_amg_isolated(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_amg_delete(self.h)
            self.h = None

    def level(self, l):
        dims = np.zeros(3, np.int32)
        lib().or_amg_level_dims(self.h, l, dims)
        n, w, nch = (int(x) for x in dims)
        nbr = np.zeros((w, n), np.int32)
        nnbr = np.zeros(n, np.int32)
        ac = np.zeros(n)
        anb = np.zeros((w, n))
        parent = np.zeros(n, np.int32)
        seed = np.zeros(n, np.int32)
        nchild = np.zeros(n, np.int32)
        cptr = np.zeros(n + 1, np.int32)
        child = np.zeros(max(nch, 1), np.int32)
        lib().or_amg_level_copy(self.h, l, nbr, nnbr, ac, anb, parent, seed, nchild, cptr, child)
        return {"n": n, "w": w, "nbr": nbr, "nnbr": nnbr, "ac": ac, "anb": anb, "parent": parent,
                "seed": seed, "nchild": nchild, "child_ptr": cptr, "child": child[:nch]}

    def update(self, ac, anb):
        lib().or_amg_update(self.h, _c(ac, np.float64), _c(anb, np.float64))

    def solve(self, bc, ires, rrf, maxcycles=30, npre=3, npost=5, cycle=0):
        """cycle 0 = V, 1 = F, 2 = W (MultiGridCycleType)."""
        dphi = np.zeros(self.n)
        hist = np.zeros(max(maxcycles, 1))
        k = lib().or_amg_solve_cycle(self.h, int(cycle), _c(bc, np.float64), ires, rrf, maxcycles, npre, npost,
                                     dphi, hist)
        return dphi, hist[:k].copy()


# ----------------------------------------------------------------------------
# mesh topology / geometry and the SIMPLE case (or_geom.c, or_simple.c)

class _OrMesh(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("npoints", "nfaces", "nif", "ncells", "npatches")] + \
               [(n, C.POINTER(C.c_double)) for n in ("x", "y", "z")] + \
               [(n, C.POINTER(C.c_int)) for n in ("fptr", "fnodes", "owner", "neigh", "patch_start")] + \
               [("w", C.c_int)] + \
               [(n, C.POINTER(C.c_int)) for n in ("nef", "gef", "enbr", "iown", "inbr", "nen", "len")] + \
               [("len_w", C.c_int)] + \
               [(n, C.POINTER(C.c_double)) for n in (
                   "Sx", "Sy", "Sz", "A", "fcx", "fcy", "fcz", "nx", "ny", "nz", "vol", "xc", "yc", "zc",
                   "dcf", "dx", "dy", "dz", "ex", "ey", "ez", "g", "Ex", "Ey", "Ez", "E", "gdiff", "Tx", "Ty",
                   "Tz", "T")]


FACE_FIELDS = ("Sx", "Sy", "Sz", "A", "fcx", "fcy", "fcz", "nx", "ny", "nz", "dcf", "dx", "dy", "dz", "ex", "ey",
               "ez", "g", "Ex", "Ey", "Ez", "E", "gdiff", "Tx", "Ty", "Tz", "T")
CELL_FIELDS = ("vol", "xc", "yc", "zc")


class OrMesh:
    """Topology and geometry of a polyMesh computed by the restatement."""

    def __init__(self, pm):
        L = lib()
        L.or_mesh_new.restype = C.POINTER(_OrMesh)
        L.or_mesh_new.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _ip, _ip, _ip, C.c_int, _ip, C.c_int, _ip]
        L.or_mesh_delete.argtypes = [C.POINTER(_OrMesh)]
        pts = np.ascontiguousarray(pm.points)
        self.ptr = L.or_mesh_new(len(pts), _c(pts[:, 0], np.float64), _c(pts[:, 1], np.float64),
                                 _c(pts[:, 2], np.float64), pm.n_faces, _c(pm.face_ptr, np.int32),
                                 _c(pm.face_nodes, np.int32), _c(pm.owner, np.int32), pm.n_internal,
                                 _c(pm.neighbour, np.int32), len(pm.patch_sizes), _c(pm.patch_sizes, np.int32))
        s = self.ptr.contents
        self.N, self.W, self.nf, self.nif = s.ncells, s.w, s.nfaces, s.nif

        def arr(p, n, dt):
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy()

        self.nef = arr(s.nef, self.N, np.int32)
        self.gef = arr(s.gef, self.N * self.W, np.int32).reshape(self.W, self.N)
        self.enbr = arr(s.enbr, self.N * self.W, np.int32).reshape(self.W, self.N)
        self.iown = arr(s.iown, self.nif, np.int32)
        self.inbr = arr(s.inbr, self.nif, np.int32)
        self.face = {k: arr(getattr(s, k), self.nf, np.float64) for k in FACE_FIELDS}
        self.cell = {k: arr(getattr(s, k), self.N, np.float64) for k in CELL_FIELDS}

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().or_mesh_delete(self.ptr)
            self.ptr = None


class _OrCase(C.Structure):
    _fields_ = [("m", C.c_void_p)] + [(n, C.c_int) for n in ("N", "Fi", "Fb", "W")] + \
               [(n, C.POINTER(C.c_double)) for n in ("bu0", "bv0", "bw0")] + \
               [(n, C.c_double) for n in ("rho", "mu", "urf_u", "urf_p")] + \
               [(n, C.c_int) for n in ("msolver", "mitmax", "psolver", "pitmax", "plmg", "ngrad", "refcell")] + \
               [(n, C.c_double) for n in ("mrrf", "prrf")] + \
               [(n, C.POINTER(C.c_double)) for n in (
                   "u", "v", "w", "p", "pc", "mdot", "gu", "gv", "gw", "gp", "gpc", "bu", "bv", "bw", "bp", "bpc",
                   "bgu", "bgv", "bgw", "bgp", "bgpc", "fgu", "fgv", "fgw", "fgp", "du", "dv", "dw", "us", "vs",
                   "ws", "visc", "fcf", "fff", "fvf", "ftf", "fce", "fte", "effdiv", "ac", "anb", "bc", "dphi")] + \
               [("res_mom", C.c_double * 12), ("res_pc", C.c_double * 4), ("mass_l1", C.c_double),
                ("it_mom", C.c_int * 3), ("it_pc", C.c_int), ("btype", C.POINTER(C.c_int)),
                ("bzq", C.POINTER(C.c_int)), ("sym_dirty", C.c_int)] + \
               [(n, C.c_int) for n in ("hr_fw", "hr_scheme", "hr_start", "lim_mom", "lim_cont", "iter")] + \
               [("bleed", C.c_double)] + \
               [(n, C.POINTER(C.c_double)) for n in ("du2", "dv2", "dw2")] + \
               [("ngrad_cont", C.c_int), ("decomp", C.c_int), ("gi_mom", C.c_int), ("gi_cont", C.c_int)] + \
               [(n, C.POINTER(C.c_int)) for n in ("pbf", "pact", "pj1")] + \
               [(n, C.POINTER(C.c_double)) for n in (
                   "pg", "pdx", "pdy", "pdz", "pdcf", "pux", "puy", "puz", "pgd", "ptx", "pty", "ptz", "bmdot")] + \
               [("tuse", C.c_double * 3)] + [(n, C.c_double) for n in ("beta", "beta_mdot", "beta_relax")] + \
               [(n, C.c_int) for n in ("beta_calc", "beta_iters", "beta_every", "nperiodic")] + \
               [("ppair", C.POINTER(C.c_int)), ("ptrans", C.POINTER(C.c_double))] + \
               [("turb", C.c_int)] + \
               [(n, C.POINTER(C.c_double)) for n in (
                   "tk", "tw", "mut", "wd", "btk", "btw", "bmut", "gk", "gom", "bgk", "bgom", "fgk", "fgom",
                   "rhok", "brhok", "grk", "bgrk", "fgrkb", "fgrkc", "F1", "bF1", "F2", "bF2", "prod", "sr", "bsr",
                   "med", "bmed")] + \
               [(n, C.c_int) for n in ("k_solver", "k_itmax", "w_solver", "w_itmax", "k_fw", "k_scheme", "w_fw",
                                       "w_scheme", "k_ngrad", "w_ngrad", "k_lim", "w_lim")] + \
               [(n, C.c_double) for n in ("urf_k", "urf_w", "urf_mut", "k_rrf", "w_rrf", "max_visc_ratio")] + \
               [("it_k", C.c_int), ("it_w", C.c_int), ("res_k", C.c_double * 4), ("res_w", C.c_double * 4)]


PATCH_CODES = {"wall": 0, "symmetry": 1, "periodic": 2}

GRADIENT_INTERPOLATION = {"average": 0, "upwind": 1, "downwind": 2, "averagecorrected": 3, "harmonic": 4}


HR_FRAMEWORKS = {"none": 0, "tvd": 1, "nvf": 2}
CONV_SCHEMES = {name: i for i, name in enumerate(
    ("upwind", "downwind", "fromm", "sou", "cd", "quick", "superbee", "minmod", "osher", "vanleer", "muscl",
     "smart", "hquick", "umist", "charm", "boundedcd", "stoic", "gamma"))}


class OrCase:
    """The SIMPLE restatement on a mesh (or_simple.c)."""

    SIZES = {"u": "N", "v": "N", "w": "N", "p": "N", "pc": "N", "mdot": "Fi", "gu": "3N", "gv": "3N",
             "gw": "3N", "gp": "3N", "gpc": "3N", "bu": "Fb", "bv": "Fb", "bw": "Fb", "bp": "Fb", "bpc": "Fb",
             "bgu": "3Fb", "bgv": "3Fb", "bgw": "3Fb", "bgp": "3Fb", "fgu": "3Fi", "fgv": "3Fi", "fgw": "3Fi",
             "fgp": "3Fi", "du": "N", "dv": "N", "dw": "N", "us": "N", "vs": "N", "ws": "N", "ac": "N",
             "anb": "NW", "bc": "N", "fcf": "F", "fff": "F", "ftf": "F", "fce": "N", "fte": "N", "bmdot": "Fb",
             "bgpc": "3Fb", "gpc": "3N", "tk": "N", "tw": "N", "mut": "N", "wd": "N", "btk": "Fb", "btw": "Fb",
             "bmut": "Fb", "gk": "3N", "gom": "3N", "bgk": "3Fb", "bgom": "3Fb", "F1": "N", "bF1": "Fb", "F2": "N",
             "bF2": "Fb", "prod": "N", "sr": "N", "bsr": "Fb", "rhok": "N", "grk": "3N", "bgrk": "3Fb"}

    def __init__(self, pm, bu, bv, bw, rho=1.0, mu=0.01, urf_u=0.7, urf_p=0.3, msolver=SOR, mrrf=0.1,
                 mitmax=20, psolver=SOR, prrf=0.01, pitmax=200):
        L = lib()
        self.mesh = OrMesh(pm)
        L.or_case_new.restype = C.POINTER(_OrCase)
        L.or_case_new.argtypes = [C.POINTER(_OrMesh), _dp, _dp, _dp] + [C.c_double] * 4 + \
            [C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int]
        L.or_case_delete.argtypes = [C.POINTER(_OrCase)]
        for f in ("or_case_iterate", "or_case_momentum_assemble", "or_case_momentum_solve"):
            getattr(L, f).argtypes = [C.POINTER(_OrCase), C.c_int]
        for f in ("or_case_momentum_gradients", "or_case_pc_assemble", "or_case_pc_solve_correct"):
            getattr(L, f).argtypes = [C.POINTER(_OrCase)]
        L.or_case_mass_imbalance.argtypes = [C.POINTER(_OrCase)]
        L.or_case_mass_imbalance.restype = C.c_double
        self.ptr = L.or_case_new(self.mesh.ptr, _c(bu, np.float64), _c(bv, np.float64), _c(bw, np.float64), rho, mu,
                                 urf_u, urf_p, msolver, mrrf, mitmax, psolver, prrf, pitmax)
        s = self.ptr.contents
        self.dims = {"N": s.N, "Fi": s.Fi, "Fb": s.Fb, "3N": 3 * s.N, "3Fb": 3 * s.Fb, "3Fi": 3 * s.Fi,
                     "NW": s.N * s.W, "F": s.Fi + s.Fb}

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().or_case_delete(self.ptr)
            self.ptr = None

    def view(self, name):
        """numpy view (no copy) of a field of the case."""
        n = self.dims[self.SIZES[name]]
        return np.ctypeslib.as_array(getattr(self.ptr.contents, name), shape=(n,))

    def get(self, name):
        return self.view(name).copy()

    def set(self, name, values):
        self.view(name)[:] = values

    def set_patch_types(self, patch_sizes, patch_types):
        """Boundary condition per patch, in patch order: "wall", "symmetry" or
        "periodic" (then set_periodic)."""
        L = lib()
        L.or_case_set_patch_types.argtypes = [C.POINTER(_OrCase), C.c_int, _ip, _ip]
        codes = np.array([PATCH_CODES[t] for t in patch_types], np.int32)
        sizes = np.ascontiguousarray(patch_sizes, dtype=np.int32)
        if L.or_case_set_patch_types(self.ptr, len(sizes), _c(sizes, np.int32), _c(codes, np.int32)) != 0:
            raise ValueError("or_case_set_patch_types: patches do not fit the boundary")

    def set_periodic(self, pairs, translations, beta=0.0, mdot=0.0, calc_beta=True, relax=1.0, iters=1, every=1):
        """Translational periodic pairs with the implicit coupling and the
        fixed-bulk-flux controls (or_case_set_periodic)."""
        L = lib()
        L.or_case_set_periodic.argtypes = [C.POINTER(_OrCase), C.c_int, _ip, _dp, C.c_double, C.c_int, C.c_double,
                                           C.c_double, C.c_int, C.c_int]
        pairs = np.ascontiguousarray(pairs, dtype=np.int32)
        tr = np.ascontiguousarray(np.asarray(translations, np.float64).ravel())
        if L.or_case_set_periodic(self.ptr, len(pairs), _c(pairs, np.int32), _c(tr, np.float64), beta,
                                  int(calc_beta), mdot, relax, iters, every) != 0:
            raise ValueError("or_case_set_periodic: the periodic patches do not pair")

    @property
    def beta(self):
        return self.ptr.contents.beta

    def set_sst(self, k, omega, mut, wd, bk, bomega, bmut, **opts):
        """The k-omega SST case (or_case_set_sst): initial cell fields k,
        omega, mu_t, the wall distance and the initial boundary-face k, omega,
        mu_t; opts as SST_DEFAULTS (solvers, iteration limits, convection,
        gradient method, limiters, relaxation, tolerances, viscosity cap)."""
        t = dict(SST_DEFAULTS, **opts)
        io = np.array([t["ksolver"], t["kitmax"], t["wsolver"], t["witmax"], HR_FRAMEWORKS[t["kframework"]],
                       CONV_SCHEMES[t["kscheme"]], HR_FRAMEWORKS[t["wframework"]], CONV_SCHEMES[t["wscheme"]],
                       *t["kgrad"], *t["wgrad"], t["klim"], t["wlim"]], np.int32)
        d = np.array([t["urf_k"], t["urf_w"], t["urf_mut"], t["krrf"], t["wrrf"], t["max_visc_ratio"]], np.float64)
        L = lib()
        L.or_case_set_sst.argtypes = [C.POINTER(_OrCase), _ip, _dp] + [_dp] * 7
        if L.or_case_set_sst(self.ptr, _c(io, np.int32), _c(d, np.float64), *(_c(a, np.float64) for a in
                                                                          (k, omega, mut, wd, bk, bomega, bmut))) != 0:
            raise ValueError("or_case_set_sst: invalid option")
        L.or_case_sst.argtypes = [C.POINTER(_OrCase)]

    def wall_distance(self, tol=1e-10, maxouter=10000, solver=ILU, rrf=1e-7, itmax=10):
        """The Poisson wall distance of the case's no-slip walls
        (or_case_wall_distance): (wd, phi, outer passes)."""
        L = lib()
        L.or_case_wall_distance.argtypes = [C.POINTER(_OrCase), C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                            _dp, _dp]
        n = self.dims["N"]
        wd, phi = np.zeros(n), np.zeros(n)
        it = L.or_case_wall_distance(self.ptr, tol, maxouter, solver, rrf, itmax, wd, phi)
        if it < 0:
            raise RuntimeError("or_case_wall_distance did not converge")
        return wd, phi, it

    def set_convection(self, framework, scheme, bleed=0.0, start_iter=0):
        """HRFramework ('none', 'tvd', 'nvf'), ConvectionScheme, Bleed and
        nIterStartApplyingHR of the momentum equations."""
        L = lib()
        L.or_case_set_convection.argtypes = [C.POINTER(_OrCase), C.c_int, C.c_int, C.c_double, C.c_int]
        if L.or_case_set_convection(self.ptr, HR_FRAMEWORKS[framework], CONV_SCHEMES[scheme], bleed,
                                    start_iter) != 0:
            raise ValueError(f"convection scheme {scheme!r} not defined in framework {framework!r}")

    def set_gradient_limiters(self, momentum=0, continuity=0):
        """LimitGradientMomentumMethod / LimitGradientContinuityMethod (0 = off)."""
        L = lib()
        L.or_case_set_gradient_limiters.argtypes = [C.POINTER(_OrCase), C.c_int, C.c_int]
        if L.or_case_set_gradient_limiters(self.ptr, momentum, continuity) != 0:
            raise ValueError("limiter method must be 0, 1 or 2")

    def set_gradient_method(self, momentum=(2, 2), continuity=(2, 2)):
        """(MethodCalcGradient, nIterGradient) for u, v, w and for p, p'."""
        L = lib()
        L.or_case_set_gradient_method.argtypes = [C.POINTER(_OrCase)] + [C.c_int] * 4
        if L.or_case_set_gradient_method(self.ptr, *momentum, *continuity) != 0:
            raise ValueError("gradient method must be 1 or 2 with iterations >= 0")

    def set_gradient_interpolation(self, momentum="average", continuity="averagecorrected"):
        """GradientInterpolationSchemeMomentum / Continuity."""
        L = lib()
        L.or_case_set_gradient_interpolation.argtypes = [C.POINTER(_OrCase), C.c_int, C.c_int]
        if L.or_case_set_gradient_interpolation(self.ptr, GRADIENT_INTERPOLATION[momentum],
                                                GRADIENT_INTERPOLATION[continuity]) != 0:
            raise ValueError((momentum, continuity))

    def set_decompose(self, method):
        """MethodDecomposeS 1, 2 or 3 (recomputes the mesh's E/T split)."""
        L = lib()
        L.or_case_set_decompose.argtypes = [C.POINTER(_OrCase), C.c_int]
        if L.or_case_set_decompose(self.ptr, method) != 0:
            raise ValueError(method)
        s = self.mesh.ptr.contents
        for k in ("Ex", "Ey", "Ez", "E", "gdiff", "Tx", "Ty", "Tz", "T"):
            self.mesh.face[k] = np.ctypeslib.as_array(getattr(s, k), shape=(self.mesh.nf,)).copy()

    def set_algorithm(self, algorithm):
        """Algorithm: "simple" or "simplec"."""
        L = lib()
        L.or_case_set_algorithm.argtypes = [C.POINTER(_OrCase), C.c_int]
        if L.or_case_set_algorithm(self.ptr, {"simple": 0, "simplec": 1}[algorithm]) != 0:
            raise ValueError(algorithm)

    def limit_gradient(self, method, phi, bphi, grad):
        """or_limit_gradient on copies: returns the limited (grad, bgrad)."""
        L = lib()
        L.or_limit_gradient.argtypes = [C.POINTER(_OrCase), C.c_int, _dp, _dp, _dp, _dp]
        g = np.array(grad, np.float64)
        bg = np.zeros(3 * self.dims["Fb"])
        L.or_limit_gradient(self.ptr, method, _c(phi, np.float64), _c(bphi, np.float64), g, bg)
        return g, bg

    def hr_correction(self, phi, grad, ftf):
        """or_hr_correction with the case's mdot and convection settings on a
        copy of ftf (Fi or more entries); returns it."""
        L = lib()
        L.or_hr_correction.argtypes = [C.POINTER(_OrCase), _dp, _dp, _dp]
        out = np.array(ftf, np.float64)
        L.or_hr_correction(self.ptr, _c(phi, np.float64), _c(grad, np.float64), out)
        return out

    def update_symmetry(self):
        lib().or_case_update_symmetry.argtypes = [C.POINTER(_OrCase)]
        lib().or_case_update_symmetry(self.ptr)

    def iterate(self, n):
        lib().or_case_iterate(self.ptr, n)

    def momentum_gradients(self):
        lib().or_case_momentum_gradients(self.ptr)

    def momentum_assemble(self, comp):
        lib().or_case_momentum_assemble(self.ptr, comp)

    def momentum_solve(self, comp):
        lib().or_case_momentum_solve(self.ptr, comp)

    def pc_assemble(self):
        lib().or_case_pc_assemble(self.ptr)

    def pc_solve_correct(self):
        lib().or_case_pc_solve_correct(self.ptr)

    def mass_imbalance(self):
        return lib().or_case_mass_imbalance(self.ptr)
