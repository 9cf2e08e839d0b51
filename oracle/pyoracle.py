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

    def solve(self, bc, ires, rrf, maxcycles=30, npre=3, npost=5):
        dphi = np.zeros(self.n)
        hist = np.zeros(max(maxcycles, 1))
        k = lib().or_amg_solve(self.h, _c(bc, np.float64), ires, rrf, maxcycles, npre, npost, dphi, hist)
        return dphi, hist[:k].copy()
