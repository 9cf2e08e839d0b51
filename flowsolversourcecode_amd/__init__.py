"""MI355X (gfx950) hot path of the FlowSolver finite-volume code.

Python view of the C ABI in ``include/fsamd.h`` (``libfsamd.so``, built by
``python -m flowsolversourcecode_amd.build``).  The entry points keep the
reference's names and array conventions:

* ``Mesh(elem_nbr, n_nbr, ...)``  -- the topology arrays ``ElementNeighbor(N,W)``
  and ``NumberofElementNeighbors(N)`` (DeclareVariables/Geometry3.F90:34-35),
  passed as a (W, N) C-ordered int32 array, i.e. the Fortran column-major
  memory image, 1-based with 0 for a boundary slot.
* ``Mesh.solve_equation(...)`` -- SUBROUTINE SolveEquation(Variable, FiT, rrF,
  itmax, solver, LMultigrid) (SolveAlegbraicEquations.for:4) with the module
  inputs ac, anb, bc passed explicitly; ``solver`` and ``variable`` take the
  reference's strings.

There is no CPU fallback: without the built library every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfsamd.so")

SOLVERS = {"sor": 0, "ilu": 1, "pbcg": 2, "direct": 3, "gmres": 4, "bicgs": 5}
MG_OFF, MG_SOLVER, MG_PRECOND = 0, 1, 2
# GetNFfromVariable (AssignNFvaluesToVariables.for:20-108)
NF_CODES = {"velx": 1, "vely": 2, "velz": 3, "pressc": 4, "press": 5, "temp": 6, "htotal": 7,
            "tke": 8, "ted": 9, "tomega": 10, "med": 11, "tkl": 12, "tgamma": 13,
            "tretheta": 14, "tv2": 15, "tzeta": 16, "frelax": 17, "lambda": 18}
OPTIONS = {"mg_presweep": 1, "mg_postsweep": 2, "mg_maxcycles": 3, "mg_maxlevels": 4,
           "mg_minparents": 5, "mg_reagglomerate": 6, "mg_fp32": 7, "mg_omega": 8,
           "mg_smoother": 9, "gmres_restart": 10, "check_every": 11}


class FsError(RuntimeError):
    pass


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("levels", C.c_int), ("initial_l1", C.c_double),
                ("final_l1", C.c_double), ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


def lib():
    """The loaded libfsamd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FsError(f"{LIB_PATH} is missing: build it with python -m flowsolversourcecode_amd.build")
        L = C.CDLL(LIB_PATH)
        L.fs_last_error.restype = C.c_char_p
        L.fs_version.restype = C.c_char_p
        L.fs_mesh_create.argtypes = [C.c_int, C.c_int, _ip, _ip, C.c_int, _ip, _ip, _ip, _ip,
                                     C.POINTER(C.c_void_p)]
        L.fs_mesh_destroy.argtypes = [C.c_void_p]
        L.fs_set_option.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.fs_solve_equation.argtypes = [C.c_void_p, C.c_int, _dp, C.c_double, C.c_int, C.c_int, C.c_int,
                                        _dp, _dp, _dp, _dp, _dp, C.POINTER(SolveStats)]
        L.fs_solve_equation_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_int,
                                               C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p, _dp, C.POINTER(SolveStats)]
        L.fs_amg_export.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp]
        _lib = L
    return _lib


def last_error():
    return lib().fs_last_error().decode()


def _iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def _dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _check(rc, what):
    if rc < 0:
        raise FsError(f"{what} failed ({rc}): {last_error()}")
    return rc


class Mesh:
    """Device context for one mesh (fs_mesh_create)."""

    def __init__(self, elem_nbr, n_nbr, owner=None, neighbour=None, i_own_nbr=None, i_nbr_own=None):
        L = lib()
        self.nbr = np.ascontiguousarray(elem_nbr, dtype=np.int32)
        self.w, self.n = self.nbr.shape
        self.nnbr = np.ascontiguousarray(n_nbr, dtype=np.int32)
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.int32)
                for a in (owner, neighbour, i_own_nbr, i_nbr_own)]
        self._keep = arrs
        nif = 0 if arrs[0] is None else len(arrs[0])
        h = C.c_void_p()
        _check(L.fs_mesh_create(self.n, self.w, _iptr(self.nbr), _iptr(self.nnbr), nif,
                                *[_iptr(a) for a in arrs], C.byref(h)), "fs_mesh_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fs_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, key, value):
        k = OPTIONS[key] if isinstance(key, str) else int(key)
        _check(lib().fs_set_option(self.h, k, float(value)), "fs_set_option")

    def solve_equation(self, variable, fit, rrf, itmax, solver, lmultigrid, ac, anb, bc):
        """SolveEquation on host arrays.  Returns (rc, fit_new, dphi, res4, stats)."""
        nf = NF_CODES[variable] if isinstance(variable, str) else int(variable)
        sv = SOLVERS[solver] if isinstance(solver, str) else int(solver)
        fit = np.array(fit, dtype=np.float64, copy=True)
        ac = np.ascontiguousarray(ac, dtype=np.float64)
        anb = np.ascontiguousarray(anb, dtype=np.float64)
        bc = np.ascontiguousarray(bc, dtype=np.float64)
        assert ac.shape == (self.n,) and anb.shape == (self.w, self.n) and bc.shape == (self.n,)
        dphi = np.zeros(self.n)
        res4 = np.zeros(4)
        st = SolveStats()
        rc = _check(lib().fs_solve_equation(self.h, nf, _dptr(fit), float(rrf), int(itmax), sv,
                                            int(lmultigrid), _dptr(ac), _dptr(anb), _dptr(bc),
                                            _dptr(dphi), _dptr(res4), C.byref(st)), "fs_solve_equation")
        return rc, fit, dphi, res4, st.as_dict()

    def solve_equation_device(self, variable, fit, rrf, itmax, solver, lmultigrid, ac, anb, bc, dphi):
        """SolveEquation on device tensors (torch, float64, on the current
        device; anb shaped (W, N)).  fit and dphi are updated in place.
        Returns (rc, res4, stats)."""
        import torch
        nf = NF_CODES[variable] if isinstance(variable, str) else int(variable)
        sv = SOLVERS[solver] if isinstance(solver, str) else int(solver)
        for t in (fit, ac, anb, bc, dphi):
            assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
        assert fit.numel() == self.n and ac.numel() == self.n and bc.numel() == self.n
        assert anb.numel() == self.n * self.w and dphi.numel() == self.n
        torch.cuda.synchronize()
        res4 = np.zeros(4)
        st = SolveStats()
        rc = _check(lib().fs_solve_equation_device(self.h, nf, fit.data_ptr(), float(rrf), int(itmax), sv,
                                                   int(lmultigrid), ac.data_ptr(), anb.data_ptr(),
                                                   bc.data_ptr(), dphi.data_ptr(), _dptr(res4),
                                                   C.byref(st)), "fs_solve_equation_device")
        return rc, res4, st.as_dict()

    def amg_level(self, level):
        """Maps and operators of AMG level `level` (reference numbering)."""
        dims = np.zeros(5, np.int32)
        _check(lib().fs_amg_export(self.h, level, _iptr(dims), None, None, None, None, None, None, None,
                                   None, None), "fs_amg_export")
        n, w, nch, nlev, ncol = (int(x) for x in dims)
        out = {"n": n, "w": w, "levels": nlev, "colors": ncol,
               "parent": np.zeros(n, np.int32), "seed": np.zeros(n, np.int32),
               "nchild": np.zeros(n, np.int32), "child_ptr": np.zeros(n + 1, np.int32),
               "child": np.zeros(max(nch, 1), np.int32), "nbr": np.zeros((w, n), np.int32),
               "nnbr": np.zeros(n, np.int32), "ac": np.zeros(n), "anb": np.zeros((w, n))}
        _check(lib().fs_amg_export(self.h, level, _iptr(dims), _iptr(out["parent"]), _iptr(out["seed"]),
                                   _iptr(out["nchild"]), _iptr(out["child_ptr"]), _iptr(out["child"]),
                                   _iptr(out["nbr"]), _iptr(out["nnbr"]), _dptr(out["ac"]),
                                   _dptr(out["anb"])), "fs_amg_export")
        out["child"] = out["child"][:nch]
        return out
