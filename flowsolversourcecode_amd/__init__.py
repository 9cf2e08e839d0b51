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

SOLVERS = {"sor": 0, "ilu": 1, "pbcg": 2, "direct": 3, "gmres": 4, "bicgs": 5, "cg": 6}
MG_OFF, MG_SOLVER, MG_PRECOND = 0, 1, 2
# GetNFfromVariable (AssignNFvaluesToVariables.for:20-108)
NF_CODES = {"velx": 1, "vely": 2, "velz": 3, "pressc": 4, "press": 5, "temp": 6, "htotal": 7,
            "tke": 8, "ted": 9, "tomega": 10, "med": 11, "tkl": 12, "tgamma": 13,
            "tretheta": 14, "tv2": 15, "tzeta": 16, "frelax": 17, "lambda": 18}
OPTIONS = {"mg_presweep": 1, "mg_postsweep": 2, "mg_maxcycles": 3, "mg_maxlevels": 4,
           "mg_minparents": 5, "mg_reagglomerate": 6, "mg_fp32": 7, "mg_omega": 8,
           "mg_smoother": 9, "gmres_restart": 10, "check_every": 11,
           "mg_tail_cells": 12, "mg_coarse_sweeps": 13, "mg_alpha": 14, "mg_coarse_presweep": 15,
           "mg_coarse_postsweep": 16, "mg_sell_rows": 17, "mg_aggregation": 18,
           "mg_fuse_prolong": 19, "warm_previous": 20, "mg_cycle": 21}


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
        L.fs_solve_equation_device_from_guess.argtypes = L.fs_solve_equation_device.argtypes
        L.fs_amg_export.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp]
        L.fs_time_device_copy.argtypes = [C.c_longlong, C.c_int, _dp]
        L.fs_kernel_clock_start.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.fs_kernel_clock_read.argtypes = [C.c_void_p, C.c_int, _dp]
        L.fs_kernel_clock_clear.argtypes = [C.c_void_p]
        L.fs_kernel_clock_stride.argtypes = [C.c_void_p, C.c_int]
        _lib = L
    return _lib


def last_error():
    return lib().fs_last_error().decode()


CLOCK_SITES = {"off": 0, "spmv": 1, "fine_sweep": 2}


def _clock_start(mesh_handle, sites, capacity, stride=1):
    _check(lib().fs_kernel_clock_stride(mesh_handle, int(stride)), "fs_kernel_clock_stride")
    mask = 0
    for st in ([sites] if isinstance(sites, (str, int)) else sites):
        mask |= 1 << (CLOCK_SITES[st] if isinstance(st, str) else int(st))
    _check(lib().fs_kernel_clock_start(mesh_handle, mask & ~1, int(capacity)), "fs_kernel_clock_start")
    _check(lib().fs_kernel_clock_clear(mesh_handle), "fs_kernel_clock_clear")


def _clock_read(mesh_handle, site):
    out = np.zeros(4)
    st = CLOCK_SITES[site] if isinstance(site, str) else int(site)
    _check(lib().fs_kernel_clock_read(mesh_handle, st, _dptr(out)), "fs_kernel_clock_read")
    return {"launches": int(out[0]), "avg_ms": float(out[1]), "bytes": float(out[2]), "untimed": int(out[3])}


def read_polymesh(directory):
    """Reads an OpenFOAM polyMesh directory (ASCII) with the library's reader
    (fs_polymesh_read) into a meshgen.PolyMesh."""
    from .meshgen import PolyMesh
    L = lib()
    L.fs_polymesh_read.argtypes = [C.c_char_p, _ip, C.POINTER(C.c_void_p)]
    L.fs_polymesh_get.argtypes = [C.c_void_p, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _ip, _ip, C.c_char_p, C.c_int]
    L.fs_polymesh_destroy.argtypes = [C.c_void_p]
    counts = np.zeros(5, np.int32)
    h = C.c_void_p()
    _check(L.fs_polymesh_read(str(directory).encode(), _iptr(counts), C.byref(h)), "fs_polymesh_read")
    try:
        npts, nf, nfn, nint, npat = (int(v) for v in counts)
        x, y, z = (np.zeros(npts) for _ in range(3))
        fp = np.zeros(nf + 1, np.int32)
        fn = np.zeros(nfn, np.int32)
        ow = np.zeros(nf, np.int32)
        nb = np.zeros(nint, np.int32)
        ps = np.zeros(npat, np.int32)
        pn = np.zeros(npat, np.int32)
        names = C.create_string_buffer(64 * max(npat, 1))
        _check(L.fs_polymesh_get(h, _dptr(x), _dptr(y), _dptr(z), _iptr(fp), _iptr(fn), _iptr(ow), _iptr(nb),
                                 _iptr(ps), _iptr(pn), names, 64), "fs_polymesh_get")
    finally:
        L.fs_polymesh_destroy(h)
    patch_names = [names.raw[64 * k:64 * (k + 1)].split(b"\0")[0].decode() for k in range(npat)]
    return PolyMesh(np.stack([x, y, z], 1), fp, fn, ow, nb, patch_names, pn, None)


def time_device_copy(nbytes, reps=3):
    """Average ms of one device-to-device copy of nbytes (streaming kernel)."""
    ms = np.zeros(1)
    _check(lib().fs_time_device_copy(int(nbytes), int(reps), _dptr(ms)), "fs_time_device_copy")
    return float(ms[0])


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

    def kernel_clock_start(self, sites, capacity=4096, stride=1):
        """Arm per-launch HIP-event timing of solve-loop kernel sites ("spmv",
        "fine_sweep"), every stride-th launch of each site."""
        _clock_start(self.h, sites, capacity, stride)

    def kernel_clock_read(self, site):
        """{launches, avg_ms, bytes, untimed} of one site since the clock was started."""
        return _clock_read(self.h, site)

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

    def solve_equation_device(self, variable, fit, rrf, itmax, solver, lmultigrid, ac, anb, bc, dphi,
                              from_guess=False):
        """SolveEquation on device tensors (torch, float64, on the current
        device; anb shaped (W, N)).  fit and dphi are updated in place.  With
        from_guess the solve starts from the guess in dphi
        (fs_solve_equation_device_from_guess).  Returns (rc, res4, stats)."""
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
        fn = lib().fs_solve_equation_device_from_guess if from_guess else lib().fs_solve_equation_device
        rc = _check(fn(self.h, nf, fit.data_ptr(), float(rrf), int(itmax), sv, int(lmultigrid), ac.data_ptr(),
                       anb.data_ptr(), bc.data_ptr(), dphi.data_ptr(), _dptr(res4), C.byref(st)),
                    "fs_solve_equation_device")
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


# ----------------------------------------------------------------------------
# topology / geometry (host) and the device SIMPLE case (include/fsamd_simple.h)

class SimpleStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("momentum_iterations", C.c_int), ("pc_iterations", C.c_int),
                ("pc_unconverged", C.c_int), ("pc_levels", C.c_int), ("seconds", C.c_double),
                ("momentum_seconds", C.c_double), ("pc_seconds", C.c_double), ("pc_initial_l1", C.c_double),
                ("pc_final_l1", C.c_double), ("res_momentum", C.c_double * 3), ("mass_l1", C.c_double),
                ("turbulence_iterations", C.c_int)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "res_momentum"}
        d["res_momentum"] = list(self.res_momentum)
        return d


PHASES = {"momentum_gradients": 0, "assemble_u": 1, "assemble_v": 2, "assemble_w": 3, "solve_u": 4,
          "solve_v": 5, "solve_w": 6, "pc_assemble": 7, "pc_solve_correct": 8, "update_symmetry": 9,
          "turbulence": 10}
TOPOLOGY = {"nef": 0, "gef": 1, "enbr": 2, "iown": 3, "inbr": 4}


def _mesh_args(pm):
    pts = np.ascontiguousarray(pm.points, dtype=np.float64)
    keep = [np.ascontiguousarray(pts[:, k]) for k in range(3)]
    keep += [np.ascontiguousarray(a, dtype=np.int32) for a in (pm.face_ptr, pm.face_nodes, pm.owner, pm.neighbour)]
    return keep


def _setup_case_symbols(L):
    if getattr(L, "_case_ready", False):
        return
    L.fs_geom_create.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _ip, _ip, _ip, C.c_int, _ip,
                                 C.POINTER(C.c_void_p)]
    L.fs_geom_destroy.argtypes = [C.c_void_p]
    L.fs_geom_dims.argtypes = [C.c_void_p, _ip]
    L.fs_geom_get.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_longlong]
    L.fs_geom_topology.argtypes = [C.c_void_p, C.c_int, _ip, C.c_longlong]
    L.fs_case_create.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _ip, _ip, _ip, C.c_int, _ip, _dp, _dp, _dp,
                                 C.c_double, C.c_double, C.POINTER(C.c_void_p)]
    L.fs_case_destroy.argtypes = [C.c_void_p]
    L.fs_case_configure.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                    C.c_int, C.c_double, C.c_int, C.c_int]
    L.fs_case_mesh.argtypes = [C.c_void_p]
    L.fs_case_mesh.restype = C.c_void_p
    L.fs_case_iterate.argtypes = [C.c_void_p, C.c_int, C.POINTER(SimpleStats)]
    L.fs_case_phase.argtypes = [C.c_void_p, C.c_int]
    L.fs_case_get.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_longlong]
    L.fs_case_set.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_longlong]
    L.fs_case_dims.argtypes = [C.c_void_p, _ip]
    L.fs_case_topology.argtypes = [C.c_void_p, C.c_int, _ip, C.c_longlong]
    L.fs_case_set_patch_types.argtypes = [C.c_void_p, C.c_int, _ip, _ip]
    L.fs_case_set_convection.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int]
    L.fs_case_set_gradient_limiters.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.fs_case_set_algorithm.argtypes = [C.c_void_p, C.c_int]
    L.fs_case_set_gradient_method.argtypes = [C.c_void_p] + [C.c_int] * 4
    L.fs_case_set_decompose.argtypes = [C.c_void_p, C.c_int]
    L.fs_case_set_gradient_interpolation.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.fs_case_set_restart_capture.argtypes = [C.c_void_p, C.c_int]
    L.fs_case_write_solution.argtypes = [C.c_void_p, C.c_char_p, C.c_int, _ip]
    L.fs_case_read_solution.argtypes = [C.c_void_p, C.c_char_p, C.c_int, _ip]
    L.fs_case_time_spmv.argtypes = [C.c_void_p, C.c_int, _dp]
    L._case_ready = True


class Geometry:
    """Host-side topology and geometry of a polyMesh (fs_geom_*)."""

    def __init__(self, pm):
        L = lib()
        _setup_case_symbols(L)
        x, y, z, fp, fn, ow, nb = _mesh_args(pm)
        h = C.c_void_p()
        _check(L.fs_geom_create(len(x), _dptr(x), _dptr(y), _dptr(z), pm.n_faces, _iptr(fp), _iptr(fn), _iptr(ow),
                                pm.n_internal, _iptr(nb), C.byref(h)), "fs_geom_create")
        self.h = h
        d = np.zeros(4, np.int32)
        L.fs_geom_dims(h, _iptr(d))
        self.N, self.W, self.F, self.Fi = (int(v) for v in d)

    def __del__(self):
        if getattr(self, "h", None):
            lib().fs_geom_destroy(self.h)
            self.h = None

    def get(self, name):
        n = self.N if name in ("vol", "xc", "yc", "zc") else self.F
        out = np.zeros(n)
        _check(lib().fs_geom_get(self.h, name.encode(), _dptr(out), n), f"fs_geom_get({name})")
        return out

    def topology(self, name):
        n = {"nef": self.N, "gef": self.N * self.W, "enbr": self.N * self.W, "iown": self.Fi, "inbr": self.Fi}[name]
        out = np.zeros(n, np.int32)
        _check(lib().fs_geom_topology(self.h, TOPOLOGY[name], _iptr(out), n), f"fs_geom_topology({name})")
        return out.reshape(self.W, self.N) if name in ("gef", "enbr") else out


BC_TYPES = {"wall": 0, "symmetry": 1, "periodic": 2}
HR_FRAMEWORKS = {"none": 0, "tvd": 1, "nvf": 2}
CONV_SCHEMES = {name: i for i, name in enumerate(
    ("upwind", "downwind", "fromm", "sou", "cd", "quick", "superbee", "minmod", "osher", "vanleer", "muscl",
     "smart", "hquick", "umist", "charm", "boundedcd", "stoic", "gamma"))}


class Case:
    """Device-resident SIMPLE case (fs_case_*) on a polyMesh bounded by no-slip
    walls and symmetry planes."""

    SIZES = {"u": "N", "v": "N", "w": "N", "p": "N", "pc": "N", "mdot": "Fi", "gu": "3N", "gv": "3N",
             "gw": "3N", "gp": "3N", "gpc": "3N", "bu": "Fb", "bv": "Fb", "bw": "Fb", "bp": "Fb", "bgp": "3Fb", "du": "N", "dv": "N", "dw": "N", "du2": "N", "dv2": "N", "dw2": "N", "us": "N", "vs": "N", "ws": "N", "ac": "N",
             "anb": "NW", "bc": "N", "cf": "F", "ff": "F", "tf": "F", "vol": "N", "Sx": "F", "Sy": "F",
             "Sz": "F", "g": "F", "gdiff": "F", "Tx": "F", "Ty": "F", "Tz": "F", "dcf": "F", "A": "F",
             "kx": "Fi", "ky": "Fi", "kz": "Fi", "bmdot": "Fb", "beta": "1", "k": "N", "omega": "N", "mut": "N",
             "wd": "N", "bk": "Fb", "bomega": "Fb", "bmut": "Fb", "F1": "N", "F2": "N", "grk": "3N"}

    def __init__(self, pm, wall_u, wall_v, wall_w, rho=1.0, mu=0.01):
        L = lib()
        _setup_case_symbols(L)
        x, y, z, fp, fn, ow, nb = _mesh_args(pm)
        bu, bv, bw = (np.ascontiguousarray(a, dtype=np.float64) for a in (wall_u, wall_v, wall_w))
        h = C.c_void_p()
        _check(L.fs_case_create(len(x), _dptr(x), _dptr(y), _dptr(z), pm.n_faces, _iptr(fp), _iptr(fn), _iptr(ow),
                                pm.n_internal, _iptr(nb), _dptr(bu), _dptr(bv), _dptr(bw), float(rho), float(mu),
                                C.byref(h)), "fs_case_create")
        self.h = h
        d = np.zeros(4, np.int32)
        L.fs_case_dims(h, _iptr(d))
        N, W, F, Fi = (int(v) for v in d)
        self.N, self.W, self.F, self.Fi, self.Fb = N, W, F, Fi, F - Fi
        self.dims = {"N": N, "3N": 3 * N, "Fi": Fi, "12Fi": 12 * Fi, "Fb": F - Fi, "3Fb": 3 * (F - Fi),
                     "F": F, "NW": N * W, "1": 1}

    def close(self):
        if getattr(self, "h", None):
            lib().fs_case_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, urf_u=0.7, urf_p=0.3, msolver="sor", mrrf=0.1, mitmax=20, mlmg=0, psolver="bicgs",
                  prrf=1e-8, pitmax=500, plmg=MG_PRECOND):
        ms = SOLVERS[msolver] if isinstance(msolver, str) else int(msolver)
        ps = SOLVERS[psolver] if isinstance(psolver, str) else int(psolver)
        _check(lib().fs_case_configure(self.h, urf_u, urf_p, ms, mrrf, int(mitmax), int(mlmg), ps, prrf,
                                       int(pitmax), int(plmg)), "fs_case_configure")

    def set_option(self, key, value):
        k = OPTIONS[key] if isinstance(key, str) else int(key)
        _check(lib().fs_set_option(lib().fs_case_mesh(self.h), k, float(value)), "fs_set_option")

    def kernel_clock_start(self, sites, capacity=4096, stride=1):
        """Arm per-launch HIP-event timing of solve-loop kernel sites ("spmv",
        "fine_sweep"), every stride-th launch of each site."""
        _clock_start(lib().fs_case_mesh(self.h), sites, capacity, stride)

    def kernel_clock_read(self, site):
        """{launches, avg_ms, bytes, untimed} of one site since the clock was started."""
        return _clock_read(lib().fs_case_mesh(self.h), site)

    def iterate(self, n):
        st = SimpleStats()
        _check(lib().fs_case_iterate(self.h, int(n), C.byref(st)), "fs_case_iterate")
        return st.as_dict()

    def time_spmv(self, reps=50):
        """Average ms of one ELL product with the last assembled system."""
        ms = np.zeros(1)
        _check(lib().fs_case_time_spmv(self.h, int(reps), _dptr(ms)), "fs_case_time_spmv")
        return float(ms[0])

    def phase(self, name):
        _check(lib().fs_case_phase(self.h, PHASES[name] if isinstance(name, str) else int(name)), "fs_case_phase")

    def get(self, name):
        n = self.dims[self.SIZES[name]]
        out = np.zeros(n)
        _check(lib().fs_case_get(self.h, name.encode(), _dptr(out), n), f"fs_case_get({name})")
        return out

    def set(self, name, values):
        n = self.dims[self.SIZES[name]]
        a = np.ascontiguousarray(values, dtype=np.float64)
        assert a.size == n
        _check(lib().fs_case_set(self.h, name.encode(), _dptr(a), n), f"fs_case_set({name})")

    def set_patch_types(self, patch_sizes, patch_types):
        """Boundary condition per patch, the patches in face order: "wall"
        (no-slip, the velocity given at creation) or "symmetry"."""
        sizes = np.ascontiguousarray(patch_sizes, dtype=np.int32)
        codes = np.array([BC_TYPES[t] if isinstance(t, str) else int(t) for t in patch_types], np.int32)
        _check(lib().fs_case_set_patch_types(self.h, len(sizes), _iptr(sizes), _iptr(codes)),
               "fs_case_set_patch_types")

    def set_periodic(self, patch_sizes, pairs, translations, beta=0.0, calc_beta=False, mdot=0.0, relax=1.0,
                     iters=1, every=1):
        """Translational periodic pairs with the implicit coupling on the
        patches set_patch_types marked "periodic" (fs_case_set_periodic):
        pairs[q] the partner patch (-1 none), translations[q] its (x, y, z)
        translation; beta the pressure gradient along the translation
        (periodicBeta), updated towards the outward mass flow `mdot` of the
        lower periodic patch when calc_beta."""
        L = lib()
        L.fs_case_set_periodic.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp, C.c_double, C.c_int, C.c_double,
                                           C.c_double, C.c_int, C.c_int]
        sizes = np.ascontiguousarray(patch_sizes, dtype=np.int32)
        pr = np.ascontiguousarray(pairs, dtype=np.int32)
        tr = np.ascontiguousarray(np.asarray(translations, np.float64).ravel())
        _check(L.fs_case_set_periodic(self.h, len(sizes), _iptr(sizes), _iptr(pr), _dptr(tr), float(beta),
                                      int(bool(calc_beta)), float(mdot), float(relax), int(iters), int(every)),
               "fs_case_set_periodic")

    def set_sst(self, k, omega, mut, wd, bk, bomega, bmut, ksolver="ilu", kitmax=20, wsolver="ilu", witmax=20,
                kframework="none", kscheme="upwind", wframework="none", wscheme="upwind", kgrad=(2, 2),
                wgrad=(2, 2), klim=0, wlim=0, urf_k=0.5, urf_w=0.5, urf_mut=0.5, krrf=0.1, wrrf=0.1,
                max_visc_ratio=1.0e6):
        """The k-omega SST model with low-Reynolds-number walls
        (fs_case_set_sst): initial cell fields k, omega, mu_t and the wall
        distance, boundary-face k, omega, mu_t; the k and omega solvers,
        iteration limits, convection, gradient methods and limiters,
        relaxation factors (urfTKE, urfTOmega, urfTViscosity), tolerances and
        MaximumViscosityRatio."""
        L = lib()
        L.fs_case_set_sst.argtypes = [C.c_void_p, _ip, _dp] + [_dp] * 7
        sv = lambda x: SOLVERS[x] if isinstance(x, str) else int(x)
        io = np.array([sv(ksolver), kitmax, sv(wsolver), witmax, HR_FRAMEWORKS[kframework], CONV_SCHEMES[kscheme],
                       HR_FRAMEWORKS[wframework], CONV_SCHEMES[wscheme], *kgrad, *wgrad, klim, wlim], np.int32)
        d = np.array([urf_k, urf_w, urf_mut, krrf, wrrf, max_visc_ratio], np.float64)
        n, nb = self.dims["N"], self.dims["Fb"]
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (k, omega, mut, wd, bk, bomega, bmut)]
        for a, want in zip(arrs, (n, n, n, n, nb, nb, nb)):
            assert a is None or a.shape == (want,)
        _check(L.fs_case_set_sst(self.h, _iptr(io), _dptr(d), *(None if a is None else _dptr(a) for a in arrs)),
               "fs_case_set_sst")

    def wall_distance(self, tol=1e-10, maxouter=2000):
        """The Poisson wall distance of the no-slip walls on the device
        (fs_case_wall_distance); kept for set_sst(wd=None).  Returns (wd,
        passes); raises when the passes do not reach tol."""
        L = lib()
        L.fs_case_wall_distance.argtypes = [C.c_void_p, C.c_double, C.c_int, _dp, C.POINTER(C.c_int)]
        wd = np.zeros(self.dims["N"])
        passes = C.c_int(0)
        rc = _check(L.fs_case_wall_distance(self.h, float(tol), int(maxouter), _dptr(wd), C.byref(passes)),
                    "fs_case_wall_distance")
        if rc != 0:
            raise FsError(f"fs_case_wall_distance: {maxouter} passes did not reach {tol}")
        return wd, passes.value

    def set_convection(self, framework, scheme, bleed=0.0, start_iter=0):
        """HRFrameworkMomentum ("none", "tvd", "nvf"), ConvectionSchemeMomentum
        (e.g. "sou", "minmod"), BleedMomentum and nIterStartApplyingHR."""
        _check(lib().fs_case_set_convection(self.h, HR_FRAMEWORKS[framework], CONV_SCHEMES[scheme], bleed,
                                            start_iter), "fs_case_set_convection")

    def set_gradient_method(self, momentum=(2, 2), continuity=(2, 2)):
        """(MethodCalcGradient, nIterGradient) of u, v, w and of p, p'."""
        _check(lib().fs_case_set_gradient_method(self.h, *momentum, *continuity), "fs_case_set_gradient_method")

    def set_gradient_interpolation(self, momentum="average", continuity="averagecorrected"):
        """GradientInterpolationSchemeMomentum / Continuity: "average", "upwind",
        "downwind", "averagecorrected" or "harmonic"."""
        codes = {"average": 0, "upwind": 1, "downwind": 2, "averagecorrected": 3, "harmonic": 4}
        _check(lib().fs_case_set_gradient_interpolation(self.h, codes[momentum], codes[continuity]),
               "fs_case_set_gradient_interpolation")

    def set_restart_capture(self, on=True):
        """Keep the state the restart file holds beyond the fields (on for
        the iteration before write_solution)."""
        _check(lib().fs_case_set_restart_capture(self.h, 1 if on else 0), "fs_case_set_restart_capture")

    def write_solution(self, path, patch_sizes):
        """The reference's restart file (writeSolution) of the current state."""
        sizes = np.ascontiguousarray(patch_sizes, dtype=np.int32)
        _check(lib().fs_case_write_solution(self.h, os.fsencode(path), len(sizes), _iptr(sizes)),
               "fs_case_write_solution")

    def read_solution(self, path, patch_sizes):
        """Resume from a restart file in the reference's layout."""
        sizes = np.ascontiguousarray(patch_sizes, dtype=np.int32)
        _check(lib().fs_case_read_solution(self.h, os.fsencode(path), len(sizes), _iptr(sizes)),
               "fs_case_read_solution")

    def set_decompose(self, method):
        """MethodDecomposeS 1, 2 or 3 (default 3)."""
        _check(lib().fs_case_set_decompose(self.h, method), "fs_case_set_decompose")

    def set_algorithm(self, algorithm):
        """Algorithm: "simple" (default) or "simplec"."""
        _check(lib().fs_case_set_algorithm(self.h, {"simple": 0, "simplec": 1}[algorithm]),
               "fs_case_set_algorithm")
        self._simplec = algorithm == "simplec"

    def set_gradient_limiters(self, momentum=0, continuity=0):
        """LimitGradientMomentumMethod / LimitGradientContinuityMethod: 0 off,
        1 or 2 (LimitGradients.for)."""
        _check(lib().fs_case_set_gradient_limiters(self.h, momentum, continuity),
               "fs_case_set_gradient_limiters")

    def topology(self, name):
        """Readit's topology arrays of the case's mesh (reference numbering):
        "nef", "gef", "enbr" ((W, N) = ElementNeighbor), "iown", "inbr"."""
        n = {"nef": self.N, "gef": self.N * self.W, "enbr": self.N * self.W, "iown": self.Fi, "inbr": self.Fi}[name]
        out = np.zeros(n, np.int32)
        _check(lib().fs_case_topology(self.h, TOPOLOGY[name], _iptr(out), n), f"fs_case_topology({name})")
        return out.reshape(self.W, self.N) if name in ("gef", "enbr") else out
