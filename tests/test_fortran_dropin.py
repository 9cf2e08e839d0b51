"""The Fortran drop-in (flowsolversourcecode_amd/fortran/SolveEquationDevice.f90)
linked with the reference's own solver sources in place of
SolveAlegbraicEquations.for (oracle/build_ref.sh -> oracle/_ref/ref_solve_device.x).

Without a device the shim must fall back to the reference's host code and
reproduce SolveEquation bitwise for every solver string and multigrid cycle
(oracle/_ref/ref_solve.x is the reference's own SolveEquation linked with the
same driver); on the MI355X (test_gpu_fortran_dropin.py) it must solve
through the C ABI.  FSAMD_SOLVE_LOG names a file where the shim records which
side ran each call."""
import os

import numpy as np
import pytest

import pyoracle as po
import systems
from conftest import golden

DEV = po.REF_SOLVE.replace("ref_solve.x", "ref_solve_device.x")
pytestmark = pytest.mark.skipif(not os.path.exists(DEV), reason="reference build not present (build container only)")


def _run(exe, g, solver, **kw):
    old = po.REF_SOLVE
    po.REF_SOLVE = exe
    try:
        return po.run_reference(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], g["fit"], 1e-8, 200, solver,
                                mode=0, **kw)
    finally:
        po.REF_SOLVE = old


@pytest.fixture
def no_device(monkeypatch, tmp_path):
    monkeypatch.setenv("HIP_VISIBLE_DEVICES", "-1")  # no device: fs_mesh_create fails, the shim falls back
    log = tmp_path / "solve.log"
    monkeypatch.setenv("FSAMD_SOLVE_LOG", str(log))
    return log


def _same(a, b):
    for k in ("dphi", "fit", "res"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("case", ["pneumann", "momentum"])
@pytest.mark.parametrize("solver,tag", [(po.SOR, "sor"), (po.ILU, "ilu"), (po.PBCG, "pbcg")])
def test_host_fallback_is_the_reference(case, solver, tag, no_device):
    g = golden(f"solve_{case}.npz")
    r = _run(DEV, g, solver)[0]
    assert np.array_equal(r["dphi"], g[f"{tag}_dphi"])
    assert np.array_equal(r["fit"], g[f"{tag}_fit"])
    assert np.array_equal(r["res"], g[f"{tag}_res"])
    assert no_device.read_text().split() == ["host", "pressc", tag, "0"]


@pytest.mark.parametrize("solver,tag", [(3, "direct"), (4, "gmres"), (5, "bicgs")])
def test_host_fallback_direct_and_parallel_solvers_are_the_reference(solver, tag, no_device, monkeypatch):
    """'direct' (SetupMatrixforDirectSolver, ludcmpF77, lubksbF77) and the
    OpenMP 'gmres'/'bicgs' (SolveAlgebraicSystemInParallel, one thread so
    that its reductions have one order): the shim's host branch equals the
    reference's SolveEquation bitwise."""
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    g = golden("solve_pneumann.npz" if tag == "direct" else "solve_pshift.npz")
    ref = _run(po.REF_SOLVE, g, solver)[0]
    _same(_run(DEV, g, solver)[0], ref)
    assert no_device.read_text().split()[:3] == ["host", "pressc", tag]
    if tag == "direct":
        A = systems.dense(g["nbr"], g["nnbr"], g["ac"], g["anb"])
        np.testing.assert_allclose(A @ ref["dphi"], g["bc"], rtol=0, atol=1e-9 * np.abs(g["bc"]).max())


@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_host_fallback_multigrid_is_the_reference(cycle, no_device):
    """LMultigrid with the algebraic agglomeration, V (0), F (1) and W (2)
    cycles, two calls (the second reuses the hierarchy): the reference's
    AlgebraicMultigrid, bitwise."""
    g = golden("solve_pshift.npz")
    kw = dict(lmg=1, cycle=cycle, npre=2, npost=3, minparents=2, ncalls=2)
    ref = _run(po.REF_SOLVE, g, po.SOR, **kw)
    mine = _run(DEV, g, po.SOR, **kw)
    for a, b in zip(mine, ref):
        _same(a, b)
    assert [ln.split()[0] for ln in no_device.read_text().splitlines()] == ["host", "host"]
