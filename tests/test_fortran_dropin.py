"""The Fortran drop-in (flowsolversourcecode_amd/fortran/SolveEquationDevice.f90)
linked with the reference's own solver sources in place of
SolveAlegbraicEquations.for (oracle/build_ref.sh -> oracle/_ref/ref_solve_device.x).

Without a device the shim must fall back to the reference's host solvers and
reproduce SolveEquation bitwise; on the MI355X (test_gpu_fortran_dropin.py)
it must solve through the C ABI."""
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import golden

DEV = po.REF_SOLVE.replace("ref_solve.x", "ref_solve_device.x")
pytestmark = pytest.mark.skipif(not os.path.exists(DEV), reason="reference build not present (build container only)")


def _run(exe, g, solver, **kw):
    old = po.REF_SOLVE
    po.REF_SOLVE = exe
    try:
        return po.run_reference(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], g["fit"], 1e-8, 200, solver,
                                mode=0, **kw)[0]
    finally:
        po.REF_SOLVE = old


@pytest.mark.parametrize("case", ["pneumann", "momentum"])
@pytest.mark.parametrize("solver,tag", [(po.SOR, "sor"), (po.ILU, "ilu"), (po.PBCG, "pbcg")])
def test_host_fallback_is_the_reference(case, solver, tag, monkeypatch):
    monkeypatch.setenv("HIP_VISIBLE_DEVICES", "-1")  # no device: fs_mesh_create fails, the shim falls back
    g = golden(f"solve_{case}.npz")
    r = _run(DEV, g, solver)
    assert np.array_equal(r["dphi"], g[f"{tag}_dphi"])
    assert np.array_equal(r["fit"], g[f"{tag}_fit"])
    assert np.array_equal(r["res"], g[f"{tag}_res"])
