"""The Fortran drop-in SolveEquation on the MI355X: the reference's driver
code calls the device through ISO_C_BINDING and gets the solution the
reference's own solvers converge to."""
import os

import numpy as np
import pytest

import pyoracle as po
import systems
from conftest import golden

DEV = po.REF_SOLVE.replace("ref_solve.x", "ref_solve_device.x")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DEV), reason="drop-in binary not built")]


def _run(g, solver, lmg):
    old = po.REF_SOLVE
    po.REF_SOLVE = DEV
    try:
        return po.run_reference(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], g["fit"], 1e-8, 200, solver,
                                mode=0, lmg=lmg, minparents=2)[0]
    finally:
        po.REF_SOLVE = old


@pytest.mark.parametrize("case", ["pshift", "momentum"])
def test_device_bicgstab_amg_through_fortran(case):
    g = golden(f"solve_{case}.npz")
    r = _run(g, 5, 1)  # 'bicgs' with LMultigrid -> AMG-preconditioned BiCGSTAB on the device
    A = systems.dense(g["nbr"], g["nnbr"], g["ac"], g["anb"])
    x = np.linalg.solve(A, g["bc"])
    assert np.abs(g["bc"] - A @ r["dphi"]).sum() <= 1.0001e-8 * np.abs(g["bc"]).sum()
    np.testing.assert_allclose(r["dphi"], x, rtol=0, atol=1e-6 * np.abs(x).max())
    np.testing.assert_allclose(r["fit"], g["fit"] + r["dphi"], rtol=1e-15)
    np.testing.assert_allclose(r["res"], g["sor_res"], rtol=1e-13)  # initial norms do not depend on the solver


def test_device_pbcg_through_fortran_follows_reference():
    """The reference's PBCG stagnates on this system near 1e-7 (its 1e-10
    step-length regularisers, PBCGMethod.for:183,204); the device runs the same
    algorithm and must end at the same residual level and solution."""
    g = golden("solve_momentum.npz")
    r = _run(g, 2, 0)
    A = systems.dense(g["nbr"], g["nnbr"], g["ac"], g["anb"])
    l1 = np.abs(g["bc"] - A @ r["dphi"]).sum()
    assert l1 <= 2.0 * g["pbcg_hist"][-1]
    x = np.linalg.solve(A, g["bc"])
    np.testing.assert_allclose(r["dphi"], x, rtol=0, atol=1e-5 * np.abs(x).max())
