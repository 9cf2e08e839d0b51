"""The Fortran drop-in SolveEquation on the MI355X: the reference's driver
code calls the device through ISO_C_BINDING and gets the solution the
reference's own solvers converge to.  FSAMD_SOLVE_LOG makes the shim record
which side ran each call: the plain solvers and the algebraic multigrid with
V cycles must run on the device, 'direct' and the F and W cycles on the
host, where they are the reference's SolveEquation bitwise."""
import os

import numpy as np
import pytest

import pyoracle as po
import systems
from conftest import golden

DEV = po.REF_SOLVE.replace("ref_solve.x", "ref_solve_device.x")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DEV), reason="drop-in binary not built")]


def _run(g, solver, lmg, exe=DEV, **kw):
    old = po.REF_SOLVE
    po.REF_SOLVE = exe
    try:
        return po.run_reference(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], g["fit"], 1e-8, 200, solver,
                                mode=0, lmg=lmg, minparents=2, **kw)
    finally:
        po.REF_SOLVE = old


@pytest.fixture
def log(monkeypatch, tmp_path):
    p = tmp_path / "solve.log"
    monkeypatch.setenv("FSAMD_SOLVE_LOG", str(p))
    return p


def _sides(log):
    return [ln.split()[0] for ln in log.read_text().splitlines()]


@pytest.mark.parametrize("case", ["pshift", "momentum"])
def test_device_bicgstab_amg_through_fortran(case, log):
    g = golden(f"solve_{case}.npz")
    r = _run(g, 5, 1)[0]  # 'bicgs' with LMultigrid -> AMG-preconditioned BiCGSTAB on the device
    A = systems.dense(g["nbr"], g["nnbr"], g["ac"], g["anb"])
    x = np.linalg.solve(A, g["bc"])
    assert np.abs(g["bc"] - A @ r["dphi"]).sum() <= 1.0001e-8 * np.abs(g["bc"]).sum()
    np.testing.assert_allclose(r["dphi"], x, rtol=0, atol=1e-6 * np.abs(x).max())
    np.testing.assert_allclose(r["fit"], g["fit"] + r["dphi"], rtol=1e-15)
    np.testing.assert_allclose(r["res"], g["sor_res"], rtol=1e-13)  # initial norms do not depend on the solver
    side, var, solver, its = log.read_text().split()
    assert (side, var, solver) == ("device", "pressc", "bicgs") and int(its) > 0


def test_device_pbcg_through_fortran_follows_reference(log):
    """The reference's PBCG stagnates on this system near 1e-7 (its 1e-10
    step-length regularisers, PBCGMethod.for:183,204); the device runs the same
    algorithm and must end at the same residual level and solution."""
    g = golden("solve_momentum.npz")
    r = _run(g, 2, 0)[0]
    A = systems.dense(g["nbr"], g["nnbr"], g["ac"], g["anb"])
    l1 = np.abs(g["bc"] - A @ r["dphi"]).sum()
    assert l1 <= 2.0 * g["pbcg_hist"][-1]
    x = np.linalg.solve(A, g["bc"])
    np.testing.assert_allclose(r["dphi"], x, rtol=0, atol=1e-5 * np.abs(x).max())
    assert _sides(log) == ["device"]


def test_vcycle_multigrid_solver_runs_on_the_device(log):
    """LMultigrid with 'sor' and V cycles: the device's AMG as the solver
    (with the reference's pre/post sweep counts and cycle cap) on both calls
    ends at the reference's criterion or below the residual the reference's
    own AlgebraicMultigrid leaves (about 1e-3 of the initial one here)."""
    g = golden("solve_pshift.npz")
    kw = dict(npre=2, npost=3, ncalls=2)
    rs = _run(g, 0, 1, **kw)
    ref = _run(g, 0, 1, exe=po.REF_SOLVE, **kw)
    A = systems.dense(g["nbr"], g["nnbr"], g["ac"], g["anb"])
    for k, (r, q) in enumerate(zip(rs, ref)):
        b = g["bc"] * (1.0 + 0.25 * k)
        l1, l1_ref = np.abs(b - A @ r["dphi"]).sum(), np.abs(b - A @ q["dphi"]).sum()
        print(f"call {k}: device {l1 / np.abs(b).sum():.2e}, reference {l1_ref / np.abs(b).sum():.2e}", flush=True)
        assert l1 <= max(1.0001e-8 * np.abs(b).sum(), l1_ref)
    assert _sides(log) == ["device", "device"]


@pytest.mark.parametrize("cycle", [1, 2])
def test_f_and_w_cycles_run_the_reference_host_code(cycle, log):
    """F (1) and W (2) cycles are not on the device: the shim calls the
    reference's AlgebraicMultigrid, bitwise its own SolveEquation."""
    g = golden("solve_pshift.npz")
    kw = dict(cycle=cycle, npre=2, npost=3, ncalls=2)
    mine = _run(g, 0, 1, **kw)
    ref = _run(g, 0, 1, exe=po.REF_SOLVE, **kw)
    for a, b in zip(mine, ref):
        for k in ("dphi", "fit", "res"):
            assert np.array_equal(a[k], b[k]), k
    assert _sides(log) == ["host", "host"]


def test_direct_runs_the_reference_host_code(log):
    g = golden("solve_pneumann.npz")
    mine = _run(g, 3, 0)[0]
    ref = _run(g, 3, 0, exe=po.REF_SOLVE)[0]
    for k in ("dphi", "fit", "res"):
        assert np.array_equal(mine[k], ref[k]), k
    assert _sides(log) == ["host"]
